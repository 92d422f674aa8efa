"""Map directories on the device: build -> save -> load gives bit-identical
maps and bit-identical lookups (receptor_cache::save / restore,
engine/src/core/grid.cpp:305-431), loading refuses maps of another receptor,
and the bdgen entry point (engine/src/cli/bdgen.cpp:28-186) writes the maps
the library builds directly."""
import json
import math

import numpy as np
import pytest

from pxdock_amd import bdgen, data, lib, synth

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def built():
    import torch
    assert torch.cuda.is_available(), "a GPU is required"
    rec = synth.make_receptor(3, n_heavy=500, radius=16.0, pocket_offset=7.0, n_ring6=3,
                              n_ring5=1)
    ligs = [synth.make_ligand(900 + i, n_heavy=12 + 4 * i, n_torsions=i + 1, n_rings=1)
            for i in range(4)]
    types = sorted(set(int(t) for m in ligs for t in m.vdw_type))
    box = lib.decide_grid_dims(rec.pocket_center, [12.0, 13.5, 15.0], 0.375)
    R = lib.Receptor(rec)
    G = lib.Grid(R, box, types, slope=2.5e5)
    return dict(rec=rec, ligs=ligs, types=types, box=box, R=R, G=G)


def _poses(b, n_per=6, seed=5):
    rng = np.random.default_rng(seed)
    lo, hi = np.array(b["box"].init_xyz), np.array(b["box"].oppo_xyz)
    pose_lig, xyz = [], []
    for l, lig in enumerate(b["ligs"]):
        for k in range(n_per):
            c = rng.uniform(lo - 1.0, hi + 1.0)  # some atoms outside: the wall term
            Rm = synth.quat_to_rot(synth.random_quaternion(rng))
            t = rng.uniform(-math.pi, math.pi, lig.tree.n_torsions)
            xyz.append(synth.ligand_apply(lig.tree, lig.xyz[0], c, Rm, t))
            pose_lig.append(l)
    nl = b["rec"].leaves.n_leaves
    tors = rng.uniform(-math.pi, math.pi, (len(xyz), nl))
    return lib.PoseSet(lib.LigandBatch(b["ligs"]), pose_lig, xyz, nl, tors=tors)


def test_gpu_grid_build_is_deterministic(built):
    ref = built["G"].download().tobytes()
    for _ in range(3):
        again = lib.Grid(built["R"], built["box"], built["types"], slope=2.5e5)
        assert again.download().tobytes() == ref


def test_gpu_maps_save_load_round_trip(built, tmp_path):
    d = str(tmp_path / "maps")
    built["G"].save(d)
    print(f"\n[maps] saved {len(built['types']) + 1} maps of {built['box'].npts}", flush=True)
    meta, ex = lib.maps_info(d)
    assert meta.type_list == built["types"] and meta.slope == 2.5e5
    assert list(ex) == sorted(built["rec"].leaves.movable())
    assert meta.receptor_key == lib.receptor_key(built["rec"])
    L = lib.Grid.load(built["R"], d)
    assert list(L.types) == built["types"]
    assert L.download().tobytes() == built["G"].download().tobytes()
    # lookups through the loaded grid equal those through the built one, bit for bit
    P = _poses(built)
    e0, g0, t0 = lib.score_grad(built["R"], built["G"], P)
    e1, g1, t1 = lib.score_grad(built["R"], L, P)
    assert e0.cpu().numpy().tobytes() == e1.cpu().numpy().tobytes()
    assert g0.cpu().numpy().tobytes() == g1.cpu().numpy().tobytes()
    assert t0.cpu().numpy().tobytes() == t1.cpu().numpy().tobytes()
    # a subset of the saved types, and another slope
    sub = built["types"][::2]
    S = lib.Grid.load(built["R"], d, types=sub, slope=1e6)
    full = built["G"].download()
    idx = [0] + [1 + built["types"].index(t) for t in sub]
    assert S.download().tobytes() == full[idx].tobytes()
    assert lib.maps_info(d)[0].slope == 2.5e5


def test_gpu_maps_of_another_receptor_are_refused(built, tmp_path):
    d = str(tmp_path / "maps")
    built["G"].save(d)
    other = synth.make_receptor(3, n_heavy=500, radius=16.0, pocket_offset=7.0, n_ring6=3,
                                n_ring5=1)
    mov = set(other.leaves.movable())
    i = next(a for a in range(other.n_atoms) if a not in mov)
    other.xyz[i, 0] += 0.01
    with pytest.raises(lib.PxdError) as e:
        lib.Grid.load(lib.Receptor(other), d)
    assert e.value.status == 2 and "another receptor" in str(e.value)
    with pytest.raises(lib.PxdError) as e:
        lib.Grid.load(built["R"], d, types=[t for t in range(99) if t not in built["types"]][:1])
    assert e.value.status == 6


def test_gpu_bdgen_writes_the_library_maps(built, tmp_path):
    rec_file = tmp_path / "receptor.json"
    rec_file.write_text(json.dumps(data.receptor_json(built["rec"])))
    c = built["rec"].pocket_center
    out = str(tmp_path / "docked")
    tags = lib.type_tags()
    argv = ["--receptor_file", str(rec_file), "--output_dir", out, "--verbosity", "0",
            "--center_x", repr(float(c[0])), "--center_y", repr(float(c[1])),
            "--center_z", repr(float(c[2])), "--size_x", "12", "--size_y", "13.5",
            "--size_z", "15", "--types", ",".join(tags[t] for t in built["types"])]
    assert bdgen.main(argv) == 0
    meta, _ = lib.maps_info(out)
    assert meta.box.npts == built["box"].npts and meta.type_list == built["types"]
    # the JSON round trip of the receptor keeps its FP64 values: same key, same maps
    assert meta.receptor_key == lib.receptor_key(built["rec"])
    L = lib.Grid.load(built["R"], out, slope=2.5e5)
    assert L.download().tobytes() == built["G"].download().tobytes()
    # --max_size widens every dimension to at least that size (entrance.cpp:105-107)
    out2 = str(tmp_path / "wide")
    assert bdgen.main(argv[:-2] + ["--types", tags[built["types"][0]], "--output_dir", out2,
                                   "--max_size", "14"]) == 0
    wide = lib.maps_info(out2)[0].box
    want = lib.decide_grid_dims(c, [14.0, 14.0, 15.0], 0.375)
    assert wide.npts == want.npts
