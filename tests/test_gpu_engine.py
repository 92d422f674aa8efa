"""The batch pipelines end to end on the device (ReusableEngine,
engine/src/core/entrance.cpp:148-420; the bdscr / bdopt / bdock tools and the
ProtenixDock facade): reference-format JSON in, reference-format CSV / JSON
out, every written score checked against the FP64 oracle on the written pose.

Tolerances: pscore 1e-4 max(|s|, 1) (the pscore kernel's); bscore the same
plus 1e-9 of the GBSA energies (FP64)."""
import csv
import dataclasses
import itertools
import json
import math
import os

import numpy as np
import pytest

import pyoracle as O
from pxdock_amd import cli, data, docking, engine, scoring, synth

pytestmark = pytest.mark.gpu


def _f32(x):
    return np.asarray(np.asarray(x, np.float32), np.float64)


@pytest.fixture(scope="module")
def world(tmp_path_factory):
    import torch
    assert torch.cuda.is_available(), "a GPU is required"
    d = tmp_path_factory.mktemp("engine")
    rec = synth.make_receptor(3, n_heavy=500, radius=16.0, pocket_offset=7.0, n_ring6=3,
                              n_ring5=1)
    rec_file = d / "receptor.json"
    rec_file.write_text(json.dumps(data.receptor_json(rec)))
    rng = np.random.default_rng(7)
    ligs, files = [], []
    for i in range(3):
        lig = synth.make_ligand(700 + i, n_heavy=14 + 6 * i, n_torsions=i + 2, n_rings=1)
        poses = []
        for _ in range(3):
            c = rec.pocket_center + rng.normal(0.0, 1.0, 3)
            R = synth.quat_to_rot(synth.random_quaternion(rng))
            t = rng.uniform(-math.pi, math.pi, lig.tree.n_torsions)
            poses.append(synth.ligand_apply(lig.tree, lig.xyz[0], c, R, t))
        lig = dataclasses.replace(lig, xyz=np.stack(poses))
        f = d / f"lig_{i}.json"
        f.write_text(json.dumps(data.ligand_json(lig)))
        ligs.append(data.parse_ligand(str(f)))
        files.append(str(f))
    index = d / "ligands.index"
    index.write_text("\n".join(files) + "\n\n" + str(d / "missing.json") + "\n")
    return dict(d=d, rec=data.parse_receptor(str(rec_file)), rec_file=str(rec_file),
                center=[float(v) for v in rec.pocket_center],
                ligs=ligs, files=files, index=str(index),
                sp=scoring.load_pose_selection(), bp=scoring.load_affinity_ranking())


def _check_scores(w, lig, xyz, tors, pscore, bscore):
    rec = w["rec"]
    rx = O.receptor_apply(rec.leaves, rec.xyz, tors) if len(tors) else rec.xyz
    ps, _ = O.pscore(w["sp"], rec, rx, lig, _f32(xyz))
    assert abs(pscore - ps) <= 1e-4 * max(abs(ps), 1.0), (pscore, ps)
    if bscore is not None:
        bs, bt = O.bscore(w["bp"], rec, rx, lig, _f32(xyz))
        tol = 1e-4 * max(abs(bs), 1.0) * 10 + 1e-9 * float(np.abs(bt[-3:]).sum())
        assert abs(bscore - bs) <= tol, (bscore, bs)


def test_gpu_evaluate_and_bdscr(world, tmp_path):
    out = str(tmp_path / "scores.csv")
    eng = engine.ReusableEngine(world["rec_file"])
    assert eng.evaluate(world["index"], out) == 9
    rows = list(csv.reader(open(out)))
    assert rows[0] == ["ligand_file", "pose_id", "pscore", "bscore"]
    assert [(r[0], int(r[1])) for r in rows[1:]] == \
        [(f"lig_{i}", p) for i in range(3) for p in range(3)]
    for r in rows[1:]:
        lig = world["ligs"][int(r[0][-1])]
        _check_scores(world, lig, lig.xyz[int(r[1])], [], float(r[2]), float(r[3]))
    out2 = str(tmp_path / "cli.csv")
    assert cli.main(["bdscr", "--receptor_file", world["rec_file"], "--ligand_index",
                     world["index"], "--output_file", out2, "--verbosity", "0"]) == 0
    assert open(out2).read() == open(out).read()


def _leaf_tors(w, doc_tors):
    L = w["rec"].leaves
    order = L.bond_order if L.bond_order is not None else np.arange(L.n_leaves)
    return np.array([doc_tors[int(order[i])] for i in range(L.n_leaves)])


def test_gpu_optimize_on_maps(world, tmp_path):
    eng = engine.ReusableEngine(world["rec_file"])
    eng.set_box(*world["center"], 14.0, 14.0, 14.0)
    eng.generate_maps(0.375)
    out = str(tmp_path / "opt")
    files = eng.optimize(world["index"], out, max_niters=60)
    assert sorted(os.path.basename(f) for f in files) == [f"lig_{i}_out.json" for i in range(3)]
    for i, lig in enumerate(world["ligs"]):
        doc = json.load(open(os.path.join(out, f"lig_{i}_out.json")))
        assert doc["mapped_smiles"] == lig.smiles
        poses = doc["poses"]
        assert [p["offset"] for p in poses] == [0, 1, 2]
        ps = [p["pscore"] for p in poses]
        assert doc["best_pose"]["index"] == int(np.argmin(ps))
        for k, p in enumerate(poses):
            tors = _leaf_tors(world, p["receptor"]["torsions"])
            best = k == doc["best_pose"]["index"]
            _check_scores(world, lig, np.array(p["ligand"]["xyz"]), tors, p["pscore"],
                          doc["best_pose"]["bscore"] if best else None)
            assert p["nevals"] >= 1 and np.isfinite(p["energy"])


def test_gpu_search_and_facade(world, tmp_path):
    d = docking.ProtenixDock(world["rec_file"], work_dir=str(tmp_path))
    d.set_box(world["center"], [14.0, 14.0, 14.0])
    d.generate_cache_maps(spacing=0.375, output_dir=str(tmp_path / "maps"))
    d.drop_cache_maps()
    d.restore_cache_maps(str(tmp_path / "maps"))
    eng = d._engine
    out = str(tmp_path / "dock")
    files = eng.search(world["index"], out, seed=3, exhaustiveness=8, max_nsteps=4,
                       relax_nsteps=10, num_modes=4, min_rmsd=0.5)
    assert len(files) == 3
    for i, lig in enumerate(world["ligs"]):
        doc = json.load(open(os.path.join(out, f"lig_{i}_out.json")))
        poses = doc["poses"]
        assert 1 <= len(poses) <= 4 and doc["best_pose"]["index"] == 0
        ps = [p["pscore"] for p in poses]
        assert ps == sorted(ps)
        xs = [np.array(p["ligand"]["xyz"]) for p in poses]
        for a, b in itertools.combinations(xs, 2):
            assert np.sqrt(((a - b) ** 2).sum(1).mean()) >= 0.5
        assert len({p["offset"] for p in poses}) == len(poses)
        _check_scores(world, lig, xs[0], _leaf_tors(world, poses[0]["receptor"]["torsions"]),
                      ps[0], doc["best_pose"]["bscore"])
    out2 = d.run_docking(world["files"][:1], num_walker=8, relax_nsteps=10)
    assert os.listdir(out2) == ["lig_0_out.json"]
