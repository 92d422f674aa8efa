"""Map directories on the host: the format round trip, the failure cases of
receptor_cache::restore (engine/src/core/grid.cpp:346-431: metadata missing,
version mismatch, a map file missing) and this format's own checks
(truncation, checksums, another receptor, other movable atoms, incomplete
metadata), the receptor fingerprint, and the bdgen option checks
(engine/src/cli/bdgen.cpp:28-186).  No device is used."""
import os
import re
import struct

import numpy as np
import pytest

from pxdock_amd import bdgen, ffi, lib, synth

REF_CONSTANT = "/root/reference/engine/include/bytedock/core/constant.h"

IO, CONF, ARG = 6, 2, 1


def _maps(box, n_types, seed=0):
    rng = np.random.default_rng(seed)
    m = rng.normal(0.0, 50.0, (n_types + 1, box.n_points(), 4)).astype(np.float32)
    m[0, 0] = [3.0e38, -3.0e38, 0.0, -0.0]  # values the build clamps to keep finite
    return m


@pytest.fixture()
def saved(tmp_path):
    box = lib.decide_grid_dims([1.0, -2.0, 3.5], [3.0, 2.25, 1.5], 0.375)
    types = [4, 5, 33, 62]  # c3, ca, hx, ni: ascending GAFF2 indices
    maps = _maps(box, len(types))
    excluded = [3, 7, 8, 20]
    key = 0x0123456789ABCDEF
    d = str(tmp_path / "maps")
    lib.maps_write(d, box, types, maps, 40, excluded, key, slope=1e6)
    return dict(dir=d, box=box, types=types, maps=maps, excluded=excluded, key=key)


def _status(fn, *a, **k):
    with pytest.raises(lib.PxdError) as e:
        fn(*a, **k)
    return e.value.status, str(e.value)


def test_type_tags_follow_the_reference_table():
    tags = lib.type_tags()
    assert len(tags) == ffi.NUM_VDW_TYPES and len(set(tags)) == len(tags)
    assert tags[:5] == ["c", "cs", "c1", "c2", "c3"] and tags[-1] == "sq"
    assert lib.load().pxd_vdw_type_tag(ffi.NUM_VDW_TYPES) == b"coul"
    assert lib.load().pxd_vdw_type_tag(-1) is None
    if not os.path.exists(REF_CONSTANT):
        pytest.skip("reference tree absent")
    text = open(REF_CONSTANT).read()
    body = text[text.index("kVdwTypeTags"):]
    ref = re.findall(r'"([^"]*)"', body[:body.index("};")])
    assert ref == tags + ["coul"]


def test_round_trip_is_bit_identical(saved):
    meta, ex = lib.maps_info(saved["dir"])
    assert meta.type_list == saved["types"]
    assert list(ex) == saved["excluded"]
    assert meta.receptor_key == saved["key"] and meta.n_atoms == 40 and meta.slope == 1e6
    b = saved["box"]
    assert meta.box.npts == b.npts and meta.box.spacing == b.spacing
    for d in range(3):  # FP64 box fields survive the text record exactly
        assert meta.box.init_xyz[d] == b.init_xyz[d]
        assert meta.box.oppo_xyz[d] == b.oppo_xyz[d]
    got = lib.maps_read(saved["dir"])
    assert got.tobytes() == saved["maps"].tobytes()
    # a subset in a requested order: Coulomb first, then the types as asked
    sub = lib.maps_read(saved["dir"], types=[33, 4])
    want = saved["maps"][[0, 3, 1]]
    assert sub.tobytes() == want.tobytes()
    only = lib.maps_read(saved["dir"], types=[])
    assert only.tobytes() == saved["maps"][:1].tobytes()
    files = sorted(os.listdir(saved["dir"]))
    assert files == sorted(["coul.pxm", "c3.pxm", "ca.pxm", "hx.pxm", "ni.pxm", "meta.pxd"])


def test_map_file_layout_is_as_documented(saved):
    path = os.path.join(saved["dir"], "c3.pxm")
    raw = open(path, "rb").read()
    magic, ver, typ, nx, ny, nz, ch, nbytes, csum = struct.unpack_from("<8sIi3iiQQ", raw, 0)
    assert magic == b"PXDMAP\x00\x01" and ver == ffi.MAPS_VERSION and typ == 4
    assert (nx, ny, nz) == saved["box"].npts and ch == 4
    P = saved["box"].n_points()
    assert nbytes == P * 16 and len(raw) == 64 + nbytes
    # FNV-1a 64 over little-endian 64-bit words
    h = 0xcbf29ce484222325
    for (w,) in struct.iter_unpack("<Q", raw[64:]):
        h = ((h ^ w) * 0x100000001b3) & 0xFFFFFFFFFFFFFFFF
    assert h == csum
    pts = np.frombuffer(raw[64:], np.float32).reshape(P, 4)
    assert np.array_equal(pts, saved["maps"][1])
    meta = open(os.path.join(saved["dir"], "meta.pxd")).read().splitlines()
    assert meta[0] == f"pxdock-maps {ffi.MAPS_VERSION}" and meta[-1] == "end"
    assert f"map 4 c3.pxm {csum:016x}" in meta


def test_missing_metadata(saved):
    os.remove(os.path.join(saved["dir"], "meta.pxd"))
    st, msg = _status(lib.maps_read, saved["dir"])
    assert st == IO and "metadata file of maps is missing" in msg


def test_version_mismatch(saved):
    p = os.path.join(saved["dir"], "meta.pxd")
    text = open(p).read().replace(f"pxdock-maps {ffi.MAPS_VERSION}", "pxdock-maps 20241030", 1)
    open(p, "w").write(text)
    st, msg = _status(lib.maps_info, saved["dir"])
    assert st == CONF and "Expected [1] vs. Got [20241030]" in msg


def test_map_file_version_mismatch(saved):
    p = os.path.join(saved["dir"], "ca.pxm")
    raw = bytearray(open(p, "rb").read())
    raw[8:12] = struct.pack("<I", 2)
    open(p, "wb").write(bytes(raw))
    st, msg = _status(lib.maps_read, saved["dir"])
    assert st == CONF and "not compatible" in msg


def test_missing_map_file(saved):
    os.remove(os.path.join(saved["dir"], "ca.pxm"))
    st, msg = _status(lib.maps_read, saved["dir"])
    assert st == IO and "Map file for atom type [ca] is missing" in msg
    # the other maps still load
    assert lib.maps_read(saved["dir"], types=[4]).shape[0] == 2


def test_type_not_saved(saved):
    st, msg = _status(lib.maps_read, saved["dir"], types=[4, 0])
    assert st == IO and "[c]" in msg


def test_truncated_and_corrupted_maps(saved):
    p = os.path.join(saved["dir"], "hx.pxm")
    raw = open(p, "rb").read()
    open(p, "wb").write(raw[:-16])
    st, msg = _status(lib.maps_read, saved["dir"])
    assert st == IO and "truncated" in msg
    bad = bytearray(raw)
    bad[64 + 1000] ^= 0x40
    open(p, "wb").write(bytes(bad))
    st, msg = _status(lib.maps_read, saved["dir"])
    assert st == IO and "checksum" in msg
    open(p, "wb").write(raw + b"\0")
    st, msg = _status(lib.maps_read, saved["dir"])
    assert st == IO and "trailing" in msg
    open(p, "wb").write(raw[:40])
    st, msg = _status(lib.maps_read, saved["dir"])
    assert st == IO and "no header" in msg


def test_incomplete_or_malformed_metadata(saved):
    p = os.path.join(saved["dir"], "meta.pxd")
    lines = open(p).read().splitlines()
    open(p, "w").write("\n".join(lines[:-1]) + "\n")  # no end record
    st, msg = _status(lib.maps_info, saved["dir"])
    assert st == CONF and "incomplete" in msg
    body = [ln for ln in lines if not ln.startswith("map coul")]
    open(p, "w").write("\n".join(body) + "\n")  # a listed map without its record
    st, _ = _status(lib.maps_info, saved["dir"])
    assert st == CONF
    swapped = lines[:]
    swapped[1], swapped[2] = swapped[2], swapped[1]
    open(p, "w").write("\n".join(swapped) + "\n")
    st, _ = _status(lib.maps_info, saved["dir"])
    assert st == CONF
    sums = [ln for ln in lines]
    i = next(k for k, ln in enumerate(sums) if ln.startswith("map 5 "))
    head, tail = sums[i].rsplit(" ", 1)
    sums[i] = head + " " + ("0" * 16 if tail != "0" * 16 else "1" * 16)
    open(p, "w").write("\n".join(sums) + "\n")
    st, msg = _status(lib.maps_read, saved["dir"])
    assert st == CONF and "disagrees" in msg


def test_maps_of_another_receptor_are_refused(saved):
    d = saved["dir"]
    ok = lib.maps_read(d, expect=(40, saved["excluded"], saved["key"]))
    assert ok.tobytes() == saved["maps"].tobytes()
    st, msg = _status(lib.maps_read, d, expect=(40, saved["excluded"], saved["key"] ^ 1))
    assert st == CONF and "another receptor" in msg
    st, _ = _status(lib.maps_read, d, expect=(41, saved["excluded"], saved["key"]))
    assert st == CONF
    st, msg = _status(lib.maps_read, d, expect=(40, [3, 7, 9, 20], saved["key"]))
    assert st == CONF and "movable" in msg
    st, _ = _status(lib.maps_read, d, expect=(40, [3, 7, 8], saved["key"]))
    assert st == CONF


def test_saving_again_replaces_the_directory(saved):
    d = saved["dir"]
    box = saved["box"]
    maps2 = _maps(box, 1, seed=5)
    lib.maps_write(d, box, [9], maps2, 40, saved["excluded"], saved["key"])
    assert lib.maps_info(d)[0].type_list == [9]
    assert lib.maps_read(d).tobytes() == maps2.tobytes()
    assert not os.path.exists(os.path.join(d, "meta.pxd.tmp"))


def test_write_rejects_bad_metadata(tmp_path):
    box = lib.decide_grid_dims([0, 0, 0], [1.5] * 3, 0.375)
    m = _maps(box, 2)
    d = str(tmp_path / "x")
    st, _ = _status(lib.maps_write, d, box, [5, 4], m, 10, [], 1)  # not ascending
    assert st == ARG
    st, _ = _status(lib.maps_write, d, box, [4, 5], m, 10, [3, 3], 1)
    assert st == ARG
    st, _ = _status(lib.maps_write, d, box, [4, 5], m, 10, [10], 1)
    assert st == ARG
    assert not os.path.exists(os.path.join(d, "meta.pxd"))


def test_receptor_key_tracks_what_the_maps_depend_on():
    rec = synth.make_receptor(3, n_heavy=200, radius=11.0, pocket_offset=5.0, n_ring6=2)
    k0 = lib.receptor_key(rec)
    assert k0 == lib.receptor_key(rec)
    mov = rec.leaves.movable()
    fixed = [i for i in range(rec.n_atoms) if i not in set(mov)]
    assert len(mov) > 0
    # a movable atom's position does not enter the maps
    rec.xyz[mov[0], 0] += 0.5
    assert lib.receptor_key(rec) == k0
    rec.xyz[mov[0], 0] -= 0.5
    for arr, i in ((rec.xyz, (fixed[3], 1)), (rec.charge, fixed[5]), (rec.sigma, fixed[7]),
                   (rec.epsilon, fixed[9])):
        old = arr[i]
        arr[i] = old + 1e-12 * max(abs(old), 1.0)
        assert lib.receptor_key(rec) != k0
        arr[i] = old
    assert lib.receptor_key(rec) == k0


def test_bdgen_option_checks(tmp_path):
    base = ["--receptor_file", "r.json", "--center_x", "1", "--center_y", "2", "--center_z", "3",
            "--size_x", "10", "--size_y", "10", "--size_z", "12"]
    _, a = bdgen.parse_args(base)
    assert a.spacing == 0.375 and a.max_size == 0.0 and a.output_dir == "docked"
    assert a.type_list == list(range(99))
    for drop in ("--receptor_file", "--center_y", "--size_z"):
        i = base.index(drop)
        with pytest.raises(ValueError, match="required"):
            bdgen.parse_args(base[:i] + base[i + 2:])
    with pytest.raises(ValueError, match="spacing"):
        bdgen.parse_args(base + ["--spacing", "0.01"])
    with pytest.raises(ValueError, match="size_x"):
        bdgen.parse_args([x if x != "10" else "0" for x in base])
    with pytest.raises(ValueError, match="verbosity"):
        bdgen.parse_args(base + ["--verbosity", "3"])
    _, a = bdgen.parse_args(base + ["--types", "c3,ca,4,n+"])
    assert a.type_list == [4, 5, lib.type_tags().index("n+")]
    with pytest.raises(ValueError, match="unknown"):
        bdgen.parse_args(base + ["--types", "zz"])
    cfg = tmp_path / "gen.conf"
    cfg.write_text("# box\ncenter_x = 1\ncenter_y = 2\ncenter_z = 3\nsize_x = 10\n"
                   "size_y = 10\nsize_z = 12\nspacing = 0.5\nreceptor_file = r.json\n")
    _, a = bdgen.parse_args(["--config", str(cfg), "--spacing", "0.25"])
    assert a.center_z == 3.0 and a.size_z == 12.0 and a.spacing == 0.25
    assert a.receptor_file == "r.json"
    assert bdgen.main(["--version"]) == 0
    assert bdgen.main(base[2:]) == 1
