"""Host logic of the pipelines (no device): the ligand index reader
(source.cpp:27-44), file stems, torsion reordering (data.h:53-61), the
result JSON layout (writer.cpp:40-81), the bdock / bdopt / bdscr option
checks (engine/src/cli/bdock.cpp, bdopt.cpp, bdscr.cpp) and the facade's
argument handling."""
import json
import logging

import numpy as np
import pytest

from pxdock_amd import cli, data, docking, engine


def test_read_index_skips_empty_lines_and_ghost_files(tmp_path, caplog):
    a = tmp_path / "a.json"
    a.write_text("{}")
    b = tmp_path / "b.x.json"
    b.write_text("{}")
    idx = tmp_path / "ligands.index"
    idx.write_text(f"{a}\n\n{tmp_path / 'ghost.json'}\n{b}\n")
    with caplog.at_level(logging.WARNING, logger="pxdock"):
        assert engine.read_index(str(idx)) == [str(a), str(b)]
    assert "line#2 with empty text" in caplog.text and "line#3 with ghost file" in caplog.text
    assert engine.file_stem(str(b)) == "b.x" and engine.file_stem("/x/y/lig_1.json") == "lig_1"


def test_torsions_go_back_to_the_json_bond_order():
    lv = data.Leaves(origin=np.zeros((3, 3)), axis=np.eye(3), atom_off=np.array([0, 1, 2, 3]),
                     atoms=np.array([5, 6, 7]), bond_order=np.array([2, 0, 4]))
    assert engine._aligned_torsions(lv, [0.1, 0.2, 0.3]) == [0.2, 0.0, 0.1, 0.0, 0.3]
    assert engine._aligned_torsions(None, []) == []


def test_result_json_layout(tmp_path):
    poses = [dict(offset=3, energy=-12.5, pscore=-7.25, nevals=40, torsions=[0.5, -1.0],
                  xyz=np.arange(6.0).reshape(2, 3))]
    path = engine.write_result(str(tmp_path), "lig", "[C:1][O:2]", 0, -9.5, poses)
    doc = json.load(open(path))
    assert list(doc) == ["mapped_smiles", "best_pose", "poses"]
    assert doc["best_pose"] == {"index": 0, "bscore": -9.5}
    p = doc["poses"][0]
    assert list(p) == ["offset", "energy", "pscore", "nevals", "receptor", "ligand"]
    assert p["receptor"]["torsions"] == [0.5, -1.0] and p["ligand"]["xyz"][1] == [3.0, 4.0, 5.0]
    no_smiles = engine.write_result(str(tmp_path), "lig2", "", 0, 1.0, poses)
    assert "mapped_smiles" not in json.load(open(no_smiles))


BOX = ["--center_x", "1", "--center_y", "2", "--center_z", "3", "--size_x", "20",
       "--size_y", "20", "--size_z", "20"]


def test_bdock_options():
    base = ["--receptor_file", "r.json", "--ligand_index", "l.index"]
    _, a = cli.parse_args("bdock", base + BOX)
    assert (a.seed, a.exhaustiveness, a.max_nsteps, a.relax_nsteps, a.num_modes) == \
        (0, 256, 200, 40, 8)
    assert a.min_rmsd == 0.5 and a.mc_prune_energy_threshold == -1.0 and a.spacing == 0.175
    assert a.penalty == 1e6 and a.output_dir == "docked"
    with pytest.raises(ValueError, match="ligand_index"):
        cli.parse_args("bdock", ["--receptor_file", "r.json"] + BOX)
    with pytest.raises(ValueError, match="center_y"):
        cli.parse_args("bdock", base + BOX[:2] + BOX[4:])
    # a map folder replaces the box
    _, a = cli.parse_args("bdock", base + ["--map_dir", "maps"])
    assert a.map_dir == "maps"
    with pytest.raises(ValueError, match="exhaustiveness"):
        cli.parse_args("bdock", base + BOX + ["--exhaustiveness", "0"])


def test_bdopt_and_bdscr_options(tmp_path):
    base = ["--receptor_file", "r.json", "--ligand_index", "l.index"]
    _, a = cli.parse_args("bdopt", base + BOX)
    assert a.max_niters == 240 and a.penalty == 1e6
    # no maps: no box needed (bdopt.cpp:215-220)
    _, a = cli.parse_args("bdopt", base + ["--spacing", "0"])
    assert a.center_x is None
    _, a = cli.parse_args("bdscr", base)
    assert a.output_file == "scores.csv"
    cfg = tmp_path / "opt.conf"
    cfg.write_text("receptor_file = r.json\nligand_index = l.index\nmax_niters = 120\n"
                   "spacing = 0\n")
    _, a = cli.parse_args("bdopt", ["--config", str(cfg), "--max_niters", "60"])
    assert a.max_niters == 60 and a.spacing == 0.0 and a.ligand_index == "l.index"
    assert cli.main(["bdscr", "--version"]) == 0
    assert cli.main(["bdscr", "--receptor_file", "r.json"]) == 1
    assert cli.main(["nothing"]) == 1


def test_facade_needs_a_prepared_receptor(tmp_path):
    d = docking.ProtenixDock(str(tmp_path / "protein.pdb"), work_dir=str(tmp_path))
    with pytest.raises(RuntimeError, match="not part of this package"):
        d.set_box([0, 0, 0], [10, 10, 10])
    with pytest.raises(RuntimeError, match="set_box"):
        d.generate_cache_maps()
    idx = d._index([str(tmp_path / "a.json"), str(tmp_path / "b.json")], "mc")
    assert open(idx).read().splitlines() == [str(tmp_path / "a.json"), str(tmp_path / "b.json")]
    assert d._index(str(tmp_path / "given.index"), "mc") == str(tmp_path / "given.index")
