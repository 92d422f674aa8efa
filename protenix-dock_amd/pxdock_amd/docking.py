"""ProtenixDock: the reference's Python docking facade
(pxdock/dock/docking.py:30-180) over ReusableEngine on the GPU.

Same methods, arguments and defaults.  Input preparation (PDB -> receptor
JSON with AmberTools, SDF -> ligand JSON with RDKit, Vina pre-poses) needs
tools this package does not carry: the facade takes prepared JSON files.  A
receptor given as anything but a JSON file needs a `receptor_preparer`
callable (pdb, pocket_center, box_size, cofactor_fpath) -> JSON path, and
ligands are prepared ligand JSON files (a list, or an index file with one
path per line) instead of an SDF.
"""
import logging
import os
import tempfile

from .engine import ReusableEngine
from .scoring import DEFAULT_CONFIG

log = logging.getLogger("pxdock")


class ProtenixDock:
    def __init__(self, receptor_pdb_file, cofactor_fpath=None, sf_file=None, nthreads=0,
                 receptor_preparer=None, work_dir=None):
        self._receptor_pdb = receptor_pdb_file
        self._cofactor_fpath = cofactor_fpath
        self._sf_file = sf_file or DEFAULT_CONFIG
        self._nthreads = nthreads
        self._prepare = receptor_preparer
        self._work_dir = work_dir or tempfile.mkdtemp(prefix="pxdock_")
        self._engine = None

    def set_box(self, pocket_center, box_size, buffer=0):
        self.pocket_center = [float(c) for c in pocket_center]
        self.box_size = [float(b) + buffer for b in box_size]
        log.info("pocket_center: %s", ",".join(str(round(a, 3)) for a in self.pocket_center))
        log.info("box_size: %s", ",".join(str(round(a, 3)) for a in self.box_size))
        if str(self._receptor_pdb).endswith(".json"):
            receptor_json = self._receptor_pdb
        elif self._prepare is not None:
            receptor_json = self._prepare(self._receptor_pdb, self.pocket_center, self.box_size,
                                          self._cofactor_fpath)
        else:
            raise RuntimeError("receptor preparation (PDB -> JSON) is not part of this package: "
                               "give a prepared receptor JSON or a receptor_preparer")
        if not receptor_json:
            raise RuntimeError(f"prepare_receptor failed {self._receptor_pdb}")
        self._engine = ReusableEngine(receptor_json, self._sf_file, self._nthreads)
        self._engine.set_box(*self.pocket_center, *self.box_size)

    def _require(self):
        if self._engine is None:
            raise RuntimeError("Please call `set_box(...)` first!")
        return self._engine

    def generate_cache_maps(self, spacing=0.175, cache_max_size=0.0, output_dir=None):
        e = self._require()
        e.generate_maps(spacing, cache_max_size)
        if output_dir is not None:
            e.dump_maps(output_dir)

    def restore_cache_maps(self, output_dir=None):
        self._require().load_maps(output_dir)

    def drop_cache_maps(self):
        self._require().drop_maps()

    def _index(self, ligands, tag):
        if isinstance(ligands, str):
            if ligands.endswith(".json"):
                ligands = [ligands]
            else:
                return ligands  # an index file already
        path = os.path.join(self._work_dir, f"{tag}_ligands.index")
        with open(path, "w") as fh:
            for p in ligands:
                fh.write(os.path.abspath(p) + "\n")
        return path

    def run_pose_opt(self, ligand_jsons, out_dir=None, penalty=1e6, vina_nposes=50):
        """optimize the prepared ligands' poses; vina_nposes is accepted for the
        reference's signature (its Vina pre-posing is external here)."""
        index = self._index(ligand_jsons, "pose_opt")
        if out_dir is None:
            out_dir = tempfile.mkdtemp(prefix="pose_opt_out_", dir=self._work_dir)
        self._require().optimize(index, out_dir, slope=penalty)
        return out_dir

    def run_docking(self, ligand_jsons, out_dir=None, penalty=1e6, seed=101, num_walker=256,
                    relax_nsteps=40, mc_prune_energy_threshold=-1):
        index = self._index(ligand_jsons, "mc")
        if out_dir is None:
            out_dir = tempfile.mkdtemp(prefix="docking_out_", dir=self._work_dir)
        self._require().search(index, out_dir, seed, num_walker, relax_nsteps=relax_nsteps,
                               mc_mmenergy_threshold=mc_prune_energy_threshold, slope=penalty)
        return out_dir
