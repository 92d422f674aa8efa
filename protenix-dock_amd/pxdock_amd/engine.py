"""ReusableEngine on the GPU: the reference's batch pipelines (receptor ->
maps -> evaluate / optimize / search over a ligand index file -> scores CSV
or per-ligand result JSON) over this library's kernels.

Names, arguments and defaults are those of the reference's Python binding
(engine/src/pyb/api.cpp:34-63) and its ReusableEngine
(engine/src/core/entrance.cpp:103-420).  The reference runs readers,
worker threads and writers over blocking queues; here the ligands of the
index file are read in chunks, every chunk is one batched GPU call per stage,
and results are written in input order (the reference's output order depends
on its thread timing).  cpu_nthreads is accepted and unused: the work runs
on the GPU.

Outputs keep the reference's formats:
  * evaluate: CSV "ligand_file,pose_id,pscore,bscore" (report_writer,
    writer.cpp:210-230), ligand_file the JSON file's stem, 16 significant
    digits;
  * optimize / search: <stem>_out.json per ligand (pose_writer::consume,
    writer.cpp:40-81) with mapped_smiles, best_pose {index, bscore} and the
    poses {offset, energy, pscore, nevals, receptor {torsions}, ligand {xyz}},
    receptor torsions in the JSON's rot_bond_order (restore_bond_orders,
    data.h:53-61).
"""
import json
import logging
import os

import numpy as np

from . import data, ffi, lib, scoring

log = logging.getLogger("pxdock")

CHUNK_LIGANDS = 512  # ligands per batched call


def read_index(path):
    """index_file_source::fill (source.cpp:27-44): one JSON path per line;
    empty lines and missing files are skipped with a warning."""
    out = []
    with open(path) as fh:
        for n, line in enumerate(fh, 1):
            line = line.rstrip("\n")
            if not line:
                log.warning("Skip illegal line#%d with empty text.", n)
            elif not os.path.isfile(line):
                log.warning("Skip illegal line#%d with ghost file: %s", n, line)
            else:
                out.append(line)
    return out


def file_stem(path):
    """boost::filesystem::path::stem (reader.cpp:28-31)."""
    return os.path.splitext(os.path.basename(path))[0]


def _parse_ligands(paths, geometry):
    """free_ligand::parse per file; unreadable ligands are skipped with a
    warning (reader.cpp:36-44,76-84)."""
    out = []
    for p in paths:
        try:
            out.append((p, data.parse_ligand(p, geometry=geometry)))
        except Exception as e:  # noqa: BLE001 - any parse failure skips the file
            log.warning("Due to invalid format or content, skip ligand: %s with reason: %s", p, e)
    return out


def _chunks(items, n):
    for i in range(0, len(items), n):
        yield items[i:i + n]


def _aligned_torsions(leaves, tors):
    """restore_bond_orders (data.h:53-61): torsion i goes to slot bond_order[i]."""
    if leaves is None or leaves.n_leaves == 0:
        return []
    order = leaves.bond_order if leaves.bond_order is not None else np.arange(leaves.n_leaves)
    out = np.zeros(int(np.max(order)) + 1)
    for i in range(leaves.n_leaves):
        out[int(order[i])] = float(tors[i])
    return out.tolist()


def write_result(output_dir, name, smiles, best_index, bscore, poses):
    """pose_writer::consume (writer.cpp:40-81): <name>_out.json."""
    doc = {}
    if smiles:
        doc["mapped_smiles"] = smiles
    doc["best_pose"] = {"index": int(best_index), "bscore": float(bscore)}
    doc["poses"] = [{"offset": int(p["offset"]), "energy": float(p["energy"]),
                     "pscore": float(p["pscore"]), "nevals": int(p["nevals"]),
                     "receptor": {"torsions": [float(t) for t in p["torsions"]]},
                     "ligand": {"xyz": np.asarray(p["xyz"], np.float64).tolist()}}
                    for p in poses]
    path = os.path.join(output_dir, name + "_out.json")
    with open(path, "w") as fh:
        json.dump(doc, fh, indent=2)
    return path


class ReusableEngine:
    """ReusableEngine(receptor_file, sf_file="", cpu_nthreads=0) (api.cpp:38-40)."""

    def __init__(self, receptor_file, sf_file="", cpu_nthreads=0):
        self.receptor_mol = data.parse_receptor(receptor_file)
        self.receptor = lib.Receptor(self.receptor_mol)
        sf = sf_file or None
        self.pose_params = scoring.load_pose_selection(sf)
        self.affinity_params = scoring.load_affinity_ranking(sf)
        self.cpu_nthreads = cpu_nthreads
        self.box = None
        self.grid = None
        self._rec_gb = None

    # ---------------------------------------------------------------- box / maps

    def set_box(self, center_x, center_y, center_z, size_x, size_y, size_z):
        self.box = [float(center_x), float(center_y), float(center_z),
                    float(size_x), float(size_y), float(size_z)]

    def _require_box(self):
        if self.box is None:
            raise RuntimeError("Please call `set_box(...)` first!")

    def _require_maps(self):
        if self.grid is None:
            raise RuntimeError(
                "Please call either `generate_cache_maps(...)` or `load_maps(...)` first!")

    def generate_maps(self, grid_spacing, max_size=0.0, types=None):
        """generate_maps (entrance.cpp:103-121); types (an extension) limits the
        LJ maps to those GAFF2 types, default all 99 as the reference."""
        self._require_box()
        size = [max(s, max_size) for s in self.box[3:]]
        bc = lib.decide_grid_dims(self.box[:3], size, grid_spacing)
        log.info("Generating receptor cache with grid settings:\n- Front-left-bottom point: "
                 "[%.3f, %.3f, %.3f]\n- Sample point count per dimension: [%d, %d, %d]",
                 *bc.init_xyz, *bc.npts)
        self.grid = lib.Grid(self.receptor, bc, range(ffi.NUM_VDW_TYPES) if types is None
                             else types)

    def dump_maps(self, cache_dir):
        self._require_maps()
        self.grid.save(cache_dir)

    def load_maps(self, cache_dir):
        self.grid = lib.Grid.load(self.receptor, cache_dir)
        bc = self.grid.box
        log.info("Previous grid settings:\n- Front-left-bottom point: [%.3f, %.3f, %.3f]\n"
                 "- Spacing: %g\n- Sample point count per dimension: [%d, %d, %d]",
                 *bc.init_xyz, bc.spacing, *bc.npts)

    def drop_maps(self):
        self.grid = None

    # ---------------------------------------------------------------- helpers

    def _gb_params(self, ligs):
        if self._rec_gb is None:
            self._rec_gb = lib.gbsa_atom_params(self.receptor_mol)
        parts = [lib.gbsa_atom_params(m) for m in ligs]
        return self._rec_gb, (np.concatenate([p[0] for p in parts]),
                              np.concatenate([p[1] for p in parts]))

    def _check_types(self, ligs):
        if self.grid is None:
            return
        have = set(int(t) for t in self.grid.types)
        missing = sorted(set(int(t) for m in ligs for t in m.vdw_type) - have)
        if missing:
            tags = lib.type_tags()
            raise RuntimeError("the loaded maps lack atom types: " +
                               ", ".join(tags[t] for t in missing))

    def _bscores(self, batch, pose_lig, xyz_list, tors):
        nl = self.receptor.n_leaves
        P = lib.PoseSet(batch, pose_lig, xyz_list, nl, tors=tors)
        rec_gb, lig_gb = self._gb_params(batch.ligands)
        s, _ = lib.bscore(self.receptor, P, self.affinity_params, with_terms=False,
                          rec_gb=rec_gb, lig_gb=lig_gb)
        return s.cpu().numpy()

    # ---------------------------------------------------------------- pipelines

    def evaluate(self, ligand_index_file, output_dir):
        """evaluate (entrance.cpp:148-197): pscore and bscore of every pose of
        every ligand at the receptor's input coordinates; output_dir is the
        CSV file written (bdscr's output_file)."""
        paths = read_index(ligand_index_file)
        nl = self.receptor.n_leaves
        n_rows = 0
        with open(output_dir, "w") as out:
            out.write("ligand_file,pose_id,pscore,bscore\n")
            for part in _chunks(_parse_ligands(paths, geometry=False), CHUNK_LIGANDS):
                ligs = [m for _, m in part]
                batch = lib.LigandBatch(ligs)
                pose_lig, xyz = [], []
                for l, m in enumerate(ligs):
                    for x in m.xyz:
                        pose_lig.append(l)
                        xyz.append(x)
                P = lib.PoseSet(batch, pose_lig, xyz, nl)
                ps, _ = lib.pscore(self.receptor, P, self.pose_params, with_terms=False)
                ps = ps.cpu().numpy()
                bs = self._bscores(batch, pose_lig, xyz, None)
                k = 0
                for (path, m) in part:
                    for pid in range(len(m.xyz)):
                        out.write(f"{file_stem(path)},{pid},{ps[k]:.16g},{bs[k]:.16g}\n")
                        k += 1
                n_rows += k
        log.info("Number of successful evaluations is: %d", n_rows)
        return n_rows

    def optimize(self, ligand_index_file, output_dir, max_niters=240, slope=1e6):
        """optimize (entrance.cpp:199-327): L-BFGS of every pose (lbfgs_step
        with max_niters, on the maps when present), then pose_ranker
        (writer.cpp:179-208): pscore of each optimised pose, bscore of the best."""
        paths = read_index(ligand_index_file)
        if self.grid is None:
            log.info("Receptor cache is disabled in this run.")
        else:
            self.grid.set_slope(slope)
        os.makedirs(output_dir, exist_ok=True)
        nl = self.receptor.n_leaves
        prm = ffi.lbfgs_params(max_niters)
        files = []
        for part in _chunks(_parse_ligands(paths, geometry=False), CHUNK_LIGANDS):
            ligs = [m for _, m in part]
            self._check_types(ligs)
            batch = lib.LigandBatch(ligs)
            pose_lig, xyz = [], []
            for l, m in enumerate(ligs):
                for x in m.xyz:
                    pose_lig.append(l)
                    xyz.append(x)
            P = lib.PoseSet(batch, pose_lig, xyz, nl)
            e, ne, cv = lib.lbfgs(self.receptor, self.grid, P, prm)
            ps, _ = lib.pscore(self.receptor, P, self.pose_params, with_terms=False)
            e, ne, cv, ps = (t.cpu().numpy() for t in (e, ne, cv, ps))
            tors = P.tors.cpu().numpy()
            pose_xyz = [P.xyz_of(p) for p in range(P.n_poses)]
            starts = np.concatenate([[0], np.cumsum([len(m.xyz) for m in ligs])])
            best = [int(starts[l] + np.argmin(ps[starts[l]:starts[l + 1]]))
                    for l in range(len(ligs))]
            bs = self._bscores(batch, list(range(len(ligs))), [pose_xyz[b] for b in best],
                               tors[best])
            for l, (path, m) in enumerate(part):
                name = file_stem(path)
                poses = []
                for p in range(starts[l], starts[l + 1]):
                    if not cv[p]:
                        log.warning("Minimization of [%s/pose-%d] failed to converge!", name,
                                    p - starts[l])
                    poses.append(dict(offset=p - starts[l], energy=e[p], pscore=ps[p],
                                      nevals=ne[p], torsions=_aligned_torsions(
                                          self.receptor_mol.leaves, tors[p][:nl]),
                                      xyz=pose_xyz[p]))
                files.append(write_result(output_dir, name, m.smiles, best[l] - starts[l],
                                          bs[l], poses))
        log.info("Number of successful optimizations is: %d", len(files))
        return files

    def search(self, ligand_index_file, output_dir, seed=0, exhaustiveness=256, max_nsteps=200,
               relax_nsteps=40, num_modes=8, min_rmsd=0.5, mc_mmenergy_threshold=-1.0,
               slope=1e6):
        """search (entrance.cpp:329-420): exhaustiveness Monte-Carlo walkers per
        ligand (monte_carlo_searcher, each walker starting from the ligand's
        conformer of its index when the JSON holds one), then pose_cluster
        (writer.cpp:124-177): the num_modes best distinct poses by pscore and
        the bscore of the best."""
        paths = read_index(ligand_index_file)
        if self.grid is None:
            self._require_box()
            box = lib.describe_boundary(self.box[:3], self.box[3:])
            log.info("Receptor cache is disabled in this run.")
        else:
            self.grid.set_slope(slope)
            box = self.grid.box
        os.makedirs(output_dir, exist_ok=True)
        prm = lib.mc_default_params()
        prm.seed = int(seed)
        prm.max_nsteps = int(max_nsteps)
        prm.relax = ffi.lbfgs_params(int(relax_nsteps))
        prm.mmenergy_threshold = float(mc_mmenergy_threshold)
        nl = self.receptor.n_leaves
        W = int(exhaustiveness)
        files = []
        for part in _chunks(_parse_ligands(paths, geometry=True), CHUNK_LIGANDS):
            ligs = [m for _, m in part]
            self._check_types(ligs)
            batch = lib.LigandBatch(ligs)
            res = lib.mc_search(self.receptor, self.grid, batch, None, self.pose_params, prm, box,
                                W, num_modes=num_modes, min_rmsd=min_rmsd,
                                conformers=[list(m.xyz) for m in ligs])
            ci, cc = res.cluster_idx.cpu().numpy(), res.cluster_count.cpu().numpy()
            be, bp, ne = (t.cpu().numpy() for t in (res.best_energy, res.best_pscore, res.nevals))
            bt = res.best_tors.cpu().numpy()
            kept = [(l, [int(w) for w in ci[l, :cc[l]]]) for l in range(len(ligs)) if cc[l] > 0]
            if not kept:
                continue
            bs = self._bscores(batch, [l for l, _ in kept],
                               [res.walker_xyz(l * W + ws[0]) for l, ws in kept],
                               np.stack([bt[l * W + ws[0]] for l, ws in kept]))
            for k, (l, ws) in enumerate(kept):
                path, m = part[l]
                poses = [dict(offset=w, energy=be[l * W + w], pscore=bp[l * W + w],
                              nevals=ne[l * W + w],
                              torsions=_aligned_torsions(self.receptor_mol.leaves,
                                                         bt[l * W + w][:nl]),
                              xyz=res.walker_xyz(l * W + w)) for w in ws]
                files.append(write_result(output_dir, file_stem(path), m.smiles, 0, bs[k], poses))
            for l in range(len(ligs)):
                if cc[l] == 0:
                    log.debug("Tasks with prefix [%s] has failed.", file_stem(part[l][0]))
        log.info("Number of successful dockings is: %d", len(files))
        return files
