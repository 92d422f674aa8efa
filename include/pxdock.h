/*
 * pxdock.h — C ABI of the MI355X (gfx950) docking hot path.
 *
 * Everything crossing this boundary is a plain C type: integers, doubles,
 * pointers and sizes.  No exceptions cross it: every entry point returns a
 * pxd_status and pxd_last_error() gives the text of the last failure on the
 * calling thread.  Objects are opaque handles with an explicit destroy.
 *
 * The same descriptor structs (pxd_mol, pxd_frag_tree, pxd_leaves, pxd_box,
 * pxd_score_params, pxd_lbfgs_params) are consumed by the FP64 CPU oracle that
 * the tests use as the checker, so the GPU and the oracle read identical
 * inputs.
 *
 * Reference interfaces replaced (paths relative to the Protenix-Dock tree):
 *   receptor_cache::populate / get         engine/include/bytedock/core/grid.h:66-74
 *   binding_system_interactions::put_gradients
 *                                          engine/include/bytedock/core/interaction.h:105-108
 *   optimization_step::apply (lbfgs_step)  engine/include/bytedock/core/optimizer.h:26-67
 *   monte_carlo_searcher::fill             engine/include/bytedock/core/searcher.h:43
 *   root_scorer::combine (pose_selection)  engine/include/bytedock/core/scorer.h:96-117
 *   pose_cluster::preprocess               engine/include/bytedock/core/writer.h:53-67
 */
#ifndef PXDOCK_H_
#define PXDOCK_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PXD_NUM_VDW_TYPES 99 /* engine/include/bytedock/core/constant.h:27 */
#define PXD_VDW_ID_OFFSET 1000 /* engine/include/bytedock/core/constant.h:28 */
#define PXD_DIH_TERMS 6 /* engine/include/bytedock/core/system.h:48 */
#define PXD_MAX_LIGAND_ATOMS 128 /* per-pose LDS budget of the batched kernels */
#define PXD_MAX_FRAGMENTS 32 /* root + up to 31 rotatable bonds per ligand */
#define PXD_MAX_LEAVES 64 /* receptor rotatable leaves */

typedef enum pxd_status {
  PXD_OK = 0,
  PXD_ERR_ARG = 1,   /* invalid argument, shape or size */
  PXD_ERR_CONF = 2,  /* malformed input data (failed_conf_error in the reference) */
  PXD_ERR_HIP = 3,   /* HIP runtime error */
  PXD_ERR_STATE = 4, /* call order misuse (e.g. no maps built) */
  PXD_ERR_RANGE = 5, /* a size exceeds a documented maximum */
  PXD_ERR_IO = 6     /* a file is missing, unreadable or truncated (file_system_error) */
} pxd_status;

/* Force-field and interaction-set data of one molecule.
 * Mirrors force_field_params (engine/include/bytedock/core/system.h:85-109).
 * All index arrays are 0-based atom indices of this molecule.
 * Ligand: vdw_type set (GAFF2 index 0..98, the JSON id minus 1000), sigma/epsilon NULL.
 * Receptor: sigma/epsilon set, vdw_type NULL. */
typedef struct pxd_mol {
  int32_t n_atoms;
  const int32_t* vdw_type;
  const double* sigma;
  const double* epsilon;
  const double* charge;
  /* harmonic bonds: ids[n][2], k, length (interaction.cpp:27-47) */
  int32_t n_bonds;
  const int32_t* bond_ij;
  const double* bond_k;
  const double* bond_r0;
  /* harmonic angles: ids[n][3], k, angle in radians (interaction.cpp:49-74) */
  int32_t n_angles;
  const int32_t* angle_ijk;
  const double* angle_k;
  const double* angle_t0;
  /* Fourier dihedrals: impropers first, then propers (interaction.cpp:104-117).
   * ids[n][4]; periodicity[n][6] (a 0 ends the series), phase[n][6] radians, k[n][6] */
  int32_t n_dihedrals;
  const int32_t* dih_ijkl;
  const int32_t* dih_n;
  const double* dih_phase;
  const double* dih_k;
  /* non-bonded pairs: 1-4 (Coulomb x0.833333333, LJ x0.5) and all others (x1) */
  int32_t n_pairs14;
  const int32_t* pairs14;
  int32_t n_pairs;
  const int32_t* pairs;
  /* pscore interaction sets (data.cpp:260-272) */
  int32_t n_hydrophobic;
  const int32_t* hydrophobic;
  int32_t n_cation;
  const int32_t* cation;
  int32_t n_anion;
  const int32_t* anion;
  int32_t n_ring5;
  const int32_t* ring5; /* [n][5] */
  int32_t n_ring6;
  const int32_t* ring6; /* [n][6] */
  int32_t n_don_charged;
  const int32_t* don_charged; /* [n][2] = (heavy donor, hydrogen) */
  int32_t n_don_neutral;
  const int32_t* don_neutral;
  int32_t n_acc_charged;
  const int32_t* acc_charged;
  int32_t n_acc_neutral;
  const int32_t* acc_neutral;
  /* torsion-strain penalties (ligand only, data.cpp:204-223):
   * coefs[n][4]; matches of penalty p are ijkl[match_off[p] .. match_off[p+1]) */
  int32_t n_tstrain;
  const double* tstrain_coef;
  const int32_t* tstrain_match_off;
  const int32_t* tstrain_ijkl;
  int32_t n_rotatable;
} pxd_mol;

/* Rigid-fragment tree of a ligand (data.cpp:410-482), in a flat form.
 * Atom positions are permuted so every fragment is a contiguous range.
 * Torsion t (the order of the torsion vector in free_ligand::apply_parameters,
 * data.h:78-81) rotates fragment tors_frag[t]; tors_frag lists the non-root
 * fragments in the reference's pre-order traversal. */
typedef struct pxd_frag_tree {
  int32_t n_frags;
  const int32_t* perm;        /* [n_atoms]: permuted slot p holds original atom perm[p] */
  const int32_t* frag_start;  /* [n_frags] range [start, end) in permuted slots */
  const int32_t* frag_end;
  const int32_t* frag_parent; /* [n_frags] parent fragment, -1 for a root */
  const int32_t* frag_bond;   /* [n_frags][2] (parent atom, mine atom) as permuted slots */
  int32_t n_torsions;         /* = n_frags - number of roots */
  const int32_t* tors_frag;   /* [n_torsions] */
} pxd_frag_tree;

/* Rotatable receptor leaves (data.cpp:348-363): leaf k rotates its movable
 * atoms about origin[k] along the unit vector axis[k]. */
typedef struct pxd_leaves {
  int32_t n_leaves;
  const double* origin;     /* [n][3] */
  const double* axis;       /* [n][3] unit vectors */
  const int32_t* atom_off;  /* [n+1] */
  const int32_t* atoms;     /* movable receptor atoms */
} pxd_leaves;

/* Grid geometry: the fields of box_config (grid.h:31-40) in its order, with
 * the point counts as int32 (the reference's index_t is size_t, dtype.h:29-30;
 * the structs are not layout-compatible: convert field by field). */
typedef struct pxd_box {
  double init_xyz[3];
  int32_t npoints[3];
  double spacing;
  double oppo_xyz[3];
  int32_t npoints_minus_one[3];
  double spacing_inv;
} pxd_box;

/* pose_selection parameters (conf.h:137-170).  A term is active when
 * |coef| > 1e-5 (utility.h:24-26).  Inter-parameter sextets are
 * (dist_min, dist_max, angle_min, angle_max, dist_center|sigma_min,
 *  angle_center|sigma_max) in the order named per field below. */
typedef struct pxd_score_params {
  double coul_coef, vdw_coef, ionic_coef, pication_coef, f2f_coef, e2f_coef;
  double hbond_coef, hydrophobic_coef, tstrain_coef;
  /* interaction_vdw_energy */
  double vdw_scale, vdw_cutoff;
  double vdw_sigma_scale[PXD_NUM_VDW_TYPES];
  double vdw_epsilon_scale[PXD_NUM_VDW_TYPES];
  /* ionic_energy: coefs (lc, pc); ranges [2] = (dist_min, dist_max) */
  double ionic_lc_coef, ionic_pc_coef;
  double ionic_lc[2], ionic_pc[2];
  /* pi_cationic_energy: order lc5, lc6, pc5, pc6 */
  double pication_sub_coef[4];
  double pication_range[4][2];
  /* f2f / e2f pi stacking: order l5r5, l5r6, l6r5, l6r6;
   * sextet = (dist_min, dist_max, angle_min, angle_max, sigma_min, sigma_max) */
  double f2f_sub_coef[4];
  double f2f_inter[4][6];
  double e2f_sub_coef[4];
  double e2f_inter[4][6];
  /* hbond: scale; order hldcc, hldcn, hldnc, hldnn, hpdcc, hpdcn, hpdnc, hpdnn;
   * sextet = (dist_min, dist_max, angle_min, angle_max, dist_center, angle_center) */
  double hbond_scale;
  double hbond_sub_coef[8];
  double hbond_inter[8][6];
  /* hydrophobic_energy */
  double hydro_scale, hydro_dist_min, hydro_dist_max, hydro_scale_factor;
} pxd_score_params;

/* Number of pscore leaf terms reported per pose, in this order. */
enum {
  PXD_TERM_COULOMB = 0,
  PXD_TERM_VDW,
  PXD_TERM_IONIC,
  PXD_TERM_PICATION,
  PXD_TERM_F2F,
  PXD_TERM_E2F,
  PXD_TERM_HBOND,
  PXD_TERM_HYDROPHOBIC,
  PXD_TERM_TSTRAIN,
  PXD_NUM_TERMS
};

/* L-BFGS with strong-Wolfe line search (optimizer.h:50-67, optimizer.cpp:235-560). */
typedef struct pxd_lbfgs_params {
  int32_t history;     /* reference: 100 */
  int32_t max_iter;    /* reference: 240 (full) / relax_nsteps (relax) */
  int32_t max_eval;    /* reference: max_iter * 5 / 4 */
  double lr;           /* 1 */
  double tol_grad;     /* infinity norm, 0.5 */
  double tol_change;   /* 1e-4 */
  double c1, c2;       /* 1e-4, 0.9 */
  double ls_tol_change; /* 1e-9 */
  double gtd_threshold; /* 1e5 */
  int32_t max_ls;      /* 25 */
  int32_t dof_mode;    /* 0: the reference's variables [ligand xyz | receptor torsions]
                          (optimizer.cpp:56-76); 1: [translation | rotation vector |
                          ligand torsions | receptor torsions] about the input pose */
} pxd_lbfgs_params;

/* Fill the reference defaults (optimizer.h:52-66) for the given iteration cap. */
void pxd_lbfgs_default_params(pxd_lbfgs_params* p, int32_t max_iter);

/* ------------------------------------------------------------------------ */
/* Library lifecycle and errors                                              */

/* Text of the last error raised on the calling thread ("" if none). */
const char* pxd_last_error(void);
/* Library version string. */
const char* pxd_version(void);
/* Number of visible HIP devices (0 when none), without creating a context. */
pxd_status pxd_device_count(int32_t* out);

/* Grid geometry helpers (grid.cpp:42-86). */
void pxd_describe_boundary(const double center[3], const double size[3], pxd_box* out);
void pxd_decide_grid_dims(const double center[3], const double size[3], double spacing,
                          pxd_box* out);

/* ------------------------------------------------------------------------ */
/* Receptor (torsional_receptor, data.h:34-71)                               */

typedef struct pxd_receptor pxd_receptor;

/* Copies the receptor's coordinates, force field, interaction sets and leaves
 * to device memory.  xyz is [n_atoms][3].
 * Leaf energy tables: the receptor terms whose movable atoms all belong to
 * one leaf depend on that leaf's angle only; their sum is tabulated per leaf
 * in FP64 (value and derivative, cubic Hermite lookup) when the table
 * reproduces direct FP64 evaluation at every interval midpoint within 1e-8
 * (value) and 1e-6 (derivative), relative and floored at 1; otherwise these
 * terms stay evaluated pair by pair (binding_system_interactions::
 * put_gradients, interaction.cpp:271-298, gives the same energy either way). */
pxd_status pxd_receptor_create(const pxd_mol* mol, const double* xyz,
                               const pxd_leaves* leaves, pxd_receptor** out);
void pxd_receptor_destroy(pxd_receptor* r);

/* Switches the leaf tables off (0) or back on (1) for later launches; the
 * switch is not synchronised with launches already queued on the receptor.
 * PXD_ERR_STATE when asked to enable tables that were not built.
 * No counterpart in the reference: it selects between two evaluations of
 * interaction.cpp:316-350 that agree within the bounds above. */
pxd_status pxd_receptor_leaf_tables(pxd_receptor* r, int32_t enable);

/* Leaves, table nodes per leaf (0: direct evaluation) and the worst midpoint
 * errors measured when the tables were built. */
pxd_status pxd_receptor_info(const pxd_receptor* r, int32_t* n_leaves, int32_t* table_nodes,
                             double* table_err_e, double* table_err_d);

/* ------------------------------------------------------------------------ */
/* Grid maps (receptor_cache, grid.h:60-133)                                 */

typedef struct pxd_grid pxd_grid;

/* Builds the Coulomb map and one LJ map per requested GAFF2 type on the GPU,
 * each point holding (value, dE/dx, dE/dy, dE/dz) in FP32 (grid.cpp:115-195).
 * Movable receptor atoms are excluded (grid.cpp:118-123).  slope is the
 * out-of-box wall (grid.cpp:246-251; set_slope grid.h:79). */
pxd_status pxd_grid_build(const pxd_receptor* r, const pxd_box* box, const int32_t* types,
                          int32_t n_types, double slope, void* stream, pxd_grid** out);
/* Wraps maps computed elsewhere (e.g. loaded from disk).  maps is host FP32
 * [(n_types+1)][npoints][4], map 0 = Coulomb. */
pxd_status pxd_grid_upload(const pxd_receptor* r, const pxd_box* box, const int32_t* types,
                           int32_t n_types, double slope, const float* maps, pxd_grid** out);
/* Copies the maps back to host FP32 in the upload layout. */
pxd_status pxd_grid_download(const pxd_grid* g, float* maps, size_t n_floats);
pxd_status pxd_grid_info(const pxd_grid* g, pxd_box* box, int32_t* n_types);
/* Memory layout of the maps read by the lookups (receptor_cache::get,
 * grid.cpp:197-303, gives identical results in all three):
 *   0  point-major, each point one (value, dE/dx, dE/dy, dE/dz) record, as
 *      built and as uploaded/downloaded (default);
 *   1  every cell's 8 corner points stored together, so one lookup reads one
 *      128-byte line per map; costs 8x the map memory, kept beside layout 0;
 *   2  point-major component planes (value, dE/dx, dE/dy, dE/dz each a
 *      separate array); costs 1x the map memory, kept beside layout 0.
 * Waits for the stream; the choice applies to later launches. */
pxd_status pxd_grid_set_layout(pxd_grid* g, int32_t layout, void* stream);
pxd_status pxd_grid_set_slope(pxd_grid* g, double slope);
void pxd_grid_destroy(pxd_grid* g);

/* ------------------------------------------------------------------------ */
/* Map directories (receptor_cache::save / restore, grid.cpp:305-431)        */
/*
 * A map directory holds one file per map and a metadata file written last,
 * whose presence marks the directory complete (the role of meta.dat in
 * grid.cpp:336-343).  The reference's Boost archives of FP64 maps are not
 * read or written; the format is this library's own:
 *
 *   <tag>.pxm  one per GAFF2 type, the tag of pxdock_vdw_table.h ("c3",
 *              "n+", ...); coul.pxm for the unit-probe Coulomb map (the tag
 *              list of constant.h kVdwTypeTags, "coul" last).  A 64-byte
 *              little-endian header
 *                0  char[8]  "PXDMAP\0\1"
 *                8  uint32   format version (PXD_MAPS_VERSION)
 *               12  int32    GAFF2 type index, -1 for Coulomb
 *               16  int32[3] npoints (x, y, z)
 *               28  int32    values per point (4)
 *               32  uint64   payload bytes = npoints product * 16
 *               40  uint64   payload checksum (FNV-1a 64 over little-endian
 *                            64-bit words: h = 0xcbf29ce484222325, per word
 *                            h ^= w, h *= 0x100000001b3)
 *               48  zero padding
 *              then FP32 (value, dE/dx, dE/dy, dE/dz) per point, point
 *              (i, j, k) at (i * ny + j) * nz + k (grid.h:120-122).
 *   meta.pxd   text, one record per line:
 *                pxdock-maps <version>
 *                init_xyz <x> <y> <z>        (%.17g, exact for FP64)
 *                npoints <nx> <ny> <nz>
 *                spacing <h>
 *                slope <s>
 *                receptor <n_atoms> <key, 16 hex digits>
 *                excluded <n> <atom> ...     (movable receptor atoms, ascending)
 *                types <n> <type> ...        (ascending GAFF2 indices)
 *                map <type|coul> <file> <checksum, 16 hex digits>   (one per map)
 *                end
 *
 * Saving removes an old meta.pxd first and renames meta.pxd.tmp into place
 * only after every map file is complete.  Loading fails with PXD_ERR_IO when
 * meta.pxd or a map file is missing, unreadable, truncated or fails its
 * checksum, and with PXD_ERR_CONF on a version mismatch (grid.cpp:356-363),
 * malformed metadata, or maps saved for another receptor or another set of
 * movable atoms. */
#define PXD_MAPS_VERSION 1

/* GAFF2 tag of type t (0..98, kVdwTypeTags, constant.h:131-231), "coul" for
 * t = PXD_NUM_VDW_TYPES, NULL otherwise. */
const char* pxd_vdw_type_tag(int32_t t);

typedef struct pxd_maps_meta {
  pxd_box box;
  double slope;
  int32_t n_atoms;        /* receptor atoms, movable ones included */
  int32_t n_excluded;     /* movable receptor atoms, left out of the maps */
  uint64_t receptor_key;  /* pxd_receptor_key of the receptor the maps belong to */
  int32_t n_types;
  int32_t types[PXD_NUM_VDW_TYPES]; /* ascending GAFF2 indices */
} pxd_maps_meta;

/* Fingerprint of what the maps depend on: FNV-1a 64 (as above, over bytes)
 * of n_atoms and, for every receptor atom outside the leaves in index order,
 * its index, x, y, z, charge, sigma and epsilon (int32 / FP64 little-endian).
 * Host only; no device is touched. */
pxd_status pxd_receptor_key(const pxd_mol* mol, const double* xyz, const pxd_leaves* leaves,
                            uint64_t* key);
/* Host-only directory I/O.  maps is FP32 [(n_types+1)][npoints][4] in the
 * layout of pxd_grid_upload: map 0 = Coulomb, map t+1 = meta->types[t].
 * excluded holds meta->n_excluded ascending atom indices. */
pxd_status pxd_maps_write(const char* dir, const pxd_maps_meta* meta, const int32_t* excluded,
                          const float* maps);
/* Reads and checks meta.pxd (not the map payloads).  excluded (capacity cap)
 * may be NULL. */
pxd_status pxd_maps_info(const char* dir, pxd_maps_meta* meta, int32_t* excluded, int32_t cap);
/* Reads the Coulomb map and the maps of types[0..n_types) (NULL: every saved
 * type, in meta order) into maps, checking each payload's checksum.  With
 * expect non-NULL the saved receptor (n_atoms, receptor_key, n_excluded and
 * the excluded atoms) must equal expect's and excluded's. */
pxd_status pxd_maps_read(const char* dir, const pxd_maps_meta* expect, const int32_t* excluded,
                         const int32_t* types, int32_t n_types, float* maps, size_t n_floats);
/* receptor_cache::save (grid.cpp:309-344): every map of g and its metadata,
 * for the receptor g was built on.  Creates dir when it does not exist. */
pxd_status pxd_grid_save(const pxd_grid* g, const char* dir);
/* receptor_cache::restore (grid.cpp:346-391) into a device grid of r: the
 * Coulomb map and the maps of types[0..n_types) (NULL: every saved type).
 * The maps must have been saved for r (same key and movable atoms).  slope < 0
 * takes the saved slope. */
pxd_status pxd_grid_load(const pxd_receptor* r, const char* dir, const int32_t* types,
                         int32_t n_types, double slope, pxd_grid** out);

/* ------------------------------------------------------------------------ */
/* Ligand batch (free_ligand, data.h:73-99) packed for uneven sizes          */

typedef struct pxd_ligand_batch pxd_ligand_batch;

/* Packs n ligands (term lists, types, charges, interaction sets and fragment
 * trees; trees may be NULL) into device CSR arrays.  Every ligand must use a
 * type present in the grid it is scored against.  A tree is rejected with
 * PXD_ERR_CONF before any device work unless: perm is a permutation of the
 * atoms; the fragment ranges partition the permuted slots; every parent is a
 * fragment of the tree and every bond joins an atom of the parent to an atom
 * of the child; n_torsions equals n_frags minus the number of roots; and
 * tors_frag lists every non-root fragment exactly once, each after its
 * parent (free_ligand's pre-order, data.h:78-81). */
pxd_status pxd_ligand_batch_create(const pxd_mol* ligands, const pxd_frag_tree* trees,
                                   int32_t n, pxd_ligand_batch** out);
void pxd_ligand_batch_destroy(pxd_ligand_batch* b);
/* atom count of ligand i */
pxd_status pxd_ligand_batch_natoms(const pxd_ligand_batch* b, int32_t i, int32_t* out);

/* ------------------------------------------------------------------------ */
/* Poses. A pose set is CSR: pose p is ligand pose_lig[p], its coordinates    */
/* are xyz[3*pose_atom_off[p] ...] (FP32, device memory), its receptor       */
/* torsions tors[p*n_leaves ...].  All pose arrays below are DEVICE pointers. */

/* Batched energy + analytic gradient (binding_system_interactions::put_gradients,
 * interaction.cpp:271-298) for n_poses poses, one wavefront per pose.
 * grid may be NULL: then the inter term is the brute-force sum over all
 * receptor atoms (interaction.cpp:241-254).  Outputs are OVERWRITTEN:
 * energy[p] (FP64), lig_grad (FP32, same CSR as xyz), tors_grad[p*n_leaves+k] (FP32). */
pxd_status pxd_score_grad_batch(const pxd_receptor* r, const pxd_grid* grid,
                                const pxd_ligand_batch* b, int32_t n_poses,
                                const int32_t* pose_lig, const int32_t* pose_atom_off,
                                const float* xyz, const float* tors, double* energy,
                                float* lig_grad, float* tors_grad, void* stream);

/* Batched persistent L-BFGS (lbfgs_step::apply, optimizer.cpp:429-560):
 * xyz and tors are optimised IN PLACE.  Outputs per pose: energy (final
 * evaluation), nevals, converged (1/0).  Many uneven instances run in one
 * launch from a device work queue.  The library keeps the history storage of
 * its work slots between calls (grown on demand, one per process). */
pxd_status pxd_lbfgs_batch(const pxd_receptor* r, const pxd_grid* grid,
                           const pxd_ligand_batch* b, const pxd_lbfgs_params* prm,
                           int32_t n_poses, const int32_t* pose_lig,
                           const int32_t* pose_atom_off, float* xyz, float* tors,
                           double* energy, int32_t* nevals, int32_t* converged,
                           void* stream);

/* Work totals of the most recent pxd_lbfgs_batch call in this process (all
 * zero before the first).  No counterpart in the reference; they give the
 * algorithmic bytes of a run for its roofline.  A history entry is the n
 * floats of s and of y plus its FP64 1 / (s.y), n the pose's variable count;
 * the two-loop recursion reads every stored entry twice per iteration. */
typedef struct pxd_lbfgs_stats {
  uint64_t iterations;    /* L-BFGS iterations, all poses */
  uint64_t evaluations;   /* score+grad evaluations, all poses */
  uint64_t history_bytes; /* history entries read and written x (8 n + 8) bytes */
  uint64_t map_bytes;     /* map lookups: 256 bytes per ligand atom and evaluation */
} pxd_lbfgs_stats;
pxd_status pxd_lbfgs_last_stats(pxd_lbfgs_stats* out);

/* pxd_lbfgs_batch for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * No counterpart in the reference.  Same arguments, results and checks as
 * pxd_lbfgs_batch, but nothing is allocated or freed and the call does not
 * wait for the stream: the work-queue heads and L-BFGS histories live in a
 * fixed scratch array of PXD_NOALLOC_SCRATCH_MB MiB per device, held until the
 * process ends.  The work slots are as many as the tuning asks for and the
 * scratch holds (history 8: 2,335; history 100: 186), so long
 * histories run fewer poses at a time; results are identical either way.
 * Every call on a device shares that device's scratch: calls whose launches
 * overlap on one device (two streams, or two host threads) race, so order
 * them on one stream. */
#define PXD_NOALLOC_SCRATCH_MB 64
pxd_status pxd_lbfgs_batch_noalloc(const pxd_receptor* r, const pxd_grid* grid,
                                   const pxd_ligand_batch* b, const pxd_lbfgs_params* prm,
                                   int32_t n_poses, const int32_t* pose_lig,
                                   const int32_t* pose_atom_off, float* xyz, float* tors,
                                   double* energy, int32_t* nevals, int32_t* converged,
                                   void* stream);

/* The plain GPU form of pxd_lbfgs_batch, kept as the comparison baseline:
 * the optimiser runs on the host (FP64) and every energy evaluation is one
 * launch of the same score+grad kernel for one pose, followed by a copy back
 * and a stream synchronisation.  Same arguments and results as
 * pxd_lbfgs_batch (lbfgs_step::apply, optimizer.cpp:429-560). */
pxd_status pxd_lbfgs_plain(const pxd_receptor* r, const pxd_grid* grid,
                           const pxd_ligand_batch* b, const pxd_lbfgs_params* prm,
                           int32_t n_poses, const int32_t* pose_lig,
                           const int32_t* pose_atom_off, float* xyz, float* tors,
                           double* energy, int32_t* nevals, int32_t* converged,
                           void* stream);

/* Monte-Carlo search (monte_carlo_searcher, searcher.h:27-82). */
typedef struct pxd_mc_params {
  uint64_t seed;            /* walker k of ligand l draws Philox4x32-10 blocks with key
                               (seed_lo, seed_hi + k) and counter (draw, l, 0, 0) */
  int32_t max_nsteps;       /* Monte-Carlo steps per walker (bdock default 200) */
  int32_t mutation;         /* 0: reference mutate_inplace (searcher.cpp:189-223);
                               1: Gaussian translation + rotation and one torsion per step */
  double amplitude;         /* 2 A (searcher.h:33) */
  double temperature;       /* 1.2 (searcher.h:33) */
  double trans_sigma;       /* mutation 1 only */
  double rot_sigma;         /* mutation 1 only */
  int32_t max_tries;        /* randomize attempts (searcher.cpp:73-77): 2000 */
  int32_t steps_per_launch; /* steps per kernel launch of the captured step loop */
  pxd_lbfgs_params relax;   /* after each mutation: relax_nsteps (40) iterations */
  pxd_lbfgs_params full;    /* initial and refine: 240 iterations */
  double mmenergy_threshold; /* > 0: a mutated pose whose energy exceeds the lowest
                                pre-optimisation energy of its ligand's walkers by more
                                is skipped (searcher.cpp:141-157); <= 0 (default -1,
                                api.cpp:61): off */
} pxd_mc_params;

/* Reference defaults (bdock.cpp:60-78, searcher.h:33). */
void pxd_mc_default_params(pxd_mc_params* p);

/* Docking search for every ligand of the batch with `walkers` walkers each
 * (exhaustiveness); walker 0 of a ligand optimises the input pose and does no
 * Monte-Carlo (searcher.cpp:105-123).  All buffers are DEVICE memory:
 *   lig_xyz0   [sum_l natoms_l][3] input conformation of each ligand (batch order)
 *   box        search box (check_atom_in_box, grid.h:42-46): the grid box when
 *              maps are used, describe_boundary otherwise
 *   best_xyz   [sum_l walkers*natoms_l][3] best pose per walker (ligand-major)
 *   best_tors  [n_walkers][n_leaves], best_energy / best_pscore [n_walkers] (FP64),
 *   nevals     [n_walkers] (0 marks a failed walker, searcher.cpp:178-181)
 *   cluster_idx [n_ligands][num_modes] walker index within the ligand (-1 pad),
 *   cluster_count [n_ligands]: pose_cluster::preprocess (writer.cpp:124-177).
 * The step loop is captured once in a hipGraph and replayed. */
pxd_status pxd_mc_search(const pxd_receptor* r, const pxd_grid* grid,
                         const pxd_ligand_batch* b, const pxd_score_params* sp,
                         const pxd_mc_params* prm, const pxd_box* box, int32_t walkers,
                         const float* lig_xyz0, float* best_xyz, float* best_tors,
                         double* best_energy, double* best_pscore, int32_t* nevals,
                         int32_t num_modes, double min_rmsd, int32_t* cluster_idx,
                         int32_t* cluster_count, void* stream);

/* pxd_mc_search with pre-generated conformers (searcher.cpp:107-113):
 * ligand l brings conf_count[l] >= 1 conformations (HOST array), stored in
 * the DEVICE array conformers ligand-major, [sum_l conf_count[l]*natoms_l][3].
 * Walker 0 optimises conformer 0; walker k >= 1 is placed at random in the
 * box from its own conformer k when k < conf_count[l] (keeping its
 * torsions), else from conformer 0 with random torsions.  Everything else as
 * pxd_mc_search, which is this call with one conformer per ligand. */
pxd_status pxd_mc_search_conformers(const pxd_receptor* r, const pxd_grid* grid,
                                    const pxd_ligand_batch* b, const pxd_score_params* sp,
                                    const pxd_mc_params* prm, const pxd_box* box,
                                    int32_t walkers, const int32_t* conf_count,
                                    const float* conformers, float* best_xyz, float* best_tors,
                                    double* best_energy, double* best_pscore, int32_t* nevals,
                                    int32_t num_modes, double min_rmsd, int32_t* cluster_idx,
                                    int32_t* cluster_count, void* stream);

/* free_ligand::apply_parameters (data.cpp:592-641) for a batch of poses:
 * pose p of ligand pose_lig[p] is rebuilt from xyz_in (its input conformation)
 * with centre center[p][3], orientation matrix rot[p][9] (row major) and the
 * torsion rotations tors[tors_off[p] .. tors_off[p+1]) in the tree's
 * pre-order.  DEVICE buffers; xyz_out uses the xyz_in layout. */
pxd_status pxd_apply_poses(const pxd_ligand_batch* b, int32_t n_poses, const int32_t* pose_lig,
                           const int32_t* pose_atom_off, const float* xyz_in,
                           const float* center, const float* rot, const int32_t* tors_off,
                           const float* tors, float* xyz_out, void* stream);

/* Gradient of the energy with respect to rigid-body + torsion variables,
 * from the Cartesian ligand gradient of pxd_score_grad_batch (chain rule
 * through free_ligand::apply_parameters, data.cpp:592-641):
 * dof_grad[p][PXD_DOF_STRIDE] = force (3), torque about the pose's geometric
 * centre (3), one torsion torque per ligand torsion in tree pre-order
 * (zero-padded).  Receptor torsion gradients are pxd_score_grad_batch's.
 * The reference optimises Cartesian coordinates and has no such gradient;
 * this is the gradient of pxd_lbfgs_params.dof_mode 1 at zero variables. */
#define PXD_DOF_STRIDE (6 + PXD_MAX_FRAGMENTS - 1)
pxd_status pxd_dof_grad_batch(const pxd_ligand_batch* b, int32_t n_poses, const int32_t* pose_lig,
                              const int32_t* pose_atom_off, const float* xyz,
                              const float* lig_grad, float* dof_grad, void* stream);

/* Scoring-function parameters differentiated by pxd_pscore_grad_batch, in
 * this order: the nine term coefficients (PXD_TERM_* order), then
 * vdw_sigma_scale[99], then vdw_epsilon_scale[99]. */
#define PXD_NUM_SCORE_PARAMS (PXD_NUM_TERMS + 2 * PXD_NUM_VDW_TYPES)

/* pscore of each pose (as pxd_pscore_batch) with its analytic derivatives:
 * dscore_dxyz (FP32, the CSR layout of xyz) with respect to the ligand
 * coordinates, and dscore_dparams[n_poses][PXD_NUM_SCORE_PARAMS] (FP64) with
 * respect to the parameters above.  Derivatives of the piecewise leaves are
 * those of the piece the pose is on (the ramps, the vdW clamp and the
 * stacking cut-offs are not differentiable at their breakpoints).  A
 * coefficient's derivative is its unweighted leaf value when the term is
 * active (|coef| > 1e-5) and 0 otherwise.  The reference scores but never
 * differentiates the score (scorer.cpp:117-724); this is the backward pass of
 * the scoring function for fitting it on the GPU.  dscore_dxyz and
 * dscore_dparams may be NULL. */
pxd_status pxd_pscore_grad_batch(const pxd_receptor* r, const pxd_ligand_batch* b,
                                 const pxd_score_params* sp, int32_t n_poses,
                                 const int32_t* pose_lig, const int32_t* pose_atom_off,
                                 const float* xyz, const float* tors, double* score,
                                 float* dscore_dxyz, double* dscore_dparams, void* stream);

/* Batched pscore (root_scorer::combine on the pose_selection tree,
 * scorer.cpp:117-205 and leaves :207-613) for n_poses poses.  Receptor
 * coordinates come from applying tors to the receptor's leaves
 * (data.cpp:366-385); tors may be NULL (all zero).  score[p] (FP64) is the
 * weighted sum; terms (may be NULL) receives [n_poses][PXD_NUM_TERMS] unweighted
 * leaf values (FP64). */
pxd_status pxd_pscore_batch(const pxd_receptor* r, const pxd_ligand_batch* b,
                            const pxd_score_params* sp, int32_t n_poses,
                            const int32_t* pose_lig, const int32_t* pose_atom_off,
                            const float* xyz, const float* tors, double* score,
                            double* terms, void* stream);

/* ------------------------------------------------------------------------ */
/* Affinity score (bscore: root_scorer on the affinity_ranking tree,         */
/* scorer.cpp:53-115, with gbsa_total_energy, scorer.cpp:735-866)            */

/* leaves: the pscore leaves with the affinity tree's parameters, the
 * torsion-strain weight being what the reference's coefficient map holds
 * (the rotatable-energy coefficient is stored under the torsion-strain key,
 * scorer.cpp:102-109: when both are active it replaces the torsion-strain
 * weight, and the rotatable-energy leaf itself is never weighted).
 * gbsa_coef: gbsa_energy_delta_coef; the term is E_com - E_lig - E_pro. */
typedef struct pxd_bscore_params {
  pxd_score_params leaves;
  double gbsa_coef;
} pxd_bscore_params;

#define PXD_NUM_BSCORE_TERMS (PXD_NUM_TERMS + 3) /* leaves, then GB E_pro, E_lig, E_com */

/* fill_atom_properties (scorer.cpp:735-758), host only: Born radius and
 * screening factor of each atom from its atomic number, an H bonded (bonds
 * [n_bonds][2], the JSON bond_index) to an N taking 1.3 A. */
pxd_status pxd_gbsa_atom_params(int32_t n_atoms, const int32_t* atomic_numbers,
                                int32_t n_bonds, const int32_t* bonds, double* radius,
                                double* screen);

/* Affinity score of each pose: the leaves (as pxd_pscore_batch with
 * p->leaves) plus gbsa_coef * (E_com - E_lig - E_pro), the generalized-Born
 * + surface-area energies (gbsa_total_energy::calculate, scorer.cpp:798-866)
 * of the receptor (with the pose's receptor torsions), the ligand, and the
 * complex (ligand atoms first, then the receptor's), in FP64 from FP64
 * parameters and the pose's coordinates.  rec_radius / rec_screen: host
 * [n_rec_atoms]; lig_radius / lig_screen: host, every atom of the batch in
 * batch order (pxd_gbsa_atom_params per ligand, concatenated).  score:
 * device [n_poses] (FP64); terms (may be NULL): device
 * [n_poses][PXD_NUM_BSCORE_TERMS].  Allocates per call and returns when the
 * results are in the caller's buffers. */
pxd_status pxd_bscore_batch(const pxd_receptor* r, const double* rec_radius,
                            const double* rec_screen, const pxd_ligand_batch* b,
                            const double* lig_radius, const double* lig_screen,
                            const pxd_bscore_params* p, int32_t n_poses,
                            const int32_t* pose_lig, const int32_t* pose_atom_off,
                            const float* xyz, const float* tors, double* score,
                            double* terms, void* stream);

/* ------------------------------------------------------------------------ */
/* Autotuning of launch parameters (no counterpart in the reference, whose    */
/* CPU path has no launch parameters).                                       */

typedef struct pxd_tuning {
  int32_t grid_types_per_pass; /* type maps per pass of the grid build: 4, 8, 16 */
  int32_t map_layout;          /* lookup layout applied to new grids (pxd_grid_set_layout) */
  int32_t lbfgs_slots_per_cu;  /* resident L-BFGS work slots per compute unit */
  int32_t score_poses_per_block; /* poses (one wave each) per workgroup of the batched
                                    score+grad: 1, 2, 4 */
  double grid_ms[3];           /* measured build times for 4, 8, 16 types per pass */
  double score_ms[3];          /* measured score+grad times for layouts 0, 1, 2 */
  double lbfgs_ms[4];          /* measured L-BFGS times for 2, 4, 8, 16 slots per CU */
  double score_ppb_ms[3];      /* measured score+grad times for 1, 2, 4 poses per block */
} pxd_tuning;

/* The process-wide choice in force (defaults: 8, 0, 8, 1; timings zero). */
void pxd_get_tuning(pxd_tuning* out);
/* Sets the four choices (timings ignored). */
pxd_status pxd_set_tuning(const pxd_tuning* t);
/* Times the variants on the caller's receptor, box, types, ligands and poses
 * (device arrays, left unchanged) and keeps the fastest of each: grid build
 * types per pass, map layout for score+grad, poses per workgroup of
 * score+grad (on the chosen layout), L-BFGS slots per CU (40 iterations
 * from the given poses).  Results of every variant are identical; only
 * their speed differs. */
pxd_status pxd_autotune(const pxd_receptor* r, const pxd_box* box, const int32_t* types,
                        int32_t n_types, const pxd_ligand_batch* b, int32_t n_poses,
                        const int32_t* pose_lig, const int32_t* pose_atom_off, const float* xyz,
                        const float* tors, void* stream, pxd_tuning* out);

#ifdef __cplusplus
}
#endif

#endif /* PXDOCK_H_ */
