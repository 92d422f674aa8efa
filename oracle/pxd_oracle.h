/*
 * pxd_oracle.h — FP64 CPU restatement of the Protenix-Dock docking hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP path:
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load it.  The product (libpxdock.so) never links or calls it.
 *
 * Every function cites the reference file:line it restates (paths relative
 * to the Protenix-Dock tree).  Inputs are the descriptor structs of
 * include/pxdock.h, so the oracle and the GPU read the same data.
 */
#ifndef PXD_ORACLE_H_
#define PXD_ORACLE_H_

#include <stdint.h>

#include "pxdock.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ---- constants (system.h:83, math.h:128, constant.h:29-130) ---- */
extern const double or_coulomb_factor;
extern const double or_gaff2_sigma[PXD_NUM_VDW_TYPES];
extern const double or_gaff2_epsilon[PXD_NUM_VDW_TYPES];

/* ---- lib/math (math.h, math.cpp) ---- */
void or_matmul3(const double* a, const double* b, double* out);
void or_rotmat_quat(const double q[4], double out[9]);
void or_rotmat_axis(const double axis[3], double theta, double out[9]);
void or_rotmat_axis_grad(const double axis[3], double theta, double out[9], double jac[9]);
double or_angle_grad(const double* pi, const double* pj, const double* pk, double* gi,
                     double* gk);
double or_dihedral(const double* pi, const double* pj, const double* pk, const double* pl);
double or_dihedral_grad(const double* pi, const double* pj, const double* pk,
                        const double* pl, double* gi, double* gj, double* gk, double* gl);
double or_rev_linear(double v, double minv, double maxv);
double or_rev_linear_c(double v, double center, double minv, double maxv);

/* ---- pair terms (pair.h:25-60) ---- */
double or_coul_pair(double qi, double qj, const double* rij, double d, double scale,
                    double* gi, double* gj);
double or_vdw_pair(double si, double ei, double sj, double ej, const double* rij, double d,
                   double scale, double* gi, double* gj);

/* ---- intra-molecular terms (interaction.cpp:27-196) ---- */
double or_bond_energy(const pxd_mol* m, const double* xyz, double* grad);
double or_angle_energy(const pxd_mol* m, const double* xyz, double* grad);
double or_dihedral_energy(const pxd_mol* m, const double* xyz, double* grad);
double or_nonbonded_energy(const pxd_mol* m, const double* xyz, double* grad);
double or_intra_energy(const pxd_mol* m, const double* xyz, double* grad);

/* ---- grid maps (grid.cpp:115-303) ---- */
typedef struct or_grid {
  pxd_box box;
  int32_t n_types;
  const int32_t* types;   /* GAFF2 index of map t+1 */
  const double* maps;     /* [(n_types+1)][P][4], map 0 = Coulomb */
  double slope;
  int32_t n_excluded;     /* movable receptor atoms, computed brute force */
  const int32_t* excluded;
} or_grid;

/* Values of the Coulomb map and the requested type maps at one point:
 * out[(n_types+1)*4] (grid.cpp:164-194 for a single point). */
void or_grid_point(const pxd_mol* rec, const double* rec_xyz, const uint8_t* excluded,
                   const double* p, const int32_t* types, int32_t n_types, double* out);
/* Full build, threaded over x-slabs (grid.cpp:115-148). */
void or_grid_build(const pxd_mol* rec, const double* rec_xyz, const uint8_t* excluded,
                   const pxd_box* box, const int32_t* types, int32_t n_types, double* maps,
                   int32_t nthreads);
/* receptor_cache::get (grid.cpp:197-262): adds into energy/grad, returns hit. */
int32_t or_grid_get(const or_grid* g, int32_t type, double q, const double* xyz,
                    double* energy, double* grad);

/* ---- inter + total energy (interaction.cpp:202-350) ---- */
double or_inter_energy(const pxd_mol* rec, const double* rec_xyz, const or_grid* grid,
                       const pxd_mol* lig, const double* lig_xyz, double* rec_grad,
                       double* lig_grad);

typedef struct or_system {
  const pxd_mol* rec;
  const double* rec_xyz;     /* initial receptor coordinates */
  const pxd_leaves* leaves;
  const pxd_mol* lig;
  const or_grid* grid;       /* NULL: brute-force inter */
  const pxd_frag_tree* tree; /* ligand fragment tree (rigid+torsion L-BFGS only) */
} or_system;

/* binding_system_interactions::put_gradients (interaction.cpp:271-298).
 * Gradients are OVERWRITTEN (the caller-side reset of interaction.cpp:309-314
 * is folded in).  tors/tors_grad have leaves->n_leaves entries. */
double or_system_energy(const or_system* s, const double* tors, const double* lig_xyz,
                        double* tors_grad, double* lig_grad);

/* ---- L-BFGS (optimizer.cpp:205-560) ---- */
/* Test-only: multiply every energy / gradient component the L-BFGS closure
 * sees by (1 + rel * u), u uniform in [-1, 1) from a seeded stream, to model
 * FP32 evaluation error.  rel = 0 (the default) disables it.  Per thread. */
void or_set_eval_noise(double rel, uint64_t seed);
/* Optimises (lig_xyz, tors) in place; returns converged; energy/nevals out.
 * p->dof_mode 0: the reference's variables, [ligand xyz | receptor torsions]
 * (optimizer.cpp:56-76).  1: rigid-body + torsion variables
 * [translation | rotation vector | ligand torsions | receptor torsions] about
 * the input pose, mapped to coordinates by free_ligand::apply_parameters
 * (data.cpp:592-641) with s->tree; the gradient is the chain rule of the
 * Cartesian one (or_dof_grad, with the SO(3) left Jacobian for the rotation
 * vector). */
int32_t or_lbfgs(const or_system* s, const pxd_lbfgs_params* p, double* lig_xyz,
                 double* tors, double* energy, int32_t* nevals);

/* Rigid-body + torsion gradient of a pose from its Cartesian gradient g:
 * out[0..2] force sum(g), out[3..5] torque sum((x - center) x g),
 * out[6 + k] torsion k (tree pre-order): axis_k . sum over the atoms the
 * torsion moves of (x - x_mine) x g, axis_k the unit bond vector from the
 * parent-side atom to the moving-side atom. */
void or_dof_grad(const pxd_frag_tree* t, int32_t n_atoms, const double* xyz, const double* g,
                 const double center[3], double* out);
/* The energy of or_lbfgs's rigid-body + torsion variables x about xyz0 and its
 * gradient; pose (may be NULL) receives the coordinates. */
double or_dof_energy(const or_system* s, const double* xyz0, const double* x, double* grad,
                     double* pose);

/* Threaded batch forms for the CPU baseline: systems[] one per ligand. */
void or_score_many(const or_system* systems, int32_t n_poses, const int32_t* pose_lig,
                   const int32_t* pose_atom_off, const double* xyz, const double* tors,
                   int32_t n_leaves, double* energy, double* lig_grad, double* tors_grad,
                   int32_t nthreads);
void or_lbfgs_many(const or_system* systems, const pxd_lbfgs_params* p, int32_t n_poses,
                   const int32_t* pose_lig, const int32_t* pose_atom_off, double* xyz,
                   double* tors, int32_t n_leaves, double* energy, int32_t* nevals,
                   int32_t* converged, int32_t nthreads);

/* ---- conformations (data.cpp:366-385, 522-641) ---- */
void or_receptor_apply(const pxd_leaves* L, int32_t n_atoms, const double* xyz,
                       const double* tors, double* out);
void or_ligand_apply(const pxd_frag_tree* t, int32_t n_atoms, const double* xyz,
                     const double center[3], const double rot[9], const double* tors,
                     double* out);

/* ---- pscore (scorer.cpp:117-613) ---- */
double or_pscore(const pxd_score_params* sp, const pxd_mol* rec, const double* rec_xyz,
                 const pxd_mol* lig, const double* lig_xyz, double* terms);
/* the vdW leaf alone, as used by randomize (searcher.cpp:75-77) */
double or_pscore_vdw(const pxd_score_params* sp, const pxd_mol* rec, const double* rec_xyz,
                     const pxd_mol* lig, const double* lig_xyz);

/* ---- affinity score: GBSA (scorer.cpp:735-866) and bscore ---- */
/* fill_atom_properties (scorer.cpp:735-758) */
void or_gbsa_atom_params(int32_t n, const int32_t* atomic_numbers, int32_t n_bonds,
                         const int32_t* bonds, double* radius, double* screen);
/* gbsa_total_energy::calculate (scorer.cpp:798-866) of one system */
double or_gbsa_energy(int32_t n, const double* xyz, const double* charge, const double* radius,
                      const double* screen);
/* root_scorer on the affinity tree: the leaves of or_pscore with bp->leaves
 * plus gbsa_coef * (E_com - E_lig - E_pro) (gbsa_total_energy::summary,
 * scorer.cpp:760-796); terms[PXD_NUM_BSCORE_TERMS] = leaves, E_pro, E_lig,
 * E_com (the GB entries 0 when the term is inactive). */
double or_bscore(const pxd_bscore_params* bp, const pxd_mol* rec, const double* rec_xyz,
                 const double* rec_radius, const double* rec_screen, const pxd_mol* lig,
                 const double* lig_xyz, const double* lig_radius, const double* lig_screen,
                 double* terms);

/* ---- counter-based RNG shared with the GPU (Philox4x32-10) ---- */
void or_philox4x32(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);

/* ---- top-k clustering (writer.cpp:99-177) ---- */
int32_t or_pose_cluster(int32_t n_cand, int32_t n_atoms, const double* xyz,
                        const double* pscore, const int32_t* nevals, int32_t num_modes,
                        double min_rmsd, int32_t* out_index);

/* ---- Monte-Carlo walker (searcher.cpp:46-223) ----
 * Walker `walker` of ligand `ligand` draws Philox4x32-10 blocks with key
 * (seed_lo, seed_hi + walker) and counter (draw, ligand, 0, 0); walker 0
 * optimises xyz0 and takes no Monte-Carlo step.  xyz0 is the conformer the
 * walker starts from: its own pre-generated one (own_conformer 1: randomize
 * draws no torsions) or the ligand's conformer 0 (searcher.cpp:107-113).
 * emin_io: the ligand's lowest pre-optimisation energy, shared by its
 * walkers (start at 1e8), used when prm->mmenergy_threshold > 0; may be
 * NULL.  best_tors has n_leaves entries; n_pruned (may be NULL) receives the
 * steps the prune skipped.  Returns the number of Monte-Carlo steps taken. */
int32_t or_mc_walker(const or_system* s, const pxd_frag_tree* t, const pxd_score_params* sp,
                     const pxd_mc_params* prm, const pxd_box* box, int32_t walker,
                     int32_t ligand, const double* xyz0, int32_t own_conformer,
                     double* emin_io, double* best_xyz, double* best_tors, double* best_e,
                     double* best_ps, int32_t* nevals, int32_t* n_pruned);

/* One mutation of the walker (mutate_inplace, searcher.cpp:189-223, or the
 * Gaussian variant of pxd_mc_params.mutation 1) applied to xyz in place,
 * drawing from the walker's stream at counter *ctr (advanced).  Returns the
 * mutation drawn: 0 translation, 1 rotation, k + 2 a change of torsion k. */
int32_t or_mc_mutate(const pxd_frag_tree* t, const pxd_mc_params* prm, int32_t walker,
                     int32_t ligand, uint32_t* ctr, int32_t n, double* xyz);
/* count draws of the walker stream's distributions (random.cpp:22-49):
 * kind 0 uniform [0, 1); 1 a point in the unit ball (3 doubles each);
 * 2 a unit quaternion (4 each); 3 a standard normal; k >= 4 a uniform
 * integer in [0, k - 4] inclusive. */
void or_mc_draws(int32_t kind, uint64_t seed, int32_t walker, int32_t ligand, int32_t count,
                 double* out);

#ifdef __cplusplus
}
#endif

#endif /* PXD_ORACLE_H_ */
