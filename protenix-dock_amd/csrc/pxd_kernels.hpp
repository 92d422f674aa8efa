// pxd_kernels.hpp — launcher interface between the C-ABI and the kernels.
#pragma once

#include <hip/hip_runtime.h>

#include "pxd_internal.hpp"

namespace pxd {

constexpr int kMaxHistory = 128;  // L-BFGS history cap (LDS for rho/alpha)
constexpr int kLbfgsVpl = 7;      // vector elements per lane: 3*128 + 64 <= 64*7

struct GridBuildArgs {
  const float4* atoms;  // fixed receptor atoms x, y, z, q
  const float2* alj;    // sigma, sqrt(eps)
  const double4* atoms64;  // the same in FP64 (terms of near atoms)
  const double2* alj64;
  int n_atoms;
  const float2* tlj;    // per requested type: sigma, 4 * sqrt(eps)
  const double2* tlj64;
  int n_types;
  float init[3];
  double init64[3];
  int np[3];
  float spacing;
  double spacing64;
  long long P;
  float4* maps;
};

struct ScoreArgs {
  int n_poses;
  const int* pose_lig;
  const int* pose_off;
  const float* xyz;
  const float* tors;
  double* energy;
  float* lig_grad;
  float* tors_grad;
  int max_atoms;  // largest ligand of the batch (sizes the per-pose LDS); <= 0: the maximum
  int poses_per_block = 1;  // poses (waves) per workgroup: 1, 2 or 4
};

struct LbfgsParamsDev {
  int history, max_iter, max_eval, max_ls;
  double lr, tol_grad, tol_change, c1, c2, ls_tol_change, gtd_threshold;
  int dof_mode;  // 0: [ligand xyz | receptor torsions]; 1: rigid-body + torsions
};

struct LbfgsArgs {
  LbfgsParamsDev prm;
  int n_poses;
  const int* pose_lig;
  const int* pose_off;
  float* xyz;
  float* tors;
  double* energy;
  int* nevals;
  int* converged;
  int* counter;   // work-queue heads (one per size class), zero before launch
  float* hist;    // [n_slots][lbfgs_slot_floats(history)]
  int max_atoms;  // largest ligand of the batch (sizes the per-pose LDS); <= 0: the maximum
  // nullable, zero before launch: totals over the poses of iterations,
  // evaluations, history bytes and map-lookup bytes (pxd_lbfgs_stats)
  unsigned long long* stats;
  // nullable: the poses of one size class (indices into the pose arrays),
  // n_order of them; the launch then runs exactly these
  const int* order;
  int n_order;
};

// Per L-BFGS work slot: the y and s histories, [history][64*kLbfgsVpl] each,
// then rho = 1 / (s.y) per history entry (FP64).
__host__ __device__ inline size_t lbfgs_slot_floats(int history) {
  return (size_t)2 * history * kWave * kLbfgsVpl + (size_t)2 * history;
}

// L-BFGS size classes: a launch per class takes the poses whose variable
// count n satisfies 64 * (lower class width) < n <= 64 * width, so its LDS
// vector pool (and its occupancy) follows the poses it runs.
constexpr int kLbfgsClasses = 3;
constexpr int kLbfgsClassVpl[kLbfgsClasses] = {2, 4, kLbfgsVpl};

// Occupancy target of the Monte-Carlo walker kernel (waves per SIMD); the
// search launches this many walker slots per SIMD.
// Occupancy target of the batched L-BFGS kernel (waves per SIMD) and the
// depth of its two-loop history prefetch per size class (vector width 2 / 4 / 7).
#ifndef PXD_LBFGS_WAVES
#define PXD_LBFGS_WAVES 3
#endif
#ifndef PXD_LBFGS_PRE2
#define PXD_LBFGS_PRE2 8
#endif
#ifndef PXD_LBFGS_PRE4
#define PXD_LBFGS_PRE4 6
#endif
#ifndef PXD_LBFGS_PRE7
#define PXD_LBFGS_PRE7 2
#endif
constexpr int kLbfgsWaves = PXD_LBFGS_WAVES;

#ifndef PXD_MC_WAVES
#define PXD_MC_WAVES 2
#endif
constexpr int kMcWaves = PXD_MC_WAVES;

struct PscoreArgs {
  int n_poses;
  const int* pose_lig;
  const int* pose_off;
  const float* xyz;
  const float* tors;
  double* score;
  double* terms;
  int max_atoms;  // largest ligand of the batch (sizes the per-pose LDS); <= 0: the maximum
};

struct McParamsDev {
  uint32_t seed_lo, seed_hi;
  int max_nsteps;
  int mutation;        // 0: reference mutate_inplace, 1: Gaussian (translation + rotation + torsion)
  float amplitude, temperature, trans_sigma, rot_sigma;
  int max_tries;
  LbfgsParamsDev relax, full;
  double mm_threshold;  // > 0: the energy prune of searcher.cpp:141-157
};

// Per-walker scalars kept on the device between launches.
struct WalkerState {
  int mode;       // 0: needs init, 1: stepping, 2: done
  int step;
  uint32_t rng_ctr;
  int nevals;
  double cur_e, cur_ps, best_e, best_ps;
};

struct McArgs {
  McParamsDev prm;
  const pxd_score_params* sp;  // device copy
  float box_lo[3], box_hi[3];
  int n_walkers;
  const int* walker_lig;
  const int* walker_pid;   // walker index within its ligand (0: no Monte-Carlo)
  const int* walker_off;   // atom offset of the walker's conformations
  const float* lig_xyz0;   // conformers: ligand l's k-th at atom lig_conf_off[l] + k * natoms
  const int* lig_conf_off; // [n_ligands]
  const int* lig_conf_n;   // [n_ligands] conformers of each ligand (>= 1)
  unsigned long long* lig_emin;  // [n_ligands] lowest pre-optimisation energy (ordered key)
  float* cur_xyz;
  float* best_xyz;
  float* cur_tors;
  float* best_tors;
  WalkerState* st;
  int* counter;    // queue heads, one per L-BFGS size class
  float* hist;
  int hist_m;      // history slots per wave in hist
  int step_end;
  int max_atoms;   // largest ligand of the batch (sizes the dynamic LDS)
  int max_frags;   // most fragments of a ligand of the batch
  // nullable: the walkers of one size class (indices), n_order of them; the
  // launch then runs exactly these
  const int* order;
  int n_order;
};

struct ClusterArgs {
  int n_ligands;
  const int* lig_walker_off;  // [n_ligands+1] walkers of ligand l
  const int* walker_off;
  const int* walker_lig;
  const float* best_xyz;
  const WalkerState* st;
  int num_modes;
  float min_rmsd;
  int* out_idx;    // [n_ligands * num_modes] walker indices (-1 padded)
  int* out_count;  // [n_ligands]
};

// types_per_pass: 4, 8 or 16 type maps per pass over the receptor atoms
void launch_grid_build(const GridBuildArgs& a, int types_per_pass, hipStream_t st);
void launch_score(const DevReceptor& R, const DevLigands& L, const DevGrid* G,
                  const ScoreArgs& a, hipStream_t st);
// One size class (0, 1, 2: vector width kLbfgsClassVpl[cls]) over the poses
// a.order lists; the launches of different classes may run concurrently.
void launch_lbfgs_class(const DevReceptor& R, const DevLigands& L, const DevGrid* G,
                        const LbfgsArgs& a, int cls, int n_slots, hipStream_t st);
void launch_lbfgs(const DevReceptor& R, const DevLigands& L, const DevGrid* G,
                  const LbfgsArgs& a, int n_slots, hipStream_t st);
void launch_pscore(const DevReceptor& R, const DevLigands& L, const pxd_score_params& sp,
                   const PscoreArgs& a, hipStream_t st);
void launch_pscore_grad(const DevReceptor& R, const DevLigands& L, const pxd_score_params& sp,
                        const PscoreArgs& a, float* dxyz, double* dparams, hipStream_t st);
// One size class (0, 1, 2) over the walkers a.order lists.
void launch_mc_class(const DevReceptor& R, const DevLigands& L, const DevGrid* G,
                     const McArgs& a, int cls, int n_slots, hipStream_t st);
void launch_mc(const DevReceptor& R, const DevLigands& L, const DevGrid* G, const McArgs& a,
               int n_slots, hipStream_t st);
void launch_fill_int(int* p, int n, int v, hipStream_t st);
void launch_fill_u64(unsigned long long* p, int n, unsigned long long v, hipStream_t st);

// A total order on doubles as unsigned 64-bit keys (a < b iff key(a) < key(b)
// for non-NaN values), so a minimum over walkers is one atomicMin.
__host__ __device__ inline unsigned long long order_key(double v) {
  unsigned long long b;
  __builtin_memcpy(&b, &v, 8);
  return (b >> 63) ? ~b : (b | 0x8000000000000000ull);
}
__host__ __device__ inline double order_key_value(unsigned long long k) {
  unsigned long long b = (k >> 63) ? (k & 0x7fffffffffffffffull) : ~k;
  double v;
  __builtin_memcpy(&v, &b, 8);
  return v;
}
void launch_cluster(const ClusterArgs& a, int n_atoms_max, hipStream_t st);
struct ApplyArgs {
  int n_poses;
  const int* pose_lig;
  const int* pose_off;
  const float* xyz_in;
  const float* center;   // [n][3]
  const float* rot;      // [n][9]
  const int* tors_off;   // [n+1] into tors
  const float* tors;
  float* xyz_out;
};
void launch_apply(const DevLigands& L, const ApplyArgs& a, hipStream_t st);
// Receptor leaf energy tables: out[k*n + i] = (T_k, dT_k/dtheta) at
// theta = 2*pi*(i + shift)/n.
void launch_leaf_table(const DevReceptor& R, int n, double shift, double2* out,
                       hipStream_t st);
// Map layout 1: every cell's 8 corners of every map stored together.
void launch_expand_cells(const DevGrid& G, int n_maps, float4* cm, hipStream_t st);
void launch_split_planes(const DevGrid& G, int n_maps, float* sp, hipStream_t st);
// Rigid-body + torsion gradient of poses from their Cartesian gradient:
// out[p][6 + kMaxFrags - 1] = force, torque about the centre, torsion torques.
void launch_dof_grad(const DevLigands& L, int n_poses, const int* pose_lig, const int* pose_off,
                     const float* xyz, const float* grad, float* out, hipStream_t st);

}  // namespace pxd

namespace pxd {
// Shared by pxd_lbfgs_batch and pxd_lbfgs_batch_noalloc (pxd_api.cpp).
pxd_status lbfgs_prepare(const pxd_receptor* r, const pxd_grid* g, const pxd_ligand_batch* b,
                         const pxd_lbfgs_params* prm, int32_t n_poses, const int32_t* pose_lig,
                         const int32_t* pose_atom_off, float* xyz, float* tors, double* energy,
                         int32_t* nevals, int32_t* converged, LbfgsArgs* a, int* n_slots);
const DevGrid* grid_device(const pxd_grid* g);
}  // namespace pxd
