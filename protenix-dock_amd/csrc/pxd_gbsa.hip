// pxd_gbsa.hip — the affinity score (bscore): the pscore leaves with the
// affinity tree's parameters plus the generalized-Born / surface-area delta
// (gbsa_total_energy, engine/src/core/scorer.cpp:735-866), in FP64.
//
// Per pose three systems are scored: the receptor (its atoms with the pose's
// receptor torsions), the ligand, and the complex (ligand atoms first, then
// the receptor's).  All three are contiguous atom ranges of one per-pose
// "complex" array, so each pass is one launch over (atom block, system,
// pose):
//   k_gb_fixed  once per call: for every fixed receptor atom, its
//               descreening sum over the other fixed atoms (no pose moves
//               them);
//   k_gb_setup  complex coordinates and per-atom parameters of every pose;
//   k_gb_born   the descreening sums I_j (LDS-tiled passes: the sources that
//               move with the pose, then the fixed ones or their
//               precomputed sum) and the OBC Born radii;
//   k_gb_egb    the generalized-Born row sums (each unordered pair once,
//               counted twice) plus each atom's surface term;
//   k_gb_sum    the rows of each system summed in a fixed order.
// Every sum has a fixed order, so the results are the same bits on every run.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <vector>

#include "pxd_internal.hpp"
#include "pxd_kernels.hpp"
#include "pxdock.h"
#include "pxdock_gbsa_table.h"

using namespace pxd;

extern "C" pxd_status pxd_set_error_(pxd_status st, const char* msg);

namespace {

constexpr int kGbBlock = 256;
constexpr double kGbCoulomb = 332.0637141491396;  // kCoulombFactor (system.h:83)
constexpr double kGbEps = 1e-20;                  // kParamEpsilon (math.h:128)

struct GbArgs {
  int n_chunk;   // poses in this launch
  int pose0;     // first pose of the launch
  int n_rec;
  int n_fix;     // receptor atoms [0, n_rec - n_fix) are movable, the rest fixed
  const double* sfix;  // [n_fix] descreening sum of each fixed atom over the other fixed atoms
  int stride;    // atoms per pose slot: largest ligand + receptor
  const int* pose_lig;
  const int* pose_off;
  const float* xyz;
  const float* tors;
  int n_leaves;
  const LigInfo* info;
  const double4* lig_p;    // charge, radius, screen, 0 per batch atom
  const double4* rec_x;    // x, y, z, charge per receptor atom (input coordinates)
  const double2* rec_p;    // radius, screen
  const int* mov_slot;     // receptor atom -> leaf, -1 when fixed
  const double4* leaf_o;   // origin of each leaf
  const double4* leaf_a;   // unit axis of each leaf
  double4* cx;             // [n_chunk][stride] x, y, z, charge
  double4* cp;             // [n_chunk][stride] radius - offset, (radius - offset) * screen, radius
  double* born;            // [n_chunk][3][stride]
  double* row;             // [n_chunk][3][stride]
  double* energy;          // [n_poses][3] E_pro, E_lig, E_com
};

// atom range [a0, a1) of system s of a pose with n ligand atoms:
// 0 receptor, 1 ligand, 2 complex
__device__ inline void gb_range(int s, int n, int n_rec, int* a0, int* a1) {
  *a0 = s == 0 ? n : 0;
  *a1 = s == 1 ? n : n + n_rec;
}

__global__ __launch_bounds__(kGbBlock) void k_gb_setup(GbArgs A) {
  const int p = blockIdx.y;
  const int k = blockIdx.x * kGbBlock + threadIdx.x;
  const int pg = A.pose0 + p;
  const LigInfo li = A.info[A.pose_lig[pg]];
  const int n = li.n_atoms;
  if (k >= n + A.n_rec) return;
  double4 x, q;
  if (k < n) {
    const float* c = A.xyz + 3ll * (A.pose_off[pg] + k);
    const double4 lp = A.lig_p[li.atom_off + k];
    x = make_double4((double)c[0], (double)c[1], (double)c[2], lp.x);
    q = make_double4(lp.y, lp.z, 0.0, 0.0);
  } else {
    const int a = k - n;
    x = A.rec_x[a];
    const int leaf = A.mov_slot[a];
    if (leaf >= 0) {
      // torsional_receptor::apply_parameters (data.cpp:366-385): rotation by
      // the pose's torsion about the leaf axis through its origin
      // (get_rotation_matrix, math.cpp:62-94)
      const double t = A.tors ? (double)A.tors[(long long)pg * A.n_leaves + leaf] : 0.0;
      const double4 o = A.leaf_o[leaf], ax = A.leaf_a[leaf];
      const double s = sin(t), c = cos(t), omc = 1.0 - c;
      const double vx = x.x - o.x, vy = x.y - o.y, vz = x.z - o.z;
      const double X = ax.x, Y = ax.y, Z = ax.z;
      const double rx = (c + X * X * omc) * vx + (X * Y * omc - Z * s) * vy + (Y * s + X * Z * omc) * vz;
      const double ry = (Z * s + X * Y * omc) * vx + (c + Y * Y * omc) * vy + (Y * Z * omc - X * s) * vz;
      const double rz = (X * Z * omc - Y * s) * vx + (X * s + Y * Z * omc) * vy + (c + Z * Z * omc) * vz;
      x = make_double4(rx + o.x, ry + o.y, rz + o.z, x.w);
    }
    const double2 rp = A.rec_p[a];
    q = make_double4(rp.x, rp.y, 0.0, 0.0);
  }
  const double off = q.x - PXD_GB_OFFSET;
  A.cx[(size_t)p * A.stride + k] = x;
  A.cp[(size_t)p * A.stride + k] = make_double4(off, off * q.y, q.x, 0.0);
}

// The descreening integral of source atom i (screened radius sr) on target
// atom j (offset radius orj) at distance r12 (scorer.cpp:810-829); 0 when j
// lies inside i's screened sphere.
__device__ inline double gb_descreen(double r12, double sr, double orj) {
  const double U = r12 + sr;
  if (U < orj) return 0.0;
  const double r12_inv = 1.0 / r12;
  double D = fabs(r12 - sr);
  const double L = orj > D ? orj : D;
  const double U_inv = 1.0 / U, L_inv = 1.0 / L;
  D = 0.25 * (r12 - sr * sr * r12_inv) * (U_inv * U_inv - L_inv * L_inv);
  D += L_inv - U_inv + 0.5 * log(L * U_inv + kGbEps) * r12_inv;
  return 0.5 * D;
}

// Descreening sum of each fixed receptor atom j over the other fixed atoms
// (scorer.cpp:810-829 restricted to sources that never move): one thread per
// target, the sources staged through LDS.  Launched once per call.
__global__ __launch_bounds__(kGbBlock) void k_gb_fixed(const double4* __restrict__ rec_x,
                                                       const double2* __restrict__ rec_p,
                                                       int n_fix, double* __restrict__ sfix) {
  __shared__ double4 t_x[kGbBlock];
  __shared__ double t_sr[kGbBlock];
  const int j = blockIdx.x * kGbBlock + threadIdx.x;
  const bool act = j < n_fix;
  const double4 xj = act ? rec_x[j] : make_double4(0.0, 0.0, 0.0, 0.0);
  const double orj = act ? rec_p[j].x - PXD_GB_OFFSET : 1.0;
  double sum = 0.0;
  for (int base = 0; base < n_fix; base += kGbBlock) {
    __syncthreads();
    const int i = base + threadIdx.x;
    if (i < n_fix) {
      t_x[threadIdx.x] = rec_x[i];
      const double2 pi = rec_p[i];
      t_sr[threadIdx.x] = (pi.x - PXD_GB_OFFSET) * pi.y;
    }
    __syncthreads();
    const int cnt = min(kGbBlock, n_fix - base);
    if (act) {
      for (int u = 0; u < cnt; ++u) {
        if (base + u == j) continue;
        const double4 xi = t_x[u];
        const double dx = xi.x - xj.x, dy = xi.y - xj.y, dz = xi.z - xj.z;
        sum += gb_descreen(sqrt(dx * dx + dy * dy + dz * dz), t_sr[u], orj);
      }
    }
  }
  if (act) sfix[j] = sum;
}

// Accumulates into *sum the descreening of target j (position xj, offset
// radius orj) by the sources src(v), v in [0, nv), staged through LDS in
// tiles; every thread of the block calls it with the same nv.
template <class Src>
__device__ inline void gb_descreen_tiles(const double4* X, const double4* P, int nv, Src src,
                                         bool act, int j, double4 xj, double orj, double* sum,
                                         double4* t_x, double* t_sr, int* t_i) {
  for (int base = 0; base < nv; base += kGbBlock) {
    __syncthreads();
    const int v = base + threadIdx.x;
    if (v < nv) {
      const int i = src(v);
      t_x[threadIdx.x] = X[i];
      t_sr[threadIdx.x] = P[i].y;
      t_i[threadIdx.x] = i;
    }
    __syncthreads();
    const int cnt = min(kGbBlock, nv - base);
    if (act) {
      for (int u = 0; u < cnt; ++u) {
        if (t_i[u] == j) continue;
        const double4 xi = t_x[u];
        const double dx = xi.x - xj.x, dy = xi.y - xj.y, dz = xi.z - xj.z;
        *sum += gb_descreen(sqrt(dx * dx + dy * dy + dz * dz), t_sr[u], orj);
      }
    }
  }
}

__global__ __launch_bounds__(kGbBlock) void k_gb_born(GbArgs A) {
  __shared__ double4 t_x[kGbBlock];
  __shared__ double t_sr[kGbBlock];
  __shared__ int t_i[kGbBlock];
  const int p = blockIdx.z, s = blockIdx.y;
  const LigInfo li = A.info[A.pose_lig[A.pose0 + p]];
  const int n = li.n_atoms;
  int a0, a1;
  gb_range(s, n, A.n_rec, &a0, &a1);
  const int b0 = a0 + (int)blockIdx.x * kGbBlock;
  if (b0 >= a1) return;  // the whole block is past the range
  const int j = b0 + threadIdx.x;
  const bool act = j < a1;
  const double4* X = A.cx + (size_t)p * A.stride;
  const double4* P = A.cp + (size_t)p * A.stride;
  const double4 xj = act ? X[j] : make_double4(0.0, 0.0, 0.0, 0.0);
  const double4 pj = act ? P[j] : make_double4(1.0, 1.0, 1.0, 0.0);
  // the system's sources that move with the pose: its ligand atoms (ligand
  // and complex systems) and the movable receptor atoms [n, fix0), which
  // follow the ligand's
  const int nmov_rec = A.n_rec - A.n_fix, fix0 = n + nmov_rec;
  const int v0 = s == 0 ? n : 0, v1 = s == 1 ? n : fix0;
  const int nfix = s == 1 ? 0 : A.n_fix;
  double sum = 0.0;
  gb_descreen_tiles(X, P, v1 - v0, [&](int v) { return v0 + v; }, act, j, xj, pj.x, &sum, t_x,
                    t_sr, t_i);
  // the fixed receptor atoms [fix0, n + n_rec): a fixed target takes its
  // precomputed sum, the others (ligand and movable atoms, all in the
  // system's first block) sum them here; a block of fixed targets skips the
  // pass
  const bool fixed_target = act && j >= fix0;
  if (nfix > 0) {
    if (b0 < fix0)
      gb_descreen_tiles(X, P, nfix, [&](int v) { return fix0 + v; }, act && !fixed_target, j, xj,
                        pj.x, &sum, t_x, t_sr, t_i);
    if (fixed_target) sum += A.sfix[j - fix0];
  }
  if (!act) return;
  // Born radius (scorer.cpp:832-839): psi = I * (radius - offset)
  const double psi = sum * pj.x, psi2 = psi * psi;
  const double th = tanh(PXD_GB_ALPHA * psi - PXD_GB_BETA * psi2 + PXD_GB_GAMMA * psi * psi2);
  A.born[((size_t)p * 3 + s) * A.stride + j] = 1.0 / (1.0 / pj.x - th / pj.z);
}

// The GB pair term is symmetric in i and j and only each system's total is
// used, so row i sums the pairs (i, j > i) twice and its own diagonal term
// once: each block starts at its own tile, half the ordered pairs overall.
__global__ __launch_bounds__(kGbBlock) void k_gb_egb(GbArgs A) {
  __shared__ double4 t_x[kGbBlock];
  __shared__ double t_b[kGbBlock];
  __shared__ double t_ib[kGbBlock];  // 1 / (4 B_j): the exponent's divisor
  const int p = blockIdx.z, s = blockIdx.y;
  const LigInfo li = A.info[A.pose_lig[A.pose0 + p]];
  int a0, a1;
  gb_range(s, li.n_atoms, A.n_rec, &a0, &a1);
  const int b0 = a0 + (int)blockIdx.x * kGbBlock;
  if (b0 >= a1) return;
  const int i = b0 + threadIdx.x;
  const bool act = i < a1;
  const double4* X = A.cx + (size_t)p * A.stride;
  const double* B = A.born + ((size_t)p * 3 + s) * A.stride;
  const double4 xi = act ? X[i] : make_double4(0.0, 0.0, 0.0, 0.0);
  const double bi = act ? B[i] : 1.0;
  const double ibi = 1.0 / bi;
  const double pref = 0.5 * kGbCoulomb;  // 0.5 k (1 / solute - exp(-kappa f) / solvent)
  double egb = 0.0;
  for (int base = b0; base < a1; base += kGbBlock) {
    __syncthreads();
    const int j = base + threadIdx.x;
    if (j < a1) {
      t_x[threadIdx.x] = X[j];
      const double bj = B[j];
      t_b[threadIdx.x] = bj;
      t_ib[threadIdx.x] = 0.25 / bj;
    }
    __syncthreads();
    const int cnt = min(kGbBlock, a1 - base);
    if (act) {
      for (int u = 0; u < cnt; ++u) {
        const int jj = base + u;
        if (jj < i) continue;
        const double4 xj = t_x[u];
        double f, w;
        if (jj == i) {
          f = bi;
          w = 1.0;
        } else {
          const double dx = xi.x - xj.x, dy = xi.y - xj.y, dz = xi.z - xj.z;
          const double r2 = dx * dx + dy * dy + dz * dz;
          const double bp = bi * t_b[u];
          // r^2 / (4 B_i B_j) by the atoms' reciprocals: no division per pair
          f = sqrt(r2 + bp * exp(-(r2 * ibi) * t_ib[u]) + kGbEps);
          w = 2.0;
        }
        const double L = pref * (1.0 / PXD_GB_SOLUTE_DIELECTRIC -
                                 exp(-PXD_GB_KAPPA * f) / PXD_GB_SOLVENT_DIELECTRIC);
        egb -= w * xi.w * xj.w / f * L;
      }
    }
  }
  if (!act) return;
  // the surface term of atom i (scorer.cpp:860-864)
  const double r = A.cp[(size_t)p * A.stride + i].z;
  const double ratio = r / bi, r2 = ratio * ratio;
  const double esa = (r + PXD_GB_PROBE_RADIUS) * (r + PXD_GB_PROBE_RADIUS) * (r2 * r2 * r2);
  A.row[((size_t)p * 3 + s) * A.stride + i] = egb + esa * PXD_GB_SURFACE_ENERGY;
}

__global__ __launch_bounds__(kGbBlock) void k_gb_sum(GbArgs A) {
  __shared__ double red[kGbBlock];
  const int p = blockIdx.y, s = blockIdx.x;
  const LigInfo li = A.info[A.pose_lig[A.pose0 + p]];
  int a0, a1;
  gb_range(s, li.n_atoms, A.n_rec, &a0, &a1);
  const double* R = A.row + ((size_t)p * 3 + s) * A.stride;
  double acc = 0.0;
  for (int i = a0 + threadIdx.x; i < a1; i += kGbBlock) acc += R[i];
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int w = kGbBlock / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) A.energy[(size_t)(A.pose0 + p) * 3 + s] = red[0];
}

// score += coef * (E_com - E_lig - E_pro); the leaf terms and the three
// energies into terms[n][PXD_NUM_BSCORE_TERMS]
__global__ void k_gb_combine(int n_poses, double coef, const double* energy,
                             const double* leaf_terms, double* score, double* terms) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n_poses) return;
  const double ep = energy[3 * p], el = energy[3 * p + 1], ec = energy[3 * p + 2];
  if (coef != 0.0) score[p] += coef * (ec - el - ep);
  if (terms) {
    for (int t = 0; t < PXD_NUM_TERMS; ++t)
      terms[(size_t)p * PXD_NUM_BSCORE_TERMS + t] = leaf_terms[(size_t)p * PXD_NUM_TERMS + t];
    terms[(size_t)p * PXD_NUM_BSCORE_TERMS + PXD_NUM_TERMS] = ep;
    terms[(size_t)p * PXD_NUM_BSCORE_TERMS + PXD_NUM_TERMS + 1] = el;
    terms[(size_t)p * PXD_NUM_BSCORE_TERMS + PXD_NUM_TERMS + 2] = ec;
  }
}

pxd_status gb_fail(pxd_status st, const char* what, hipError_t e) {
  char buf[256];
  snprintf(buf, sizeof(buf), "bscore: %s: %s", what, hipGetErrorString(e));
  return pxd_set_error_(st, buf);
}

// Device buffers of one call, freed on every exit path.
struct GbBuffers {
  std::vector<void*> ptrs;
  ~GbBuffers() {
    for (void* p : ptrs) (void)hipFree(p);
  }
  template <class T>
  hipError_t alloc(T** out, size_t count) {
    void* p = nullptr;
    hipError_t e = hipMalloc(&p, std::max<size_t>(count, 1) * sizeof(T));
    if (e == hipSuccess) {
      ptrs.push_back(p);
      *out = static_cast<T*>(p);
    }
    return e;
  }
};

}  // namespace

extern "C" {

pxd_status pxd_gbsa_atom_params(int32_t n, const int32_t* an, int32_t n_bonds,
                                const int32_t* bonds, double* radius, double* screen) {
  if (n < 0 || (n > 0 && (!an || !radius || !screen)) || n_bonds < 0 || (n_bonds > 0 && !bonds))
    return pxd_set_error_(PXD_ERR_ARG, "gbsa_atom_params: bad argument");
  static const double kRadii[] = PXD_GB_RADII;
  static const double kScreens[] = PXD_GB_SCREENS;
  for (int i = 0; i < n; ++i) {
    if (an[i] < 0) return pxd_set_error_(PXD_ERR_CONF, "gbsa_atom_params: negative atomic number");
    radius[i] = -1.0;
  }
  for (int b = 0; b < n_bonds; ++b) {
    const int i = bonds[2 * b], j = bonds[2 * b + 1];
    if (i < 0 || i >= n || j < 0 || j >= n)
      return pxd_set_error_(PXD_ERR_CONF, "gbsa_atom_params: bond atom out of range");
    if (an[i] == 1 && an[j] == 7)
      radius[i] = PXD_GB_RADIUS_HN;
    else if (an[i] == 7 && an[j] == 1)
      radius[j] = PXD_GB_RADIUS_HN;
  }
  for (int i = 0; i < n; ++i) {
    if (radius[i] < 0.0) radius[i] = an[i] <= PXD_GB_NUM_RADII ? kRadii[an[i]] : kRadii[0];
    screen[i] = an[i] <= PXD_GB_NUM_SCREENS ? kScreens[an[i]] : kScreens[0];
  }
  return PXD_OK;
}

pxd_status pxd_bscore_batch(const pxd_receptor* r, const double* rec_radius,
                            const double* rec_screen, const pxd_ligand_batch* b,
                            const double* lig_radius, const double* lig_screen,
                            const pxd_bscore_params* bp, int32_t n_poses,
                            const int32_t* pose_lig, const int32_t* pose_atom_off,
                            const float* xyz, const float* tors, double* score,
                            double* terms, void* stream) {
  if (!r || !b || !bp) return pxd_set_error_(PXD_ERR_ARG, "bscore: null argument");
  if (n_poses < 0) return pxd_set_error_(PXD_ERR_ARG, "bscore: n_poses < 0");
  if (n_poses == 0) return PXD_OK;
  if (!pose_lig || !pose_atom_off || !xyz || !score)
    return pxd_set_error_(PXD_ERR_ARG, "bscore: null pose array");
  if (r->n_leaves > 0 && !tors) return pxd_set_error_(PXD_ERR_ARG, "bscore: tors required");
  const bool gb = bp->gbsa_coef > 1e-5 || bp->gbsa_coef < -1e-5;  // utility.h:24-26
  if (gb && (!rec_radius || !rec_screen || !lig_radius || !lig_screen))
    return pxd_set_error_(PXD_ERR_ARG, "bscore: GBSA needs the atom radii and screens");
  hipStream_t st = static_cast<hipStream_t>(stream);
  GbBuffers buf;
  hipError_t e;
  double* leaf_terms = nullptr;
  if ((e = buf.alloc(&leaf_terms, (size_t)n_poses * PXD_NUM_TERMS)) != hipSuccess)
    return gb_fail(PXD_ERR_HIP, "hipMalloc", e);
  // the leaves: the pscore kernel with the affinity parameters
  pxd_status ps = pxd_pscore_batch(r, b, &bp->leaves, n_poses, pose_lig, pose_atom_off, xyz,
                                   tors, score, leaf_terms, stream);
  if (ps != PXD_OK) return ps;
  double* energy = nullptr;
  if ((e = buf.alloc(&energy, (size_t)n_poses * 3)) != hipSuccess)
    return gb_fail(PXD_ERR_HIP, "hipMalloc", e);
  if ((e = hipMemsetAsync(energy, 0, sizeof(double) * 3 * n_poses, st)) != hipSuccess)
    return gb_fail(PXD_ERR_HIP, "hipMemsetAsync", e);
  if (gb) {
    const int N = r->n_atoms, nl = r->n_leaves;
    // receptor atoms in the device arrays: the movable ones first (input
    // order), right after a pose's ligand atoms, then the fixed ones; every
    // sum over them is order-fixed
    std::vector<int> leaf_of(N, -1), ord;
    ord.reserve(N);
    for (size_t m = 0; m < r->h_mov_atom.size(); ++m) leaf_of[r->h_mov_atom[m]] = r->h_mov_leaf[m];
    for (int a = 0; a < N; ++a)
      if (leaf_of[a] >= 0) ord.push_back(a);
    const int n_mov = (int)ord.size(), n_fix = N - n_mov;
    for (int a = 0; a < N; ++a)
      if (leaf_of[a] < 0) ord.push_back(a);
    std::vector<double4> rx(N);
    std::vector<double2> rp(N);
    std::vector<int> slot(N, -1);
    for (int k = 0; k < N; ++k) {
      const int a = ord[k];
      rx[k] = make_double4(r->h_xyz64[3 * a], r->h_xyz64[3 * a + 1], r->h_xyz64[3 * a + 2],
                           r->h_q64[a]);
      rp[k] = make_double2(rec_radius[a], rec_screen[a]);
      slot[k] = leaf_of[a];
    }
    std::vector<double4> lo(std::max(nl, 1)), la(std::max(nl, 1));
    for (int k = 0; k < nl; ++k) {
      lo[k] = make_double4(r->h_leaf_o[3 * k], r->h_leaf_o[3 * k + 1], r->h_leaf_o[3 * k + 2], 0.0);
      la[k] = make_double4(r->h_leaf_a[3 * k], r->h_leaf_a[3 * k + 1], r->h_leaf_a[3 * k + 2], 0.0);
    }
    const int nb = b->h_atom_off[b->n];
    std::vector<double4> lp(std::max(nb, 1));
    for (int a = 0; a < nb; ++a) lp[a] = make_double4(b->h_q64[a], lig_radius[a], lig_screen[a], 0.0);
    double4 *d_rx, *d_lo, *d_la, *d_lp, *d_cx, *d_cp;
    double2* d_rp;
    int* d_slot;
    double *d_born, *d_row, *d_sfix;
    const int stride = b->max_atoms + N;
    // poses per launch: about 192 MiB of per-pose workspace
    const size_t per_pose = (size_t)stride * (2 * sizeof(double4) + 6 * sizeof(double));
    const int chunk = (int)std::max<size_t>(1, std::min<size_t>(n_poses, (192ull << 20) / per_pose));
    if ((e = buf.alloc(&d_rx, N)) != hipSuccess || (e = buf.alloc(&d_rp, N)) != hipSuccess ||
        (e = buf.alloc(&d_slot, N)) != hipSuccess || (e = buf.alloc(&d_lo, lo.size())) != hipSuccess ||
        (e = buf.alloc(&d_la, la.size())) != hipSuccess ||
        (e = buf.alloc(&d_lp, lp.size())) != hipSuccess ||
        (e = buf.alloc(&d_cx, (size_t)chunk * stride)) != hipSuccess ||
        (e = buf.alloc(&d_cp, (size_t)chunk * stride)) != hipSuccess ||
        (e = buf.alloc(&d_born, (size_t)chunk * 3 * stride)) != hipSuccess ||
        (e = buf.alloc(&d_row, (size_t)chunk * 3 * stride)) != hipSuccess ||
        (e = buf.alloc(&d_sfix, n_fix)) != hipSuccess)
      return gb_fail(PXD_ERR_HIP, "hipMalloc", e);
    if ((e = hipMemcpyAsync(d_rx, rx.data(), sizeof(double4) * N, hipMemcpyHostToDevice, st)) != hipSuccess ||
        (e = hipMemcpyAsync(d_rp, rp.data(), sizeof(double2) * N, hipMemcpyHostToDevice, st)) != hipSuccess ||
        (e = hipMemcpyAsync(d_slot, slot.data(), sizeof(int) * N, hipMemcpyHostToDevice, st)) != hipSuccess ||
        (e = hipMemcpyAsync(d_lo, lo.data(), sizeof(double4) * lo.size(), hipMemcpyHostToDevice, st)) != hipSuccess ||
        (e = hipMemcpyAsync(d_la, la.data(), sizeof(double4) * la.size(), hipMemcpyHostToDevice, st)) != hipSuccess ||
        (e = hipMemcpyAsync(d_lp, lp.data(), sizeof(double4) * lp.size(), hipMemcpyHostToDevice, st)) != hipSuccess)
      return gb_fail(PXD_ERR_HIP, "hipMemcpyAsync", e);
    if (n_fix > 0)
      hipLaunchKernelGGL(k_gb_fixed, dim3((n_fix + kGbBlock - 1) / kGbBlock), dim3(kGbBlock), 0,
                         st, d_rx + n_mov, d_rp + n_mov, n_fix, d_sfix);
    GbArgs A{};
    A.n_rec = N;
    A.n_fix = n_fix;
    A.sfix = d_sfix;
    A.stride = stride;
    A.pose_lig = pose_lig;
    A.pose_off = pose_atom_off;
    A.xyz = xyz;
    A.tors = nl > 0 ? tors : nullptr;
    A.n_leaves = nl;
    A.info = b->dev.info;
    A.lig_p = d_lp;
    A.rec_x = d_rx;
    A.rec_p = d_rp;
    A.mov_slot = d_slot;
    A.leaf_o = d_lo;
    A.leaf_a = d_la;
    A.cx = d_cx;
    A.cp = d_cp;
    A.born = d_born;
    A.row = d_row;
    A.energy = energy;
    const int blocks = (stride + kGbBlock - 1) / kGbBlock;
    for (int p0 = 0; p0 < n_poses; p0 += chunk) {
      A.pose0 = p0;
      A.n_chunk = std::min(chunk, n_poses - p0);
      hipLaunchKernelGGL(k_gb_setup, dim3(blocks, A.n_chunk), dim3(kGbBlock), 0, st, A);
      hipLaunchKernelGGL(k_gb_born, dim3(blocks, 3, A.n_chunk), dim3(kGbBlock), 0, st, A);
      hipLaunchKernelGGL(k_gb_egb, dim3(blocks, 3, A.n_chunk), dim3(kGbBlock), 0, st, A);
      hipLaunchKernelGGL(k_gb_sum, dim3(3, A.n_chunk), dim3(kGbBlock), 0, st, A);
      if ((e = hipGetLastError()) != hipSuccess) return gb_fail(PXD_ERR_HIP, "launch", e);
    }
  }
  hipLaunchKernelGGL(k_gb_combine, dim3((n_poses + 255) / 256), dim3(256), 0, st, n_poses,
                     gb ? bp->gbsa_coef : 0.0, energy, leaf_terms, score, terms);
  if ((e = hipGetLastError()) != hipSuccess) return gb_fail(PXD_ERR_HIP, "launch", e);
  // the workspace is freed on return: finish first
  if ((e = hipStreamSynchronize(st)) != hipSuccess) return gb_fail(PXD_ERR_HIP, "sync", e);
  return PXD_OK;
}

}  // extern "C"
