// pxd_kernels.hip — grid build, batched score+grad and persistent batched
// L-BFGS kernels for gfx950, with their host launchers.
#include <hip/hip_runtime.h>

#include "pxd_device.hpp"
#include "pxd_kernels.hpp"
#include "pxd_pscore.hpp"
#include "pxd_pscore_grad.hpp"
#include "pxd_mc.hpp"
#include "pxd_leaf.hpp"

namespace pxd {

// Blocks are handed to the 8 XCDs round-robin; give each XCD a contiguous
// range of logical blocks so consecutive poses (same ligand, same term
// lists) share one XCD's L2.
__device__ inline int xcd_block(int b, int nb) {
  const int nx = 8;
  int q = nb / nx, r = nb % nx, x = b % nx, i = b / nx;
  return x * q + (x < r ? x : r) + i;
}

// ============================================================== K1: grid build
// receptor_cache::partial_worker::populate (grid.cpp:150-195): for every grid
// point, Coulomb for a unit probe and LJ per ligand type, each with the probe
// gradient, summed over the fixed (non-movable) receptor atoms.  One thread
// per point, TG types per pass, receptor atoms staged through LDS tiles.
// Distances and reciprocals use correctly rounded sqrt/division.
//
// Near atoms in FP64.  A map value or gradient is a sum of ~N_rec terms, and
// next to an atom single terms reach 1e5-1e6 while the channel they land in
// may sum to a few units: an FP32 term (and the FP32 coordinates and
// parameters it is computed from) then carries more error than the whole
// sum may have.  So an atom closer to a point than R = (sigma_atom +
// largest type sigma) / 2 -- every term of size above ~eps of the pair, and
// the steep Coulomb core -- is left out of the FP32 sweep for that point and
// summed afterwards in FP64 from FP64 coordinates, charges and LJ
// parameters.  The candidates are the atoms within R of the block's bounding
// box, collected per block in LDS while the tiles stream past (a few dozen;
// beyond kNearCap the rest stay FP32).  Sweep and FP64 pass test "closer
// than R" with the same uncontracted FP32 expression, so every term is
// counted exactly once.

constexpr int kGridTile = 256;
constexpr int kNearCap = 1024;

__device__ inline float dist2_rn(float vx, float vy, float vz) {
  return __fadd_rn(__fadd_rn(__fmul_rn(vx, vx), __fmul_rn(vy, vy)), __fmul_rn(vz, vz));
}

template <int TG>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) void k_grid(GridBuildArgs A, int type0,
                                                                                 int with_coulomb) {
  constexpr int kTile = kGridTile;
  __shared__ float4 s_at[kTile];
  __shared__ float2 s_lj[kTile];
  __shared__ float s_r2[kTile];      // R^2 of a near candidate, -1 otherwise
  __shared__ float2 s_tlj[TG];
  __shared__ float4 s_nat[kNearCap]; // candidate x, y, z, R^2
  __shared__ int s_nidx[kNearCap];
  __shared__ int s_nn;
  __shared__ int s_wcnt[kTile / kWave];  // candidates per wave of the current tile
  const int ny = A.np[1], nz = A.np[2];
  const float ix = A.init[0], iy = A.init[1], iz = A.init[2], spacing = A.spacing;
  long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  int nt = A.n_types - type0 < TG ? A.n_types - type0 : TG;
  if (threadIdx.x < TG)
    s_tlj[threadIdx.x] = threadIdx.x < nt ? A.tlj[type0 + threadIdx.x] : make_float2(0.f, 0.f);
  if (threadIdx.x == 0) s_nn = 0;
  int k = (int)(p % nz);
  int j = (int)((p / nz) % ny);
  int i = (int)(p / ((long long)nz * ny));
  // grid.cpp:165-169 (init + spacing * index)
  const float px = ix + spacing * i, py = iy + spacing * j, pz = iz + spacing * k;
  // the block's bounding box (its points are consecutive in x-major order)
  // and the largest type sigma, in LDS: only the tile loads read them
  __shared__ float s_bb[7];
  if (threadIdx.x == 0) {
    long long p0 = (long long)blockIdx.x * blockDim.x;
    long long p1 = p0 + blockDim.x - 1 < A.P - 1 ? p0 + blockDim.x - 1 : A.P - 1;
    long long nyz = (long long)ny * nz;
    int i0 = (int)(p0 / nyz), i1 = (int)(p1 / nyz), j0 = 0, j1 = ny - 1, k0 = 0, k1 = nz - 1;
    if (i0 == i1) {
      j0 = (int)((p0 / nz) % ny);
      j1 = (int)((p1 / nz) % ny);
      if (j0 == j1) {
        k0 = (int)(p0 % nz);
        k1 = (int)(p1 % nz);
      }
    }
    s_bb[0] = ix + spacing * i0; s_bb[3] = ix + spacing * i1;
    s_bb[1] = iy + spacing * j0; s_bb[4] = iy + spacing * j1;
    s_bb[2] = iz + spacing * k0; s_bb[5] = iz + spacing * k1;
    // over all requested types, so every pass (and every types-per-pass
    // choice) draws the same near/far line and the maps are bit-identical
    float smax = 0.f;
    for (int t = 0; t < A.n_types; ++t) smax = fmaxf(smax, A.tlj[t].x);
    s_bb[6] = smax;
  }
  double ce = 0.0, cg0 = 0.0, cg1 = 0.0, cg2 = 0.0;
  float acc[TG][4];
#pragma unroll
  for (int t = 0; t < TG; ++t) acc[t][0] = acc[t][1] = acc[t][2] = acc[t][3] = 0.f;
  for (int base = 0; base < A.n_atoms; base += kTile) {
    __syncthreads();
    int a = base + threadIdx.x;
    bool cand = false;
    float4 at = make_float4(0.f, 0.f, 0.f, 0.f);
    float r2 = 0.f;
    if (a < A.n_atoms) {
      at = A.atoms[a];
      float2 lj = A.alj[a];
      s_at[threadIdx.x] = at;
      s_lj[threadIdx.x] = lj;
      const float R = 0.5f * (lj.x + s_bb[6]);
      r2 = R * R;
      float dx = fmaxf(fmaxf(s_bb[0] - at.x, at.x - s_bb[3]), 0.f);
      float dy = fmaxf(fmaxf(s_bb[1] - at.y, at.y - s_bb[4]), 0.f);
      float dz = fmaxf(fmaxf(s_bb[2] - at.z, at.z - s_bb[5]), 0.f);
      cand = dx * dx + dy * dy + dz * dz < 1.01f * r2 + 1e-3f;
    }
    // candidates take list slots in atom order (a prefix count, not an
    // atomic counter), so the FP64 near sums below add in a fixed order and
    // the maps are the same bits on every build
    const uint64_t bal = __ballot(cand);
    const int wave = threadIdx.x / kWave;
    if ((threadIdx.x & (kWave - 1)) == 0) s_wcnt[wave] = __popcll(bal);
    __syncthreads();
    int slot = s_nn + __builtin_amdgcn_mbcnt_hi((uint32_t)(bal >> 32),
                                                __builtin_amdgcn_mbcnt_lo((uint32_t)bal, 0u));
    for (int w = 0; w < wave; ++w) slot += s_wcnt[w];
    if (a < A.n_atoms) {
      float keep = -1.f;
      if (cand && slot < kNearCap) {
        s_nat[slot] = make_float4(at.x, at.y, at.z, r2);
        s_nidx[slot] = a;
        keep = r2;
      }
      s_r2[threadIdx.x] = keep;
    }
    __syncthreads();  // every thread has read s_nn and the tile is staged
    if (threadIdx.x == 0) {
      int tot = 0;
      for (int w = 0; w < kTile / kWave; ++w) tot += s_wcnt[w];
      s_nn += tot;  // read again only after the next barrier
    }
    int cnt = A.n_atoms - base < kTile ? A.n_atoms - base : kTile;
    for (int u = 0; u < cnt; ++u) {
      float4 at = s_at[u];
      float2 lj = s_lj[u];
      float vx = at.x - px, vy = at.y - py, vz = at.z - pz;
      float d2 = dist2_rn(vx, vy, vz);
      const bool near = d2 < s_r2[u];  // summed in FP64 after the sweep
      float d = sqrtf(d2);
      float inv_d = near ? 0.f : 1.f / (d + kEps);
      float irsqr = 1.f / (d2 + kEps);
      if (with_coulomb) {
        float q = near ? 0.f : at.w;
        float e = q * inv_d * kCoulomb;
        float de = -(q / (d2 * d + kEps)) * kCoulomb;
        ce += e;
        cg0 -= (double)(vx * de);
        cg1 -= (double)(vy * de);
        cg2 -= (double)(vz * de);
      }
#pragma unroll
      for (int t = 0; t < TG; ++t) {
        float2 tl = s_tlj[t];
        float uu = (tl.x + lj.x) * 0.5f * inv_d;
        float u2 = uu * uu;
        float u6 = u2 * u2 * u2;
        float u12 = u6 * u6;
        float e4 = tl.y * lj.y;
        acc[t][0] += (u12 - u6) * e4;
        float de = (-12.f * u12 + 6.f * u6) * irsqr * e4;
        acc[t][1] -= vx * de;
        acc[t][2] -= vy * de;
        acc[t][3] -= vz * de;
      }
    }
  }
  __syncthreads();  // the last tile's count is in s_nn
  if (p >= A.P) return;
  const int nn = s_nn < kNearCap ? s_nn : kNearCap;
  const double qx = A.init64[0] + A.spacing64 * i, qy = A.init64[1] + A.spacing64 * j,
               qz = A.init64[2] + A.spacing64 * k;
  const float big = 3.0e37f;  // FP32 maps saturate instead of overflowing
  auto sat = [&](double v) { return (float)fmin(fmax(v, -(double)big), (double)big); };
  auto near_term = [&](int c, double* dv, double* d2o, double* dd) -> bool {
    float4 na = s_nat[c];
    float vx = na.x - px, vy = na.y - py, vz = na.z - pz;
    if (!(dist2_rn(vx, vy, vz) < na.w)) return false;
    double4 at = A.atoms64[s_nidx[c]];
    dv[0] = at.x - qx;
    dv[1] = at.y - qy;
    dv[2] = at.z - qz;
    *d2o = dv[0] * dv[0] + dv[1] * dv[1] + dv[2] * dv[2];
    *dd = sqrt(*d2o);
    return true;
  };
  if (with_coulomb) {
    for (int c = 0; c < nn; ++c) {
      double v[3], d2, d;
      if (!near_term(c, v, &d2, &d)) continue;
      const double q = A.atoms64[s_nidx[c]].w, kc = 332.0637141491396;
      ce += q / (d + 1e-20) * kc;
      const double de = -(q / (d2 * d + 1e-20)) * kc;
      cg0 -= v[0] * de;
      cg1 -= v[1] * de;
      cg2 -= v[2] * de;
    }
    A.maps[p] = make_float4(sat(ce), sat(cg0), sat(cg1), sat(cg2));
  }
  // the type maps: the FP32 sums of the sweep are stored first (which frees
  // their registers), then every pair of maps is read back, completed with
  // the FP64 near terms and stored again
#pragma unroll
  for (int t = 0; t < TG; ++t)
    if (t < nt)
      A.maps[(long long)(1 + type0 + t) * A.P + p] =
          make_float4(sat(acc[t][0]), sat(acc[t][1]), sat(acc[t][2]), sat(acc[t][3]));
  if (nn == 0) return;
  constexpr int kC = 2;  // maps completed per pass over the near list
  for (int t0 = 0; t0 < nt; t0 += kC) {
    double s4[kC][4];
#pragma unroll
    for (int t = 0; t < kC; ++t) {
      float4 m = t0 + t < nt ? A.maps[(long long)(1 + type0 + t0 + t) * A.P + p]
                             : make_float4(0.f, 0.f, 0.f, 0.f);
      s4[t][0] = m.x;
      s4[t][1] = m.y;
      s4[t][2] = m.z;
      s4[t][3] = m.w;
    }
    bool any = false;
    for (int c = 0; c < nn; ++c) {
      double v[3], d2, d;
      if (!near_term(c, v, &d2, &d)) continue;
      any = true;
      const double2 alj = A.alj64[s_nidx[c]];
      const double inv = 1.0 / (d + 1e-20), irs = 1.0 / (d2 + 1e-20);
#pragma unroll
      for (int t = 0; t < kC; ++t) {
        const double2 tl = A.tlj64[type0 + (t0 + t < nt ? t0 + t : t0)];
        const double uu = (tl.x + alj.x) * 0.5 * inv;
        const double u2 = uu * uu, u6 = u2 * u2 * u2, u12 = u6 * u6;
        const double e4 = tl.y * alj.y;
        s4[t][0] += (u12 - u6) * e4;
        const double de = (-12.0 * u12 + 6.0 * u6) * irs * e4;
        s4[t][1] -= v[0] * de;
        s4[t][2] -= v[1] * de;
        s4[t][3] -= v[2] * de;
      }
    }
    if (!any) return;  // no atom of the list is near this point
#pragma unroll
    for (int t = 0; t < kC; ++t)
      if (t0 + t < nt)
        A.maps[(long long)(1 + type0 + t0 + t) * A.P + p] =
            make_float4(sat(s4[t][0]), sat(s4[t][1]), sat(s4[t][2]), sat(s4[t][3]));
  }
}

template <int TG>
static void grid_passes(const GridBuildArgs& a, hipStream_t st) {
  const int threads = 256;
  long long blocks = (a.P + threads - 1) / threads;
  for (int t0 = 0, first = 1; t0 < a.n_types || first; t0 += TG, first = 0) {
    hipLaunchKernelGGL(k_grid<TG>, dim3((unsigned)blocks), dim3(threads), 0, st, a, t0, first);
  }
}

void launch_grid_build(const GridBuildArgs& a, int types_per_pass, hipStream_t st) {
  if (types_per_pass <= 4)
    grid_passes<4>(a, st);
  else if (types_per_pass <= 8)
    grid_passes<8>(a, st);
  else
    grid_passes<16>(a, st);
}

// Dynamic LDS of a launch (sized by the host per launch): the pose arrays of
// k_score, the pscore ring geometry of the pscore kernels and MC walkers.
extern __shared__ __align__(16) float dyn_lds[];
static_assert(pose_dims(kMaxAtoms, kMaxMovable, kMaxLeaves).total == PXD_POSE_LDS_FLOATS,
              "pose LDS layout at the maxima");

// ============================================================== K2: score+grad
// One wavefront per pose; W poses (waves) per block, each wave with its own
// pose arrays in the block's dynamic LDS.  The per-pose code synchronises
// within its wave only (wave_sync), so the waves of a block are independent.

// Occupancy target 6 waves per SIMD: the compiler then schedules within
// 85 VGPRs (77 used, no spills); measured 1.91 ms at the default 4 waves
// (101 VGPRs), 1.70 ms at 6, 1.74 ms at 5 and 7 on the bench batch.
template <int W>
__global__ __launch_bounds__(64 * W) __attribute__((amdgpu_waves_per_eu(6))) void k_score(
    DevReceptor R, DevLigands L, const DevGrid* __restrict__ G, int n_poses,
    const int* __restrict__ pose_lig, const int* __restrict__ pose_off,
    const float* __restrict__ xyz, const float* __restrict__ tors, double* __restrict__ energy,
    float* __restrict__ lig_grad, float* __restrict__ tors_grad, int max_atoms) {
  // pose arrays sized for this launch's largest ligand and this receptor
  const PoseDims dims = pose_dims(max_atoms, R.n_mov, R.n_leaves);
  const int wv = W > 1 ? (int)(threadIdx.x >> 6) : 0;
  PoseLds s = carve_pose_lds(dyn_lds + wv * dims.total, dims);
  const int p = xcd_block(blockIdx.x, gridDim.x) * W + wv;
  if (p >= n_poses) return;
  const int lane = lane_id();
  const LigInfo li = L.info[pose_lig[p]];
  load_ligand_params(L, li, G, s);
  const long long off = 3ll * pose_off[p];
  for (int k = lane; k < li.n_atoms * 3; k += kWave) s.xyz[k] = xyz[off + k];
  for (int k = lane; k < R.n_leaves; k += kWave)
    s.tors[k] = tors ? tors[(long long)p * R.n_leaves + k] : 0.f;
  wave_sync();
  double e = eval_pose(R, L, G, li, s);
  if (lane == 0) energy[p] = e;
  for (int k = lane; k < li.n_atoms * 3; k += kWave) lig_grad[off + k] = s.grad[k];
  if (tors_grad)
    for (int k = lane; k < R.n_leaves; k += kWave)
      tors_grad[(long long)p * R.n_leaves + k] = s.torsg[k];
}

template <int W>
static void launch_score_w(const DevReceptor& R, const DevLigands& L, const DevGrid* G,
                           const ScoreArgs& a, int ma, hipStream_t st) {
  const size_t lds = sizeof(float) * W * pose_dims(ma, R.n_mov, R.n_leaves).total;
  const int blocks = (a.n_poses + W - 1) / W;
  hipLaunchKernelGGL(k_score<W>, dim3(blocks), dim3(kWave * W), lds, st, R, L, G, a.n_poses,
                     a.pose_lig, a.pose_off, a.xyz, a.tors, a.energy, a.lig_grad, a.tors_grad,
                     ma);
}

void launch_score(const DevReceptor& R, const DevLigands& L, const DevGrid* G,
                  const ScoreArgs& a, hipStream_t st) {
  if (a.n_poses <= 0) return;
  const int ma = a.max_atoms > 0 && a.max_atoms <= kMaxAtoms ? a.max_atoms : kMaxAtoms;
  if (a.poses_per_block == 4)
    launch_score_w<4>(R, L, G, a, ma, st);
  else if (a.poses_per_block == 2)
    launch_score_w<2>(R, L, G, a, ma, st);
  else
    launch_score_w<1>(R, L, G, a, ma, st);
}

// ============================================================== K3: L-BFGS
// lbfgs_step::apply + strong_wolfe (optimizer.cpp:205-560), one pose per
// wavefront, many poses per launch from a device work queue.  The flat
// variable [ligand xyz | receptor torsions] is distributed over the lanes
// (element e = lane + 64 k); each lane only ever touches its own elements of
// the LDS-resident vectors, so vector algebra needs no synchronisation.  The
// s/y history lives in the slot's global scratch; scalar control (f, t, gtd,
// curvature s.y / y.y, Armijo) is FP64.

// Per-wave LDS vector pool: the variables, gradient, previous gradient,
// direction and three line-search buffers (the bracketing phase holds the
// new and the previous point's gradients, the zoom the bracket's two ends
// and the new point's: never more than three at once).
enum { V_X = 0, V_G, V_GP, V_D, V_W0, V_W1, V_W2, V_COUNT };

// SW: floats per bonded-term stage slot of the pose arrays (pose_dims).
template <int VPL, int SW = 3>
struct LbfgsCtx {
  const DevReceptor* R;
  const DevLigands* L;
  const DevGrid* G;
  LigInfo li;
  PoseLds s;
  float* v;  // V_COUNT vectors of kWave*VPL floats
  int n3, n;  // n3: ligand variables (3 N, or 6 + n_tors), n: all variables
  int nk;     // vector elements per lane in use: ceil(n / 64) <= VPL
  int nevals;
  int n_iter;          // iterations of the last lbfgs_solve
  long long hist_io;   // history entries read (two-loop) and written by it
  // rigid-body + torsion variables: the input pose, its centre, LDS scratch
  int mode;
  float* xyz0;   // [kMaxAtoms*3]
  float c0[3];
  ApplyLds ap;
  float* tq;     // [kMaxFrags*3]
  float* dv;     // [6 + kMaxFrags] current variables / chain-rule output
  __device__ float* vec(int i) const { return v + i * (kWave * VPL); }
};

#define PXD_DOF_LDS_FLOATS \
  (pxd::kMaxAtoms * 3 + PXD_APPLY_LDS_FLOATS + pxd::kMaxFrags * 3 + 6 + pxd::kMaxFrags)

// Pose of the rigid-body + torsion variables held in c.dv (translation,
// rotation vector, ligand torsions; receptor torsions already in c.s.tors):
// free_ligand::apply_parameters (data.cpp:592-641) of the input pose about
// its centre.  cen receives the rotation centre.
template <int VPL, int SW>
__device__ inline void dof_pose(LbfgsCtx<VPL, SW>& c, float* cen) {
  const int lane = lane_id();
  const LigInfo& li = c.li;
  cen[0] = c.c0[0] + c.dv[0];
  cen[1] = c.c0[1] + c.dv[1];
  cen[2] = c.c0[2] + c.dv[2];
  float R[9];
  rotvec_matrix(c.dv + 3, R);
  for (int f = lane; f < li.n_frags; f += kWave) c.ap.th[f] = 0.f;
  for (int k = lane; k < 3 * li.n_atoms; k += kWave) c.s.xyz[k] = c.xyz0[k];
  __syncthreads();
  const int* tf = c.L->tors_frag + li.frag_off;
  for (int k = lane; k < li.n_tors; k += kWave) c.ap.th[tf[k]] = c.dv[6 + k];
  __syncthreads();
  apply_ligand(*c.L, li, c.s.xyz, cen, R, c.ap);
}

// Vector algebra over the nk elements per lane in use (nk is wave-uniform).
template <int VPL>
__device__ inline double vdot(int nk, const float* a, const float* b) {
  const int lane = lane_id();
  double acc = 0.0;
#pragma unroll
  for (int k = 0; k < VPL; ++k)
    if (k < nk) acc += (double)a[lane + kWave * k] * (double)b[lane + kWave * k];
  return wave_sum(acc);
}

template <int VPL>
__device__ inline float vamax(int nk, const float* a) {
  const int lane = lane_id();
  float m = 0.f;
#pragma unroll
  for (int k = 0; k < VPL; ++k)
    if (k < nk) m = fmaxf(m, fabsf(a[lane + kWave * k]));
  return wave_max(m);
}

template <int VPL>
__device__ inline void vcopy(int nk, float* dst, const float* src) {
  const int lane = lane_id();
#pragma unroll
  for (int k = 0; k < VPL; ++k)
    if (k < nk) dst[lane + kWave * k] = src[lane + kWave * k];
}

// Two-element bracket state held in scalars (a runtime-indexed local array
// would live in scratch memory).
struct Pair2 {
  double a, b;
  __device__ double get(int i) const { return i ? b : a; }
  __device__ void set(int i, double v) {
    if (i)
      b = v;
    else
      a = v;
  }
};
struct IPair2 {
  int a, b;
  __device__ int get(int i) const { return i ? b : a; }
  __device__ void set(int i, int v) {
    if (i)
      b = v;
    else
      a = v;
  }
};

// obj_func_closure::evaluate_direction (optimizer.cpp:155-167): f and g at x + t d.
template <int VPL, int SW>
__device__ inline double eval_dir(LbfgsCtx<VPL, SW>& c, double t, const float* d, float* g) {
  const int lane = lane_id();
  const float* x = c.vec(V_X);
  float tf = (float)t;
  if (c.mode) {
#pragma unroll
    for (int k = 0; k < VPL; ++k) if (k < c.nk) {
      int e = lane + kWave * k;
      float xv = x[e] + d[e] * tf;
      if (e < c.n3)
        c.dv[e] = xv;
      else if (e < c.n)
        c.s.tors[e - c.n3] = xv;
    }
    __syncthreads();
    float cen[3], w[3] = {c.dv[3], c.dv[4], c.dv[5]};
    dof_pose(c, cen);
    double f = eval_pose<SW>(*c.R, *c.L, c.G, c.li, c.s);
    dof_grad(*c.L, c.li, c.s.xyz, c.s.grad, cen, c.tq, c.dv);
    float tau[3] = {c.dv[3], c.dv[4], c.dv[5]}, jw[3];
    so3_jlt(w, tau, jw);
#pragma unroll
    for (int k = 0; k < VPL; ++k) if (k < c.nk) {
      int e = lane + kWave * k;
      g[e] = e < 3 ? c.dv[e]
                   : (e < 6 ? jw[e - 3]
                            : (e < c.n3 ? c.dv[e] : (e < c.n ? c.s.torsg[e - c.n3] : 0.f)));
    }
    __syncthreads();
    c.nevals++;
    return f;
  }
#pragma unroll
  for (int k = 0; k < VPL; ++k) if (k < c.nk) {
    int e = lane + kWave * k;
    float xv = x[e] + d[e] * tf;
    if (e < c.n3)
      c.s.xyz[e] = xv;
    else if (e < c.n)
      c.s.tors[e - c.n3] = xv;
  }
  __syncthreads();
  double f = eval_pose<SW>(*c.R, *c.L, c.G, c.li, c.s);
#pragma unroll
  for (int k = 0; k < VPL; ++k) if (k < c.nk) {
    int e = lane + kWave * k;
    g[e] = e < c.n3 ? c.s.grad[e] : (e < c.n ? c.s.torsg[e - c.n3] : 0.f);
  }
  __syncthreads();
  c.nevals++;
  return f;
}

// optimizer.cpp:205-233
__device__ inline double cubic_interpolate(double x1, double f1, double g1, double x2, double f2,
                                           double g2, double lo, double hi) {
  double d1 = g1 + g2 - 3.0 * (f1 - f2) / (x1 - x2);
  double d2sq = d1 * d1 - g1 * g2;
  if (d2sq >= 0.0) {
    double d2 = sqrt(d2sq), mp;
    if (x1 <= x2)
      mp = x2 - (x2 - x1) * ((g2 + d2 - d1) / (g2 - g1 + 2.0 * d2));
    else
      mp = x1 - (x1 - x2) * ((g1 + d2 - d1) / (g1 - g2 + 2.0 * d2));
    double t = mp > lo ? mp : lo;
    return t < hi ? t : hi;
  }
  return (lo + hi) / 2.0;
}

// strong_wolfe (optimizer.cpp:235-427).  Gradient vectors are named buffers
// in the LDS pool; the reference's std::move of vectors becomes a swap of
// buffer indices (uniform scalars), so no vector is copied in the loop.
template <int VPL, int SW>
__device__ void strong_wolfe(LbfgsCtx<VPL, SW>& c, const LbfgsParamsDev& p, double* t_io,
                             const float* d, double* f_io, double gtd) {
  double t = *t_io, f = *f_io;
  int b_new = V_W0, b_prev = V_W1;
  Pair2 br{0, 0}, bf{0, 0}, bgtd{0, 0};
  // buffers of bracket_g[0], bracket_g[1]: always assigned below before the
  // zoom, from the two bracketing buffers V_W0 / V_W1
  IPair2 idx{V_W0, V_W1};
  double f_new = eval_dir(c, t, d, c.vec(b_new));
  if (isnan(f_new)) {
    *f_io = f_new;
    return;
  }
  double gtd_new = vdot<VPL>(c.nk, c.vec(b_new), d);
  if (fabs(gtd_new) > p.gtd_threshold) {
    *f_io = f_new;
    vcopy<VPL>(c.nk, c.vec(V_G), c.vec(b_new));
    return;
  }
  double t_prev = 0.0, f_prev = f, gtd_prev = gtd;
  vcopy<VPL>(c.nk, c.vec(b_prev), c.vec(V_G));
  bool done = false;
  int ls_iter = 0;
  while (ls_iter < p.max_ls) {
    if ((f_new > f + p.c1 * t * gtd) || (ls_iter > 1 && f_new >= f_prev)) {
      br = {t_prev, t}; bf = {f_prev, f_new}; bgtd = {gtd_prev, gtd_new};
      idx = {b_prev, b_new};
      break;
    }
    if (fabs(gtd_new) <= -p.c2 * gtd) {
      br.a = t; bf.a = f_new;
      idx = {b_new, b_prev};
      done = true;
      break;
    }
    if (gtd_new >= 0) {
      br = {t_prev, t}; bf = {f_prev, f_new}; bgtd = {gtd_prev, gtd_new};
      idx = {b_prev, b_new};
      break;
    }
    double lo = t + 0.01 * (t - t_prev), hi = t * 10.0, tmp = t;
    t = cubic_interpolate(t_prev, f_prev, gtd_prev, t, f_new, gtd_new, lo, hi);
    t_prev = tmp;
    f_prev = f_new;
    int sw = b_prev;  // g_prev <- g_new (a move): swap buffer roles
    b_prev = b_new;
    b_new = sw;
    gtd_prev = gtd_new;
    f_new = eval_dir(c, t, d, c.vec(b_new));
    gtd_new = vdot<VPL>(c.nk, c.vec(b_new), d);
    ls_iter++;
  }
  double d_norm = 0.0;
  if (ls_iter == p.max_ls) {
    br = {0.0, t}; bf = {f, f_new};
    vcopy<VPL>(c.nk, c.vec(b_prev), c.vec(V_G));  // bracket_g[0] = g
    idx = {b_prev, b_new};
  } else {
    d_norm = (double)vamax<VPL>(c.nk, d);
  }
  // a pool buffer outside the bracket receives each zoom evaluation
  int z_new = V_W0;
  while (z_new == idx.a || z_new == idx.b) ++z_new;
  bool insuf = false;
  int low = 0, high = 1;
  if (!done && bf.a > bf.b) {
    low = 1;
    high = 0;
  }
  while (!done && ls_iter < p.max_ls) {
    double lo, hi;
    if (br.a > br.b) { lo = br.b; hi = br.a; } else { lo = br.a; hi = br.b; }
    double eps = hi - lo;
    if (eps * d_norm < p.ls_tol_change) break;
    t = cubic_interpolate(br.a, bf.a, bgtd.a, br.b, bf.b, bgtd.b, lo, hi);
    eps *= 0.1;
    double m1 = hi - t, m2 = t - lo;
    if ((m1 < m2 ? m1 : m2) < eps) {
      if (insuf || t >= hi || t <= lo) {
        t = fabs(t - hi) < fabs(t - lo) ? hi - eps : lo + eps;
        insuf = false;
      } else {
        insuf = true;
      }
    } else {
      insuf = false;
    }
    f_new = eval_dir(c, t, d, c.vec(z_new));
    gtd_new = vdot<VPL>(c.nk, c.vec(z_new), d);
    ls_iter++;
    if (f_new > (f + p.c1 * t * gtd) || f_new >= bf.get(low)) {
      // bracket[high] <- new point; its old buffer becomes the zoom scratch
      br.set(high, t); bf.set(high, f_new); bgtd.set(high, gtd_new);
      int old = idx.get(high);
      idx.set(high, z_new);
      z_new = old;
      if (bf.a > bf.b) { low = 1; high = 0; } else { low = 0; high = 1; }
    } else {
      if (fabs(gtd_new) <= -p.c2 * gtd) {
        done = true;
      } else if (gtd_new * (br.get(high) - br.get(low)) >= 0.0) {
        // old high becomes new low
        br.set(high, br.get(low)); bf.set(high, bf.get(low)); bgtd.set(high, bgtd.get(low));
        int old = idx.get(high);
        idx.set(high, idx.get(low));
        idx.set(low, old);
      }
      br.set(low, t); bf.set(low, f_new); bgtd.set(low, gtd_new);
      int old = idx.get(low);
      idx.set(low, z_new);
      z_new = old;
    }
  }
  *t_io = br.get(low);
  *f_io = bf.get(low);
  vcopy<VPL>(c.nk, c.vec(V_G), c.vec(idx.get(low)));
}

// lbfgs_step::apply (optimizer.cpp:429-560) on the pose held in the context:
// x = c.vec(V_X) is the start (ligand xyz then receptor torsions, c.n3 / c.n
// set, c.nevals = 0); on return x is the result.  hy / hs: this wave's history
// scratch (history x 64*VPL floats each).
// s_al: the two-loop's alphas, history + 1 entries (one slot for masked steps).
template <int VPL, int SW>
__device__ void lbfgs_solve(LbfgsCtx<VPL, SW>& c, const LbfgsParamsDev& P, float* hy, float* hs,
                            double* hr, double* s_al, double* energy_out,
                            bool* converged_out, bool evaluate_only = false) {
  const int lane = lane_id();
  const int m = P.history;
  c.n_iter = 0;
  c.hist_io = 0;
  // history entries are stored at the class's full width (VPL elements per
  // lane; m entries fill the caller's history*64*VPL floats of hy and of hs)
  constexpr int NV = kWave * VPL;
  float* x = c.vec(V_X);
  float* g = c.vec(V_G);
  float* gp = c.vec(V_GP);
  float* d = c.vec(V_D);
  float* y = c.vec(V_W0);
  float* sv = c.vec(V_W1);
#pragma unroll
  for (int k = 0; k < VPL; ++k) if (k < c.nk) d[lane + kWave * k] = 0.f;
  // initial f(x), df/dx (optimizer.cpp:435-441)
  double loss = eval_dir(c, 0.0, d, g);
  double energy = loss;  // kept unless a step is taken (optimizer.cpp:554-558)
  bool converged = false, updated = false;
  if (evaluate_only) {
    // the energy at x alone (no iteration, x unchanged)
  } else if (vamax<VPL>(c.nk, g) < (float)P.tol_grad) {
    converged = true;
  } else {
    int head = 0, num_old = 0, n_iter = 0;
    long long hist_io = 0;
    bool high_change = false;
    double H = 1.0, t = 0.0, prev_loss;
    while (n_iter < P.max_iter && c.nevals < P.max_eval) {
      n_iter++;
      if (n_iter > 1) {
#pragma unroll
        for (int k = 0; k < VPL; ++k) if (k < c.nk) y[lane + kWave * k] = g[lane + kWave * k] - gp[lane + kWave * k];
        high_change = vamax<VPL>(c.nk, y) > 1e8f;
      }
      if (n_iter == 1 || high_change) {
#pragma unroll
        for (int k = 0; k < VPL; ++k) if (k < c.nk) d[lane + kWave * k] = -g[lane + kWave * k];
        num_old = 0;
        head = 0;
        H = 1.0;
      } else {
        float tf = (float)t;
#pragma unroll
        for (int k = 0; k < VPL; ++k) if (k < c.nk) sv[lane + kWave * k] = d[lane + kWave * k] * tf;
        double ys = vdot<VPL>(c.nk, y, sv);
        if (ys > 1e-10) {  // optimizer.cpp:473-487
          int slot;
          if (num_old == m) {
            slot = head;
            head = (head + 1) % m;
          } else {
            slot = (head + num_old) % m;
            num_old++;
          }
          H = ys / vdot<VPL>(c.nk, y, y);
          // the slot's full class width is written, zeros past the pose's
          // nk elements, so the two-loop needs no masks
#pragma unroll
          for (int k = 0; k < VPL; ++k) {
            const bool in = k < c.nk;
            hy[(size_t)slot * NV + lane + kWave * k] = in ? y[lane + kWave * k] : 0.f;
            hs[(size_t)slot * NV + lane + kWave * k] = in ? sv[lane + kWave * k] : 0.f;
          }
          hr[slot] = 1.0 / ys;  // every lane stores (and later reads) the same value
          hist_io += 1;
          __syncthreads();
        }
        // two-loop recursion (optimizer.cpp:493-507), its dot products in FP32
        // (FP64 stays for s.y, y.y, the Armijo test and the line search's
        // scalars).  q stays in registers;
        // history entries are loaded kPre steps ahead of their use into a
        // ring of register buffers (statically indexed by unrolling kPre
        // steps).  Every load is unconditional — indices are clamped, and a
        // step past the last entry is masked out of the arithmetic instead of
        // branched around — so the compiler can count the loads in flight and
        // only waits for the one entry it needs.
        // The arithmetic does not depend on kPre or VPL: every size class
        // gives the same result for the same pose.
        constexpr int kPre = VPL <= 2 ? PXD_LBFGS_PRE2 : (VPL <= 4 ? PXD_LBFGS_PRE4 : PXD_LBFGS_PRE7);
        hist_io += 2 * num_old;  // each loop reads every stored entry once
        float qr[VPL], bs[kPre][VPL], by[kPre][VPL];
        double br[kPre];
#pragma unroll
        for (int k = 0; k < VPL; ++k) qr[k] = k < c.nk ? -g[lane + kWave * k] : 0.f;
        // A slot holds the class's full width (VPL elements per lane), so
        // every element of a slot is read from one wave-uniform base with
        // constant offsets.  Elements past the pose's nk are zero in every
        // entry of this pose (and in q), so they add exact zeros: the
        // arithmetic needs no masks and is the same in every size class.
        auto fetch = [&](int h, float* ds, float* dy, double& rho) {  // h clamped to [0, num_old)
          h = min(max(h, 0), max(num_old - 1, 0));
          unsigned sl = (unsigned)head + (unsigned)h;  // head, h < m: one wrap at most
          sl = sl >= (unsigned)m ? sl - (unsigned)m : sl;
          const float* ps = hs + (size_t)sl * NV;
          const float* py = hy + (size_t)sl * NV;
#pragma unroll
          for (int k = 0; k < VPL; ++k) {
            ds[k] = ps[lane + kWave * k];
            dy[k] = py[lane + kWave * k];
          }
          rho = hr[sl];
        };
        const int ngrp = (num_old + kPre - 1) / kPre;
#pragma unroll
        for (int j = 0; j < kPre; ++j) fetch(num_old - 1 - j, bs[j], by[j], br[j]);
        for (int grp = 0; grp < ngrp; ++grp) {
#pragma unroll
          for (int j = 0; j < kPre; ++j) {
            const int h = num_old - 1 - (grp * kPre + j);
            float acc = 0.f;
#pragma unroll
            for (int k = 0; k < VPL; ++k) acc += bs[j][k] * qr[k];
            double al = (double)wave_sum(acc) * br[j];
            al = h >= 0 ? al : 0.0;
            const float alf = (float)al;
            // every lane stores the same value; a masked step writes slot m
            s_al[h >= 0 ? h : m] = al;
#pragma unroll
            for (int k = 0; k < VPL; ++k) qr[k] -= alf * by[j][k];
            fetch(h - kPre, bs[j], by[j], br[j]);
          }
        }
        __syncthreads();
        float Hf = (float)H;
#pragma unroll
        for (int k = 0; k < VPL; ++k) qr[k] *= Hf;  // r = H q, kept in qr
        double ba[kPre];  // alpha of the ring's entries, read ahead like the vectors
#pragma unroll
        for (int j = 0; j < kPre; ++j) {
          fetch(j, bs[j], by[j], br[j]);
          ba[j] = s_al[max(min(j, num_old - 1), 0)];
        }
        for (int grp = 0; grp < ngrp; ++grp) {
#pragma unroll
          for (int j = 0; j < kPre; ++j) {
            const int h = grp * kPre + j;
            float acc = 0.f;
#pragma unroll
            for (int k = 0; k < VPL; ++k) acc += by[j][k] * qr[k];
            const double be = (double)wave_sum(acc) * br[j];
            float cf = (float)(ba[j] - be);
            cf = h < num_old ? cf : 0.f;
#pragma unroll
            for (int k = 0; k < VPL; ++k) qr[k] += bs[j][k] * cf;
            fetch(h + kPre, bs[j], by[j], br[j]);
            ba[j] = s_al[max(min(h + kPre, num_old - 1), 0)];
          }
        }
#pragma unroll
        for (int k = 0; k < VPL; ++k)
          if (k < c.nk) d[lane + kWave * k] = qr[k];
      }
      vcopy<VPL>(c.nk, gp, g);
      prev_loss = loss;
      if (n_iter == 1 || high_change) {
        double l1 = 0.0;
#pragma unroll
        for (int k = 0; k < VPL; ++k) if (k < c.nk) l1 += fabs((double)g[lane + kWave * k]);
        l1 = wave_sum(l1);
        double inv = 1.0 / l1;
        t = (1.0 < inv ? 1.0 : inv) * P.lr * 0.1;
      } else {
        t = P.lr;
      }
      double gtd = vdot<VPL>(c.nk, g, d);
      if (gtd > -P.tol_change) {
        converged = true;
        break;
      }
      strong_wolfe<VPL>(c, P, &t, d, &loss, gtd);
      if (isnan(loss)) break;
      float tf = (float)t;
      float smax = 0.f;
#pragma unroll
      for (int k = 0; k < VPL; ++k) if (k < c.nk) {
        x[lane + kWave * k] += d[lane + kWave * k] * tf;
        smax = fmaxf(smax, fabsf(d[lane + kWave * k] * tf));
      }
      updated = true;
      if (vamax<VPL>(c.nk, g) <= (float)P.tol_grad) {
        converged = true;
        break;
      }
      if (wave_max(smax) <= (float)P.tol_change) {
        converged = true;
        break;
      }
      if (fabs(loss - prev_loss) < P.tol_change) {
        converged = true;
        break;
      }
    }
    // optimizer.cpp:553-558: one final evaluation only after a step
    if (updated) {
#pragma unroll
      for (int k = 0; k < VPL; ++k) if (k < c.nk) d[lane + kWave * k] = 0.f;
      energy = eval_dir(c, 0.0, d, g);
    }
    c.n_iter = n_iter;
    c.hist_io = hist_io;
  }
  *energy_out = energy;
  *converged_out = converged;
}

// Occupancy target 3 waves per SIMD (168 VGPRs, a few spilled): measured
// 0.139 s against 0.145 s at the default 2 waves on the bench L-BFGS.
template <int VPL, int MODE>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(kLbfgsWaves))) void k_lbfgs(DevReceptor R, DevLigands L,
                                              const DevGrid* __restrict__ G, LbfgsArgs A,
                                              int vpl_lo) {
  __shared__ float vpool[V_COUNT * kWave * VPL];
  __shared__ __align__(16) float dofl[MODE ? PXD_DOF_LDS_FLOATS : 1];
  __shared__ int s_pose;
  const int lane = lane_id();
  constexpr int NV = kWave * VPL;
  float* hy = A.hist + (size_t)blockIdx.x * lbfgs_slot_floats(A.prm.history);
  float* hs = hy + (size_t)A.prm.history * NV;
  double* hr = reinterpret_cast<double*>(hs + (size_t)A.prm.history * NV);
  LbfgsCtx<VPL> c;
  c.R = &R;
  c.L = &L;
  c.G = G;
  // pose arrays in dynamic LDS, sized for the batch's largest ligand
  const PoseDims pdims = pose_dims(A.max_atoms, R.n_mov, R.n_leaves);
  c.s = carve_pose_lds(dyn_lds, pdims);
  // the two-loop's alphas (history + 1 doubles: one slot for masked steps)
  // after the pose arrays in the launch-sized LDS
  double* s_al = reinterpret_cast<double*>(dyn_lds + pdims.total);
  c.v = vpool;
  c.mode = MODE;
  if (MODE) {
    c.xyz0 = dofl;
    c.ap = carve_apply_lds(dofl + kMaxAtoms * 3);
    c.tq = dofl + kMaxAtoms * 3 + PXD_APPLY_LDS_FLOATS;
    c.dv = c.tq + kMaxFrags * 3;
  }
  float* x = c.vec(V_X);
  for (;;) {
    if (lane == 0) s_pose = atomicAdd(A.counter, 1);
    __syncthreads();
    const int qi = s_pose;
    __syncthreads();
    if (qi >= (A.order ? A.n_order : A.n_poses)) break;
    const int pi = A.order ? A.order[qi] : qi;
    c.li = L.info[A.pose_lig[pi]];
    c.n3 = MODE ? 6 + c.li.n_tors : 3 * c.li.n_atoms;
    c.n = c.n3 + R.n_leaves;
    c.nk = (c.n + kWave - 1) / kWave;
    // without a class list every launch scans all poses and takes its own
    if (!A.order && (c.nk <= vpl_lo || c.nk > VPL)) continue;
    load_ligand_params(L, c.li, G, c.s);
    c.nevals = 0;
    const long long off = 3ll * A.pose_off[pi];
    if (MODE) {
      // variables start at zero: the input pose about its centre
      for (int k = lane; k < 3 * c.li.n_atoms; k += kWave) c.xyz0[k] = A.xyz[off + k];
      __syncthreads();
      centroid(c.xyz0, c.li.n_atoms, c.c0);
    }
#pragma unroll
    for (int k = 0; k < VPL; ++k) {
      int e = lane + kWave * k;
      x[e] = e < c.n3 ? (MODE ? 0.f : A.xyz[off + e])
                      : (e < c.n ? A.tors[(long long)pi * R.n_leaves + (e - c.n3)] : 0.f);
    }
    double energy;
    bool converged;
    lbfgs_solve<VPL>(c, A.prm, hy, hs, hr, s_al, &energy, &converged);
    if (MODE) {  // the pose of the final variables
#pragma unroll
      for (int k = 0; k < VPL; ++k) {
        int e = lane + kWave * k;
        if (e < c.n3) c.dv[e] = x[e];
      }
      __syncthreads();
      float cen[3];
      dof_pose(c, cen);
      for (int k = lane; k < 3 * c.li.n_atoms; k += kWave) A.xyz[off + k] = c.s.xyz[k];
    }
#pragma unroll
    for (int k = 0; k < VPL; ++k) {
      int e = lane + kWave * k;
      if (e < c.n3) {
        if (!MODE) A.xyz[off + e] = x[e];
      } else if (e < c.n) {
        A.tors[(long long)pi * R.n_leaves + (e - c.n3)] = x[e];
      }
    }
    if (lane == 0) {
      A.energy[pi] = energy;
      A.nevals[pi] = c.nevals;
      A.converged[pi] = converged ? 1 : 0;
      if (A.stats) {
        // a history entry is n floats of s and of y plus its FP64 1 / (s.y)
        atomicAdd(&A.stats[0], (unsigned long long)c.n_iter);
        atomicAdd(&A.stats[1], (unsigned long long)c.nevals);
        atomicAdd(&A.stats[2], (unsigned long long)c.hist_io * (8ull * c.n + 8ull));
        atomicAdd(&A.stats[3], (unsigned long long)c.nevals * 256ull * c.li.n_atoms);
      }
    }
    __syncthreads();
  }
}

// Dynamic LDS of a k_lbfgs launch: the pose arrays for the batch's largest
// ligand, then the two-loop's alphas (history + 1 doubles).
static size_t lbfgs_lds_bytes(const DevReceptor& R, const LbfgsArgs& c) {
  return sizeof(float) * pose_dims(c.max_atoms, R.n_mov, R.n_leaves).total +
         sizeof(double) * (c.prm.history + 1);
}

void launch_lbfgs_class(const DevReceptor& R, const DevLigands& L, const DevGrid* G,
                        const LbfgsArgs& a, int cls, int n_slots, hipStream_t st) {
  LbfgsArgs c = a;
  c.max_atoms = a.max_atoms > 0 && a.max_atoms <= kMaxAtoms ? a.max_atoms : kMaxAtoms;
  const size_t lds = lbfgs_lds_bytes(R, c);
  static_assert(kLbfgsClasses == 3, "one template per class below");
  if (cls == 2)
    hipLaunchKernelGGL((k_lbfgs<kLbfgsClassVpl[2], 0>), dim3(n_slots), dim3(kWave), lds, st, R,
                       L, G, c, 0);
  else if (cls == 1)
    hipLaunchKernelGGL((k_lbfgs<kLbfgsClassVpl[1], 0>), dim3(n_slots), dim3(kWave), lds, st, R,
                       L, G, c, 0);
  else
    hipLaunchKernelGGL((k_lbfgs<kLbfgsClassVpl[0], 0>), dim3(n_slots), dim3(kWave), lds, st, R,
                       L, G, c, 0);
}

void launch_lbfgs(const DevReceptor& R, const DevLigands& L, const DevGrid* G,
                  const LbfgsArgs& a, int n_slots, hipStream_t st) {
  LbfgsArgs c = a;
  c.max_atoms = a.max_atoms > 0 && a.max_atoms <= kMaxAtoms ? a.max_atoms : kMaxAtoms;
  const size_t lds = lbfgs_lds_bytes(R, c);
  // rigid-body + torsion variables: n <= 6 + 31 + 64 always fits width 2
  if (a.prm.dof_mode) {
    hipLaunchKernelGGL((k_lbfgs<2, 1>), dim3(n_slots), dim3(kWave), lds, st, R, L, G, c, 0);
    return;
  }
  static_assert(kLbfgsClasses == 3, "one launch per class below");
  hipLaunchKernelGGL((k_lbfgs<kLbfgsClassVpl[2], 0>), dim3(n_slots), dim3(kWave), lds, st, R, L,
                     G, c, kLbfgsClassVpl[1]);
  c.counter = a.counter + 1;
  hipLaunchKernelGGL((k_lbfgs<kLbfgsClassVpl[1], 0>), dim3(n_slots), dim3(kWave), lds, st, R, L,
                     G, c, kLbfgsClassVpl[0]);
  c.counter = a.counter + 2;
  hipLaunchKernelGGL((k_lbfgs<kLbfgsClassVpl[0], 0>), dim3(n_slots), dim3(kWave), lds, st, R, L,
                     G, c, 0);
}

// Ring centroids / normals of the pscore leaves in dynamic LDS, sized by the
// receptor's ring count at launch (ring_lds_bytes) rather than the maximum.

__device__ inline PscoreLds carve_ring_lds(float* base, int n_rec_rings) {
  PscoreLds ps;
  ps.rc = base;
  ps.rn = base + 3 * n_rec_rings;
  ps.lc = base + 6 * n_rec_rings;
  ps.ln = ps.lc + 3 * kMaxLigRings;
  return ps;
}

static size_t ring_lds_bytes(const DevReceptor& R) {
  return sizeof(float) * 6 * (size_t)(R.n_ring5 + R.n_ring6 + kMaxLigRings);
}

// ============================================================== K5: pscore
// root_scorer::combine for a batch of poses, one wavefront per pose.

__global__ __launch_bounds__(64) void k_pscore(DevReceptor R, DevLigands L, pxd_score_params sp,
                                               int n_poses, const int* __restrict__ pose_lig,
                                               const int* __restrict__ pose_off,
                                               const float* __restrict__ xyz,
                                               const float* __restrict__ tors,
                                               double* __restrict__ score,
                                               double* __restrict__ terms, int max_atoms) {
  __shared__ double s_terms[PXD_NUM_TERMS];
  // dynamic LDS: the pose arrays (no gradient area) sized for the batch's
  // largest ligand, then the ring geometry
  const PoseDims pd = pose_dims(max_atoms, R.n_mov, R.n_leaves);
  PoseLds s = carve_pose_lds(dyn_lds, pd);
  PscoreLds ps;
  ps = carve_ring_lds(dyn_lds + pd.grad, R.n_ring5 + R.n_ring6);
  ps.terms = s_terms;
  const int p = xcd_block(blockIdx.x, gridDim.x);
  if (p >= n_poses) return;
  const int lane = lane_id();
  const LigInfo li = L.info[pose_lig[p]];
  load_ligand_params(L, li, nullptr, s);
  const long long off = 3ll * pose_off[p];
  for (int k = lane; k < li.n_atoms * 3; k += kWave) s.xyz[k] = xyz[off + k];
  for (int k = lane; k < R.n_leaves; k += kWave)
    s.tors[k] = tors ? tors[(long long)p * R.n_leaves + k] : 0.f;
  __syncthreads();
  receptor_torsion2xyz(R, s);
  __syncthreads();
  double v = pscore_pose(R, L, li, sp, s, ps);
  if (lane == 0) score[p] = v;
  if (terms && lane < PXD_NUM_TERMS) terms[(long long)p * PXD_NUM_TERMS + lane] = s_terms[lane];
}

void launch_pscore(const DevReceptor& R, const DevLigands& L, const pxd_score_params& sp,
                   const PscoreArgs& a, hipStream_t st) {
  if (a.n_poses <= 0) return;
  const int ma = a.max_atoms > 0 && a.max_atoms <= kMaxAtoms ? a.max_atoms : kMaxAtoms;
  const size_t lds = sizeof(float) * pose_dims(ma, R.n_mov, R.n_leaves).grad + ring_lds_bytes(R);
  hipLaunchKernelGGL(k_pscore, dim3(a.n_poses), dim3(kWave), lds, st, R, L, sp, a.n_poses,
                     a.pose_lig, a.pose_off, a.xyz, a.tors, a.score, a.terms, ma);
}

// pscore with its derivatives (pxd_pscore_grad.hpp), one wavefront per pose.
// Occupancy target 5 waves per SIMD (96 VGPRs, a few spilled): 6.7 ms
// against 7.9 ms at the default 4 waves on the training batch.
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(5))) void k_pscore_grad(DevReceptor R, DevLigands L,
                                                    pxd_score_params sp, int n_poses,
                                                    const int* __restrict__ pose_lig,
                                                    const int* __restrict__ pose_off,
                                                    const float* __restrict__ xyz,
                                                    const float* __restrict__ tors,
                                                    double* __restrict__ score,
                                                    float* __restrict__ dxyz,
                                                    double* __restrict__ dparams, int max_atoms) {
  __shared__ float s_gx[kMaxAtoms * 3];
  __shared__ double s_gp[PXD_NUM_SCORE_PARAMS];
  const PoseDims pd = pose_dims(max_atoms, R.n_mov, R.n_leaves);
  PoseLds s = carve_pose_lds(dyn_lds, pd);
  PscoreLds ps;
  ps = carve_ring_lds(dyn_lds + pd.grad, R.n_ring5 + R.n_ring6);
  ps.terms = nullptr;
  PscoreGradLds g{s_gx, s_gp};
  const int p = xcd_block(blockIdx.x, gridDim.x);
  if (p >= n_poses) return;
  const int lane = lane_id();
  const LigInfo li = L.info[pose_lig[p]];
  load_ligand_params(L, li, nullptr, s);
  const long long off = 3ll * pose_off[p];
  for (int k = lane; k < li.n_atoms * 3; k += kWave) s.xyz[k] = xyz[off + k];
  for (int k = lane; k < R.n_leaves; k += kWave)
    s.tors[k] = tors ? tors[(long long)p * R.n_leaves + k] : 0.f;
  __syncthreads();
  receptor_torsion2xyz(R, s);
  __syncthreads();
  double v = pscore_pose_grad(R, L, li, sp, s, ps, g);
  if (lane == 0) score[p] = v;
  if (dxyz)
    for (int k = lane; k < li.n_atoms * 3; k += kWave) dxyz[off + k] = s_gx[k];
  if (dparams)
    for (int k = lane; k < PXD_NUM_SCORE_PARAMS; k += kWave)
      dparams[(long long)p * PXD_NUM_SCORE_PARAMS + k] = s_gp[k];
}

void launch_pscore_grad(const DevReceptor& R, const DevLigands& L, const pxd_score_params& sp,
                        const PscoreArgs& a, float* dxyz, double* dparams, hipStream_t st) {
  if (a.n_poses <= 0) return;
  const int ma = a.max_atoms > 0 && a.max_atoms <= kMaxAtoms ? a.max_atoms : kMaxAtoms;
  const size_t lds = sizeof(float) * pose_dims(ma, R.n_mov, R.n_leaves).grad + ring_lds_bytes(R);
  hipLaunchKernelGGL(k_pscore_grad, dim3(a.n_poses), dim3(kWave), lds, st, R, L, sp, a.n_poses,
                     a.pose_lig, a.pose_off, a.xyz, a.tors, a.score, dxyz, dparams, ma);
}

// ============================================================== K4: Monte-Carlo search
// monte_carlo_searcher::mutate_and_optimize (searcher.cpp:98-187) for many
// walkers, one wavefront per walker from a device work queue.  Walker state
// (current and best conformations, step, RNG counter) stays in HBM between
// launches; a launch advances every walker to step_end.  Each iteration of
// the walker loop makes exactly one L-BFGS run and one pscore evaluation,
// selected by the walker's phase (init / relax / refine).

enum { PH_INIT = 0, PH_MUTATE = 1, PH_REFINE = 2 };

// Dynamic LDS of k_mc (floats), sized at launch for the batch's largest
// ligand: pose arrays, ring geometry, apply_parameters scratch, then the
// walker's current conformation (coordinates, then leaf torsions at cur_t).
struct McDims {
  int ring, apply, cur, cur_t, total;
};

__host__ __device__ constexpr McDims mc_dims(int ma, int mf, int nm, int nlv, int n_rings) {
  auto r4 = [](int x) { return (x + 3) & ~3; };
  McDims d{};
  d.ring = pose_dims(ma, nm, nlv, 4).total;
  d.apply = d.ring + r4(6 * (n_rings + kMaxLigRings));
  d.cur = d.apply + r4(apply_lds_floats(ma, mf));
  d.cur_t = d.cur + r4(3 * ma);
  d.total = d.cur_t + r4(nlv);
  return d;
}

template <int VPL>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(kMcWaves))) void k_mc(
    DevReceptor R, DevLigands L, const DevGrid* __restrict__ G, McArgs A, int vpl_lo) {
  __shared__ float vpool[V_COUNT * kWave * VPL];
  __shared__ double s_al[kMaxHistory + 1];  // + one slot for masked steps
  __shared__ double s_terms[PXD_NUM_TERMS];
  __shared__ int s_w;
  __shared__ double s_emin;
  const int lane = lane_id();
  constexpr int NV = kWave * VPL;
  const McParamsDev& P = A.prm;
  const pxd_score_params& sp = *A.sp;
  const int nl = R.n_leaves;
  float* hy = A.hist + (size_t)blockIdx.x * lbfgs_slot_floats(A.hist_m);
  float* hs = hy + (size_t)A.hist_m * NV;
  double* hr = reinterpret_cast<double*>(hs + (size_t)A.hist_m * NV);
  LbfgsCtx<VPL, 4> c;  // 16-byte stage slots: one LDS access each (registers, not LDS, bound this kernel)
  c.R = &R;
  c.L = &L;
  c.G = G;
  const McDims md = mc_dims(A.max_atoms, A.max_frags, R.n_mov, nl, R.n_ring5 + R.n_ring6);
  c.s = carve_pose_lds(dyn_lds, pose_dims(A.max_atoms, R.n_mov, nl, 4));
  c.v = vpool;
  c.mode = 0;
  PscoreLds ps;
  ps = carve_ring_lds(dyn_lds + md.ring, R.n_ring5 + R.n_ring6);
  ps.terms = s_terms;
  ApplyLds ap = carve_apply_lds(dyn_lds + md.apply, A.max_atoms, A.max_frags);
  float* curx = dyn_lds + md.cur;
  const int ct = md.cur_t - md.cur;  // leaf torsions follow the coordinates
  BoxF box;
  for (int d = 0; d < 3; ++d) {
    box.lo[d] = A.box_lo[d];
    box.hi[d] = A.box_hi[d];
  }
  float* x = c.vec(V_X);
  float* curt = curx + ct;
  for (;;) {
    if (lane == 0) s_w = atomicAdd(A.counter, 1);
    __syncthreads();
    const int qi = s_w;
    __syncthreads();
    if (qi >= (A.order ? A.n_order : A.n_walkers)) break;
    const int w = A.order ? A.order[qi] : qi;
    c.li = L.info[A.walker_lig[w]];
    const LigInfo& li = c.li;
    const int n = li.n_atoms, n3 = 3 * n;
    c.n3 = n3;
    c.n = n3 + nl;
    c.nk = (c.n + kWave - 1) / kWave;
    // without a class list every launch scans all walkers and takes its own
    if (!A.order && (c.nk <= vpl_lo || c.nk > VPL)) continue;
    WalkerState st = A.st[w];
    if (st.mode == 2) continue;
    const int pid = A.walker_pid[w];
    load_ligand_params(L, li, G, c.s);
    Rng rng{P.seed_lo, P.seed_hi + (uint32_t)pid, (uint32_t)A.walker_lig[w], st.rng_ctr};
    const long long off = 3ll * A.walker_off[w];
    for (int k = lane; k < n3; k += kWave) curx[k] = A.cur_xyz[off + k];
    for (int k = lane; k < nl; k += kWave) curt[k] = A.cur_tors[(long long)w * nl + k];
    int phase;
    if (st.mode == 0) {
      // initial conformer (searcher.cpp:105-115): walker 0 keeps conformer
      // 0; walker pid starts from its own pre-generated conformer pid when
      // there is one (keeping its torsions), else from conformer 0 with
      // random torsions; both are placed at random (searcher.cpp:46-87)
      const int lig = A.walker_lig[w];
      const int src = pid > 0 && pid < A.lig_conf_n[lig] ? pid : 0;
      const float* x0 = A.lig_xyz0 + 3ll * (A.lig_conf_off[lig] + (long long)src * n);
      for (int k = lane; k < n3; k += kWave) c.s.xyz[k] = x0[k];
      __syncthreads();
      if (pid != 0) {
        for (int tries = 1;; ++tries) {
          float cen[3], rot[9];
          cen[0] = (float)rng.uniform(box.lo[0], box.hi[0]);
          cen[1] = (float)rng.uniform(box.lo[1], box.hi[1]);
          cen[2] = (float)rng.uniform(box.lo[2], box.hi[2]);
          double q[4];
          rng.quaternion(q);
          for (int f = lane; f < li.n_frags; f += kWave) ap.th[f] = 0.f;
          __syncthreads();
          // torsion draws are wave-uniform: every lane draws, lane 0 stores
          for (int t = 0; src == 0 && t < li.n_tors; ++t) {
            float v = (float)rng.uniform(-3.141592653589793, 3.141592653589793);
            if (lane == 0) ap.th[L.tors_frag[li.frag_off + t]] = v;
          }
          quat_to_rot(q, rot);
          for (int k = lane; k < n3; k += kWave) c.s.xyz[k] = x0[k];
          __syncthreads();
          apply_ligand(L, li, c.s.xyz, cen, rot, ap);
          bool inbox = all_in_box(c.s.xyz, n, box);
          if (tries < P.max_tries && !inbox && tries < 4 * P.max_tries) continue;
          float cc[3];
          centroid(c.s.xyz, n, cc);
          if (!point_in_box(cc, box) && tries < 4 * P.max_tries) continue;
          if (tries >= P.max_tries) break;
          if (pscore_vdw_initial(R, L, li, sp, c.s) < 100.0) break;
        }
      }
      for (int k = lane; k < n3; k += kWave) x[k] = c.s.xyz[k];
      for (int k = lane; k < nl; k += kWave) x[n3 + k] = 0.f;
      phase = PH_INIT;
    } else {
      phase = PH_MUTATE;
    }
    const int step_end = A.step_end < P.max_nsteps ? A.step_end : P.max_nsteps;
    // the energy prune (searcher.cpp:141-157) evaluates the mutated pose
    // before its relax: that is the L-BFGS site below asked for its initial
    // evaluation only (not counted in nevals, as the reference's), so the
    // kernel keeps a single inlined evaluation
    bool probe = false, mutated = false;
    for (;;) {
      if (phase == PH_MUTATE && mutated) {
        mutated = false;  // the pose passed the prune: relax it
      } else if (phase == PH_MUTATE) {
        if (st.step >= step_end) break;
        // mutate_inplace (searcher.cpp:189-223) on a copy of the current state
        for (int k = lane; k < n3; k += kWave) c.s.xyz[k] = curx[k];
        for (int f = lane; f < li.n_frags; f += kWave) ap.th[f] = 0.f;
        __syncthreads();
        float cen[3], rot[9] = {1.f, 0.f, 0.f, 0.f, 1.f, 0.f, 0.f, 0.f, 1.f};
        centroid(c.s.xyz, n, cen);
        bool apply = true;
        if (P.mutation == 0) {
          int which = rng.uniform_int(0, li.n_tors + 1);
          if (which == 0) {
            double dl[3];
            rng.in_sphere(dl);
            for (int d = 0; d < 3; ++d) cen[d] += (float)(dl[d] * P.amplitude);
          } else if (which == 1) {
            float gr = 0.f;
            for (int a = lane; a < n; a += kWave)
              for (int d = 0; d < 3; ++d) {
                float v = c.s.xyz[3 * a + d] - cen[d];
                gr += v * v;
              }
            gr = sqrtf(wave_sum(gr) / (float)n);
            if (gr < 1e-6f) {
              apply = false;  // searcher.cpp:207: mutation skipped
            } else {
              double dl[3];
              rng.in_sphere(dl);
              double sc = P.amplitude / gr;
              double v[3] = {dl[0] * sc, dl[1] * sc, dl[2] * sc};
              double ang = sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
              rot_axis((float)(v[0] / ang), (float)(v[1] / ang), (float)(v[2] / ang), (float)ang, rot);
            }
          } else {
            float v = (float)rng.uniform(-3.141592653589793, 3.141592653589793);
            if (lane == 0) ap.th[L.tors_frag[li.frag_off + which - 2]] = v;
          }
        } else {
          for (int d = 0; d < 3; ++d) cen[d] += (float)(rng.normal() * P.trans_sigma);
          double v[3] = {rng.normal() * P.rot_sigma, rng.normal() * P.rot_sigma,
                         rng.normal() * P.rot_sigma};
          double ang = sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
          if (ang > 1e-12)
            rot_axis((float)(v[0] / ang), (float)(v[1] / ang), (float)(v[2] / ang), (float)ang, rot);
          if (li.n_tors > 0) {
            int t = rng.uniform_int(0, li.n_tors - 1);
            float th = (float)rng.uniform(-3.141592653589793, 3.141592653589793);
            if (lane == 0) ap.th[L.tors_frag[li.frag_off + t]] = th;
          }
        }
        __syncthreads();
        if (apply) apply_ligand(L, li, c.s.xyz, cen, rot, ap);
        for (int k = lane; k < n3; k += kWave) x[k] = c.s.xyz[k];
        for (int k = lane; k < nl; k += kWave) x[n3 + k] = curt[k];
        probe = P.mm_threshold > 0.0;
      } else if (phase == PH_REFINE) {
        for (int k = lane; k < n3; k += kWave) x[k] = curx[k];
        for (int k = lane; k < nl; k += kWave) x[n3 + k] = curt[k];
      }
      __syncthreads();
      // the single L-BFGS site: full step for init / refine, relax after a
      // mutation, no iteration for the prune's evaluation
      c.nevals = 0;
      double e;
      bool conv;
      lbfgs_solve<VPL>(c, phase == PH_MUTATE ? P.relax : P.full, hy, hs, hr, s_al, &e, &conv,
                       probe);
      if (probe) {
        // the mutated pose's energy updates the ligand's lowest
        // pre-optimisation energy, shared by its walkers (starts at 1e8); a
        // pose above it by more than the threshold skips this step: no
        // relax and no Metropolis draw
        probe = false;
        if (lane == 0) {
          unsigned long long* slot = A.lig_emin + A.walker_lig[w];
          unsigned long long old = e == e ? atomicMin(slot, order_key(e)) : *slot;
          double m = order_key_value(old);
          s_emin = e < m ? e : m;
        }
        __syncthreads();
        const double emin = s_emin;
        __syncthreads();
        if (e > emin + P.mm_threshold) {
          st.step++;
        } else {
          mutated = true;  // x still holds the mutated pose
        }
        continue;
      }
      // the single pscore site, at the optimised pose
      for (int k = lane; k < n3; k += kWave) c.s.xyz[k] = x[k];
      for (int k = lane; k < nl; k += kWave) c.s.tors[k] = x[n3 + k];
      __syncthreads();
      receptor_torsion2xyz(R, c.s);
      __syncthreads();
      double psv = pscore_pose(R, L, li, sp, c.s, ps);
      if (phase == PH_INIT) {
        st.nevals = c.nevals;
        for (int k = lane; k < c.n; k += kWave) curx[k < n3 ? k : ct + k - n3] = x[k];
        st.cur_e = st.best_e = e;
        st.cur_ps = st.best_ps = psv;
        for (int k = lane; k < n3; k += kWave) A.best_xyz[off + k] = x[k];
        for (int k = lane; k < nl; k += kWave) A.best_tors[(long long)w * nl + k] = x[n3 + k];
        __syncthreads();
        if (pid == 0) {  // searcher.cpp:123: walker 0 does no Monte-Carlo
          st.mode = 2;
          break;
        }
        if (!all_in_box(curx, n, box)) {  // searcher.cpp:178-185
          if (!coords_sane(curx, n)) st.nevals = 0;
          st.mode = 2;
          break;
        }
        st.mode = 1;
        st.step = 0;
        phase = PH_MUTATE;
      } else if (phase == PH_MUTATE) {
        st.nevals += c.nevals;
        // Metropolis on pscore (searcher.cpp:164-165, searcher.h:53-56)
        bool acc = st.step == 0;
        if (!acc) {
          if (psv < st.cur_ps)
            acc = true;
          else
            acc = rng.u01() < exp((st.cur_ps - psv) / (double)P.temperature);
        }
        bool inbox = all_in_box(x, n, box);
        if (acc && inbox) {
          for (int k = lane; k < c.n; k += kWave) curx[k < n3 ? k : ct + k - n3] = x[k];
          st.cur_e = e;
          st.cur_ps = psv;
          __syncthreads();
          if (psv < st.best_ps) {
            phase = PH_REFINE;
            continue;
          }
        }
        st.step++;
      } else {  // PH_REFINE (searcher.cpp:167-174)
        st.nevals += c.nevals;
        for (int k = lane; k < c.n; k += kWave) curx[k < n3 ? k : ct + k - n3] = x[k];
        st.cur_e = e;
        st.cur_ps = psv;
        if (psv < st.best_ps) {
          st.best_ps = psv;
          st.best_e = e;
          for (int k = lane; k < n3; k += kWave) A.best_xyz[off + k] = x[k];
          for (int k = lane; k < nl; k += kWave) A.best_tors[(long long)w * nl + k] = x[n3 + k];
        }
        __syncthreads();
        st.step++;
        phase = PH_MUTATE;
      }
    }
    if (st.mode == 1 && st.step >= P.max_nsteps) st.mode = 2;
    st.rng_ctr = rng.ctr;
    for (int k = lane; k < n3; k += kWave) A.cur_xyz[off + k] = curx[k];
    for (int k = lane; k < nl; k += kWave) A.cur_tors[(long long)w * nl + k] = curt[k];
    if (lane == 0) A.st[w] = st;
    __syncthreads();
  }
}

void launch_mc_class(const DevReceptor& R, const DevLigands& L, const DevGrid* G,
                     const McArgs& a, int cls, int n_slots, hipStream_t st) {
  McArgs c = a;
  c.max_atoms = a.max_atoms > 0 && a.max_atoms <= kMaxAtoms ? a.max_atoms : kMaxAtoms;
  c.max_frags = a.max_frags > 0 && a.max_frags <= kMaxFrags ? a.max_frags : kMaxFrags;
  const size_t dyn = sizeof(float) * mc_dims(c.max_atoms, c.max_frags, R.n_mov, R.n_leaves,
                                             R.n_ring5 + R.n_ring6).total;
  static_assert(kLbfgsClasses == 3, "one template per class below");
  if (cls == 2)
    hipLaunchKernelGGL(k_mc<kLbfgsClassVpl[2]>, dim3(n_slots), dim3(kWave), dyn, st, R, L, G, c, 0);
  else if (cls == 1)
    hipLaunchKernelGGL(k_mc<kLbfgsClassVpl[1]>, dim3(n_slots), dim3(kWave), dyn, st, R, L, G, c, 0);
  else
    hipLaunchKernelGGL(k_mc<kLbfgsClassVpl[0]>, dim3(n_slots), dim3(kWave), dyn, st, R, L, G, c, 0);
}

void launch_mc(const DevReceptor& R, const DevLigands& L, const DevGrid* G, const McArgs& a,
               int n_slots, hipStream_t st) {
  // one launch per L-BFGS size class (the walker's ligand decides), each
  // with its own queue head: a.counter[0..kLbfgsClasses)
  McArgs c = a;
  c.max_atoms = a.max_atoms > 0 && a.max_atoms <= kMaxAtoms ? a.max_atoms : kMaxAtoms;
  c.max_frags = a.max_frags > 0 && a.max_frags <= kMaxFrags ? a.max_frags : kMaxFrags;
  const size_t dyn = sizeof(float) * mc_dims(c.max_atoms, c.max_frags, R.n_mov, R.n_leaves,
                                             R.n_ring5 + R.n_ring6).total;
  hipLaunchKernelGGL(k_mc<kLbfgsClassVpl[2]>, dim3(n_slots), dim3(kWave), dyn, st, R, L, G, c,
                     kLbfgsClassVpl[1]);
  c.counter = a.counter + 1;
  hipLaunchKernelGGL(k_mc<kLbfgsClassVpl[1]>, dim3(n_slots), dim3(kWave), dyn, st, R, L, G, c,
                     kLbfgsClassVpl[0]);
  c.counter = a.counter + 2;
  hipLaunchKernelGGL(k_mc<kLbfgsClassVpl[0]>, dim3(n_slots), dim3(kWave), dyn, st, R, L, G, c, 0);
}

__global__ void k_fill_u64(unsigned long long* p, int n, unsigned long long v) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

void launch_fill_u64(unsigned long long* p, int n, unsigned long long v, hipStream_t st) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_fill_u64, dim3((n + 255) / 256), dim3(256), 0, st, p, n, v);
}

__global__ void k_fill_int(int* p, int n, int v) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

void launch_fill_int(int* p, int n, int v, hipStream_t st) {
  hipLaunchKernelGGL(k_fill_int, dim3((n + 255) / 256), dim3(256), 0, st, p, n, v);
}

// ============================================================== K6: top-k clustering
// pose_cluster::preprocess (writer.cpp:124-177) per ligand, one wave per ligand;
// RMSDs (writer.cpp:99-106) reduced across lanes.

__global__ __launch_bounds__(64) void k_cluster(ClusterArgs A) {
  const int l = blockIdx.x;
  if (l >= A.n_ligands) return;
  const int lane = lane_id();
  __shared__ int kept[64];
  int cnt = 0;
  const int w0 = A.lig_walker_off[l], w1 = A.lig_walker_off[l + 1];
  for (int w = w0; w < w1; ++w) {
    const WalkerState st = A.st[w];
    if (st.nevals == 0 || st.mode == 0) continue;  // failure mark (writer.cpp:135)
    if (cnt == 0) {
      if (lane == 0) kept[0] = w;
      cnt = 1;
      __syncthreads();
      continue;
    }
    int best = 0;
    float rbest = 0.f;
    const float* xw = A.best_xyz + 3ll * A.walker_off[w];
    for (int k = 0; k < cnt; ++k) {
      const float* xk = A.best_xyz + 3ll * A.walker_off[kept[k]];
      int na = A.walker_off[w + 1] - A.walker_off[w];
      float acc = 0.f;
      for (int i = lane; i < 3 * na; i += kWave) {
        float dv = xk[i] - xw[i];
        acc += dv * dv;
      }
      float r = sqrtf(wave_sum(acc) / (float)na);
      if (k == 0 || r < rbest) {
        best = k;
        rbest = r;
      }
    }
    bool updated = false;
    double ps = st.best_ps;
    if (rbest < A.min_rmsd) {
      if (ps < A.st[kept[best]].best_ps) {
        if (lane == 0) kept[best] = w;
        updated = true;
      }
    } else if (cnt < A.num_modes) {
      if (lane == 0) kept[cnt] = w;
      cnt++;
      updated = true;
    } else if (ps < A.st[kept[cnt - 1]].best_ps) {
      if (lane == 0) kept[cnt - 1] = w;
      updated = true;
    }
    __syncthreads();
    if (updated && lane == 0) {  // stable insertion sort by pscore
      for (int a = 1; a < cnt; ++a) {
        int v = kept[a], b = a - 1;
        while (b >= 0 && A.st[kept[b]].best_ps > A.st[v].best_ps) {
          kept[b + 1] = kept[b];
          --b;
        }
        kept[b + 1] = v;
      }
    }
    __syncthreads();
  }
  for (int k = lane; k < A.num_modes; k += kWave)
    A.out_idx[l * A.num_modes + k] = k < cnt ? kept[k] - w0 : -1;
  if (lane == 0) A.out_count[l] = cnt;
}

void launch_cluster(const ClusterArgs& a, int n_atoms_max, hipStream_t st) {
  (void)n_atoms_max;
  hipLaunchKernelGGL(k_cluster, dim3(a.n_ligands), dim3(kWave), 0, st, a);
}

// ============================================================== pose parameterisation
// free_ligand::apply_parameters for a batch of poses (one wave per pose):
// new centre, orientation matrix and torsion rotations applied to the input
// conformation of each pose.

__global__ __launch_bounds__(64) void k_apply(DevLigands L, ApplyArgs A) {
  __shared__ float xyz[kMaxAtoms * 3];
  __shared__ float apl[PXD_APPLY_LDS_FLOATS];
  const int p = blockIdx.x;
  if (p >= A.n_poses) return;
  const int lane = lane_id();
  const LigInfo li = L.info[A.pose_lig[p]];
  ApplyLds ap = carve_apply_lds(apl);
  const long long off = 3ll * A.pose_off[p];
  for (int k = lane; k < 3 * li.n_atoms; k += kWave) xyz[k] = A.xyz_in[off + k];
  for (int f = lane; f < li.n_frags; f += kWave) ap.th[f] = 0.f;
  __syncthreads();
  const int t0 = A.tors_off[p], t1 = A.tors_off[p + 1];
  for (int t = lane; t < t1 - t0 && t < li.n_tors; t += kWave)
    ap.th[L.tors_frag[li.frag_off + t]] = A.tors[t0 + t];
  float cen[3] = {A.center[3 * p], A.center[3 * p + 1], A.center[3 * p + 2]};
  float rot[9];
  for (int k = 0; k < 9; ++k) rot[k] = A.rot[9 * p + k];
  __syncthreads();
  apply_ligand(L, li, xyz, cen, rot, ap);
  for (int k = lane; k < 3 * li.n_atoms; k += kWave) A.xyz_out[off + k] = xyz[k];
}

void launch_apply(const DevLigands& L, const ApplyArgs& a, hipStream_t st) {
  if (a.n_poses <= 0) return;
  hipLaunchKernelGGL(k_apply, dim3(a.n_poses), dim3(kWave), 0, st, L, a);
}

// ============================================================== receptor leaf tables
// One wave per (leaf, sample): T_k and dT_k/dtheta at theta = 2*pi*(i + shift)/n
// (shift 0: table nodes, 0.5: midpoints used to measure the interpolation error).

__global__ __launch_bounds__(64) void k_leaf_table(DevReceptor R, int n, double shift,
                                                   double2* out) {
  const int k = blockIdx.y, i = blockIdx.x;
  if (k >= R.n_leaves || i >= n) return;
  double e, d;
  leaf_energy(R, k, 6.283185307179586 * ((double)i + shift) / n, &e, &d);
  if (lane_id() == 0) out[(long long)k * n + i] = make_double2(e, d);
}

void launch_leaf_table(const DevReceptor& R, int n, double shift, double2* out,
                       hipStream_t st) {
  if (R.n_leaves <= 0 || n <= 0) return;
  hipLaunchKernelGGL(k_leaf_table, dim3(n, R.n_leaves), dim3(kWave), 0, st, R, n, shift, out);
}

// ============================================================== rigid-body + torsion gradient
// One wave per pose: force, torque about the pose's geometric centre and the
// torsion torques from the Cartesian gradient (dof_grad).

__global__ __launch_bounds__(64) void k_dof_grad(DevLigands L, int n_poses,
                                                 const int* __restrict__ pose_lig,
                                                 const int* __restrict__ pose_off,
                                                 const float* __restrict__ xyz,
                                                 const float* __restrict__ grad,
                                                 float* __restrict__ out) {
  __shared__ float sx[kMaxAtoms * 3];
  __shared__ float sg[kMaxAtoms * 3];
  __shared__ float tq[kMaxFrags * 3];
  __shared__ float dv[6 + kMaxFrags];
  const int p = blockIdx.x;
  if (p >= n_poses) return;
  const int lane = lane_id();
  const LigInfo li = L.info[pose_lig[p]];
  const long long off = 3ll * pose_off[p];
  for (int k = lane; k < 3 * li.n_atoms; k += kWave) {
    sx[k] = xyz[off + k];
    sg[k] = grad[off + k];
  }
  __syncthreads();
  float cen[3];
  centroid(sx, li.n_atoms, cen);
  dof_grad(L, li, sx, sg, cen, tq, dv);
  constexpr int W = 6 + kMaxFrags - 1;
  for (int k = lane; k < W; k += kWave) out[(long long)p * W + k] = k < 6 + li.n_tors ? dv[k] : 0.f;
}

void launch_dof_grad(const DevLigands& L, int n_poses, const int* pose_lig, const int* pose_off,
                     const float* xyz, const float* grad, float* out, hipStream_t st) {
  if (n_poses <= 0) return;
  hipLaunchKernelGGL(k_dof_grad, dim3(n_poses), dim3(kWave), 0, st, L, n_poses, pose_lig,
                     pose_off, xyz, grad, out);
}

// ============================================================== map layout 1
// Copies every cell's 8 corner points of every map next to each other:
// cm[(m*C + cell)*8 + c] = maps[m*P + corner c of cell], c = x | y<<1 | z<<2.
__global__ __launch_bounds__(256) void k_expand_cells(const float4* __restrict__ maps, long long P,
                                                      int nx, int ny, int nz, int n_maps,
                                                      float4* __restrict__ cm) {
  const long long C = (long long)(nx - 1) * (ny - 1) * (nz - 1);
  long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= C * 8 * n_maps) return;
  const int c = (int)(t & 7);
  const long long cell = (t >> 3) % C;
  const int m = (int)((t >> 3) / C);
  const int k = (int)(cell % (nz - 1));
  const int j = (int)((cell / (nz - 1)) % (ny - 1));
  const int i = (int)(cell / ((long long)(nz - 1) * (ny - 1)));
  const long long p = ((long long)(i + (c & 1)) * ny + (j + ((c >> 1) & 1))) * nz + k + (c >> 2);
  cm[t] = maps[(long long)m * P + p];
}

// ============================================================== map layout 2
// Splits every map's float4 points into four planes:
// sp[(m*4 + c)*P + p] = component c of maps[m*P + p].
__global__ __launch_bounds__(256) void k_split_planes(const float4* __restrict__ maps, long long P,
                                                      int n_maps, float* __restrict__ sp) {
  long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= P * n_maps) return;
  const long long m = t / P, p = t - m * P;
  const float4 v = maps[t];
  float* o = sp + m * 4 * P + p;
  o[0] = v.x;
  o[P] = v.y;
  o[2 * P] = v.z;
  o[3 * P] = v.w;
}

void launch_split_planes(const DevGrid& G, int n_maps, float* sp, hipStream_t st) {
  const long long total = G.P * n_maps;
  hipLaunchKernelGGL(k_split_planes, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, st,
                     G.maps, G.P, n_maps, sp);
}

void launch_expand_cells(const DevGrid& G, int n_maps, float4* cm, hipStream_t st) {
  const long long C = (long long)(G.np[0] - 1) * (G.np[1] - 1) * (G.np[2] - 1);
  const long long total = C * 8 * n_maps;
  hipLaunchKernelGGL(k_expand_cells, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, st,
                     G.maps, G.P, G.np[0], G.np[1], G.np[2], n_maps, cm);
}

}  // namespace pxd
