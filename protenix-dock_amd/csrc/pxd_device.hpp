// pxd_device.hpp — device functions of the docking hot path (gfx950, wave64).
//
// One wavefront evaluates one pose: lanes run in parallel over atoms and
// over the term lists of the pose's ligand; per-pose coordinates, gradients
// and parameters sit in LDS; energies accumulate per lane in FP64 and are
// reduced across the wave.  Every function names the reference code it
// restates (paths relative to the Protenix-Dock tree).
#pragma once

#include <hip/hip_runtime.h>

#include "pxd_internal.hpp"

namespace pxd {

// ---------------------------------------------------------------- reductions

// Wave-wide sums with DPP lane moves (no LDS traffic): pairs, quads, the
// half-row and row mirrors give every lane its 16-lane row sum; the row
// broadcasts of lanes 15 and 31 then accumulate the four rows into lane 63,
// which is read back as a uniform value.  GFX9 DPP controls: quad_perm 0xB1
// (xor 1) and 0x4E (xor 2), row_half_mirror 0x141, row_mirror 0x140,
// row_bcast:15 0x142 (rows 1 and 3), row_bcast:31 0x143 (rows 2 and 3).
// With all rows enabled every lane is written from a valid source lane, so
// bound_ctrl changes nothing; setting it lets the compiler fold the lane move
// into the add (one v_add_f32_dpp per level instead of a move and an add).
// With rows masked off the disabled rows keep old = 0 (update_dpp), so the
// following add leaves them unchanged.
template <int CTRL, int ROWS>
__device__ inline float dpp_f(float v) {
  if (ROWS == 0xF)
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, 0xF, 0xF, true));
  return __int_as_float(
      __builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, ROWS, 0xF, false));
}

template <int CTRL, int ROWS>
__device__ inline double dpp_d(double v) {
  const long long b = __double_as_longlong(v);
  int lo, hi;
  if (ROWS == 0xF) {
    lo = __builtin_amdgcn_mov_dpp((int)(b & 0xffffffffll), CTRL, 0xF, 0xF, false);
    hi = __builtin_amdgcn_mov_dpp((int)(b >> 32), CTRL, 0xF, 0xF, false);
  } else {
    lo = __builtin_amdgcn_update_dpp(0, (int)(b & 0xffffffffll), CTRL, ROWS, 0xF, false);
    hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), CTRL, ROWS, 0xF, false);
  }
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

__device__ inline float wave_sum(float v) {
  v += dpp_f<0xB1, 0xF>(v);
  v += dpp_f<0x4E, 0xF>(v);
  v += dpp_f<0x141, 0xF>(v);
  v += dpp_f<0x140, 0xF>(v);
  v += dpp_f<0x142, 0xA>(v);
  v += dpp_f<0x143, 0xC>(v);
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 63));
}

__device__ inline double wave_sum(double v) {
  v += dpp_d<0xB1, 0xF>(v);
  v += dpp_d<0x4E, 0xF>(v);
  v += dpp_d<0x141, 0xF>(v);
  v += dpp_d<0x140, 0xF>(v);
  v += dpp_d<0x142, 0xA>(v);
  v += dpp_d<0x143, 0xC>(v);
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffffll), 63);
  const int hi = __builtin_amdgcn_readlane((int)(b >> 32), 63);
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// Sums over each half-wave: *lo of lanes 0-31, *hi of lanes 32-63 (both
// wave-uniform).
__device__ inline void half_sums(float v, float* lo, float* hi) {
  v += dpp_f<0xB1, 0xF>(v);
  v += dpp_f<0x4E, 0xF>(v);
  v += dpp_f<0x141, 0xF>(v);
  v += dpp_f<0x140, 0xF>(v);
  v += dpp_f<0x142, 0xA>(v);
  *lo = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 31));
  *hi = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 63));
}

template <int CTRL, int ROWS>
__device__ inline float dpp_fmax_src(float v) {  // disabled rows read -inf
  return __int_as_float(__builtin_amdgcn_update_dpp((int)0xff800000u, __float_as_int(v), CTRL,
                                                    ROWS, 0xF, false));
}

__device__ inline float wave_max(float v) {
  v = fmaxf(v, dpp_fmax_src<0xB1, 0xF>(v));
  v = fmaxf(v, dpp_fmax_src<0x4E, 0xF>(v));
  v = fmaxf(v, dpp_fmax_src<0x141, 0xF>(v));
  v = fmaxf(v, dpp_fmax_src<0x140, 0xF>(v));
  v = fmaxf(v, dpp_fmax_src<0x142, 0xA>(v));
  v = fmaxf(v, dpp_fmax_src<0x143, 0xC>(v));
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 63));
}

__device__ inline int lane_id() { return threadIdx.x & (kWave - 1); }

// Ordering point of the per-pose code.  A pose is evaluated by one wave
// whose lanes exchange data through LDS with lanes of the same wave only;
// a wave's LDS (and global) accesses are performed in the order it issues
// them, so no hardware barrier is needed -- the wavefront-scope fences keep
// the compiler from moving accesses across the point.  Blocks that hold
// several poses (one per wave) thus never wait for each other's waves.
__device__ inline void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// ---------------------------------------------------------------- geometry

// math.h:176-178
__device__ inline float safe_div(float a, float b) { return a / (b + kEps); }

// Hardware forms for the bonded terms, whose inputs are FP32 geometry with
// relative errors near 1e-7 already: the reciprocal and square root are the
// 1-ulp V_RCP_F32 / V_SQRT_F32 (no correctly-rounded refinement); 1/(b + eps)
// saturates at 1e20 for b = 0 as safe_divide does.
__device__ inline float safe_rcp(float b) { return fminf(__builtin_amdgcn_rcpf(b + kEps), 1e20f); }
__device__ inline float hw_sqrt(float x) { return __builtin_amdgcn_sqrtf(x); }

// atan2(y, x) for y >= 0 (every caller passes a vector norm as y): reduced to
// t = min(|x|, y) / max(|x|, y) in [0, 1], atan(t) = t P(t^2) with P of degree
// 7, fit to atan on [0, 1] by least squares on 4,001 Chebyshev nodes with
// iteratively reweighted residuals (max error 1.3e-7 rad evaluated in FP32),
// then unfolded: pi/2 - r when y > |x|, pi - r when x < 0.
__device__ inline float atan2_ypos(float y, float x) {
  float ax = fabsf(x);
  float mx = fmaxf(ax, y), mn = fminf(ax, y);
  float t = mx > 0.f ? mn * __builtin_amdgcn_rcpf(mx) : 0.f;
  float s = t * t;
  float p = -0.00405454f;
  p = fmaf(p, s, 0.02186289f);
  p = fmaf(p, s, -0.05591225f);
  p = fmaf(p, s, 0.09642195f);
  p = fmaf(p, s, -0.1390863f);
  p = fmaf(p, s, 0.19946566f);
  p = fmaf(p, s, -0.33329862f);
  p = fmaf(p, s, 0.99999934f);
  float r = p * t;
  r = y > ax ? 1.57079632679489662f - r : r;
  return x < 0.f ? 3.14159265358979324f - r : r;
}

// sin and cos of x radians from V_SIN_F32 / V_COS_F32, which take their
// argument in revolutions: x / 2 pi is reduced to [-1/2, 1/2] first.
__device__ inline void sincos_hw(float x, float* s, float* c) {
  float r = x * 0.159154943091895336f;
  r -= rintf(r);
  *s = __builtin_amdgcn_sinf(r);
  *c = __builtin_amdgcn_cosf(r);
}

// math.cpp:62-94: Rodrigues matrix and d/dtheta
__device__ inline void rot_axis_grad(float kx, float ky, float kz, float t, float* m,
                                     float* j) {
  float s, c;
  sincosf(t, &s, &c);
  float o = 1.f - c;
  m[0] = c + kx * kx * o;
  m[1] = kx * ky * o - kz * s;
  m[2] = ky * s + kx * kz * o;
  m[3] = kz * s + kx * ky * o;
  m[4] = c + ky * ky * o;
  m[5] = ky * kz * o - kx * s;
  m[6] = kx * kz * o - ky * s;
  m[7] = kx * s + ky * kz * o;
  m[8] = c + kz * kz * o;
  j[0] = -s + kx * kx * s;
  j[1] = -kz * c + kx * ky * s;
  j[2] = ky * c + kx * kz * s;
  j[3] = kz * c + kx * ky * s;
  j[4] = -s + ky * ky * s;
  j[5] = -kx * c + ky * kz * s;
  j[6] = -ky * c + kx * kz * s;
  j[7] = kx * c + ky * kz * s;
  j[8] = -s + kz * kz * s;
}

__device__ inline void rot_axis(float kx, float ky, float kz, float t, float* m) {
  float s, c;
  sincosf(t, &s, &c);
  float o = 1.f - c;
  m[0] = c + kx * kx * o;
  m[1] = kx * ky * o - kz * s;
  m[2] = ky * s + kx * kz * o;
  m[3] = kz * s + kx * ky * o;
  m[4] = c + ky * ky * o;
  m[5] = ky * kz * o - kx * s;
  m[6] = kx * kz * o - ky * s;
  m[7] = kx * s + ky * kz * o;
  m[8] = c + kz * kz * o;
}

__device__ inline void mv3(const float* m, const float* v, float* o) {
  o[0] = m[0] * v[0] + m[1] * v[1] + m[2] * v[2];
  o[1] = m[3] * v[0] + m[4] * v[1] + m[5] * v[2];
  o[2] = m[6] * v[0] + m[7] * v[1] + m[8] * v[2];
}

__device__ inline void cross3(const float* a, const float* b, float* o) {
  o[0] = a[1] * b[2] - a[2] * b[1];
  o[1] = a[2] * b[0] - a[0] * b[2];
  o[2] = a[0] * b[1] - a[1] * b[0];
}

__device__ inline float dot3(const float* a, const float* b) {
  return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}

// math.cpp:96-110
__device__ inline float angle_grad(const float* pi, const float* pj, const float* pk,
                                   float* gi, float* gk) {
  float v0[3] = {pj[0] - pi[0], pj[1] - pi[1], pj[2] - pi[2]};
  float v1[3] = {pj[0] - pk[0], pj[1] - pk[1], pj[2] - pk[2]};
  float cr[3];
  cross3(v0, v1, cr);
  float rp = hw_sqrt(dot3(cr, cr));
  cross3(v0, cr, gi);
  cross3(cr, v1, gk);
  float s0 = -safe_rcp(dot3(v0, v0) * rp);
  float s1 = -safe_rcp(dot3(v1, v1) * rp);
  for (int d = 0; d < 3; ++d) {
    gi[d] *= s0;
    gk[d] *= s1;
  }
  return atan2_ypos(rp, dot3(v0, v1) + kEps);
}

// math.cpp:130-167
__device__ inline float dihedral_grad(const float* pi, const float* pj, const float* pk,
                                      const float* pl, float* gi, float* gj, float* gk,
                                      float* gl) {
  float rij[3] = {pj[0] - pi[0], pj[1] - pi[1], pj[2] - pi[2]};
  float rkj[3] = {pj[0] - pk[0], pj[1] - pk[1], pj[2] - pk[2]};
  float rkl[3] = {pl[0] - pk[0], pl[1] - pk[1], pl[2] - pk[2]};
  float m[3], n[3], w[3];
  cross3(rij, rkj, m);
  cross3(rkj, rkl, n);
  cross3(m, n, w);
  float wlen = hw_sqrt(dot3(w, w));
  float phi = atan2_ypos(wlen, dot3(m, n) + kEps);
  float theta = dot3(rij, n) < 0.f ? phi : -phi;
  float nrkj2 = dot3(rkj, rkj);
  float nrkj_1 = safe_rcp(hw_sqrt(nrkj2));
  float nrkj_2 = nrkj_1 * nrkj_1;
  float nrkj = nrkj2 * nrkj_1;
  float a = -nrkj * __builtin_amdgcn_rcpf(dot3(m, m));
  float b = nrkj * __builtin_amdgcn_rcpf(dot3(n, n));
  float p = dot3(rij, rkj) * nrkj_2;
  float q = dot3(rkj, rkl) * nrkj_2;
  for (int d = 0; d < 3; ++d) {
    gi[d] = -a * m[d];
    gl[d] = -b * n[d];
  }
  for (int d = 0; d < 3; ++d) {
    float sv = p * gi[d] - q * gl[d];
    gj[d] = -(gi[d] - sv);
    gk[d] = -(gl[d] + sv);
  }
  return theta;
}

// math.cpp:112-128 (angle only)
__device__ inline float dihedral(const float* pi, const float* pj, const float* pk,
                                 const float* pl) {
  float rij[3] = {pj[0] - pi[0], pj[1] - pi[1], pj[2] - pi[2]};
  float rkj[3] = {pj[0] - pk[0], pj[1] - pk[1], pj[2] - pk[2]};
  float rkl[3] = {pl[0] - pk[0], pl[1] - pk[1], pl[2] - pk[2]};
  float m[3], n[3], w[3];
  cross3(rij, rkj, m);
  cross3(rkj, rkl, n);
  cross3(m, n, w);
  float phi = atan2_ypos(hw_sqrt(dot3(w, w)), dot3(m, n) + kEps);
  return dot3(rij, n) < 0.f ? phi : -phi;
}

// Coulomb + LJ of one pair (pair.h:25-60) with r = x_j - x_i.  qq carries
// q_i q_j scale, e4 carries sqrt(eps_i eps_j) * 4 * scale.  Returns the energy
// and writes de_dr_r (the gradient on i is -r * de_dr_r, on j +r * de_dr_r).
// qk = qq * kCoulomb given by the caller (pair_energy below forms it)
__device__ inline float pair_energy_qk(float rx, float ry, float rz, float qk, float sig,
                                       float e4, float* de_dr_r) {
  float d2 = rx * rx + ry * ry + rz * rz;
  // 1/(d + eps) from one hardware reciprocal square root (1 ulp); at d = 0
  // it saturates at 1/eps = 1e20 as safe_divide does (math.h:176-178).
  // 1/(d^3 + eps) and 1/(d^2 + eps) become powers of it: equal to FP32
  // rounding for every d > 1e-6 A.
  float rd = fminf(__builtin_amdgcn_rsqf(d2), 1e20f);
  float rd2 = rd * rd;
  float ec = qk * rd;
  float u = sig * rd;
  float u2 = u * u;
  float u6 = u2 * u2 * u2;
  float u12 = u6 * u6;
  // dE/dd / d = (-qk / d^3) + (6 u^6 - 12 u^12) e4 / d^2, with 1/d^2 taken out
  const float w = fmaf(-12.f, u12, 6.f * u6);
  *de_dr_r = rd2 * fmaf(w, e4, -ec);
  return fmaf(u12 - u6, e4, ec);
}

__device__ inline float pair_energy(float rx, float ry, float rz, float qq, float sig,
                                    float e4, float* de_dr_r) {
  return pair_energy_qk(rx, ry, rz, qq * kCoulomb, sig, e4, de_dr_r);
}

// ---------------------------------------------------------------- per-pose LDS

struct PoseLds {
  float* xyz;    // [kMaxAtoms*3]
  float* q;      // [kMaxAtoms]
  float2* lj;    // [kMaxAtoms] sigma, sqrt(eps)
  int* slot;     // [kMaxAtoms] grid map index
  float* mov;    // [kMaxMovable*3]
  float* movg;   // [kMaxMovable*3]
  float* jac;    // [kMaxMovable*3]
  float* tors;   // [kMaxLeaves]
  float* torsg;  // [kMaxLeaves]
  float4* movp;  // [kMaxMovable] charge, sigma, sqrt(epsilon) of each movable slot
  float* grad;   // [kMaxAtoms*3], overlaid by stage while the bonded terms run
  float* stage;  // [kStageSlots][SW] bonded-term gradient vectors of one term chunk
                 // (SW = 3 or 4 floats per slot: the evaluation's template parameter)
};

constexpr int kStageSlots = 4 * kWave;  // 64 dihedrals x 4 atoms

#define PXD_POSE_LDS_FLOATS                                                   \
  (pxd::kMaxAtoms * 3 + pxd::kMaxAtoms * 4 + pxd::kMaxMovable * 13 +         \
   pxd::kMaxLeaves * 2 +                                                      \
   (pxd::kMaxAtoms * 3 > pxd::kStageSlots * 3 ? pxd::kMaxAtoms * 3 : pxd::kStageSlots * 3))

// Offsets (floats) of the per-pose LDS arrays for ligands of up to ma atoms,
// nm movable receptor slots and nlv leaves; every array starts on a 16-byte
// boundary.  With the maxima this is the PXD_POSE_LDS_FLOATS layout.
struct PoseDims {
  int q, lj, slot, mov, movg, jac, tors, torsg, movp, grad, total;
};

// sw: floats per bonded-term stage slot, 3 (compact: the launch-sized LDS
// of the LDS-bound kernels) or 4 (one 16-byte LDS access per slot).
__host__ __device__ constexpr PoseDims pose_dims(int ma, int nm, int nlv, int sw = 3) {
  auto r4 = [](int x) { return (x + 3) & ~3; };
  PoseDims d{};
  d.q = r4(3 * ma);
  d.lj = d.q + r4(ma);
  d.slot = d.lj + r4(2 * ma);
  d.mov = d.slot + r4(ma);
  d.movg = d.mov + r4(3 * nm);
  d.jac = d.movg + r4(3 * nm);
  d.tors = d.jac + r4(3 * nm);
  d.torsg = d.tors + r4(nlv);
  d.movp = d.torsg + r4(nlv);
  d.grad = d.movp + 4 * nm;
  const int g = r4(3 * ma) > sw * kStageSlots ? r4(3 * ma) : sw * kStageSlots;
  d.total = d.grad + g;
  return d;
}

__device__ inline PoseLds carve_pose_lds(float* base, const PoseDims& d) {
  PoseLds s;
  s.xyz = base;
  s.q = base + d.q;
  s.lj = reinterpret_cast<float2*>(base + d.lj);
  s.slot = reinterpret_cast<int*>(base + d.slot);
  s.mov = base + d.mov;
  s.movg = base + d.movg;
  s.jac = base + d.jac;
  s.tors = base + d.tors;
  s.torsg = base + d.torsg;
  s.movp = reinterpret_cast<float4*>(base + d.movp);
  s.grad = base + d.grad;
  s.stage = s.grad;
  return s;
}

__device__ inline PoseLds carve_pose_lds(float* base) {
  return carve_pose_lds(base, pose_dims(kMaxAtoms, kMaxMovable, kMaxLeaves));
}

// Per-atom parameters of a ligand into LDS (pose independent).
__device__ inline void load_ligand_params(const DevLigands& L, const LigInfo& li,
                                          const DevGrid* G, PoseLds& s) {
  for (int a = lane_id(); a < li.n_atoms; a += kWave) {
    int g = li.atom_off + a;
    s.q[a] = L.q[g];
    s.lj[a] = L.lj[g];
    s.slot[a] = G ? G->type_slot[L.type[g]] : 0;
  }
}

// Receptor atom position for a term reference (fixed atom or movable slot).
__device__ inline void rec_pos(const DevReceptor& R, const PoseLds& s, int ref, float* p) {
  if (ref >= 0) {
    float4 v = R.xyzq[ref];
    p[0] = v.x;
    p[1] = v.y;
    p[2] = v.z;
  } else {
    int m = -ref - 1;
    p[0] = s.mov[3 * m];
    p[1] = s.mov[3 * m + 1];
    p[2] = s.mov[3 * m + 2];
  }
}

__device__ inline void rec_grad_add(PoseLds& s, int ref, const float* g, float f) {
  if (ref < 0) {
    int m = -ref - 1;
    atomicAdd(&s.movg[3 * m], g[0] * f);
    atomicAdd(&s.movg[3 * m + 1], g[1] * f);
    atomicAdd(&s.movg[3 * m + 2], g[2] * f);
  }
}

__device__ inline void lig_grad_add(PoseLds& s, int a, const float* g, float f) {
  atomicAdd(&s.grad[3 * a], g[0] * f);
  atomicAdd(&s.grad[3 * a + 1], g[1] * f);
  atomicAdd(&s.grad[3 * a + 2], g[2] * f);
}

// Fourier series of one dihedral (interaction.cpp:76-102): energy and dE/dtheta.
__device__ inline float fourier(uint32_t per, const float4* kp, float theta, float* dedt) {
  float4 a = kp[0], b = kp[1], c = kp[2];
  float k[6] = {a.x, a.y, a.z, a.w, b.x, b.y};
  float ph[6] = {b.z, b.w, c.x, c.y, c.z, c.w};
  float e = 0.f, de = 0.f;
  for (int n = 0; n < PXD_DIH_TERMS; ++n) {
    int pn = (per >> (4 * n)) & 15;
    if (pn == 0) break;
    float dl = theta * pn - ph[n];
    float sn, cs;
    sincos_hw(dl, &sn, &cs);
    e += k[n] * (1.f + cs);
    de += k[n] * pn * sn;
  }
  *dedt = de;
  return e;
}

// ---------------------------------------------------------------- energy terms

// receptor_torsion2xyz (interaction.cpp:316-336): movable positions and
// Jacobians from the pose's receptor torsions (s.tors).
__device__ inline void receptor_torsion2xyz(const DevReceptor& R, PoseLds& s) {
  for (int m = lane_id(); m < R.n_mov; m += kWave) {
    int leaf = R.mov_leaf[m];
    float4 o = R.leaf_origin[leaf], ax = R.leaf_axis[leaf];
    float4 p0 = R.xyzq[R.mov_atom[m]];
    float rm[9], jm[9];
    rot_axis_grad(ax.x, ax.y, ax.z, s.tors[leaf], rm, jm);
    float t[3] = {p0.x - o.x, p0.y - o.y, p0.z - o.z}, r[3], j[3];
    mv3(rm, t, r);
    mv3(jm, t, j);
    s.mov[3 * m] = r[0] + o.x;
    s.mov[3 * m + 1] = r[1] + o.y;
    s.mov[3 * m + 2] = r[2] + o.z;
    s.jac[3 * m] = j[0];
    s.jac[3 * m + 1] = j[1];
    s.jac[3 * m + 2] = j[2];
    s.movg[3 * m] = s.movg[3 * m + 1] = s.movg[3 * m + 2] = 0.f;
    float2 l = R.lj[R.mov_atom[m]];
    s.movp[m] = make_float4(p0.w, l.x, l.y, 0.f);
  }
}

// Receptor intra terms that touch movable atoms (interaction.cpp:27-168 on
// the pruned receptor lists).  Gradients on fixed atoms are dropped: only the
// movable atoms feed the torsion chain rule (interaction.cpp:338-350).
__device__ inline double receptor_intra(const DevReceptor& R, PoseLds& s) {
  double e = 0.0;
  const int lane = lane_id();
  // with leaf tables only the terms spanning two or more leaves remain
  const bool tab = R.tab != nullptr;
  const int L = R.n_leaves;
  const int b0 = tab ? R.bond_seg[L] : 0, a0 = tab ? R.angle_seg[L] : 0,
            t0 = tab ? R.dih_seg[L] : 0, m0 = tab ? R.mm_seg[L] : 0;
  for (int b = b0 + lane; b < R.n_bonds; b += kWave) {
    int2 ab = R.bond_ab[b];
    float2 kr = R.bond_kr[b];
    float pa[3], pb[3];
    rec_pos(R, s, ab.x, pa);
    rec_pos(R, s, ab.y, pb);
    float d[3] = {pb[0] - pa[0], pb[1] - pa[1], pb[2] - pa[2]};
    float r = hw_sqrt(dot3(d, d));
    float dl = r - kr.y;
    e += 0.5f * kr.x * dl * dl;
    float de = kr.x * (1.f - kr.y * safe_rcp(r));
    rec_grad_add(s, ab.x, d, -de);
    rec_grad_add(s, ab.y, d, de);
  }
  for (int a = a0 + lane; a < R.n_angles; a += kWave) {
    int4 id = R.angle_abc[a];
    float2 kt = R.angle_kt[a];
    float pi[3], pj[3], pk[3], fi[3], fk[3];
    rec_pos(R, s, id.x, pi);
    rec_pos(R, s, id.y, pj);
    rec_pos(R, s, id.z, pk);
    float th = angle_grad(pi, pj, pk, fi, fk);
    float dl = th - kt.y;
    e += 0.5f * kt.x * dl * dl;
    float de = kt.x * dl;
    float fj[3] = {fi[0] + fk[0], fi[1] + fk[1], fi[2] + fk[2]};
    rec_grad_add(s, id.x, fi, de);
    rec_grad_add(s, id.z, fk, de);
    rec_grad_add(s, id.y, fj, -de);
  }
  for (int t = t0 + lane; t < R.n_dih; t += kWave) {
    int4 id = R.dih_abcd[t];
    float p[4][3], g[4][3];
    rec_pos(R, s, id.x, p[0]);
    rec_pos(R, s, id.y, p[1]);
    rec_pos(R, s, id.z, p[2]);
    rec_pos(R, s, id.w, p[3]);
    float th = dihedral_grad(p[0], p[1], p[2], p[3], g[0], g[1], g[2], g[3]);
    float dedt;
    e += fourier(R.dih_per[t], R.dih_kp + 3 * t, th, &dedt);
    rec_grad_add(s, id.x, g[0], -dedt);
    rec_grad_add(s, id.y, g[1], -dedt);
    rec_grad_add(s, id.z, g[2], -dedt);
    rec_grad_add(s, id.w, g[3], -dedt);
  }
  // movable-vs-movable pairs (parameters formed on the host).  A slot's
  // gradient only enters through its leaf's dE/dtheta = jac . g, so each
  // pair adds jac . g of its two slots to the two leaves' torsion gradients
  // directly (two LDS atomics instead of six on the slot gradients).
  for (int k = m0 + lane; k < R.n_mm; k += kWave) {
    const int4 mm = R.mm[k];
    const float4 pp = R.mm_par[k];
    const float* pa = s.mov + 3 * mm.x;
    const float* pb = s.mov + 3 * mm.y;
    const float r[3] = {pb[0] - pa[0], pb[1] - pa[1], pb[2] - pa[2]};
    float de;
    e += pair_energy_qk(r[0], r[1], r[2], pp.x, pp.y, pp.z, &de);
    const float* ja = s.jac + 3 * mm.x;
    const float* jb = s.jac + 3 * mm.y;
    atomicAdd(&s.torsg[mm.w & 0xffff], -de * dot3(ja, r));
    atomicAdd(&s.torsg[mm.w >> 16], de * dot3(jb, r));
  }
  wave_sync();
  if (tab) return e;  // movable-vs-fixed pairs are all inside the leaf tables
  // movable-vs-fixed pairs: per-slot segments padded to the wave width.  A
  // slot's gradient only enters through its leaf's dE/dtheta = jac . g, so
  // each lane accumulates jac . g over the leaf's slots (grouped by leaf) and
  // one wave sum per leaf adds the leaf's share to s.torsg
  float tl = 0.f;
  for (int m = 0; m < R.n_mov; ++m) {
    float px = s.mov[3 * m], py = s.mov[3 * m + 1], pz = s.mov[3 * m + 2];
    int atom = R.mov_atom[m];
    float qm = R.xyzq[atom].w;
    float2 lm = R.lj[atom];
    float gx = 0.f, gy = 0.f, gz = 0.f;
    double em = 0.0;
    for (int p = R.rp_seg[m] + lane; p < R.rp_seg[m + 1]; p += kWave) {
      float4 pj = R.rp_pos[p];
      float2 lj = R.rp_lj[p];
      float rx = pj.x - px, ry = pj.y - py, rz = pj.z - pz, de;
      em += pair_energy(rx, ry, rz, qm * pj.w, (lm.x + lj.x) * 0.5f, lm.y * lj.y, &de);
      gx -= rx * de;
      gy -= ry * de;
      gz -= rz * de;
    }
    e += em;
    tl += s.jac[3 * m] * gx + s.jac[3 * m + 1] * gy + s.jac[3 * m + 2] * gz;
    const int leaf = R.mov_leaf[m];
    if (m + 1 == R.n_mov || R.mov_leaf[m + 1] != leaf) {
      const float t = wave_sum(tl);
      if (lane == 0) s.torsg[leaf] += t;
      tl = 0.f;
    }
  }
  return e;
}

// Bonded-term gradients without atomics: a chunk of up to 64 terms (one per
// lane) writes each (term, atom) gradient vector to its own stage slot; the
// slots are numbered atom by atom, so lane then sums the contiguous slots of
// atoms lane and lane + 64 into its registers g0 / g1 (fixed order: the sums
// are deterministic).  gw: the chunk's gather word of this lane.
// SW floats per slot (3, or 4 with one 16-byte LDS access per slot).
template <int SW>
__device__ inline float3 stage_get(const PoseLds& s, int slot) {
  if (SW == 4) {
    const float4 v = reinterpret_cast<const float4*>(s.stage)[slot];
    return make_float3(v.x, v.y, v.z);
  }
  const float* v = s.stage + 3 * slot;
  return make_float3(v[0], v[1], v[2]);
}

template <int SW>
__device__ inline void stage_gather(const PoseLds& s, uint32_t gw, float* g0, float* g1) {
  wave_sync();
  const int o0 = gw & 255, c0 = (gw >> 8) & 255, o1 = (gw >> 16) & 255, c1 = gw >> 24;
  for (int k = 0; k < c0; ++k) {
    const float3 v = stage_get<SW>(s, o0 + k);
    g0[0] += v.x;
    g0[1] += v.y;
    g0[2] += v.z;
  }
  for (int k = 0; k < c1; ++k) {
    const float3 v = stage_get<SW>(s, o1 + k);
    g1[0] += v.x;
    g1[1] += v.y;
    g1[2] += v.z;
  }
  wave_sync();
}

template <int SW>
__device__ inline void stage_put(const PoseLds& s, uint32_t slots, int r, const float* v, float f) {
  const int slot = (slots >> (8 * r)) & 255;
  if (SW == 4) {
    reinterpret_cast<float4*>(s.stage)[slot] = make_float4(v[0] * f, v[1] * f, v[2] * f, 0.f);
  } else {
    float* p = s.stage + 3 * slot;
    p[0] = v[0] * f;
    p[1] = v[1] * f;
    p[2] = v[2] * f;
  }
}

// intra_molecular_interactions::put_gradients on the ligand (interaction.cpp:170-196):
// bonds, angles and dihedrals by term chunks (descriptors of the next chunk
// are loaded before the current one is evaluated), their gradients gathered
// into g0 / g1; non-bonded pairs by the list (atomics into s.grad, which is
// cleared here) only for ligands without partner masks.
template <int SW>
__device__ inline double ligand_intra(const DevLigands& L, const LigInfo& li, PoseLds& s,
                                      float* g0, float* g1) {
  double e = 0.0;
  const int lane = lane_id();
  const uint32_t* bij = L.bond_ij + li.bond_off;
  const uint32_t* bsl = L.bond_slot + li.bond_off;
  const float2* bkr = L.bond_kr + li.bond_off;
  const uint32_t* aijk = L.angle_ijk + li.angle_off;
  const uint32_t* asl = L.angle_slot + li.angle_off;
  const float2* akt = L.angle_kt + li.angle_off;
  const uint32_t* dq = L.dih_ijkl + li.dih_off;
  const uint32_t* dsl = L.dih_slot + li.dih_off;
  const uint32_t* dp = L.dih_per + li.dih_off;
  const float4* dkp = L.dih_kp + 3 * li.dih_off;
  const uint32_t* gat = L.gat + li.gat_off + lane;
  const int nbc = (li.n_bonds + kWave - 1) / kWave;
  const int nac = (li.n_angles + kWave - 1) / kWave;
  const int ndc = (li.n_dih + kWave - 1) / kWave;
  int b = lane, a = lane, t = lane;
  uint32_t b_w = 0u, b_s = 0u, a_w = 0u, a_s = 0u, d_w = 0u, d_s = 0u, d_p = 0u;
  float2 b_k = make_float2(0.f, 0.f), a_k = b_k;
  float4 d_k[3];
  if (b < li.n_bonds) {
    b_w = bij[b];
    b_s = bsl[b];
    b_k = bkr[b];
  }
  if (a < li.n_angles) {
    a_w = aijk[a];
    a_s = asl[a];
    a_k = akt[a];
  }
  if (t < li.n_dih) {
    d_w = dq[t];
    d_s = dsl[t];
    d_p = dp[t];
    d_k[0] = dkp[3 * t];
    d_k[1] = dkp[3 * t + 1];
    d_k[2] = dkp[3 * t + 2];
  }
  int chunk = 0;
  for (int c = 0; c < nbc; ++c, ++chunk, b += kWave) {
    const uint32_t gw = gat[kWave * chunk];
    const uint32_t ij = b_w, sl = b_s;
    const float2 kr = b_k;
    if (b + kWave < li.n_bonds) {
      b_w = bij[b + kWave];
      b_s = bsl[b + kWave];
      b_k = bkr[b + kWave];
    }
    if (b < li.n_bonds) {
      int i = ij & 255, j = (ij >> 8) & 255;
      float d[3] = {s.xyz[3 * j] - s.xyz[3 * i], s.xyz[3 * j + 1] - s.xyz[3 * i + 1],
                    s.xyz[3 * j + 2] - s.xyz[3 * i + 2]};
      float r = hw_sqrt(dot3(d, d));
      float dl = r - kr.y;
      e += 0.5f * kr.x * dl * dl;
      float de = kr.x * (1.f - kr.y * safe_rcp(r));
      stage_put<SW>(s, sl, 0, d, -de);
      stage_put<SW>(s, sl, 1, d, de);
    }
    stage_gather<SW>(s, gw, g0, g1);
  }
  for (int c = 0; c < nac; ++c, ++chunk, a += kWave) {
    const uint32_t gw = gat[kWave * chunk];
    const uint32_t ijk = a_w, sl = a_s;
    const float2 kt = a_k;
    if (a + kWave < li.n_angles) {
      a_w = aijk[a + kWave];
      a_s = asl[a + kWave];
      a_k = akt[a + kWave];
    }
    if (a < li.n_angles) {
      int i = ijk & 255, j = (ijk >> 8) & 255, k = (ijk >> 16) & 255;
      float fi[3], fk[3];
      float th = angle_grad(s.xyz + 3 * i, s.xyz + 3 * j, s.xyz + 3 * k, fi, fk);
      float dl = th - kt.y;
      e += 0.5f * kt.x * dl * dl;
      float de = kt.x * dl;
      float fj[3] = {fi[0] + fk[0], fi[1] + fk[1], fi[2] + fk[2]};
      stage_put<SW>(s, sl, 0, fi, de);
      stage_put<SW>(s, sl, 1, fj, -de);
      stage_put<SW>(s, sl, 2, fk, de);
    }
    stage_gather<SW>(s, gw, g0, g1);
  }
  for (int c = 0; c < ndc; ++c, ++chunk, t += kWave) {
    const uint32_t gw = gat[kWave * chunk];
    const uint32_t q = d_w, sl = d_s, per = d_p;
    float4 kp[3] = {d_k[0], d_k[1], d_k[2]};
    if (t + kWave < li.n_dih) {
      d_w = dq[t + kWave];
      d_s = dsl[t + kWave];
      d_p = dp[t + kWave];
      d_k[0] = dkp[3 * (t + kWave)];
      d_k[1] = dkp[3 * (t + kWave) + 1];
      d_k[2] = dkp[3 * (t + kWave) + 2];
    }
    if (t < li.n_dih) {
      int i = q & 255, j = (q >> 8) & 255, k = (q >> 16) & 255, l = (q >> 24) & 255;
      float g[4][3];
      float th = dihedral_grad(s.xyz + 3 * i, s.xyz + 3 * j, s.xyz + 3 * k, s.xyz + 3 * l,
                               g[0], g[1], g[2], g[3]);
      float dedt;
      e += fourier(per, kp, th, &dedt);
      stage_put<SW>(s, sl, 0, g[0], -dedt);
      stage_put<SW>(s, sl, 1, g[1], -dedt);
      stage_put<SW>(s, sl, 2, g[2], -dedt);
      stage_put<SW>(s, sl, 3, g[3], -dedt);
    }
    stage_gather<SW>(s, gw, g0, g1);
  }
  // pairs by partner masks (ligand_pairs_ring / _ring2): the gradient
  // stays in registers and is stored at the end of inter, so the region the
  // stage overlays is left to the ring's partner records
  if (li.pmode) {
    wave_sync();
    return e;
  }
  // the stage is free again: clear the gradient array it overlays
  for (int k = lane; k < li.n_atoms * 3; k += kWave) s.grad[k] = 0.f;
  wave_sync();
  for (int p = lane; p < li.n_pairs; p += kWave) {
    uint32_t w = L.pairs[li.pair_off + p];
    int i = w & 255, j = (w >> 8) & 255;
    bool is14 = (w >> 16) & 1;
    float cs = is14 ? kScale14Coul : 1.f, vs = is14 ? kScale14Vdw : 1.f;
    float r[3] = {s.xyz[3 * j] - s.xyz[3 * i], s.xyz[3 * j + 1] - s.xyz[3 * i + 1],
                  s.xyz[3 * j + 2] - s.xyz[3 * i + 2]};
    float2 la = s.lj[i], lb = s.lj[j];
    float de;
    e += pair_energy(r[0], r[1], r[2], s.q[i] * s.q[j] * cs, (la.x + lb.x) * 0.5f,
                     la.y * lb.y * 4.f * vs, &de);
    lig_grad_add(s, i, r, -de);
    lig_grad_add(s, j, r, de);
  }
  return e;
}

// Ligand non-bonded pairs (interaction.cpp:127-168) by partner masks (no
// atomics); a pair listed as all-pairs and/or 1-4 contributes with the sum
// of its scales.  Ligands of at most 64 atoms: each pair met
// once: lane a owns atom a; in step k = 1 .. n/2 it takes the pair
// (a, a + k mod n) — every unordered pair once, the last step of an even n
// only from its lower half.  Lane a keeps its own gradient in g0; the
// partner's reaction goes into a travelling accumulator t, which holds atom
// a + k's sum in step k and moves one lane down the ring (ds_bpermute) after
// each step; at the end it is sent home.  The partner bits come from the
// owned atom's masks rotated so that bit k is partner a + k.  Partner data
// come from 24-byte records (x, y, z, q, sigma / 2, sqrt(eps)) written twice,
// at j and j + n, into the gradient region (free here: this mode keeps the
// gradient in registers), so partner a + k is one pointer step away.
template <int SW>
__device__ inline double ligand_pairs_ring(const DevLigands& L, const LigInfo& li,
                                           const PoseLds& s, float* g0) {
  const int lane = lane_id(), n = li.n_atoms;
  // Ligands of at most 32 atoms: both half-waves own the atoms (lane mod 32)
  // and split the ring steps, the lower half taking steps 1 .. K0, the upper
  // K0 + 1 .. K; each half runs its own ring of travelling accumulators and
  // the halves' sums are added at the end.
  const bool split = n <= kWave / 2;
  const int half = split ? lane >> 5 : 0, base = half << 5;
  const int al = split ? (lane & 31) : lane;
  const bool act = al < n;
  const int a = act ? al : 0;
  const uint4* pm = L.pmask + 2ll * li.atom_off;
  uint64_t ra = 0, rf = 0;
  if (act) {
    const uint4 ma = pm[2 * a], mf = pm[2 * a + 1];
    const uint64_t full = n == 64 ? ~0ull : ((1ull << n) - 1);
    const uint64_t wa = ((uint64_t)ma.y << 32 | ma.x) & full, wf = ((uint64_t)mf.y << 32 | mf.x) & full;
    // rotate right by a within n bits: bit k <- bit (a + k) mod n
    ra = a ? ((wa >> a) | (wa << (n - a))) & full : wa;
    rf = a ? ((wf >> a) | (wf << (n - a))) & full : wf;
  }
  const float x0 = s.xyz[3 * a], y0 = s.xyz[3 * a + 1], z0 = s.xyz[3 * a + 2], q0 = s.q[a];
  const float2 l0 = s.lj[a];
  // 2n records of float2 (x, y), (z, q), (sigma / 2, sqrt eps) (and, with
  // 16-byte stage slots, a pad pair: 32-byte records), n <= 64
  constexpr int RW = SW == 4 ? 4 : 3;  // float2 per record
  static_assert(2 * kWave * 2 * RW <= SW * kStageSlots, "2n records fit the gradient region");
  float2* rec = reinterpret_cast<float2*>(s.grad);
  if (act && !half) {
    const float2 r0 = make_float2(x0, y0), r1 = make_float2(z0, q0),
                 r2 = make_float2(0.5f * l0.x, l0.y);
    if (RW == 4) {  // 32-byte records: one 16-byte access for (x, y, z, q)
      const float4 r01 = make_float4(x0, y0, z0, q0);
      *reinterpret_cast<float4*>(rec + RW * a) = r01;
      *reinterpret_cast<float4*>(rec + RW * (a + n)) = r01;
    } else {
      rec[RW * a] = r0;
      rec[RW * a + 1] = r1;
      rec[RW * (a + n)] = r0;
      rec[RW * (a + n) + 1] = r1;
    }
    rec[RW * a + 2] = r2;
    rec[RW * (a + n) + 2] = r2;
  }
  wave_sync();
  // own factors of the pair parameters: q_a * kCoulomb, sigma_a / 2, 4 sqrt(eps_a)
  const float q0k = q0 * kCoulomb, sh0 = 0.5f * l0.x, e40 = 4.f * l0.y;
  // ring successor (source of t) within the half
  const int nxt = act ? base + (a + 1 == n ? 0 : a + 1) : lane;
  const int K = n / 2;
  const int K0 = split ? (K + 1) / 2 : K;  // steps per half: k0 + 1 .. k0 + K0
  const int k0 = half * K0;
  float gr[3] = {0.f, 0.f, 0.f}, t[3] = {0.f, 0.f, 0.f}, es = 0.f;
  const float2* pr = rec + RW * (a + 1 + k0);  // partner a + k's record in step k
  // bits 1 .. 32 of the rotated masks (K <= 32 steps), bit k - 1 = step k;
  // with n even, the pair (a, a + n/2) is met from a < n/2 only: lanes
  // a >= n/2 drop step n/2 (inactive lanes have empty masks).  The upper
  // half shifts its steps to bit 0 and drops those past K.
  uint32_t ba = (uint32_t)(ra >> 1), bf = (uint32_t)(rf >> 1);
  if (K > 0 && 2 * K == n && a >= K) {
    ba &= ~(1u << (K - 1));
    bf &= ~(1u << (K - 1));
  }
  if (half) {
    const uint32_t keep = (1u << (K - k0)) - 1;  // K - k0 <= 8 here
    ba = (ba >> k0) & keep;
    bf = (bf >> k0) & keep;
  }
  for (int k = 1; k <= K0; ++k, pr += RW) {
    const uint32_t ia = __builtin_amdgcn_ubfe(ba, k - 1, 1), i4 = __builtin_amdgcn_ubfe(bf, k - 1, 1);
    if (ia | i4) {
      const float fa = (float)ia, f4 = (float)i4;  // scales: 1, the 1-4 scale, or their sum
      const float cs = fmaf(f4, kScale14Coul, fa), vs = fmaf(f4, kScale14Vdw, fa);
      float2 pxy, pzq;
      if (RW == 4) {
        const float4 p4 = *reinterpret_cast<const float4*>(pr);
        pxy = make_float2(p4.x, p4.y);
        pzq = make_float2(p4.z, p4.w);
      } else {
        pxy = pr[0];
        pzq = pr[1];
      }
      const float2 lb = pr[2];
      const float rx = pxy.x - x0, ry = pxy.y - y0, rz = pzq.x - z0;
      float de;
      es += pair_energy_qk(rx, ry, rz, q0k * pzq.y * cs, sh0 + lb.x, e40 * lb.y * vs, &de);
      gr[0] -= rx * de;
      gr[1] -= ry * de;
      gr[2] -= rz * de;
      t[0] += rx * de;
      t[1] += ry * de;
      t[2] += rz * de;
    }
    if (k < K0) {
#pragma unroll
      for (int d = 0; d < 3; ++d)
        t[d] = __int_as_float(__builtin_amdgcn_ds_bpermute(4 * nxt, __float_as_int(t[d])));
    }
  }
  // t of lane a belongs to atom a + k0 + K0: atom j's sum sits in lane
  // j - k0 - K0 of the same half
  if (K0 > 0) {
    int home = a - k0 - K0;
    if (home < 0) home += n;
    home = act ? base + home : lane;
#pragma unroll
    for (int d = 0; d < 3; ++d)
      gr[d] += __int_as_float(__builtin_amdgcn_ds_bpermute(4 * home, __float_as_int(t[d])));
  }
  if (split) {  // the upper half's sums to the owning lane of the lower half
#pragma unroll
    for (int d = 0; d < 3; ++d)
      gr[d] += __int_as_float(__builtin_amdgcn_ds_bpermute(4 * (lane ^ 32), __float_as_int(gr[d])));
  }
  if (!half) {
    g0[0] += gr[0];
    g0[1] += gr[1];
    g0[2] += gr[2];
  }
  return (double)es;
}

// Ligand non-bonded pairs for ligands of 65 .. 128 atoms, each pair met once
// on the same ring schedule: lane l owns atoms l (slot 0) and l + 64 (slot
// 1, when < n), owner index i = l + 64 * slot; in step k = 1 .. n/2 owner i
// takes the pair (i, i + k mod n).  Each slot keeps its own gradient in
// registers and carries a travelling accumulator for its partner; after a
// step every accumulator moves from owner i to owner i - 1 mod n (slot 0's
// lane 63 takes slot 1's lane 0, slot 1's last owner takes owner 0), and at
// the end atom j's sum is fetched from owner j - n/2 mod n.
__device__ inline uint64_t ring_bits(uint4 m, int i, int n) {
  // bits 1 .. 64 of the n-bit mask m rotated right by i (bit k <- bit (i + k) mod n)
  typedef unsigned __int128 u128;
  const u128 full = n == 128 ? ~(u128)0 : (((u128)1 << n) - 1);
  const u128 w = (((u128)m.w << 96) | ((u128)m.z << 64) | ((u128)m.y << 32) | (u128)m.x) & full;
  const u128 r = i ? ((w >> i) | (w << (n - i))) & full : w;
  return (uint64_t)(r >> 1);
}

__device__ inline double ligand_pairs_ring2(const DevLigands& L, const LigInfo& li,
                                            const PoseLds& s, float* g0, float* g1) {
  const int lane = lane_id(), n = li.n_atoms;  // 64 < n <= 128
  const int i0 = lane, i1 = lane + kWave;
  const bool act1 = i1 < n;
  const int j1 = act1 ? i1 : i0;  // inactive slot 1 reads atom i0 (never live)
  const uint4* pm = L.pmask + 2ll * li.atom_off;
  uint64_t a0 = ring_bits(pm[2 * i0], i0, n), f0 = ring_bits(pm[2 * i0 + 1], i0, n);
  uint64_t a1 = 0, f1 = 0;
  if (act1) {
    a1 = ring_bits(pm[2 * i1], i1, n);
    f1 = ring_bits(pm[2 * i1 + 1], i1, n);
  }
  const float x0 = s.xyz[3 * i0], y0 = s.xyz[3 * i0 + 1], z0 = s.xyz[3 * i0 + 2];
  const float x1 = s.xyz[3 * j1], y1 = s.xyz[3 * j1 + 1], z1 = s.xyz[3 * j1 + 2];
  const float2 l0 = s.lj[i0], l1 = s.lj[j1];
  // own factors of the pair parameters: q * kCoulomb, sigma / 2, 4 sqrt(eps)
  const float q0k = s.q[i0] * kCoulomb, sh0 = 0.5f * l0.x, e40 = 4.f * l0.y;
  const float q1k = s.q[j1] * kCoulomb, sh1 = 0.5f * l1.x, e41 = 4.f * l1.y;
  const int K = n / 2;
  const int src = (lane + 1) & (kWave - 1);  // owner i + 1 of slot 0 (lane 63: slot 1's lane 0)
  const bool last1 = i1 + 1 == n;            // slot 1's last owner: its successor is owner 0
  float t0[3] = {0.f, 0.f, 0.f}, t1[3] = {0.f, 0.f, 0.f}, es = 0.f;
  double e = 0.0;
  // with n even, the pair (i, i + n/2) is met from i < n/2 only: owners
  // i >= n/2 drop step n/2 (bit K - 1; inactive slot 1 has empty masks)
  if (2 * K == n) {
    const uint64_t drop = ~(1ull << (K - 1));
    if (i0 >= K) {
      a0 &= drop;
      f0 &= drop;
    }
    if (i1 >= K) {
      a1 &= drop;
      f1 &= drop;
    }
  }
  // the 32-bit half of the masks that holds step k's bit (k - 1)
  uint32_t wa0 = (uint32_t)a0, wf0 = (uint32_t)f0, wa1 = (uint32_t)a1, wf1 = (uint32_t)f1;
  int b0 = i0, b1 = j1;
  for (int k = 1; k <= K; ++k) {
    b0 = b0 + 1 == n ? 0 : b0 + 1;
    b1 = b1 + 1 == n ? 0 : b1 + 1;
    if (k == 33) {
      wa0 = (uint32_t)(a0 >> 32);
      wf0 = (uint32_t)(f0 >> 32);
      wa1 = (uint32_t)(a1 >> 32);
      wf1 = (uint32_t)(f1 >> 32);
    }
    const int bit = (k - 1) & 31;
    {
      const bool ia = __builtin_amdgcn_ubfe(wa0, bit, 1), i4 = __builtin_amdgcn_ubfe(wf0, bit, 1);
      if (ia || i4) {
        const float cs = (ia ? 1.f : 0.f) + (i4 ? kScale14Coul : 0.f);
        const float vs = (ia ? 1.f : 0.f) + (i4 ? kScale14Vdw : 0.f);
        const float2 lb = s.lj[b0];
        const float rx = s.xyz[3 * b0] - x0, ry = s.xyz[3 * b0 + 1] - y0,
                    rz = s.xyz[3 * b0 + 2] - z0;
        float de;
        es += pair_energy_qk(rx, ry, rz, q0k * s.q[b0] * cs, fmaf(lb.x, 0.5f, sh0),
                             e40 * lb.y * vs, &de);
        g0[0] -= rx * de;
        g0[1] -= ry * de;
        g0[2] -= rz * de;
        t0[0] += rx * de;
        t0[1] += ry * de;
        t0[2] += rz * de;
      }
    }
    {
      const bool ia = __builtin_amdgcn_ubfe(wa1, bit, 1), i4 = __builtin_amdgcn_ubfe(wf1, bit, 1);
      if (ia || i4) {  // zero masks when slot 1 is inactive
        const float cs = (ia ? 1.f : 0.f) + (i4 ? kScale14Coul : 0.f);
        const float vs = (ia ? 1.f : 0.f) + (i4 ? kScale14Vdw : 0.f);
        const float2 lb = s.lj[b1];
        const float rx = s.xyz[3 * b1] - x1, ry = s.xyz[3 * b1 + 1] - y1,
                    rz = s.xyz[3 * b1 + 2] - z1;
        float de;
        es += pair_energy_qk(rx, ry, rz, q1k * s.q[b1] * cs, fmaf(lb.x, 0.5f, sh1),
                             e41 * lb.y * vs, &de);
        g1[0] -= rx * de;
        g1[1] -= ry * de;
        g1[2] -= rz * de;
        t1[0] += rx * de;
        t1[1] += ry * de;
        t1[2] += rz * de;
      }
    }
    if ((k & 31) == 0) {  // FP32 pair energies per block of 32 steps, then FP64
      e += (double)es;
      es = 0.f;
    }
    if (k < K) {
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        const float h0 = __int_as_float(__builtin_amdgcn_ds_bpermute(4 * src, __float_as_int(t0[d])));
        const float h1 = __int_as_float(__builtin_amdgcn_ds_bpermute(4 * src, __float_as_int(t1[d])));
        const float z = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(t0[d]), 0));
        t0[d] = lane == kWave - 1 ? h1 : h0;
        t1[d] = last1 ? z : h1;
      }
    }
  }
  e += (double)es;
  // atom j's sum sits at owner (j - K) mod n
#pragma unroll
  for (int slot = 0; slot < 2; ++slot) {
    const int j = slot ? j1 : i0;
    int o = j - K;
    if (o < 0) o += n;
    const int sl = o & (kWave - 1);
    float* g = slot ? g1 : g0;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      const float h0 = __int_as_float(__builtin_amdgcn_ds_bpermute(4 * sl, __float_as_int(t0[d])));
      const float h1 = __int_as_float(__builtin_amdgcn_ds_bpermute(4 * sl, __float_as_int(t1[d])));
      if (slot == 0 || act1) g[d] += o < kWave ? h0 : h1;
    }
  }
  return e;
}

// Trilinear value + stored-gradient interpolation (grid.cpp:273-303).
__device__ inline float4 interp(const float4* map, const int* off, const float* w) {
  // the maps live in global memory: global (not flat) loads, which do not
  // hold up the waits for LDS results
  typedef float vf4 __attribute__((ext_vector_type(4)));
  typedef const __attribute__((address_space(1))) vf4 gvf4;
  const gvf4* gm = (const gvf4*)map;
  float4 acc = make_float4(0.f, 0.f, 0.f, 0.f);
#pragma unroll
  for (int c = 0; c < 8; ++c) {
    const vf4 v = gm[off[c]];
    acc.x += v.x * w[c];
    acc.y += v.y * w[c];
    acc.z += v.z * w[c];
    acc.w += v.w * w[c];
  }
  return acc;
}

// The same interpolation from component planes (map layout 2): value and
// gradient components P floats apart, the same order of operations.
__device__ inline float4 interp_planes(const float* map, long long P, const int* off,
                                       const float* w) {
  typedef const __attribute__((address_space(1))) float gf;
  const gf* gm = (const gf*)map;
  float4 acc = make_float4(0.f, 0.f, 0.f, 0.f);
#pragma unroll
  for (int c = 0; c < 8; ++c) {
    acc.x += gm[off[c]] * w[c];
    acc.y += gm[P + off[c]] * w[c];
    acc.z += gm[2 * P + off[c]] * w[c];
    acc.w += gm[3 * P + off[c]] * w[c];
  }
  return acc;
}

// receptor_cache::get (grid.cpp:197-262) for one ligand atom: adds into e/g.
__device__ inline void grid_get(const DevGrid& G, int slot, float q, const float* x,
                                double* e, float* g) {
  // The cell coordinate is formed in FP64: in FP32 its fractional part would
  // carry ~1e-6 absolute error, i.e. a large relative error on the small
  // trilinear weights that multiply steep (clash) corner values.
  float f[3], gt[3], pen = 0.f;
  int lo[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    double fd = ((double)x[d] - G.init_d[d]) * G.inv_spacing_d;
    int nm1 = G.np[d] - 1;
    if (fd < 0.0) {
      lo[d] = 0;
      gt[d] = -1.f;
      pen -= (float)fd;
      fd = 0.0;
    } else if (fd < (double)nm1) {
      lo[d] = (int)fd;
      gt[d] = 0.f;
    } else {
      lo[d] = nm1 - 1;
      gt[d] = 1.f;
      pen += (float)(fd - (double)nm1);
      fd = (double)nm1;
    }
    f[d] = (float)(fd - (double)lo[d]);
  }
  float x0 = f[0], y0 = f[1], z0 = f[2], mx = 1.f - x0, my = 1.f - y0, mz = 1.f - z0;
  float w[8] = {mx * my * mz, x0 * my * mz, mx * y0 * mz, x0 * y0 * mz,
                mx * my * z0, x0 * my * z0, mx * y0 * z0, x0 * y0 * z0};
  bool hit = true;
  if (pen > 0.f) {
    *e += (double)(pen * G.spacing * G.slope);
    g[0] += gt[0] * G.slope;
    g[1] += gt[1] * G.slope;
    g[2] += gt[2] * G.slope;
    hit = false;
  }
  // layout 0: point-major maps, the corners 1, nz, ny*nz points apart;
  // layout 1: the cell's 8 corners contiguous (one 128-byte line per map);
  // layout 2: point-major component planes (value, dE/dx, dE/dy, dE/dz).
  // One base address per map; the corners are constant (layout 1) or shared
  // 32-bit (layouts 0, 2) offsets from it.
  float4 c, v;
  if (G.layout == 2) {
    const int nz = G.np[2], nyz = G.np[1] * G.np[2];
    const long long base = ((long long)lo[0] * G.np[1] + lo[1]) * nz + lo[2];
    const float* p0 = G.smaps + base;
    const float* p1 = G.smaps + 4 * G.P * slot + base;
    const int o[8] = {0, nyz, nz, nyz + nz, 1, nyz + 1, nz + 1, nyz + nz + 1};
    c = interp_planes(p0, G.P, o, w);
    v = interp_planes(p1, G.P, o, w);
  } else if (G.layout) {
    const long long cell =
        (((long long)lo[0] * (G.np[1] - 1) + lo[1]) * (G.np[2] - 1) + lo[2]) * 8;
    const float4* p0 = G.cmaps + cell;
    const float4* p1 = G.cmaps + G.C * 8 * slot + cell;
    const int o[8] = {0, 1, 2, 3, 4, 5, 6, 7};
    c = interp(p0, o, w);
    v = interp(p1, o, w);
  } else {
    const int nz = G.np[2], nyz = G.np[1] * G.np[2];
    const long long base = ((long long)lo[0] * G.np[1] + lo[1]) * nz + lo[2];
    const float4* p0 = G.maps + base;
    const float4* p1 = G.maps + G.P * slot + base;
    const int o[8] = {0, nyz, nz, nyz + nz, 1, nyz + 1, nz + 1, nyz + nz + 1};
    c = interp(p0, o, w);
    v = interp(p1, o, w);
  }
  *e += (double)(c.x * q + v.x);
  if (hit) {
    g[0] += c.y * q + v.y;
    g[1] += c.z * q + v.z;
    g[2] += c.w * q + v.w;
  }
}

// inter_molecular_interactions::put_gradients (interaction.cpp:202-257).
// With a grid: map lookup + pairs against movable receptor atoms; without:
// pairs against every receptor atom (fixed atoms by broadcast loads).
__device__ inline double inter(const DevReceptor& R, const DevGrid* G, const LigInfo& li,
                               PoseLds& s, float* g0, float* g1) {
  double e = 0.0;
  const int lane = lane_id();
  const int n = li.n_atoms;
  // lane owns atoms lane and lane + 64; g0 / g1 carry their gradients
  const int a0 = lane, a1 = lane + kWave;
  if (G) {
    if (a0 < n) grid_get(*G, s.slot[a0], s.q[a0], s.xyz + 3 * a0, &e, g0);
    if (a1 < n) grid_get(*G, s.slot[a1], s.q[a1], s.xyz + 3 * a1, &e, g1);
  } else {
    for (int k = 0; k < R.n_fixed; ++k) {
      int j = R.fixed[k];
      float4 pj = R.xyzq[j];
      float2 lj = R.lj[j];
      for (int h = 0; h < 2; ++h) {
        int a = h ? a1 : a0;
        if (a >= n) continue;
        float* gg = h ? g1 : g0;
        float rx = pj.x - s.xyz[3 * a], ry = pj.y - s.xyz[3 * a + 1], rz = pj.z - s.xyz[3 * a + 2];
        float2 la = s.lj[a];
        float de;
        e += pair_energy(rx, ry, rz, s.q[a] * pj.w, (la.x + lj.x) * 0.5f, la.y * lj.y * 4.f, &de);
        gg[0] -= rx * de;
        gg[1] -= ry * de;
        gg[2] -= rz * de;
      }
    }
  }
  // movable receptor atoms: explicit pairs; as in receptor_intra, the slots'
  // gradients enter through jac . g summed per leaf (one wave sum per leaf).
  // The movable atom's factors of the pair parameters (q * kCoulomb,
  // sigma / 2, 4 sqrt(eps)) are formed once per atom for the wave; the pair
  // energies are summed in FP32 per lane, then added to the FP64 total.
  float tl = 0.f, em = 0.f;
  if (n <= kWave / 2) {
    // Ligands of at most 32 atoms: both half-waves own the atoms (lane mod
    // 32) and take alternate movable slots, the lower half slot 2i and the
    // upper 2i + 1; one pass of half-wave sums gives both slots' jac . g.  A
    // half's running sum is flushed whenever the next slot of either half
    // starts another leaf (a partial flush adds to the same leaf later).
    const int half = lane >> 5, am = lane & 31, M = R.n_mov;
    const bool act = am < n;
    const int ac = act ? am : 0;
    const float xa = s.xyz[3 * ac], ya = s.xyz[3 * ac + 1], za = s.xyz[3 * ac + 2];
    const float2 la = s.lj[ac];
    const float qa = s.q[ac], sa = 0.5f * la.x;
    float gm[3] = {0.f, 0.f, 0.f};
    for (int i = 0; 2 * i < M; ++i) {
      const int mlo = 2 * i, mhi = 2 * i + 1;
      const bool has_hi = mhi < M;
      const int m = half && has_hi ? mhi : mlo;
      float gx = 0.f, gy = 0.f, gz = 0.f;
      if (act && (!half || has_hi)) {
        const float4 mp = s.movp[m];
        const float rx = s.mov[3 * m] - xa, ry = s.mov[3 * m + 1] - ya, rz = s.mov[3 * m + 2] - za;
        float de;
        em += pair_energy_qk(rx, ry, rz, qa * (mp.x * kCoulomb), fmaf(mp.y, 0.5f, sa),
                             la.y * (4.f * mp.z), &de);
        gx = rx * de;
        gy = ry * de;
        gz = rz * de;
        gm[0] -= gx;
        gm[1] -= gy;
        gm[2] -= gz;
      }
      tl += s.jac[3 * m] * gx + s.jac[3 * m + 1] * gy + s.jac[3 * m + 2] * gz;
      const int leaf_lo = R.mov_leaf[mlo], leaf_hi = has_hi ? R.mov_leaf[mhi] : -1;
      const bool flush = mlo + 2 >= M || R.mov_leaf[mlo + 2] != leaf_lo ||
                         (has_hi && (mhi + 2 >= M || R.mov_leaf[mhi + 2] != leaf_hi));
      if (flush) {
        float tlo, thi;
        half_sums(tl, &tlo, &thi);
        if (lane == 0) {
          s.torsg[leaf_lo] += tlo;
          if (has_hi) s.torsg[leaf_hi] += thi;
        }
        tl = 0.f;
      }
    }
#pragma unroll
    for (int d = 0; d < 3; ++d)
      gm[d] += __int_as_float(__builtin_amdgcn_ds_bpermute(4 * (lane ^ 32), __float_as_int(gm[d])));
    if (!half) {
      g0[0] += gm[0];
      g0[1] += gm[1];
      g0[2] += gm[2];
    }
  } else
  for (int m = 0; m < R.n_mov; ++m) {
    const float4 mp = s.movp[m];
    const float qmk = mp.x * kCoulomb, smh = 0.5f * mp.y, em4 = 4.f * mp.z;
    float px = s.mov[3 * m], py = s.mov[3 * m + 1], pz = s.mov[3 * m + 2];
    float gx = 0.f, gy = 0.f, gz = 0.f;
    for (int h = 0; h < 2; ++h) {
      int a = h ? a1 : a0;
      if (a >= n) continue;
      float* gg = h ? g1 : g0;
      float rx = px - s.xyz[3 * a], ry = py - s.xyz[3 * a + 1], rz = pz - s.xyz[3 * a + 2];
      float2 la = s.lj[a];
      float de;
      em += pair_energy_qk(rx, ry, rz, s.q[a] * qmk, fmaf(la.x, 0.5f, smh), la.y * em4, &de);
      gg[0] -= rx * de;
      gg[1] -= ry * de;
      gg[2] -= rz * de;
      gx += rx * de;
      gy += ry * de;
      gz += rz * de;
    }
    tl += s.jac[3 * m] * gx + s.jac[3 * m + 1] * gy + s.jac[3 * m + 2] * gz;
    const int leaf = R.mov_leaf[m];
    if (m + 1 == R.n_mov || R.mov_leaf[m + 1] != leaf) {
      const float t = wave_sum(tl);
      if (lane == 0) s.torsg[leaf] += t;
      tl = 0.f;
    }
  }
  e += (double)em;
  if (li.pmode) {  // nothing else wrote s.grad: store (the region held scratch)
    for (int d = 0; d < 3; ++d) {
      if (a0 < n) s.grad[3 * a0 + d] = g0[d];
      if (a1 < n) s.grad[3 * a1 + d] = g1[d];
    }
  } else {
    if (a0 < n) lig_grad_add(s, a0, g0, 1.f);
    if (a1 < n) lig_grad_add(s, a1, g1, 1.f);
  }
  return e;
}

// Cubic Hermite lookup of leaf k's table at theta (any finite theta; NaN or
// infinite theta gives NaN).  The table index stays in [0, n) for every input.
__device__ inline double leaf_reduce(double th) {
  const double two_pi = 6.283185307179586;
  // |theta| <= 2^20: reduction by floor (a few FP64 instructions; the
  // rounding of theta / 2 pi moves q by ~1e-15 rad, far below a table step).
  // Beyond, the exact remainder: theta / 2 pi then carries no fraction bits
  // an FMA-contracted floor form could keep inside one revolution.
  if (fabs(th) <= 1048576.0) return th - two_pi * floor(th * (1.0 / two_pi));
  return fmod(th, two_pi);
}

__device__ inline void leaf_lookup(const DevReceptor& R, int k, float theta, double* e,
                                   double* d) {
  const int n = R.tab_n;
  const double two_pi = 6.283185307179586, h = two_pi / n, inv_h = n * (1.0 / two_pi);
  double q = leaf_reduce((double)theta);
  if (!(q == q)) {  // NaN (or infinite theta) in, NaN out
    *e = *d = q;
    return;
  }
  if (q < 0.0) q += two_pi;
  double u = q * inv_h;
  double fl = floor(u);
  double t = u - fl;
  int i0 = (int)fl;
  if (i0 >= n) i0 -= n;
  i0 = i0 < 0 ? 0 : (i0 > n - 1 ? n - 1 : i0);  // bound the read whatever the rounding
  int i1 = i0 + 1 == n ? 0 : i0 + 1;
  double2 a = R.tab[(long long)k * n + i0], b = R.tab[(long long)k * n + i1];
  double t2 = t * t, t3 = t2 * t;
  double h00 = 2 * t3 - 3 * t2 + 1, h10 = t3 - 2 * t2 + t, h01 = -2 * t3 + 3 * t2,
         h11 = t3 - t2;
  *e = h00 * a.x + h10 * h * a.y + h01 * b.x + h11 * h * b.y;
  double g00 = 6 * t2 - 6 * t, g10 = 3 * t2 - 4 * t + 1, g01 = -6 * t2 + 6 * t,
         g11 = 3 * t2 - 2 * t;
  *d = (g00 * a.x + g01 * b.x) * inv_h + g10 * a.y + g11 * b.y;
}

// binding_system_interactions::put_gradients (interaction.cpp:271-298) for
// the pose whose ligand coordinates and receptor torsions are in s.xyz /
// s.tors.  Gradients are written to s.grad / s.torsg (overwrite).  Returns
// the wave-uniform FP64 energy.  One wave per pose (wave_sync: no block barrier).
template <int SW = 3>
__device__ inline double eval_pose(const DevReceptor& R, const DevLigands& L, const DevGrid* G,
                                   const LigInfo& li, PoseLds& s) {
  const int lane = lane_id();
  for (int k = lane; k < R.n_leaves; k += kWave) s.torsg[k] = 0.f;
  receptor_torsion2xyz(R, s);
  wave_sync();
  double e = receptor_intra(R, s);
  if (lane == 0) e += R.e_const;
  wave_sync();
  float g0[3] = {0.f, 0.f, 0.f}, g1[3] = {0.f, 0.f, 0.f};
  e += ligand_intra<SW>(L, li, s, g0, g1);
  if (li.pmode) {
    if (li.n_atoms <= kWave)
      e += ligand_pairs_ring<SW>(L, li, s, g0);
    else
      e += ligand_pairs_ring2(L, li, s, g0, g1);
  }
  e += inter(R, G, li, s, g0, g1);
  wave_sync();
  // receptor_xyz2torsion (interaction.cpp:338-350): slots are grouped by leaf;
  // a leaf table contributes its energy and dT/dtheta directly
  for (int k = lane; k < R.n_leaves; k += kWave) {
    double te = 0.0, td = 0.0;
    if (R.tab) leaf_lookup(R, k, s.tors[k], &te, &td);
    e += te;
    s.torsg[k] += (float)td;
  }
  wave_sync();
  // the slot gradients (s.movg) hold the receptor bonded terms' shares only
  const int nl = R.n_leaves;
  const bool movg_used = R.tab ? (R.bond_seg[nl] < R.n_bonds || R.angle_seg[nl] < R.n_angles ||
                                  R.dih_seg[nl] < R.n_dih)
                               : (R.n_bonds > 0 || R.n_angles > 0 || R.n_dih > 0);
  if (movg_used)
    for (int m = lane; m < R.n_mov; m += kWave) {
      float v = s.jac[3 * m] * s.movg[3 * m] + s.jac[3 * m + 1] * s.movg[3 * m + 1] +
                s.jac[3 * m + 2] * s.movg[3 * m + 2];
      atomicAdd(&s.torsg[R.mov_leaf[m]], v);
    }
  wave_sync();
  return wave_sum(e);
}

}  // namespace pxd
