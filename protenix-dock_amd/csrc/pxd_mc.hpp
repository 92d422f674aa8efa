// pxd_mc.hpp — conformation operators and random numbers for the device
// Monte-Carlo search: free_ligand::apply_parameters on a wavefront, the
// receptor-box checks and the counter-based generator.
#pragma once

#include "pxd_device.hpp"

namespace pxd {

// ---------------------------------------------------------------- RNG
// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11).  Key = 64-bit search
// seed, counter = (draw index, walker id, 0, 0).  Every lane computes the
// same block, so draws are wave-uniform.

__device__ inline void philox(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, uint32_t k0,
                              uint32_t k1, uint32_t* out) {
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    uint64_t p0 = (uint64_t)0xD2511F53u * c0;
    uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
    uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
    uint32_t n1 = (uint32_t)p1;
    uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
    uint32_t n3 = (uint32_t)p0;
    c0 = n0;
    c1 = n1;
    c2 = n2;
    c3 = n3;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
  out[0] = c0;
  out[1] = c1;
  out[2] = c2;
  out[3] = c3;
}

struct Rng {
  uint32_t k0, k1, walker, ctr;
  // one Philox block per draw; returns 4 words
  __device__ void block(uint32_t* w) { philox(ctr++, walker, 0u, 0u, k0, k1, w); }
  // uniform in [0, 1) with 32 random bits (exact in double)
  __device__ double u01() {
    uint32_t w[4];
    block(w);
    return (double)w[0] * 2.3283064365386963e-10;
  }
  // uniform_real_distribution(a, b) (random.h:44-47)
  __device__ double uniform(double a, double b) { return a + (b - a) * u01(); }
  // uniform_int_distribution(a, b), inclusive (random.h:39-42)
  __device__ int uniform_int(int a, int b) {
    int v = a + (int)floor(u01() * (double)(b - a + 1));
    return v > b ? b : v;
  }
  // random_in_sphere (random.cpp:42-49): rejection from the [-1,1)^3 cube
  __device__ void in_sphere(double* v) {
    for (;;) {
      uint32_t w[4];
      block(w);
      for (int d = 0; d < 3; ++d) v[d] = (double)w[d] * 4.656612873077393e-10 - 1.0;
      if (v[0] * v[0] + v[1] * v[1] + v[2] * v[2] < 1.0) return;
    }
  }
  // random_orientation (random.cpp:29-40)
  __device__ void quaternion(double* q) {
    uint32_t w[4];
    block(w);
    double u1 = (double)w[0] * 2.3283064365386963e-10;
    double u2 = (double)w[1] * 2.3283064365386963e-10 * 6.283185307179586;
    double u3 = (double)w[2] * 2.3283064365386963e-10 * 6.283185307179586;
    double s1 = sqrt(1.0 - u1), s2 = sqrt(u1);
    q[0] = s1 * sin(u2);
    q[1] = s1 * cos(u2);
    q[2] = s2 * sin(u3);
    q[3] = s2 * cos(u3);
  }
  // standard normal (Box-Muller) for the Gaussian mutation mode
  __device__ double normal() {
    uint32_t w[4];
    block(w);
    double u1 = ((double)w[0] + 1.0) * 2.3283064365386963e-10;
    double u2 = (double)w[1] * 2.3283064365386963e-10;
    return sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
  }
};

// math.cpp:22-42
__device__ inline void quat_to_rot(const double* q, float* m) {
  double aa = q[0] * q[0], ab = q[0] * q[1], ac = q[0] * q[2], ad = q[0] * q[3];
  double bb = q[1] * q[1], bc = q[1] * q[2], bd = q[1] * q[3];
  double cc = q[2] * q[2], cd = q[2] * q[3], dd = q[3] * q[3];
  m[0] = (float)(aa + bb - cc - dd);
  m[1] = (float)(2 * (bc - ad));
  m[2] = (float)(2 * (ac + bd));
  m[3] = (float)(2 * (ad + bc));
  m[4] = (float)(aa - bb + cc - dd);
  m[5] = (float)(2 * (cd - ab));
  m[6] = (float)(2 * (bd - ac));
  m[7] = (float)(2 * (ab + cd));
  m[8] = (float)(aa - bb - cc + dd);
}

// ---------------------------------------------------------------- apply_parameters

struct ApplyLds {
  float* cp;    // [kMaxAtoms*3] permuted input
  float* loc;   // [kMaxAtoms*3] local coordinates
  float* M;     // [kMaxFrags*9]
  float* O;     // [kMaxFrags*3]
  float* ax;    // [kMaxFrags*3]
  float* th;    // [kMaxFrags]
};

#define PXD_APPLY_LDS_FLOATS (pxd::kMaxAtoms * 6 + pxd::kMaxFrags * 16)

// apply_parameters scratch for ligands of up to ma atoms and mf fragments
__host__ __device__ constexpr int apply_lds_floats(int ma, int mf) { return 6 * ma + 16 * mf; }

__device__ inline ApplyLds carve_apply_lds(float* b, int ma, int mf) {
  ApplyLds a;
  a.cp = b;
  a.loc = a.cp + ma * 3;
  a.M = a.loc + ma * 3;
  a.O = a.M + mf * 9;
  a.ax = a.O + mf * 3;
  a.th = a.ax + mf * 3;
  return a;
}

__device__ inline ApplyLds carve_apply_lds(float* b) {
  return carve_apply_lds(b, kMaxAtoms, kMaxFrags);
}

__device__ inline void mat3mul(const float* a, const float* b, float* o) {
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j)
      o[3 * i + j] = a[3 * i] * b[j] + a[3 * i + 1] * b[3 + j] + a[3 * i + 2] * b[6 + j];
}

// free_ligand::apply_parameters (data.cpp:522-641) for one ligand on one wave.
// xyz (original atom order, LDS) is replaced by the new conformation.
// tors[t] is the rotation of torsion t (pre-order); a.th must have been
// filled by the caller through set_torsion / zero_torsions.
__device__ inline void apply_ligand(const DevLigands& L, const LigInfo& li, float* xyz,
                                    const float* center, const float* rot, ApplyLds& a) {
  const int lane = lane_id();
  const int n = li.n_atoms, nf = li.n_frags;
  const int* F = L.frag + 6 * li.frag_off;  // start, end, parent, bond_parent, bond_mine
  const int* perm = L.perm + li.atom_off;
  for (int p = lane; p < n; p += kWave) {
    int o = perm[p];
    a.cp[3 * p] = xyz[3 * o];
    a.cp[3 * p + 1] = xyz[3 * o + 1];
    a.cp[3 * p + 2] = xyz[3 * o + 2];
  }
  __syncthreads();
  // rotatable-bond axes in the input frame (data.cpp:530-536)
  for (int f = lane; f < nf; f += kWave) {
    if (F[6 * f + 2] < 0) continue;
    int pa = F[6 * f + 3], mi = F[6 * f + 4];
    float v[3] = {a.cp[3 * mi] - a.cp[3 * pa], a.cp[3 * mi + 1] - a.cp[3 * pa + 1],
                  a.cp[3 * mi + 2] - a.cp[3 * pa + 2]};
    float nn = sqrtf(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    a.ax[3 * f] = v[0] / nn;
    a.ax[3 * f + 1] = v[1] / nn;
    a.ax[3 * f + 2] = v[2] / nn;
  }
  // geometric centre (data.cpp:606)
  float c[3] = {0.f, 0.f, 0.f};
  for (int p = lane; p < n; p += kWave) {
    c[0] += a.cp[3 * p];
    c[1] += a.cp[3 * p + 1];
    c[2] += a.cp[3 * p + 2];
  }
  for (int d = 0; d < 3; ++d) c[d] = wave_sum(c[d]) / (float)n;
  // local coordinates (data.cpp:538-554, 605-615)
  for (int p = lane; p < n; p += kWave) {
    int f = 0;
    while (f < nf - 1 && !(p >= F[6 * f] && p < F[6 * f + 1])) ++f;
    int par = F[6 * f + 2];
    const float* ref;
    if (par < 0) {
      ref = c;
    } else {
      int mi = F[6 * f + 4];
      if (p != mi) {
        ref = a.cp + 3 * mi;
      } else {
        ref = F[6 * par + 2] < 0 ? c : a.cp + 3 * F[6 * par + 4];
      }
    }
    a.loc[3 * p] = a.cp[3 * p] - ref[0];
    a.loc[3 * p + 1] = a.cp[3 * p + 1] - ref[1];
    a.loc[3 * p + 2] = a.cp[3 * p + 2] - ref[2];
  }
  __syncthreads();
  // fragment frames, sequential in pre-order on one lane (data.cpp:556-590)
  if (lane == 0) {
    for (int f = 0; f < nf; ++f)
      if (F[6 * f + 2] < 0) {
        for (int k = 0; k < 9; ++k) a.M[9 * f + k] = rot[k];
        for (int k = 0; k < 3; ++k) a.O[3 * f + k] = center[k];
      }
    const int* tf = L.tors_frag + li.frag_off;
    for (int k = 0; k < li.n_tors; ++k) {
      int f = tf[k], par = F[6 * f + 2], mi = F[6 * f + 4];
      float axp[3], y[3], R[9];
      mv3(a.M + 9 * par, a.ax + 3 * f, axp);
      mv3(a.M + 9 * par, a.loc + 3 * mi, y);
      for (int d = 0; d < 3; ++d) a.O[3 * f + d] = y[d] + a.O[3 * par + d];
      rot_axis(axp[0], axp[1], axp[2], a.th[f], R);
      mat3mul(R, a.M + 9 * par, a.M + 9 * f);
    }
  }
  __syncthreads();
  // every atom from its fragment frame; a bond ("mine") atom is its own origin
  for (int p = lane; p < n; p += kWave) {
    int f = 0;
    while (f < nf - 1 && !(p >= F[6 * f] && p < F[6 * f + 1])) ++f;
    float y[3];
    if (F[6 * f + 2] >= 0 && p == F[6 * f + 4]) {
      y[0] = a.O[3 * f];
      y[1] = a.O[3 * f + 1];
      y[2] = a.O[3 * f + 2];
    } else {
      mv3(a.M + 9 * f, a.loc + 3 * p, y);
      y[0] += a.O[3 * f];
      y[1] += a.O[3 * f + 1];
      y[2] += a.O[3 * f + 2];
    }
    int o = perm[p];
    xyz[3 * o] = y[0];
    xyz[3 * o + 1] = y[1];
    xyz[3 * o + 2] = y[2];
  }
  __syncthreads();
}

// ---------------------------------------------------------------- rigid-body + torsion variables
// Chain rule of a pose's Cartesian gradient g onto translation, rotation and
// ligand torsions: out[0..2] = sum g, out[3..5] = sum (x - cen) x g, out[6 + k]
// = axis_k . sum over the atoms torsion k moves of (x - x_mine) x g, axis_k the
// unit bond vector parent-side atom -> moving-side atom in the current pose
// (free_ligand::apply_parameters rotates the subtree of a fragment about it,
// data.cpp:556-590).  tq: LDS scratch [kMaxFrags*3]; out: LDS [6 + kMaxFrags].
__device__ inline void dof_grad(const DevLigands& L, const LigInfo& li, const float* xyz,
                                const float* g, const float* cen, float* tq, float* out) {
  const int lane = lane_id();
  const int n = li.n_atoms, nf = li.n_frags;
  const int* F = L.frag + 6 * li.frag_off;
  const int* perm = L.perm + li.atom_off;
  for (int k = lane; k < 3 * nf; k += kWave) tq[k] = 0.f;
  __syncthreads();
  float fs[3] = {0.f, 0.f, 0.f}, ts[3] = {0.f, 0.f, 0.f};
  for (int p = lane; p < n; p += kWave) {
    const int a = perm[p];
    const float xa[3] = {xyz[3 * a], xyz[3 * a + 1], xyz[3 * a + 2]};
    const float ga[3] = {g[3 * a], g[3 * a + 1], g[3 * a + 2]};
    float r[3] = {xa[0] - cen[0], xa[1] - cen[1], xa[2] - cen[2]}, c[3];
    cross3(r, ga, c);
    for (int d = 0; d < 3; ++d) {
      fs[d] += ga[d];
      ts[d] += c[d];
    }
    int f = 0;
    while (f < nf - 1 && !(p >= F[6 * f] && p < F[6 * f + 1])) ++f;
    for (; F[6 * f + 2] >= 0; f = F[6 * f + 2]) {
      const int om = perm[F[6 * f + 4]];
      float rr[3] = {xa[0] - xyz[3 * om], xa[1] - xyz[3 * om + 1], xa[2] - xyz[3 * om + 2]};
      cross3(rr, ga, c);
      atomicAdd(&tq[3 * f], c[0]);
      atomicAdd(&tq[3 * f + 1], c[1]);
      atomicAdd(&tq[3 * f + 2], c[2]);
    }
  }
  for (int d = 0; d < 3; ++d) {
    fs[d] = wave_sum(fs[d]);
    ts[d] = wave_sum(ts[d]);
  }
  if (lane == 0)
    for (int d = 0; d < 3; ++d) {
      out[d] = fs[d];
      out[3 + d] = ts[d];
    }
  __syncthreads();
  const int* tf = L.tors_frag + li.frag_off;
  for (int k = lane; k < li.n_tors; k += kWave) {
    const int f = tf[k];
    const int am = perm[F[6 * f + 4]], ap = perm[F[6 * f + 3]];
    float ax[3] = {xyz[3 * am] - xyz[3 * ap], xyz[3 * am + 1] - xyz[3 * ap + 1],
                   xyz[3 * am + 2] - xyz[3 * ap + 2]};
    float inv = 1.f / sqrtf(dot3(ax, ax));
    out[6 + k] = (ax[0] * tq[3 * f] + ax[1] * tq[3 * f + 1] + ax[2] * tq[3 * f + 2]) * inv;
  }
  __syncthreads();
}

// J_l(w)^T tau for the rotation vector w (left Jacobian of SO(3)):
// tau - a (w x tau) + b (w x (w x tau)), a = (1 - cos t)/t^2, b = (t - sin t)/t^3.
__device__ inline void so3_jlt(const float* w, const float* tau, float* out) {
  float t2 = dot3(w, w), a, b;
  if (t2 < 1e-8f) {
    a = 0.5f - t2 / 24.f;
    b = 1.f / 6.f - t2 / 120.f;
  } else {
    float t = sqrtf(t2), sn, cs;
    sincosf(t, &sn, &cs);
    a = (1.f - cs) / t2;
    b = (t - sn) / (t2 * t);
  }
  float c1[3], c2[3];
  cross3(w, tau, c1);
  cross3(w, c1, c2);
  for (int d = 0; d < 3; ++d) out[d] = tau[d] - a * c1[d] + b * c2[d];
}

// Rotation matrix of the rotation vector w (identity at w = 0).
__device__ inline void rotvec_matrix(const float* w, float* R) {
  float th = sqrtf(dot3(w, w));
  if (th > 0.f) {
    float inv = 1.f / th;
    rot_axis(w[0] * inv, w[1] * inv, w[2] * inv, th, R);
  } else {
    for (int k = 0; k < 9; ++k) R[k] = (k % 4 == 0) ? 1.f : 0.f;
  }
}

// check_atom_in_box (grid.h:42-46) for every ligand atom / the centroid.
struct BoxF {
  float lo[3], hi[3];
};

__device__ inline bool all_in_box(const float* xyz, int n, const BoxF& b) {
  const int lane = lane_id();
  int bad = 0;
  for (int a = lane; a < n; a += kWave)
    for (int d = 0; d < 3; ++d)
      if (!(xyz[3 * a + d] > b.lo[d] && xyz[3 * a + d] < b.hi[d])) bad = 1;
  return wave_max((float)bad) == 0.f;
}

__device__ inline void centroid(const float* xyz, int n, float* c) {
  const int lane = lane_id();
  c[0] = c[1] = c[2] = 0.f;
  for (int a = lane; a < n; a += kWave) {
    c[0] += xyz[3 * a];
    c[1] += xyz[3 * a + 1];
    c[2] += xyz[3 * a + 2];
  }
  for (int d = 0; d < 3; ++d) c[d] = wave_sum(c[d]) / (float)n;
}

__device__ inline bool point_in_box(const float* c, const BoxF& b) {
  for (int d = 0; d < 3; ++d)
    if (!(c[d] > b.lo[d] && c[d] < b.hi[d])) return false;
  return true;
}

// check_ligand_xyz (searcher.cpp:89-96)
__device__ inline bool coords_sane(const float* xyz, int n) {
  const int lane = lane_id();
  int bad = 0;
  for (int a = lane; a < 3 * n; a += kWave)
    if (!(xyz[a] <= 1e8f && xyz[a] >= -1e8f)) bad = 1;
  return wave_max((float)bad) == 0.f;
}

// The vdW leaf of pose_selection alone (mm_interaction_vdw_energy::report,
// scorer.cpp:222-252) against the receptor's initial coordinates, as used by
// randomize (searcher.cpp:75-77).
__device__ inline double pscore_vdw_initial(const DevReceptor& R, const DevLigands& L,
                                            const LigInfo& li, const pxd_score_params& sp,
                                            const PoseLds& s) {
  const int lane = lane_id();
  const float cut = (float)sp.vdw_cutoff;
  double ev = 0.0;
  for (int h = 0; h < 2; ++h) {
    int a = lane + kWave * h;
    if (a >= li.n_atoms) continue;
    int ty = L.type[li.atom_off + a];
    float ls = s.lj[a].x * (float)sp.vdw_sigma_scale[ty];
    float le = s.lj[a].y * sqrtf((float)sp.vdw_epsilon_scale[ty]);
    float px = s.xyz[3 * a], py = s.xyz[3 * a + 1], pz = s.xyz[3 * a + 2];
    for (int j = 0; j < R.n_atoms; ++j) {
      float4 v = R.xyzq[j];
      float2 lj = R.lj[j];
      float dx = v.x - px, dy = v.y - py, dz = v.z - pz;
      float d = sqrtf(dx * dx + dy * dy + dz * dz);
      float u = safe_div((ls + lj.x) * 0.5f, d);
      float u2 = u * u;
      float u6 = u2 * u2 * u2;
      float e = (u6 * u6 - u6) * (le * lj.y * 4.f);
      ev += (double)fminf(fmaxf(e, -cut), cut);
    }
  }
  return wave_sum(ev) * sp.vdw_scale;
}

}  // namespace pxd
