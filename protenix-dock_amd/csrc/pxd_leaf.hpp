// pxd_leaf.hpp — receptor leaf energy tables.
//
// A receptor leaf rotates its movable atoms about a fixed axis
// (torsional_receptor::apply_parameters, data.cpp:366-385).  Every receptor
// term whose movable atoms all belong to one leaf k (bonds, angles,
// dihedrals, non-bonded pairs against fixed atoms or atoms of the same leaf,
// interaction.cpp:27-168 on the pruned receptor lists) therefore depends on
// theta_k alone, so their sum T_k is a 2*pi-periodic function of one angle.
// It is tabulated once per receptor in FP64 (value and derivative) and
// looked up by cubic Hermite interpolation, which replaces thousands of
// pair evaluations per pose evaluation by two 16-byte loads per leaf.
// dT_k/dtheta equals the chain-rule sum of receptor_xyz2torsion
// (interaction.cpp:338-350) over those terms.
#pragma once

#include "pxd_device.hpp"

namespace pxd {

// ---------------------------------------------------------------- FP64 term helpers

__device__ inline double dsafe_div(double a, double b) { return a / (b + 1e-20); }

__device__ inline void dcross(const double* a, const double* b, double* o) {
  o[0] = a[1] * b[2] - a[2] * b[1];
  o[1] = a[2] * b[0] - a[0] * b[2];
  o[2] = a[0] * b[1] - a[1] * b[0];
}

__device__ inline double ddot(const double* a, const double* b) {
  return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}

// math.cpp:96-110
__device__ inline double dangle_grad(const double* pi, const double* pj, const double* pk,
                                     double* gi, double* gk) {
  double v0[3] = {pj[0] - pi[0], pj[1] - pi[1], pj[2] - pi[2]};
  double v1[3] = {pj[0] - pk[0], pj[1] - pk[1], pj[2] - pk[2]};
  double cr[3];
  dcross(v0, v1, cr);
  double rp = sqrt(ddot(cr, cr));
  dcross(v0, cr, gi);
  dcross(cr, v1, gk);
  double s0 = -dsafe_div(1.0, ddot(v0, v0) * rp);
  double s1 = -dsafe_div(1.0, ddot(v1, v1) * rp);
  for (int d = 0; d < 3; ++d) {
    gi[d] *= s0;
    gk[d] *= s1;
  }
  return atan2(rp, ddot(v0, v1) + 1e-20);
}

// math.cpp:130-167
__device__ inline double ddihedral_grad(const double* pi, const double* pj, const double* pk,
                                        const double* pl, double* gi, double* gj, double* gk,
                                        double* gl) {
  double rij[3] = {pj[0] - pi[0], pj[1] - pi[1], pj[2] - pi[2]};
  double rkj[3] = {pj[0] - pk[0], pj[1] - pk[1], pj[2] - pk[2]};
  double rkl[3] = {pl[0] - pk[0], pl[1] - pk[1], pl[2] - pk[2]};
  double m[3], n[3], w[3];
  dcross(rij, rkj, m);
  dcross(rkj, rkl, n);
  dcross(m, n, w);
  double phi = atan2(sqrt(ddot(w, w)), ddot(m, n) + 1e-20);
  double theta = ddot(rij, n) < 0.0 ? phi : -phi;
  double nrkj2 = ddot(rkj, rkj);
  double nrkj_1 = dsafe_div(1.0, sqrt(nrkj2));
  double nrkj_2 = nrkj_1 * nrkj_1;
  double nrkj = nrkj2 * nrkj_1;
  double a = -nrkj / ddot(m, m);
  double b = nrkj / ddot(n, n);
  double p = ddot(rij, rkj) * nrkj_2;
  double q = ddot(rkj, rkl) * nrkj_2;
  for (int d = 0; d < 3; ++d) {
    gi[d] = -a * m[d];
    gl[d] = -b * n[d];
  }
  for (int d = 0; d < 3; ++d) {
    double sv = p * gi[d] - q * gl[d];
    gj[d] = -(gi[d] - sv);
    gk[d] = -(gl[d] + sv);
  }
  return theta;
}

// pair.h:25-60 with r = x_j - x_i; gradient on i is -r * de_dr_r
__device__ inline double dpair_energy(const double* r, double qq, double sig, double e4,
                                      double* de_dr_r) {
  double d2 = ddot(r, r);
  double d = sqrt(d2);
  double ec = dsafe_div(qq, d) * 332.0637141491396;
  double dc = -dsafe_div(qq, d2 * d) * 332.0637141491396;
  double u = dsafe_div(sig, d);
  double u2 = u * u;
  double u6 = u2 * u2 * u2;
  double u12 = u6 * u6;
  double dv = (-12.0 * u12 + 6.0 * u6) * dsafe_div(1.0, d2) * e4;
  *de_dr_r = dc + dv;
  return ec + (u12 - u6) * e4;
}

// interaction.cpp:76-102
__device__ inline double dfourier(uint32_t per, const float4* kp, double theta, double* dedt) {
  float4 a = kp[0], b = kp[1], c = kp[2];
  double k[6] = {a.x, a.y, a.z, a.w, b.x, b.y};
  double ph[6] = {b.z, b.w, c.x, c.y, c.z, c.w};
  double e = 0.0, de = 0.0;
  for (int n = 0; n < PXD_DIH_TERMS; ++n) {
    int pn = (per >> (4 * n)) & 15;
    if (pn == 0) break;
    double dl = theta * pn - ph[n];
    e += k[n] * (1.0 + cos(dl));
    de += k[n] * pn * sin(dl);
  }
  *dedt = de;
  return e;
}

// Position and d(position)/d(theta) of a term atom when leaf `leaf` is at
// theta and every other atom is at its initial coordinates.
struct LeafFrame {
  double o[3], m[9], j[9];
};

__device__ inline void leaf_frame(const DevReceptor& R, int leaf, double theta, LeafFrame& f) {
  float4 o = R.leaf_origin[leaf], ax = R.leaf_axis[leaf];
  f.o[0] = o.x;
  f.o[1] = o.y;
  f.o[2] = o.z;
  double kx = ax.x, ky = ax.y, kz = ax.z, s, c;
  sincos(theta, &s, &c);
  double t = 1.0 - c;
  // math.cpp:62-94
  f.m[0] = c + kx * kx * t;
  f.m[1] = kx * ky * t - kz * s;
  f.m[2] = ky * s + kx * kz * t;
  f.m[3] = kz * s + kx * ky * t;
  f.m[4] = c + ky * ky * t;
  f.m[5] = ky * kz * t - kx * s;
  f.m[6] = kx * kz * t - ky * s;
  f.m[7] = kx * s + ky * kz * t;
  f.m[8] = c + kz * kz * t;
  f.j[0] = -s + kx * kx * s;
  f.j[1] = -kz * c + kx * ky * s;
  f.j[2] = ky * c + kx * kz * s;
  f.j[3] = kz * c + kx * ky * s;
  f.j[4] = -s + ky * ky * s;
  f.j[5] = -kx * c + ky * kz * s;
  f.j[6] = -ky * c + kx * kz * s;
  f.j[7] = kx * c + ky * kz * s;
  f.j[8] = -s + kz * kz * s;
}

// Returns true for a movable reference (jac filled), false for a fixed atom.
__device__ inline bool leaf_pos(const DevReceptor& R, const LeafFrame& f, int ref, double* p,
                                double* jac) {
  if (ref >= 0) {
    float4 v = R.xyzq[ref];
    p[0] = v.x;
    p[1] = v.y;
    p[2] = v.z;
    return false;
  }
  float4 v = R.xyzq[R.mov_atom[-ref - 1]];
  double t[3] = {v.x - f.o[0], v.y - f.o[1], v.z - f.o[2]};
  for (int d = 0; d < 3; ++d) {
    p[d] = f.m[3 * d] * t[0] + f.m[3 * d + 1] * t[1] + f.m[3 * d + 2] * t[2] + f.o[d];
    jac[d] = f.j[3 * d] * t[0] + f.j[3 * d + 1] * t[1] + f.j[3 * d + 2] * t[2];
  }
  return true;
}

// T_k(theta) and dT_k/dtheta for one leaf, summed over the wave (every lane
// returns the total).
__device__ inline void leaf_energy(const DevReceptor& R, int k, double theta, double* e_out,
                                   double* d_out) {
  const int lane = lane_id();
  LeafFrame f;
  leaf_frame(R, k, theta, f);
  double e = 0.0, de = 0.0;
  for (int b = R.bond_seg[k] + lane; b < R.bond_seg[k + 1]; b += kWave) {
    int2 ab = R.bond_ab[b];
    float2 kr = R.bond_kr[b];
    double pa[3], pb[3], ja[3], jb[3];
    bool ma = leaf_pos(R, f, ab.x, pa, ja), mb = leaf_pos(R, f, ab.y, pb, jb);
    double d[3] = {pb[0] - pa[0], pb[1] - pa[1], pb[2] - pa[2]};
    double r = sqrt(ddot(d, d));
    double dl = r - (double)kr.y;
    e += 0.5 * kr.x * dl * dl;
    double g = kr.x * (1.0 - dsafe_div(kr.y, r));
    if (ma) de -= g * ddot(d, ja);
    if (mb) de += g * ddot(d, jb);
  }
  for (int a = R.angle_seg[k] + lane; a < R.angle_seg[k + 1]; a += kWave) {
    int4 id = R.angle_abc[a];
    float2 kt = R.angle_kt[a];
    double pi[3], pj[3], pk[3], ji[3], jj[3], jk[3], fi[3], fk[3];
    bool mi = leaf_pos(R, f, id.x, pi, ji), mj = leaf_pos(R, f, id.y, pj, jj),
         mk = leaf_pos(R, f, id.z, pk, jk);
    double th = dangle_grad(pi, pj, pk, fi, fk);
    double dl = th - (double)kt.y;
    e += 0.5 * kt.x * dl * dl;
    double g = kt.x * dl;
    if (mi) de += g * ddot(fi, ji);
    if (mk) de += g * ddot(fk, jk);
    if (mj) de -= g * (ddot(fi, jj) + ddot(fk, jj));
  }
  for (int t = R.dih_seg[k] + lane; t < R.dih_seg[k + 1]; t += kWave) {
    int4 id = R.dih_abcd[t];
    int refs[4] = {id.x, id.y, id.z, id.w};
    double p[4][3], jv[4][3], g[4][3];
    bool mv[4];
    for (int q = 0; q < 4; ++q) mv[q] = leaf_pos(R, f, refs[q], p[q], jv[q]);
    double th = ddihedral_grad(p[0], p[1], p[2], p[3], g[0], g[1], g[2], g[3]);
    double dedt;
    e += dfourier(R.dih_per[t], R.dih_kp + 3 * t, th, &dedt);
    for (int q = 0; q < 4; ++q)
      if (mv[q]) de -= dedt * ddot(g[q], jv[q]);
  }
  for (int q = R.mm_seg[k] + lane; q < R.mm_seg[k + 1]; q += kWave) {
    int4 mm = R.mm[q];
    double cs = mm.z ? 0.833333333 : 1.0, vs = mm.z ? 0.5 : 1.0;
    int ia = R.mov_atom[mm.x], ib = R.mov_atom[mm.y];
    float4 qa = R.xyzq[ia], qb = R.xyzq[ib];
    float2 la = R.lj[ia], lb = R.lj[ib];
    double pa[3], pb[3], ja[3], jb[3];
    leaf_pos(R, f, mov_ref(mm.x), pa, ja);
    leaf_pos(R, f, mov_ref(mm.y), pb, jb);
    double r[3] = {pb[0] - pa[0], pb[1] - pa[1], pb[2] - pa[2]}, g;
    e += dpair_energy(r, (double)qa.w * qb.w * cs, ((double)la.x + lb.x) * 0.5,
                      (double)la.y * lb.y * 4.0 * vs, &g);
    de += g * (ddot(r, jb) - ddot(r, ja));
  }
  for (int m = R.leaf_slot[k]; m < R.leaf_slot[k + 1]; ++m) {
    double pm[3], jm[3];
    leaf_pos(R, f, mov_ref(m), pm, jm);
    int atom = R.mov_atom[m];
    double qm = R.xyzq[atom].w;
    float2 lm = R.lj[atom];
    for (int p = R.rp_seg[m] + lane; p < R.rp_seg[m + 1]; p += kWave) {
      float4 pj = R.rp_pos[p];
      float2 lj = R.rp_lj[p];
      double r[3] = {pj.x - pm[0], pj.y - pm[1], pj.z - pm[2]}, g;
      e += dpair_energy(r, qm * pj.w, ((double)lm.x + lj.x) * 0.5, (double)lm.y * lj.y, &g);
      de -= g * ddot(r, jm);
    }
  }
  *e_out = wave_sum(e);
  *d_out = wave_sum(de);
}

}  // namespace pxd
