// pxd_pscore_grad.hpp — gradient of the pose score (pose_selection tree) with
// respect to the ligand coordinates and to the scoring parameters, for one
// pose per wavefront.  The reference scores poses but never differentiates
// the score (scorer.cpp:117-724 has no gradient); this is the analytic
// derivative of the same expressions, piecewise where the score is piecewise
// (ramps, the vdW clamp, the stacking cut-offs), used to fit the scoring
// function on the GPU.
//
// Parameter vector (PXD_NUM_SCORE_PARAMS doubles): the nine term
// coefficients in PXD_TERM_* order, then vdw_sigma_scale[99], then
// vdw_epsilon_scale[99].  The coefficient derivative is the unweighted leaf
// value for an active term and 0 for an inactive one (|coef| <= 1e-5 terms
// are not evaluated, utility.h:26).
#pragma once

#include "pxd_pscore.hpp"

namespace pxd {

struct PscoreGradLds {
  float* gx;     // [kMaxAtoms*3] dS/dx of the ligand atoms (original order)
  double* gp;    // [PXD_NUM_SCORE_PARAMS] dS/dparameter
};

// d rev_linear(v, lo, hi) / dv
__device__ inline float rev_linear_d(float v, float lo, float hi) {
  return (v < lo || v > hi) ? 0.f : -1.f / (hi - lo);
}
// d rev_linear_c(v, c, lo, hi) / dv
__device__ inline float rev_linear_c_d(float v, float c, float lo, float hi) {
  return v < c ? -rev_linear_d(c - v, lo, hi) : rev_linear_d(v - c, lo, hi);
}

__device__ inline void add3_lds(float* g, int a, const float* v, float f) {
  atomicAdd(&g[3 * a], v[0] * f);
  atomicAdd(&g[3 * a + 1], v[1] * f);
  atomicAdd(&g[3 * a + 2], v[2] * f);
}

// Ring centroid c and unit normal n = (p1-p0)x(p1-p2)/|.| (scorer.cpp:128-162)
// and the adjoint: given dS/dc and dS/dn, add dS/dp_i to the ring atoms.
template <class AddFn>
__device__ inline void ring_adjoint(const float* p0, const float* p1, const float* p2, int size,
                                    const int* ids, const float* dc, const float* dn,
                                    AddFn add) {
  const float inv = 1.f / size;
  for (int i = 0; i < size; ++i) add(ids[i], dc, inv);
  if (!dn) return;
  float e1[3] = {p1[0] - p0[0], p1[1] - p0[1], p1[2] - p0[2]};
  float e2[3] = {p1[0] - p2[0], p1[1] - p2[1], p1[2] - p2[2]};
  float m[3];
  cross3(e1, e2, m);
  float mn = sqrtf(dot3(m, m));
  float nn[3] = {m[0] / mn, m[1] / mn, m[2] / mn};
  // dn/dm = (I - n n^T)/|m|  ->  dS/dm
  float pr = dot3(dn, nn);
  float dm[3] = {(dn[0] - pr * nn[0]) / mn, (dn[1] - pr * nn[1]) / mn, (dn[2] - pr * nn[2]) / mn};
  // m = e1 x e2: dS/de1 = e2 x dm, dS/de2 = dm x e1
  float de1[3], de2[3];
  cross3(e2, dm, de1);
  cross3(dm, e1, de2);
  // e1 = p1 - p0, e2 = p1 - p2
  float g0[3] = {-de1[0], -de1[1], -de1[2]};
  float g1[3] = {de1[0] + de2[0], de1[1] + de2[1], de1[2] + de2[2]};
  float g2[3] = {-de2[0], -de2[1], -de2[2]};
  add(ids[0], g0, 1.f);
  add(ids[1], g1, 1.f);
  add(ids[2], g2, 1.f);
}

// pscore of one pose (as pscore_pose) and its gradients.  Receptor movable
// atoms must already be in s.mov.  g.gx and g.gp are zeroed here.
__device__ inline double pscore_pose_grad(const DevReceptor& R, const DevLigands& L,
                                          const LigInfo& li, const pxd_score_params& sp,
                                          PoseLds& s, PscoreLds& ps, PscoreGradLds& g) {
  const int lane = lane_id();
  const int n = li.n_atoms;
  for (int k = lane; k < 3 * n; k += kWave) g.gx[k] = 0.f;
  for (int k = lane; k < PXD_NUM_SCORE_PARAMS; k += kWave) g.gp[k] = 0.0;
  double t[PXD_NUM_TERMS];
  for (int k = 0; k < PXD_NUM_TERMS; ++k) t[k] = 0.0;
  auto lpos = [&](int a, float* p) {
    p[0] = s.xyz[3 * a];
    p[1] = s.xyz[3 * a + 1];
    p[2] = s.xyz[3 * a + 2];
  };
  auto rpos = [&](int j, float* p) { rec_atom_pos(R, s, j, p); };
  auto gadd = [&](int a, const float* v, float f) { add3_lds(g.gx, a, v, f); };
  const int nr5 = R.n_ring5, nr6 = R.n_ring6, nl5 = li.n_ring5, nl6 = li.n_ring6;
  for (int r = lane; r < nr5 + nr6; r += kWave) {
    const int* ids = r < nr5 ? R.ring5 + 5 * r : R.ring6 + 6 * (r - nr5);
    ring_geom(ids, r < nr5 ? 5 : 6, rpos, ps.rc + 3 * r, ps.rn + 3 * r);
  }
  for (int r = lane; r < nl5 + nl6; r += kWave) {
    const int* ids = r < nl5 ? L.rings + li.ring5_off + 5 * r : L.rings + li.ring6_off + 6 * (r - nl5);
    ring_geom(ids, r < nl5 ? 5 : 6, lpos, ps.lc + 3 * r, ps.ln + 3 * r);
  }
  __syncthreads();
  const double coef[PXD_NUM_TERMS] = {sp.coul_coef, sp.vdw_coef, sp.ionic_coef,
                                      sp.pication_coef, sp.f2f_coef, sp.e2f_coef,
                                      sp.hbond_coef, sp.hydrophobic_coef, sp.tstrain_coef};
  auto w_of = [&](int k) { return active_coef(coef[k]) ? (float)coef[k] : 0.f; };
  // --- Coulomb and clamped vdW: lanes own ligand atoms, gradients in registers
  const bool do_coul = active_coef(sp.coul_coef), do_vdw = active_coef(sp.vdw_coef);
  if (do_coul || do_vdw) {
    const float cut = (float)sp.vdw_cutoff;
    const float wc = w_of(PXD_TERM_COULOMB), wv = w_of(PXD_TERM_VDW) * (float)sp.vdw_scale;
    float lx[2][3], lq[2], ls[2], le[2], sg[2], es[2], ies[2];
    int ty[2];
    bool has[2];
    float gx[2][3] = {{0.f, 0.f, 0.f}, {0.f, 0.f, 0.f}};
    double dsig[2] = {0.0, 0.0}, deps[2] = {0.0, 0.0};
    for (int h = 0; h < 2; ++h) {
      int a = lane + kWave * h;
      has[h] = a < n;
      ty[h] = 0;
      if (has[h]) {
        lpos(a, lx[h]);
        lq[h] = s.q[a];
        ty[h] = L.type[li.atom_off + a];
        sg[h] = s.lj[a].x;
        es[h] = (float)sp.vdw_epsilon_scale[ty[h]];
        ies[h] = 1.f / es[h];
        ls[h] = sg[h] * (float)sp.vdw_sigma_scale[ty[h]];
        le[h] = s.lj[a].y * sqrtf(es[h]);
      }
    }
    // per-lane sums in FP32 over blocks of 64 receptor atoms, each block
    // added to the FP64 totals
    double ec = 0.0, ev = 0.0;
    float ecf = 0.f, evf = 0.f, dsf[2] = {0.f, 0.f}, def[2] = {0.f, 0.f};
    auto one = [&](const float* pj, float qj, float2 lj) {
      for (int h = 0; h < 2; ++h) {
        if (!has[h]) continue;
        float r[3] = {lx[h][0] - pj[0], lx[h][1] - pj[1], lx[h][2] - pj[2]};
        // 1/(d + eps) as in the forward pass; 1/d for the direction (0 at d = 0)
        float rd = fminf(__builtin_amdgcn_rsqf(dot3(r, r)), 1e20f);
        float rinv = rd < 1e20f ? rd : 0.f;
        float dedd = 0.f;
        if (do_coul) {
          float e = lq[h] * qj * rd * kCoulomb;
          ecf += e;
          dedd += wc * (-e * rd);
        }
        if (do_vdw) {
          float sgm = (ls[h] + lj.x) * 0.5f;
          float u = sgm * rd;
          float u2 = u * u;
          float u6 = u2 * u2 * u2;
          float u12 = u6 * u6;
          float e4 = le[h] * lj.y * 4.f;
          float e = (u12 - u6) * e4;
          evf += fminf(fmaxf(e, -cut), cut);
          if (e > -cut && e < cut) {
            // de/du = e4 (12 u^11 - 6 u^5); du/dd = -u/d; du/dsigma_scale = sigma_i / (2 d)
            float dedu_u = e4 * (12.f * u12 - 6.f * u6);  // de/du * u
            dedd += wv * (-dedu_u * rinv);
            float inv_u = u > 0.f ? __builtin_amdgcn_rcpf(u) : 0.f;
            dsf[h] += dedu_u * inv_u * sg[h] * 0.5f * rinv;
            def[h] += 0.5f * e * ies[h];
          }
        }
        float f = dedd * rinv;
        gx[h][0] += r[0] * f;
        gx[h][1] += r[1] * f;
        gx[h][2] += r[2] * f;
      }
    };
    auto flush = [&]() {
      ec += (double)ecf;
      ev += (double)evf;
      ecf = evf = 0.f;
      for (int h = 0; h < 2; ++h) {
        dsig[h] += (double)dsf[h];
        deps[h] += (double)def[h];
        dsf[h] = def[h] = 0.f;
      }
    };
    // compacted fixed atoms: one stream, the next atom's (uniform) loads
    // issued before the current atom is evaluated
    float4 vn = R.n_fixed > 0 ? R.fixed_xyzq[0] : make_float4(0.f, 0.f, 0.f, 0.f);
    float2 ln = R.n_fixed > 0 ? R.fixed_lj[0] : make_float2(0.f, 0.f);
    for (int k0 = 0; k0 < R.n_fixed; k0 += kWave) {
      const int k1 = min(R.n_fixed, k0 + kWave);
      for (int k = k0; k < k1; ++k) {
        const float4 v = vn;
        const float2 lj = ln;
        const int kn = min(k + 1, R.n_fixed - 1);
        vn = R.fixed_xyzq[kn];
        ln = R.fixed_lj[kn];
        float pj[3] = {v.x, v.y, v.z};
        one(pj, v.w, lj);
      }
      flush();
    }
    for (int m = 0; m < R.n_mov; ++m) {
      int j = R.mov_atom[m];
      one(s.mov + 3 * m, R.xyzq[j].w, R.lj[j]);
    }
    flush();
    t[PXD_TERM_COULOMB] = ec;
    t[PXD_TERM_VDW] = ev * sp.vdw_scale;
    const double wvd = do_vdw ? sp.vdw_coef * sp.vdw_scale : 0.0;
    for (int h = 0; h < 2; ++h) {
      if (!has[h]) continue;
      gadd(lane + kWave * h, gx[h], 1.f);
      if (do_vdw) {
        atomicAdd(&g.gp[PXD_NUM_TERMS + ty[h]], wvd * dsig[h]);
        atomicAdd(&g.gp[PXD_NUM_TERMS + PXD_NUM_VDW_TYPES + ty[h]], wvd * deps[h]);
      }
    }
  }
  // --- hydrophobic (scorer.cpp:278-311)
  if (active_coef(sp.hydrophobic_coef)) {
    double e = 0.0;
    const float w = w_of(PXD_TERM_HYDROPHOBIC) * (float)sp.hydro_scale;
    const float sf = (float)sp.hydro_scale_factor, dmin = (float)sp.hydro_dist_min,
                dmax = (float)sp.hydro_dist_max;
    // lanes over the receptor's hydrophobic atoms, the ligand's in turn; a
    // ligand atom's gradient is reduced across the wave and added once
    for (int a = 0; a < li.n_hydro; ++a) {
      int i = L.sets[li.hydro_off + a];
      float pi[3], gi[3] = {0.f, 0.f, 0.f};
      lpos(i, pi);
      float s1 = s.lj[i].x;
      float ef = 0.f;  // terms in [0, 1]: FP32 per ligand atom and lane, then FP64
      for (int b = lane; b < R.n_hydro; b += kWave) {
        int j = R.hydro[b];
        float pj[3];
        rpos(j, pj);
        float r1 = (s1 + R.lj[j].x) * sf + dmin;
        float r[3] = {pi[0] - pj[0], pi[1] - pj[1], pi[2] - pj[2]};
        float d = sqrtf(dot3(r, r));
        ef += rev_linear(d, r1, r1 + dmax);
        float f = -w * rev_linear_d(d, r1, r1 + dmax) * __builtin_amdgcn_rcpf(d);
        gi[0] += r[0] * f;
        gi[1] += r[1] * f;
        gi[2] += r[2] * f;
      }
      e -= (double)ef;
      for (int k = 0; k < 3; ++k) gi[k] = wave_sum(gi[k]);
      if (lane == 0) gadd(i, gi, 1.f);
    }
    t[PXD_TERM_HYDROPHOBIC] = e * sp.hydro_scale;
  }
  // --- ionic (scorer.cpp:322-368)
  if (active_coef(sp.ionic_coef)) {
    double elc = 0.0, epc = 0.0;
    const float w = w_of(PXD_TERM_IONIC);
    for (int a = lane; a < li.n_cation; a += kWave) {
      int i = L.sets[li.cation_off + a];
      float pi[3], gi[3] = {0.f, 0.f, 0.f};
      lpos(i, pi);
      const float lo = (float)sp.ionic_lc[0], hi = (float)sp.ionic_lc[1];
      const float wl = w * (float)sp.ionic_lc_coef;
      for (int b = 0; b < R.n_anion; ++b) {
        float pj[3];
        rpos(R.anion[b], pj);
        float r[3] = {pi[0] - pj[0], pi[1] - pj[1], pi[2] - pj[2]};
        float d = sqrtf(dot3(r, r));
        elc -= (double)rev_linear(d, lo, hi);
        float f = -wl * rev_linear_d(d, lo, hi) / d;
        gi[0] += r[0] * f;
        gi[1] += r[1] * f;
        gi[2] += r[2] * f;
      }
      gadd(i, gi, 1.f);
    }
    for (int a = lane; a < li.n_anion; a += kWave) {
      int i = L.sets[li.anion_off + a];
      float pi[3], gi[3] = {0.f, 0.f, 0.f};
      lpos(i, pi);
      const float lo = (float)sp.ionic_pc[0], hi = (float)sp.ionic_pc[1];
      const float wp = w * (float)sp.ionic_pc_coef;
      for (int b = 0; b < R.n_cation; ++b) {
        float pj[3];
        rpos(R.cation[b], pj);
        float r[3] = {pi[0] - pj[0], pi[1] - pj[1], pi[2] - pj[2]};
        float d = sqrtf(dot3(r, r));
        epc -= (double)rev_linear(d, lo, hi);
        float f = -wp * rev_linear_d(d, lo, hi) / d;
        gi[0] += r[0] * f;
        gi[1] += r[1] * f;
        gi[2] += r[2] * f;
      }
      gadd(i, gi, 1.f);
    }
    t[PXD_TERM_IONIC] = sp.ionic_lc_coef * elc + sp.ionic_pc_coef * epc;
  }
  // --- cation-pi (scorer.cpp:383-462)
  if (active_coef(sp.pication_coef)) {
    double e4[4] = {0, 0, 0, 0};
    const float w = w_of(PXD_TERM_PICATION);
    for (int a = lane; a < li.n_cation; a += kWave) {
      int i = L.sets[li.cation_off + a];
      float pi[3], gi[3] = {0.f, 0.f, 0.f};
      lpos(i, pi);
      for (int r = 0; r < nr5 + nr6; ++r) {
        int sub = r < nr5 ? 0 : 1;
        const float lo = (float)sp.pication_range[sub][0], hi = (float)sp.pication_range[sub][1];
        const float* c = ps.rc + 3 * r;
        float v[3] = {pi[0] - c[0], pi[1] - c[1], pi[2] - c[2]};
        float d = sqrtf(dot3(v, v));
        e4[sub] -= (double)rev_linear(d, lo, hi);
        float f = -w * (float)sp.pication_sub_coef[sub] * rev_linear_d(d, lo, hi) / d;
        gi[0] += v[0] * f;
        gi[1] += v[1] * f;
        gi[2] += v[2] * f;
      }
      gadd(i, gi, 1.f);
    }
    // receptor cations vs ligand ring centroids: d/dcentroid spreads over the ring
    for (int b = lane; b < R.n_cation; b += kWave) {
      float pj[3];
      rpos(R.cation[b], pj);
      for (int r = 0; r < nl5 + nl6; ++r) {
        int sub = r < nl5 ? 2 : 3;
        const float lo = (float)sp.pication_range[sub][0], hi = (float)sp.pication_range[sub][1];
        const float* c = ps.lc + 3 * r;
        float v[3] = {c[0] - pj[0], c[1] - pj[1], c[2] - pj[2]};
        float d = sqrtf(dot3(v, v));
        e4[sub] -= (double)rev_linear(d, lo, hi);
        float f = -w * (float)sp.pication_sub_coef[sub] * rev_linear_d(d, lo, hi) / d;
        if (f != 0.f) {
          const int size = r < nl5 ? 5 : 6;
          const int* ids = r < nl5 ? L.rings + li.ring5_off + 5 * r
                                   : L.rings + li.ring6_off + 6 * (r - nl5);
          float dc[3] = {v[0] * f, v[1] * f, v[2] * f};
          for (int q = 0; q < size; ++q) gadd(ids[q], dc, 1.f / size);
        }
      }
    }
    double e = 0.0;
    for (int k = 0; k < 4; ++k) e += sp.pication_sub_coef[k] * e4[k];
    t[PXD_TERM_PICATION] = e;
  }
  // --- pi stacking (scorer.cpp:651-724): lanes over (receptor ring, ligand ring)
  for (int kind = 0; kind < 2; ++kind) {
    const int term = kind == 0 ? PXD_TERM_F2F : PXD_TERM_E2F;
    double cf = kind == 0 ? sp.f2f_coef : sp.e2f_coef;
    if (!active_coef(cf)) continue;
    const double* wsub = kind == 0 ? sp.f2f_sub_coef : sp.e2f_sub_coef;
    const double(*in)[6] = kind == 0 ? sp.f2f_inter : sp.e2f_inter;
    const float center = kind == 0 ? 0.f : 90.f;
    const float w = w_of(term);
    double e4[4] = {0, 0, 0, 0};
    int total = (nr5 + nr6) * (nl5 + nl6);
    for (int q = lane; q < total; q += kWave) {
      int i = q / (nl5 + nl6), j = q % (nl5 + nl6);
      bool r6 = i >= nr5, l6 = j >= nl5;
      int sub = (l6 ? 2 : 0) + (r6 ? 1 : 0);
      const double* s6 = in[sub];
      const float* rc = ps.rc + 3 * i;
      const float* rn = ps.rn + 3 * i;
      const float* lc = ps.lc + 3 * j;
      const float* ln = ps.ln + 3 * j;
      float dv[3] = {lc[0] - rc[0], lc[1] - rc[1], lc[2] - rc[2]};
      float dd = sqrtf(dot3(dv, dv));
      if (dd > (float)s6[1]) continue;
      float fd = rev_linear(dd, (float)s6[0], (float)s6[1]);
      // projected offsets: s1 of lc from the receptor plane, s2 of rc from the ligand plane
      float p1 = dot3(rn, dv), p2 = -dot3(ln, dv);
      float s1 = sqrtf(fmaxf(dd * dd - p1 * p1, 0.f)), s2 = sqrtf(fmaxf(dd * dd - p2 * p2, 0.f));
      bool use1 = s1 < s2;
      float sgv = use1 ? s1 : s2;
      if (sgv > (float)s6[5]) continue;
      float fs = rev_linear(sgv, (float)s6[4], (float)s6[5]);
      float cth = fminf(fmaxf(dot3(rn, ln), -1.f), 1.f);
      float th = acosf(cth) / 3.141592653589793f * 180.f;
      bool flip = !(th < 180.f - th);
      float tha = flip ? 180.f - th : th;
      float ft = rev_linear_c(tha, center, (float)s6[2], (float)s6[3]);
      e4[sub] -= (double)(fd * fs * ft);
      // derivatives of -w_sub * fd * fs * ft * w with respect to lc and ln
      const float scale = -w * (float)wsub[sub];
      float dfd = rev_linear_d(dd, (float)s6[0], (float)s6[1]);
      float dfs = rev_linear_d(sgv, (float)s6[4], (float)s6[5]);
      float dft = rev_linear_c_d(tha, center, (float)s6[2], (float)s6[3]);
      float d_lc[3] = {0.f, 0.f, 0.f}, d_ln[3] = {0.f, 0.f, 0.f};
      // fd(dd): d dd / d lc = dv / dd
      float a1 = scale * dfd * fs * ft / dd;
      for (int k = 0; k < 3; ++k) d_lc[k] += a1 * dv[k];
      // fs(s): s^2 = |dv|^2 - (n.dv)^2
      if (dfs != 0.f && sgv > 0.f) {
        float a2 = scale * fd * dfs * ft / sgv;
        if (use1) {  // n = rn fixed: ds/dlc = (dv - p1 rn) / s
          for (int k = 0; k < 3; ++k) d_lc[k] += a2 * (dv[k] - p1 * rn[k]);
        } else {     // s2 with p2 = -ln.dv: ds/dlc = (dv - p2 (-ln)) / s, ds/dln = -p2 (-dv) / s
          for (int k = 0; k < 3; ++k) {
            d_lc[k] += a2 * (dv[k] + p2 * ln[k]);
            d_ln[k] += a2 * (p2 * dv[k]);
          }
        }
      }
      // ft(theta): theta = acos(rn.ln) in degrees (mirrored above 90)
      if (dft != 0.f && cth > -1.f && cth < 1.f) {
        float dth_dc = -180.f / 3.141592653589793f / sqrtf(1.f - cth * cth);
        float a3 = scale * fd * fs * dft * (flip ? -1.f : 1.f) * dth_dc;
        for (int k = 0; k < 3; ++k) d_ln[k] += a3 * rn[k];
      }
      const int size = j < nl5 ? 5 : 6;
      const int* ids = j < nl5 ? L.rings + li.ring5_off + 5 * j
                               : L.rings + li.ring6_off + 6 * (j - nl5);
      float p0[3], pp1[3], pp2[3];
      lpos(ids[0], p0);
      lpos(ids[1], pp1);
      lpos(ids[2], pp2);
      ring_adjoint(p0, pp1, pp2, size, ids, d_lc, d_ln, gadd);
    }
    double e = 0.0;
    for (int k = 0; k < 4; ++k) e += wsub[k] * e4[k];
    t[term] = e;
  }
  // --- hydrogen bonds (scorer.cpp:486-577)
  if (active_coef(sp.hbond_coef)) {
    double e8[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    const float w = w_of(PXD_TERM_HBOND) * (float)sp.hbond_scale;
    // value and gradient of one donor-hydrogen-acceptor triple
    auto hb = [&](const float* dp, const float* hp, const float* ap, const double* s6,
                  float* gd, float* gh, float* ga) -> float {
      gd[0] = gd[1] = gd[2] = gh[0] = gh[1] = gh[2] = ga[0] = ga[1] = ga[2] = 0.f;
      float ra[3] = {ap[0] - hp[0], ap[1] - hp[1], ap[2] - hp[2]};
      float dha = sqrtf(dot3(ra, ra));
      float dr = rev_linear_c(dha, (float)s6[4], (float)s6[0], (float)s6[1]);
      if (dr < kEps) return 0.f;
      float v1[3] = {dp[0] - hp[0], dp[1] - hp[1], dp[2] - hp[2]};
      float n1 = sqrtf(dot3(v1, v1)), n2 = dha;
      float cs = safe_div(dot3(v1, ra), n1 * n2);
      float csc = fminf(fmaxf(cs, -1.f), 1.f);
      float ang = acosf(csc) / 3.141592653589793f * 180.f;
      float fa = rev_linear_c(ang, (float)s6[5], (float)s6[2], (float)s6[3]);
      float ddr = rev_linear_c_d(dha, (float)s6[4], (float)s6[0], (float)s6[1]);
      float dfa = rev_linear_c_d(ang, (float)s6[5], (float)s6[2], (float)s6[3]);
      // distance part: d dha / d a = ra / dha
      for (int k = 0; k < 3; ++k) {
        float v = ddr * fa * ra[k] / dha;
        ga[k] += v;
        gh[k] -= v;
      }
      // angle part
      if (dfa != 0.f && cs > -1.f && cs < 1.f) {
        float dang = -180.f / 3.141592653589793f / sqrtf(1.f - cs * cs) * dr * dfa;
        for (int k = 0; k < 3; ++k) {
          float dc1 = (ra[k] / (n1 * n2) - cs * v1[k] / (n1 * n1));  // d cs / d v1
          float dc2 = (v1[k] / (n1 * n2) - cs * ra[k] / (n2 * n2));  // d cs / d ra
          gd[k] += dang * dc1;
          ga[k] += dang * dc2;
          gh[k] -= dang * (dc1 + dc2);
        }
      }
      return dr * fa;
    };
    // Lanes run over the receptor acceptors (hundreds), the ligand's few
    // donors in turn; a donor's and its hydrogen's gradients are summed in
    // registers, reduced across the wave and added to the shared array once.
    int nd = li.n_donc + li.n_donn;
    for (int a = 0; a < nd; ++a) {
      bool charged = a < li.n_donc;
      int2 dh = L.dons[(charged ? li.donc_off : li.donn_off) + (charged ? a : a - li.n_donc)];
      float dp[3], hp[3], gd[3], gh[3], ga[3];
      lpos(dh.x, dp);
      lpos(dh.y, hp);
      float sd[3] = {0.f, 0.f, 0.f}, sh[3] = {0.f, 0.f, 0.f};
      for (int b = lane; b < R.n_acc_c + R.n_acc_n; b += kWave) {
        bool acc_c = b < R.n_acc_c;
        int j = acc_c ? R.acc_c[b] : R.acc_n[b - R.n_acc_c];
        int sub = (charged ? 0 : 2) + (acc_c ? 0 : 1);
        float ap[3];
        rpos(j, ap);
        e8[sub] -= (double)hb(dp, hp, ap, sp.hbond_inter[sub], gd, gh, ga);
        const float f = -w * (float)sp.hbond_sub_coef[sub];
        for (int k = 0; k < 3; ++k) {
          sd[k] += gd[k] * f;
          sh[k] += gh[k] * f;
        }
      }
      for (int k = 0; k < 3; ++k) {
        sd[k] = wave_sum(sd[k]);
        sh[k] = wave_sum(sh[k]);
      }
      if (lane == 0) {
        gadd(dh.x, sd, 1.f);
        gadd(dh.y, sh, 1.f);
      }
    }
    int na = li.n_accc + li.n_accn;
    for (int a = 0; a < na; ++a) {  // lanes over the receptor donors
      bool acc_c = a < li.n_accc;
      int i = L.sets[(acc_c ? li.accc_off : li.accn_off) + (acc_c ? a : a - li.n_accc)];
      float ap[3], gd[3], gh[3], ga[3];
      lpos(i, ap);
      float sa[3] = {0.f, 0.f, 0.f};  // the acceptor's gradient over the receptor donors
      for (int b = lane; b < R.n_don_c + R.n_don_n; b += kWave) {
        bool charged = b < R.n_don_c;
        int2 dh = charged ? R.don_c[b] : R.don_n[b - R.n_don_c];
        int sub = 4 + (charged ? 0 : 2) + (acc_c ? 0 : 1);
        float dp[3], hp[3];
        rpos(dh.x, dp);
        rpos(dh.y, hp);
        e8[sub] -= (double)hb(dp, hp, ap, sp.hbond_inter[sub], gd, gh, ga);
        const float f = -w * (float)sp.hbond_sub_coef[sub];
        for (int k = 0; k < 3; ++k) sa[k] += ga[k] * f;
      }
      for (int k = 0; k < 3; ++k) sa[k] = wave_sum(sa[k]);
      if (lane == 0) gadd(i, sa, 1.f);
    }
    double e = 0.0;
    for (int k = 0; k < 8; ++k) e += sp.hbond_sub_coef[k] * e8[k];
    t[PXD_TERM_HBOND] = e * sp.hbond_scale;
  }
  // --- torsion strain (scorer.cpp:579-613)
  if (active_coef(sp.tstrain_coef)) {
    double e = 0.0;
    const float w = w_of(PXD_TERM_TSTRAIN);
    for (int p = lane; p < li.n_ts; p += kWave) {
      float4 cf = L.ts_coef[li.ts_off + p];
      int m0 = L.ts_moff[li.tsm_off + p], m1 = L.ts_moff[li.tsm_off + p + 1];
      int cnt = 0;
      for (; cnt < m1 - m0; ++cnt) {
        uint32_t q = L.ts_ijkl[m0 + cnt];
        if ((q & 255) == ((q >> 8) & 255)) break;
      }
      if (cnt == 0) continue;
      float tmp = 0.f;
      for (int j = 0; j < cnt; ++j) {
        uint32_t q = L.ts_ijkl[m0 + j];
        int a0 = q & 255, a1 = (q >> 8) & 255, a2 = (q >> 16) & 255, a3 = (q >> 24) & 255;
        float gg[4][3];
        float th = dihedral_grad(s.xyz + 3 * a0, s.xyz + 3 * a1, s.xyz + 3 * a2, s.xyz + 3 * a3,
                                 gg[0], gg[1], gg[2], gg[3]);
        tmp += (1.f + cosf(th)) * cf.x + (1.f + cosf(2.f * th - 3.141592653589793f)) * cf.y +
               (1.f + cosf(3.f * th)) * cf.z + (1.f + cosf(4.f * th - 3.141592653589793f)) * cf.w;
        float dth = -(sinf(th) * cf.x + 2.f * sinf(2.f * th - 3.141592653589793f) * cf.y +
                      3.f * sinf(3.f * th) * cf.z + 4.f * sinf(4.f * th - 3.141592653589793f) * cf.w);
        float f = w * dth / cnt;
        gadd(a0, gg[0], f);
        gadd(a1, gg[1], f);
        gadd(a2, gg[2], f);
        gadd(a3, gg[3], f);
      }
      e += (double)(tmp / cnt);
    }
    t[PXD_TERM_TSTRAIN] = e;
  }
  double total = 0.0;
  for (int k = 0; k < PXD_NUM_TERMS; ++k) {
    double v = wave_sum(t[k]);
    if (lane == 0) {
      if (ps.terms) ps.terms[k] = v;
      g.gp[k] = active_coef(coef[k]) ? v : 0.0;
    }
    if (active_coef(coef[k])) total += coef[k] * v;
  }
  __syncthreads();
  return total;
}

}  // namespace pxd
