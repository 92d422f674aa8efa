// pxd_pscore.hpp — the empirical pose score (pose_selection tree) for one pose
// per wavefront.  Lanes own ligand-side items (atoms, donors, rings);
// receptor-side items are walked in a uniform loop (broadcast loads), so no
// pair-distance matrix is materialised (the reference allocates the full
// N_lig x N_rec map per call, scorer.cpp:176-205).  Each leaf is
// accumulated per lane in FP64.
#pragma once

#include "pxd_device.hpp"

namespace pxd {

struct PscoreLds {
  float* rc;   // receptor ring centroids [n_ring5 + n_ring6][3]
  float* rn;   // receptor ring normals
  float* lc;   // ligand ring centroids [n_ring5 + n_ring6][3]
  float* ln;   // ligand ring normals
  double* terms;  // [PXD_NUM_TERMS]
};

constexpr int kMaxRecRings = 256;
constexpr int kMaxLigRings = 16;

// math.h:187-215 ramps
__device__ inline float rev_linear(float v, float lo, float hi) {
  if (v < lo) return 1.f;
  if (v <= hi) return 1.f - (v - lo) / (hi - lo);
  return 0.f;
}
__device__ inline float rev_linear_c(float v, float c, float lo, float hi) {
  return v < c ? rev_linear(c - v, lo, hi) : rev_linear(v - c, lo, hi);
}

__device__ inline void rec_atom_pos(const DevReceptor& R, const PoseLds& s, int j, float* p) {
  int m = R.atom_mov_slot[j];
  if (m >= 0) {
    p[0] = s.mov[3 * m];
    p[1] = s.mov[3 * m + 1];
    p[2] = s.mov[3 * m + 2];
  } else {
    float4 v = R.xyzq[j];
    p[0] = v.x;
    p[1] = v.y;
    p[2] = v.z;
  }
}

__device__ inline float dist3(const float* a, const float* b) {
  float dx = a[0] - b[0], dy = a[1] - b[1], dz = a[2] - b[2];
  return sqrtf(dx * dx + dy * dy + dz * dz);
}

// scorer.cpp:128-162: centroid and normalised (p2 - p1) x (p2 - p3)
template <class PosFn>
__device__ inline void ring_geom(const int* ids, int size, PosFn pos, float* c, float* nrm) {
  float p[6][3];
  c[0] = c[1] = c[2] = 0.f;
  for (int i = 0; i < size; ++i) {
    pos(ids[i], p[i]);
    c[0] += p[i][0];
    c[1] += p[i][1];
    c[2] += p[i][2];
  }
  c[0] /= size;
  c[1] /= size;
  c[2] /= size;
  float e1[3] = {p[1][0] - p[0][0], p[1][1] - p[0][1], p[1][2] - p[0][2]};
  float e2[3] = {p[1][0] - p[2][0], p[1][1] - p[2][1], p[1][2] - p[2][2]};
  cross3(e1, e2, nrm);
  float nn = sqrtf(dot3(nrm, nrm));
  nrm[0] /= nn;
  nrm[1] /= nn;
  nrm[2] /= nn;
}

// scorer.cpp:615-636
__device__ inline float proj_dist(const float* n1, const float* c1, const float* c2) {
  float t[3] = {c2[0] - c1[0], c2[1] - c1[1], c2[2] - c1[2]};
  float dp = dot3(n1, t);
  return sqrtf(dot3(t, t) - dp * dp);
}

__device__ inline bool active_coef(double c) { return c > 1e-5 || c < -1e-5; }  // utility.h:26

// root_scorer::combine on the pose_selection tree (scorer.cpp:53-126) with the
// leaves of scorer.cpp:207-613.  Receptor movable atoms must already be in
// s.mov (receptor_torsion2xyz).  Returns the weighted score; unweighted leaf
// values go to ps.terms (lane 0 writes).
__device__ inline double pscore_pose(const DevReceptor& R, const DevLigands& L, const LigInfo& li,
                                     const pxd_score_params& sp, PoseLds& s, PscoreLds& ps) {
  const int lane = lane_id();
  const int n = li.n_atoms;
  double t[PXD_NUM_TERMS];
  for (int k = 0; k < PXD_NUM_TERMS; ++k) t[k] = 0.0;
  auto lpos = [&](int a, float* p) {
    p[0] = s.xyz[3 * a];
    p[1] = s.xyz[3 * a + 1];
    p[2] = s.xyz[3 * a + 2];
  };
  auto rpos = [&](int j, float* p) { rec_atom_pos(R, s, j, p); };
  // ring centroids / normals (scorer.cpp:190-203)
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
  // --- all-pair leaves: Coulomb (scorer.cpp:254-276) and clamped vdW (:222-252)
  const bool do_coul = active_coef(sp.coul_coef), do_vdw = active_coef(sp.vdw_coef);
  if (do_coul || do_vdw) {
    const float cut = (float)sp.vdw_cutoff;
    float lx[2][3], lq[2], ls[2], le[2];
    bool has[2];
    for (int h = 0; h < 2; ++h) {
      int a = lane + kWave * h;
      has[h] = a < n;
      if (has[h]) {
        lpos(a, lx[h]);
        lq[h] = s.q[a];
        int ty = L.type[li.atom_off + a];
        ls[h] = s.lj[a].x * (float)sp.vdw_sigma_scale[ty];
        // sqrt(eps * eps_scale) of the scaled type (scorer.cpp:213-217)
        le[h] = s.lj[a].y * sqrtf((float)sp.vdw_epsilon_scale[ty]);
      }
    }
    // per-lane sums in FP32 over blocks of 64 receptor atoms, each block
    // added to the FP64 totals
    double ec = 0.0, ev = 0.0;
    float ecf = 0.f, evf = 0.f;
    auto one = [&](const float* pj, float qj, float2 lj) {
      for (int h = 0; h < 2; ++h) {
        if (!has[h]) continue;
        float dx = lx[h][0] - pj[0], dy = lx[h][1] - pj[1], dz = lx[h][2] - pj[2];
        // 1/(d + eps) by one hardware reciprocal square root (saturating at
        // 1/eps for d = 0, as safe_divide, math.h:176-178)
        float rd = fminf(__builtin_amdgcn_rsqf(dx * dx + dy * dy + dz * dz), 1e20f);
        if (do_coul) ecf += lq[h] * qj * rd * kCoulomb;
        if (do_vdw) {
          float u = (ls[h] + lj.x) * 0.5f * rd;
          float u2 = u * u;
          float u6 = u2 * u2 * u2;
          float e = (u6 * u6 - u6) * (le[h] * lj.y * 4.f);
          evf += fminf(fmaxf(e, -cut), cut);
        }
      }
    };
    auto flush = [&]() {
      ec += (double)ecf;
      ev += (double)evf;
      ecf = evf = 0.f;
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
  }
  // --- hydrophobic (scorer.cpp:278-311): ligand sigma is the unscaled type sigma
  if (active_coef(sp.hydrophobic_coef)) {
    double e = 0.0;
    const float sf = (float)sp.hydro_scale_factor, dmin = (float)sp.hydro_dist_min,
                dmax = (float)sp.hydro_dist_max;
    // lanes over the receptor's hydrophobic atoms (about a thousand), the
    // ligand's in turn
    for (int a = 0; a < li.n_hydro; ++a) {
      int i = L.sets[li.hydro_off + a];
      float pi[3];
      lpos(i, pi);
      float s1 = s.lj[i].x;
      float ef = 0.f;  // terms in [0, 1]: FP32 per ligand atom and lane, then FP64
      for (int b = lane; b < R.n_hydro; b += kWave) {
        int j = R.hydro[b];
        float pj[3];
        rpos(j, pj);
        float r1 = (s1 + R.lj[j].x) * sf + dmin;
        ef += rev_linear(dist3(pi, pj), r1, r1 + dmax);
      }
      e -= (double)ef;
    }
    t[PXD_TERM_HYDROPHOBIC] = e * sp.hydro_scale;
  }
  // --- ionic (scorer.cpp:322-368)
  if (active_coef(sp.ionic_coef)) {
    double elc = 0.0, epc = 0.0;
    for (int a = lane; a < li.n_cation; a += kWave) {
      float pi[3];
      lpos(L.sets[li.cation_off + a], pi);
      for (int b = 0; b < R.n_anion; ++b) {
        float pj[3];
        rpos(R.anion[b], pj);
        elc -= (double)rev_linear(dist3(pi, pj), (float)sp.ionic_lc[0], (float)sp.ionic_lc[1]);
      }
    }
    for (int a = lane; a < li.n_anion; a += kWave) {
      float pi[3];
      lpos(L.sets[li.anion_off + a], pi);
      for (int b = 0; b < R.n_cation; ++b) {
        float pj[3];
        rpos(R.cation[b], pj);
        epc -= (double)rev_linear(dist3(pj, pi), (float)sp.ionic_pc[0], (float)sp.ionic_pc[1]);
      }
    }
    t[PXD_TERM_IONIC] = sp.ionic_lc_coef * elc + sp.ionic_pc_coef * epc;
  }
  // --- cation-pi (scorer.cpp:383-462, centroid form)
  if (active_coef(sp.pication_coef)) {
    double e4[4] = {0, 0, 0, 0};
    for (int a = lane; a < li.n_cation; a += kWave) {
      float pi[3];
      lpos(L.sets[li.cation_off + a], pi);
      for (int r = 0; r < nr5 + nr6; ++r) {
        int sub = r < nr5 ? 0 : 1;
        e4[sub] -= (double)rev_linear(dist3(ps.rc + 3 * r, pi), (float)sp.pication_range[sub][0],
                                      (float)sp.pication_range[sub][1]);
      }
    }
    for (int b = lane; b < R.n_cation; b += kWave) {
      float pj[3];
      rpos(R.cation[b], pj);
      for (int r = 0; r < nl5 + nl6; ++r) {
        int sub = r < nl5 ? 2 : 3;
        e4[sub] -= (double)rev_linear(dist3(ps.lc + 3 * r, pj), (float)sp.pication_range[sub][0],
                                      (float)sp.pication_range[sub][1]);
      }
    }
    double e = 0.0;
    for (int k = 0; k < 4; ++k) e += sp.pication_sub_coef[k] * e4[k];
    t[PXD_TERM_PICATION] = e;
  }
  // --- pi stacking f2f (centre 0) / e2f (centre 90) (scorer.cpp:651-724)
  for (int kind = 0; kind < 2; ++kind) {
    double coef = kind == 0 ? sp.f2f_coef : sp.e2f_coef;
    if (!active_coef(coef)) continue;
    const double* w = kind == 0 ? sp.f2f_sub_coef : sp.e2f_sub_coef;
    const double(*in)[6] = kind == 0 ? sp.f2f_inter : sp.e2f_inter;
    const float center = kind == 0 ? 0.f : 90.f;
    double e4[4] = {0, 0, 0, 0};
    // pairs (receptor ring i, ligand ring j): lanes over ligand rings x receptor rings
    int total = (nr5 + nr6) * (nl5 + nl6);
    for (int q = lane; q < total; q += kWave) {
      int i = q / (nl5 + nl6), j = q % (nl5 + nl6);
      bool r6 = i >= nr5, l6 = j >= nl5;
      int sub = (l6 ? 2 : 0) + (r6 ? 1 : 0);  // l5r5, l5r6, l6r5, l6r6
      const double* s6 = in[sub];
      const float* rc = ps.rc + 3 * i;
      const float* rn = ps.rn + 3 * i;
      const float* lc = ps.lc + 3 * j;
      const float* ln = ps.ln + 3 * j;
      float d = dist3(rc, lc);
      if (d > (float)s6[1]) continue;
      d = rev_linear(d, (float)s6[0], (float)s6[1]);
      float s1 = proj_dist(rn, rc, lc), s2 = proj_dist(ln, lc, rc);
      float sg = s1 < s2 ? s1 : s2;
      if (sg > (float)s6[5]) continue;
      sg = rev_linear(sg, (float)s6[4], (float)s6[5]);
      float cth = fminf(fmaxf(dot3(rn, ln), -1.f), 1.f);
      float th = acosf(cth) / 3.141592653589793f * 180.f;
      th = th < 180.f - th ? th : 180.f - th;
      th = rev_linear_c(th, center, (float)s6[2], (float)s6[3]);
      e4[sub] -= (double)(d * sg * th);
    }
    double e = 0.0;
    for (int k = 0; k < 4; ++k) e += w[k] * e4[k];
    t[kind == 0 ? PXD_TERM_F2F : PXD_TERM_E2F] = e;
  }
  // --- hydrogen bonds (scorer.cpp:486-577)
  if (active_coef(sp.hbond_coef)) {
    double e8[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    auto hb = [&](const float* dp, const float* hp, const float* ap, const double* s6) -> float {
      float dr = rev_linear_c(dist3(hp, ap), (float)s6[4], (float)s6[0], (float)s6[1]);
      if (dr < kEps) return 0.f;
      float v1[3] = {dp[0] - hp[0], dp[1] - hp[1], dp[2] - hp[2]};
      float v2[3] = {ap[0] - hp[0], ap[1] - hp[1], ap[2] - hp[2]};
      float cs = safe_div(dot3(v1, v2), sqrtf(dot3(v1, v1)) * sqrtf(dot3(v2, v2)));
      float ang = acosf(fminf(fmaxf(cs, -1.f), 1.f)) / 3.141592653589793f * 180.f;
      return dr * rev_linear_c(ang, (float)s6[5], (float)s6[2], (float)s6[3]);
    };
    // ligand donors x receptor acceptors: hldcc, hldcn, hldnc, hldnn.  Lanes
    // run over the receptor acceptors (hundreds), the ligand's few donors in
    // turn; the per-lane sums are reduced with the other terms.
    int nd = li.n_donc + li.n_donn;
    for (int a = 0; a < nd; ++a) {
      bool charged = a < li.n_donc;
      int2 dh = L.dons[(charged ? li.donc_off : li.donn_off) + (charged ? a : a - li.n_donc)];
      float dp[3], hp[3];
      lpos(dh.x, dp);
      lpos(dh.y, hp);
      for (int b = lane; b < R.n_acc_c + R.n_acc_n; b += kWave) {
        bool acc_c = b < R.n_acc_c;
        int j = acc_c ? R.acc_c[b] : R.acc_n[b - R.n_acc_c];
        int sub = (charged ? 0 : 2) + (acc_c ? 0 : 1);
        float ap[3];
        rpos(j, ap);
        e8[sub] -= (double)hb(dp, hp, ap, sp.hbond_inter[sub]);
      }
    }
    // receptor donors x ligand acceptors: hpdcc, hpdcn, hpdnc, hpdnn (lanes
    // over the receptor donors)
    int na = li.n_accc + li.n_accn;
    for (int a = 0; a < na; ++a) {
      bool acc_c = a < li.n_accc;
      int i = L.sets[(acc_c ? li.accc_off : li.accn_off) + (acc_c ? a : a - li.n_accc)];
      float ap[3];
      lpos(i, ap);
      for (int b = lane; b < R.n_don_c + R.n_don_n; b += kWave) {
        bool charged = b < R.n_don_c;
        int2 dh = charged ? R.don_c[b] : R.don_n[b - R.n_don_c];
        int sub = 4 + (charged ? 0 : 2) + (acc_c ? 0 : 1);
        float dp[3], hp[3];
        rpos(dh.x, dp);
        rpos(dh.y, hp);
        e8[sub] -= (double)hb(dp, hp, ap, sp.hbond_inter[sub]);
      }
    }
    double e = 0.0;
    for (int k = 0; k < 8; ++k) e += sp.hbond_sub_coef[k] * e8[k];
    t[PXD_TERM_HBOND] = e * sp.hbond_scale;
  }
  // --- torsion strain (scorer.cpp:579-613): one penalty per lane
  if (active_coef(sp.tstrain_coef)) {
    double e = 0.0;
    for (int p = lane; p < li.n_ts; p += kWave) {
      float4 cf = L.ts_coef[li.ts_off + p];
      int m0 = L.ts_moff[li.tsm_off + p], m1 = L.ts_moff[li.tsm_off + p + 1];
      float tmp = 0.f;
      int j = 0;
      for (; j < m1 - m0; ++j) {
        uint32_t w = L.ts_ijkl[m0 + j];
        int a0 = w & 255, a1 = (w >> 8) & 255, a2 = (w >> 16) & 255, a3 = (w >> 24) & 255;
        if (a0 == a1) break;
        float th = dihedral(s.xyz + 3 * a0, s.xyz + 3 * a1, s.xyz + 3 * a2, s.xyz + 3 * a3);
        tmp += (1.f + cosf(th)) * cf.x + (1.f + cosf(2.f * th - 3.141592653589793f)) * cf.y +
               (1.f + cosf(3.f * th)) * cf.z + (1.f + cosf(4.f * th - 3.141592653589793f)) * cf.w;
      }
      if (j > 0) e += (double)(tmp / j);
    }
    t[PXD_TERM_TSTRAIN] = e;
  }
  const double coef[PXD_NUM_TERMS] = {sp.coul_coef, sp.vdw_coef, sp.ionic_coef,
                                      sp.pication_coef, sp.f2f_coef, sp.e2f_coef,
                                      sp.hbond_coef, sp.hydrophobic_coef, sp.tstrain_coef};
  double total = 0.0;
  for (int k = 0; k < PXD_NUM_TERMS; ++k) {
    double v = wave_sum(t[k]);
    if (lane == 0 && ps.terms) ps.terms[k] = v;
    if (active_coef(coef[k])) total += coef[k] * v;
  }
  __syncthreads();
  return total;
}

}  // namespace pxd
