// pxd_api.cpp — the C ABI (include/pxdock.h): validation, packing of the
// receptor / ligand descriptors into device arrays, and kernel launches.
// No exception crosses the boundary; failures set a thread-local message.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "pxd_internal.hpp"
#include "pxd_kernels.hpp"
#include "pxdock.h"
#include "pxdock_vdw_table.h"

using namespace pxd;

namespace {

thread_local std::string g_error;

pxd_status fail(pxd_status st, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_error = buf;
  return st;
}

#define HIP_TRY(expr)                                                              \
  do {                                                                             \
    hipError_t e_ = (expr);                                                        \
    if (e_ != hipSuccess)                                                          \
      return fail(PXD_ERR_HIP, "%s failed: %s", #expr, hipGetErrorString(e_));     \
  } while (0)

#define SIGMA_ROW(i, tag, s, e) s,
#define EPSILON_ROW(i, tag, s, e) e,
const double kGaffSigma[PXD_NUM_VDW_TYPES] = {PXD_VDW_TABLE(SIGMA_ROW)};
const double kGaffEpsilon[PXD_NUM_VDW_TYPES] = {PXD_VDW_TABLE(EPSILON_ROW)};

// Host staging buffer that lays several typed arrays into one device allocation.
struct Packer {
  std::vector<char> host;
  size_t add(const void* src, size_t bytes) {
    size_t off = (host.size() + 255) & ~size_t(255);
    host.resize(off + bytes + 1);
    if (bytes && src) memcpy(host.data() + off, src, bytes);
    return off;
  }
  template <class T>
  size_t add(const std::vector<T>& v) {
    return add(v.data(), v.size() * sizeof(T));
  }
};

template <class T>
const T* dptr(void* base, size_t off) {
  return reinterpret_cast<const T*>(static_cast<char*>(base) + off);
}

uint32_t pack_periods(const int32_t* per, int* bad) {
  uint32_t v = 0;
  for (int n = 0; n < PXD_DIH_TERMS; ++n) {
    int p = per[n];
    if (p < 0 || p > 15) *bad = 1;
    if (p == 0) break;
    v |= uint32_t(p & 15) << (4 * n);
  }
  return v;
}

void pack_kp(const double* k, const double* ph, const int32_t* per, float4* out) {
  float kk[6], pp[6];
  for (int n = 0; n < 6; ++n) {
    kk[n] = (float)k[n];
    pp[n] = (float)ph[n];
  }
  out[0] = make_float4(kk[0], kk[1], kk[2], kk[3]);
  out[1] = make_float4(kk[4], kk[5], pp[0], pp[1]);
  out[2] = make_float4(pp[2], pp[3], pp[4], pp[5]);
  (void)per;
}

// ---- host FP64 energies of receptor terms that touch no movable atom: they
// are pose-independent and enter every pose as one constant.
double h_dist(const double* a, const double* b) {
  double dx = a[0] - b[0], dy = a[1] - b[1], dz = a[2] - b[2];
  return std::sqrt(dx * dx + dy * dy + dz * dz);
}

double h_dihedral(const double* pi, const double* pj, const double* pk, const double* pl) {
  double rij[3], rkj[3], rkl[3], m[3], n[3], w[3];
  for (int d = 0; d < 3; ++d) {
    rij[d] = pj[d] - pi[d];
    rkj[d] = pj[d] - pk[d];
    rkl[d] = pl[d] - pk[d];
  }
  auto cr = [](const double* a, const double* b, double* o) {
    o[0] = a[1] * b[2] - a[2] * b[1];
    o[1] = a[2] * b[0] - a[0] * b[2];
    o[2] = a[0] * b[1] - a[1] * b[0];
  };
  cr(rij, rkj, m);
  cr(rkj, rkl, n);
  cr(m, n, w);
  double wl = std::sqrt(w[0] * w[0] + w[1] * w[1] + w[2] * w[2]);
  double phi = std::atan2(wl, m[0] * n[0] + m[1] * n[1] + m[2] * n[2] + 1e-20);
  return (rij[0] * n[0] + rij[1] * n[1] + rij[2] * n[2]) < 0 ? phi : -phi;
}

double h_pair(double qq, double sig, double e4, double d) {
  double u6 = std::pow(sig / (d + 1e-20), 6.0);
  return qq / (d + 1e-20) * 332.0637141491396 + (u6 * u6 - u6) * e4;
}

// Leaf energy tables (pxd_leaf.hpp): nodes at 2*pi*i/n, n doubled from 2048
// until cubic Hermite interpolation reproduces the directly evaluated FP64
// value and derivative at every midpoint within kTabTolE / kTabTolD
// (relative, floored at 1), else the receptor keeps direct evaluation.
constexpr double kTabTolE = 1e-8, kTabTolD = 1e-6;
constexpr int kTabMinNodes = 2048, kTabMaxNodes = 65536;

pxd_status build_leaf_tables(pxd_receptor* r) {
  const int L = r->n_leaves;
  if (L == 0) return PXD_OK;
  const double two_pi = 6.283185307179586;
  double2* buf = nullptr;
  const size_t cap = (size_t)L * kTabMaxNodes;
  HIP_TRY(hipMalloc(&buf, 2 * cap * sizeof(double2)));
  std::vector<double2> nodes, mids;
  for (int n = kTabMinNodes; n <= kTabMaxNodes; n *= 2) {
    double2* dn = buf;
    double2* dm = buf + (size_t)L * n;
    launch_leaf_table(r->dev, n, 0.0, dn, nullptr);
    launch_leaf_table(r->dev, n, 0.5, dm, nullptr);
    nodes.resize((size_t)L * n);
    mids.resize((size_t)L * n);
    hipError_t e1 = hipMemcpy(nodes.data(), dn, nodes.size() * sizeof(double2),
                              hipMemcpyDeviceToHost);
    hipError_t e2 = hipMemcpy(mids.data(), dm, mids.size() * sizeof(double2),
                              hipMemcpyDeviceToHost);
    if (e1 != hipSuccess || e2 != hipSuccess) {
      (void)hipFree(buf);
      return fail(PXD_ERR_HIP, "leaf tables: %s",
                  hipGetErrorString(e1 != hipSuccess ? e1 : e2));
    }
    const double h = two_pi / n;
    double we = 0.0, wd = 0.0;
    bool ok = true;
    for (int k = 0; k < L && ok; ++k)
      for (int i = 0; i < n; ++i) {
        double2 a = nodes[(size_t)k * n + i], b = nodes[(size_t)k * n + (i + 1) % n];
        double2 m = mids[(size_t)k * n + i];
        double eh = 0.5 * (a.x + b.x) + 0.125 * h * (a.y - b.y);
        double dh = 1.5 * (b.x - a.x) / h - 0.25 * (a.y + b.y);
        double re = std::fabs(eh - m.x) / std::max(std::fabs(m.x), 1.0);
        double rd = std::fabs(dh - m.y) / std::max(std::fabs(m.y), 1.0);
        if (!(re <= kTabTolE) || !(rd <= kTabTolD)) {
          ok = false;
          we = re;
          wd = rd;
          break;
        }
        we = std::max(we, re);
        wd = std::max(wd, rd);
      }
    r->tab_err_e = we;
    r->tab_err_d = wd;
    if (!ok) continue;
    // keep the nodes in their own allocation
    double2* tab = nullptr;
    hipError_t he = hipMalloc(&tab, (size_t)L * n * sizeof(double2));
    if (he == hipSuccess)
      he = hipMemcpy(tab, nodes.data(), (size_t)L * n * sizeof(double2), hipMemcpyHostToDevice);
    (void)hipFree(buf);
    if (he != hipSuccess) {
      if (tab) (void)hipFree(tab);
      return fail(PXD_ERR_HIP, "leaf tables: %s", hipGetErrorString(he));
    }
    r->tab_mem = tab;
    r->tab_n_built = n;
    r->dev.tab = tab;
    r->dev.tab_n = n;
    return PXD_OK;
  }
  (void)hipFree(buf);
  return PXD_OK;  // no table met the bound: direct evaluation
}

// Launch parameters chosen by pxd_autotune (or pxd_set_tuning).
struct Tuning {
  std::mutex mu;
  int grid_tpp = 8;
  int map_layout = 0;
  int lbfgs_slots_per_cu = 8;
  int score_ppb = 1;
};
Tuning g_tune;

struct LbfgsWorkspace {
  std::mutex mu;
  float* hist = nullptr;
  size_t hist_floats = 0;
  int* counter = nullptr;
  pxd_lbfgs_stats last{};  // totals of the last pxd_lbfgs_batch
  int* order = nullptr;    // poses grouped by size class
  size_t order_n = 0;
  // events of the class launches side by side (on side_streams())
  hipEvent_t fork = nullptr, join[kLbfgsClasses - 1] = {};
};
LbfgsWorkspace g_ws;

}  // namespace

extern "C" {

const char* pxd_last_error(void) { return g_error.c_str(); }

// internal: lets the other translation units record an error
pxd_status pxd_set_error_(pxd_status st, const char* msg) { return fail(st, "%s", msg); }

const char* pxd_version(void) { return "pxdock-mi355x 0.1 (gfx950)"; }

pxd_status pxd_device_count(int32_t* out) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) n = 0;
  *out = n;
  return PXD_OK;
}

void pxd_lbfgs_default_params(pxd_lbfgs_params* p, int32_t max_iter) {
  p->history = 100;
  p->max_iter = max_iter;
  p->max_eval = max_iter * 5 / 4;
  p->lr = 1.0;
  p->tol_grad = 0.5;
  p->tol_change = 1e-4;
  p->c1 = 1e-4;
  p->c2 = 0.9;
  p->ls_tol_change = 1e-9;
  p->gtd_threshold = 1e5;
  p->max_ls = 25;
  p->dof_mode = 0;
}

// grid.cpp:42-55
void pxd_describe_boundary(const double c[3], const double s[3], pxd_box* b) {
  memset(b, 0, sizeof(*b));
  for (int d = 0; d < 3; ++d) {
    double h = s[d] * 0.5;
    b->init_xyz[d] = c[d] - h;
    b->oppo_xyz[d] = c[d] + h;
  }
}

// grid.cpp:57-86
void pxd_decide_grid_dims(const double c[3], const double s[3], double spacing, pxd_box* b) {
  memset(b, 0, sizeof(*b));
  for (int d = 0; d < 3; ++d) {
    long nv = (long)std::ceil(s[d] / spacing);
    if (nv % 2 == 1) nv += 1;
    double len = nv * spacing;
    b->npoints[d] = (int32_t)(nv + 1);
    b->init_xyz[d] = c[d] - len * 0.5;
  }
  b->spacing = spacing;
  b->spacing_inv = 1.0 / spacing;
  for (int d = 0; d < 3; ++d) {
    b->npoints_minus_one[d] = b->npoints[d] - 1;
    b->oppo_xyz[d] = b->init_xyz[d] + b->npoints_minus_one[d] * spacing;
  }
}

// ------------------------------------------------------------------ receptor

pxd_status pxd_receptor_create(const pxd_mol* m, const double* xyz, const pxd_leaves* lv,
                               pxd_receptor** out) {
  if (!m || !xyz || !out) return fail(PXD_ERR_ARG, "receptor: null argument");
  *out = nullptr;
  const int n = m->n_atoms;
  if (n <= 0) return fail(PXD_ERR_ARG, "receptor: no atoms");
  if (!m->sigma || !m->epsilon || !m->charge)
    return fail(PXD_ERR_CONF, "receptor: sigma/epsilon/charge required");
  const int nl = lv ? lv->n_leaves : 0;
  if (nl > PXD_MAX_LEAVES) return fail(PXD_ERR_RANGE, "receptor: %d leaves > %d", nl, PXD_MAX_LEAVES);
  // movable slots ordered by leaf (receptor_torsion2xyz order, interaction.cpp:323-335)
  std::vector<int> slot_of(n, -1), mov_leaf, mov_atom;
  for (int k = 0; k < nl; ++k)
    for (int a = lv->atom_off[k]; a < lv->atom_off[k + 1]; ++a) {
      int j = lv->atoms[a];
      if (j < 0 || j >= n) return fail(PXD_ERR_CONF, "receptor: leaf atom %d out of range", j);
      if (slot_of[j] >= 0)
        return fail(PXD_ERR_CONF, "receptor: atom %d belongs to two leaves (unsupported)", j);
      slot_of[j] = (int)mov_atom.size();
      mov_leaf.push_back(k);
      mov_atom.push_back(j);
    }
  const int M = (int)mov_atom.size();
  if (M > kMaxMovable) return fail(PXD_ERR_RANGE, "receptor: %d movable atoms > %d", M, kMaxMovable);
  auto ref = [&](int a) { return slot_of[a] >= 0 ? mov_ref(slot_of[a]) : a; };
  auto chk = [&](const int32_t* ids, int cnt, int w) -> bool {
    for (int i = 0; i < cnt * w; ++i)
      if (ids[i] < 0 || ids[i] >= n) return false;
    return true;
  };
  if (!chk(m->bond_ij, m->n_bonds, 2) || !chk(m->angle_ijk, m->n_angles, 3) ||
      !chk(m->dih_ijkl, m->n_dihedrals, 4) || !chk(m->pairs14, m->n_pairs14, 2) ||
      !chk(m->pairs, m->n_pairs, 2))
    return fail(PXD_ERR_CONF, "receptor: term atom index out of range");

  std::vector<float4> xyzq(n);
  std::vector<float2> lj(n);
  std::vector<int> fixed;
  std::vector<float4> fxyzq;
  std::vector<float2> flj;
  std::vector<double4> fxyzq64;
  std::vector<double2> flj64;
  for (int i = 0; i < n; ++i) {
    xyzq[i] = make_float4((float)xyz[3 * i], (float)xyz[3 * i + 1], (float)xyz[3 * i + 2],
                          (float)m->charge[i]);
    lj[i] = make_float2((float)m->sigma[i], (float)std::sqrt(m->epsilon[i]));
    if (slot_of[i] < 0) {
      fixed.push_back(i);
      fxyzq.push_back(xyzq[i]);
      flj.push_back(lj[i]);
      fxyzq64.push_back(make_double4(xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2], m->charge[i]));
      flj64.push_back(make_double2(m->sigma[i], std::sqrt(m->epsilon[i])));
    }
  }
  // constant part: terms with no movable atom (FP64 on the host)
  double e_const = 0.0;
  std::vector<int2> bab;
  std::vector<float2> bkr;
  for (int b = 0; b < m->n_bonds; ++b) {
    int i = m->bond_ij[2 * b], j = m->bond_ij[2 * b + 1];
    if (slot_of[i] < 0 && slot_of[j] < 0) {
      double r = h_dist(xyz + 3 * i, xyz + 3 * j) - m->bond_r0[b];
      e_const += 0.5 * m->bond_k[b] * r * r;
      continue;
    }
    bab.push_back(make_int2(ref(i), ref(j)));
    bkr.push_back(make_float2((float)m->bond_k[b], (float)m->bond_r0[b]));
  }
  std::vector<int4> aabc;
  std::vector<float2> akt;
  for (int a = 0; a < m->n_angles; ++a) {
    const int* id = m->angle_ijk + 3 * a;
    if (slot_of[id[0]] < 0 && slot_of[id[1]] < 0 && slot_of[id[2]] < 0) {
      const double *pi = xyz + 3 * id[0], *pj = xyz + 3 * id[1], *pk = xyz + 3 * id[2];
      double v0[3], v1[3], c[3];
      for (int d = 0; d < 3; ++d) {
        v0[d] = pj[d] - pi[d];
        v1[d] = pj[d] - pk[d];
      }
      c[0] = v0[1] * v1[2] - v0[2] * v1[1];
      c[1] = v0[2] * v1[0] - v0[0] * v1[2];
      c[2] = v0[0] * v1[1] - v0[1] * v1[0];
      double th = std::atan2(std::sqrt(c[0] * c[0] + c[1] * c[1] + c[2] * c[2]),
                             v0[0] * v1[0] + v0[1] * v1[1] + v0[2] * v1[2] + 1e-20);
      double dl = th - m->angle_t0[a];
      e_const += 0.5 * m->angle_k[a] * dl * dl;
      continue;
    }
    aabc.push_back(make_int4(ref(id[0]), ref(id[1]), ref(id[2]), 0));
    akt.push_back(make_float2((float)m->angle_k[a], (float)m->angle_t0[a]));
  }
  std::vector<int4> dabcd;
  std::vector<uint32_t> dper;
  std::vector<float4> dkp;
  int bad = 0;
  for (int t = 0; t < m->n_dihedrals; ++t) {
    const int* id = m->dih_ijkl + 4 * t;
    const int32_t* per = m->dih_n + 6 * t;
    if (slot_of[id[0]] < 0 && slot_of[id[1]] < 0 && slot_of[id[2]] < 0 && slot_of[id[3]] < 0) {
      double th = h_dihedral(xyz + 3 * id[0], xyz + 3 * id[1], xyz + 3 * id[2], xyz + 3 * id[3]);
      for (int k = 0; k < 6 && per[k] != 0; ++k)
        e_const += m->dih_k[6 * t + k] * (1.0 + std::cos(th * per[k] - m->dih_phase[6 * t + k]));
      continue;
    }
    dabcd.push_back(make_int4(ref(id[0]), ref(id[1]), ref(id[2]), ref(id[3])));
    dper.push_back(pack_periods(per, &bad));
    float4 kp[3];
    pack_kp(m->dih_k + 6 * t, m->dih_phase + 6 * t, per, kp);
    dkp.insert(dkp.end(), kp, kp + 3);
  }
  if (bad) return fail(PXD_ERR_CONF, "receptor: dihedral periodicity outside 0..15");
  // non-bonded pairs
  std::vector<std::vector<float4>> seg_pos(M);
  std::vector<std::vector<float2>> seg_lj(M);
  std::vector<int4> mm;
  std::vector<float4> mm_par;
  for (int pass = 0; pass < 2; ++pass) {
    int cnt = pass == 0 ? m->n_pairs14 : m->n_pairs;
    const int32_t* ij = pass == 0 ? m->pairs14 : m->pairs;
    double cs = pass == 0 ? 0.833333333 : 1.0, vs = pass == 0 ? 0.5 : 1.0;
    for (int p = 0; p < cnt; ++p) {
      int i = ij[2 * p], j = ij[2 * p + 1];
      int si = slot_of[i], sj = slot_of[j];
      if (si < 0 && sj < 0) {
        double sig = (m->sigma[i] + m->sigma[j]) * 0.5;
        double e4 = std::sqrt(m->epsilon[i] * m->epsilon[j]) * vs * 4.0;
        e_const += h_pair(m->charge[i] * m->charge[j] * cs, sig, e4, h_dist(xyz + 3 * i, xyz + 3 * j));
      } else if (si >= 0 && sj >= 0) {
        mm.push_back(make_int4(si, sj, pass == 0 ? 1 : 0, mov_leaf[si] | mov_leaf[sj] << 16));
        mm_par.push_back(make_float4((float)(m->charge[i] * m->charge[j] * cs * 332.0637141491396),
                                     (float)((m->sigma[i] + m->sigma[j]) * 0.5),
                                     (float)(std::sqrt(m->epsilon[i] * m->epsilon[j]) * 4.0 * vs),
                                     0.f));
      } else {
        int mv = si >= 0 ? si : sj, fx = si >= 0 ? j : i;
        seg_pos[mv].push_back(make_float4((float)xyz[3 * fx], (float)xyz[3 * fx + 1],
                                          (float)xyz[3 * fx + 2], (float)(m->charge[fx] * cs)));
        seg_lj[mv].push_back(make_float2((float)m->sigma[fx],
                                         (float)(std::sqrt(m->epsilon[fx]) * 4.0 * vs)));
      }
    }
  }
  std::vector<int> rp_seg(M + 1, 0);
  std::vector<float4> rp_pos;
  std::vector<float2> rp_lj;
  for (int s = 0; s < M; ++s) {
    rp_pos.insert(rp_pos.end(), seg_pos[s].begin(), seg_pos[s].end());
    rp_lj.insert(rp_lj.end(), seg_lj[s].begin(), seg_lj[s].end());
    while (rp_pos.size() % kWave) {  // inert padding: no charge, no LJ well
      rp_pos.push_back(make_float4(1e4f, 1e4f, 1e4f, 0.f));
      rp_lj.push_back(make_float2(0.f, 0.f));
    }
    rp_seg[s + 1] = (int)rp_pos.size();
  }
  // order every term list by leaf class (one leaf k: class k; several: nl)
  auto leaf_of_ref = [&](int rf) { return rf < 0 ? mov_leaf[-rf - 1] : -1; };
  auto term_class = [&](const int* refs, int cnt) {
    int c = -1;
    for (int q = 0; q < cnt; ++q) {
      int l = leaf_of_ref(refs[q]);
      if (l < 0) continue;
      if (c < 0) c = l;
      else if (c != l) return nl;
    }
    return c < 0 ? nl : c;
  };
  auto by_class = [&](const std::vector<int>& cls, std::vector<int>& seg) {
    std::vector<int> idx(cls.size());
    for (size_t i = 0; i < idx.size(); ++i) idx[i] = (int)i;
    std::stable_sort(idx.begin(), idx.end(), [&](int a, int b) { return cls[a] < cls[b]; });
    seg.assign(nl + 2, 0);
    for (int c : cls) seg[c + 1]++;
    for (int k = 0; k <= nl; ++k) seg[k + 1] += seg[k];
    seg.pop_back();  // [nl+1]: seg[nl] is the first cross term
    return idx;
  };
  auto permute = [](auto& v, const std::vector<int>& idx, int width) {
    auto old = v;
    for (size_t i = 0; i < idx.size(); ++i)
      for (int w = 0; w < width; ++w) v[i * width + w] = old[(size_t)idx[i] * width + w];
  };
  std::vector<int> bond_seg, angle_seg, dih_seg, mm_seg, leaf_slot(nl + 1, 0);
  {
    std::vector<int> cls(bab.size());
    for (size_t i = 0; i < bab.size(); ++i) {
      int r2[2] = {bab[i].x, bab[i].y};
      cls[i] = term_class(r2, 2);
    }
    auto idx = by_class(cls, bond_seg);
    permute(bab, idx, 1);
    permute(bkr, idx, 1);
  }
  {
    std::vector<int> cls(aabc.size());
    for (size_t i = 0; i < aabc.size(); ++i) {
      int r3[3] = {aabc[i].x, aabc[i].y, aabc[i].z};
      cls[i] = term_class(r3, 3);
    }
    auto idx = by_class(cls, angle_seg);
    permute(aabc, idx, 1);
    permute(akt, idx, 1);
  }
  {
    std::vector<int> cls(dabcd.size());
    for (size_t i = 0; i < dabcd.size(); ++i) {
      int r4[4] = {dabcd[i].x, dabcd[i].y, dabcd[i].z, dabcd[i].w};
      cls[i] = term_class(r4, 4);
    }
    auto idx = by_class(cls, dih_seg);
    permute(dabcd, idx, 1);
    permute(dper, idx, 1);
    permute(dkp, idx, 3);
  }
  {
    std::vector<int> cls(mm.size());
    for (size_t i = 0; i < mm.size(); ++i) {
      int r2[2] = {mov_ref(mm[i].x), mov_ref(mm[i].y)};
      cls[i] = term_class(r2, 2);
    }
    auto idx = by_class(cls, mm_seg);
    permute(mm, idx, 1);
    permute(mm_par, idx, 1);
  }
  for (int sl = 0; sl < M; ++sl) leaf_slot[mov_leaf[sl] + 1] = sl + 1;
  for (int k = 0; k < nl; ++k) leaf_slot[k + 1] = std::max(leaf_slot[k + 1], leaf_slot[k]);
  std::vector<float4> lorg(nl), lax(nl);
  for (int k = 0; k < nl; ++k) {
    lorg[k] = make_float4((float)lv->origin[3 * k], (float)lv->origin[3 * k + 1],
                          (float)lv->origin[3 * k + 2], 0.f);
    lax[k] = make_float4((float)lv->axis[3 * k], (float)lv->axis[3 * k + 1],
                         (float)lv->axis[3 * k + 2], 0.f);
  }
  // pscore sets
  auto ivec = [](const int32_t* p, int cnt) { return std::vector<int>(p, p + cnt); };
  std::vector<int> hydro = ivec(m->hydrophobic, m->n_hydrophobic),
                   cat = ivec(m->cation, m->n_cation), an = ivec(m->anion, m->n_anion),
                   r5 = ivec(m->ring5, 5 * m->n_ring5), r6 = ivec(m->ring6, 6 * m->n_ring6),
                   accc = ivec(m->acc_charged, m->n_acc_charged),
                   accn = ivec(m->acc_neutral, m->n_acc_neutral);
  std::vector<int2> donc(m->n_don_charged), donn(m->n_don_neutral);
  for (int i = 0; i < m->n_don_charged; ++i)
    donc[i] = make_int2(m->don_charged[2 * i], m->don_charged[2 * i + 1]);
  for (int i = 0; i < m->n_don_neutral; ++i)
    donn[i] = make_int2(m->don_neutral[2 * i], m->don_neutral[2 * i + 1]);
  for (auto* v : {&hydro, &cat, &an, &r5, &r6, &accc, &accn})
    for (int a : *v)
      if (a < 0 || a >= n) return fail(PXD_ERR_CONF, "receptor: interaction-set atom out of range");

  Packer P;
  size_t o_xyzq = P.add(xyzq), o_lj = P.add(lj), o_fixed = P.add(fixed), o_fxyzq = P.add(fxyzq),
         o_flj = P.add(flj), o_fxyzq64 = P.add(fxyzq64), o_flj64 = P.add(flj64),
         o_lorg = P.add(lorg), o_lax = P.add(lax),
         o_movleaf = P.add(mov_leaf), o_movatom = P.add(mov_atom), o_bab = P.add(bab),
         o_bkr = P.add(bkr), o_aabc = P.add(aabc), o_akt = P.add(akt), o_dabcd = P.add(dabcd),
         o_dper = P.add(dper), o_dkp = P.add(dkp), o_rpseg = P.add(rp_seg),
         o_rppos = P.add(rp_pos), o_rplj = P.add(rp_lj), o_mm = P.add(mm), o_mmpar = P.add(mm_par),
         o_hydro = P.add(hydro), o_cat = P.add(cat), o_an = P.add(an), o_r5 = P.add(r5),
         o_r6 = P.add(r6), o_donc = P.add(donc), o_donn = P.add(donn), o_accc = P.add(accc),
         o_accn = P.add(accn), o_slot = P.add(slot_of), o_bseg = P.add(bond_seg),
         o_aseg = P.add(angle_seg), o_dseg = P.add(dih_seg), o_mseg = P.add(mm_seg),
         o_lslot = P.add(leaf_slot);
  pxd_receptor* r = new pxd_receptor();
  hipError_t he = hipMalloc(&r->mem, P.host.size());
  if (he != hipSuccess) {
    delete r;
    return fail(PXD_ERR_HIP, "receptor: hipMalloc: %s", hipGetErrorString(he));
  }
  he = hipMemcpy(r->mem, P.host.data(), P.host.size(), hipMemcpyHostToDevice);
  if (he != hipSuccess) {
    (void)hipFree(r->mem);
    delete r;
    return fail(PXD_ERR_HIP, "receptor: hipMemcpy: %s", hipGetErrorString(he));
  }
  r->mem_bytes = P.host.size();
  DevReceptor& D = r->dev;
  void* B = r->mem;
  D.n_atoms = n;
  D.xyzq = dptr<float4>(B, o_xyzq);
  D.lj = dptr<float2>(B, o_lj);
  D.n_fixed = (int)fixed.size();
  D.fixed = dptr<int>(B, o_fixed);
  D.n_leaves = nl;
  D.n_mov = M;
  D.leaf_origin = dptr<float4>(B, o_lorg);
  D.leaf_axis = dptr<float4>(B, o_lax);
  D.mov_leaf = dptr<int>(B, o_movleaf);
  D.mov_atom = dptr<int>(B, o_movatom);
  D.n_bonds = (int)bab.size();
  D.bond_ab = dptr<int2>(B, o_bab);
  D.bond_kr = dptr<float2>(B, o_bkr);
  D.n_angles = (int)aabc.size();
  D.angle_abc = dptr<int4>(B, o_aabc);
  D.angle_kt = dptr<float2>(B, o_akt);
  D.n_dih = (int)dabcd.size();
  D.dih_abcd = dptr<int4>(B, o_dabcd);
  D.dih_per = dptr<uint32_t>(B, o_dper);
  D.dih_kp = dptr<float4>(B, o_dkp);
  D.rp_seg = dptr<int>(B, o_rpseg);
  D.rp_pos = dptr<float4>(B, o_rppos);
  D.rp_lj = dptr<float2>(B, o_rplj);
  D.n_mm = (int)mm.size();
  D.mm = dptr<int4>(B, o_mm);
  D.mm_par = dptr<float4>(B, o_mmpar);
  D.n_hydro = (int)hydro.size();
  D.n_cation = (int)cat.size();
  D.n_anion = (int)an.size();
  D.n_ring5 = m->n_ring5;
  D.n_ring6 = m->n_ring6;
  D.n_don_c = m->n_don_charged;
  D.n_don_n = m->n_don_neutral;
  D.n_acc_c = m->n_acc_charged;
  D.n_acc_n = m->n_acc_neutral;
  D.hydro = dptr<int>(B, o_hydro);
  D.cation = dptr<int>(B, o_cat);
  D.anion = dptr<int>(B, o_an);
  D.ring5 = dptr<int>(B, o_r5);
  D.ring6 = dptr<int>(B, o_r6);
  D.don_c = dptr<int2>(B, o_donc);
  D.don_n = dptr<int2>(B, o_donn);
  D.acc_c = dptr<int>(B, o_accc);
  D.acc_n = dptr<int>(B, o_accn);
  D.atom_mov_slot = dptr<int>(B, o_slot);
  D.e_const = e_const;
  D.fixed_xyzq = dptr<float4>(B, o_fxyzq);
  D.fixed_lj = dptr<float2>(B, o_flj);
  D.fixed_xyzq64 = dptr<double4>(B, o_fxyzq64);
  D.fixed_lj64 = dptr<double2>(B, o_flj64);
  D.bond_seg = dptr<int>(B, o_bseg);
  D.angle_seg = dptr<int>(B, o_aseg);
  D.dih_seg = dptr<int>(B, o_dseg);
  D.mm_seg = dptr<int>(B, o_mseg);
  D.leaf_slot = dptr<int>(B, o_lslot);
  D.tab_n = 0;
  D.tab = nullptr;
  r->tab_mem = nullptr;
  r->tab_n_built = 0;
  r->tab_err_e = r->tab_err_d = 0.0;
  r->n_atoms = n;
  r->n_leaves = nl;
  r->n_mov = M;
  r->h_xyz = nullptr;
  r->movable.assign(mov_atom.begin(), mov_atom.end());
  std::sort(r->movable.begin(), r->movable.end());
  r->h_xyz64.assign(xyz, xyz + 3 * (size_t)n);
  r->h_q64.assign(m->charge, m->charge + n);
  r->h_mov_atom = mov_atom;
  r->h_mov_leaf = mov_leaf;
  for (int k = 0; k < nl; ++k)
    for (int d = 0; d < 3; ++d) {
      r->h_leaf_o.push_back(lv->origin[3 * k + d]);
      r->h_leaf_a.push_back(lv->axis[3 * k + d]);
    }
  (void)pxd_receptor_key(m, xyz, lv, &r->key);
  pxd_status st = build_leaf_tables(r);
  if (st != PXD_OK) {
    pxd_receptor_destroy(r);
    return st;
  }
  *out = r;
  return PXD_OK;
}

void pxd_receptor_destroy(pxd_receptor* r) {
  if (!r) return;
  (void)hipFree(r->mem);
  if (r->tab_mem) (void)hipFree(r->tab_mem);
  delete r;
}

pxd_status pxd_receptor_leaf_tables(pxd_receptor* r, int32_t enable) {
  if (!r) return fail(PXD_ERR_ARG, "leaf_tables: null receptor");
  if (enable && !r->tab_mem && r->n_leaves > 0)
    return fail(PXD_ERR_STATE, "leaf_tables: no table met the accuracy bound at build");
  r->dev.tab = enable && r->tab_mem ? static_cast<const double2*>(r->tab_mem) : nullptr;
  r->dev.tab_n = r->dev.tab ? r->tab_n_built : 0;
  return PXD_OK;
}

pxd_status pxd_receptor_info(const pxd_receptor* r, int32_t* n_leaves, int32_t* table_nodes,
                             double* table_err_e, double* table_err_d) {
  if (!r) return fail(PXD_ERR_ARG, "receptor_info: null receptor");
  if (n_leaves) *n_leaves = r->n_leaves;
  if (table_nodes) *table_nodes = r->dev.tab ? r->tab_n_built : 0;
  if (table_err_e) *table_err_e = r->tab_err_e;
  if (table_err_d) *table_err_d = r->tab_err_d;
  return PXD_OK;
}

// ------------------------------------------------------------------ grid

// the DevGrid record heads the grid allocation; maps start 256-byte aligned
constexpr size_t kGridHeader = (sizeof(DevGrid) + 255) & ~size_t(255);

static pxd_status grid_alloc(const pxd_receptor* r, const pxd_box* box, const int32_t* types,
                             int32_t n_types, double slope, pxd_grid** out) {
  if (!r || !box || !out || (n_types > 0 && !types)) return fail(PXD_ERR_ARG, "grid: null argument");
  if (n_types < 0 || n_types > PXD_NUM_VDW_TYPES) return fail(PXD_ERR_ARG, "grid: n_types=%d", n_types);
  for (int d = 0; d < 3; ++d)
    if (box->npoints[d] < 2) return fail(PXD_ERR_ARG, "grid: npoints[%d] < 2", d);
  if (!(box->spacing > 0)) return fail(PXD_ERR_ARG, "grid: spacing must be > 0");
  pxd_grid* g = new pxd_grid();
  g->box = *box;
  g->n_types = n_types;
  g->receptor = r;
  for (int t = 0; t < PXD_NUM_VDW_TYPES; ++t) g->dev.type_slot[t] = -1;
  for (int t = 0; t < n_types; ++t) {
    if (types[t] < 0 || types[t] >= PXD_NUM_VDW_TYPES) {
      delete g;
      return fail(PXD_ERR_ARG, "grid: type %d out of range", types[t]);
    }
    g->types[t] = types[t];
    g->dev.type_slot[types[t]] = t + 1;
  }
  long long P = (long long)box->npoints[0] * box->npoints[1] * box->npoints[2];
  size_t bytes = (size_t)(n_types + 1) * P * sizeof(float4);
  hipError_t he = hipMalloc(&g->mem, bytes + kGridHeader);
  if (he != hipSuccess) {
    delete g;
    return fail(PXD_ERR_HIP, "grid: hipMalloc(%zu): %s", bytes, hipGetErrorString(he));
  }
  DevGrid& D = g->dev;
  for (int d = 0; d < 3; ++d) {
    D.init[d] = (float)box->init_xyz[d];
    D.init_d[d] = box->init_xyz[d];
    D.np[d] = box->npoints[d];
  }
  D.spacing = (float)box->spacing;
  D.inv_spacing = (float)(1.0 / box->spacing);
  D.inv_spacing_d = 1.0 / box->spacing;
  D.slope = (float)slope;
  g->slope = slope;
  D.P = P;
  D.maps = reinterpret_cast<float4*>(static_cast<char*>(g->mem) + kGridHeader);
  D.layout = 0;
  D.C = (long long)(box->npoints[0] - 1) * (box->npoints[1] - 1) * (box->npoints[2] - 1);
  D.cmaps = nullptr;
  D.smaps = nullptr;
  g->cmem = nullptr;
  *out = g;
  return PXD_OK;
}

static DevGrid* grid_dev_ptr(const pxd_grid* g) { return static_cast<DevGrid*>(g->mem); }

static pxd_status grid_sync_struct(pxd_grid* g) {
  HIP_TRY(hipMemcpy(g->mem, &g->dev, sizeof(DevGrid), hipMemcpyHostToDevice));
  return PXD_OK;
}

pxd_status pxd_grid_build(const pxd_receptor* r, const pxd_box* box, const int32_t* types,
                          int32_t n_types, double slope, void* stream, pxd_grid** out) {
  pxd_grid* g = nullptr;
  pxd_status st = grid_alloc(r, box, types, n_types, slope, &g);
  if (st != PXD_OK) return st;
  const int nt1 = n_types > 0 ? n_types : 1;
  std::vector<float2> tlj(nt1);
  std::vector<double2> tlj64(nt1);
  for (int t = 0; t < n_types; ++t) {
    tlj64[t] = make_double2(kGaffSigma[types[t]], 4.0 * std::sqrt(kGaffEpsilon[types[t]]));
    tlj[t] = make_float2((float)tlj64[t].x, (float)tlj64[t].y);
  }
  // one staging allocation: FP64 type parameters (256-byte aligned), then FP32
  char* d_tlj_mem = nullptr;
  const size_t o32 = (sizeof(double2) * nt1 + 255) & ~size_t(255);
  hipError_t he = hipMalloc(&d_tlj_mem, o32 + sizeof(float2) * nt1);
  double2* d_tlj64 = reinterpret_cast<double2*>(d_tlj_mem);
  float2* d_tlj = reinterpret_cast<float2*>(d_tlj_mem + o32);
  if (he == hipSuccess)
    he = hipMemcpy(d_tlj64, tlj64.data(), sizeof(double2) * nt1, hipMemcpyHostToDevice);
  if (he == hipSuccess) he = hipMemcpy(d_tlj, tlj.data(), sizeof(float2) * nt1, hipMemcpyHostToDevice);
  if (he != hipSuccess) {
    pxd_grid_destroy(g);
    return fail(PXD_ERR_HIP, "grid: %s", hipGetErrorString(he));
  }
  GridBuildArgs a;
  a.atoms = r->dev.fixed_xyzq;
  a.alj = r->dev.fixed_lj;
  a.atoms64 = r->dev.fixed_xyzq64;
  a.alj64 = r->dev.fixed_lj64;
  a.n_atoms = r->dev.n_fixed;
  a.tlj = d_tlj;
  a.tlj64 = d_tlj64;
  a.n_types = n_types;
  for (int d = 0; d < 3; ++d) {
    a.init[d] = g->dev.init[d];
    a.init64[d] = g->box.init_xyz[d];
    a.np[d] = g->dev.np[d];
  }
  a.spacing = g->dev.spacing;
  a.spacing64 = g->box.spacing;
  a.P = g->dev.P;
  a.maps = const_cast<float4*>(g->dev.maps);
  hipStream_t s = static_cast<hipStream_t>(stream);
  int tpp, layout;
  {
    std::lock_guard<std::mutex> lk(g_tune.mu);
    tpp = g_tune.grid_tpp;
    layout = g_tune.map_layout;
  }
  launch_grid_build(a, tpp, s);
  he = hipGetLastError();
  if (he == hipSuccess) he = hipStreamSynchronize(s);
  (void)hipFree(d_tlj_mem);
  if (he != hipSuccess) {
    pxd_grid_destroy(g);
    return fail(PXD_ERR_HIP, "grid build: %s", hipGetErrorString(he));
  }
  st = grid_sync_struct(g);
  if (st == PXD_OK && layout) st = pxd_grid_set_layout(g, layout, stream);
  if (st != PXD_OK) {
    pxd_grid_destroy(g);
    return st;
  }
  *out = g;
  return PXD_OK;
}

pxd_status pxd_grid_upload(const pxd_receptor* r, const pxd_box* box, const int32_t* types,
                           int32_t n_types, double slope, const float* maps, pxd_grid** out) {
  if (!maps) return fail(PXD_ERR_ARG, "grid upload: null maps");
  pxd_grid* g = nullptr;
  pxd_status st = grid_alloc(r, box, types, n_types, slope, &g);
  if (st != PXD_OK) return st;
  size_t bytes = (size_t)(n_types + 1) * g->dev.P * sizeof(float4);
  hipError_t he = hipMemcpy(const_cast<float4*>(g->dev.maps), maps, bytes, hipMemcpyHostToDevice);
  if (he != hipSuccess) {
    pxd_grid_destroy(g);
    return fail(PXD_ERR_HIP, "grid upload: %s", hipGetErrorString(he));
  }
  st = grid_sync_struct(g);
  if (st != PXD_OK) {
    pxd_grid_destroy(g);
    return st;
  }
  *out = g;
  return PXD_OK;
}

pxd_status pxd_grid_download(const pxd_grid* g, float* maps, size_t n_floats) {
  if (!g || !maps) return fail(PXD_ERR_ARG, "grid download: null argument");
  size_t need = (size_t)(g->n_types + 1) * g->dev.P * 4;
  if (n_floats < need) return fail(PXD_ERR_ARG, "grid download: buffer %zu < %zu floats", n_floats, need);
  HIP_TRY(hipMemcpy(maps, g->dev.maps, need * sizeof(float), hipMemcpyDeviceToHost));
  return PXD_OK;
}

pxd_status pxd_grid_set_layout(pxd_grid* g, int32_t layout, void* stream) {
  if (!g) return fail(PXD_ERR_ARG, "grid layout: null grid");
  if (layout < 0 || layout > 2) return fail(PXD_ERR_ARG, "grid layout: %d (0, 1 or 2)", layout);
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (layout == g->dev.layout) return PXD_OK;
  // the lookup copy of the new layout (layouts 1 and 2; layout 0 reads the maps)
  void* mem = nullptr;
  if (layout != 0) {
    const int nm = g->n_types + 1;
    size_t bytes = layout == 1 ? (size_t)nm * g->dev.C * 8 * sizeof(float4)
                               : (size_t)nm * 4 * g->dev.P * sizeof(float);
    HIP_TRY(hipMalloc(&mem, bytes));
    if (layout == 1)
      launch_expand_cells(g->dev, nm, static_cast<float4*>(mem), s);
    else
      launch_split_planes(g->dev, nm, static_cast<float*>(mem), s);
    hipError_t he = hipGetLastError();
    if (he == hipSuccess) he = hipStreamSynchronize(s);
    if (he != hipSuccess) {
      (void)hipFree(mem);
      return fail(PXD_ERR_HIP, "grid layout: %s", hipGetErrorString(he));
    }
  }
  HIP_TRY(hipStreamSynchronize(s));  // no launch may still read the old descriptor
  void* old = g->cmem;
  g->cmem = mem;
  g->dev.layout = layout;
  g->dev.cmaps = layout == 1 ? static_cast<const float4*>(mem) : nullptr;
  g->dev.smaps = layout == 2 ? static_cast<const float*>(mem) : nullptr;
  pxd_status st = grid_sync_struct(g);
  if (old) HIP_TRY(hipFree(old));
  return st;
}

pxd_status pxd_grid_info(const pxd_grid* g, pxd_box* box, int32_t* n_types) {
  if (!g) return fail(PXD_ERR_ARG, "grid info: null grid");
  if (box) *box = g->box;
  if (n_types) *n_types = g->n_types;
  return PXD_OK;
}

pxd_status pxd_grid_set_slope(pxd_grid* g, double slope) {
  if (!g) return fail(PXD_ERR_ARG, "grid slope: null grid");
  g->dev.slope = (float)slope;
  g->slope = slope;
  return grid_sync_struct(g);
}

void pxd_grid_destroy(pxd_grid* g) {
  if (!g) return;
  (void)hipFree(g->mem);
  if (g->cmem) (void)hipFree(g->cmem);
  delete g;
}

// ------------------------------------------------------------------ ligands

pxd_status pxd_ligand_batch_create(const pxd_mol* ligs, const pxd_frag_tree* trees, int32_t nlig,
                                   pxd_ligand_batch** out) {
  if (!ligs || nlig <= 0 || !out) return fail(PXD_ERR_ARG, "ligand batch: empty");
  *out = nullptr;
  std::vector<LigInfo> info(nlig);
  std::vector<int> type;
  std::vector<float> q;
  std::vector<double> q64;
  std::vector<float2> lj;
  std::vector<uint32_t> bij, aijk, dijkl, dper, pairs, tsijkl;
  std::vector<float2> bkr, akt;
  std::vector<float4> dkp, tscoef;
  std::vector<int> frag, torsf, perm, sets, rings, tsmoff;
  std::vector<int2> dons;
  std::vector<uint4> pmask;
  std::vector<uint32_t> bslot, aslot, dslot, gat;
  std::vector<int> natoms(nlig), ntors(nlig), nfrags(nlig);
  uint8_t used[PXD_NUM_VDW_TYPES] = {0};
  for (int l = 0; l < nlig; ++l) {
    const pxd_mol& m = ligs[l];
    LigInfo& I = info[l];
    memset(&I, 0, sizeof(I));
    const int n = m.n_atoms;
    if (n <= 0 || n > kMaxAtoms)
      return fail(PXD_ERR_RANGE, "ligand %d: %d atoms (1..%d supported)", l, n, kMaxAtoms);
    if (!m.vdw_type || !m.charge) return fail(PXD_ERR_CONF, "ligand %d: vdw_type/charge required", l);
    auto ok = [&](const int32_t* ids, int cnt, int w) {
      for (int i = 0; i < cnt * w; ++i)
        if (ids[i] < 0 || ids[i] >= n) return false;
      return true;
    };
    if (!ok(m.bond_ij, m.n_bonds, 2) || !ok(m.angle_ijk, m.n_angles, 3) ||
        !ok(m.dih_ijkl, m.n_dihedrals, 4) || !ok(m.pairs14, m.n_pairs14, 2) ||
        !ok(m.pairs, m.n_pairs, 2) || !ok(m.hydrophobic, m.n_hydrophobic, 1) ||
        !ok(m.cation, m.n_cation, 1) || !ok(m.anion, m.n_anion, 1) ||
        !ok(m.ring5, m.n_ring5, 5) || !ok(m.ring6, m.n_ring6, 6) ||
        !ok(m.don_charged, m.n_don_charged, 2) || !ok(m.don_neutral, m.n_don_neutral, 2) ||
        !ok(m.acc_charged, m.n_acc_charged, 1) || !ok(m.acc_neutral, m.n_acc_neutral, 1))
      return fail(PXD_ERR_CONF, "ligand %d: term atom index out of range", l);
    I.n_atoms = n;
    I.atom_off = (int)type.size();
    for (int i = 0; i < n; ++i) {
      int t = m.vdw_type[i];
      if (t < 0 || t >= PXD_NUM_VDW_TYPES)
        return fail(PXD_ERR_CONF, "ligand %d: VDW type [%d] of atom#%d is illegal!", l, t, i);
      used[t] = 1;
      type.push_back(t);
      q.push_back((float)m.charge[i]);
      q64.push_back(m.charge[i]);
      lj.push_back(make_float2((float)kGaffSigma[t], (float)std::sqrt(kGaffEpsilon[t])));
    }
    I.bond_off = (int)bij.size();
    I.n_bonds = m.n_bonds;
    for (int b = 0; b < m.n_bonds; ++b) {
      bij.push_back(uint32_t(m.bond_ij[2 * b]) | uint32_t(m.bond_ij[2 * b + 1]) << 8);
      bkr.push_back(make_float2((float)m.bond_k[b], (float)m.bond_r0[b]));
    }
    I.angle_off = (int)aijk.size();
    I.n_angles = m.n_angles;
    for (int a = 0; a < m.n_angles; ++a) {
      const int* id = m.angle_ijk + 3 * a;
      aijk.push_back(uint32_t(id[0]) | uint32_t(id[1]) << 8 | uint32_t(id[2]) << 16);
      akt.push_back(make_float2((float)m.angle_k[a], (float)m.angle_t0[a]));
    }
    I.dih_off = (int)dijkl.size();
    I.n_dih = m.n_dihedrals;
    int bad = 0;
    for (int t = 0; t < m.n_dihedrals; ++t) {
      const int* id = m.dih_ijkl + 4 * t;
      dijkl.push_back(uint32_t(id[0]) | uint32_t(id[1]) << 8 | uint32_t(id[2]) << 16 |
                      uint32_t(id[3]) << 24);
      dper.push_back(pack_periods(m.dih_n + 6 * t, &bad));
      float4 kp[3];
      pack_kp(m.dih_k + 6 * t, m.dih_phase + 6 * t, m.dih_n + 6 * t, kp);
      dkp.insert(dkp.end(), kp, kp + 3);
    }
    if (bad) return fail(PXD_ERR_CONF, "ligand %d: dihedral periodicity outside 0..15", l);
    // bonded-gradient gather tables (chunks of 64 terms, slots atom by atom)
    I.gat_off = (int)gat.size();
    auto chunks = [&](const int32_t* ids, int cnt, int w, std::vector<uint32_t>& slotw) {
      for (int c0 = 0; c0 < cnt; c0 += kWave) {
        const int c1 = std::min(cnt, c0 + kWave);
        std::vector<int> num(n, 0), first(n, 0);
        for (int t = c0; t < c1; ++t)
          for (int r = 0; r < w; ++r) ++num[ids[w * t + r]];
        int acc = 0;
        for (int a = 0; a < n; ++a) {
          if (num[a] > 255) return false;
          first[a] = acc;
          acc += num[a];
        }
        std::vector<int> next = first;
        for (int t = c0; t < c1; ++t) {
          uint32_t sw = 0;
          for (int r = 0; r < w; ++r) sw |= uint32_t(next[ids[w * t + r]]++) << (8 * r);
          slotw.push_back(sw);
        }
        for (int lane = 0; lane < kWave; ++lane) {
          uint32_t g = 0;
          for (int h = 0; h < 2; ++h) {
            int a = lane + h * kWave;
            if (a < n && num[a]) g |= (uint32_t(first[a]) | uint32_t(num[a]) << 8) << (16 * h);
          }
          gat.push_back(g);
        }
      }
      return true;
    };
    if (!chunks(m.bond_ij, m.n_bonds, 2, bslot) || !chunks(m.angle_ijk, m.n_angles, 3, aslot) ||
        !chunks(m.dih_ijkl, m.n_dihedrals, 4, dslot))
      return fail(PXD_ERR_CONF, "ligand %d: an atom in over 255 of 64 consecutive bonded terms", l);
    I.pair_off = (int)pairs.size();
    I.n_pairs = m.n_pairs14 + m.n_pairs;
    for (int p = 0; p < m.n_pairs14; ++p)
      pairs.push_back(uint32_t(m.pairs14[2 * p]) | uint32_t(m.pairs14[2 * p + 1]) << 8 | 1u << 16);
    for (int p = 0; p < m.n_pairs; ++p)
      pairs.push_back(uint32_t(m.pairs[2 * p]) | uint32_t(m.pairs[2 * p + 1]) << 8);
    // partner masks: each listed pair marked from both ends; a list that names
    // a pair twice (or an atom with itself) keeps the list form
    {
      std::vector<uint32_t> mk((size_t)n * 8, 0u);
      bool mask_ok = true;
      for (int pass = 0; pass < 2 && mask_ok; ++pass) {
        int cnt = pass == 0 ? m.n_pairs : m.n_pairs14;
        const int32_t* ij = pass == 0 ? m.pairs : m.pairs14;
        for (int p = 0; p < cnt; ++p) {
          int i = ij[2 * p], j = ij[2 * p + 1];
          uint32_t* wi = &mk[(size_t)i * 8 + 4 * pass];
          uint32_t* wj = &mk[(size_t)j * 8 + 4 * pass];
          if (i == j || (wi[j >> 5] >> (j & 31) & 1u)) {
            mask_ok = false;
            break;
          }
          wi[j >> 5] |= 1u << (j & 31);
          wj[i >> 5] |= 1u << (i & 31);
        }
      }
      I.pmode = mask_ok ? 1 : 0;
      for (int a = 0; a < n; ++a)
        for (int h = 0; h < 2; ++h) {
          const uint32_t* w = &mk[(size_t)a * 8 + 4 * h];
          pmask.push_back(mask_ok ? make_uint4(w[0], w[1], w[2], w[3]) : make_uint4(0, 0, 0, 0));
        }
    }
    // fragment tree
    I.frag_off = (int)torsf.size();
    if (trees) {
      const pxd_frag_tree& T = trees[l];
      if (T.n_frags < 1 || T.n_frags > kMaxFrags)
        return fail(PXD_ERR_RANGE, "ligand %d: %d fragments (1..%d)", l, T.n_frags, kMaxFrags);
      I.n_frags = T.n_frags;
      I.n_tors = T.n_torsions;
      for (int f = 0; f < T.n_frags; ++f) {
        int s = T.frag_start[f], e = T.frag_end[f], par = T.frag_parent[f];
        if (s < 0 || e > n || s >= e || par >= T.n_frags)
          return fail(PXD_ERR_CONF, "ligand %d: bad fragment %d", l, f);
        if (par >= 0 && (T.frag_bond[2 * f] < 0 || T.frag_bond[2 * f] >= n ||
                         T.frag_bond[2 * f + 1] < 0 || T.frag_bond[2 * f + 1] >= n))
          return fail(PXD_ERR_CONF, "ligand %d: fragment %d bond atom out of range", l, f);
        frag.insert(frag.end(), {s, e, par, T.frag_bond[2 * f], T.frag_bond[2 * f + 1], 0});
      }
      // the fragment ranges partition the permuted slots [0, n), and each
      // bond joins a parent-side atom of the parent to a mine atom of the child
      std::vector<int> slot_frag(n, -1);
      int n_roots = 0;
      for (int f = 0; f < T.n_frags; ++f) {
        for (int p = T.frag_start[f]; p < T.frag_end[f]; ++p) {
          if (slot_frag[p] >= 0)
            return fail(PXD_ERR_CONF, "ligand %d: fragments %d and %d overlap", l, slot_frag[p], f);
          slot_frag[p] = f;
        }
        if (T.frag_parent[f] < 0) ++n_roots;
      }
      for (int p = 0; p < n; ++p)
        if (slot_frag[p] < 0) return fail(PXD_ERR_CONF, "ligand %d: slot %d in no fragment", l, p);
      for (int f = 0; f < T.n_frags; ++f) {
        const int par = T.frag_parent[f];
        if (par >= 0 && (slot_frag[T.frag_bond[2 * f]] != par || slot_frag[T.frag_bond[2 * f + 1]] != f))
          return fail(PXD_ERR_CONF, "ligand %d: fragment %d bond does not join it to its parent", l, f);
      }
      if (T.n_torsions != T.n_frags - n_roots)
        return fail(PXD_ERR_CONF, "ligand %d: %d torsions for %d fragments with %d roots", l,
                    T.n_torsions, T.n_frags, n_roots);
      // every child fragment once, after its parent (the frames are composed
      // in this order: a parent's frame must exist before its child's)
      std::vector<char> placed(T.n_frags, 0);
      for (int f = 0; f < T.n_frags; ++f) placed[f] = T.frag_parent[f] < 0;
      for (int k = 0; k < T.n_torsions; ++k) {
        const int f = T.tors_frag[k];
        if (f < 0 || f >= T.n_frags || T.frag_parent[f] < 0)
          return fail(PXD_ERR_CONF, "ligand %d: torsion %d moves fragment %d, not a child fragment",
                      l, k, f);
        if (placed[f])
          return fail(PXD_ERR_CONF, "ligand %d: fragment %d listed twice in the torsions", l, f);
        if (!placed[T.frag_parent[f]])
          return fail(PXD_ERR_CONF, "ligand %d: torsion %d (fragment %d) precedes its parent's", l,
                      k, f);
        placed[f] = 1;
        torsf.push_back(f);
      }
      torsf.resize(I.frag_off + T.n_frags, 0);
      // perm must be a permutation of the atoms
      std::vector<char> seen(n, 0);
      for (int i = 0; i < n; ++i) {
        const int o = T.perm[i];
        if (o < 0 || o >= n || seen[o])
          return fail(PXD_ERR_CONF, "ligand %d: perm is not a permutation of 0..%d", l, n - 1);
        seen[o] = 1;
        perm.push_back(o);
      }
    } else {
      frag.insert(frag.end(), {0, n, -1, 0, 0, 0});
      torsf.push_back(0);
      I.n_frags = 1;
      for (int i = 0; i < n; ++i) perm.push_back(i);
    }
    // pscore sets
    auto put = [&](const int32_t* p, int cnt, int* off, int* num) {
      *off = (int)sets.size();
      *num = cnt;
      sets.insert(sets.end(), p, p + cnt);
    };
    put(m.hydrophobic, m.n_hydrophobic, &I.hydro_off, &I.n_hydro);
    put(m.cation, m.n_cation, &I.cation_off, &I.n_cation);
    put(m.anion, m.n_anion, &I.anion_off, &I.n_anion);
    put(m.acc_charged, m.n_acc_charged, &I.accc_off, &I.n_accc);
    put(m.acc_neutral, m.n_acc_neutral, &I.accn_off, &I.n_accn);
    if (m.n_ring5 + m.n_ring6 > 16)
      return fail(PXD_ERR_RANGE, "ligand %d: more than 16 pi rings", l);
    I.ring5_off = (int)rings.size();
    I.n_ring5 = m.n_ring5;
    rings.insert(rings.end(), m.ring5, m.ring5 + 5 * m.n_ring5);
    I.ring6_off = (int)rings.size();
    I.n_ring6 = m.n_ring6;
    rings.insert(rings.end(), m.ring6, m.ring6 + 6 * m.n_ring6);
    I.donc_off = (int)dons.size();
    I.n_donc = m.n_don_charged;
    for (int i = 0; i < m.n_don_charged; ++i)
      dons.push_back(make_int2(m.don_charged[2 * i], m.don_charged[2 * i + 1]));
    I.donn_off = (int)dons.size();
    I.n_donn = m.n_don_neutral;
    for (int i = 0; i < m.n_don_neutral; ++i)
      dons.push_back(make_int2(m.don_neutral[2 * i], m.don_neutral[2 * i + 1]));
    I.ts_off = (int)tscoef.size();
    I.n_ts = m.n_tstrain;
    I.tsm_off = (int)tsmoff.size();
    for (int p = 0; p < m.n_tstrain; ++p) {
      const double* c = m.tstrain_coef + 4 * p;
      tscoef.push_back(make_float4((float)c[0], (float)c[1], (float)c[2], (float)c[3]));
      tsmoff.push_back((int)tsijkl.size());
      for (int k = m.tstrain_match_off[p]; k < m.tstrain_match_off[p + 1]; ++k) {
        const int* a = m.tstrain_ijkl + 4 * k;
        for (int d = 0; d < 4; ++d)
          if (a[d] < 0 || a[d] >= n) return fail(PXD_ERR_CONF, "ligand %d: torsion-strain atom out of range", l);
        tsijkl.push_back(uint32_t(a[0]) | uint32_t(a[1]) << 8 | uint32_t(a[2]) << 16 | uint32_t(a[3]) << 24);
      }
    }
    tsmoff.push_back((int)tsijkl.size());
    I.n_rot = m.n_rotatable;
    natoms[l] = n;
    ntors[l] = I.n_tors;
    nfrags[l] = I.n_frags;
  }
  Packer P;
  size_t o_info = P.add(info), o_type = P.add(type), o_q = P.add(q), o_lj = P.add(lj),
         o_bij = P.add(bij), o_bkr = P.add(bkr), o_aijk = P.add(aijk), o_akt = P.add(akt),
         o_dijkl = P.add(dijkl), o_dper = P.add(dper), o_dkp = P.add(dkp), o_pairs = P.add(pairs),
         o_frag = P.add(frag), o_torsf = P.add(torsf), o_perm = P.add(perm), o_sets = P.add(sets),
         o_rings = P.add(rings), o_dons = P.add(dons), o_tsc = P.add(tscoef),
         o_tsm = P.add(tsmoff), o_tsi = P.add(tsijkl), o_pmask = P.add(pmask),
         o_bslot = P.add(bslot), o_aslot = P.add(aslot), o_dslot = P.add(dslot),
         o_gat = P.add(gat);
  pxd_ligand_batch* b = new pxd_ligand_batch();
  hipError_t he = hipMalloc(&b->mem, P.host.size());
  if (he == hipSuccess) he = hipMemcpy(b->mem, P.host.data(), P.host.size(), hipMemcpyHostToDevice);
  if (he != hipSuccess) {
    if (b->mem) (void)hipFree(b->mem);
    delete b;
    return fail(PXD_ERR_HIP, "ligand batch: %s", hipGetErrorString(he));
  }
  void* B = b->mem;
  DevLigands& D = b->dev;
  D.n = nlig;
  D.info = dptr<LigInfo>(B, o_info);
  D.type = dptr<int>(B, o_type);
  D.q = dptr<float>(B, o_q);
  D.lj = dptr<float2>(B, o_lj);
  D.bond_ij = dptr<uint32_t>(B, o_bij);
  D.bond_kr = dptr<float2>(B, o_bkr);
  D.angle_ijk = dptr<uint32_t>(B, o_aijk);
  D.angle_kt = dptr<float2>(B, o_akt);
  D.dih_ijkl = dptr<uint32_t>(B, o_dijkl);
  D.dih_per = dptr<uint32_t>(B, o_dper);
  D.dih_kp = dptr<float4>(B, o_dkp);
  D.pairs = dptr<uint32_t>(B, o_pairs);
  D.frag = dptr<int>(B, o_frag);
  D.tors_frag = dptr<int>(B, o_torsf);
  D.perm = dptr<int>(B, o_perm);
  D.sets = dptr<int>(B, o_sets);
  D.rings = dptr<int>(B, o_rings);
  D.dons = dptr<int2>(B, o_dons);
  D.ts_coef = dptr<float4>(B, o_tsc);
  D.ts_moff = dptr<int>(B, o_tsm);
  D.ts_ijkl = dptr<uint32_t>(B, o_tsi);
  D.pmask = dptr<uint4>(B, o_pmask);
  D.bond_slot = dptr<uint32_t>(B, o_bslot);
  D.angle_slot = dptr<uint32_t>(B, o_aslot);
  D.dih_slot = dptr<uint32_t>(B, o_dslot);
  D.gat = dptr<uint32_t>(B, o_gat);
  b->mem_bytes = P.host.size();
  b->n = nlig;
  b->h_natoms = new int[nlig];
  b->h_ntors = new int[nlig];
  memcpy(b->h_natoms, natoms.data(), sizeof(int) * nlig);
  b->max_atoms = *std::max_element(natoms.begin(), natoms.end());
  b->max_frags = *std::max_element(nfrags.begin(), nfrags.end());
  memcpy(b->h_ntors, ntors.data(), sizeof(int) * nlig);
  memcpy(b->type_used, used, sizeof(used));
  b->has_trees = trees ? 1 : 0;
  b->h_q64 = std::move(q64);
  b->h_atom_off.resize(nlig + 1, 0);
  for (int l = 0; l < nlig; ++l) b->h_atom_off[l + 1] = b->h_atom_off[l] + natoms[l];
  *out = b;
  return PXD_OK;
}

void pxd_ligand_batch_destroy(pxd_ligand_batch* b) {
  if (!b) return;
  (void)hipFree(b->mem);
  delete[] b->h_natoms;
  delete[] b->h_ntors;
  delete b;
}

pxd_status pxd_ligand_batch_natoms(const pxd_ligand_batch* b, int32_t i, int32_t* out) {
  if (!b || i < 0 || i >= b->n || !out) return fail(PXD_ERR_ARG, "natoms: bad index");
  *out = b->h_natoms[i];
  return PXD_OK;
}

// ------------------------------------------------------------------ scoring

static pxd_status check_grid_types(const pxd_grid* g, const pxd_ligand_batch* b) {
  if (!g) return PXD_OK;
  for (int t = 0; t < PXD_NUM_VDW_TYPES; ++t)
    if (b->type_used[t] && g->dev.type_slot[t] < 0)
      return fail(PXD_ERR_STATE, "ligand type %d has no map in the grid", t);
  return PXD_OK;
}

}  // extern "C"

namespace pxd {

namespace {
constexpr int kMaxSideDevices = 64;
std::mutex g_side_mu;
hipStream_t g_side[kMaxSideDevices][kSideStreams] = {};
}  // namespace

static_assert(kLbfgsClasses - 1 <= kSideStreams,
              "every size class but one needs a side stream");

std::mutex& side_stream_mutex() { return g_side_mu; }

hipError_t side_streams(hipStream_t out[kSideStreams]) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  if (dev < 0 || dev >= kMaxSideDevices) return hipErrorInvalidDevice;
  for (int k = 0; k < kSideStreams; ++k) {
    if (!g_side[dev][k]) {
      e = hipStreamCreateWithFlags(&g_side[dev][k], hipStreamNonBlocking);
      if (e != hipSuccess) {
        g_side[dev][k] = nullptr;
        return e;
      }
    }
    out[k] = g_side[dev][k];
  }
  return hipSuccess;
}

// Argument checks of pxd_lbfgs_batch and its launch arguments (all but the
// queue heads and the history storage), shared with the entry point that
// allocates nothing.  *n_slots: the work slots the tuning asks for (at most
// one per pose).  n_poses == 0 returns PXD_OK with nothing to launch.
pxd_status lbfgs_prepare(const pxd_receptor* r, const pxd_grid* g, const pxd_ligand_batch* b,
                         const pxd_lbfgs_params* prm, int32_t n_poses, const int32_t* pose_lig,
                         const int32_t* pose_atom_off, float* xyz, float* tors, double* energy,
                         int32_t* nevals, int32_t* converged, LbfgsArgs* a, int* n_slots) {
  if (!r || !b || !prm) return fail(PXD_ERR_ARG, "lbfgs: null argument");
  if (n_poses < 0) return fail(PXD_ERR_ARG, "lbfgs: n_poses < 0");
  if (n_poses == 0) return PXD_OK;
  if (!pose_lig || !pose_atom_off || !xyz || !energy || !nevals || !converged)
    return fail(PXD_ERR_ARG, "lbfgs: null pose array");
  if (r->n_leaves > 0 && !tors) return fail(PXD_ERR_ARG, "lbfgs: tors required (receptor has leaves)");
  if (prm->history < 1 || prm->history > kMaxHistory)
    return fail(PXD_ERR_RANGE, "lbfgs: history %d outside 1..%d", prm->history, kMaxHistory);
  if (prm->max_iter < 0 || prm->max_eval < 0 || prm->max_ls < 0)
    return fail(PXD_ERR_ARG, "lbfgs: negative iteration limits");
  if (prm->dof_mode && !b->has_trees)
    return fail(PXD_ERR_STATE, "lbfgs: rigid-body + torsion variables need fragment trees");
  pxd_status st = check_grid_types(g, b);
  if (st != PXD_OK) return st;
  int dev = 0, ncu = 0;
  HIP_TRY(hipGetDevice(&dev));
  HIP_TRY(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
  int spc;
  {
    std::lock_guard<std::mutex> lk(g_tune.mu);
    spc = g_tune.lbfgs_slots_per_cu;
  }
  *n_slots = std::min(n_poses, std::max(1, ncu) * spc);
  a->prm = {prm->history, prm->max_iter,  prm->max_eval,      prm->max_ls,
            prm->lr,      prm->tol_grad,  prm->tol_change,    prm->c1,
            prm->c2,      prm->ls_tol_change, prm->gtd_threshold, prm->dof_mode ? 1 : 0};
  a->n_poses = n_poses;
  a->pose_lig = pose_lig;
  a->pose_off = pose_atom_off;
  a->xyz = xyz;
  a->tors = tors;
  a->energy = energy;
  a->nevals = nevals;
  a->converged = converged;
  a->counter = nullptr;
  a->hist = nullptr;
  a->max_atoms = b->max_atoms;
  a->stats = nullptr;
  a->order = nullptr;
  a->n_order = 0;
  return PXD_OK;
}

const DevGrid* grid_device(const pxd_grid* g) { return g ? grid_dev_ptr(g) : nullptr; }

}  // namespace pxd

extern "C" {

pxd_status pxd_score_grad_batch(const pxd_receptor* r, const pxd_grid* g,
                                const pxd_ligand_batch* b, int32_t n_poses,
                                const int32_t* pose_lig, const int32_t* pose_atom_off,
                                const float* xyz, const float* tors, double* energy,
                                float* lig_grad, float* tors_grad, void* stream) {
  if (!r || !b) return fail(PXD_ERR_ARG, "score: null receptor/batch");
  if (n_poses < 0) return fail(PXD_ERR_ARG, "score: n_poses < 0");
  if (n_poses == 0) return PXD_OK;
  if (!pose_lig || !pose_atom_off || !xyz || !energy || !lig_grad)
    return fail(PXD_ERR_ARG, "score: null pose array");
  if (r->n_leaves > 0 && tors_grad == nullptr)
    return fail(PXD_ERR_ARG, "score: tors_grad required when the receptor has leaves");
  pxd_status st = check_grid_types(g, b);
  if (st != PXD_OK) return st;
  ScoreArgs a{n_poses, pose_lig, pose_atom_off, xyz, tors, energy, lig_grad, tors_grad,
              b->max_atoms};
  {
    std::lock_guard<std::mutex> lk(g_tune.mu);
    a.poses_per_block = g_tune.score_ppb;
  }
  launch_score(r->dev, b->dev, g ? grid_dev_ptr(g) : nullptr, a, static_cast<hipStream_t>(stream));
  HIP_TRY(hipGetLastError());
  return PXD_OK;
}

pxd_status pxd_lbfgs_batch(const pxd_receptor* r, const pxd_grid* g, const pxd_ligand_batch* b,
                           const pxd_lbfgs_params* prm, int32_t n_poses, const int32_t* pose_lig,
                           const int32_t* pose_atom_off, float* xyz, float* tors,
                           double* energy, int32_t* nevals, int32_t* converged, void* stream) {
  LbfgsArgs a;
  int n_slots = 0;
  pxd_status st = pxd::lbfgs_prepare(r, g, b, prm, n_poses, pose_lig, pose_atom_off, xyz, tors,
                                     energy, nevals, converged, &a, &n_slots);
  if (st != PXD_OK || n_poses == 0) return st;
  hipStream_t s = static_cast<hipStream_t>(stream);
  size_t per_slot = lbfgs_slot_floats(prm->history);
  std::lock_guard<std::mutex> lock(g_ws.mu);
  std::lock_guard<std::mutex> side_lock(side_stream_mutex());
  // Cartesian variables: the poses grouped by size class on the host (stable
  // order), one launch per class, the classes side by side on streams of
  // their own so that one class's tail overlaps the others' work; work
  // slots shared out by each class's poses x vector width.
  int cls_n[kLbfgsClasses] = {0, 0, 0}, cls_slots[kLbfgsClasses] = {0, 0, 0};
  std::vector<int32_t> order;
  if (!prm->dof_mode) {
    HIP_TRY(hipStreamSynchronize(s));  // pose_lig may be written by earlier work on s
    std::vector<int32_t> hl(n_poses);
    HIP_TRY(hipMemcpy(hl.data(), pose_lig, 4 * (size_t)n_poses, hipMemcpyDeviceToHost));
    std::vector<int8_t> cls(n_poses);
    for (int p = 0; p < n_poses; ++p) {
      if (hl[p] < 0 || hl[p] >= b->n) return fail(PXD_ERR_ARG, "lbfgs: pose ligand out of range");
      const int n = 3 * b->h_natoms[hl[p]] + r->n_leaves;
      const int nk = (n + kWave - 1) / kWave;
      cls[p] = nk <= kLbfgsClassVpl[0] ? 0 : (nk <= kLbfgsClassVpl[1] ? 1 : 2);
      cls_n[cls[p]]++;
    }
    order.reserve(n_poses);
    for (int c = 0; c < kLbfgsClasses; ++c)
      for (int p = 0; p < n_poses; ++p)
        if (cls[p] == c) order.push_back(p);
    double wsum = 0.0;
    for (int c = 0; c < kLbfgsClasses; ++c) wsum += (double)cls_n[c] * kLbfgsClassVpl[c];
    int total = 0;
    for (int c = 0; c < kLbfgsClasses; ++c) {
      if (!cls_n[c]) continue;
      const int k = (int)std::lround((double)n_slots * cls_n[c] * kLbfgsClassVpl[c] / wsum);
      cls_slots[c] = std::max(1, std::min(k, cls_n[c]));
      total += cls_slots[c];
    }
    n_slots = total;
  }
  if (g_ws.hist_floats < per_slot * n_slots) {
    if (g_ws.hist) HIP_TRY(hipFree(g_ws.hist));
    g_ws.hist = nullptr;
    g_ws.hist_floats = 0;
    HIP_TRY(hipMalloc(&g_ws.hist, per_slot * n_slots * sizeof(float)));
    g_ws.hist_floats = per_slot * n_slots;
  }
  if (!g_ws.counter) HIP_TRY(hipMalloc(&g_ws.counter, sizeof(int) * 64));
  // queue heads at ints 0..2, the statistics (4 x uint64) at ints 16..23
  HIP_TRY(hipMemsetAsync(g_ws.counter, 0, sizeof(int) * 24, s));
  a.counter = g_ws.counter;
  a.hist = g_ws.hist;
  a.stats = reinterpret_cast<unsigned long long*>(g_ws.counter + 16);
  if (prm->dof_mode) {
    launch_lbfgs(r->dev, b->dev, g ? grid_dev_ptr(g) : nullptr, a, n_slots, s);
  } else {
    if (g_ws.order_n < (size_t)n_poses) {
      if (g_ws.order) HIP_TRY(hipFree(g_ws.order));
      g_ws.order = nullptr;
      g_ws.order_n = 0;
      HIP_TRY(hipMalloc(&g_ws.order, sizeof(int) * (size_t)n_poses));
      g_ws.order_n = n_poses;
    }
    HIP_TRY(hipMemcpyAsync(g_ws.order, order.data(), sizeof(int) * (size_t)n_poses,
                           hipMemcpyHostToDevice, s));
    if (!g_ws.fork) {
      HIP_TRY(hipEventCreateWithFlags(&g_ws.fork, hipEventDisableTiming));
      for (int k = 0; k < kLbfgsClasses - 1; ++k)
        HIP_TRY(hipEventCreateWithFlags(&g_ws.join[k], hipEventDisableTiming));
    }
    hipStream_t side_s[kSideStreams];
    HIP_TRY(side_streams(side_s));
    HIP_TRY(hipEventRecord(g_ws.fork, s));
    // the class with the most work on the caller's stream, the others beside it
    int main_c = 0;
    for (int c = 1; c < kLbfgsClasses; ++c)
      if ((double)cls_n[c] * kLbfgsClassVpl[c] > (double)cls_n[main_c] * kLbfgsClassVpl[main_c])
        main_c = c;
    int off = 0, side = 0;
    bool joined[kLbfgsClasses - 1] = {};
    size_t hist_off = 0;
    for (int c = 0; c < kLbfgsClasses; ++c) {
      LbfgsArgs ac = a;
      ac.order = g_ws.order + off;
      ac.n_order = cls_n[c];
      ac.counter = g_ws.counter + c;
      ac.hist = g_ws.hist + hist_off;
      off += cls_n[c];
      hist_off += per_slot * (size_t)cls_slots[c];
      if (!cls_n[c]) continue;
      hipStream_t cs = s;
      if (c != main_c) {
        cs = side_s[side];
        HIP_TRY(hipStreamWaitEvent(cs, g_ws.fork, 0));
      }
      launch_lbfgs_class(r->dev, b->dev, g ? grid_dev_ptr(g) : nullptr, ac, c, cls_slots[c], cs);
      if (c != main_c) {
        HIP_TRY(hipEventRecord(g_ws.join[side], cs));
        joined[side] = true;
        side++;
      }
    }
    for (int k = 0; k < kLbfgsClasses - 1; ++k)
      if (joined[k]) HIP_TRY(hipStreamWaitEvent(s, g_ws.join[k], 0));
  }
  HIP_TRY(hipGetLastError());
  // the shared workspace may be reused by the next call: finish this one
  HIP_TRY(hipStreamSynchronize(s));
  unsigned long long st4[4];
  HIP_TRY(hipMemcpy(st4, a.stats, sizeof(st4), hipMemcpyDeviceToHost));
  g_ws.last = {st4[0], st4[1], st4[2], st4[3]};
  return PXD_OK;
}

pxd_status pxd_lbfgs_last_stats(pxd_lbfgs_stats* out) {
  if (!out) return fail(PXD_ERR_ARG, "lbfgs stats: null argument");
  std::lock_guard<std::mutex> lock(g_ws.mu);
  *out = g_ws.last;
  return PXD_OK;
}

pxd_status pxd_dof_grad_batch(const pxd_ligand_batch* b, int32_t n_poses, const int32_t* pose_lig,
                              const int32_t* pose_atom_off, const float* xyz,
                              const float* lig_grad, float* dof_grad, void* stream) {
  if (!b || n_poses < 0) return fail(PXD_ERR_ARG, "dof_grad: bad argument");
  if (n_poses == 0) return PXD_OK;
  if (!pose_lig || !pose_atom_off || !xyz || !lig_grad || !dof_grad)
    return fail(PXD_ERR_ARG, "dof_grad: null array");
  if (!b->has_trees) return fail(PXD_ERR_STATE, "dof_grad: the batch has no fragment trees");
  launch_dof_grad(b->dev, n_poses, pose_lig, pose_atom_off, xyz, lig_grad, dof_grad,
                  static_cast<hipStream_t>(stream));
  HIP_TRY(hipGetLastError());
  return PXD_OK;
}

pxd_status pxd_pscore_grad_batch(const pxd_receptor* r, const pxd_ligand_batch* b,
                                 const pxd_score_params* sp, int32_t n_poses,
                                 const int32_t* pose_lig, const int32_t* pose_atom_off,
                                 const float* xyz, const float* tors, double* score,
                                 float* dscore_dxyz, double* dscore_dparams, void* stream) {
  if (!r || !b || !sp) return fail(PXD_ERR_ARG, "pscore_grad: null argument");
  if (n_poses < 0) return fail(PXD_ERR_ARG, "pscore_grad: n_poses < 0");
  if (n_poses == 0) return PXD_OK;
  if (!pose_lig || !pose_atom_off || !xyz || !score)
    return fail(PXD_ERR_ARG, "pscore_grad: null pose array");
  if (r->dev.n_ring5 + r->dev.n_ring6 > 256)
    return fail(PXD_ERR_RANGE, "pscore_grad: receptor has more than 256 pi rings");
  PscoreArgs a{n_poses, pose_lig, pose_atom_off, xyz, tors, score, nullptr, b->max_atoms};
  launch_pscore_grad(r->dev, b->dev, *sp, a, dscore_dxyz, dscore_dparams,
                     static_cast<hipStream_t>(stream));
  HIP_TRY(hipGetLastError());
  return PXD_OK;
}

pxd_status pxd_pscore_batch(const pxd_receptor* r, const pxd_ligand_batch* b,
                            const pxd_score_params* sp, int32_t n_poses, const int32_t* pose_lig,
                            const int32_t* pose_atom_off, const float* xyz, const float* tors,
                            double* score, double* terms, void* stream) {
  if (!r || !b || !sp) return fail(PXD_ERR_ARG, "pscore: null argument");
  if (n_poses < 0) return fail(PXD_ERR_ARG, "pscore: n_poses < 0");
  if (n_poses == 0) return PXD_OK;
  if (!pose_lig || !pose_atom_off || !xyz || !score) return fail(PXD_ERR_ARG, "pscore: null pose array");
  if (r->dev.n_ring5 + r->dev.n_ring6 > 256)
    return fail(PXD_ERR_RANGE, "pscore: receptor has more than 256 pi rings");
  PscoreArgs a{n_poses, pose_lig, pose_atom_off, xyz, tors, score, terms, b->max_atoms};
  launch_pscore(r->dev, b->dev, *sp, a, static_cast<hipStream_t>(stream));
  HIP_TRY(hipGetLastError());
  return PXD_OK;
}

// ------------------------------------------------------------------ autotuning

void pxd_get_tuning(pxd_tuning* out) {
  if (!out) return;
  memset(out, 0, sizeof(*out));
  std::lock_guard<std::mutex> lk(g_tune.mu);
  out->grid_types_per_pass = g_tune.grid_tpp;
  out->map_layout = g_tune.map_layout;
  out->lbfgs_slots_per_cu = g_tune.lbfgs_slots_per_cu;
  out->score_poses_per_block = g_tune.score_ppb;
}

pxd_status pxd_set_tuning(const pxd_tuning* t) {
  if (!t) return fail(PXD_ERR_ARG, "tuning: null argument");
  if (t->grid_types_per_pass != 4 && t->grid_types_per_pass != 8 && t->grid_types_per_pass != 16)
    return fail(PXD_ERR_ARG, "tuning: types per pass %d (4, 8 or 16)", t->grid_types_per_pass);
  if (t->map_layout < 0 || t->map_layout > 2)
    return fail(PXD_ERR_ARG, "tuning: map layout %d (0, 1 or 2)", t->map_layout);
  if (t->lbfgs_slots_per_cu < 1 || t->lbfgs_slots_per_cu > 64)
    return fail(PXD_ERR_ARG, "tuning: slots per CU %d (1..64)", t->lbfgs_slots_per_cu);
  if (t->score_poses_per_block != 1 && t->score_poses_per_block != 2 &&
      t->score_poses_per_block != 4)
    return fail(PXD_ERR_ARG, "tuning: poses per block %d (1, 2 or 4)", t->score_poses_per_block);
  std::lock_guard<std::mutex> lk(g_tune.mu);
  g_tune.grid_tpp = t->grid_types_per_pass;
  g_tune.map_layout = t->map_layout;
  g_tune.lbfgs_slots_per_cu = t->lbfgs_slots_per_cu;
  g_tune.score_ppb = t->score_poses_per_block;
  return PXD_OK;
}

}  // extern "C"

namespace {

struct DevBuf {
  void* p = nullptr;
  ~DevBuf() {
    if (p) (void)hipFree(p);
  }
};

// milliseconds of fn() on stream s, best of reps after one warm-up
template <class Fn>
pxd_status time_best(hipStream_t s, int reps, Fn fn, double* ms) {
  hipEvent_t a = nullptr, b = nullptr;
  HIP_TRY(hipEventCreate(&a));
  HIP_TRY(hipEventCreate(&b));
  double best = 1e300;
  pxd_status st = PXD_OK;
  for (int r = 0; r <= reps && st == PXD_OK; ++r) {
    if (hipEventRecord(a, s) != hipSuccess) st = PXD_ERR_HIP;
    if (st == PXD_OK) st = fn();
    if (st == PXD_OK && hipEventRecord(b, s) != hipSuccess) st = PXD_ERR_HIP;
    if (st == PXD_OK && hipEventSynchronize(b) != hipSuccess) st = PXD_ERR_HIP;
    float t = 0.f;
    if (st == PXD_OK && hipEventElapsedTime(&t, a, b) != hipSuccess) st = PXD_ERR_HIP;
    if (st == PXD_OK && r > 0) best = std::min(best, (double)t);
  }
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  if (st == PXD_ERR_HIP) return fail(PXD_ERR_HIP, "autotune: event timing failed");
  *ms = best;
  return st;
}

}  // namespace

extern "C" {

pxd_status pxd_autotune(const pxd_receptor* r, const pxd_box* box, const int32_t* types,
                        int32_t n_types, const pxd_ligand_batch* b, int32_t n_poses,
                        const int32_t* pose_lig, const int32_t* pose_atom_off, const float* xyz,
                        const float* tors, void* stream, pxd_tuning* out) {
  if (!r || !box || !b || !out || n_poses <= 0 || !pose_lig || !pose_atom_off || !xyz)
    return fail(PXD_ERR_ARG, "autotune: bad argument");
  if (r->n_leaves > 0 && !tors) return fail(PXD_ERR_ARG, "autotune: tors required");
  hipStream_t s = static_cast<hipStream_t>(stream);
  pxd_tuning t;
  pxd_get_tuning(&t);
  // atoms spanned by the poses (for scratch copies)
  std::vector<int32_t> hl(n_poses), ho(n_poses);
  HIP_TRY(hipMemcpy(hl.data(), pose_lig, 4 * (size_t)n_poses, hipMemcpyDeviceToHost));
  HIP_TRY(hipMemcpy(ho.data(), pose_atom_off, 4 * (size_t)n_poses, hipMemcpyDeviceToHost));
  size_t atoms = 0;
  for (int p = 0; p < n_poses; ++p) {
    if (hl[p] < 0 || hl[p] >= b->n) return fail(PXD_ERR_ARG, "autotune: pose ligand out of range");
    atoms = std::max(atoms, (size_t)ho[p] + (size_t)b->h_natoms[hl[p]]);
  }
  const int L = r->n_leaves;
  pxd_status st = PXD_OK;
  // 1. grid build: type maps per pass
  const int tpps[3] = {4, 8, 16};
  double best = 1e300;
  for (int k = 0; k < 3; ++k) {
    {
      std::lock_guard<std::mutex> lk(g_tune.mu);
      g_tune.grid_tpp = tpps[k];
      g_tune.map_layout = 0;
    }
    st = time_best(s, 1, [&]() -> pxd_status {
      pxd_grid* g = nullptr;
      pxd_status e = pxd_grid_build(r, box, types, n_types, 1e6, stream, &g);
      pxd_grid_destroy(g);
      return e;
    }, &t.grid_ms[k]);
    if (st != PXD_OK) return st;
    if (t.grid_ms[k] < best) {
      best = t.grid_ms[k];
      t.grid_types_per_pass = tpps[k];
    }
  }
  {
    std::lock_guard<std::mutex> lk(g_tune.mu);
    g_tune.grid_tpp = t.grid_types_per_pass;
  }
  pxd_grid* g = nullptr;
  st = pxd_grid_build(r, box, types, n_types, 1e6, stream, &g);
  if (st != PXD_OK) return st;
  // 2. score+grad: map layout
  DevBuf e, gx, tg, x2, t2, ne, cv;
  HIP_TRY(hipMalloc(&e.p, 8 * (size_t)n_poses));
  HIP_TRY(hipMalloc(&gx.p, 12 * atoms));
  HIP_TRY(hipMalloc(&tg.p, 4 * (size_t)n_poses * (L + 1)));
  auto time_score = [&](double* ms) {
    return time_best(s, 3, [&]() {
      return pxd_score_grad_batch(r, g, b, n_poses, pose_lig, pose_atom_off, xyz, tors,
                                  (double*)e.p, (float*)gx.p, (float*)tg.p, stream);
    }, ms);
  };
  {
    std::lock_guard<std::mutex> lk(g_tune.mu);
    g_tune.score_ppb = 1;
  }
  best = 1e300;
  for (int lay = 0; lay < 3 && st == PXD_OK; ++lay) {
    st = pxd_grid_set_layout(g, lay, stream);
    if (st == PXD_OK) st = time_score(&t.score_ms[lay]);
    if (st == PXD_OK && t.score_ms[lay] < best) {
      best = t.score_ms[lay];
      t.map_layout = lay;
    }
  }
  if (st == PXD_OK) st = pxd_grid_set_layout(g, t.map_layout, stream);
  // poses (waves) per workgroup of the batched score, on the chosen layout
  const int ppbs[3] = {1, 2, 4};
  best = 1e300;
  for (int k = 0; k < 3 && st == PXD_OK; ++k) {
    {
      std::lock_guard<std::mutex> lk(g_tune.mu);
      g_tune.score_ppb = ppbs[k];
    }
    st = time_score(&t.score_ppb_ms[k]);
    if (st == PXD_OK && t.score_ppb_ms[k] < best) {
      best = t.score_ppb_ms[k];
      t.score_poses_per_block = ppbs[k];
    }
  }
  {
    std::lock_guard<std::mutex> lk(g_tune.mu);
    g_tune.score_ppb = t.score_poses_per_block;
  }
  // 3. L-BFGS: resident work slots per CU (relax step, 40 iterations)
  const int spcs[4] = {2, 4, 8, 16};
  if (st == PXD_OK) {
    HIP_TRY(hipMalloc(&x2.p, 12 * atoms));
    HIP_TRY(hipMalloc(&t2.p, 4 * (size_t)n_poses * (L + 1)));
    HIP_TRY(hipMalloc(&ne.p, 4 * (size_t)n_poses));
    HIP_TRY(hipMalloc(&cv.p, 4 * (size_t)n_poses));
  }
  pxd_lbfgs_params lp;
  pxd_lbfgs_default_params(&lp, 40);
  best = 1e300;
  for (int k = 0; k < 4 && st == PXD_OK; ++k) {
    {
      std::lock_guard<std::mutex> lk(g_tune.mu);
      g_tune.lbfgs_slots_per_cu = spcs[k];
    }
    st = time_best(s, 1, [&]() -> pxd_status {
      if (hipMemcpyAsync(x2.p, xyz, 12 * atoms, hipMemcpyDeviceToDevice, s) != hipSuccess ||
          (L > 0 && hipMemcpyAsync(t2.p, tors, 4 * (size_t)n_poses * L, hipMemcpyDeviceToDevice,
                                   s) != hipSuccess))
        return fail(PXD_ERR_HIP, "autotune: copy failed");
      return pxd_lbfgs_batch(r, g, b, &lp, n_poses, pose_lig, pose_atom_off, (float*)x2.p,
                             L > 0 ? (float*)t2.p : nullptr, (double*)e.p, (int32_t*)ne.p,
                             (int32_t*)cv.p, stream);
    }, &t.lbfgs_ms[k]);
    if (st == PXD_OK && t.lbfgs_ms[k] < best) {
      best = t.lbfgs_ms[k];
      t.lbfgs_slots_per_cu = spcs[k];
    }
  }
  pxd_grid_destroy(g);
  if (st != PXD_OK) return st;
  st = pxd_set_tuning(&t);
  *out = t;
  return st;
}

}  // extern "C"
