// pxd_plain.cpp — the plain GPU form of L-BFGS: one pose per kernel launch
// with the optimiser on the host.  Same score+grad kernel as the batched path;
// this is the comparison baseline for the persistent batched kernel, not a
// production path.  The host loop restates lbfgs_step::apply + strong_wolfe
// (engine/src/core/optimizer.cpp:205-560) in FP64.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "pxd_internal.hpp"
#include "pxd_kernels.hpp"
#include "pxdock.h"

using namespace pxd;

extern "C" pxd_status pxd_set_error_(pxd_status st, const char* msg);

namespace {

struct PlainEval {
  const pxd_receptor* r;
  const DevGrid* G;
  const pxd_ligand_batch* b;
  hipStream_t st;
  int lig;
  int n3, nt, n;
  // device scratch for one pose
  int *d_lig, *d_off;
  float *d_xyz, *d_tors, *d_g, *d_tg;
  double* d_e;
  std::vector<float> hx, hg;
  int nevals = 0;
  hipError_t err = hipSuccess;

  double operator()(const std::vector<double>& x, std::vector<double>& g) {
    for (int i = 0; i < n3; ++i) hx[i] = (float)x[i];
    for (int i = 0; i < nt; ++i) hx[n3 + i] = (float)x[n3 + i];
    hipError_t e = hipMemcpyAsync(d_xyz, hx.data(), sizeof(float) * n3, hipMemcpyHostToDevice, st);
    if (e == hipSuccess && nt)
      e = hipMemcpyAsync(d_tors, hx.data() + n3, sizeof(float) * nt, hipMemcpyHostToDevice, st);
    ScoreArgs a{1, d_lig, d_off, d_xyz, d_tors, d_e, d_g, d_tg, b->max_atoms};
    if (e == hipSuccess) {
      launch_score(r->dev, b->dev, G, a, st);
      e = hipGetLastError();
    }
    double energy = NAN;
    if (e == hipSuccess) e = hipMemcpyAsync(hg.data(), d_g, sizeof(float) * n3, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess && nt)
      e = hipMemcpyAsync(hg.data() + n3, d_tg, sizeof(float) * nt, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipMemcpyAsync(&energy, d_e, sizeof(double), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) {
      err = e;
      energy = NAN;
    }
    for (int i = 0; i < n; ++i) g[i] = hg[i];
    ++nevals;
    return energy;
  }
};

double vdot(const std::vector<double>& a, const std::vector<double>& b) {
  double v = 0.0;
  for (size_t i = 0; i < a.size(); ++i) v += a[i] * b[i];
  return v;
}
double vamax(const std::vector<double>& a) {
  double v = 0.0;
  for (double x : a) v = std::max(v, std::fabs(x));
  return v;
}

double cubic(double x1, double f1, double g1, double x2, double f2, double g2, double lo, double hi) {
  double d1 = g1 + g2 - 3.0 * (f1 - f2) / (x1 - x2);
  double d2sq = d1 * d1 - g1 * g2;
  if (d2sq >= 0.0) {
    double d2 = std::sqrt(d2sq), mp;
    if (x1 <= x2)
      mp = x2 - (x2 - x1) * ((g2 + d2 - d1) / (g2 - g1 + 2.0 * d2));
    else
      mp = x1 - (x1 - x2) * ((g1 + d2 - d1) / (g1 - g2 + 2.0 * d2));
    return std::min(std::max(mp, lo), hi);
  }
  return (lo + hi) / 2.0;
}

struct Dir {
  PlainEval& ev;
  const std::vector<double>& x0;
  std::vector<double> xt;
  double operator()(double t, const std::vector<double>& d, std::vector<double>& g) {
    for (size_t i = 0; i < x0.size(); ++i) xt[i] = x0[i] + d[i] * t;
    return ev(xt, g);
  }
};

void strong_wolfe(Dir& obj, const pxd_lbfgs_params& p, double& t, const std::vector<double>& d,
                  double& f, std::vector<double>& g, double gtd) {
  size_t n = d.size();
  std::vector<double> g_new(n), g_prev, bg[2];
  double br[2] = {0, 0}, bf[2] = {0, 0}, bgtd[2] = {0, 0};
  double f_new = obj(t, d, g_new);
  if (std::isnan(f_new)) {
    f = f_new;
    return;
  }
  double gtd_new = vdot(g_new, d);
  if (std::fabs(gtd_new) > p.gtd_threshold) {
    f = f_new;
    g = g_new;
    return;
  }
  double t_prev = 0.0, f_prev = f, gtd_prev = gtd;
  g_prev = g;
  bool done = false;
  int ls = 0;
  while (ls < p.max_ls) {
    if ((f_new > f + p.c1 * t * gtd) || (ls > 1 && f_new >= f_prev)) {
      br[0] = t_prev; br[1] = t; bf[0] = f_prev; bf[1] = f_new;
      bg[0] = g_prev; bg[1] = g_new; bgtd[0] = gtd_prev; bgtd[1] = gtd_new;
      break;
    }
    if (std::fabs(gtd_new) <= -p.c2 * gtd) {
      br[0] = t; bf[0] = f_new; bg[0] = g_new;
      done = true;
      break;
    }
    if (gtd_new >= 0) {
      br[0] = t_prev; br[1] = t; bf[0] = f_prev; bf[1] = f_new;
      bg[0] = g_prev; bg[1] = g_new; bgtd[0] = gtd_prev; bgtd[1] = gtd_new;
      break;
    }
    double lo = t + 0.01 * (t - t_prev), hi = t * 10.0, tmp = t;
    t = cubic(t_prev, f_prev, gtd_prev, t, f_new, gtd_new, lo, hi);
    t_prev = tmp;
    f_prev = f_new;
    g_prev = g_new;
    gtd_prev = gtd_new;
    f_new = obj(t, d, g_new);
    gtd_new = vdot(g_new, d);
    ls++;
  }
  double d_norm = 0.0;
  if (ls == p.max_ls) {
    br[0] = 0; br[1] = t; bf[0] = f; bf[1] = f_new; bg[0] = g; bg[1] = g_new;
  } else {
    d_norm = vamax(d);
  }
  bool insuf = false;
  int low = 0, high = 1;
  if (!done && bf[0] > bf[1]) {
    low = 1;
    high = 0;
  }
  while (!done && ls < p.max_ls) {
    double lo = std::min(br[0], br[1]), hi = std::max(br[0], br[1]);
    double eps = hi - lo;
    if (eps * d_norm < p.ls_tol_change) break;
    t = cubic(br[0], bf[0], bgtd[0], br[1], bf[1], bgtd[1], lo, hi);
    eps *= 0.1;
    if (std::min(hi - t, t - lo) < eps) {
      if (insuf || t >= hi || t <= lo) {
        t = std::fabs(t - hi) < std::fabs(t - lo) ? hi - eps : lo + eps;
        insuf = false;
      } else {
        insuf = true;
      }
    } else {
      insuf = false;
    }
    f_new = obj(t, d, g_new);
    gtd_new = vdot(g_new, d);
    ls++;
    if (f_new > (f + p.c1 * t * gtd) || f_new >= bf[low]) {
      br[high] = t; bf[high] = f_new; bg[high] = g_new; bgtd[high] = gtd_new;
      if (bf[0] > bf[1]) { low = 1; high = 0; } else { low = 0; high = 1; }
    } else {
      if (std::fabs(gtd_new) <= -p.c2 * gtd) {
        done = true;
      } else if (gtd_new * (br[high] - br[low]) >= 0.0) {
        br[high] = br[low]; bf[high] = bf[low]; bg[high] = bg[low]; bgtd[high] = bgtd[low];
      }
      br[low] = t; bf[low] = f_new; bg[low] = g_new; bgtd[low] = gtd_new;
    }
  }
  t = br[low];
  f = bf[low];
  g = bg[low];
}

bool lbfgs_host(PlainEval& ev, const pxd_lbfgs_params& p, std::vector<double>& x, double& energy) {
  const size_t n = x.size();
  const int m = p.history;
  std::vector<double> g(n), gp(n), y(n), d(n), s(n), q(n);
  std::vector<std::vector<double>> dirs, stps;
  std::vector<double> ro, al(m);
  double loss = ev(x, g);
  energy = loss;
  if (vamax(g) < p.tol_grad) return true;
  bool high = false, conv = false, updated = false;
  int n_iter = 0;
  double H = 1.0, t = 0.0, prev;
  while (n_iter < p.max_iter && ev.nevals < p.max_eval) {
    n_iter++;
    if (n_iter > 1) {
      for (size_t i = 0; i < n; ++i) y[i] = g[i] - gp[i];
      high = vamax(y) > 1e8;
    }
    if (n_iter == 1 || high) {
      for (size_t i = 0; i < n; ++i) d[i] = -g[i];
      dirs.clear();
      stps.clear();
      ro.clear();
      H = 1.0;
    } else {
      for (size_t i = 0; i < n; ++i) s[i] = d[i] * t;
      double ys = vdot(y, s);
      if (ys > 1e-10) {
        if ((int)dirs.size() == m) {
          dirs.erase(dirs.begin());
          stps.erase(stps.begin());
          ro.erase(ro.begin());
        }
        H = ys / vdot(y, y);
        dirs.push_back(y);
        stps.push_back(s);
        ro.push_back(1.0 / ys);
      }
      int k = (int)dirs.size();
      for (size_t i = 0; i < n; ++i) q[i] = -g[i];
      for (int h = k - 1; h >= 0; --h) {
        al[h] = vdot(stps[h], q) * ro[h];
        for (size_t i = 0; i < n; ++i) q[i] -= al[h] * dirs[h][i];
      }
      for (size_t i = 0; i < n; ++i) d[i] = q[i] * H;
      for (int h = 0; h < k; ++h) {
        double be = vdot(dirs[h], d) * ro[h];
        for (size_t i = 0; i < n; ++i) d[i] += stps[h][i] * (al[h] - be);
      }
    }
    gp = g;
    prev = loss;
    if (n_iter == 1 || high) {
      double l1 = 0.0;
      for (double v : g) l1 += std::fabs(v);
      t = std::min(1.0, 1.0 / l1) * p.lr * 0.1;
    } else {
      t = p.lr;
    }
    double gtd = vdot(g, d);
    if (gtd > -p.tol_change) {
      conv = true;
      break;
    }
    Dir obj{ev, x, std::vector<double>(n)};
    strong_wolfe(obj, p, t, d, loss, g, gtd);
    if (std::isnan(loss)) break;
    for (size_t i = 0; i < n; ++i) x[i] += d[i] * t;
    updated = true;
    if (vamax(g) <= p.tol_grad) {
      conv = true;
      break;
    }
    double smax = 0.0;
    for (size_t i = 0; i < n; ++i) smax = std::max(smax, std::fabs(d[i] * t));
    if (smax <= p.tol_change) {
      conv = true;
      break;
    }
    if (std::fabs(loss - prev) < p.tol_change) {
      conv = true;
      break;
    }
  }
  if (updated) energy = ev(x, g);
  return conv;
}

}  // namespace

// Declared in pxdock.h.
extern "C" pxd_status pxd_lbfgs_plain(const pxd_receptor* r, const pxd_grid* g,
                                      const pxd_ligand_batch* b, const pxd_lbfgs_params* prm,
                                      int32_t n_poses, const int32_t* pose_lig,
                                      const int32_t* pose_atom_off, float* xyz, float* tors,
                                      double* energy, int32_t* nevals, int32_t* converged,
                                      void* stream) {
  if (!r || !b || !prm || n_poses < 0 || (n_poses > 0 && (!pose_lig || !pose_atom_off || !xyz)))
    return pxd_set_error_(PXD_ERR_ARG, "lbfgs_plain: bad argument");
  if (prm->dof_mode)
    return pxd_set_error_(PXD_ERR_ARG, "lbfgs_plain: Cartesian variables only (dof_mode 0)");
  if (n_poses == 0) return PXD_OK;
  if (g) {
    for (int t = 0; t < PXD_NUM_VDW_TYPES; ++t)
      if (b->type_used[t] && g->dev.type_slot[t] < 0)
        return pxd_set_error_(PXD_ERR_STATE, "lbfgs_plain: ligand type without a map");
  }
  PlainEval ev;
  ev.r = r;
  ev.G = g ? static_cast<const DevGrid*>(g->mem) : nullptr;
  ev.b = b;
  ev.st = static_cast<hipStream_t>(stream);
  const int nt = r->n_leaves;
  int maxn = 0;
  for (int l = 0; l < b->n; ++l) maxn = std::max(maxn, b->h_natoms[l]);
  hipError_t he = hipMalloc(&ev.d_lig, sizeof(int) * 2);
  if (he == hipSuccess) he = hipMalloc(&ev.d_xyz, sizeof(float) * (3 * maxn * 2 + 2 * (nt + 1)));
  if (he == hipSuccess) he = hipMalloc(&ev.d_e, sizeof(double));
  if (he != hipSuccess) return pxd_set_error_(PXD_ERR_HIP, hipGetErrorString(he));
  ev.d_off = ev.d_lig + 1;
  ev.d_g = ev.d_xyz + 3 * maxn;
  ev.d_tors = ev.d_g + 3 * maxn;
  ev.d_tg = ev.d_tors + (nt + 1);
  // pose-local copies of the inputs (the device buffers belong to the caller)
  std::vector<int> plig(n_poses), poff(n_poses);
  he = hipMemcpy(plig.data(), pose_lig, sizeof(int) * n_poses, hipMemcpyDeviceToHost);
  if (he == hipSuccess) he = hipMemcpy(poff.data(), pose_atom_off, sizeof(int) * n_poses, hipMemcpyDeviceToHost);
  pxd_status st = PXD_OK;
  for (int p = 0; p < n_poses && he == hipSuccess; ++p) {
    int l = plig[p];
    if (l < 0 || l >= b->n) {
      st = pxd_set_error_(PXD_ERR_ARG, "lbfgs_plain: pose ligand out of range");
      break;
    }
    int na = b->h_natoms[l];
    ev.lig = l;
    ev.n3 = 3 * na;
    ev.nt = nt;
    ev.n = ev.n3 + nt;
    ev.hx.assign(ev.n, 0.f);
    ev.hg.assign(ev.n, 0.f);
    ev.nevals = 0;
    int zero = 0;
    he = hipMemcpy(ev.d_lig, &l, sizeof(int), hipMemcpyHostToDevice);
    if (he == hipSuccess) he = hipMemcpy(ev.d_off, &zero, sizeof(int), hipMemcpyHostToDevice);
    std::vector<float> fx(ev.n);
    if (he == hipSuccess)
      he = hipMemcpy(fx.data(), xyz + 3ll * poff[p], sizeof(float) * ev.n3, hipMemcpyDeviceToHost);
    if (he == hipSuccess && nt)
      he = hipMemcpy(fx.data() + ev.n3, tors + (long long)p * nt, sizeof(float) * nt, hipMemcpyDeviceToHost);
    if (he != hipSuccess) break;
    std::vector<double> x(fx.begin(), fx.end());
    double e = 0.0;
    bool conv = lbfgs_host(ev, *prm, x, e);
    if (ev.err != hipSuccess) {
      he = ev.err;
      break;
    }
    for (int i = 0; i < ev.n; ++i) fx[i] = (float)x[i];
    he = hipMemcpy(xyz + 3ll * poff[p], fx.data(), sizeof(float) * ev.n3, hipMemcpyHostToDevice);
    if (he == hipSuccess && nt)
      he = hipMemcpy(tors + (long long)p * nt, fx.data() + ev.n3, sizeof(float) * nt, hipMemcpyHostToDevice);
    int ne = ev.nevals, cv = conv ? 1 : 0;
    if (he == hipSuccess) he = hipMemcpy(energy + p, &e, sizeof(double), hipMemcpyHostToDevice);
    if (he == hipSuccess) he = hipMemcpy(nevals + p, &ne, sizeof(int), hipMemcpyHostToDevice);
    if (he == hipSuccess) he = hipMemcpy(converged + p, &cv, sizeof(int), hipMemcpyHostToDevice);
  }
  (void)hipFree(ev.d_lig);
  (void)hipFree(ev.d_xyz);
  (void)hipFree(ev.d_e);
  if (he != hipSuccess) return pxd_set_error_(PXD_ERR_HIP, hipGetErrorString(he));
  return st;
}
