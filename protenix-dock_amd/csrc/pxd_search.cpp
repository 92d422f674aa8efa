// pxd_search.cpp — C ABI of the device Monte-Carlo search, top-k clustering
// and batched pose parameterisation.  The step loop (one k_mc launch per
// segment of steps, then clustering) is captured once into a hipGraph on a
// private stream and replayed; walker state never leaves the device.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <vector>

#include "pxd_internal.hpp"
#include "pxd_kernels.hpp"
#include "pxdock.h"

using namespace pxd;

extern "C" pxd_status pxd_set_error_(pxd_status st, const char* msg);

namespace {

LbfgsParamsDev to_dev(const pxd_lbfgs_params& p) {
  return {p.history, p.max_iter, p.max_eval, p.max_ls, p.lr, p.tol_grad,
          p.tol_change, p.c1, p.c2, p.ls_tol_change, p.gtd_threshold, 0};
}

bool bad_lbfgs(const pxd_lbfgs_params& p) {
  return p.history < 1 || p.history > kMaxHistory || p.max_iter < 0 || p.max_eval < 0 ||
         p.max_ls < 0;
}

#define HIP_OK(expr)                                                   \
  do {                                                                 \
    hipError_t e_ = (expr);                                            \
    if (e_ != hipSuccess) {                                            \
      err = e_;                                                        \
      goto done;                                                       \
    }                                                                  \
  } while (0)

}  // namespace

extern "C" {

void pxd_mc_default_params(pxd_mc_params* p) {
  memset(p, 0, sizeof(*p));
  p->mmenergy_threshold = -1.0;
  p->seed = 0;
  p->max_nsteps = 200;
  p->mutation = 0;
  p->amplitude = 2.0;
  p->temperature = 1.2;
  p->trans_sigma = 1.0;
  p->rot_sigma = 0.3;
  p->max_tries = 2000;
  p->steps_per_launch = 10;
  pxd_lbfgs_default_params(&p->relax, 40);
  pxd_lbfgs_default_params(&p->full, 240);
}

}  // extern "C"

namespace {

// pxd_mc_search / pxd_mc_search_conformers; conf_count == nullptr: one
// conformer per ligand.
pxd_status mc_search(const pxd_receptor* r, const pxd_grid* g, const pxd_ligand_batch* b,
                     const pxd_score_params* sp, const pxd_mc_params* prm, const pxd_box* box,
                     int32_t walkers, const int32_t* conf_count, const float* lig_xyz0,
                     float* best_xyz, float* best_tors, double* best_energy,
                     double* best_pscore, int32_t* nevals, int32_t num_modes, double min_rmsd,
                     int32_t* cluster_idx, int32_t* cluster_count, void* stream) {
  if (!r || !b || !sp || !prm || !box || !lig_xyz0 || !best_xyz || !best_energy ||
      !best_pscore || !nevals)
    return pxd_set_error_(PXD_ERR_ARG, "mc_search: null argument");
  if (conf_count)
    for (int l = 0; l < b->n; ++l)
      if (conf_count[l] < 1)
        return pxd_set_error_(PXD_ERR_ARG, "mc_search: every ligand needs a conformer");
  if (walkers < 1) return pxd_set_error_(PXD_ERR_ARG, "mc_search: walkers < 1");
  if (r->n_leaves > 0 && !best_tors) return pxd_set_error_(PXD_ERR_ARG, "mc_search: best_tors required");
  if (bad_lbfgs(prm->relax) || bad_lbfgs(prm->full))
    return pxd_set_error_(PXD_ERR_RANGE, "mc_search: L-BFGS parameters out of range");
  if (prm->relax.dof_mode || prm->full.dof_mode)
    return pxd_set_error_(PXD_ERR_ARG, "mc_search: the search optimises Cartesian variables "
                                       "(dof_mode 0)");
  if (prm->max_nsteps < 0 || prm->max_tries < 1 || prm->steps_per_launch < 1)
    return pxd_set_error_(PXD_ERR_ARG, "mc_search: bad step parameters");
  if (num_modes < 0 || num_modes > 64)
    return pxd_set_error_(PXD_ERR_RANGE, "mc_search: num_modes outside 0..64");
  if (num_modes > 0 && (!cluster_idx || !cluster_count))
    return pxd_set_error_(PXD_ERR_ARG, "mc_search: cluster outputs required");
  if (r->dev.n_ring5 + r->dev.n_ring6 > 256)
    return pxd_set_error_(PXD_ERR_RANGE, "mc_search: receptor has more than 256 pi rings");
  if (g) {
    for (int t = 0; t < PXD_NUM_VDW_TYPES; ++t)
      if (b->type_used[t] && g->dev.type_slot[t] < 0)
        return pxd_set_error_(PXD_ERR_STATE, "mc_search: ligand type without a map");
  }
  for (int l = 0; l < b->n; ++l)
    if (b->h_ntors[l] + 1 > PXD_MAX_FRAGMENTS)
      return pxd_set_error_(PXD_ERR_RANGE, "mc_search: too many torsions");

  const int nl = b->n, W = nl * walkers, L = r->n_leaves;
  std::vector<int> wlig(W), wpid(W), woff(W + 1), lwoff(nl + 1);
  long long atoms = 0;
  for (int l = 0, w = 0; l < nl; ++l) {
    lwoff[l] = w;
    for (int k = 0; k < walkers; ++k, ++w) {
      wlig[w] = l;
      wpid[w] = k;
      woff[w] = (int)atoms;
      atoms += b->h_natoms[l];
    }
  }
  lwoff[nl] = W;
  woff[W] = (int)atoms;
  std::vector<int> cnum(nl), coff(nl);
  for (int l = 0, a = 0; l < nl; ++l) {
    cnum[l] = conf_count ? conf_count[l] : 1;
    coff[l] = a;
    a += cnum[l] * b->h_natoms[l];
  }
  if (atoms > (1ll << 30)) return pxd_set_error_(PXD_ERR_RANGE, "mc_search: too many atoms");
  int dev = 0, ncu = 0;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
    return pxd_set_error_(PXD_ERR_HIP, "mc_search: no current HIP device");
  const int n_slots = std::min(W, std::max(1, ncu) * 4 * kMcWaves);
  const int hist_m = std::max(prm->relax.history, prm->full.history);
  const int n_launch = (prm->max_nsteps + prm->steps_per_launch - 1) / prm->steps_per_launch + 1;
  // walkers grouped by L-BFGS size class (their ligand's variable count), in
  // walker order, so that a class's launches run its own walkers only; each
  // class's step loop is one captured graph, the graphs replayed side by side
  // on streams of their own; work slots shared out by each class's walkers
  // x vector width
  int cls_n[kLbfgsClasses] = {0, 0, 0}, cls_slots[kLbfgsClasses] = {0, 0, 0};
  std::vector<int> order;
  order.reserve(W);
  {
    std::vector<int8_t> lcls(nl);
    for (int l = 0; l < nl; ++l) {
      const int nk = (3 * b->h_natoms[l] + L + kWave - 1) / kWave;
      lcls[l] = nk <= kLbfgsClassVpl[0] ? 0 : (nk <= kLbfgsClassVpl[1] ? 1 : 2);
      cls_n[lcls[l]] += walkers;
    }
    for (int c = 0; c < kLbfgsClasses; ++c)
      for (int w = 0; w < W; ++w)
        if (lcls[wlig[w]] == c) order.push_back(w);
    double wsum = 0.0;
    for (int c = 0; c < kLbfgsClasses; ++c) wsum += (double)cls_n[c] * kLbfgsClassVpl[c];
    for (int c = 0; c < kLbfgsClasses; ++c)
      if (cls_n[c])
        cls_slots[c] = std::max(1, std::min(cls_n[c], (int)std::lround((double)n_slots * cls_n[c] *
                                                                       kLbfgsClassVpl[c] / wsum)));
  }
  const size_t slot_floats = lbfgs_slot_floats(hist_m);
  const size_t hist_floats = (size_t)(cls_slots[0] + cls_slots[1] + cls_slots[2]) * slot_floats;
  hipError_t err = hipSuccess;
  hipStream_t user = static_cast<hipStream_t>(stream);
  hipEvent_t evf = nullptr, evj[kLbfgsClasses] = {};
  hipStream_t side[kLbfgsClasses] = {}, pool[kSideStreams] = {};
  hipGraph_t graph[kLbfgsClasses] = {};
  hipGraphExec_t exec[kLbfgsClasses] = {};
  // the class with the most work replays on the caller's stream, the others
  // on the device's two shared side streams (no more streams than the
  // process's hardware queues)
  int main_c = 0;
  for (int c = 1; c < kLbfgsClasses; ++c)
    if ((double)cls_n[c] * kLbfgsClassVpl[c] > (double)cls_n[main_c] * kLbfgsClassVpl[main_c])
      main_c = c;
  char* mem = nullptr;
  pxd_score_params* d_sp = nullptr;
  int *d_wlig, *d_wpid, *d_woff, *d_lwoff, *d_cnt, *d_coff, *d_cnum, *d_order;
  unsigned long long* d_emin;
  float *d_cur, *d_curt, *d_hist;
  WalkerState* d_st;
  size_t off = 0;
  auto take = [&](size_t bytes) {
    size_t o = off;
    off += (bytes + 255) & ~size_t(255);
    return o;
  };
  const size_t o_sp = take(sizeof(pxd_score_params)), o_wlig = take(4 * W), o_wpid = take(4 * W),
               o_woff = take(4 * (W + 1)), o_lwoff = take(4 * (nl + 1)),
               o_cnt = take(4 * (size_t)n_launch * kLbfgsClasses), o_cur = take(12 * (size_t)atoms),
               o_curt = take(4 * (size_t)W * (L + 1)), o_st = take(sizeof(WalkerState) * W),
               o_hist = take(4 * hist_floats), o_coff = take(4 * nl), o_cnum = take(4 * nl),
               o_emin = take(8 * (size_t)nl), o_order = take(4 * (size_t)W);
  std::unique_lock<std::mutex> side_lock(side_stream_mutex());
  HIP_OK(hipMalloc(&mem, off));
  d_sp = reinterpret_cast<pxd_score_params*>(mem + o_sp);
  d_wlig = reinterpret_cast<int*>(mem + o_wlig);
  d_wpid = reinterpret_cast<int*>(mem + o_wpid);
  d_woff = reinterpret_cast<int*>(mem + o_woff);
  d_lwoff = reinterpret_cast<int*>(mem + o_lwoff);
  d_cnt = reinterpret_cast<int*>(mem + o_cnt);
  d_cur = reinterpret_cast<float*>(mem + o_cur);
  d_curt = reinterpret_cast<float*>(mem + o_curt);
  d_st = reinterpret_cast<WalkerState*>(mem + o_st);
  d_hist = reinterpret_cast<float*>(mem + o_hist);
  d_coff = reinterpret_cast<int*>(mem + o_coff);
  d_cnum = reinterpret_cast<int*>(mem + o_cnum);
  d_emin = reinterpret_cast<unsigned long long*>(mem + o_emin);
  d_order = reinterpret_cast<int*>(mem + o_order);
  HIP_OK(hipMemcpy(d_sp, sp, sizeof(*sp), hipMemcpyHostToDevice));
  HIP_OK(hipMemcpy(d_wlig, wlig.data(), 4 * W, hipMemcpyHostToDevice));
  HIP_OK(hipMemcpy(d_wpid, wpid.data(), 4 * W, hipMemcpyHostToDevice));
  HIP_OK(hipMemcpy(d_woff, woff.data(), 4 * (W + 1), hipMemcpyHostToDevice));
  HIP_OK(hipMemcpy(d_lwoff, lwoff.data(), 4 * (nl + 1), hipMemcpyHostToDevice));
  HIP_OK(hipMemcpy(d_coff, coff.data(), 4 * nl, hipMemcpyHostToDevice));
  HIP_OK(hipMemcpy(d_cnum, cnum.data(), 4 * nl, hipMemcpyHostToDevice));
  HIP_OK(hipMemcpy(d_order, order.data(), 4 * (size_t)W, hipMemcpyHostToDevice));
  HIP_OK(side_streams(pool));
  HIP_OK(hipEventCreateWithFlags(&evf, hipEventDisableTiming));
  for (int c = 0, k = 0; c < kLbfgsClasses; ++c)
    if (cls_n[c] && c != main_c) {
      side[c] = pool[k++];
      HIP_OK(hipEventCreateWithFlags(&evj[c], hipEventDisableTiming));
    }
  {
    McArgs a;
    memset(&a, 0, sizeof(a));
    a.prm.seed_lo = (uint32_t)prm->seed;
    a.prm.seed_hi = (uint32_t)(prm->seed >> 32);
    a.prm.max_nsteps = prm->max_nsteps;
    a.prm.mutation = prm->mutation;
    a.prm.amplitude = (float)prm->amplitude;
    a.prm.temperature = (float)prm->temperature;
    a.prm.trans_sigma = (float)prm->trans_sigma;
    a.prm.rot_sigma = (float)prm->rot_sigma;
    a.prm.max_tries = prm->max_tries;
    a.prm.relax = to_dev(prm->relax);
    a.prm.full = to_dev(prm->full);
    a.prm.mm_threshold = prm->mmenergy_threshold;
    a.sp = d_sp;
    for (int d = 0; d < 3; ++d) {
      a.box_lo[d] = (float)box->init_xyz[d];
      a.box_hi[d] = (float)box->oppo_xyz[d];
    }
    a.n_walkers = W;
    a.walker_lig = d_wlig;
    a.walker_pid = d_wpid;
    a.walker_off = d_woff;
    a.lig_xyz0 = lig_xyz0;
    a.lig_conf_off = d_coff;
    a.lig_conf_n = d_cnum;
    a.lig_emin = d_emin;
    a.cur_xyz = d_cur;
    a.best_xyz = best_xyz;
    a.cur_tors = d_curt;
    a.best_tors = best_tors ? best_tors : d_curt;
    a.st = d_st;
    a.hist = d_hist;
    a.hist_m = hist_m;
    a.max_atoms = b->max_atoms;
    a.max_frags = b->max_frags;
    // each class's step loop captured once: one k_mc launch per segment of
    // steps over the class's walkers, with its own queue heads, work slots
    // and history scratch
    size_t hist_off = 0;
    for (int c = 0, off = 0; c < kLbfgsClasses; ++c) {
      McArgs ac = a;
      ac.order = d_order + off;
      ac.n_order = cls_n[c];
      ac.hist = d_hist + hist_off;
      off += cls_n[c];
      hist_off += slot_floats * (size_t)cls_slots[c];
      if (!cls_n[c]) continue;
      // captured on a shared side stream (the caller's may be the null
      // stream, which cannot capture); replayed below
      hipStream_t cap = pool[0];
      HIP_OK(hipStreamBeginCapture(cap, hipStreamCaptureModeThreadLocal));
      for (int k = 0; k < n_launch; ++k) {
        ac.step_end = std::min(prm->max_nsteps, k * prm->steps_per_launch);
        ac.counter = d_cnt + k * kLbfgsClasses + c;
        launch_mc_class(r->dev, b->dev, g ? static_cast<const DevGrid*>(g->mem) : nullptr, ac, c,
                        cls_slots[c], cap);
      }
      err = hipStreamEndCapture(cap, &graph[c]);
      if (err != hipSuccess) goto done;
      HIP_OK(hipGraphInstantiate(&exec[c], graph[c], nullptr, nullptr, 0));
    }
  }
  // reset state and queue heads with fill kernels (no memset), then the
  // class graphs side by side, then the clustering
  launch_fill_int(reinterpret_cast<int*>(d_st), (int)(sizeof(WalkerState) / 4 * W), 0, user);
  launch_fill_int(d_cnt, n_launch * kLbfgsClasses, 0, user);
  // the ligands' lowest pre-optimisation energies start at 1e8 (reader.cpp:51-53)
  launch_fill_u64(d_emin, nl, order_key(1e8), user);
  HIP_OK(hipEventRecord(evf, user));
  for (int c = 0; c < kLbfgsClasses; ++c) {
    if (!cls_n[c] || c == main_c) continue;
    HIP_OK(hipStreamWaitEvent(side[c], evf, 0));
    HIP_OK(hipGraphLaunch(exec[c], side[c]));
    HIP_OK(hipEventRecord(evj[c], side[c]));
  }
  if (cls_n[main_c]) HIP_OK(hipGraphLaunch(exec[main_c], user));
  for (int c = 0; c < kLbfgsClasses; ++c)
    if (cls_n[c] && c != main_c) HIP_OK(hipStreamWaitEvent(user, evj[c], 0));
  if (num_modes > 0) {
    ClusterArgs ca{nl, d_lwoff, d_woff, d_wlig, best_xyz, d_st, num_modes, (float)min_rmsd,
                   cluster_idx, cluster_count};
    launch_cluster(ca, 0, user);
  }
  HIP_OK(hipGetLastError());
  HIP_OK(hipStreamSynchronize(user));
  {
    // per-walker scalars out of the walker state
    std::vector<WalkerState> st(W);
    HIP_OK(hipMemcpy(st.data(), d_st, sizeof(WalkerState) * W, hipMemcpyDeviceToHost));
    std::vector<double> be(W), bp(W);
    std::vector<int> ne(W);
    for (int w = 0; w < W; ++w) {
      be[w] = st[w].best_e;
      bp[w] = st[w].best_ps;
      ne[w] = st[w].nevals;
    }
    HIP_OK(hipMemcpy(best_energy, be.data(), 8 * W, hipMemcpyHostToDevice));
    HIP_OK(hipMemcpy(best_pscore, bp.data(), 8 * W, hipMemcpyHostToDevice));
    HIP_OK(hipMemcpy(nevals, ne.data(), 4 * W, hipMemcpyHostToDevice));
  }
done:
  for (int c = 0; c < kLbfgsClasses; ++c) {
    if (exec[c]) (void)hipGraphExecDestroy(exec[c]);
    if (graph[c]) (void)hipGraphDestroy(graph[c]);
    // the side streams are kept; their work must finish before the memory goes
    if (side[c]) (void)hipStreamSynchronize(side[c]);
    if (evj[c]) (void)hipEventDestroy(evj[c]);
  }
  if (evf) (void)hipEventDestroy(evf);
  (void)hipStreamSynchronize(user);

  if (mem) (void)hipFree(mem);
  if (err != hipSuccess) return pxd_set_error_(PXD_ERR_HIP, hipGetErrorString(err));
  return PXD_OK;
}

}  // namespace

extern "C" {

pxd_status pxd_mc_search(const pxd_receptor* r, const pxd_grid* g, const pxd_ligand_batch* b,
                         const pxd_score_params* sp, const pxd_mc_params* prm, const pxd_box* box,
                         int32_t walkers, const float* lig_xyz0, float* best_xyz,
                         float* best_tors, double* best_energy, double* best_pscore,
                         int32_t* nevals, int32_t num_modes, double min_rmsd,
                         int32_t* cluster_idx, int32_t* cluster_count, void* stream) {
  return mc_search(r, g, b, sp, prm, box, walkers, nullptr, lig_xyz0, best_xyz, best_tors,
                   best_energy, best_pscore, nevals, num_modes, min_rmsd, cluster_idx,
                   cluster_count, stream);
}

pxd_status pxd_mc_search_conformers(const pxd_receptor* r, const pxd_grid* g,
                                    const pxd_ligand_batch* b, const pxd_score_params* sp,
                                    const pxd_mc_params* prm, const pxd_box* box,
                                    int32_t walkers, const int32_t* conf_count,
                                    const float* conformers, float* best_xyz, float* best_tors,
                                    double* best_energy, double* best_pscore, int32_t* nevals,
                                    int32_t num_modes, double min_rmsd, int32_t* cluster_idx,
                                    int32_t* cluster_count, void* stream) {
  if (!conf_count) return pxd_set_error_(PXD_ERR_ARG, "mc_search_conformers: null conf_count");
  return mc_search(r, g, b, sp, prm, box, walkers, conf_count, conformers, best_xyz, best_tors,
                   best_energy, best_pscore, nevals, num_modes, min_rmsd, cluster_idx,
                   cluster_count, stream);
}

pxd_status pxd_apply_poses(const pxd_ligand_batch* b, int32_t n_poses, const int32_t* pose_lig,
                           const int32_t* pose_atom_off, const float* xyz_in, const float* center,
                           const float* rot, const int32_t* tors_off, const float* tors,
                           float* xyz_out, void* stream) {
  if (!b || n_poses < 0) return pxd_set_error_(PXD_ERR_ARG, "apply_poses: bad argument");
  if (n_poses == 0) return PXD_OK;
  if (!pose_lig || !pose_atom_off || !xyz_in || !center || !rot || !tors_off || !xyz_out)
    return pxd_set_error_(PXD_ERR_ARG, "apply_poses: null array");
  ApplyArgs a{n_poses, pose_lig, pose_atom_off, xyz_in, center, rot, tors_off, tors, xyz_out};
  launch_apply(b->dev, a, static_cast<hipStream_t>(stream));
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return pxd_set_error_(PXD_ERR_HIP, hipGetErrorString(e));
  return PXD_OK;
}

}  // extern "C"
