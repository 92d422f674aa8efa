// ref_glue.cpp — C entry points over the reference's own compiled sources.
//
// TEST INFRASTRUCTURE ONLY.  Built by oracle/Makefile together with the
// reference files that compile with the C++ standard library alone
// (engine/src/core/scorer.cpp, engine/src/lib/math.cpp,
// engine/src/lib/random.cpp and the header-only engine/include/bytedock/core/pair.h),
// straight from where they lie under the reference tree, into oracle/_ref/.
// The rest of the reference path (interaction.cpp, grid.cpp, optimizer.cpp,
// data.cpp) includes Boost and is not built.  This glue only converts the
// pxdock.h descriptors into the reference's own structs and calls the
// reference code; it restates no algorithm.
#include <cstdio>
#include <cstring>
#include <string>
#include <unordered_map>

#include "bytedock/core/constant.h"
#include "bytedock/core/pair.h"
#include "bytedock/core/scorer.h"
#include "bytedock/lib/math.h"
#include "bytedock/lib/random.h"
#include "json.hpp"
#include "pxdock.h"

using namespace bytedock;

namespace {

template <class T>
void fill_ids(std::vector<T>& dst, const int32_t* src, int n, int width) {
  dst.resize(n);
  for (int i = 0; i < n; ++i)
    for (int k = 0; k < width; ++k) dst[i].ids[k] = src[width * i + k];
}

void fill_list(std::vector<index_t>& dst, const int32_t* src, int n) {
  dst.assign(src, src + n);
}

force_field_params to_ffp(const pxd_mol* m) {
  force_field_params f;
  if (m->vdw_type) {
    f.vdw_types.assign(m->vdw_type, m->vdw_type + m->n_atoms);
  } else {
    f.vdw_params.resize(m->n_atoms);
    for (int i = 0; i < m->n_atoms; ++i) f.vdw_params[i] = {m->sigma[i], m->epsilon[i]};
  }
  f.partial_charges.assign(m->charge, m->charge + m->n_atoms);
  fill_list(f.hydrophobic_atoms, m->hydrophobic, m->n_hydrophobic);
  fill_list(f.cation_atoms, m->cation, m->n_cation);
  fill_list(f.anion_atoms, m->anion, m->n_anion);
  f.pi5_rings.resize(m->n_ring5);
  for (int r = 0; r < m->n_ring5; ++r) {
    for (int k = 0; k < 5; ++k) f.pi5_rings[r].ids[k] = m->ring5[5 * r + k];
    f.pi5_rings[r].size = 5;
  }
  f.pi6_rings.resize(m->n_ring6);
  for (int r = 0; r < m->n_ring6; ++r) {
    for (int k = 0; k < 6; ++k) f.pi6_rings[r].ids[k] = m->ring6[6 * r + k];
    f.pi6_rings[r].size = 6;
  }
  fill_ids(f.hbonddon_charged, m->don_charged, m->n_don_charged, 2);
  fill_ids(f.hbonddon_neutral, m->don_neutral, m->n_don_neutral, 2);
  fill_list(f.hbondacc_charged, m->acc_charged, m->n_acc_charged);
  fill_list(f.hbondacc_neutral, m->acc_neutral, m->n_acc_neutral);
  f.bad_torsions.resize(m->n_tstrain);
  for (int p = 0; p < m->n_tstrain; ++p) {
    for (int k = 0; k < 4; ++k) f.bad_torsions[p].coefs[k] = m->tstrain_coef[4 * p + k];
    int a = m->tstrain_match_off[p], b = m->tstrain_match_off[p + 1];
    fill_ids(f.bad_torsions[p].matched, m->tstrain_ijkl + 4 * a, b - a, 4);
  }
  return f;
}

molecule_pose to_pose(const double* xyz, int n) {
  molecule_pose p(n);
  for (int i = 0; i < n; ++i)
    for (int k = 0; k < 3; ++k) p[i].xyz[k] = xyz[3 * i + k];
  return p;
}

void set_hb(HbondInteractionParams& h, const double* s) {
  h.dist_min = s[0];
  h.dist_max = s[1];
  h.angle_min = s[2];
  h.angle_max = s[3];
  h.dist_center = s[4];
  h.angle_center = s[5];
}

void set_ps(PiStackingInteractionParams& h, const double* s) {
  h.dist_min = s[0];
  h.dist_max = s[1];
  h.angle_min = s[2];
  h.angle_max = s[3];
  h.sigma_min = s[4];
  h.sigma_max = s[5];
}

void set_stack(PiStackingEnergyParams& e, const double* w, const double (*s)[6]) {
  e.l5r5_coef = w[0];
  e.l5r6_coef = w[1];
  e.l6r5_coef = w[2];
  e.l6r6_coef = w[3];
  set_ps(e.l5r5_inter, s[0]);
  set_ps(e.l5r6_inter, s[1]);
  set_ps(e.l6r5_inter, s[2]);
  set_ps(e.l6r6_inter, s[3]);
}

ScoringFunctionParams to_sfp(const pxd_score_params* sp) {
  ScoringFunctionParams f{};
  f.ligand_ff = ByteFF1;
  f.interaction_coulomb_energy_coef = sp->coul_coef;
  f.interaction_vdw_energy_coef = sp->vdw_coef;
  f.ionic_energy_coef = sp->ionic_coef;
  f.pi_cationic_energy_coef = sp->pication_coef;
  f.f2f_pi_stacking_energy_coef = sp->f2f_coef;
  f.e2f_pi_stacking_energy_coef = sp->e2f_coef;
  f.hbond_energy_coef = sp->hbond_coef;
  f.hydrophobic_energy_coef = sp->hydrophobic_coef;
  f.torsion_strain_penalty_coef = sp->tstrain_coef;
  f.interaction_vdw_energy.scale = sp->vdw_scale;
  f.interaction_vdw_energy.cutoff = sp->vdw_cutoff;
  f.interaction_vdw_energy.types.resize(kNumVdwTypes);
  for (size_t t = 0; t < kNumVdwTypes; ++t)
    f.interaction_vdw_energy.types[t] = {sp->vdw_sigma_scale[t], sp->vdw_epsilon_scale[t]};
  f.ionic_energy.lc_coef = sp->ionic_lc_coef;
  f.ionic_energy.pc_coef = sp->ionic_pc_coef;
  f.ionic_energy.lc_inter = {sp->ionic_lc[0], sp->ionic_lc[1]};
  f.ionic_energy.pc_inter = {sp->ionic_pc[0], sp->ionic_pc[1]};
  auto& pc = f.pi_cationic_energy;
  pc.lc5_coef = sp->pication_sub_coef[0];
  pc.lc6_coef = sp->pication_sub_coef[1];
  pc.pc5_coef = sp->pication_sub_coef[2];
  pc.pc6_coef = sp->pication_sub_coef[3];
  pc.lc5_inter = {sp->pication_range[0][0], sp->pication_range[0][1]};
  pc.lc6_inter = {sp->pication_range[1][0], sp->pication_range[1][1]};
  pc.pc5_inter = {sp->pication_range[2][0], sp->pication_range[2][1]};
  pc.pc6_inter = {sp->pication_range[3][0], sp->pication_range[3][1]};
  set_stack(f.f2f_pi_stacking_energy, sp->f2f_sub_coef, sp->f2f_inter);
  set_stack(f.e2f_pi_stacking_energy, sp->e2f_sub_coef, sp->e2f_inter);
  auto& hb = f.hbond_energy;
  hb.scale = sp->hbond_scale;
  hb.hldcc_coef = sp->hbond_sub_coef[0];
  hb.hldcn_coef = sp->hbond_sub_coef[1];
  hb.hldnc_coef = sp->hbond_sub_coef[2];
  hb.hldnn_coef = sp->hbond_sub_coef[3];
  hb.hpdcc_coef = sp->hbond_sub_coef[4];
  hb.hpdcn_coef = sp->hbond_sub_coef[5];
  hb.hpdnc_coef = sp->hbond_sub_coef[6];
  hb.hpdnn_coef = sp->hbond_sub_coef[7];
  set_hb(hb.hldcc_inter, sp->hbond_inter[0]);
  set_hb(hb.hldcn_inter, sp->hbond_inter[1]);
  set_hb(hb.hldnc_inter, sp->hbond_inter[2]);
  set_hb(hb.hldnn_inter, sp->hbond_inter[3]);
  set_hb(hb.hpdcc_inter, sp->hbond_inter[4]);
  set_hb(hb.hpdcn_inter, sp->hbond_inter[5]);
  set_hb(hb.hpdnc_inter, sp->hbond_inter[6]);
  set_hb(hb.hpdnn_inter, sp->hbond_inter[7]);
  f.hydrophobic_energy.scale = sp->hydro_scale;
  f.hydrophobic_energy.dist_min = sp->hydro_dist_min;
  f.hydrophobic_energy.dist_max = sp->hydro_dist_max;
  f.hydrophobic_energy.scale_factor = sp->hydro_scale_factor;
  return f;
}

// ScoringFunctionParams straight from a pose_selection JSON tree.  Every
// field is read under the key that spells its member name, which is how the
// reference's loader names them (factory.cpp:33-211): the stringified member
// name is the key, so this mapping cannot disagree with the reference's
// structs.  Sub-objects are read only for active terms (utility.h:24-26).
#define PXD_RD(node, dst, name) (dst).name = (node)[#name].num()
#define PXD_RD0(node, dst, name) (dst).name = (node).has(#name) ? (node)[#name].num() : 0.0

bool active_coef(double c) { return c > 1e-5 || c < -1e-5; }

void rd_range2(const pxdock::json& n, IonicInteractionParams& d) {
  PXD_RD(n, d, dist_min);
  PXD_RD(n, d, dist_max);
}
void rd_range2(const pxdock::json& n, CationPiInteractionParams& d) {
  PXD_RD(n, d, dist_min);
  PXD_RD(n, d, dist_max);
}
void rd_stack_inter(const pxdock::json& n, PiStackingInteractionParams& d) {
  PXD_RD(n, d, dist_min);
  PXD_RD(n, d, dist_max);
  PXD_RD(n, d, angle_min);
  PXD_RD(n, d, angle_max);
  PXD_RD(n, d, sigma_min);
  PXD_RD(n, d, sigma_max);
}
void rd_stack(const pxdock::json& n, PiStackingEnergyParams& d) {
  PXD_RD(n, d, l5r5_coef);
  PXD_RD(n, d, l5r6_coef);
  PXD_RD(n, d, l6r5_coef);
  PXD_RD(n, d, l6r6_coef);
  rd_stack_inter(n["l5r5_inter"], d.l5r5_inter);
  rd_stack_inter(n["l5r6_inter"], d.l5r6_inter);
  rd_stack_inter(n["l6r5_inter"], d.l6r5_inter);
  rd_stack_inter(n["l6r6_inter"], d.l6r6_inter);
}
void rd_hb_inter(const pxdock::json& n, HbondInteractionParams& d) {
  PXD_RD(n, d, dist_min);
  PXD_RD(n, d, dist_max);
  PXD_RD(n, d, angle_min);
  PXD_RD(n, d, angle_max);
  PXD_RD(n, d, dist_center);
  PXD_RD(n, d, angle_center);
}

ScoringFunctionParams sfp_from_json(const pxdock::json& n) {
  ScoringFunctionParams f{};
  f.ligand_ff = ByteFF1;
  PXD_RD0(n, f, interaction_coulomb_energy_coef);
  PXD_RD0(n, f, interaction_vdw_energy_coef);
  PXD_RD0(n, f, ionic_energy_coef);
  PXD_RD0(n, f, pi_cationic_energy_coef);
  PXD_RD0(n, f, f2f_pi_stacking_energy_coef);
  PXD_RD0(n, f, e2f_pi_stacking_energy_coef);
  PXD_RD0(n, f, hbond_energy_coef);
  PXD_RD0(n, f, hydrophobic_energy_coef);
  PXD_RD0(n, f, torsion_strain_penalty_coef);
  PXD_RD0(n, f, rotatable_energy_coef);
  PXD_RD0(n, f, gbsa_energy_delta_coef);
  if (active_coef(f.interaction_vdw_energy_coef)) {
    const auto& v = n["interaction_vdw_energy"];
    PXD_RD(v, f.interaction_vdw_energy, scale);
    PXD_RD(v, f.interaction_vdw_energy, cutoff);
    const auto& ty = v["types"];
    // the leaf's constructor reads 99 entries (scorer.cpp:210-217); a shorter
    // list (the default affinity tree has 38) is padded with scale 1 here so
    // the reference code reads defined values
    const size_t nt = ty.size() < 99 ? 99 : ty.size();
    f.interaction_vdw_energy.types.resize(nt);
    for (size_t t = 0; t < nt; ++t) {
      if (t < ty.size()) {
        PXD_RD(ty[t], f.interaction_vdw_energy.types[t], sigma_scale);
        PXD_RD(ty[t], f.interaction_vdw_energy.types[t], epsilon_scale);
      } else {
        f.interaction_vdw_energy.types[t].sigma_scale = 1.0;
        f.interaction_vdw_energy.types[t].epsilon_scale = 1.0;
      }
    }
  }
  if (active_coef(f.ionic_energy_coef)) {
    const auto& v = n["ionic_energy"];
    PXD_RD(v, f.ionic_energy, lc_coef);
    PXD_RD(v, f.ionic_energy, pc_coef);
    rd_range2(v["lc_inter"], f.ionic_energy.lc_inter);
    rd_range2(v["pc_inter"], f.ionic_energy.pc_inter);
  }
  if (active_coef(f.pi_cationic_energy_coef)) {
    const auto& v = n["pi_cationic_energy"];
    auto& d = f.pi_cationic_energy;
    PXD_RD(v, d, lc5_coef);
    PXD_RD(v, d, lc6_coef);
    PXD_RD(v, d, pc5_coef);
    PXD_RD(v, d, pc6_coef);
    rd_range2(v["lc5_inter"], d.lc5_inter);
    rd_range2(v["lc6_inter"], d.lc6_inter);
    rd_range2(v["pc5_inter"], d.pc5_inter);
    rd_range2(v["pc6_inter"], d.pc6_inter);
  }
  if (active_coef(f.f2f_pi_stacking_energy_coef))
    rd_stack(n["f2f_pi_stacking_energy"], f.f2f_pi_stacking_energy);
  if (active_coef(f.e2f_pi_stacking_energy_coef))
    rd_stack(n["e2f_pi_stacking_energy"], f.e2f_pi_stacking_energy);
  if (active_coef(f.hbond_energy_coef)) {
    const auto& v = n["hbond_energy"];
    auto& d = f.hbond_energy;
    PXD_RD(v, d, scale);
    PXD_RD(v, d, hldcc_coef);
    PXD_RD(v, d, hldcn_coef);
    PXD_RD(v, d, hldnc_coef);
    PXD_RD(v, d, hldnn_coef);
    PXD_RD(v, d, hpdcc_coef);
    PXD_RD(v, d, hpdcn_coef);
    PXD_RD(v, d, hpdnc_coef);
    PXD_RD(v, d, hpdnn_coef);
    rd_hb_inter(v["hldcc_inter"], d.hldcc_inter);
    rd_hb_inter(v["hldcn_inter"], d.hldcn_inter);
    rd_hb_inter(v["hldnc_inter"], d.hldnc_inter);
    rd_hb_inter(v["hldnn_inter"], d.hldnn_inter);
    rd_hb_inter(v["hpdcc_inter"], d.hpdcc_inter);
    rd_hb_inter(v["hpdcn_inter"], d.hpdcn_inter);
    rd_hb_inter(v["hpdnc_inter"], d.hpdnc_inter);
    rd_hb_inter(v["hpdnn_inter"], d.hpdnn_inter);
  }
  if (active_coef(f.hydrophobic_energy_coef)) {
    const auto& v = n["hydrophobic_energy"];
    PXD_RD(v, f.hydrophobic_energy, scale);
    PXD_RD(v, f.hydrophobic_energy, dist_min);
    PXD_RD(v, f.hydrophobic_energy, dist_max);
    PXD_RD(v, f.hydrophobic_energy, scale_factor);
  }
  return f;
}

// Field-by-field comparison of two parameter sets of the active terms; the
// name of the first mismatch goes to msg.
#define PXD_CMP(a, b, path)                                              \
  do {                                                                   \
    if (!((a).path == (b).path)) {                                       \
      if (!nbad) std::snprintf(msg, msg_len, "%s", #path);              \
      ++nbad;                                                            \
    }                                                                    \
  } while (0)

int compare_sfp(const ScoringFunctionParams& a, const ScoringFunctionParams& b, char* msg,
                size_t msg_len) {
  int nbad = 0;
  PXD_CMP(a, b, interaction_coulomb_energy_coef);
  PXD_CMP(a, b, interaction_vdw_energy_coef);
  PXD_CMP(a, b, ionic_energy_coef);
  PXD_CMP(a, b, pi_cationic_energy_coef);
  PXD_CMP(a, b, f2f_pi_stacking_energy_coef);
  PXD_CMP(a, b, e2f_pi_stacking_energy_coef);
  PXD_CMP(a, b, hbond_energy_coef);
  PXD_CMP(a, b, hydrophobic_energy_coef);
  PXD_CMP(a, b, torsion_strain_penalty_coef);
  if (active_coef(a.interaction_vdw_energy_coef)) {
    PXD_CMP(a, b, interaction_vdw_energy.scale);
    PXD_CMP(a, b, interaction_vdw_energy.cutoff);
    for (size_t t = 0; t < kNumVdwTypes; ++t) {
      PXD_CMP(a, b, interaction_vdw_energy.types[t].sigma_scale);
      PXD_CMP(a, b, interaction_vdw_energy.types[t].epsilon_scale);
    }
  }
  if (active_coef(a.ionic_energy_coef)) {
    PXD_CMP(a, b, ionic_energy.lc_coef);
    PXD_CMP(a, b, ionic_energy.pc_coef);
    PXD_CMP(a, b, ionic_energy.lc_inter.dist_min);
    PXD_CMP(a, b, ionic_energy.lc_inter.dist_max);
    PXD_CMP(a, b, ionic_energy.pc_inter.dist_min);
    PXD_CMP(a, b, ionic_energy.pc_inter.dist_max);
  }
  if (active_coef(a.pi_cationic_energy_coef)) {
    PXD_CMP(a, b, pi_cationic_energy.lc5_coef);
    PXD_CMP(a, b, pi_cationic_energy.lc6_coef);
    PXD_CMP(a, b, pi_cationic_energy.pc5_coef);
    PXD_CMP(a, b, pi_cationic_energy.pc6_coef);
    PXD_CMP(a, b, pi_cationic_energy.lc5_inter.dist_min);
    PXD_CMP(a, b, pi_cationic_energy.lc5_inter.dist_max);
    PXD_CMP(a, b, pi_cationic_energy.lc6_inter.dist_min);
    PXD_CMP(a, b, pi_cationic_energy.lc6_inter.dist_max);
    PXD_CMP(a, b, pi_cationic_energy.pc5_inter.dist_min);
    PXD_CMP(a, b, pi_cationic_energy.pc5_inter.dist_max);
    PXD_CMP(a, b, pi_cationic_energy.pc6_inter.dist_min);
    PXD_CMP(a, b, pi_cationic_energy.pc6_inter.dist_max);
  }
#define PXD_CMP_STACK(E, K)                  \
  PXD_CMP(a, b, E.K##_coef);                 \
  PXD_CMP(a, b, E.K##_inter.dist_min);       \
  PXD_CMP(a, b, E.K##_inter.dist_max);       \
  PXD_CMP(a, b, E.K##_inter.angle_min);      \
  PXD_CMP(a, b, E.K##_inter.angle_max);      \
  PXD_CMP(a, b, E.K##_inter.sigma_min);      \
  PXD_CMP(a, b, E.K##_inter.sigma_max)
  if (active_coef(a.f2f_pi_stacking_energy_coef)) {
    PXD_CMP_STACK(f2f_pi_stacking_energy, l5r5);
    PXD_CMP_STACK(f2f_pi_stacking_energy, l5r6);
    PXD_CMP_STACK(f2f_pi_stacking_energy, l6r5);
    PXD_CMP_STACK(f2f_pi_stacking_energy, l6r6);
  }
  if (active_coef(a.e2f_pi_stacking_energy_coef)) {
    PXD_CMP_STACK(e2f_pi_stacking_energy, l5r5);
    PXD_CMP_STACK(e2f_pi_stacking_energy, l5r6);
    PXD_CMP_STACK(e2f_pi_stacking_energy, l6r5);
    PXD_CMP_STACK(e2f_pi_stacking_energy, l6r6);
  }
#define PXD_CMP_HB(K)                              \
  PXD_CMP(a, b, hbond_energy.K##_coef);            \
  PXD_CMP(a, b, hbond_energy.K##_inter.dist_min);  \
  PXD_CMP(a, b, hbond_energy.K##_inter.dist_max);  \
  PXD_CMP(a, b, hbond_energy.K##_inter.angle_min); \
  PXD_CMP(a, b, hbond_energy.K##_inter.angle_max); \
  PXD_CMP(a, b, hbond_energy.K##_inter.dist_center); \
  PXD_CMP(a, b, hbond_energy.K##_inter.angle_center)
  if (active_coef(a.hbond_energy_coef)) {
    PXD_CMP(a, b, hbond_energy.scale);
    PXD_CMP_HB(hldcc);
    PXD_CMP_HB(hldcn);
    PXD_CMP_HB(hldnc);
    PXD_CMP_HB(hldnn);
    PXD_CMP_HB(hpdcc);
    PXD_CMP_HB(hpdcn);
    PXD_CMP_HB(hpdnc);
    PXD_CMP_HB(hpdnn);
  }
  if (active_coef(a.hydrophobic_energy_coef)) {
    PXD_CMP(a, b, hydrophobic_energy.scale);
    PXD_CMP(a, b, hydrophobic_energy.dist_min);
    PXD_CMP(a, b, hydrophobic_energy.dist_max);
    PXD_CMP(a, b, hydrophobic_energy.scale_factor);
  }
  return nbad;
}

}  // namespace

extern "C" {

// The parameters the reference would read from a pose_selection JSON tree
// (json_text) against the pxd_score_params the host loader produced from the
// same text: the number of fields that differ (-1 on a JSON error); msg gets
// the first differing field or the error.
int ref_score_params_compare(const char* json_text, const pxd_score_params* sp, char* msg,
                             size_t msg_len) {
  try {
    ScoringFunctionParams a = sfp_from_json(pxdock::json::parse(json_text));
    return compare_sfp(a, to_sfp(sp), msg, msg_len);
  } catch (const std::exception& e) {
    std::snprintf(msg, msg_len, "%s", e.what());
    return -1;
  }
}

// ref_pscore with the parameters read by the reference's member names from
// a pose_selection JSON tree (see sfp_from_json).
double ref_pscore_json(const char* json_text, const pxd_mol* rec, const double* rec_xyz,
                       const pxd_mol* lig, const double* lig_xyz) {
  force_field_params rf = to_ffp(rec), lf = to_ffp(lig);
  molecule_pose rx = to_pose(rec_xyz, rec->n_atoms), lx = to_pose(lig_xyz, lig->n_atoms);
  root_scorer root("pscore", sfp_from_json(pxdock::json::parse(json_text)));
  instant_cache memo = root.precalculate(rx, rf, lx, lf);
  return root.combine(root.summary(rx, rf, lx, lf, &memo));
}

// root_scorer("bscore") on an affinity_ranking JSON tree (scorer.cpp:53-115,
// gbsa_total_energy::summary scorer.cpp:760-796) with the molecules' atomic
// numbers and bond lists (the GBSA radii, scorer.cpp:735-758); gb receives
// Gbsa_Pro_Energy, Gbsa_Lig_Energy, Gbsa_Com_Energy (0 when inactive).
double ref_bscore_json(const char* json_text, const pxd_mol* rec, const double* rec_xyz,
                       const int32_t* rec_an, const int32_t* rec_bonds, int n_rec_bonds,
                       const pxd_mol* lig, const double* lig_xyz, const int32_t* lig_an,
                       const int32_t* lig_bonds, int n_lig_bonds, double* gb) {
  force_field_params rf = to_ffp(rec), lf = to_ffp(lig);
  rf.atomic_numbers.assign(rec_an, rec_an + rec->n_atoms);
  lf.atomic_numbers.assign(lig_an, lig_an + lig->n_atoms);
  fill_ids(rf.all_bonds, rec_bonds, n_rec_bonds, 2);
  fill_ids(lf.all_bonds, lig_bonds, n_lig_bonds, 2);
  lf.rotatable_bonds.resize(lig->n_rotatable);
  molecule_pose rx = to_pose(rec_xyz, rec->n_atoms), lx = to_pose(lig_xyz, lig->n_atoms);
  root_scorer root("bscore", sfp_from_json(pxdock::json::parse(json_text)));
  instant_cache memo = root.precalculate(rx, rf, lx, lf);
  auto scores = root.summary(rx, rf, lx, lf, &memo);
  static const char* names[3] = {"Gbsa_Pro_Energy", "Gbsa_Lig_Energy", "Gbsa_Com_Energy"};
  for (int k = 0; k < 3; ++k) {
    auto it = scores.find(names[k]);
    gb[k] = it == scores.end() ? 0.0 : it->second;
  }
  return root.combine(scores);
}

// kVdwTypeDatabase (constant.h:29-130)
void ref_vdw_table(double* sigma, double* epsilon) {
  for (size_t t = 0; t < kNumVdwTypes; ++t) {
    sigma[t] = kVdwTypeDatabase[t].sigma;
    epsilon[t] = kVdwTypeDatabase[t].epsilon;
  }
}

// pair.h:25-40
double ref_coul_pair(double qi, double qj, const double* rij, double d, double scale,
                     double* gi, double* gj) {
  atom_position a{{gi[0], gi[1], gi[2]}}, b{{gj[0], gj[1], gj[2]}};
  double e = calculate_coul_pair(qi, qj, rij, d, scale, a, b);
  std::memcpy(gi, a.xyz, 24);
  std::memcpy(gj, b.xyz, 24);
  return e;
}

// pair.h:42-60
double ref_vdw_pair(double si, double ei, double sj, double ej, const double* rij, double d,
                    double scale, double* gi, double* gj) {
  atom_position a{{gi[0], gi[1], gi[2]}}, b{{gj[0], gj[1], gj[2]}};
  lj_vdw ti{si, ei}, tj{sj, ej};
  double e = calculate_vdw_pair(ti, tj, rij, d, scale, a, b);
  std::memcpy(gi, a.xyz, 24);
  std::memcpy(gj, b.xyz, 24);
  return e;
}

// math.cpp:96-110
double ref_angle_grad(const double* pi, const double* pj, const double* pk, double* gi,
                      double* gk) {
  return get_angle_and_gradient_3d(pi, pj, pk, gi, gk);
}

// math.cpp:130-167
double ref_dihedral_grad(const double* pi, const double* pj, const double* pk,
                         const double* pl, double* gi, double* gj, double* gk, double* gl) {
  return get_dihedral_angle_and_gradient_3d(pi, pj, pk, pl, gi, gj, gk, gl);
}

// math.cpp:62-94
void ref_rotmat_axis_grad(const double* axis, double theta, double* out, double* jac) {
  vector_3d a{{axis[0], axis[1], axis[2]}};
  matrix_3x3 m, j;
  get_rotation_matrix_and_gradient(a, theta, m, j);
  std::memcpy(out, m.data, 72);
  std::memcpy(jac, j.data, 72);
}

// math.cpp:22-42
void ref_rotmat_quat(const double* q, double* out) {
  vector_4d v{{q[0], q[1], q[2], q[3]}};
  matrix_3x3 m;
  get_rotation_matrix(v, m);
  std::memcpy(out, m.data, 72);
}

// root_scorer("pscore").combine and its leaf summary (scorer.h:108-117).
// terms[PXD_NUM_TERMS] in pxdock.h order; inactive terms are 0.
double ref_pscore(const pxd_score_params* sp, const pxd_mol* rec, const double* rec_xyz,
                  const pxd_mol* lig, const double* lig_xyz, double* terms) {
  force_field_params rf = to_ffp(rec), lf = to_ffp(lig);
  molecule_pose rx = to_pose(rec_xyz, rec->n_atoms), lx = to_pose(lig_xyz, lig->n_atoms);
  root_scorer root("pscore", to_sfp(sp));
  instant_cache memo = root.precalculate(rx, rf, lx, lf);
  auto scores = root.summary(rx, rf, lx, lf, &memo);
  static const char* names[PXD_NUM_TERMS] = {
      "InteractionCoulomb_oriEnergy", "InteractionVDW_Energy", "Ionic_Energy",
      "Pi_Cationic_Energy", "F2FPiStacking_Energy", "E2FPiStacking_Energy",
      "Hbond_Energy", "Hydrophobic_Energy", "TorsionStrain_Penalty"};
  if (terms) {
    for (int k = 0; k < PXD_NUM_TERMS; ++k) {
      auto it = scores.find(names[k]);
      terms[k] = it == scores.end() ? 0.0 : it->second;
    }
  }
  return root.combine(scores);
}

// random_generator (random.h:33-63, random.cpp:22-49): n draws of each kind
// for property checks of the restated samplers.
void ref_random_draws(unsigned int seed, int n, double* in_sphere, double* quats) {
  random_generator rg(seed);
  for (int i = 0; i < n; ++i) rg.random_in_sphere(in_sphere + 3 * i);
  for (int i = 0; i < n; ++i) rg.random_orientation(quats + 4 * i);
}

}  // extern "C"
