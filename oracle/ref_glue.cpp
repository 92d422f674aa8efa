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
#include <cstring>
#include <string>
#include <unordered_map>

#include "bytedock/core/constant.h"
#include "bytedock/core/pair.h"
#include "bytedock/core/scorer.h"
#include "bytedock/lib/math.h"
#include "bytedock/lib/random.h"
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

}  // namespace

extern "C" {

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
