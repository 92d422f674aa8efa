// pxdock.hpp — C++ host layer over the C ABI (include/pxdock.h), with the
// names, argument meaning and error behaviour of the reference engine's
// classes on the docking hot path (engine/include/bytedock/core/*.h), so that
// engine code written against them runs on the MI355X kernels:
//
//   torsional_receptor / free_ligand   data.h:34-99     (reference JSON)
//   receptor_cache                     grid.h:60-133    (maps built on the GPU)
//   binding_system_interactions        interaction.h:97-131
//   lbfgs_step                         optimizer.h:50-67
//   pose_selection (root_scorer)       scorer.h:96-117
//   monte_carlo_searcher               searcher.h:27-82
//
// Coordinates are FP64 on the host (param_t = double, the reference's default
// build) and FP32 on the device.  Errors throw: conf_error for malformed
// input (the reference's failed_conf_error), std::invalid_argument for bad
// arguments, std::runtime_error for HIP failures.
#pragma once

#include <hip/hip_runtime_api.h>

#include <cmath>
#include <cstring>
#include <fstream>
#include <memory>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "json.hpp"
#include "pxdock.h"

namespace pxdock {

using param_t = double;
struct atom_position {
  param_t xyz[3];
};
using molecule_pose = std::vector<atom_position>;
using box_config = pxd_box;

inline void check(pxd_status st) {
  if (st == PXD_OK) return;
  std::string m = pxd_last_error();
  if (st == PXD_ERR_CONF) throw conf_error(m);
  if (st == PXD_ERR_ARG || st == PXD_ERR_RANGE) throw std::invalid_argument(m);
  throw std::runtime_error(m);
}

inline void hip_check(hipError_t e, const char* what) {
  if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

// Owning device buffer.
template <class T>
class dev_buffer {
 public:
  dev_buffer() = default;
  explicit dev_buffer(size_t n) { resize(n); }
  ~dev_buffer() { release(); }
  dev_buffer(const dev_buffer&) = delete;
  dev_buffer& operator=(const dev_buffer&) = delete;
  void resize(size_t n) {
    release();
    n_ = n;
    if (n) hip_check(hipMalloc(&p_, n * sizeof(T)), "hipMalloc");
  }
  void upload(const std::vector<T>& v) {
    if (v.size() != n_) resize(v.size());
    if (n_) hip_check(hipMemcpy(p_, v.data(), n_ * sizeof(T), hipMemcpyHostToDevice), "upload");
  }
  std::vector<T> download() const {
    std::vector<T> v(n_);
    if (n_) hip_check(hipMemcpy(v.data(), p_, n_ * sizeof(T), hipMemcpyDeviceToHost), "download");
    return v;
  }
  T* get() const { return p_; }
  size_t size() const { return n_; }

 private:
  void release() {
    if (p_) (void)hipFree(p_);
    p_ = nullptr;
    n_ = 0;
  }
  T* p_ = nullptr;
  size_t n_ = 0;
};

// ---------------------------------------------------------------- grid geometry

// describe_boundary / decide_grid_dims (grid.cpp:42-86)
inline box_config describe_boundary(double cx, double cy, double cz, double sx, double sy,
                                    double sz) {
  box_config b;
  const double c[3] = {cx, cy, cz}, s[3] = {sx, sy, sz};
  pxd_describe_boundary(c, s, &b);
  return b;
}
inline box_config decide_grid_dims(double cx, double cy, double cz, double sx, double sy,
                                   double sz, double spacing) {
  box_config b;
  const double c[3] = {cx, cy, cz}, s[3] = {sx, sy, sz};
  pxd_decide_grid_dims(c, s, spacing, &b);
  return b;
}
// check_atom_in_box (grid.h:42-46)
inline bool check_atom_in_box(const atom_position& a, const box_config& bc) {
  for (int d = 0; d < 3; ++d)
    if (!(a.xyz[d] > bc.init_xyz[d] && a.xyz[d] < bc.oppo_xyz[d])) return false;
  return true;
}

// ---------------------------------------------------------------- molecule data

// force_field_params (system.h:85-109) owned in flat vectors, viewed as pxd_mol.
struct force_field_params {
  std::vector<int32_t> vdw_type;
  std::vector<double> sigma, epsilon, charge;
  std::vector<int32_t> bond_ij;
  std::vector<double> bond_k, bond_r0;
  std::vector<int32_t> angle_ijk;
  std::vector<double> angle_k, angle_t0;
  std::vector<int32_t> dih_ijkl, dih_n;
  std::vector<double> dih_phase, dih_k;
  std::vector<int32_t> pairs14, pairs;
  std::vector<int32_t> hydrophobic, cation, anion, ring5, ring6;
  std::vector<int32_t> don_charged, don_neutral, acc_charged, acc_neutral;
  std::vector<double> tstrain_coef;
  std::vector<int32_t> tstrain_off{0}, tstrain_ijkl;
  int32_t n_rotatable = 0;

  pxd_mol view(int32_t n_atoms) const {
    pxd_mol m;
    std::memset(&m, 0, sizeof(m));
    auto ip = [](const std::vector<int32_t>& v) { return v.empty() ? nullptr : v.data(); };
    auto dp = [](const std::vector<double>& v) { return v.empty() ? nullptr : v.data(); };
    m.n_atoms = n_atoms;
    m.vdw_type = ip(vdw_type);
    m.sigma = dp(sigma);
    m.epsilon = dp(epsilon);
    m.charge = dp(charge);
    m.n_bonds = (int32_t)bond_ij.size() / 2;
    m.bond_ij = ip(bond_ij);
    m.bond_k = dp(bond_k);
    m.bond_r0 = dp(bond_r0);
    m.n_angles = (int32_t)angle_ijk.size() / 3;
    m.angle_ijk = ip(angle_ijk);
    m.angle_k = dp(angle_k);
    m.angle_t0 = dp(angle_t0);
    m.n_dihedrals = (int32_t)dih_ijkl.size() / 4;
    m.dih_ijkl = ip(dih_ijkl);
    m.dih_n = ip(dih_n);
    m.dih_phase = dp(dih_phase);
    m.dih_k = dp(dih_k);
    m.n_pairs14 = (int32_t)pairs14.size() / 2;
    m.pairs14 = ip(pairs14);
    m.n_pairs = (int32_t)pairs.size() / 2;
    m.pairs = ip(pairs);
    m.n_hydrophobic = (int32_t)hydrophobic.size();
    m.hydrophobic = ip(hydrophobic);
    m.n_cation = (int32_t)cation.size();
    m.cation = ip(cation);
    m.n_anion = (int32_t)anion.size();
    m.anion = ip(anion);
    m.n_ring5 = (int32_t)ring5.size() / 5;
    m.ring5 = ip(ring5);
    m.n_ring6 = (int32_t)ring6.size() / 6;
    m.ring6 = ip(ring6);
    m.n_don_charged = (int32_t)don_charged.size() / 2;
    m.don_charged = ip(don_charged);
    m.n_don_neutral = (int32_t)don_neutral.size() / 2;
    m.don_neutral = ip(don_neutral);
    m.n_acc_charged = (int32_t)acc_charged.size();
    m.acc_charged = ip(acc_charged);
    m.n_acc_neutral = (int32_t)acc_neutral.size();
    m.acc_neutral = ip(acc_neutral);
    m.n_tstrain = (int32_t)tstrain_coef.size() / 4;
    m.tstrain_coef = dp(tstrain_coef);
    m.tstrain_match_off = ip(tstrain_off);
    m.tstrain_ijkl = ip(tstrain_ijkl);
    m.n_rotatable = n_rotatable;
    return m;
  }
};

namespace detail {

inline std::string read_file(const std::string& path) {
  std::ifstream f(path);
  if (!f) throw conf_error("cannot open " + path);
  std::stringstream ss;
  ss << f.rdbuf();
  return ss.str();
}

inline void ids(const json& node, int width, std::vector<int32_t>& out) {
  out.clear();
  for (const json& row : node.items()) {
    if (row.size() != (size_t)width) throw conf_error("atom index row of wrong width");
    for (const json& v : row.items()) out.push_back((int32_t)v.integer());
  }
}
inline void flat_ints(const json& node, std::vector<int32_t>& out) {
  out.clear();
  for (const json& v : node.items()) out.push_back((int32_t)v.integer());
}
inline void flat_nums(const json& node, std::vector<double>& out) {
  out.clear();
  for (const json& v : node.items()) out.push_back(v.num());
}
// MATH_TO_RADIAN (math.h:229-230)
inline double to_radian(double deg) { return deg / 180.0 * 3.141592653589793; }

// fill(node, prefix, fourier_dihedral) (data.cpp:83-128): series padded to 6
inline void dihedrals(const json& ff, const std::string& prefix, force_field_params& p) {
  std::vector<int32_t> id;
  ids(ff[prefix + "_atomidx"], 4, id);
  const json& per = ff[prefix + "_periodicity"];
  const json& k = ff[prefix + "_k"];
  const json& ph = ff[prefix + "_phase"];
  const size_t n = id.size() / 4;
  if (per.size() != n || k.size() != n || ph.size() != n)
    throw conf_error(prefix + ": series count mismatches atom quadruples");
  for (size_t r = 0; r < n; ++r) {
    for (int q = 0; q < 4; ++q) p.dih_ijkl.push_back(id[4 * r + q]);
    if (per[r].size() > PXD_DIH_TERMS) throw conf_error(prefix + ": more than 6 terms");
    for (int t = 0; t < PXD_DIH_TERMS; ++t) {
      bool has = (size_t)t < per[r].size();
      p.dih_n.push_back(has ? (int32_t)per[r][t].integer() : 0);
      p.dih_k.push_back(has ? k[r][t].num() : 0.0);
      p.dih_phase.push_back(has ? to_radian(ph[r][t].num()) : 0.0);
    }
  }
}

// fill(node, ligand, force_field_params) (data.cpp:246-276)
inline force_field_params fill_ff(const json& ff, bool ligand) {
  force_field_params p;
  ids(ff["FF_Bonds_atomidx"], 2, p.bond_ij);
  flat_nums(ff["FF_Bonds_k"], p.bond_k);
  flat_nums(ff["FF_Bonds_length"], p.bond_r0);
  ids(ff["FF_Angles_atomidx"], 3, p.angle_ijk);
  flat_nums(ff["FF_Angles_k"], p.angle_k);
  flat_nums(ff["FF_Angles_angle"], p.angle_t0);
  for (double& a : p.angle_t0) a = to_radian(a);
  dihedrals(ff, "FF_ImproperTorsions", p);
  dihedrals(ff, "FF_ProperTorsions", p);
  if (ligand) {
    // nolmaliz_vdw_types (data.cpp:234-244); id 99 is out of the table
    for (const json& v : ff["FF_vdW_paraidx"].items()) {
      long long t = v.integer() - PXD_VDW_ID_OFFSET;
      if (t < 0 || t >= PXD_NUM_VDW_TYPES)
        throw conf_error("VDW type [" + std::to_string(t) + "] is illegal!");
      p.vdw_type.push_back((int32_t)t);
    }
  } else {
    flat_nums(ff["FF_vdW_sigma"], p.sigma);
    flat_nums(ff["FF_vdW_epsilon"], p.epsilon);
  }
  flat_nums(ff["partial_charges"], p.charge);
  ids(ff["FF_Nonbonded14_atomidx"], 2, p.pairs14);
  ids(ff["FF_NonbondedAll_atomidx"], 2, p.pairs);
  flat_ints(ff["hydrophobic_atomidx"], p.hydrophobic);
  flat_ints(ff["cation_atomidx"], p.cation);
  flat_ints(ff["anion_atomidx"], p.anion);
  ids(ff["piring5_atomidx"], 5, p.ring5);
  ids(ff["piring6_atomidx"], 6, p.ring6);
  ids(ff["hbonddon_charged_atomidx"], 2, p.don_charged);
  ids(ff["hbonddon_neut_atomidx"], 2, p.don_neutral);
  flat_ints(ff["hbondacc_charged_atomidx"], p.acc_charged);
  flat_ints(ff["hbondacc_neut_atomidx"], p.acc_neutral);
  if (ligand && ff.has("tstrain_params_pose_selection") &&
      ff["tstrain_params_pose_selection"].size() > 0) {  // data.cpp:204-223
    for (const json& row : ff["tstrain_params_pose_selection"].items())
      for (const json& v : row.items()) p.tstrain_coef.push_back(v.num());
    for (const json& m : ff["tstrain_atomidx_pose_selection"].items()) {
      for (const json& q : m.items())
        for (const json& v : q.items()) p.tstrain_ijkl.push_back((int32_t)v.integer());
      p.tstrain_off.push_back((int32_t)p.tstrain_ijkl.size() / 4);
    }
  }
  if (ligand && ff.has("rotatable_bond_index"))
    p.n_rotatable = (int32_t)ff["rotatable_bond_index"].size();
  return p;
}

inline molecule_pose pose_of(const json& rows) {
  molecule_pose m(rows.size());
  for (size_t i = 0; i < rows.size(); ++i)
    for (int d = 0; d < 3; ++d) m[i].xyz[d] = rows[i][d].num();
  return m;
}

}  // namespace detail

// ---------------------------------------------------------------- receptor

class torsional_receptor {
 public:
  // torsional_receptor::parse (data.cpp:331-364)
  static std::shared_ptr<torsional_receptor> parse(const json& root) {
    auto r = std::make_shared<torsional_receptor>();
    r->ff_ = detail::fill_ff(root["ffdata"], false);
    r->xyz_ = detail::pose_of(root["xyz"]);
    const json& geo = root["geometry"];
    if ((size_t)geo["num_atoms"].integer() != r->xyz_.size())
      throw conf_error("Atom count mismatches! Wanted [" +
                       std::to_string(geo["num_atoms"].integer()) + "] vs. Parsed [" +
                       std::to_string(r->xyz_.size()) + "]");
    // rotatable leaves (data.cpp:292-363): origin = rot_bond_1, axis towards rot_bond_2
    r->atom_off_.push_back(0);
    const json& data = geo["data"];
    for (const std::string& key : data.keys()) {
      const json& batch = data[key];
      for (size_t i = 0; i < batch["rot_bond_1"].size(); ++i) {
        int a0 = (int)batch["rot_bond_1"][i].integer(), a1 = (int)batch["rot_bond_2"][i].integer();
        const double* o = r->xyz_.at(a0).xyz;
        const double* p = r->xyz_.at(a1).xyz;
        double v[3] = {p[0] - o[0], p[1] - o[1], p[2] - o[2]};
        double nv = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
        for (int d = 0; d < 3; ++d) {
          r->origin_.push_back(o[d]);
          r->axis_.push_back(v[d] / nv);
        }
        for (const json& a : batch["fragments"][i].items())
          if ((int)a.integer() != a1) r->atoms_.push_back((int32_t)a.integer());
        r->atom_off_.push_back((int32_t)r->atoms_.size());
      }
    }
    return r;
  }
  static std::shared_ptr<torsional_receptor> load(const std::string& path) {
    return parse(json::parse(detail::read_file(path)));
  }
  ~torsional_receptor() {
    if (h_) pxd_receptor_destroy(h_);
  }

  size_t num_atoms() const { return xyz_.size(); }
  size_t num_torsions() const { return atom_off_.size() - 1; }
  const molecule_pose& get_positions() const { return xyz_; }
  const force_field_params& get_ffdata() const { return ff_; }

  // torsional_receptor::apply_parameters (data.cpp:366-385): each leaf rotates
  // its movable atoms about its axis from the initial coordinates
  molecule_pose apply_parameters(const std::vector<param_t>& torsions) const {
    if (torsions.size() != num_torsions()) throw std::invalid_argument("torsion count mismatch");
    molecule_pose out = xyz_;
    for (size_t k = 0; k < num_torsions(); ++k) {
      const double* o = &origin_[3 * k];
      const double* a = &axis_[3 * k];
      double s = std::sin(torsions[k]), c = std::cos(torsions[k]), t = 1.0 - c;
      double R[9] = {c + a[0] * a[0] * t,        a[0] * a[1] * t - a[2] * s, a[1] * s + a[0] * a[2] * t,
                     a[2] * s + a[0] * a[1] * t, c + a[1] * a[1] * t,        a[1] * a[2] * t - a[0] * s,
                     a[0] * a[2] * t - a[1] * s, a[0] * s + a[1] * a[2] * t, c + a[2] * a[2] * t};
      for (int32_t q = atom_off_[k]; q < atom_off_[k + 1]; ++q) {
        const double* x = xyz_[atoms_[q]].xyz;
        double v[3] = {x[0] - o[0], x[1] - o[1], x[2] - o[2]};
        for (int d = 0; d < 3; ++d)
          out[atoms_[q]].xyz[d] = R[3 * d] * v[0] + R[3 * d + 1] * v[1] + R[3 * d + 2] * v[2] + o[d];
      }
    }
    return out;
  }

  // the device copy (created on first use)
  pxd_receptor* device() const {
    if (!h_) {
      pxd_mol m = ff_.view((int32_t)xyz_.size());
      pxd_leaves lv{(int32_t)num_torsions(), origin_.data(), axis_.data(), atom_off_.data(),
                    atoms_.data()};
      std::vector<double> flat(3 * xyz_.size());
      for (size_t i = 0; i < xyz_.size(); ++i)
        for (int d = 0; d < 3; ++d) flat[3 * i + d] = xyz_[i].xyz[d];
      check(pxd_receptor_create(&m, flat.data(), &lv, &h_));
    }
    return h_;
  }

 private:
  force_field_params ff_;
  molecule_pose xyz_;
  std::vector<double> origin_, axis_;
  std::vector<int32_t> atom_off_, atoms_;
  mutable pxd_receptor* h_ = nullptr;
};

// ---------------------------------------------------------------- ligand

class free_ligand {
 public:
  // free_ligand::parse (data.cpp:496-520) with the fragment tree of
  // data.cpp:396-482 in flat pre-order form
  static std::shared_ptr<free_ligand> parse(const json& root) {
    auto l = std::make_shared<free_ligand>();
    l->ff_ = detail::fill_ff(root["ffdata"], true);
    const json& xyz = root["xyz"];
    if (xyz.size() && xyz[0].size() && xyz[0][0].is_array()) {
      for (const json& p : xyz.items()) l->poses_.push_back(detail::pose_of(p));
    } else {
      l->poses_.push_back(detail::pose_of(xyz));
    }
    const json& geo = root["geometry"];
    const int n = (int)geo["num_atoms"].integer();
    if ((size_t)n != l->poses_[0].size()) throw conf_error("num_atoms mismatches xyz");
    std::vector<int32_t> split;
    detail::flat_ints(geo["frag_split_index"], split);
    const int nf = (int)split.size() + 1;
    detail::flat_ints(geo["permute_index"], l->perm_);
    if ((int)l->perm_.size() != n) throw conf_error("permute_index length mismatches num_atoms");
    std::vector<int32_t> rev(n);
    for (int p = 0; p < n; ++p) rev.at(l->perm_[p]) = p;
    l->start_.resize(nf);
    l->end_.resize(nf);
    for (int f = 0; f < nf; ++f) {
      l->start_[f] = f == 0 ? 0 : split[f - 1];
      l->end_[f] = f == nf - 1 ? n : split[f];
    }
    l->parent_.assign(nf, -2);
    std::vector<std::vector<int>> children(nf);
    std::vector<int> roots;
    for (const json& level : geo["frag_traverse_levels"].items())
      for (const std::string& key : level.keys()) {
        const int f = std::stoi(key);
        const json& node = level[key];
        if (node["parent"].size() == 0) {
          roots.push_back(f);
          l->parent_.at(f) = -1;
        }
        for (const json& c : node["children"].items()) {
          children.at(f).push_back((int)c.integer());
          l->parent_.at((int)c.integer()) = f;
        }
      }
    for (int f = 0; f < nf; ++f)
      if (l->parent_[f] == -2) throw conf_error("fragment tree does not cover every fragment");
    l->bond_.assign(2 * nf, 0);
    for (int f = 0; f < nf; ++f) {
      if (l->parent_[f] < 0) continue;
      const json& e = geo["edge_to_bond"][std::to_string(l->parent_[f]) + "->" + std::to_string(f)];
      l->bond_[2 * f] = rev.at((int)e[0].integer());
      l->bond_[2 * f + 1] = rev.at((int)e[1].integer());
    }
    // torsion order: pre-order over the children (data.h:78-81)
    std::vector<int> stack;
    for (int r : roots) {
      std::vector<std::pair<int, size_t>> st{{r, 0}};
      while (!st.empty()) {
        auto& top = st.back();
        if (top.second < children[top.first].size()) {
          int c = children[top.first][top.second++];
          l->tors_frag_.push_back(c);
          st.push_back({c, 0});
        } else {
          st.pop_back();
        }
      }
    }
    return l;
  }
  static std::shared_ptr<free_ligand> load(const std::string& path) {
    return parse(json::parse(detail::read_file(path)));
  }

  size_t num_atoms() const { return poses_[0].size(); }
  size_t num_torsions() const { return tors_frag_.size(); }
  size_t get_nposes() const { return poses_.size(); }
  const molecule_pose& get_pose(size_t i) const { return poses_.at(i); }
  const force_field_params& get_ffdata() const { return ff_; }
  pxd_mol mol_view() const { return ff_.view((int32_t)num_atoms()); }
  pxd_frag_tree tree_view() const {
    return pxd_frag_tree{(int32_t)start_.size(), perm_.data(),   start_.data(),
                         end_.data(),            parent_.data(), bond_.data(),
                         (int32_t)tors_frag_.size(), tors_frag_.data()};
  }

  // free_ligand::apply_parameters (data.cpp:592-641), on the device
  molecule_pose apply_parameters(const molecule_pose& xyz, const atom_position& center,
                                 const double orientation[9],
                                 const std::vector<param_t>& torsions) const;

 private:
  force_field_params ff_;
  std::vector<molecule_pose> poses_;
  std::vector<int32_t> perm_, start_, end_, parent_, bond_, tors_frag_;
};

// A device batch of ligands (pxd_ligand_batch) with host atom counts.
class ligand_batch {
 public:
  explicit ligand_batch(const std::vector<std::shared_ptr<free_ligand>>& ligs) : ligs_(ligs) {
    std::vector<pxd_mol> mols;
    std::vector<pxd_frag_tree> trees;
    for (auto& l : ligs) {
      mols.push_back(l->mol_view());
      trees.push_back(l->tree_view());
    }
    check(pxd_ligand_batch_create(mols.data(), trees.data(), (int32_t)mols.size(), &h_));
  }
  ~ligand_batch() {
    if (h_) pxd_ligand_batch_destroy(h_);
  }
  ligand_batch(const ligand_batch&) = delete;
  ligand_batch& operator=(const ligand_batch&) = delete;
  pxd_ligand_batch* get() const { return h_; }
  const free_ligand& ligand(size_t i) const { return *ligs_.at(i); }
  size_t size() const { return ligs_.size(); }

 private:
  std::vector<std::shared_ptr<free_ligand>> ligs_;
  pxd_ligand_batch* h_ = nullptr;
};

inline std::vector<float> to_f32(const molecule_pose& p) {
  std::vector<float> v(3 * p.size());
  for (size_t i = 0; i < p.size(); ++i)
    for (int d = 0; d < 3; ++d) v[3 * i + d] = (float)p[i].xyz[d];
  return v;
}
inline molecule_pose from_f32(const float* v, size_t n) {
  molecule_pose p(n);
  for (size_t i = 0; i < n; ++i)
    for (int d = 0; d < 3; ++d) p[i].xyz[d] = v[3 * i + d];
  return p;
}

inline molecule_pose free_ligand::apply_parameters(const molecule_pose& xyz,
                                                   const atom_position& center,
                                                   const double orientation[9],
                                                   const std::vector<param_t>& torsions) const {
  if (xyz.size() != num_atoms() || torsions.size() != num_torsions())
    throw std::invalid_argument("apply_parameters: size mismatch");
  // a one-ligand batch of this ligand (copying the descriptors it points to);
  // placing atoms needs only the fragment tree, so a ligand read without force
  // field data gets neutral placeholders
  auto self = std::make_shared<free_ligand>(*this);
  if (self->ff_.charge.empty()) self->ff_.charge.assign(num_atoms(), 0.0);
  if (self->ff_.vdw_type.empty()) self->ff_.vdw_type.assign(num_atoms(), 0);
  ligand_batch b({self});
  dev_buffer<int32_t> lig(1), off(1), toff(2);
  lig.upload({0});
  off.upload({0});
  toff.upload({0, (int32_t)torsions.size()});
  dev_buffer<float> xin, c, r, t, xo(3 * num_atoms());
  xin.upload(to_f32(xyz));
  c.upload({(float)center.xyz[0], (float)center.xyz[1], (float)center.xyz[2]});
  std::vector<float> rv(9), tv(torsions.size() + 1, 0.f);
  for (int k = 0; k < 9; ++k) rv[k] = (float)orientation[k];
  for (size_t k = 0; k < torsions.size(); ++k) tv[k] = (float)torsions[k];
  r.upload(rv);
  t.upload(tv);
  check(pxd_apply_poses(b.get(), 1, lig.get(), off.get(), xin.get(), c.get(), r.get(), toff.get(),
                        t.get(), xo.get(), nullptr));
  hip_check(hipDeviceSynchronize(), "apply_parameters");
  auto v = xo.download();
  return from_f32(v.data(), num_atoms());
}

// ---------------------------------------------------------------- maps

// receptor_cache (grid.h:60-133): populate() builds the maps of the given
// ligand types on the GPU.
class receptor_cache {
 public:
  receptor_cache(std::shared_ptr<torsional_receptor> receptor, const box_config& bc,
                 std::vector<int32_t> types, param_t slope = 1e6)
      : receptor_(std::move(receptor)), bc_(bc), types_(std::move(types)), slope_(slope) {}
  ~receptor_cache() {
    if (g_) pxd_grid_destroy(g_);
  }
  void populate() {
    if (g_) pxd_grid_destroy(g_);
    g_ = nullptr;
    check(pxd_grid_build(receptor_->device(), &bc_, types_.data(), (int32_t)types_.size(), slope_,
                         nullptr, &g_));
  }
  void set_slope(param_t s) {  // grid.h:79
    slope_ = s;
    if (g_) check(pxd_grid_set_slope(g_, s));
  }
  const box_config& get_box() const { return bc_; }
  pxd_grid* device() const { return g_; }

 private:
  std::shared_ptr<torsional_receptor> receptor_;
  box_config bc_;
  std::vector<int32_t> types_;
  param_t slope_;
  pxd_grid* g_ = nullptr;
};

// ---------------------------------------------------------------- the model

// optimized_result (task.h:32-41)
struct optimized_result {
  std::vector<param_t> torsions;
  molecule_pose receptor_xyz, ligand_xyz;
  param_t energy = 0.0, pscore = 0.0;
  size_t nevals = 0;
  size_t offset = 0;
  std::string name;
};

// binding_system_interactions (interaction.h:97-131) for one receptor and one
// ligand, with or without maps.  put_gradients ADDS into the gradients as the
// reference does (callers reset them, interaction.cpp:300-314).
class binding_system_interactions {
 public:
  binding_system_interactions(std::shared_ptr<torsional_receptor> receptor,
                              std::shared_ptr<free_ligand> ligand,
                              std::shared_ptr<receptor_cache> cache = nullptr)
      : receptor_(std::move(receptor)), ligand_(std::move(ligand)), cache_(std::move(cache)),
        batch_({ligand_}) {}

  void init_gradients(std::vector<param_t>& torsion_gradient,
                      molecule_pose& ligand_gradient) const {
    torsion_gradient.assign(receptor_->num_torsions(), 0.0);
    ligand_gradient.assign(ligand_->num_atoms(), atom_position{{0.0, 0.0, 0.0}});
  }
  void reset_gradients(std::vector<param_t>& torsion_gradient,
                       molecule_pose& ligand_gradient) const {
    for (auto& v : torsion_gradient) v = 0.0;
    for (auto& a : ligand_gradient) a.xyz[0] = a.xyz[1] = a.xyz[2] = 0.0;
  }

  param_t put_gradients(const std::vector<param_t>& receptor_torsions,
                        const molecule_pose& ligand_xyz, std::vector<param_t>& torsion_gradient,
                        molecule_pose& ligand_gradient) {
    const size_t n = ligand_->num_atoms(), L = receptor_->num_torsions();
    if (ligand_xyz.size() != n || receptor_torsions.size() != L || ligand_gradient.size() != n ||
        torsion_gradient.size() != L)
      throw std::invalid_argument("put_gradients: size mismatch");
    dev_buffer<int32_t> lig(1), off(1);
    lig.upload({0});
    off.upload({0});
    dev_buffer<float> x, t, g(3 * n), tg(L + 1);
    x.upload(to_f32(ligand_xyz));
    std::vector<float> tv(L + 1, 0.f);
    for (size_t k = 0; k < L; ++k) tv[k] = (float)receptor_torsions[k];
    t.upload(tv);
    dev_buffer<double> e(1);
    check(pxd_score_grad_batch(receptor_->device(), cache_ ? cache_->device() : nullptr,
                               batch_.get(), 1, lig.get(), off.get(), x.get(), t.get(), e.get(),
                               g.get(), tg.get(), nullptr));
    hip_check(hipDeviceSynchronize(), "put_gradients");
    auto gv = g.download();
    auto tgv = tg.download();
    for (size_t i = 0; i < n; ++i)
      for (int d = 0; d < 3; ++d) ligand_gradient[i].xyz[d] += gv[3 * i + d];
    for (size_t k = 0; k < L; ++k) torsion_gradient[k] += tgv[k];
    return e.download()[0];
  }

  const torsional_receptor& receptor() const { return *receptor_; }
  const free_ligand& ligand() const { return *ligand_; }
  receptor_cache* cache() const { return cache_.get(); }
  pxd_ligand_batch* batch() const { return batch_.get(); }

 private:
  std::shared_ptr<torsional_receptor> receptor_;
  std::shared_ptr<free_ligand> ligand_;
  std::shared_ptr<receptor_cache> cache_;
  ligand_batch batch_;
};

// ---------------------------------------------------------------- optimisation

// optimization_step (optimizer.h:24-29)
class optimization_step {
 public:
  virtual ~optimization_step() = default;
  // Update out.{ligand_xyz, torsions, energy}
  virtual bool apply(binding_system_interactions& model, optimized_result& out) = 0;
};

// lbfgs_step (optimizer.h:50-67): the device L-BFGS with the reference's
// constants (history 100, |g|inf <= 0.5, tolerance_change 1e-4).
class lbfgs_step : public optimization_step {
 public:
  explicit lbfgs_step(size_t max_niters, size_t max_nevals = 0) {
    pxd_lbfgs_default_params(&p_, (int32_t)max_niters);
    p_.max_eval = (int32_t)(max_nevals > 0 ? max_nevals : max_niters * 5 / 4);
  }
  bool apply(binding_system_interactions& model, optimized_result& out) override {
    const size_t n = model.ligand().num_atoms(), L = model.receptor().num_torsions();
    if (out.ligand_xyz.size() != n) throw std::invalid_argument("lbfgs_step: ligand size");
    if (out.torsions.size() != L) out.torsions.assign(L, 0.0);
    dev_buffer<int32_t> lig(1), off(1), ne(1), cv(1);
    lig.upload({0});
    off.upload({0});
    dev_buffer<float> x, t;
    x.upload(to_f32(out.ligand_xyz));
    std::vector<float> tv(L + 1, 0.f);
    for (size_t k = 0; k < L; ++k) tv[k] = (float)out.torsions[k];
    t.upload(tv);
    dev_buffer<double> e(1);
    check(pxd_lbfgs_batch(model.receptor().device(),
                          model.cache() ? model.cache()->device() : nullptr, model.batch(), &p_, 1,
                          lig.get(), off.get(), x.get(), t.get(), e.get(), ne.get(), cv.get(),
                          nullptr));
    hip_check(hipDeviceSynchronize(), "lbfgs_step");
    auto xv = x.download();
    auto tvv = t.download();
    out.ligand_xyz = from_f32(xv.data(), n);
    for (size_t k = 0; k < L; ++k) out.torsions[k] = tvv[k];
    out.energy = e.download()[0];
    out.nevals = (size_t)ne.download()[0];
    return cv.download()[0] != 0;
  }
  pxd_lbfgs_params& params() { return p_; }

 private:
  pxd_lbfgs_params p_;
};

// ---------------------------------------------------------------- pose score

// root_scorer::combine on the pose_selection tree (scorer.h:96-117) with the
// configuration read as scoring_function_factory does (factory.cpp:144-238):
// a term's sub-object is read only when |coef| > 1e-5 (utility.h:26).
class pose_selection {
 public:
  explicit pose_selection(const pxd_score_params& p) : p_(p) {}
  static pose_selection parse(const json& root) {
    const json& n = root.has("pose_selection") ? root["pose_selection"] : root;
    pxd_score_params p;
    std::memset(&p, 0, sizeof(p));
    for (int t = 0; t < PXD_NUM_VDW_TYPES; ++t) p.vdw_sigma_scale[t] = p.vdw_epsilon_scale[t] = 1.0;
    auto g = [&](const char* k) { return n.has(k) ? n[k].num() : 0.0; };
    auto on = [](double c) { return c > 1e-5 || c < -1e-5; };
    p.coul_coef = g("interaction_coulomb_energy_coef");
    p.vdw_coef = g("interaction_vdw_energy_coef");
    p.ionic_coef = g("ionic_energy_coef");
    p.pication_coef = g("pi_cationic_energy_coef");
    p.f2f_coef = g("f2f_pi_stacking_energy_coef");
    p.e2f_coef = g("e2f_pi_stacking_energy_coef");
    p.hbond_coef = g("hbond_energy_coef");
    p.hydrophobic_coef = g("hydrophobic_energy_coef");
    p.tstrain_coef = g("torsion_strain_penalty_coef");
    if (on(p.vdw_coef)) {
      const json& v = n["interaction_vdw_energy"];
      p.vdw_scale = v["scale"].num();
      p.vdw_cutoff = v["cutoff"].num();
      if (v["types"].size() < PXD_NUM_VDW_TYPES)
        throw conf_error("interaction_vdw_energy.types needs 99 entries");
      for (int t = 0; t < PXD_NUM_VDW_TYPES; ++t) {
        p.vdw_sigma_scale[t] = v["types"][t]["sigma_scale"].num();
        p.vdw_epsilon_scale[t] = v["types"][t]["epsilon_scale"].num();
      }
    }
    if (on(p.ionic_coef)) {
      const json& v = n["ionic_energy"];
      p.ionic_lc_coef = v["lc_coef"].num();
      p.ionic_pc_coef = v["pc_coef"].num();
      p.ionic_lc[0] = v["lc_inter"]["dist_min"].num();
      p.ionic_lc[1] = v["lc_inter"]["dist_max"].num();
      p.ionic_pc[0] = v["pc_inter"]["dist_min"].num();
      p.ionic_pc[1] = v["pc_inter"]["dist_max"].num();
    }
    if (on(p.pication_coef)) {
      const json& v = n["pi_cationic_energy"];
      const char* keys[4] = {"lc5", "lc6", "pc5", "pc6"};
      for (int i = 0; i < 4; ++i) {
        p.pication_sub_coef[i] = v[std::string(keys[i]) + "_coef"].num();
        p.pication_range[i][0] = v[std::string(keys[i]) + "_inter"]["dist_min"].num();
        p.pication_range[i][1] = v[std::string(keys[i]) + "_inter"]["dist_max"].num();
      }
    }
    const char* sk[4] = {"l5r5", "l5r6", "l6r5", "l6r6"};
    const char* sf[6] = {"dist_min", "dist_max", "angle_min", "angle_max", "sigma_min", "sigma_max"};
    for (int kind = 0; kind < 2; ++kind) {
      double c = kind ? p.e2f_coef : p.f2f_coef;
      if (!on(c)) continue;
      const json& v = n[kind ? "e2f_pi_stacking_energy" : "f2f_pi_stacking_energy"];
      for (int i = 0; i < 4; ++i) {
        (kind ? p.e2f_sub_coef : p.f2f_sub_coef)[i] = v[std::string(sk[i]) + "_coef"].num();
        const json& it = v[std::string(sk[i]) + "_inter"];
        for (int j = 0; j < 6; ++j) (kind ? p.e2f_inter : p.f2f_inter)[i][j] = it[sf[j]].num();
      }
    }
    if (on(p.hbond_coef)) {
      const json& v = n["hbond_energy"];
      p.hbond_scale = v["scale"].num();
      const char* hk[8] = {"hldcc", "hldcn", "hldnc", "hldnn", "hpdcc", "hpdcn", "hpdnc", "hpdnn"};
      const char* hf[6] = {"dist_min", "dist_max", "angle_min", "angle_max", "dist_center",
                           "angle_center"};
      for (int i = 0; i < 8; ++i) {
        p.hbond_sub_coef[i] = v[std::string(hk[i]) + "_coef"].num();
        const json& it = v[std::string(hk[i]) + "_inter"];
        for (int j = 0; j < 6; ++j) p.hbond_inter[i][j] = it[hf[j]].num();
      }
    }
    if (on(p.hydrophobic_coef)) {
      const json& v = n["hydrophobic_energy"];
      p.hydro_scale = v["scale"].num();
      p.hydro_dist_min = v["dist_min"].num();
      p.hydro_dist_max = v["dist_max"].num();
      p.hydro_scale_factor = v["scale_factor"].num();
    }
    return pose_selection(p);
  }
  static pose_selection load(const std::string& path) {
    return parse(json::parse(detail::read_file(path)));
  }
  const pxd_score_params& params() const { return p_; }

  // combine(receptor_xyz(torsions), ligand_xyz): the weighted pose score
  param_t combine(const torsional_receptor& receptor, const std::vector<param_t>& torsions,
                  const binding_system_interactions& model, const molecule_pose& ligand_xyz,
                  std::vector<param_t>* terms = nullptr) const {
    const size_t L = receptor.num_torsions();
    dev_buffer<int32_t> lig(1), off(1);
    lig.upload({0});
    off.upload({0});
    dev_buffer<float> x, t;
    x.upload(to_f32(ligand_xyz));
    std::vector<float> tv(L + 1, 0.f);
    for (size_t k = 0; k < L && k < torsions.size(); ++k) tv[k] = (float)torsions[k];
    t.upload(tv);
    dev_buffer<double> s(1), tr(PXD_NUM_TERMS);
    check(pxd_pscore_batch(receptor.device(), model.batch(), &p_, 1, lig.get(), off.get(), x.get(),
                           t.get(), s.get(), tr.get(), nullptr));
    hip_check(hipDeviceSynchronize(), "combine");
    if (terms) *terms = tr.download();
    return s.download()[0];
  }

 private:
  pxd_score_params p_;
};

// ---------------------------------------------------------------- search

// monte_carlo_searcher (searcher.h:27-82) over a batch of ligands: walker k of
// ligand l is mutate_and_optimize with pose id k; the results of every walker
// and the pose_cluster top-k (writer.cpp:124-177) per ligand.
class monte_carlo_searcher {
 public:
  monte_carlo_searcher(const pose_selection& sf, const box_config& bc, size_t max_nsteps,
                       size_t relax_nsteps, unsigned seed)
      : sf_(sf), bc_(bc) {
    pxd_mc_default_params(&p_);
    p_.seed = seed;
    p_.max_nsteps = (int32_t)max_nsteps;
    pxd_lbfgs_default_params(&p_.relax, (int32_t)relax_nsteps);
  }
  pxd_mc_params& params() { return p_; }

  struct ligand_result {
    std::vector<optimized_result> walkers;  // pose id order
    std::vector<size_t> modes;              // clustered walkers, best pscore first
  };

  std::vector<ligand_result> search(std::shared_ptr<torsional_receptor> receptor,
                                    std::shared_ptr<receptor_cache> cache,
                                    const std::vector<std::shared_ptr<free_ligand>>& ligands,
                                    size_t exhaustiveness, size_t num_modes = 8,
                                    param_t min_rmsd = 0.5) {
    ligand_batch batch(ligands);
    const size_t nl = ligands.size(), W = nl * exhaustiveness, L = receptor->num_torsions();
    std::vector<float> x0;
    size_t walker_atoms = 0;
    for (auto& l : ligands) {
      auto v = to_f32(l->get_pose(0));
      x0.insert(x0.end(), v.begin(), v.end());
      walker_atoms += l->num_atoms() * exhaustiveness;
    }
    dev_buffer<float> dx0, bx(3 * walker_atoms), bt(W * (L + 1));
    dx0.upload(x0);
    dev_buffer<double> be(W), bp(W);
    dev_buffer<int32_t> ne(W), ci(nl * (num_modes ? num_modes : 1)), cc(nl);
    check(pxd_mc_search(receptor->device(), cache ? cache->device() : nullptr, batch.get(),
                        &sf_.params(), &p_, &bc_, (int32_t)exhaustiveness, dx0.get(), bx.get(),
                        bt.get(), be.get(), bp.get(), ne.get(), (int32_t)num_modes, min_rmsd,
                        ci.get(), cc.get(), nullptr));
    hip_check(hipDeviceSynchronize(), "search");
    auto hx = bx.download();
    auto ht = bt.download();
    auto he = be.download();
    auto hp = bp.download();
    auto hn = ne.download();
    auto hci = ci.download();
    auto hcc = cc.download();
    std::vector<ligand_result> out(nl);
    size_t atom = 0, w = 0;
    for (size_t l = 0; l < nl; ++l) {
      for (size_t k = 0; k < exhaustiveness; ++k, ++w) {
        optimized_result r;
        const size_t n = ligands[l]->num_atoms();
        r.ligand_xyz = from_f32(hx.data() + 3 * atom, n);
        atom += n;
        r.torsions.assign(ht.begin() + w * L, ht.begin() + w * L + L);
        r.receptor_xyz = receptor->apply_parameters(r.torsions);
        r.energy = he[w];
        r.pscore = hp[w];
        r.nevals = (size_t)hn[w];
        r.offset = k;
        out[l].walkers.push_back(std::move(r));
      }
      for (int m = 0; m < hcc[l]; ++m) out[l].modes.push_back((size_t)hci[l * num_modes + m]);
    }
    return out;
  }

 private:
  pose_selection sf_;
  box_config bc_;
  pxd_mc_params p_;
};

}  // namespace pxdock
