// ref_adaptors_main.cpp — check program of the reference-side adaptors
// (protenix-dock_amd/cpp/pxd_reference_adaptors.hpp).
//
// TEST INFRASTRUCTURE ONLY.  Built by oracle/Makefile against the reference's
// standard-library headers (bytedock/core/system.h, data.h) into oracle/_ref/.
// It reads a receptor or ligand JSON into the reference's OWN structs
// (force_field_params, fragment_tree, rotatable_leaf), following the reader
// data.cpp:34-482 (which itself needs Boost.PropertyTree and is not built),
// converts them with the adaptors and prints the resulting pxd descriptors
// as JSON, for tests/test_reference_adaptors.py to compare with the Python
// reader's.
//
//   ref_adaptors receptor <file.json>
//   ref_adaptors ligand <file.json>
#include <cmath>
#include <cstdio>
#include <fstream>
#include <sstream>
#include <string>
#include <unordered_map>

#include "json.hpp"
#include "pxd_reference_adaptors.hpp"

using bytedock::index_t;
using pxdock::json;

namespace {

void rd_ids(const json& node, index_t* ids, size_t width) {
  for (size_t k = 0; k < width && k < node.size(); ++k) ids[k] = (index_t)node[k].integer();
}

void rd_list(const json& node, std::vector<index_t>& out) {
  out.clear();
  for (const auto& v : node.items()) out.push_back((index_t)v.integer());
}

void rd_pairs(const json& node, std::vector<bytedock::atom_pair>& out) {
  out.resize(node.size());
  for (size_t i = 0; i < node.size(); ++i) rd_ids(node[i], out[i].ids, 2);
}

void rd_rings(const json& node, std::vector<bytedock::pi_ring>& out) {
  out.resize(node.size());
  for (size_t i = 0; i < node.size(); ++i) {
    rd_ids(node[i], out[i].ids, 6);
    out[i].size = node[i].size();
  }
}

// data.cpp:83-128
void rd_dihedrals(const json& ff, const std::string& p, std::vector<bytedock::fourier_dihedral>& out) {
  const json& ids = ff[p + "_atomidx"];
  out.assign(ids.size(), bytedock::fourier_dihedral{});
  for (size_t i = 0; i < ids.size(); ++i) {
    auto& q = out[i];
    rd_ids(ids[i], q.ids, 4);
    const json& per = ff[p + "_periodicity"][i];
    for (size_t s = 0; s < (size_t)FF_DIHEDRAL_NDIMS; ++s)
      q.periodicity[s] = s < per.size() ? (index_t)per[s].integer() : 0;
    const json& k = ff[p + "_k"][i];
    for (size_t s = 0; s < k.size(); ++s) q.k[s] = k[s].num();
    const json& ph = ff[p + "_phase"][i];
    for (size_t s = 0; s < ph.size(); ++s) {
      q.phase_raw[s] = ph[s].num();
      q.phase[s] = q.phase_raw[s] * M_PI / 180.0;
    }
  }
}

// fill(node, ligand, force_field_params), data.cpp:246-276
bytedock::force_field_params rd_ffdata(const json& ff, bool ligand) {
  bytedock::force_field_params f;
  const json& bij = ff["FF_Bonds_atomidx"];
  f.bonds.resize(bij.size());
  for (size_t i = 0; i < bij.size(); ++i) {
    rd_ids(bij[i], f.bonds[i].ids, 2);
    f.bonds[i].k = ff["FF_Bonds_k"][i].num();
    f.bonds[i].length = ff["FF_Bonds_length"][i].num();
  }
  const json& aijk = ff["FF_Angles_atomidx"];
  f.angles.resize(aijk.size());
  for (size_t i = 0; i < aijk.size(); ++i) {
    rd_ids(aijk[i], f.angles[i].ids, 3);
    f.angles[i].k = ff["FF_Angles_k"][i].num();
    f.angles[i].degree = ff["FF_Angles_angle"][i].num();
    f.angles[i].radian = f.angles[i].degree * M_PI / 180.0;
  }
  rd_dihedrals(ff, "FF_ImproperTorsions", f.impropers);
  rd_dihedrals(ff, "FF_ProperTorsions", f.propers);
  if (ligand) {
    rd_list(ff["FF_vdW_paraidx"], f.vdw_types);
    for (auto& t : f.vdw_types) t -= 1000;  // nolmaliz_vdw_types, data.cpp:234-244
  } else {
    const json& s = ff["FF_vdW_sigma"];
    f.vdw_params.resize(s.size());
    for (size_t i = 0; i < s.size(); ++i)
      f.vdw_params[i] = {s[i].num(), ff["FF_vdW_epsilon"][i].num()};
  }
  for (const auto& v : ff["partial_charges"].items()) f.partial_charges.push_back(v.num());
  rd_pairs(ff["FF_Nonbonded14_atomidx"], f.pairs);
  rd_pairs(ff["FF_NonbondedAll_atomidx"], f.others);
  rd_list(ff["hydrophobic_atomidx"], f.hydrophobic_atoms);
  rd_list(ff["cation_atomidx"], f.cation_atoms);
  rd_list(ff["anion_atomidx"], f.anion_atoms);
  rd_rings(ff["piring5_atomidx"], f.pi5_rings);
  rd_rings(ff["piring6_atomidx"], f.pi6_rings);
  rd_pairs(ff["hbonddon_charged_atomidx"], f.hbonddon_charged);
  rd_pairs(ff["hbonddon_neut_atomidx"], f.hbonddon_neutral);
  rd_list(ff["hbondacc_charged_atomidx"], f.hbondacc_charged);
  rd_list(ff["hbondacc_neut_atomidx"], f.hbondacc_neutral);
  if (ligand) {
    // data.cpp:204-223
    if (ff.has("tstrain_params_pose_selection") && ff["tstrain_params_pose_selection"].size()) {
      const json& prm = ff["tstrain_params_pose_selection"];
      const json& mat = ff["tstrain_atomidx_pose_selection"];
      f.bad_torsions.resize(prm.size());
      for (size_t i = 0; i < prm.size(); ++i) {
        for (size_t k = 0; k < 4; ++k) f.bad_torsions[i].coefs[k] = prm[i][k].num();
        f.bad_torsions[i].matched.resize(mat[i].size());
        for (size_t j = 0; j < mat[i].size(); ++j) rd_ids(mat[i][j], f.bad_torsions[i].matched[j].ids, 4);
      }
    }
    if (ff.has("rotatable_bond_index")) rd_pairs(ff["rotatable_bond_index"], f.rotatable_bonds);
  }
  return f;
}

// fill(node, fragment_tree) and the bond pass, data.cpp:396-482
struct tree_holder {
  bytedock::fragment_tree t;
};

void set_bonds(bytedock::fragment& node,
               std::unordered_map<std::string, std::pair<index_t, index_t>>& repo) {
  for (auto& child : node.children) {
    auto& b = repo[std::to_string(node.id) + "->" + std::to_string(child.id)];
    child.parent = b.first;
    child.mine = b.second;
    set_bonds(child, repo);
  }
}

void rd_tree(const json& geo, bytedock::fragment_tree& t) {
  t.num_atoms = (index_t)geo["num_atoms"].integer();
  const json& split = geo["frag_split_index"];
  t.fragments.assign(split.size() + 1, nullptr);
  const json& levels = geo["frag_traverse_levels"];
  // roots first: the root vector must not reallocate once children point into it
  size_t n_roots = 0;
  for (const auto& level : levels.items())
    for (const auto& kv : level.fields())
      if (kv.second["parent"].size() == 0) ++n_roots;
  t.roots.reserve(n_roots);
  for (const auto& level : levels.items()) {
    for (const auto& key : level.keys()) {
      const json& node = level[key];
      const index_t id = (index_t)std::stoul(key);
      if (node["parent"].size() == 0) {
        t.roots.push_back({});
        t.roots.back().id = id;
        t.fragments[id] = &t.roots.back();
      }
      bytedock::fragment* out = t.fragments[id];
      const json& ch = node["children"];
      out->children.resize(ch.size());
      for (size_t c = 0; c < ch.size(); ++c) {
        out->children[c].id = (index_t)ch[c].integer();
        t.fragments[out->children[c].id] = &out->children[c];
      }
    }
  }
  index_t prev = 0;
  for (size_t f = 0; f < split.size(); ++f) {
    t.fragments[f]->start = prev;
    prev = (index_t)split[f].integer();
    t.fragments[f]->end = prev;
  }
  t.fragments[split.size()]->start = prev;
  t.fragments[split.size()]->end = t.num_atoms;
  rd_list(geo["permute_index"], t.permutation_index);
  t.reverse_index.assign(t.permutation_index.size(), 0);
  for (size_t i = 0; i < t.permutation_index.size(); ++i) t.reverse_index[t.permutation_index[i]] = i;
  std::unordered_map<std::string, std::pair<index_t, index_t>> repo;
  for (const auto& kv : geo["edge_to_bond"].fields())
    repo[kv.first] = {t.reverse_index[(size_t)kv.second[0].integer()],
                      t.reverse_index[(size_t)kv.second[1].integer()]};
  for (auto& r : t.roots) set_bonds(r, repo);
}

// the leaves of torsional_receptor::parse (data.cpp:292-364), behind the
// receptor's public accessors
struct receptor_leaves {
  std::vector<bytedock::rotatable_leaf> leaves;
  size_t num_torsions() const { return leaves.size(); }
  const bytedock::rotatable_leaf& get_leaf(size_t i) const { return leaves[i]; }
};

receptor_leaves rd_leaves(const json& geo, const json& xyz) {
  receptor_leaves r;
  for (const auto& key : geo["data"].keys()) {
    const json& b = geo["data"][key];
    for (size_t i = 0; i < b["rot_bond_1"].size(); ++i) {
      const size_t a0 = (size_t)b["rot_bond_1"][i].integer();
      const size_t a1 = (size_t)b["rot_bond_2"][i].integer();
      bytedock::rotatable_leaf lf;
      double n2 = 0.0;
      for (int k = 0; k < 3; ++k) {
        lf.axis_origin.xyz[k] = xyz[a0][k].num();
        lf.axis_vector.xyz[k] = xyz[a1][k].num() - lf.axis_origin.xyz[k];
        n2 += lf.axis_vector.xyz[k] * lf.axis_vector.xyz[k];
      }
      for (int k = 0; k < 3; ++k) lf.axis_vector.xyz[k] /= std::sqrt(n2);
      for (const auto& a : b["fragments"][i].items())
        if ((size_t)a.integer() != a1) lf.movable_atoms.push_back((index_t)a.integer());
      r.leaves.push_back(lf);
    }
  }
  return r;
}

void print_i32(const char* name, const int32_t* p, size_t n, bool comma = true) {
  std::printf("\"%s\": [", name);
  for (size_t i = 0; i < n; ++i) std::printf("%s%d", i ? ", " : "", p ? p[i] : 0);
  std::printf("]%s\n", comma ? "," : "");
}

void print_f64(const char* name, const double* p, size_t n, bool comma = true) {
  std::printf("\"%s\": [", name);
  for (size_t i = 0; i < n; ++i) std::printf("%s%.17g", i ? ", " : "", p ? p[i] : 0.0);
  std::printf("]%s\n", comma ? "," : "");
}

void print_mol(const pxd_mol& m, bool ligand) {
  std::printf("\"n_atoms\": %d,\n", m.n_atoms);
  if (ligand) {
    print_i32("vdw_type", m.vdw_type, m.n_atoms);
  } else {
    print_f64("sigma", m.sigma, m.n_atoms);
    print_f64("epsilon", m.epsilon, m.n_atoms);
  }
  print_f64("charge", m.charge, m.n_atoms);
  print_i32("bond_ij", m.bond_ij, 2 * (size_t)m.n_bonds);
  print_f64("bond_k", m.bond_k, m.n_bonds);
  print_f64("bond_r0", m.bond_r0, m.n_bonds);
  print_i32("angle_ijk", m.angle_ijk, 3 * (size_t)m.n_angles);
  print_f64("angle_k", m.angle_k, m.n_angles);
  print_f64("angle_t0", m.angle_t0, m.n_angles);
  print_i32("dih_ijkl", m.dih_ijkl, 4 * (size_t)m.n_dihedrals);
  print_i32("dih_n", m.dih_n, 6 * (size_t)m.n_dihedrals);
  print_f64("dih_phase", m.dih_phase, 6 * (size_t)m.n_dihedrals);
  print_f64("dih_k", m.dih_k, 6 * (size_t)m.n_dihedrals);
  print_i32("pairs14", m.pairs14, 2 * (size_t)m.n_pairs14);
  print_i32("pairs", m.pairs, 2 * (size_t)m.n_pairs);
  print_i32("hydrophobic", m.hydrophobic, m.n_hydrophobic);
  print_i32("cation", m.cation, m.n_cation);
  print_i32("anion", m.anion, m.n_anion);
  print_i32("ring5", m.ring5, 5 * (size_t)m.n_ring5);
  print_i32("ring6", m.ring6, 6 * (size_t)m.n_ring6);
  print_i32("don_charged", m.don_charged, 2 * (size_t)m.n_don_charged);
  print_i32("don_neutral", m.don_neutral, 2 * (size_t)m.n_don_neutral);
  print_i32("acc_charged", m.acc_charged, m.n_acc_charged);
  print_i32("acc_neutral", m.acc_neutral, m.n_acc_neutral);
  print_f64("tstrain_coef", m.tstrain_coef, 4 * (size_t)m.n_tstrain);
  print_i32("tstrain_off", m.tstrain_match_off, (size_t)m.n_tstrain + 1);
  const size_t nm = m.tstrain_match_off ? (size_t)m.tstrain_match_off[m.n_tstrain] : 0;
  print_i32("tstrain_ijkl", m.tstrain_ijkl, 4 * nm);
  std::printf("\"n_rotatable\": %d", m.n_rotatable);
}

}  // namespace

int main(int argc, char** argv) {
  if (argc != 3) {
    std::fprintf(stderr, "usage: %s receptor|ligand file.json\n", argv[0]);
    return 2;
  }
  try {
    std::ifstream in(argv[2]);
    std::stringstream ss;
    ss << in.rdbuf();
    const json doc = json::parse(ss.str());
    const bool ligand = std::string(argv[1]) == "ligand";
    bytedock::force_field_params ff = rd_ffdata(doc["ffdata"], ligand);
    auto md = pxdock_ref::to_pxd(ff, ligand);
    std::printf("{\n\"mol\": {\n");
    print_mol(md->mol, ligand);
    std::printf("\n}");
    if (ligand) {
      tree_holder th;
      rd_tree(doc["geometry"], th.t);
      auto td = pxdock_ref::to_pxd(th.t);
      const pxd_frag_tree& t = td->tree;
      std::printf(",\n\"tree\": {\n");
      print_i32("perm", t.perm, td->perm.size());
      print_i32("frag_start", t.frag_start, t.n_frags);
      print_i32("frag_end", t.frag_end, t.n_frags);
      print_i32("frag_parent", t.frag_parent, t.n_frags);
      print_i32("frag_bond", t.frag_bond, 2 * (size_t)t.n_frags);
      print_i32("tors_frag", t.tors_frag, t.n_torsions, false);
      std::printf("}");
    } else {
      receptor_leaves rl = rd_leaves(doc["geometry"], doc["xyz"]);
      auto ld = pxdock_ref::to_pxd_leaves(rl);
      const pxd_leaves& L = ld->leaves;
      std::printf(",\n\"leaves\": {\n");
      print_f64("origin", L.origin, 3 * (size_t)L.n_leaves);
      print_f64("axis", L.axis, 3 * (size_t)L.n_leaves);
      print_i32("atom_off", L.atom_off, (size_t)L.n_leaves + 1);
      print_i32("atoms", L.atoms, (size_t)L.atom_off[L.n_leaves], false);
      std::printf("}");
    }
    std::printf("\n}\n");
  } catch (const std::exception& e) {
    std::fprintf(stderr, "error: %s\n", e.what());
    return 1;
  }
  return 0;
}
