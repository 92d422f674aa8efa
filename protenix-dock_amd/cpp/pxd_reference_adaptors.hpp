// pxd_reference_adaptors.hpp — the descriptors of include/pxdock.h built from
// Protenix-Dock's own engine data structures, for a maintainer who calls the
// MI355X path from inside the reference tree.
//
// Needs only the reference's standard-library headers:
//   bytedock/core/system.h  force_field_params, fragment_tree, rotable_geometry
//   bytedock/core/data.h    rotatable_leaf, torsional_receptor, free_ligand
// (engine/include/bytedock/core/system.h:32-154, data.h:28-99).
//
// The reference keeps AoS vectors of small structs with size_t indices
// (index_t, dtype.h:29-30); the ABI takes SoA int32 arrays and FP64
// parameters.  So every adaptor COPIES into storage it owns: the returned
// object must outlive the pxd_* call that reads its descriptor (the library
// copies the data to the device at pxd_receptor_create /
// pxd_ligand_batch_create and keeps no pointer to it).  Indices above
// INT32_MAX, vdW types outside the 99-entry table and rings of another size
// than their list's throw std::out_of_range.
#pragma once

#include <cstdint>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "bytedock/core/data.h"
#include "bytedock/core/system.h"
#include "pxdock.h"

namespace pxdock_ref {

inline int32_t idx32(bytedock::index_t v, const char* what) {
  if (v > (bytedock::index_t)INT32_MAX)
    throw std::out_of_range(std::string(what) + ": index beyond int32");
  return (int32_t)v;
}

// force_field_params (system.h:85-109) -> pxd_mol.  ligand: the GAFF2 type
// ids of vdw_types (already offset by free_ligand::parse, data.cpp:234-244)
// are used; receptor: the per-atom (sigma, epsilon) of vdw_params.
struct mol_desc {
  pxd_mol mol{};
  std::vector<int32_t> vdw_type, bond_ij, angle_ijk, dih_ijkl, dih_n, pairs14, pairs, hydrophobic,
      cation, anion, ring5, ring6, don_charged, don_neutral, acc_charged, acc_neutral, ts_off,
      ts_ijkl;
  std::vector<double> sigma, epsilon, charge, bond_k, bond_r0, angle_k, angle_t0, dih_phase,
      dih_k, ts_coef;
  mol_desc() = default;
  mol_desc(const mol_desc&) = delete;  // the descriptor points into these vectors
  mol_desc& operator=(const mol_desc&) = delete;
};

inline void put_ids(std::vector<int32_t>& dst, const bytedock::index_t* ids, int width,
                    const char* what) {
  for (int k = 0; k < width; ++k) dst.push_back(idx32(ids[k], what));
}

inline std::unique_ptr<mol_desc> to_pxd(const bytedock::force_field_params& f, bool ligand) {
  auto d = std::make_unique<mol_desc>();
  const size_t n = f.partial_charges.size();
  d->charge.assign(f.partial_charges.begin(), f.partial_charges.end());
  if (ligand) {
    if (f.vdw_types.size() != n) throw std::out_of_range("vdw_types: one per atom");
    for (auto t : f.vdw_types) {
      if (t >= (bytedock::index_t)PXD_NUM_VDW_TYPES)
        throw std::out_of_range("vdw_types: outside the 99-entry table");
      d->vdw_type.push_back((int32_t)t);
    }
  } else {
    if (f.vdw_params.size() != n) throw std::out_of_range("vdw_params: one per atom");
    for (const auto& p : f.vdw_params) {
      d->sigma.push_back(p.sigma);
      d->epsilon.push_back(p.epsilon);
    }
  }
  for (const auto& b : f.bonds) {
    put_ids(d->bond_ij, b.ids, 2, "bonds");
    d->bond_k.push_back(b.k);
    d->bond_r0.push_back(b.length);
  }
  for (const auto& a : f.angles) {
    put_ids(d->angle_ijk, a.ids, 3, "angles");
    d->angle_k.push_back(a.k);
    d->angle_t0.push_back(a.radian);
  }
  // impropers first, then propers: the order intra_molecular_interactions
  // sums them in (interaction.cpp:104-117).  A series ends at its first
  // periodicity 0 (interaction.cpp:88); later slots are written as zero.
  auto put_dih = [&](const bytedock::fourier_dihedral& q) {
    put_ids(d->dih_ijkl, q.ids, 4, "dihedrals");
    bool live = true;
    for (int s = 0; s < PXD_DIH_TERMS; ++s) {
      live = live && q.periodicity[s] != 0;
      d->dih_n.push_back(live ? idx32(q.periodicity[s], "periodicity") : 0);
      d->dih_phase.push_back(live ? q.phase[s] : 0.0);
      d->dih_k.push_back(live ? q.k[s] : 0.0);
    }
  };
  for (const auto& q : f.impropers) put_dih(q);
  for (const auto& q : f.propers) put_dih(q);
  for (const auto& p : f.pairs) put_ids(d->pairs14, p.ids, 2, "1-4 pairs");
  for (const auto& p : f.others) put_ids(d->pairs, p.ids, 2, "pairs");
  for (auto a : f.hydrophobic_atoms) d->hydrophobic.push_back(idx32(a, "hydrophobic"));
  for (auto a : f.cation_atoms) d->cation.push_back(idx32(a, "cation"));
  for (auto a : f.anion_atoms) d->anion.push_back(idx32(a, "anion"));
  for (const auto& r : f.pi5_rings) {
    if (r.size != 5) throw std::out_of_range("pi5_rings: a ring of 5 atoms expected");
    put_ids(d->ring5, r.ids, 5, "pi5_rings");
  }
  for (const auto& r : f.pi6_rings) {
    if (r.size != 6) throw std::out_of_range("pi6_rings: a ring of 6 atoms expected");
    put_ids(d->ring6, r.ids, 6, "pi6_rings");
  }
  for (const auto& p : f.hbonddon_charged) put_ids(d->don_charged, p.ids, 2, "donors");
  for (const auto& p : f.hbonddon_neutral) put_ids(d->don_neutral, p.ids, 2, "donors");
  for (auto a : f.hbondacc_charged) d->acc_charged.push_back(idx32(a, "acceptors"));
  for (auto a : f.hbondacc_neutral) d->acc_neutral.push_back(idx32(a, "acceptors"));
  d->ts_off.push_back(0);
  for (const auto& t : f.bad_torsions) {
    for (int k = 0; k < 4; ++k) d->ts_coef.push_back(t.coefs[k]);
    for (const auto& m : t.matched) put_ids(d->ts_ijkl, m.ids, 4, "torsion strain");
    d->ts_off.push_back((int32_t)(d->ts_ijkl.size() / 4));
  }
  pxd_mol& m = d->mol;
  m.n_atoms = (int32_t)n;
  m.vdw_type = ligand ? d->vdw_type.data() : nullptr;
  m.sigma = ligand ? nullptr : d->sigma.data();
  m.epsilon = ligand ? nullptr : d->epsilon.data();
  m.charge = d->charge.data();
  m.n_bonds = (int32_t)f.bonds.size();
  m.bond_ij = d->bond_ij.data();
  m.bond_k = d->bond_k.data();
  m.bond_r0 = d->bond_r0.data();
  m.n_angles = (int32_t)f.angles.size();
  m.angle_ijk = d->angle_ijk.data();
  m.angle_k = d->angle_k.data();
  m.angle_t0 = d->angle_t0.data();
  m.n_dihedrals = (int32_t)(f.impropers.size() + f.propers.size());
  m.dih_ijkl = d->dih_ijkl.data();
  m.dih_n = d->dih_n.data();
  m.dih_phase = d->dih_phase.data();
  m.dih_k = d->dih_k.data();
  m.n_pairs14 = (int32_t)f.pairs.size();
  m.pairs14 = d->pairs14.data();
  m.n_pairs = (int32_t)f.others.size();
  m.pairs = d->pairs.data();
  m.n_hydrophobic = (int32_t)d->hydrophobic.size();
  m.hydrophobic = d->hydrophobic.data();
  m.n_cation = (int32_t)d->cation.size();
  m.cation = d->cation.data();
  m.n_anion = (int32_t)d->anion.size();
  m.anion = d->anion.data();
  m.n_ring5 = (int32_t)f.pi5_rings.size();
  m.ring5 = d->ring5.data();
  m.n_ring6 = (int32_t)f.pi6_rings.size();
  m.ring6 = d->ring6.data();
  m.n_don_charged = (int32_t)f.hbonddon_charged.size();
  m.don_charged = d->don_charged.data();
  m.n_don_neutral = (int32_t)f.hbonddon_neutral.size();
  m.don_neutral = d->don_neutral.data();
  m.n_acc_charged = (int32_t)d->acc_charged.size();
  m.acc_charged = d->acc_charged.data();
  m.n_acc_neutral = (int32_t)d->acc_neutral.size();
  m.acc_neutral = d->acc_neutral.data();
  m.n_tstrain = (int32_t)f.bad_torsions.size();
  m.tstrain_coef = d->ts_coef.data();
  m.tstrain_match_off = d->ts_off.data();
  m.tstrain_ijkl = d->ts_ijkl.data();
  m.n_rotatable = (int32_t)f.rotatable_bonds.size();
  return d;
}

// fragment_tree (system.h:129-154, data.cpp:410-482) -> pxd_frag_tree.
// Fragment f of the descriptor is fragment id f; the torsion order is the
// order free_ligand::apply_parameters consumes torsions in (every root's
// children, each subtree in pre-order, data.cpp:592-641).  free_ligand keeps
// its tree private: the maintainer adds an accessor returning geo_.
struct tree_desc {
  pxd_frag_tree tree{};
  std::vector<int32_t> perm, start, end, parent, bond, tors;
  tree_desc() = default;
  tree_desc(const tree_desc&) = delete;
  tree_desc& operator=(const tree_desc&) = delete;
};

inline std::unique_ptr<tree_desc> to_pxd(const bytedock::fragment_tree& t) {
  auto d = std::make_unique<tree_desc>();
  const size_t nf = t.fragments.size();
  d->start.assign(nf, 0);
  d->end.assign(nf, 0);
  d->parent.assign(nf, -1);
  d->bond.assign(2 * nf, 0);
  for (auto p : t.permutation_index) d->perm.push_back(idx32(p, "permutation_index"));
  auto put = [&](const bytedock::fragment& fr, int32_t par) {
    const int32_t id = idx32(fr.id, "fragment id");
    if ((size_t)id >= nf) throw std::out_of_range("fragment id beyond the fragment count");
    d->start[id] = idx32(fr.start, "fragment start");
    d->end[id] = idx32(fr.end, "fragment end");
    d->parent[id] = par;
    if (par >= 0) {
      d->bond[2 * id] = idx32(fr.parent, "bond atom");
      d->bond[2 * id + 1] = idx32(fr.mine, "bond atom");
    }
    return id;
  };
  // pre-order walk (explicit stack, children in their stored order)
  for (const auto& root : t.roots) {
    const int32_t rid = put(root, -1);
    std::vector<std::pair<const bytedock::fragment*, int32_t>> stack;
    for (auto it = root.children.rbegin(); it != root.children.rend(); ++it)
      stack.push_back({&*it, rid});
    while (!stack.empty()) {
      auto [fr, par] = stack.back();
      stack.pop_back();
      const int32_t id = put(*fr, par);
      d->tors.push_back(id);
      for (auto it = fr->children.rbegin(); it != fr->children.rend(); ++it)
        stack.push_back({&*it, id});
    }
  }
  pxd_frag_tree& r = d->tree;
  r.n_frags = (int32_t)nf;
  r.perm = d->perm.data();
  r.frag_start = d->start.data();
  r.frag_end = d->end.data();
  r.frag_parent = d->parent.data();
  r.frag_bond = d->bond.data();
  r.n_torsions = (int32_t)d->tors.size();
  r.tors_frag = d->tors.data();
  return d;
}

// Receptor leaves (torsional_receptor::parse, data.cpp:348-363) -> pxd_leaves.
// Works on anything with num_torsions() and get_leaf(i) returning a
// rotatable_leaf, torsional_receptor's public interface (data.h:34-71).
struct leaves_desc {
  pxd_leaves leaves{};
  std::vector<double> origin, axis;
  std::vector<int32_t> off, atoms;
  leaves_desc() = default;
  leaves_desc(const leaves_desc&) = delete;
  leaves_desc& operator=(const leaves_desc&) = delete;
};

template <class Receptor>
std::unique_ptr<leaves_desc> to_pxd_leaves(const Receptor& rec) {
  auto d = std::make_unique<leaves_desc>();
  const size_t n = rec.num_torsions();
  d->off.push_back(0);
  for (size_t i = 0; i < n; ++i) {
    const bytedock::rotatable_leaf& lf = rec.get_leaf(i);
    for (int k = 0; k < 3; ++k) {
      d->origin.push_back(lf.axis_origin.xyz[k]);
      d->axis.push_back(lf.axis_vector.xyz[k]);
    }
    for (auto a : lf.movable_atoms) d->atoms.push_back(idx32(a, "movable atom"));
    d->off.push_back((int32_t)d->atoms.size());
  }
  pxd_leaves& L = d->leaves;
  L.n_leaves = (int32_t)n;
  L.origin = d->origin.data();
  L.axis = d->axis.data();
  L.atom_off = d->off.data();
  L.atoms = d->atoms.data();
  return d;
}

// box_config (grid.h:31-40; its header needs Boost, hence the template) ->
// pxd_box: the FP64 fields copied, the size_t point counts narrowed to int32.
template <class BoxConfig>
pxd_box to_pxd_box(const BoxConfig& bc) {
  pxd_box b{};
  for (int k = 0; k < 3; ++k) {
    b.init_xyz[k] = bc.init_xyz[k];
    b.oppo_xyz[k] = bc.oppo_xyz[k];
    b.npoints[k] = idx32(bc.npoints[k], "npoints");
    b.npoints_minus_one[k] = idx32(bc.npoints_minus_one[k], "npoints_minus_one");
  }
  b.spacing = bc.spacing;
  b.spacing_inv = bc.spacing_inv;
  return b;
}

// Coordinates (molecule_pose, system.h:111-114) as the flat [n][3] FP64
// array pxd_receptor_create takes, or the FP32 one the pose buffers hold.
inline std::vector<double> flat_xyz(const bytedock::molecule_pose& p) {
  std::vector<double> out;
  out.reserve(3 * p.size());
  for (const auto& a : p) out.insert(out.end(), a.xyz, a.xyz + 3);
  return out;
}

inline std::vector<float> flat_xyz_f32(const bytedock::molecule_pose& p) {
  std::vector<float> out;
  out.reserve(3 * p.size());
  for (const auto& a : p)
    for (int k = 0; k < 3; ++k) out.push_back((float)a.xyz[k]);
  return out;
}

}  // namespace pxdock_ref
