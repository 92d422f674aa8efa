// pxd_internal.hpp — device-side data layout shared by the kernels and the
// C-ABI implementation.  All device arrays are FP32/INT32 structure-of-arrays
// resident in HBM; per-pose working sets live in LDS (one wavefront per pose).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <mutex>
#include <vector>

#include "pxdock.h"

namespace pxd {

constexpr int kWave = 64;
constexpr int kMaxAtoms = PXD_MAX_LIGAND_ATOMS;  // per ligand (LDS budget)
constexpr int kMaxFrags = PXD_MAX_FRAGMENTS;
constexpr int kMaxLeaves = PXD_MAX_LEAVES;
constexpr int kMaxMovable = 64;                  // movable receptor atoms
constexpr int kMaxTypes = PXD_NUM_VDW_TYPES;
constexpr float kCoulomb = 332.0637141491396f;   // system.h:83
constexpr float kEps = 1e-20f;                   // math.h:128
constexpr float kScale14Coul = 0.833333333f;     // interaction.h:52
constexpr float kScale14Vdw = 0.5f;              // interaction.h:53

// Receptor-term atom reference: >= 0 fixed receptor atom, < 0 movable slot (-ref-1).
__host__ __device__ inline int mov_ref(int slot) { return -slot - 1; }

struct DevReceptor {
  int n_atoms;
  const float4* xyzq;   // [n] initial x, y, z, charge
  const float2* lj;     // [n] sigma, sqrt(epsilon)
  int n_fixed;          // atoms that are not movable (grid / brute-force inter)
  const int* fixed;     // [n_fixed]
  // leaves (data.cpp:348-363)
  int n_leaves;
  int n_mov;            // movable slots
  const float4* leaf_origin;  // [L]
  const float4* leaf_axis;    // [L]
  const int* mov_leaf;        // [M] leaf of each slot
  const int* mov_atom;        // [M] receptor atom of each slot
  // receptor intra terms that touch a movable atom (receptor.py:230-265)
  int n_bonds;
  const int2* bond_ab;        // refs
  const float2* bond_kr;
  int n_angles;
  const int4* angle_abc;      // refs (x, y, z)
  const float2* angle_kt;
  int n_dih;
  const int4* dih_abcd;       // refs
  const uint32_t* dih_per;    // 6 x 4-bit periodicities
  const float4* dih_kp;       // [n][3]: (k0,k1,k2,k3),(k4,k5,p0,p1),(p2,p3,p4,p5)
  // movable-vs-fixed non-bonded pairs, segmented per slot, each segment padded
  // to a multiple of 64 with inert entries
  const int* rp_seg;          // [M+1] offsets
  const float4* rp_pos;       // partner x, y, z, q * coul_scale
  const float2* rp_lj;        // partner sigma, sqrt(eps) * 4 * lj_scale
  // movable-vs-movable pairs (x = slot a, y = slot b, z = is14, w = leaf of
  // a | leaf of b << 16) and their
  // parameters (q_a q_b scale kCoulomb, (sigma_a + sigma_b) / 2,
  // 4 sqrt(eps_a eps_b) scale, 0), formed in FP64 on the host
  int n_mm;
  const int4* mm;
  const float4* mm_par;
  // pscore interaction sets (receptor side)
  int n_hydro, n_cation, n_anion, n_ring5, n_ring6;
  int n_don_c, n_don_n, n_acc_c, n_acc_n;
  const int* hydro;
  const int* cation;
  const int* anion;
  const int* ring5;           // [n][5]
  const int* ring6;           // [n][6]
  const int2* don_c;
  const int2* don_n;
  const int* acc_c;
  const int* acc_n;
  const int* atom_mov_slot;   // [n] slot or -1
  // fixed atoms compacted for the grid build
  const float4* fixed_xyzq;
  const float2* fixed_lj;
  const double4* fixed_xyzq64;  // the same in FP64 (x, y, z, q) and (sigma, sqrt eps)
  const double2* fixed_lj64;
  // FP64 energy of receptor terms touching no movable atom (pose independent)
  double e_const;
  // Term lists above are ordered by leaf class: the terms whose movable atoms
  // all belong to leaf k occupy [seg[k], seg[k+1]); terms spanning two or more
  // leaves follow from seg[n_leaves].  Movable slots of leaf k are
  // [leaf_slot[k], leaf_slot[k+1]).
  const int* bond_seg;        // [L+1]
  const int* angle_seg;
  const int* dih_seg;
  const int* mm_seg;
  const int* leaf_slot;
  // Leaf energy tables: with every other atom fixed, the energy of the terms
  // of leaf k is a 2*pi-periodic function T_k(theta_k).  tab[k*tab_n + i]
  // holds (T_k, dT_k/dtheta) at theta = 2*pi*i/tab_n in FP64; a lookup is a
  // cubic Hermite interpolation.  tab == nullptr: every term is evaluated
  // directly.
  int tab_n;
  const double2* tab;
};

// Per-ligand record of the packed batch.
struct LigInfo {
  int n_atoms, atom_off;
  int bond_off, n_bonds;
  int angle_off, n_angles;
  int dih_off, n_dih;
  int pair_off, n_pairs;
  int frag_off, n_frags, n_tors;
  int hydro_off, n_hydro;
  int cation_off, n_cation;
  int anion_off, n_anion;
  int ring5_off, n_ring5;
  int ring6_off, n_ring6;
  int donc_off, n_donc;
  int donn_off, n_donn;
  int accc_off, n_accc;
  int accn_off, n_accn;
  int ts_off, n_ts;
  int n_rot;
  int tsm_off;   // first entry of this ligand in ts_moff (n_ts + 1 entries)
  int pmode;     // 1: non-bonded pairs as per-atom partner masks (pmask), 0: pair list
  int gat_off;   // first gather word of this ligand in gat (64 per term chunk)
};

struct DevLigands {
  int n;
  const LigInfo* info;
  const int* type;        // per atom GAFF2 index
  const float* q;         // per atom charge
  const float2* lj;       // per atom sigma, sqrt(epsilon) (GAFF2 table)
  const uint32_t* bond_ij;  // i | j << 8
  const float2* bond_kr;
  const uint32_t* angle_ijk;  // i | j << 8 | k << 16
  const float2* angle_kt;
  const uint32_t* dih_ijkl;   // i | j << 8 | k << 16 | l << 24
  const uint32_t* dih_per;
  const float4* dih_kp;       // [n][3] as DevReceptor
  const uint32_t* pairs;      // i | j << 8 | is14 << 16
  // fragment trees (permuted-slot form of pxd_frag_tree)
  const int* frag;            // [frag_off*6]: start, end, parent, bond_parent, bond_mine, pad
  const int* tors_frag;       // [frag_off]  (per ligand n_tors entries)
  const int* perm;            // [atom_off] permuted slot -> atom
  // pscore sets (ligand-local atom indices)
  const int* sets;            // hydro/cation/anion/acc lists packed
  const int* rings;           // ring5 (5 ints) / ring6 (6 ints) packed
  const int2* dons;           // donor pairs
  const float4* ts_coef;      // torsion-strain coefs
  const int* ts_moff;         // per ligand n_ts+1 absolute offsets into ts_ijkl
  const uint32_t* ts_ijkl;    // packed atoms
  // non-bonded partner masks, [atom_off + a][2]: bit b of word b/32 set when
  // (a, b) is an all-pairs entry (.x..w of [0]) or a 1-4 entry ([1])
  const uint4* pmask;
  // bonded-gradient gather: the terms of a list are taken 64 at a time (one
  // per lane); every (term, atom) incidence of a chunk owns one float4 stage
  // slot, numbered atom by atom so an atom's slots are contiguous.
  // *_slot[t] holds the slot of each of the term's atoms (8 bits each, in
  // term order); gat[gat_off + 64 c + lane] holds first slot | count << 8 of
  // atom lane and (<< 16) of atom lane + 64 in chunk c (bond chunks, then
  // angle chunks, then dihedral chunks).
  const uint32_t* bond_slot;
  const uint32_t* angle_slot;
  const uint32_t* dih_slot;
  const uint32_t* gat;
};

struct DevGrid {
  float init[3];
  int np[3];
  float spacing, inv_spacing;
  double init_d[3];       // FP64 copies for the cell coordinate of a lookup
  double inv_spacing_d;
  float slope;
  long long P;                // points per map
  const float4* maps;         // [(n_types+1)][P], map 0 = Coulomb
  int type_slot[kMaxTypes];   // GAFF2 index -> map index (-1: none)
  // layout 1: every cell's 8 corner points stored together (one 128-byte
  // line per map and lookup), [(n_types+1)][C][8], corner c = x | y << 1 | z << 2
  int layout;
  long long C;                // cells per map: (nx-1)(ny-1)(nz-1)
  const float4* cmaps;
  // layout 2: point-major component planes, [(n_types+1)][4][P] floats
  // (value, dE/dx, dE/dy, dE/dz)
  const float* smaps;
};

// The two default-priority side streams of the current device on which the
// batched L-BFGS and the search run their smaller size classes beside the
// caller's stream.  They are created once per device and kept: a process has
// only a few hardware queues (GPU_MAX_HW_QUEUES, 4 by default), and streams
// beyond them share a queue with another stream, which serialises the class
// launches that were meant to overlap.  A caller holds side_stream_mutex()
// from its first use of the streams until its work on them has finished.
constexpr int kSideStreams = 2;
std::mutex& side_stream_mutex();
hipError_t side_streams(hipStream_t out[kSideStreams]);

}  // namespace pxd

// Opaque handle definitions (C-ABI side).
struct pxd_receptor {
  pxd::DevReceptor dev;
  int n_atoms;
  void* mem;                  // single device allocation backing dev arrays
  size_t mem_bytes;
  void* tab_mem;              // leaf energy tables (nullptr: none)
  int tab_n_built;            // nodes per leaf of the built tables (0: none)
  double tab_err_e, tab_err_d; // worst midpoint errors measured at build
  // host copies needed by other entry points
  float* h_xyz;               // initial coordinates
  int n_leaves;
  int n_mov;
  uint64_t key;               // pxd_receptor_key of the inputs
  std::vector<int32_t> movable;  // atoms of the leaves, ascending
  // FP64 inputs for the host-staged FP64 terms (the affinity score's GBSA)
  std::vector<double> h_xyz64, h_q64;      // [n][3], [n]
  std::vector<double> h_leaf_o, h_leaf_a;  // [n_leaves][3] origin, unit axis
  std::vector<int> h_mov_atom, h_mov_leaf; // movable slots in leaf order
};

struct pxd_grid {
  pxd::DevGrid dev;
  void* cmem;                 // lookup copy of layout 1 or 2 (nullptr: layout 0)
  pxd_box box;
  int n_types;
  int32_t types[PXD_NUM_VDW_TYPES];
  void* mem;
  const pxd_receptor* receptor;
  double slope;               // out-of-box wall slope as given (dev.slope is its FP32)
};

struct pxd_ligand_batch {
  pxd::DevLigands dev;
  void* mem;
  size_t mem_bytes;
  int n;
  int* h_natoms;
  int* h_ntors;
  int max_atoms;              // largest ligand
  int max_frags;              // most fragments of a ligand
  int has_trees;              // fragment trees were given (pose parameterisation)
  uint8_t type_used[PXD_NUM_VDW_TYPES];
  std::vector<double> h_q64;  // FP64 charges of every atom, batch order
  std::vector<int> h_atom_off; // first atom of each ligand in h_q64
};
