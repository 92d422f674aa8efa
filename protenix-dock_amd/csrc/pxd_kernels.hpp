// pxd_kernels.hpp — launcher interface between the C-ABI and the kernels.
#pragma once

#include <hip/hip_runtime.h>

#include "pxd_internal.hpp"

namespace pxd {

constexpr int kMaxHistory = 128;  // L-BFGS history cap (LDS for rho/alpha)
constexpr int kLbfgsVpl = 7;      // vector elements per lane: 3*128 + 64 <= 64*7

struct GridBuildArgs {
  const float4* atoms;  // fixed receptor atoms x, y, z, q
  const float2* alj;    // sigma, sqrt(eps)
  int n_atoms;
  const float2* tlj;    // per requested type: sigma, 4 * sqrt(eps)
  int n_types;
  float init[3];
  int np[3];
  float spacing;
  long long P;
  float4* maps;
};

struct ScoreArgs {
  int n_poses;
  const int* pose_lig;
  const int* pose_off;
  const float* xyz;
  const float* tors;
  double* energy;
  float* lig_grad;
  float* tors_grad;
};

struct LbfgsParamsDev {
  int history, max_iter, max_eval, max_ls;
  double lr, tol_grad, tol_change, c1, c2, ls_tol_change, gtd_threshold;
};

struct LbfgsArgs {
  LbfgsParamsDev prm;
  int n_poses;
  const int* pose_lig;
  const int* pose_off;
  float* xyz;
  float* tors;
  double* energy;
  int* nevals;
  int* converged;
  int* counter;   // work-queue head, zero before launch
  float* hist;    // [n_slots][2][history][64*kLbfgsVpl]
};

struct PscoreArgs {
  int n_poses;
  const int* pose_lig;
  const int* pose_off;
  const float* xyz;
  const float* tors;
  double* score;
  double* terms;
};

void launch_grid_build(const GridBuildArgs& a, hipStream_t st);
void launch_score(const DevReceptor& R, const DevLigands& L, const DevGrid* G,
                  const ScoreArgs& a, hipStream_t st);
void launch_lbfgs(const DevReceptor& R, const DevLigands& L, const DevGrid* G,
                  const LbfgsArgs& a, int n_slots, hipStream_t st);
void launch_pscore(const DevReceptor& R, const DevLigands& L, const pxd_score_params& sp,
                   const PscoreArgs& a, hipStream_t st);

}  // namespace pxd
