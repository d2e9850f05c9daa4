// ffm_soft.h -- layouts of the differentiable soft splat shared by the host
// runtime and softsplat.hip.
#pragma once

#include "ffm_internal.h"

namespace ffm {

// One fit of a (possibly batched) soft-splat launch. Blocks of all fits are
// numbered back to back; fit i owns blocks block0 .. block0 + n_blocks - 1.
struct SoftFit {
  uint32_t table_off;   // floats into the concatenated flame tables
  int32_t n_xforms, n_vars, n_params;
  ViewConst view;
  uint64_t walkers;
  uint32_t ipt, burn_in, seed, window;
  int32_t truncate;
  int32_t f64;          // 1: the walk and its adjoint in f64
  float density_scale;
  uint64_t theta_off;   // floats into theta / grad_theta
  uint64_t pix_off;     // floats into density / grad_density
  uint64_t ckpt_off;    // bytes into the checkpoint scratch (windows x walkers states)
  uint64_t part_off;    // doubles into the partials (n_params x n_blocks)
  uint32_t block0, n_blocks;
};

// The forward splats into one private density copy per XCD, folded after.
constexpr int SOFT_XCD_COPIES = 8;

struct SoftArgs {
  const SoftFit* fits;
  int32_t n_fits;
  const float* tables;
  const float* theta;
  float* density;             // forward output
  float* dens_xcd;            // forward scratch: SOFT_XCD_COPIES x total_pixels
  uint64_t total_pixels;      // density floats of all fits
  const float* grad_density;  // adjoint input
  unsigned char* ckpt;        // adjoint scratch: window-start states
  double* partials;           // adjoint scratch: per-block sums
  float* grad_theta;          // adjoint output
};

size_t soft_forward_lds_bytes(int n_xforms);
size_t soft_adjoint_lds_bytes(int n_xforms, int n_params, int window, bool f64);
size_t soft_ckpt_bytes(uint32_t ipt, uint32_t window, uint64_t walkers, bool f64);
int soft_block();  // threads per block (one walker each) of both kernels
hipError_t launch_soft_forward(const SoftArgs& a, uint32_t blocks, size_t lds, bool any_f32,
                               bool any_f64, hipStream_t s);
// lds32 / lds64: dynamic LDS of the f32 / f64 walk's launch, 0 when no fit uses it
hipError_t launch_soft_adjoint(const SoftArgs& a, uint32_t blocks, size_t lds32, size_t lds64,
                               hipStream_t s);

}  // namespace ffm
