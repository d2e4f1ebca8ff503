// Device descriptors of the heightfield-fit problems and the launcher shared by the
// C ABI and the adjoint kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <string>

#include "wt_device.hpp"

namespace wt {

constexpr int kFitMaxWaves = 128;
constexpr int kFitMaxSnapshots = 32;
constexpr int kFitThreads = 256;
// Snapshot classes S <= 4, 8, ..., 32 and the grid points each thread of their adjoint
// kernel carries (its residuals, KP x SMAX x 3 floats, stay in registers). The kernel
// of a class steps all SMAX snapshots without branches (residuals past S are zero), so
// classes 4 snapshots wide waste at most 3 per point.
constexpr int kFitClasses = 8;
constexpr int kFitClassMaxS[kFitClasses] = {4, 8, 12, 16, 20, 24, 28, 32};
constexpr int kFitPointsPerThread[kFitClasses] = {2, 2, 1, 1, 1, 1, 1, 1};
// Wavefronts per SIMD each class's adjoint kernel is compiled for (1: no request). The
// classes listed fit one more wavefront per SIMD than the register allocator's own
// choice gives, with no spills and the same instructions (4: 84 -> 80 VGPRs, 12: 83 ->
// 80, 20: 101 -> 96); for the others the next step spills.
constexpr int kFitWavesPerSimd[kFitClasses] = {6, 1, 6, 1, 5, 1, 1, 1};
inline int fit_class(int S) { return (S - 1) / 4; }

// One problem of a batch as the kernels see it. Device pointers are owned by the
// wt_fit handle except params and target, which the caller owns.
struct FitDesc {
  double spacing, x0, z0, t0, dt, g;
  const float* target;  // [S][n*n][3]
  float* params;        // [W][5] {A, L, alpha, Q, phi0}
  float* adam_m;        // [W*5]
  float* adam_v;        // [W*5]
  float* partials;      // [W*5 + 1][nblocks]: per-block sums of r . dd/dp and of |r|^2 / 2
  float* loss;          // [1] loss of the last evaluation
  int* step;            // [1] Adam step counter
  int n, S, W, nblocks;         // nblocks: adjoint blocks of kFitThreads x KP points
  long long block_offset;      // first block of this problem in the block -> problem map
  long long row_offset;        // first row of this problem in the reduce grid
  long long param_out_offset;  // first parameter of this problem in a concatenated gradient
};

struct FitAdam {
  float lr, beta1, beta2, eps;
};

// Device tables of a batch: problems, block -> problem (blocks grouped by snapshot
// class, class c at [class_begin[c], + class_blocks[c])), reduce row -> problem.
// Streams beside the caller's on which the snapshot classes' adjoint launches run
// concurrently (one class's tail overlaps the next class's blocks), joined by events
// before the reduction.
constexpr int kFitSideStreams = 3;

struct FitLaunch {
  const FitDesc* probs;
  const int* block_prob;
  const int* row_prob;
  long long class_begin[kFitClasses];
  long long class_blocks[kFitClasses];
  long long total_rows;
  hipStream_t side[kFitSideStreams];  // null: every class on the caller's stream
  hipEvent_t fork;
  hipEvent_t join[kFitSideStreams];
};

// mode 0: gradient + loss only (grad_out receives the concatenated gradients);
// mode 1: Adam step (the adjoint launch bumps each problem's step counter first).
hipError_t launch_fit_iteration(const FitLaunch& L, const FitAdam& adam, int mode, float* grad_out, float* loss_out,
                                hipStream_t st);

// Target of a problem from its own params (the hidden wave set) plus N(0, sigma) noise.
hipError_t launch_synthesize(const FitDesc& P, float* out, float sigma, unsigned long long seed, hipStream_t st);

// Zero n 32-bit words with a kernel (graph-capture safe).
hipError_t launch_zero_words(uint32_t* p, size_t n, hipStream_t st);

// Writes one problem's descriptor (passed by value) to d_desc, fills the block and row
// maps (map_words ints at d_maps) with problem 0, and zeroes zero_words 32-bit words at
// zero_begin, in one kernel: the set-up of a single-problem evaluation in memory the
// caller did not allocate per call.
hipError_t launch_fit_setup_single(const FitDesc& d, FitDesc* d_desc, int* d_maps, size_t map_words,
                                   uint32_t* zero_begin, size_t zero_words, hipStream_t st);

// The device address of this library's fixed single-problem fit scratch on `device`
// (resolved once per device, with that device current), or null with a message.
void* static_fit_scratch(int device, std::string* error);
// Its size in bytes.
size_t static_fit_scratch_bytes();

}  // namespace wt
