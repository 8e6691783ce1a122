// Deterministic FP64 reductions for CDNA (64-lane wavefronts).
//
// wave_sum: DPP row butterflies (quad_perm, row_ror) inside each 16-lane row,
// then two cross-row exchanges; every lane ends with the wave total.
// Grid reductions are two-pass: every block writes its partial sums in block
// order (block_partials), and a one-block finalizer kernel adds them in a
// fixed order (reduce_partials).  No fences or atomics inside the producing
// kernel (on a multi-XCD part an agent-scope fence per block writes back L2),
// and the result depends only on the data and the grid size.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace lym {

constexpr int kBlock = 256;  // threads per block of every site kernel
constexpr int kNumXcd = 8;   // MI355X: 8 XCDs, workgroups dispatched round-robin

template <int CTRL>
__device__ __forceinline__ double dpp_add(double v) {
  const int lo = __double2loint(v);
  const int hi = __double2hiint(v);
  const int lo2 = __builtin_amdgcn_update_dpp(0, lo, CTRL, 0xF, 0xF, false);
  const int hi2 = __builtin_amdgcn_update_dpp(0, hi, CTRL, 0xF, 0xF, false);
  return v + __hiloint2double(hi2, lo2);
}

__device__ __forceinline__ double wave_sum(double v) {
  v = dpp_add<0xB1>(v);   // quad_perm [1,0,3,2]
  v = dpp_add<0x4E>(v);   // quad_perm [2,3,0,1]
  v = dpp_add<0x124>(v);  // row_ror:4
  v = dpp_add<0x128>(v);  // row_ror:8
  v += __shfl_xor(v, 16, 64);
  v += __shfl_xor(v, 32, 64);
  return v;
}

// Logical block index such that each XCD receives one contiguous range of the
// lattice (neighbouring sites then share that XCD's L2).  gridDim.x must be a
// multiple of kNumXcd.
__device__ __forceinline__ int64_t xcd_block() {
  const int64_t per = gridDim.x / kNumXcd;
  return (int64_t)(blockIdx.x % kNumXcd) * per + blockIdx.x / kNumXcd;
}

inline int grid_for(int64_t n, int per_block = kBlock) {
  int64_t g = (n + per_block - 1) / per_block;
  if (g < 1) g = 1;
  g = ((g + kNumXcd - 1) / kNumXcd) * kNumXcd;
  return (int)g;
}

struct RedCtl {
  double* partials;  // [NR][gridDim.x] written by the producing kernel
  unsigned* ticket;  // unused by the two-pass scheme (kept for layout)
};

// Block sums of v[NR] -> partials[r * G + idx] (thread 0 writes); by default
// G = gridDim.x, idx = blockIdx.x.  blockDim.x must be a multiple of 64,
// <= 64 * MAXW.
template <int NR, int MAXW = kBlock / 64>
__device__ __forceinline__ void block_partials(const double (&v)[NR], double* partials, int idx = -1, int G = -1) {
  if (idx < 0) idx = blockIdx.x;
  if (G < 0) G = gridDim.x;
  __shared__ double sm[NR][MAXW];
  const int lane = threadIdx.x & 63;
  const int w = threadIdx.x >> 6;
  const int nw = blockDim.x >> 6;
#pragma unroll
  for (int r = 0; r < NR; ++r) {
    const double t = wave_sum(v[r]);
    if (lane == 0) sm[r][w] = t;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int r = 0; r < NR; ++r) {
      double t = 0.0;
      for (int i = 0; i < nw; ++i) t += sm[r][i];
      partials[(int64_t)r * G + idx] = t;
    }
  }
}

// One block of kBlock threads: sums[r] = sum_b partials[r * G + b], fixed order.
// Every thread returns the totals.
template <int NR>
__device__ __forceinline__ void reduce_partials(const double* partials, int G, double (&sums)[NR]) {
  __shared__ double sm[NR][kBlock / 64];
  const int lane = threadIdx.x & 63;
  const int w = threadIdx.x >> 6;
#pragma unroll
  for (int r = 0; r < NR; ++r) {
    double t = 0.0;
    for (int i = threadIdx.x; i < G; i += kBlock) t += partials[(int64_t)r * G + i];
    t = wave_sum(t);
    if (lane == 0) sm[r][w] = t;
  }
  __syncthreads();
#pragma unroll
  for (int r = 0; r < NR; ++r) {
    double t = 0.0;
#pragma unroll
    for (int i = 0; i < kBlock / 64; ++i) t += sm[r][i];
    sums[r] = t;
  }
}

}  // namespace lym
