// Deterministic FP64 reductions for CDNA (64-lane wavefronts).
//
// wave_sum: DPP row butterflies (quad_perm, row_ror) inside each 16-lane row,
// then two cross-row exchanges; every lane ends with the wave total.
// block_reduce_last: per-block partials written in block order, an atomic
// ticket elects the last block, which sums the partials in a fixed order; the
// result never depends on which block finished last or on launch timing.
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
  double* partials;  // [NR][gridDim.x]
  unsigned* ticket;  // zero between launches
};

__device__ __forceinline__ double load_coherent(const double* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Reduces v[NR] over the grid.  Returns true in the last block, where
// sums[NR] (shared memory) holds the grid totals for every thread.
template <int NR>
__device__ __forceinline__ bool block_reduce_last(double (&v)[NR], const RedCtl& rc, double (&sums)[NR]) {
  __shared__ double sm[NR][kBlock / 64];
  __shared__ int is_last;
  const int lane = threadIdx.x & 63;
  const int w = threadIdx.x >> 6;
  constexpr int NW = kBlock / 64;
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
#pragma unroll
      for (int i = 0; i < NW; ++i) t += sm[r][i];
      rc.partials[(int64_t)r * gridDim.x + blockIdx.x] = t;
    }
    __threadfence();
    const unsigned prev = atomicAdd(rc.ticket, 1u);
    is_last = (prev == gridDim.x - 1) ? 1 : 0;
  }
  __syncthreads();
  if (!is_last) return false;
  __threadfence();
#pragma unroll
  for (int r = 0; r < NR; ++r) {
    double t = 0.0;
    for (int i = threadIdx.x; i < (int)gridDim.x; i += kBlock) t += load_coherent(&rc.partials[(int64_t)r * gridDim.x + i]);
    t = wave_sum(t);
    __syncthreads();
    if (lane == 0) sm[r][w] = t;
  }
  __syncthreads();
#pragma unroll
  for (int r = 0; r < NR; ++r) {
    double t = 0.0;
#pragma unroll
    for (int i = 0; i < NW; ++i) t += sm[r][i];
    sums[r] = t;
  }
  if (threadIdx.x == 0) atomicExch(rc.ticket, 0u);
  return true;
}

// Ticket only (no values): true in the block that finishes last.
__device__ __forceinline__ bool block_is_last(unsigned* ticket) {
  __shared__ int is_last;
  __syncthreads();
  if (threadIdx.x == 0) {
    __threadfence();
    const unsigned prev = atomicAdd(ticket, 1u);
    is_last = (prev == gridDim.x - 1) ? 1 : 0;
    if (is_last) atomicExch(ticket, 0u);
  }
  __syncthreads();
  return is_last != 0;
}

}  // namespace lym
