// Per-slice sums of propagator densities: the reduction pion_correlator makes
// (fermion_measurements.py:73-84),
//   corr[z] = sum over solutions v, sites x with x_dir = z, of
//             sum_{s,c} |v_sc(x)|^2                  (fundamental)
//             sum_{s,c,d} conj(v_scd(x)) v_sdc(x)      (adjoint matrices)
// Two deterministic passes: k_site_density adds the density of every
// solution per site (streaming, one thread per site, a fixed order), then
// k_slice_reduce sums one slice per block in a fixed order and tree.
#include "lym_internal.h"

namespace lym {

template <bool ADJ>
__global__ __launch_bounds__(kBlock) void k_site_density(int64_t V, int ns, int nc, SliceVecs sv, uint32_t bytes,
                                                         double* dens) {
  const int64_t s = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (s >= V) return;
  double acc = sv.accumulate ? dens[s] : 0.0;
  for (int p = 0; p < sv.n; ++p) {
    const cpx<double>* v = sv.v[p];
    for (int i = 0; i < ns; ++i) {
      if constexpr (!ADJ) {
        for (int c = 0; c < nc; ++c) {
          const cpx<double> a = ld<cpx<double>>(v, bytes, (int64_t)i * nc + c, V, s);
          acc += a.re * a.re + a.im * a.im;
        }
      } else {
        for (int c = 0; c < nc; ++c)
          for (int d = 0; d < nc; ++d) {
            const cpx<double> a = ld<cpx<double>>(v, bytes, ((int64_t)i * nc + c) * nc + d, V, s);
            const cpx<double> b = ld<cpx<double>>(v, bytes, ((int64_t)i * nc + d) * nc + c, V, s);
            acc += a.re * b.re + a.im * b.im;  // Re conj(a) b (the sum is real)
          }
      }
    }
  }
  dens[s] = acc;
}

// block z: the V / L sites with x_dir = z, s = (outer * L + z) * inner + in
__global__ __launch_bounds__(kBlock) void k_slice_reduce(const double* dens, int64_t nslice, int64_t inner, int L,
                                                         double* out) {
  __shared__ double part[kBlock / 64];
  const int z = blockIdx.x;
  double a = 0.0;
  for (int64_t q = threadIdx.x; q < nslice; q += kBlock) {
    const int64_t outer = q / inner;
    const int64_t in = q - outer * inner;
    a += dens[(outer * L + z) * inner + in];
  }
  a = wave_sum(a);
  if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = a;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int w = 0; w < kBlock / 64; ++w) t += part[w];
    out[z] = t;
  }
}

int slice_density(const OpDesc& d, const SliceVecs& sv, double* dens, hipStream_t st) {
  const int64_t V = d.lat.V;
  const uint32_t bytes = (uint32_t)(field_elems(d) * 16);
  const int gs = (int)((V + kBlock - 1) / kBlock);
  if (d.adj)
    hipLaunchKernelGGL(k_site_density<true>, dim3(gs), dim3(kBlock), 0, st, V, d.ns, d.nc, sv, bytes, dens);
  else
    hipLaunchKernelGGL(k_site_density<false>, dim3(gs), dim3(kBlock), 0, st, V, d.ns, d.nc, sv, bytes, dens);
  LYM_HIP(hipGetLastError());
  return LYM_OK;
}

int slice_reduce(const OpDesc& d, int dir, const double* dens, double* out_dev, hipStream_t st) {
  const int L = d.lat.L[dir];
  const int64_t inner = d.lat.stride[dir];
  hipLaunchKernelGGL(k_slice_reduce, dim3(L), dim3(kBlock), 0, st, dens, d.lat.V / L, inner, L, out_dev);
  LYM_HIP(hipGetLastError());
  return LYM_OK;
}

}  // namespace lym
