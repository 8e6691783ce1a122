// BLAS-1 kernels over flat arrays of complex numbers (complex128 / complex64).
// Reductions accumulate in FP64 with a fixed partition and order, so results
// depend only on n and the data.
#include <hip/hip_runtime.h>

#include <mutex>

#include "lym_internal.h"

namespace lym {

template <typename E>
struct ElemR;
template <>
struct ElemR<double2> {
  using R = double;
};
template <>
struct ElemR<float2> {
  using R = float;
};

// Number of blocks for a grid-stride reduction over n elements: depends on n only.
static int red_grid(int64_t n) {
  int64_t g = (n + kBlock * 4 - 1) / (kBlock * 4);
  if (g > 2048) g = 2048;
  if (g < 1) g = 1;
  return (int)g;
}

template <typename E>
__global__ __launch_bounds__(kBlock) void k_dot(int64_t n, const E* __restrict__ x, const E* __restrict__ y,
                                                RedCtl rc, double* out) {
  double v[2] = {0.0, 0.0};
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock) {
    const E a = x[i], b = y[i];
    v[0] += (double)a.x * b.x + (double)a.y * b.y;
    v[1] += (double)a.x * b.y - (double)a.y * b.x;
  }
  block_partials<2>(v, rc.partials);
  (void)out;
}

// out[0] = sum |a - b|^2, out[1] = sum |a|^2
template <typename E>
__global__ __launch_bounds__(kBlock) void k_diff_norm2(int64_t n, const E* __restrict__ a, const E* __restrict__ b,
                                                       RedCtl rc, double* out) {
  double v[2] = {0.0, 0.0};
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock) {
    const E p = a[i], q = b[i];
    const double dr = (double)p.x - (double)q.x, di = (double)p.y - (double)q.y;
    v[0] += dr * dr + di * di;
    v[1] += (double)p.x * p.x + (double)p.y * p.y;
  }
  block_partials<2>(v, rc.partials);
  (void)out;
}

template <typename E>
__global__ __launch_bounds__(kBlock) void k_axpby(int64_t n, double ar, double ai, const E* __restrict__ x, double br,
                                                  double bi, E* __restrict__ y) {
  using R = typename ElemR<E>::R;
  const R Ar = (R)ar, Ai = (R)ai, Br = (R)br, Bi = (R)bi;
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock) {
    const E a = x[i];
    E b = y[i];
    const R re = Ar * a.x - Ai * a.y + Br * b.x - Bi * b.y;
    const R im = Ar * a.y + Ai * a.x + Br * b.y + Bi * b.x;
    b.x = re;
    b.y = im;
    y[i] = b;
  }
}

template <typename EI, typename EO>
__global__ __launch_bounds__(kBlock) void k_convert(int64_t n, const EI* __restrict__ in, EO* __restrict__ out) {
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock) {
    const EI a = in[i];
    EO b;
    b.x = a.x;
    b.y = a.y;
    out[i] = b;
  }
}

// complex128 links (ndim, nc, nc, V) -> the kernels' packed complex64 layout
// (ndim, U, V) of 16-byte units, U = ceil(nc^2 / 2): unit q of (mu, s) holds
// elements 2q, 2q + 1 of the row-major matrix, zero padding after the last
// element.  One thread per (mu, unit, site).
__global__ __launch_bounds__(kBlock) void k_pack_links32(int nc, int64_t V, int64_t n, const double2* __restrict__ in,
                                                         float4* __restrict__ out) {
  const int nu = (nc * nc + 1) / 2;
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock) {
    const int64_t s = i % V;
    const int64_t pq = i / V;  // mu * nu + q
    const int mu = (int)(pq / nu), q = (int)(pq % nu);
    const int e0 = 2 * q, e1 = 2 * q + 1;
    const double2 a = in[((int64_t)mu * nc * nc + e0) * V + s];
    double2 b = make_double2(0.0, 0.0);
    if (e1 < nc * nc) b = in[((int64_t)mu * nc * nc + e1) * V + s];
    out[i] = make_float4((float)a.x, (float)a.y, (float)b.x, (float)b.y);
  }
}

template <typename E>
__global__ __launch_bounds__(kBlock) void k_gamma5(int ns, int64_t nps, const E* __restrict__ in, E* __restrict__ out) {
  const int64_t n = (int64_t)ns * nps;
  const int64_t half = (int64_t)(ns / 2) * nps;
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock) {
    E a = in[i];
    if (i >= half) {
      a.x = -a.x;
      a.y = -a.y;
    }
    out[i] = a;
  }
}

// out = a_h x + b_h y with (a_h, b_h) = (c[0..1], c[4..5]) on spins < ns/2
// and (c[2..3], c[6..7]) on the others (complex128)
__global__ __launch_bounds__(kBlock) void k_spin_axpby(int ns, int64_t nps, SpinCoef c, const double2* __restrict__ x,
                                                       const double2* __restrict__ y, double2* __restrict__ out) {
  const int64_t n = (int64_t)ns * nps;
  const int64_t half = (int64_t)(ns / 2) * nps;
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock) {
    const int h = i >= half ? 1 : 0;
    const double ar = c.v[2 * h], ai = c.v[2 * h + 1], br = c.v[4 + 2 * h], bi = c.v[5 + 2 * h];
    const double2 a = x[i], b = y[i];
    out[i] = make_double2(ar * a.x - ai * a.y + br * b.x - bi * b.y, ar * a.y + ai * a.x + br * b.y + bi * b.x);
  }
}

static int stream_grid(int64_t n) {
  int64_t g = (n + kBlock - 1) / kBlock;
  if (g > 8192) g = 8192;
  if (g < 1) g = 1;
  return (int)g;
}

int blas_dot(int prec, int64_t n, const void* x, const void* y, double* out_dev, RedCtl rc, hipStream_t st) {
  const int g = red_grid(n);
  if (prec == P64)
    hipLaunchKernelGGL(k_dot<double2>, dim3(g), dim3(kBlock), 0, st, n, (const double2*)x, (const double2*)y, rc,
                       out_dev);
  else if (prec == P32)
    hipLaunchKernelGGL(k_dot<float2>, dim3(g), dim3(kBlock), 0, st, n, (const float2*)x, (const float2*)y, rc,
                       out_dev);
  else {
    set_error("lym_dot: precision must be F64 or F32");
    return LYM_ERR_ARG;
  }
  LYM_HIP(hipGetLastError());
  return launch_finalize(FK_SUM, nullptr, rc.partials, g, 2, out_dev, st);
}

// ---------------------------------------------------------------------------
// Allocation-free dot product (lym_dot_noalloc).  Its scratch, the block
// partials of k_dot (2 per block, at most kNoAllocBlocks blocks: red_grid
// caps the grid there for any n) and nothing else, lives in this file-scope
// device array: 32 KiB that every GPU the library runs on holds until the
// process ends.  In host code `scratch` names a host-side stand-in, not the
// device memory: its device address is looked up once per device with
// hipGetSymbolAddress (lym_init, or the first call on a device) and only that
// address is handed to the kernels.  Every call on a GPU shares that GPU's
// array, so calls that overlap on one GPU race; the kernels overwrite every
// partial they read, so nothing needs clearing between calls.
// ---------------------------------------------------------------------------
constexpr int kNoAllocBlocks = 2048;
static __device__ __align__(256) double scratch[2 * kNoAllocBlocks];

constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static void* g_scratch_dev[kMaxDevices];  // device address of `scratch`, per device

// device address of `scratch` on the current device (looked up once)
int noalloc_scratch(void** out) {
  int dev = 0;
  LYM_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) {
    set_error("lym_dot_noalloc: device index above 63");
    return LYM_ERR_UNSUPPORTED;
  }
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_dev[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) return hip_fail(e != hipSuccess ? e : hipErrorInvalidSymbol, "hipGetSymbolAddress(scratch)");
    g_scratch_dev[dev] = p;
  }
  *out = g_scratch_dev[dev];
  return LYM_OK;
}

int blas_dot_noalloc(int prec, int64_t n, const void* x, const void* y, double* out_dev, hipStream_t st) {
  void* sp = nullptr;
  int rc = noalloc_scratch(&sp);
  if (rc) return rc;
  const int g = red_grid(n);
  if (g > kNoAllocBlocks) {  // red_grid caps the grid: cannot happen
    set_error("lym_dot_noalloc: reduction grid above the static scratch");
    return LYM_ERR_ARG;
  }
  double* partials = (double*)sp;
  const RedCtl red{partials, nullptr};
  if (prec == P64)
    hipLaunchKernelGGL(k_dot<double2>, dim3(g), dim3(kBlock), 0, st, n, (const double2*)x, (const double2*)y, red,
                       out_dev);
  else if (prec == P32)
    hipLaunchKernelGGL(k_dot<float2>, dim3(g), dim3(kBlock), 0, st, n, (const float2*)x, (const float2*)y, red,
                       out_dev);
  else {
    set_error("lym_dot_noalloc: precision must be F64 or F32");
    return LYM_ERR_ARG;
  }
  LYM_HIP(hipGetLastError());
  return launch_finalize(FK_SUM, nullptr, partials, g, 2, out_dev, st);
}

int blas_diff_norm2(int prec, int64_t n, const void* a, const void* b, double* out_dev, RedCtl rc, hipStream_t st) {
  const int g = red_grid(n);
  if (prec == P64)
    hipLaunchKernelGGL(k_diff_norm2<double2>, dim3(g), dim3(kBlock), 0, st, n, (const double2*)a, (const double2*)b,
                       rc, out_dev);
  else
    hipLaunchKernelGGL(k_diff_norm2<float2>, dim3(g), dim3(kBlock), 0, st, n, (const float2*)a, (const float2*)b, rc,
                       out_dev);
  LYM_HIP(hipGetLastError());
  return launch_finalize(FK_SUM, nullptr, rc.partials, g, 2, out_dev, st);
}

int blas_axpby(int prec, int64_t n, double ar, double ai, const void* x, double br, double bi, void* y,
               hipStream_t st) {
  const int g = stream_grid(n);
  if (prec == P64)
    hipLaunchKernelGGL(k_axpby<double2>, dim3(g), dim3(kBlock), 0, st, n, ar, ai, (const double2*)x, br, bi,
                       (double2*)y);
  else if (prec == P32)
    hipLaunchKernelGGL(k_axpby<float2>, dim3(g), dim3(kBlock), 0, st, n, ar, ai, (const float2*)x, br, bi,
                       (float2*)y);
  else {
    set_error("lym_axpby: precision must be F64 or F32");
    return LYM_ERR_ARG;
  }
  LYM_HIP(hipGetLastError());
  return LYM_OK;
}

int blas_convert(int pin, const void* in, int pout, void* out, int64_t n, hipStream_t st) {
  const int g = stream_grid(n);
  if (pin == P64 && pout == P32)
    hipLaunchKernelGGL((k_convert<double2, float2>), dim3(g), dim3(kBlock), 0, st, n, (const double2*)in, (float2*)out);
  else if (pin == P32 && pout == P64)
    hipLaunchKernelGGL((k_convert<float2, double2>), dim3(g), dim3(kBlock), 0, st, n, (const float2*)in, (double2*)out);
  else if (pin == P64 && pout == P64)
    hipLaunchKernelGGL((k_convert<double2, double2>), dim3(g), dim3(kBlock), 0, st, n, (const double2*)in,
                       (double2*)out);
  else if (pin == P32 && pout == P32)
    hipLaunchKernelGGL((k_convert<float2, float2>), dim3(g), dim3(kBlock), 0, st, n, (const float2*)in, (float2*)out);
  else {
    set_error("lym_convert: precisions must be F64 or F32");
    return LYM_ERR_ARG;
  }
  LYM_HIP(hipGetLastError());
  return LYM_OK;
}

int blas_pack_links32(int ndim, int nc, int64_t V, const void* in, void* out, hipStream_t st) {
  const int64_t n = (int64_t)ndim * link_units32(nc) * V;
  hipLaunchKernelGGL(k_pack_links32, dim3(stream_grid(n)), dim3(kBlock), 0, st, nc, V, n, (const double2*)in,
                     (float4*)out);
  LYM_HIP(hipGetLastError());
  return LYM_OK;
}

int blas_gamma5(int prec, int ns, int64_t nps, const void* in, void* out, hipStream_t st) {
  const int g = stream_grid((int64_t)ns * nps);
  if (prec == P64)
    hipLaunchKernelGGL(k_gamma5<double2>, dim3(g), dim3(kBlock), 0, st, ns, nps, (const double2*)in, (double2*)out);
  else if (prec == P32)
    hipLaunchKernelGGL(k_gamma5<float2>, dim3(g), dim3(kBlock), 0, st, ns, nps, (const float2*)in, (float2*)out);
  else {
    set_error("lym_gamma5: precision must be F64 or F32");
    return LYM_ERR_ARG;
  }
  LYM_HIP(hipGetLastError());
  return LYM_OK;
}

int blas_spin_axpby(int ns, int64_t nps, const SpinCoef& c, const void* x, const void* y, void* out,
                    hipStream_t st) {
  const int g = stream_grid((int64_t)ns * nps);
  hipLaunchKernelGGL(k_spin_axpby, dim3(g), dim3(kBlock), 0, st, ns, nps, c, (const double2*)x, (const double2*)y,
                     (double2*)out);
  LYM_HIP(hipGetLastError());
  return LYM_OK;
}

// ---- special-unitarity test of the links ----------------------------------
__global__ void k_zero_u64(unsigned long long* p) { *p = 0ull; }

// out = max over (mu, site) and entries of |stored last row - rebuilt last row|
// (the rebuild LinkView uses with recon = 1).  The maximum of non-negative
// doubles is an atomicMax on their bit patterns: exact and order independent.
template <typename R, int NC>
__global__ __launch_bounds__(kBlock) void k_link_dev(const cpx<R>* __restrict__ links, int64_t V, int ndim,
                                                     unsigned long long* out) {
  double m = 0.0;
  const LinkView<R> full{links, V, 0};
  const LinkView<R> su{links, V, 1};
  const int64_t n = (int64_t)ndim * V;
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock) {
    const int mu = (int)(i / V);
    const int64_t s = i - (int64_t)mu * V;
    cpx<R> a[NC][NC], b[NC][NC];
    full.template load<NC>(mu, s, a);
    su.template load<NC>(mu, s, b);
#pragma unroll
    for (int j = 0; j < NC; ++j) {
      const double dr = fabs((double)a[NC - 1][j].re - (double)b[NC - 1][j].re);
      const double di = fabs((double)a[NC - 1][j].im - (double)b[NC - 1][j].im);
      // a NaN difference (non-finite entries) fails the test
      if (!(dr <= m)) m = (dr == dr) ? dr : INFINITY;
      if (!(di <= m)) m = (di == di) ? di : INFINITY;
    }
  }
  for (int off = 32; off > 0; off >>= 1) m = fmax(m, __shfl_xor(m, off, 64));
  __shared__ double sm[kBlock / 64];
  if ((threadIdx.x & 63) == 0) sm[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < kBlock / 64; ++w) m = fmax(m, sm[w]);
    atomicMax(out, (unsigned long long)__double_as_longlong(m));
  }
}

int link_deviation(const OpDesc& d, int prec, double* out_dev, hipStream_t st) {
  const void* l = prec == P64 ? d.links64 : d.links32;
  if (!l) {
    set_error("lym_link_deviation: the operator has no links at this precision");
    return LYM_ERR_ARG;
  }
  unsigned long long* o = reinterpret_cast<unsigned long long*>(out_dev);
  hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(1), 0, st, o);
  const int g = red_grid((int64_t)d.ndim * d.lat.V);
  if (prec == P64) {
    if (d.nc == 3)
      hipLaunchKernelGGL((k_link_dev<double, 3>), dim3(g), dim3(kBlock), 0, st, (const cpx<double>*)l, d.lat.V,
                         d.ndim, o);
    else
      hipLaunchKernelGGL((k_link_dev<double, 2>), dim3(g), dim3(kBlock), 0, st, (const cpx<double>*)l, d.lat.V,
                         d.ndim, o);
  } else {
    if (d.nc == 3)
      hipLaunchKernelGGL((k_link_dev<float, 3>), dim3(g), dim3(kBlock), 0, st, (const cpx<float>*)l, d.lat.V,
                         d.ndim, o);
    else
      hipLaunchKernelGGL((k_link_dev<float, 2>), dim3(g), dim3(kBlock), 0, st, (const cpx<float>*)l, d.lat.V,
                         d.ndim, o);
  }
  LYM_HIP(hipGetLastError());
  return LYM_OK;
}

}  // namespace lym
