// Unpacking of gauge fields stored as U* (the complex conjugate of U) with
// two columns per SU(3) link in XDR floats (big-endian IEEE single), the
// format read_config reads (lattice.py:10-70): words in the order
// (site, mu, row, col < 2, re/im).  Per link: conjugate, normalise column 0,
// orthogonalise column 1 against it and normalise, column 2 = conj(col0 x
// col1) -- an exact SU(3) matrix -- written to (ndim, 3, 3, V) complex128.
// One thread per (site, mu); blockIdx.y = mu, so the link planes are written
// with consecutive sites per wave.
#include "lym_internal.h"

namespace lym {

__device__ __forceinline__ double xdr_float(const uint32_t* w, int64_t i) {
  return (double)__uint_as_float(__builtin_bswap32(w[i]));
}

__global__ __launch_bounds__(kBlock) void k_unpack_ustar(const uint32_t* raw, int64_t V, int ndim,
                                                         cpx<double>* out) {
  const int64_t s = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  const int mu = blockIdx.y;
  if (s >= V) return;
  const int64_t base = (s * ndim + mu) * 12;  // 3 rows x 2 columns x (re, im)
  cpx<double> c0[3], c1[3], c2[3];
  for (int i = 0; i < 3; ++i) {
    // U* stored: U = re - i im
    c0[i] = mk<double>(xdr_float(raw, base + i * 4 + 0), -xdr_float(raw, base + i * 4 + 1));
    c1[i] = mk<double>(xdr_float(raw, base + i * 4 + 2), -xdr_float(raw, base + i * 4 + 3));
  }
  double n0 = 0.0;
  for (int i = 0; i < 3; ++i) n0 += c0[i].re * c0[i].re + c0[i].im * c0[i].im;
  const double r0 = 1.0 / sqrt(n0);
  for (int i = 0; i < 3; ++i) c0[i] = cscale(r0, c0[i]);
  // c1 -= <c0, c1> c0 with <c0, c1> = sum c1 conj(c0)
  cpx<double> p = czero<double>();
  for (int i = 0; i < 3; ++i) cfma_c(p, c1[i], c0[i]);
  for (int i = 0; i < 3; ++i) c1[i] -= cmul(p, c0[i]);
  double n1 = 0.0;
  for (int i = 0; i < 3; ++i) n1 += c1[i].re * c1[i].re + c1[i].im * c1[i].im;
  const double r1 = 1.0 / sqrt(n1);
  for (int i = 0; i < 3; ++i) c1[i] = cscale(r1, c1[i]);
  // c2_i = conj(c0_j c1_k - c0_k c1_j), (i, j, k) cyclic
  for (int i = 0; i < 3; ++i) {
    const int j = (i + 1) % 3, k = (i + 2) % 3;
    const cpx<double> t = cmul(c0[j], c1[k]) - cmul(c0[k], c1[j]);
    c2[i] = mk<double>(t.re, -t.im);
  }
  for (int i = 0; i < 3; ++i) {
    const int64_t pl = ((int64_t)mu * 3 + i) * 3;
    out[(pl + 0) * V + s] = c0[i];
    out[(pl + 1) * V + s] = c1[i];
    out[(pl + 2) * V + s] = c2[i];
  }
}

int unpack_ustar(const void* raw, int64_t V, int ndim, void* links, hipStream_t st) {
  const dim3 grid((unsigned)((V + kBlock - 1) / kBlock), (unsigned)ndim);
  hipLaunchKernelGGL(k_unpack_ustar, grid, dim3(kBlock), 0, st, (const uint32_t*)raw, V, ndim, (cpx<double>*)links);
  LYM_HIP(hipGetLastError());
  return LYM_OK;
}

}  // namespace lym
