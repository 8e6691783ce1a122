// Gauge-field kernels off the Wilson operator: unpacking of stored fields and
// the projection of random matrices onto SU(N).
//
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

// ---- projection onto SU(N) ---------------------------------------------------
// unitary_sqrt_projection (smearing.py:18-31): U = X (X^dag X)^-1/2, the
// unitary polar factor of X, divided by det(U)^(1/N) (principal branch).
// The polar factor is computed by Higham's scaled Newton iteration
// U <- (g U + U^-dag / g) / 2, g = (|U^-1|_F / |U|_F)^(1/2), which converges
// quadratically to the same unique factor for any non-singular X.
template <int NC>
__device__ __forceinline__ cpx<double> det_nc(const cpx<double> (&a)[NC][NC]) {
  if constexpr (NC == 2) {
    return cmul(a[0][0], a[1][1]) - cmul(a[0][1], a[1][0]);
  } else {
    return cmul(a[0][0], cmul(a[1][1], a[2][2]) - cmul(a[1][2], a[2][1])) -
           cmul(a[0][1], cmul(a[1][0], a[2][2]) - cmul(a[1][2], a[2][0])) +
           cmul(a[0][2], cmul(a[1][0], a[2][1]) - cmul(a[1][1], a[2][0]));
  }
}

// inv = a^-1 (adjugate / determinant); NC in {2, 3}
template <int NC>
__device__ __forceinline__ void inv_nc(const cpx<double> (&a)[NC][NC], cpx<double> (&inv)[NC][NC]) {
  const cpx<double> d = det_nc<NC>(a);
  const double dn = d.re * d.re + d.im * d.im;
  const cpx<double> r = mk<double>(d.re / dn, -d.im / dn);
  if constexpr (NC == 2) {
    inv[0][0] = cmul(a[1][1], r);
    inv[0][1] = cmul(mk<double>(-a[0][1].re, -a[0][1].im), r);
    inv[1][0] = cmul(mk<double>(-a[1][0].re, -a[1][0].im), r);
    inv[1][1] = cmul(a[0][0], r);
  } else {
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) {
        // cofactor of a[j][i]
        const int r0 = (j + 1) % 3, r1 = (j + 2) % 3, c0 = (i + 1) % 3, c1 = (i + 2) % 3;
        const cpx<double> c = cmul(a[r0][c0], a[r1][c1]) - cmul(a[r0][c1], a[r1][c0]);
        inv[i][j] = cmul(c, r);
      }
  }
}

template <int NC>
__global__ __launch_bounds__(kBlock) void k_unitary_project(const cpx<double>* in, cpx<double>* out, int64_t V) {
  const int64_t s = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  const int mu = blockIdx.y;
  if (s >= V) return;
  cpx<double> u[NC][NC];
  for (int i = 0; i < NC; ++i)
    for (int j = 0; j < NC; ++j) u[i][j] = in[(((int64_t)mu * NC + i) * NC + j) * V + s];
  for (int it = 0; it < 60; ++it) {
    cpx<double> v[NC][NC];
    inv_nc<NC>(u, v);
    double nu = 0.0, nv = 0.0;
    for (int i = 0; i < NC; ++i)
      for (int j = 0; j < NC; ++j) {
        nu += u[i][j].re * u[i][j].re + u[i][j].im * u[i][j].im;
        nv += v[i][j].re * v[i][j].re + v[i][j].im * v[i][j].im;
      }
    // scaling while far from unitary; the plain iteration near convergence
    const double g = it < 8 ? sqrt(sqrt(nv / nu)) : 1.0;
    double diff = 0.0;
    for (int i = 0; i < NC; ++i)
      for (int j = 0; j < NC; ++j) {
        // (U^-dag)_ij = conj(U^-1_ji)
        const cpx<double> n = mk<double>(0.5 * (g * u[i][j].re + v[j][i].re / g),
                                         0.5 * (g * u[i][j].im - v[j][i].im / g));
        const double dr = n.re - u[i][j].re, di = n.im - u[i][j].im;
        diff += dr * dr + di * di;
        u[i][j] = n;
      }
    if (it >= 8 && diff <= 1e-30 * NC) break;
  }
  // divide by det(U)^(1/N), principal branch
  const cpx<double> d = det_nc<NC>(u);
  const double th = atan2(d.im, d.re) / NC;
  const double m = pow(sqrt(d.re * d.re + d.im * d.im), 1.0 / NC);
  const double c = cos(th) / m, sn = -sin(th) / m;  // 1 / (m e^{i th})
  for (int i = 0; i < NC; ++i)
    for (int j = 0; j < NC; ++j)
      out[(((int64_t)mu * NC + i) * NC + j) * V + s] = cmul(u[i][j], mk<double>(c, sn));
}

int unitary_project(int ndim, int nc, int64_t V, const void* in, void* out, hipStream_t st) {
  const dim3 grid((unsigned)((V + kBlock - 1) / kBlock), (unsigned)ndim);
  if (nc == 2)
    hipLaunchKernelGGL(k_unitary_project<2>, grid, dim3(kBlock), 0, st, (const cpx<double>*)in, (cpx<double>*)out, V);
  else if (nc == 3)
    hipLaunchKernelGGL(k_unitary_project<3>, grid, dim3(kBlock), 0, st, (const cpx<double>*)in, (cpx<double>*)out, V);
  else {
    set_error("unitary projection: nc must be 2 or 3");
    return LYM_ERR_ARG;
  }
  LYM_HIP(hipGetLastError());
  return LYM_OK;
}

}  // namespace lym
