// Per-site fermion force: the Lie derivative of the Wilson operator
// (dirac.py:525-565), host and device.
//
// The reference contracts, for every direction mu and generator T^k,
//   l_k = X^dag P+ (T^k U_mu) Y(x+mu) - Y(x+mu)^dag P- (U_mu^dag T^k) X
// (P+ = [g5](1 - g_mu), P- = [g5](1 + g_mu), U after the boundary factor) and
// returns kappa sum_k T^k_fund l_k.  Summing over the generators in closed
// form (completeness of the T^k, tr T^a T^b = delta_ab / 2):
//   fundamental: out = kappa/2 (W - tr(W)/Nc),
//                W = sum_n [ v_n (x) conj(A_n) - B_n (x) conj(v_n) ],  v_n = U Y(x+mu)_n
//   adjoint (matrix form X = x^a T^a, links U X Ur^dag):
//                out = kappa sum_n ( [Z_n, A_n^dag] - [B_n, Z_n^dag] ),  Z_n = U Y(x+mu)_n Ur^dag
// with A = P+^dag X and B = P- X, spin-local combinations of X(x).  The
// adjoint form equals the reference's contraction over the real adjoint
// links V_ab with x^a = 2 tr(T^a X): the commutators drop the trace parts of
// X and Y exactly as that basis change does.
//
// Spin structure (chiral basis, one entry i^k per row of A_mu, see Gam<D>):
// row r < Ns/2 pairs with lo = Ns/2 + perm_r; with g5l the gamma5 sign of the
// lower rows (-1 hermitian, +1 otherwise) and sA = -g5l:
//   A_r  = X_r + sA i^k X_lo          B_r  = X_r + i^k X_lo
//   A_lo = g5l (X_lo + sA i^-k X_r)   B_lo = g5l (X_lo + i^-k X_r)
#pragma once
#include "wilson_site.h"

namespace lym {

// acc += contribution of one spin: colour vectors (fundamental) or row-major
// NC x NC matrices (adjoint)
template <int NC, bool ADJ>
__host__ __device__ __forceinline__ void force_accumulate(const cpx<double>* v, const cpx<double>* a,
                                                          const cpx<double>* b, cpx<double> (&acc)[NC][NC]) {
  if constexpr (!ADJ) {
    // acc_ut += v_u conj(a_t) - b_u conj(v_t)
#pragma unroll
    for (int u = 0; u < NC; ++u)
#pragma unroll
      for (int t = 0; t < NC; ++t) {
        cfma_c(acc[u][t], v[u], a[t]);
        acc[u][t] -= cmul_c(b[u], v[t]);
      }
  } else {
    // acc_ij += sum_k v_ik conj(a_jk) - conj(a_ki) v_kj - b_ik conj(v_jk) + conj(v_ki) b_kj
#pragma unroll
    for (int i = 0; i < NC; ++i)
#pragma unroll
      for (int j = 0; j < NC; ++j) {
        cpx<double> p = czero<double>(), m = czero<double>();
#pragma unroll
        for (int k = 0; k < NC; ++k) {
          cfma_c(p, v[i * NC + k], a[j * NC + k]);
          cfmac(p, v[k * NC + i], b[k * NC + j]);
          cfmac(m, a[k * NC + i], v[k * NC + j]);
          cfma_c(m, b[i * NC + k], v[j * NC + k]);
        }
        acc[i][j] += p - m;
      }
  }
}

// acc += w * [term of L(X, Y)] for direction mu at site s; nb = site of x+mu,
// edge: the hop crosses the boundary carrying the factor (bc_re, bc_im).
// X, Y: (Ns, NCC, V) complex128 planes, `bytes` their size.
template <int D, int NC, bool ADJ>
__host__ __device__ __forceinline__ void force_term(const Lat& lat, int mu, bool edge, const cpx<double>* X,
                                                    const cpx<double>* Y, uint32_t bytes, int64_t s, int64_t nb,
                                                    double w, const cpx<double> (&u)[NC][NC],
                                                    const cpx<double> (&ur)[NC][NC], cpx<double> (&acc)[NC][NC]) {
  constexpr int NS = Gam<D>::NS;
  constexpr int H = NS / 2;
  constexpr int NCC = ADJ ? NC * NC : NC;
  const double g5l = lat.herm ? -1.0 : 1.0;
  const double sA = -g5l;
  const cpx<double> bcf = mk<double>(lat.bc_re, lat.bc_im);
#pragma unroll
  for (int r = 0; r < H; ++r) {
    const int lo = H + Gam<D>::perm(mu, r);
    const int k = Gam<D>::k(mu, r);
    cpx<double> xr[NCC], xl[NCC];
#pragma unroll
    for (int c = 0; c < NCC; ++c) {
      xr[c] = ld<cpx<double>>(X, bytes, r * NCC + c, lat.V, s);
      xl[c] = ld<cpx<double>>(X, bytes, lo * NCC + c, lat.V, s);
    }
#pragma unroll
    for (int half = 0; half < 2; ++half) {
      const int n = half ? lo : r;
      cpx<double> a[NCC], b[NCC], y[NCC], v[NCC];
#pragma unroll
      for (int c = 0; c < NCC; ++c) {
        if (half == 0) {
          const cpx<double> t = ipow_rt<double>(k, xl[c]);
          a[c] = mk<double>(w * (xr[c].re + sA * t.re), w * (xr[c].im + sA * t.im));
          b[c] = mk<double>(w * (xr[c].re + t.re), w * (xr[c].im + t.im));
        } else {
          const cpx<double> t = ipow_rt<double>(4 - k, xr[c]);
          const double f = w * g5l;
          a[c] = mk<double>(f * (xl[c].re + sA * t.re), f * (xl[c].im + sA * t.im));
          b[c] = mk<double>(f * (xl[c].re + t.re), f * (xl[c].im + t.im));
        }
        y[c] = ld<cpx<double>>(Y, bytes, n * NCC + c, lat.V, nb);
      }
      transport<NC, ADJ, true, double>(u, ur, y, v);
      if (edge) {
#pragma unroll
        for (int c = 0; c < NCC; ++c) v[c] = cmul(bcf, v[c]);
      }
      force_accumulate<NC, ADJ>(v, a, b, acc);
    }
  }
}

// out value of one direction from the accumulated sum
template <int NC, bool ADJ>
__host__ __device__ __forceinline__ void force_finish(const Lat& lat, const cpx<double> (&acc)[NC][NC],
                                                      cpx<double> (&out)[NC][NC]) {
  const double kap = lat.kappa;
  if constexpr (!ADJ) {
    cpx<double> tr = czero<double>();
#pragma unroll
    for (int i = 0; i < NC; ++i) tr += acc[i][i];
    const double h = 0.5 * kap;
#pragma unroll
    for (int i = 0; i < NC; ++i)
#pragma unroll
      for (int j = 0; j < NC; ++j) {
        cpx<double> t = acc[i][j];
        if (i == j) t -= cscale(1.0 / NC, tr);
        out[i][j] = cscale(h, t);
      }
  } else {
#pragma unroll
    for (int i = 0; i < NC; ++i)
#pragma unroll
      for (int j = 0; j < NC; ++j) out[i][j] = cscale(kap, acc[i][j]);
  }
}

// Pairs (X_p, Y_p) with real weights w_p of one force launch:
//   out (+)= sum_p w_p [ L(X_p, Y_p) + sym * L(Y_p, X_p) ]
constexpr int kForceMaxPairs = 32;
struct ForcePairs {
  const cpx<double>* x[kForceMaxPairs];
  const cpx<double>* y[kForceMaxPairs];
  double w[kForceMaxPairs];
  int n;
  int sym;
  int accumulate;
};

// The force of direction mu at site s (coordinates x): reads the links of
// (mu, s) once and loops over the pairs; writes out planes (mu, i, j).
template <int D, int NC, bool ADJ>
__host__ __device__ __forceinline__ void force_site(const Lat& lat, int mu, int64_t s, const int (&x)[D],
                                                    const LinkView<double>& lk, const LinkView<double>& lkr,
                                                    const ForcePairs& fp, uint32_t vbytes, cpx<double>* out,
                                                    uint32_t obytes) {
  const int L = lat.L[mu];
  const int64_t st = lat.stride[mu];
  const int64_t nb = (x[mu] == L - 1) ? s - (int64_t)(L - 1) * st : s + st;
  const bool edge = lat.has_bc && mu == D - 1 && x[mu] == L - 1;
  cpx<double> u[NC][NC], ur[NC][NC];
  lk.template load<NC>(mu, s, u);
  if constexpr (ADJ) lkr.template load<NC>(mu, s, ur);
  cpx<double> acc[NC][NC];
#pragma unroll
  for (int i = 0; i < NC; ++i)
#pragma unroll
    for (int j = 0; j < NC; ++j) acc[i][j] = czero<double>();
  for (int p = 0; p < fp.n; ++p) {
    force_term<D, NC, ADJ>(lat, mu, edge, fp.x[p], fp.y[p], vbytes, s, nb, fp.w[p], u, ur, acc);
    if (fp.sym) force_term<D, NC, ADJ>(lat, mu, edge, fp.y[p], fp.x[p], vbytes, s, nb, fp.w[p], u, ur, acc);
  }
  cpx<double> res[NC][NC];
  force_finish<NC, ADJ>(lat, acc, res);
#pragma unroll
  for (int i = 0; i < NC; ++i)
#pragma unroll
    for (int j = 0; j < NC; ++j) {
      const int64_t pl = ((int64_t)mu * NC + i) * NC + j;
      cpx<double> v = res[i][j];
      if (fp.accumulate) v += ld<cpx<double>>(out, obytes, pl, lat.V, s);
      put<cpx<double>>(out, obytes, pl, lat.V, s, v);
    }
}

}  // namespace lym
