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
// The fundamental sum factorises around the link: with y_n = Y(x+mu)_n,
//   W = U P - Q U^dag,  P = sum_n y_n (x) conj(A_n),  Q = sum_n B_n (x) conj(y_n)
// (times the boundary factor bc on U and conj(bc) on U^dag), so the pair loop
// only accumulates P and Q and the link is applied once per (site, mu).  When
// hermitian, P+^dag = P- (g5 anticommutes with g_mu), so A = B, Q = P^dag and
// W = M - M^dag with M = bc U P.
//
// Spin structure (chiral basis, one entry i^k per row of A_mu, see Gam<D>):
// row r < Ns/2 pairs with lo = Ns/2 + perm_r; with g5l the gamma5 sign of the
// lower rows (-1 hermitian, +1 otherwise) and sA = -g5l:
//   A_r  = X_r + sA i^k X_lo          B_r  = X_r + i^k X_lo
//   A_lo = g5l (X_lo + sA i^-k X_r)   B_lo = g5l (X_lo + i^-k X_r)
#pragma once
#include "wilson_site.h"

namespace lym {

// Adjoint: acc += [v, a^dag] - [b, v^dag] for one spin (row-major NC x NC
// matrices), i.e. acc_ij += sum_k v_ik conj(a_jk) - conj(a_ki) v_kj
//                                - b_ik conj(v_jk) + conj(v_ki) b_kj
template <int NC>
__host__ __device__ __forceinline__ void force_accumulate_adj(const cpx<double>* v, const cpx<double>* a,
                                                              const cpx<double>* b, cpx<double> (&acc)[NC][NC]) {
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

// i^k a for a k that varies across the lanes of a wave: selects, no branch
__host__ __device__ __forceinline__ cpx<double> ipow_sel(int k, cpx<double> a) {
  const cpx<double> t = (k & 1) ? mk<double>(-a.im, a.re) : a;
  return (k & 2) ? mk<double>(-t.re, -t.im) : t;
}

// Component accessors of one spinor at one site: (plane) -> value.
// GlobalAcc reads device memory (buffer loads); `order(dep)` routes its site
// index through an empty asm that consumes dep, so its loads issue after dep
// is produced.  LdsAcc reads a staged tile (plane-major, nsite per plane);
// MixAcc picks one of the two per lane.
struct GlobalAcc {
  const cpx<double>* p;
  uint32_t bytes;
  int64_t V;
  int64_t s;
  __host__ __device__ __forceinline__ cpx<double> operator()(int plane) const {
    return ld<cpx<double>>(p, bytes, plane, V, s);
  }
  __host__ __device__ __forceinline__ void order(double dep) {
    int64_t t = 0;
    hop_dep<double>(s, t, dep);
  }
};
struct LdsAcc {
  const cpx<double>* base;
  int nsite;
  int sl;
  __host__ __device__ __forceinline__ cpx<double> operator()(int plane) const { return base[plane * nsite + sl]; }
  __host__ __device__ __forceinline__ void order(double) {}
};
struct MixAcc {
  LdsAcc l;
  GlobalAcc g;
  bool lds;
  __host__ __device__ __forceinline__ cpx<double> operator()(int plane) const { return lds ? l(plane) : g(plane); }
  __host__ __device__ __forceinline__ void order(double dep) { g.order(dep); }
};

// Fundamental representation: P += w sum_n y_n (x) conj(A_n) and (not
// hermitian) Q += w sum_n B_n (x) conj(y_n) for the term of L(X, Y) at one
// site, xa = X(x), ya = Y(x+mu).  The loads of each spin pair are ordered
// after the previous pair's accumulation (hop_commit / order), which keeps
// one pair's operands live at a time; other waves hide the latency.
template <int D, int NC, bool HERM, class XA, class YA>
__host__ __device__ __forceinline__ void force_term_acc(int mu, XA xa, YA ya, double w, cpx<double> (&P)[NC][NC],
                                                        cpx<double> (&Q)[NC][NC]) {
  constexpr int NS = Gam<D>::NS;
  constexpr int H = NS / 2;
  // hermitian: g5l = -1, sA = +1 (A = B); otherwise g5l = +1, sA = -1
  constexpr double g5l = HERM ? -1.0 : 1.0;
  constexpr double sA = -g5l;
#pragma unroll
  for (int r = 0; r < H; ++r) {
    // the previous spin pair's accumulation completes before these loads
    hop_commit<NC * NC, double>(reinterpret_cast<cpx<double>(&)[NC * NC]>(P));
    if constexpr (!HERM) hop_commit<NC * NC, double>(reinterpret_cast<cpx<double>(&)[NC * NC]>(Q));
    xa.order(P[NC - 1][NC - 1].re);
    ya.order(P[NC - 1][NC - 1].re);
    const int lo = H + Gam<D>::perm(mu, r);
    const int k = Gam<D>::k(mu, r);
    cpx<double> xr[NC], xl[NC], yr[NC], yl[NC];
#pragma unroll
    for (int c = 0; c < NC; ++c) {
      xr[c] = xa(r * NC + c);
      xl[c] = xa(lo * NC + c);
      yr[c] = ya(r * NC + c);
      yl[c] = ya(lo * NC + c);
    }
    // spin r:  a = w (xr + sA i^k xl),  b = w (xr + i^k xl)
    // spin lo: a = w g5l (xl + sA i^-k xr),  b = w g5l (xl + i^-k xr)
#pragma unroll
    for (int t = 0; t < NC; ++t) {
      const cpx<double> tr = ipow_sel(k, xl[t]);
      const cpx<double> tl = ipow_sel(4 - k, xr[t]);
      const cpx<double> ar = mk<double>(w * (xr[t].re + sA * tr.re), w * (xr[t].im + sA * tr.im));
      const double f = w * g5l;
      const cpx<double> al = mk<double>(f * (xl[t].re + sA * tl.re), f * (xl[t].im + sA * tl.im));
#pragma unroll
      for (int u = 0; u < NC; ++u) {
        cfma_c(P[u][t], yr[u], ar);
        cfma_c(P[u][t], yl[u], al);
      }
      if constexpr (!HERM) {
        const cpx<double> br = mk<double>(w * (xr[t].re + tr.re), w * (xr[t].im + tr.im));
        const cpx<double> bl = mk<double>(f * (xl[t].re + tl.re), f * (xl[t].im + tl.im));
#pragma unroll
        for (int u = 0; u < NC; ++u) {
          cfma_c(Q[t][u], br, yr[u]);
          cfma_c(Q[t][u], bl, yl[u]);
        }
      }
    }
  }
}

// the same with X(x) and Y(x+mu) read from device memory at s and nb
template <int D, int NC, bool HERM>
__host__ __device__ __forceinline__ void force_term_fund(const Lat& lat, int mu, const cpx<double>* X,
                                                         const cpx<double>* Y, uint32_t bytes, int64_t s, int64_t nb,
                                                         double w, cpx<double> (&P)[NC][NC],
                                                         cpx<double> (&Q)[NC][NC]) {
  force_term_acc<D, NC, HERM>(mu, GlobalAcc{X, bytes, lat.V, s}, GlobalAcc{Y, bytes, lat.V, nb}, w, P, Q);
}

// W = bc U P - conj(bc) Q U^dag  (Q = P^dag when hermitian), in place of P
template <int NC, bool HERM>
__host__ __device__ __forceinline__ void force_close_fund(const cpx<double> (&u)[NC][NC], bool edge,
                                                          const cpx<double> bcf, cpx<double> (&P)[NC][NC],
                                                          const cpx<double> (&Q)[NC][NC]) {
  cpx<double> M[NC][NC];
#pragma unroll
  for (int i = 0; i < NC; ++i)
#pragma unroll
    for (int j = 0; j < NC; ++j) {
      cpx<double> a = czero<double>();
#pragma unroll
      for (int k = 0; k < NC; ++k) cfma(a, u[i][k], P[k][j]);
      M[i][j] = edge ? cmul(bcf, a) : a;
    }
  if constexpr (HERM) {
#pragma unroll
    for (int i = 0; i < NC; ++i)
#pragma unroll
      for (int j = 0; j < NC; ++j)
        P[i][j] = mk<double>(M[i][j].re - M[j][i].re, M[i][j].im + M[j][i].im);
  } else {
#pragma unroll
    for (int i = 0; i < NC; ++i)
#pragma unroll
      for (int j = 0; j < NC; ++j) {
        cpx<double> a = czero<double>();
#pragma unroll
        for (int k = 0; k < NC; ++k) cfma_c(a, Q[i][k], u[j][k]);  // (Q U^dag)_ij
        if (edge) a = cmul_c(a, bcf);
        P[i][j] = M[i][j] - a;
      }
  }
}

// Adjoint representation: acc += w * [term of L(X, Y)] for direction mu at
// site s; nb = site of x+mu, edge: the hop crosses the boundary carrying the
// factor (bc_re, bc_im).
// X, Y: (Ns, NCC, V) complex128 planes, `bytes` their size.
template <int D, int NC>
__host__ __device__ __forceinline__ void force_term_adj(const Lat& lat, int mu, bool edge, const cpx<double>* X,
                                                        const cpx<double>* Y, uint32_t bytes, int64_t s, int64_t nb,
                                                        double w, const cpx<double> (&u)[NC][NC],
                                                        const cpx<double> (&ur)[NC][NC],
                                                        cpx<double> (&acc)[NC][NC]) {
  constexpr int NS = Gam<D>::NS;
  constexpr int H = NS / 2;
  constexpr int NCC = NC * NC;
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
      transport<NC, true, true, double>(u, ur, y, v);
      if (edge) {
#pragma unroll
        for (int c = 0; c < NCC; ++c) v[c] = cmul(bcf, v[c]);
      }
      force_accumulate_adj<NC>(v, a, b, acc);
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
// HERM: lat.herm (compile-time for the fundamental form).
template <int D, int NC, bool ADJ, bool HERM>
__host__ __device__ __forceinline__ void force_site(const Lat& lat, int mu, int64_t s, const int (&x)[D],
                                                    const LinkView<double>& lk, const LinkView<double>& lkr,
                                                    const ForcePairs& fp, uint32_t vbytes, cpx<double>* out,
                                                    uint32_t obytes) {
  const int L = lat.L[mu];
  const int64_t st = lat.stride[mu];
  const int64_t nb = (x[mu] == L - 1) ? s - (int64_t)(L - 1) * st : s + st;
  const bool edge = lat.has_bc && mu == D - 1 && x[mu] == L - 1;
  cpx<double> acc[NC][NC];
#pragma unroll
  for (int i = 0; i < NC; ++i)
#pragma unroll
    for (int j = 0; j < NC; ++j) acc[i][j] = czero<double>();
  if constexpr (!ADJ) {
    cpx<double> Q[NC][NC];
#pragma unroll
    for (int i = 0; i < NC; ++i)
#pragma unroll
      for (int j = 0; j < NC; ++j) Q[i][j] = czero<double>();
    for (int p = 0; p < fp.n; ++p) {
      force_term_fund<D, NC, HERM>(lat, mu, fp.x[p], fp.y[p], vbytes, s, nb, fp.w[p], acc, Q);
      if (fp.sym) force_term_fund<D, NC, HERM>(lat, mu, fp.y[p], fp.x[p], vbytes, s, nb, fp.w[p], acc, Q);
    }
    cpx<double> u[NC][NC];
    lk.template load<NC>(mu, s, u);
    force_close_fund<NC, HERM>(u, edge, mk<double>(lat.bc_re, lat.bc_im), acc, Q);
  } else {
    cpx<double> u[NC][NC], ur[NC][NC];
    lk.template load<NC>(mu, s, u);
    lkr.template load<NC>(mu, s, ur);
    for (int p = 0; p < fp.n; ++p) {
      force_term_adj<D, NC>(lat, mu, edge, fp.x[p], fp.y[p], vbytes, s, nb, fp.w[p], u, ur, acc);
      if (fp.sym) force_term_adj<D, NC>(lat, mu, edge, fp.y[p], fp.x[p], vbytes, s, nb, fp.w[p], u, ur, acc);
    }
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
