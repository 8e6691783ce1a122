// Per-site Wilson hopping arithmetic (host and device).
//
//   D psi(x) = psi(x) + kappa sum_mu [ (1 - g_mu) U_mu(x) psi(x+mu)
//                                   + (1 + g_mu) U_mu(x-mu)^dag psi(x-mu) ]
//   Q = g5 D   (hermitian form)
//
// follows dirac.py:458-523 (fundamental) and dirac.py:380-456 (adjoint matrix
// form U X U^dag).  The gamma matrices are the reference's chiral basis
// (dirac.py:283-297): g_mu = [[0, A_mu], [A_mu^dag, 0]] in (Ns/2 x Ns/2)
// blocks, where every A_mu has exactly one entry i^k per row.  Hence
//   upper rows of (1 -/+ g_mu) psi :  h_r = psi_r -/+ i^k_r psi_{Ns/2+perm_r}
//   lower rows                     :  psi_{Ns/2+perm_r} part = -/+ i^-k_r h_r
// so the projection and the reconstruction are exact sign/swap operations.
#pragma once
#include "lym_reduce.h"
#include "lym_types.h"

namespace lym {

// A_mu row r: column perm, phase i^k.
//  4D: A0 = [[0,-i],[-i,0]]  A1 = [[0,-1],[1,0]]  A2 = [[-i,0],[0,i]]  A3 = 1
//  2D: A0 = 1 (g0 = sigma1)  A1 = -i (g1 = sigma2)
template <int D>
struct Gam;
template <>
struct Gam<4> {
  static constexpr int NS = 4;
  __host__ __device__ static constexpr int perm(int mu, int r) {
    return (mu <= 1) ? (1 - r) : r;
  }
  __host__ __device__ static constexpr int k(int mu, int r) {
    return mu == 0 ? 3 : mu == 1 ? (r == 0 ? 2 : 0) : mu == 2 ? (r == 0 ? 3 : 1) : 0;
  }
};
template <>
struct Gam<2> {
  static constexpr int NS = 2;
  __host__ __device__ static constexpr int perm(int, int) { return 0; }
  __host__ __device__ static constexpr int k(int mu, int) { return mu == 0 ? 0 : 3; }
};

// Orders the hops of one site: the next hop's site offsets are routed through
// an empty asm that consumes `dep` (the running accumulator), so the compiler
// cannot hoist every neighbour's loads to the top of the kernel.  One hop's
// loads are in flight per wave; occupancy hides the latency.  (No code is
// emitted; host pass: nothing.)
template <typename R>
__host__ __device__ __forceinline__ void hop_dep(int64_t& a, int64_t& b, R dep) {
#if defined(__HIP_DEVICE_COMPILE__)
  uint32_t a32 = (uint32_t)a, b32 = (uint32_t)b;
  asm volatile("" : "+v"(a32), "+v"(b32) : "v"(dep));
  a = a32;
  b = b32;
#else
  (void)a;
  (void)b;
  (void)dep;
#endif
}

// Marks every accumulator as produced here (volatile asms keep their order),
// so a hop's arithmetic completes before the next hop's loads are issued
// instead of all loaded values being held live until the end.
template <int NCOMP, typename R>
__host__ __device__ __forceinline__ void hop_commit(cpx<R> (&acc)[NCOMP]) {
#if defined(__HIP_DEVICE_COMPILE__)
#pragma unroll
  for (int c = 0; c < NCOMP; ++c) asm volatile("" : "+v"(acc[c].re), "+v"(acc[c].im));
#else
  (void)acc;
#endif
}

// multiply by i^k with k known at compile time via a switch the unroller folds
template <typename R>
__host__ __device__ __forceinline__ cpx<R> ipow_rt(int k, cpx<R> a) {
  switch (k & 3) {
    case 0: return a;
    case 1: return mk<R>(-a.im, a.re);
    case 2: return mk<R>(-a.re, -a.im);
    default: return mk<R>(a.im, -a.re);
  }
}

// colour transport of one colour object h (NCC components) by link U.
// fundamental (NCC = NC):   FWD: U h          BWD: U^dag h
// adjoint (NCC = NC*NC):    FWD: U h Ur^dag   BWD: U^dag h Ur
template <int NC, bool ADJ, bool FWD, typename R>
__host__ __device__ __forceinline__ void transport(const cpx<R> (&u)[NC][NC], const cpx<R> (&ur)[NC][NC],
                                                   const cpx<R>* h, cpx<R>* out) {
  if constexpr (!ADJ) {
#pragma unroll
    for (int i = 0; i < NC; ++i) {
      cpx<R> a = czero<R>();
#pragma unroll
      for (int j = 0; j < NC; ++j) {
        if constexpr (FWD) cfma(a, u[i][j], h[j]);
        else cfmac(a, u[j][i], h[j]);
      }
      out[i] = a;
    }
  } else {
    // t = U h (FWD) or U^dag h (BWD), then out = t Ur^dag (FWD) or t Ur (BWD)
    cpx<R> t[NC][NC];
#pragma unroll
    for (int i = 0; i < NC; ++i)
#pragma unroll
      for (int k = 0; k < NC; ++k) {
        cpx<R> a = czero<R>();
#pragma unroll
        for (int j = 0; j < NC; ++j) {
          if constexpr (FWD) cfma(a, u[i][j], h[j * NC + k]);
          else cfmac(a, u[j][i], h[j * NC + k]);
        }
        t[i][k] = a;
      }
#pragma unroll
    for (int i = 0; i < NC; ++i)
#pragma unroll
      for (int l = 0; l < NC; ++l) {
        cpx<R> a = czero<R>();
#pragma unroll
        for (int k = 0; k < NC; ++k) {
          if constexpr (FWD) cfma_c(a, t[i][k], ur[l][k]);  // t * conj(Ur)^T
          else cfma(a, t[i][k], ur[k][l]);
        }
        out[i * NC + l] = a;
      }
  }
}

// Hopping term of one site:  hop = sum_mu [ (1-g)U psi(x+mu) + (1+g)U^dag psi(x-mu) ]
// (without kappa, without the identity term, without gamma5).
//   PROJ = true : half-spinor projection + reconstruction (dirac.py:458-501)
//   PROJ = false: full projector on all Ns rows (dirac.py:503-523)
// One hop term (mu, dir: 0 = +mu, 1 = -mu) of a site with coordinate xmu
// along mu, given the neighbour spinor v and the hop's link(s) u, ur
// (forward: U_mu(x); backward: U_mu(x - mu)), accumulated into acc.
// SC: v is the neighbour scaled by 1/sc (16-bit storage), and sc enters the
// accumulating FMAs (acc += sc * transported), the hop being linear in v.
template <int D, int NC, bool ADJ, bool PROJ, typename R, bool SC = false>
__host__ __device__ __forceinline__ void wilson_hop_apply(const Lat& lat, int mu, int dir, int xmu,
                                                          const cpx<R> (&v)[Gam<D>::NS * (ADJ ? NC * NC : NC)],
                                                          const cpx<R> (&u)[NC][NC], const cpx<R> (&ur)[NC][NC],
                                                          cpx<R> (&acc)[Gam<D>::NS * (ADJ ? NC * NC : NC)],
                                                          R sc = R(1)) {
  constexpr int NS = Gam<D>::NS;
  constexpr int H = NS / 2;
  constexpr int NCC = ADJ ? NC * NC : NC;
  constexpr int NCOMP = NS * NCC;
  const bool fwd = (dir == 0);
  // forward term uses (1 - g), backward (1 + g): sg = -1 / +1
  const R sg = fwd ? R(-1) : R(1);
  const bool edge = lat.has_bc && (mu == D - 1) && (fwd ? xmu == lat.L[mu] - 1 : xmu == 0);
  const cpx<R> bcf = fwd ? mk<R>((R)lat.bc_re, (R)lat.bc_im) : mk<R>((R)lat.bc_re, -(R)lat.bc_im);
  if constexpr (PROJ) {
#pragma unroll
    for (int r = 0; r < H; ++r) {
      const int pr = Gam<D>::perm(mu, r);
      const int kr = Gam<D>::k(mu, r);
      cpx<R> h[NCC];
#pragma unroll
      for (int c = 0; c < NCC; ++c) {
        const cpx<R> w = ipow_rt<R>(kr, v[(H + pr) * NCC + c]);
        h[c] = mk<R>(v[r * NCC + c].re + sg * w.re, v[r * NCC + c].im + sg * w.im);
      }
      cpx<R> chi[NCC];
      if (fwd) transport<NC, ADJ, true, R>(u, ur, h, chi);
      else transport<NC, ADJ, false, R>(u, ur, h, chi);
      if (edge) {
#pragma unroll
        for (int c = 0; c < NCC; ++c) chi[c] = cmul(bcf, chi[c]);
      }
#pragma unroll
      for (int c = 0; c < NCC; ++c) {
        const cpx<R> w = ipow_rt<R>(4 - kr, chi[c]);
        if constexpr (SC) {
          const R ssg = sg * sc;
          acc[r * NCC + c].re = fmad(sc, chi[c].re, acc[r * NCC + c].re);
          acc[r * NCC + c].im = fmad(sc, chi[c].im, acc[r * NCC + c].im);
          acc[(H + pr) * NCC + c].re = fmad(ssg, w.re, acc[(H + pr) * NCC + c].re);
          acc[(H + pr) * NCC + c].im = fmad(ssg, w.im, acc[(H + pr) * NCC + c].im);
        } else {
          acc[r * NCC + c] += chi[c];
          acc[(H + pr) * NCC + c].re += sg * w.re;
          acc[(H + pr) * NCC + c].im += sg * w.im;
        }
      }
    }
  } else {
    // full projector rows: upper r and lower H+perm_r
    cpx<R> full[NCOMP];
#pragma unroll
    for (int r = 0; r < H; ++r) {
      const int pr = Gam<D>::perm(mu, r);
      const int kr = Gam<D>::k(mu, r);
#pragma unroll
      for (int c = 0; c < NCC; ++c) {
        const cpx<R> up = v[r * NCC + c];
        const cpx<R> dn = v[(H + pr) * NCC + c];
        const cpx<R> a = ipow_rt<R>(kr, dn);      // A psi_dn
        const cpx<R> b = ipow_rt<R>(4 - kr, up);  // A^dag psi_up
        full[r * NCC + c] = mk<R>(up.re + sg * a.re, up.im + sg * a.im);
        full[(H + pr) * NCC + c] = mk<R>(dn.re + sg * b.re, dn.im + sg * b.im);
      }
    }
#pragma unroll
    for (int sp = 0; sp < NS; ++sp) {
      cpx<R> chi[NCC];
      if (fwd) transport<NC, ADJ, true, R>(u, ur, &full[sp * NCC], chi);
      else transport<NC, ADJ, false, R>(u, ur, &full[sp * NCC], chi);
      if (edge) {
#pragma unroll
        for (int c = 0; c < NCC; ++c) chi[c] = cmul(bcf, chi[c]);
      }
#pragma unroll
      for (int c = 0; c < NCC; ++c) {
        if constexpr (SC) {
          acc[sp * NCC + c].re = fmad(sc, chi[c].re, acc[sp * NCC + c].re);
          acc[sp * NCC + c].im = fmad(sc, chi[c].im, acc[sp * NCC + c].im);
        } else {
          acc[sp * NCC + c] += chi[c];
        }
      }
    }
  }
}

// One hop: neighbour spinor at site nb (psi.load_hop), link(s) at site ls.
// SV: spinor view with load_hop<NS*NCC>(mu, dir, site, v); links: LinkView<R>.
template <int D, int NC, bool ADJ, bool PROJ, typename R, typename SV>
__host__ __device__ __forceinline__ void wilson_hop_one(const Lat& lat, int mu, int dir, int xmu, int64_t nb,
                                                        int64_t ls, const SV& psi, const LinkView<R>& links,
                                                        const LinkView<R>& links_r,
                                                        cpx<R> (&acc)[Gam<D>::NS * (ADJ ? NC * NC : NC)]) {
  constexpr int NCOMP = Gam<D>::NS * (ADJ ? NC * NC : NC);
  cpx<R> v[NCOMP];
  R sc = R(1);
  if constexpr (SV::kScaled) psi.template load_hop_scaled<NCOMP>(mu, dir, nb, v, sc);
  else psi.template load_hop<NCOMP>(mu, dir, nb, v);
  cpx<R> u[NC][NC];
  links.template load<NC>(mu, ls, u);
  cpx<R> ur[NC][NC];
  if constexpr (ADJ) links_r.template load<NC>(mu, ls, ur);
  wilson_hop_apply<D, NC, ADJ, PROJ, R, SV::kScaled>(lat, mu, dir, xmu, v, u, ur, acc, sc);
}

// Hops in the order (mu, +), (mu, -), mu = 0..D-1, in groups of HIF: the
// site offsets of a group pass through hop_dep after the previous group's
// accumulators are committed, so HIF hops' loads are in flight per wave
// (HIF = 1: one hop at a time, fewest registers; larger HIF: more memory
// parallelism for lattices too small to fill the chip).  The accumulation
// order, hence the result, does not depend on HIF.
template <int D, int NC, bool ADJ, bool PROJ, typename R, typename SV, int HIF = 1>
__host__ __device__ __forceinline__ void wilson_hop_x(const Lat& lat, int64_t s, const int (&x)[D], const SV& psi,
                                                      const LinkView<R>& links, const LinkView<R>& links_r,
                                                      cpx<R> (&acc)[Gam<D>::NS * (ADJ ? NC * NC : NC)]) {
  constexpr int NCOMP = Gam<D>::NS * (ADJ ? NC * NC : NC);
  static_assert((2 * D) % HIF == 0, "hops per group must divide 2D");
#pragma unroll
  for (int c = 0; c < NCOMP; ++c) acc[c] = czero<R>();
#pragma unroll
  for (int g = 0; g < 2 * D; g += HIF) {
    int64_t nb[HIF], ls[HIF];
#pragma unroll
    for (int q = 0; q < HIF; ++q) {
      const int mu = (g + q) >> 1;
      const bool fwd = ((g + q) & 1) == 0;
      const int L = lat.L[mu];
      const int64_t st = lat.stride[mu];
      nb[q] = fwd ? ((x[mu] == L - 1) ? s - (int64_t)(L - 1) * st : s + st)
                  : ((x[mu] == 0) ? s + (int64_t)(L - 1) * st : s - st);
      ls[q] = fwd ? s : nb[q];
    }
    hop_commit<NCOMP, R>(acc);
#pragma unroll
    for (int q = 0; q < HIF; ++q) hop_dep<R>(nb[q], ls[q], acc[0].re);
#pragma unroll
    for (int q = 0; q < HIF; ++q) {
      const int mu = (g + q) >> 1;
      wilson_hop_one<D, NC, ADJ, PROJ, R, SV>(lat, mu, (g + q) & 1, x[mu], nb[q], ls[q], psi, links, links_r, acc);
    }
  }
}

// hop of site s, coordinates computed from the site index
template <int D, int NC, bool ADJ, bool PROJ, typename R, typename SV>
__host__ __device__ __forceinline__ void wilson_hop(const Lat& lat, int64_t s, const SV& psi,
                                                    const LinkView<R>& links, const LinkView<R>& links_r,
                                                    cpx<R> (&acc)[Gam<D>::NS * (ADJ ? NC * NC : NC)]) {
  int x[D];
  coords<D>(lat, s, x);
  wilson_hop_x<D, NC, ADJ, PROJ, R, SV>(lat, s, x, psi, links, links_r, acc);
}

// Site and coordinates of this thread under the block schedule `m`; false if
// the thread has no site (tail of a linear schedule).  `sblk` is the site
// block's dispatch index among `gs` site blocks (sblk % kNumXcd = its XCD).
// `tid` / `nthr`: the thread's index among the block's `nthr` site threads.
template <int D>
__host__ __device__ __forceinline__ bool site_of(const Lat& lat, const SiteMap& m, int sblk, int gs, int tid,
                                                 int nthr, int64_t& s, int (&x)[D]) {
  if constexpr (D == 4) {
    if (m.tiled) {
      const int xcd = sblk % kNumXcd;
      const int j = sblk / kNumXcd;
      const int x0l = j % m.slab;
      const int r = j / m.slab;
      const int g = m.walk ? r / lat.L[1] : r % m.ng;
      x[0] = xcd * m.slab + x0l;
      x[1] = m.walk ? r % lat.L[1] : r / m.ng;
      x[2] = g * m.k + tid / m.T;
      x[3] = tid % m.T;
      s = (((int64_t)x[0] * lat.L[1] + x[1]) * lat.L[2] + x[2]) * m.T + x[3];
      return true;
    }
  }
  const int64_t per = gs / kNumXcd;
  const int64_t lb = (int64_t)(sblk % kNumXcd) * per + sblk / kNumXcd;
  s = lb * nthr + tid;
  if (s >= lat.V) return false;
  coords<D>(lat, s, x);
  return true;
}

template <int D>
__device__ __forceinline__ bool site_of_thread(const Lat& lat, const SiteMap& m, int sblk, int gs, int64_t& s,
                                               int (&x)[D]) {
  return site_of<D>(lat, m, sblk, gs, (int)threadIdx.x, (int)blockDim.x, s, x);
}

// (site block, right-hand side) of this block in a batched launch: all RHS of
// one site block run back to back on the same XCD (they share its links).
__device__ __forceinline__ void batch_block(int nrhs, int& sblk, int& rhs) {
  const int b = blockIdx.x;
  const int xcd = b % kNumXcd;
  const int j = b / kNumXcd;
  rhs = j % nrhs;
  sblk = (j / nrhs) * kNumXcd + xcd;
}

// res = g5^herm (center + kappa * hop)
template <int NCOMP, typename R>
__host__ __device__ __forceinline__ void wilson_finish(const Lat& lat, const cpx<R> (&center)[NCOMP],
                                                       const cpx<R> (&hop)[NCOMP], cpx<R> (&res)[NCOMP]) {
  const R k = (R)lat.kappa;
  constexpr int HALF = NCOMP / 2;
#pragma unroll
  for (int c = 0; c < NCOMP; ++c) {
    cpx<R> t = mk<R>(center[c].re + k * hop[c].re, center[c].im + k * hop[c].im);
    if (lat.herm && c >= HALF) t = mk<R>(-t.re, -t.im);
    res[c] = t;
  }
}

}  // namespace lym
