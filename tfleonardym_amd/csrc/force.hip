// Fermion force kernel: out_mu(x) (+)= sum_p w_p [L(X_p, Y_p) + sym L(Y_p, X_p)]
// with L the Lie derivative of the Wilson operator (dirac.py:525-565), the
// sum n_flavor.force forms over rational-approximation terms
// (fermion_action.py:86-91).
//
// One thread per (site, direction).  A block holds its sites once per
// direction, as D groups of whole wavefronts (direction uniform per wave):
// the D waves read the same X(x), Y(x) back to back, so those hit in cache
// after the first.  Fundamental representation: the pair loop accumulates
// colour outer products only; the link U_mu(x) is read and applied once per
// thread at the end.  Bound: HBM.
#include <cstdlib>

#include "force_site.h"
#include "lym_internal.h"

namespace lym {

constexpr int kForceSites = 64;

// LYM_FORCE_LDS=1 in the environment selects the LDS-staged kernel (opt-in:
// it halves the bytes fetched but is slower, see k_force_lds); read at every
// launch so tests can run both kernels in one process
static bool force_lds_enabled() {
  const char* e = std::getenv("LYM_FORCE_LDS");
  return e && e[0] == '1';
}

// TILED: the Wilson kernels' XCD-tiled schedule (4D): a block covers k whole
// lines of the fastest direction, XCD x owns an x0 slab and runs its blocks
// x0 fastest, so the x+mu neighbours of a block were fetched into the same
// L2 by the block before.  Otherwise kForceSites consecutive sites per block,
// block ranges contiguous per XCD.
template <int D, int NC, bool ADJ, bool HERM, bool TILED>
__global__ __launch_bounds__(TILED ? 1024 : kForceSites * 4) void k_force(Lat lat, SiteMap map, int nthr, int gs,
                                                                          LinkView<double> lk, LinkView<double> lkr,
                                                                          ForcePairs fp, uint32_t vbytes,
                                                                          cpx<double>* out, uint32_t obytes) {
  const int mu = (int)threadIdx.x / nthr;
  const int tid = (int)threadIdx.x % nthr;
  int64_t s;
  int x[D];
  const SiteMap m = TILED ? map : SiteMap{0, 0, 0, 0, 0};
  if (!site_of<D>(lat, m, (int)blockIdx.x, gs, tid, nthr, s, x)) return;
  force_site<D, NC, ADJ, HERM>(lat, mu, s, x, lk, lkr, fp, vbytes, out, obytes);
}

// LDS-staged kernel (4D fundamental, XCD-tiled schedule): a block covers
// nsite = k * T sites (k whole lines of the fastest direction, contiguous
// site indices) and the 4 directions as groups of whole waves.  Per pair the
// block copies X_p and Y_p of its sites into LDS with one coalesced read per
// component, and every thread reads its own X(x), Y(x) and the in-tile
// neighbours (x3 + 1 always, x2 + 1 unless on the last line of the block)
// from there; only the x0, x1 neighbours and the x2 face go to global memory.
// Two LDS buffers: the loads of pair p+1 are in flight while pair p is
// computed, one barrier per pair.  Measured at 24^3 x 48 (4 pairs) it reads
// 3.0 GB instead of 5.4 GB per launch but takes 886 us against 695-790 us for
// k_force: the 144 KiB of LDS allow one block (12 waves) per CU, and at every
// barrier the waves of the x0 and x1 directions, whose neighbours still come
// from global memory in four dependent round trips per pair, hold back the
// others.  Opt-in (LYM_FORCE_LDS=1).
constexpr int64_t kForceLdsMax = 150 * 1024;
// block size limit: SU(3) needs more than the 128 VGPRs of 1024-thread blocks
template <int NC>
constexpr int force_lds_threads() {
  return NC == 3 ? 768 : 1024;
}

template <int NC, bool HERM>
__global__ __launch_bounds__(force_lds_threads<NC>()) void k_force_lds(Lat lat, SiteMap map, int nsite, LinkView<double> lk,
                                                    ForcePairs fp, uint32_t vbytes, cpx<double>* out,
                                                    uint32_t obytes) {
  constexpr int D = 4;
  constexpr int NCOMP = 4 * NC;
  constexpr int NST = 2 * NCOMP / D;  // staged components per thread and pair
  extern __shared__ double2 lds_raw[];
  cpx<double>* lds = reinterpret_cast<cpx<double>*>(lds_raw);
  const int mu = (int)threadIdx.x / nsite;
  const int sl = (int)threadIdx.x % nsite;
  const int nthr = D * nsite;
  int64_t s;
  int x[D];
  site_of<D>(lat, map, (int)blockIdx.x, (int)gridDim.x, sl, nsite, s, x);
  const int64_t s0 = s - sl;  // the block's sites are s0 .. s0 + nsite - 1
  const int L = lat.L[mu];
  const int64_t st = lat.stride[mu];
  const int64_t nb = (x[mu] == L - 1) ? s - (int64_t)(L - 1) * st : s + st;
  const bool edge = lat.has_bc && mu == D - 1 && x[mu] == L - 1;
  // neighbour x + mu inside the staged tile?
  const int T = lat.L[3];
  const int xl2 = sl / T;  // line of the site inside the block
  bool in_tile = false;
  int nsl = 0;
  if (mu == 3) {
    in_tile = true;
    nsl = (x[3] == T - 1) ? sl - (T - 1) : sl + 1;
  } else if (mu == 2) {
    if (xl2 < map.k - 1) {
      in_tile = true;
      nsl = sl + T;
    } else if (map.k == lat.L[2]) {
      in_tile = true;
      nsl = sl - (map.k - 1) * T;
    }
  }
  const int tile_elems = 2 * NCOMP * nsite;  // one buffer: X then Y, plane-major
  cpx<double> stage[NST];
  auto load_pair = [&](int p) {
#pragma unroll
    for (int j = 0; j < NST; ++j) {
      const int e = (int)threadIdx.x + j * nthr;
      const int v = e / (NCOMP * nsite);
      const int rem = e - v * NCOMP * nsite;
      const int plane = rem / nsite;
      const int l = rem - plane * nsite;
      stage[j] = ld<cpx<double>>(v ? fp.y[p] : fp.x[p], vbytes, plane, lat.V, s0 + l);
    }
  };
  auto store_pair = [&](int b) {
#pragma unroll
    for (int j = 0; j < NST; ++j) lds[b * tile_elems + (int)threadIdx.x + j * nthr] = stage[j];
  };
  if (fp.n > 0) {
    load_pair(0);
    store_pair(0);
  }
  __syncthreads();
  cpx<double> P[NC][NC], Q[NC][NC];
#pragma unroll
  for (int i = 0; i < NC; ++i)
#pragma unroll
    for (int j = 0; j < NC; ++j) P[i][j] = Q[i][j] = czero<double>();
  for (int p = 0; p < fp.n; ++p) {
    const int b = p & 1;
    if (p + 1 < fp.n) load_pair(p + 1);
    const cpx<double>* tx = lds + b * tile_elems;
    const cpx<double>* ty = tx + NCOMP * nsite;
    const double w = fp.w[p];
    {
      const LdsAcc xa{tx, nsite, sl};
      const MixAcc ya{LdsAcc{ty, nsite, nsl}, GlobalAcc{fp.y[p], vbytes, lat.V, nb}, in_tile};
      force_term_acc<D, NC, HERM>(mu, xa, ya, w, P, Q);
    }
    if (fp.sym) {
      const LdsAcc ya{ty, nsite, sl};
      const MixAcc xa{LdsAcc{tx, nsite, nsl}, GlobalAcc{fp.x[p], vbytes, lat.V, nb}, in_tile};
      force_term_acc<D, NC, HERM>(mu, ya, xa, w, P, Q);
    }
    if (p + 1 < fp.n) store_pair(b ^ 1);
    __syncthreads();
  }
  cpx<double> u[NC][NC];
  lk.template load<NC>(mu, s, u);
  force_close_fund<NC, HERM>(u, edge, mk<double>(lat.bc_re, lat.bc_im), P, Q);
  cpx<double> res[NC][NC];
  force_finish<NC, false>(lat, P, res);
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

// Pair-parallel kernel (4D fundamental, hermitian): a block covers nsite
// sites (nsite / T whole lines of the fastest direction, the tiled schedule)
// x D directions x PP pair slots; slot q sums the pairs p = q, q + PP, ...
// of its (site, direction) into P, the slots are added in LDS in a fixed
// order (slot PP-1 first, slot 0 last), and slot 0 applies the link and
// writes.  A block lives for about n / PP pairs instead of n, so the blocks
// in flight on an XCD work on fewer pairs at a time and the x + mu
// neighbours a block reads were fetched by the blocks next to it.
template <int NC, int PP>
__global__ __launch_bounds__(1024) void k_force_pp(Lat lat, SiteMap map, int nsite, LinkView<double> lk,
                                                   ForcePairs fp, uint32_t vbytes, cpx<double>* out,
                                                   uint32_t obytes) {
  constexpr int D = 4;
  extern __shared__ double2 force_pp_lds[];  // [D][nsite][NC * NC]
  const int t = (int)threadIdx.x;
  const int sl = t % nsite;
  const int mu = (t / nsite) % D;
  const int pq = t / (nsite * D);
  int64_t s;
  int x[D];
  site_of<D>(lat, map, (int)blockIdx.x, (int)gridDim.x, sl, nsite, s, x);
  const int L = lat.L[mu];
  const int64_t st = lat.stride[mu];
  const int64_t nb = (x[mu] == L - 1) ? s - (int64_t)(L - 1) * st : s + st;
  cpx<double> P[NC][NC], Q[NC][NC];
#pragma unroll
  for (int i = 0; i < NC; ++i)
#pragma unroll
    for (int j = 0; j < NC; ++j) P[i][j] = Q[i][j] = czero<double>();
  for (int p = pq; p < fp.n; p += PP) {
    force_term_fund<D, NC, true>(lat, mu, fp.x[p], fp.y[p], vbytes, s, nb, fp.w[p], P, Q);
    if (fp.sym) force_term_fund<D, NC, true>(lat, mu, fp.y[p], fp.x[p], vbytes, s, nb, fp.w[p], P, Q);
  }
  double2* slot = force_pp_lds + ((int64_t)mu * nsite + sl) * NC * NC;
  if (pq == PP - 1) {
#pragma unroll
    for (int i = 0; i < NC; ++i)
#pragma unroll
      for (int j = 0; j < NC; ++j) slot[i * NC + j] = make_double2(P[i][j].re, P[i][j].im);
  }
  __syncthreads();
#pragma unroll
  for (int q = PP - 2; q >= 0; --q) {
    if (pq == q) {
#pragma unroll
      for (int i = 0; i < NC; ++i)
#pragma unroll
        for (int j = 0; j < NC; ++j) {
          const double2 v = slot[i * NC + j];
          P[i][j] = mk<double>(v.x + P[i][j].re, v.y + P[i][j].im);
          if (q > 0) slot[i * NC + j] = make_double2(P[i][j].re, P[i][j].im);
        }
    }
    __syncthreads();
  }
  if (pq != 0) return;
  const bool edge = lat.has_bc && mu == D - 1 && x[mu] == L - 1;
  cpx<double> u[NC][NC];
  lk.template load<NC>(mu, s, u);
  force_close_fund<NC, true>(u, edge, mk<double>(lat.bc_re, lat.bc_im), P, Q);
  cpx<double> res[NC][NC];
  force_finish<NC, false>(lat, P, res);
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

template <int D, int NC, bool ADJ, bool HERM>
static int launch_force(const OpDesc& d, const ForcePairs& fp, cpx<double>* out, hipStream_t st) {
  const int64_t V = d.lat.V;
  const int recon = d.link_su ? 1 : 0;
  LinkView<double> lk{(const cpx<double>*)d.links64, V, recon};
  LinkView<double> lkr{(const cpx<double>*)(d.linksr64 ? d.linksr64 : d.links64), V, recon};
  const uint32_t vbytes = (uint32_t)(field_elems(d) * 16);
  const uint32_t obytes = (uint32_t)((int64_t)D * NC * NC * V * 16);
  // (hermitian only: the non-hermitian form's second accumulator does not fit the register budget)
  if constexpr (D == 4 && !ADJ && HERM) {
    const int64_t lds = (int64_t)2 * 2 * 4 * NC * d.wthreads * 16;
    if (d.map.tiled && d.wthreads * D <= force_lds_threads<NC>() && lds <= kForceLdsMax && force_lds_enabled()) {
      static std::atomic<uint32_t> raised;  // devices whose dynamic-LDS limit is raised
      const int rc = raise_lds_limit((const void*)k_force_lds<NC, HERM>, (int)kForceLdsMax, raised);
      if (rc) return rc;
      hipLaunchKernelGGL((k_force_lds<NC, HERM>), dim3(d.wblocks), dim3(d.wthreads * D), (size_t)lds, st, d.lat,
                         d.map, d.wthreads, lk, fp, vbytes, out, obytes);
      LYM_HIP(hipGetLastError());
      return LYM_OK;
    }
  }
  // three or more pairs: two pair slots per (site, direction) over blocks of
  // two x3 lines (measured at 24^3 x 48: 4 pairs 705 against 770 us, 8 pairs
  // 1300 against 1679 us; with one or two pairs the idle slots cost more
  // than the shorter block lifetime gains, 313 / 431 against 255 / 380 us)
  if constexpr (D == 4 && !ADJ && HERM) {
    constexpr int kl = 2;
    if (fp.n >= 3 && d.map.tiled && d.lat.L[2] % kl == 0) {
      const int T = d.map.T;
      const int nsite = T * kl;
      // x1 before the x2 groups: a site's x+1 neighbour then lies a few
      // blocks later on the XCD instead of a whole x1 row of blocks later
      // (709 against 716 us at 24^3 x 48, 4 pairs)
      const SiteMap m{1, T, kl, d.map.slab, d.lat.L[2] / kl, 1};
      const size_t lds = (size_t)D * nsite * NC * NC * 16;
      const int nthr = nsite * D * 2;
      if (nthr <= 1024 && nthr % 64 == 0 && lds <= 64 * 1024) {
        hipLaunchKernelGGL((k_force_pp<NC, 2>), dim3((unsigned)(d.lat.V / nsite)), dim3(nthr), lds, st, d.lat, m,
                           nsite, lk, fp, vbytes, out, obytes);
        LYM_HIP(hipGetLastError());
        return LYM_OK;
      }
    }
  }
  // (the non-hermitian SU(3) form needs more registers than 1024-thread blocks allow)
  if constexpr (D == 4 && !ADJ && (HERM || NC == 2)) {
    if (d.map.tiled && d.wthreads * D <= 1024) {
      hipLaunchKernelGGL((k_force<D, NC, ADJ, HERM, true>), dim3(d.wblocks), dim3(d.wthreads * D), 0, st, d.lat,
                         d.map, d.wthreads, d.wblocks, lk, lkr, fp, vbytes, out, obytes);
      LYM_HIP(hipGetLastError());
      return LYM_OK;
    }
  }
  const int gs = grid_for(V, kForceSites);
  hipLaunchKernelGGL((k_force<D, NC, ADJ, HERM, false>), dim3(gs), dim3(kForceSites * D), 0, st, d.lat, d.map,
                     kForceSites, gs, lk, lkr, fp, vbytes, out, obytes);
  LYM_HIP(hipGetLastError());
  return LYM_OK;
}

template <int D, int NC, bool ADJ>
static int launch_force_h(const OpDesc& d, const ForcePairs& fp, cpx<double>* out, hipStream_t st) {
  if constexpr (ADJ) {
    return launch_force<D, NC, true, false>(d, fp, out, st);  // the adjoint path reads herm at run time
  } else {
    if (d.herm) return launch_force<D, NC, false, true>(d, fp, out, st);
    return launch_force<D, NC, false, false>(d, fp, out, st);
  }
}

int force_launch(const OpDesc& d, const ForcePairs& fp, void* out, hipStream_t st) {
  cpx<double>* o = (cpx<double>*)out;
  const int key = d.ndim * 100 + d.nc * 10 + d.adj;
  switch (key) {
    case 420: return launch_force_h<4, 2, false>(d, fp, o, st);
    case 430: return launch_force_h<4, 3, false>(d, fp, o, st);
    case 220: return launch_force_h<2, 2, false>(d, fp, o, st);
    case 230: return launch_force_h<2, 3, false>(d, fp, o, st);
    case 421: return launch_force_h<4, 2, true>(d, fp, o, st);
    case 431: return launch_force_h<4, 3, true>(d, fp, o, st);
    case 221: return launch_force_h<2, 2, true>(d, fp, o, st);
    case 231: return launch_force_h<2, 3, true>(d, fp, o, st);
  }
  set_error("force: unsupported (ndim, nc, representation)");
  return LYM_ERR_ARG;
}

}  // namespace lym
