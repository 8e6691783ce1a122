// Wilson hopping kernels for gfx950: one site per lane, site-coalesced
// component planes, XCD-contiguous block ranges, fused epilogues that store,
// combine and reduce (FP64) in the same pass.
#include <hip/hip_runtime.h>

#include <type_traits>

#include "lym_fin.h"
#include "lym_internal.h"
#include "wilson_site.h"

namespace lym {

template <int NCOMP, typename R>
__device__ __forceinline__ double norm2(const cpx<R> (&v)[NCOMP]) {
  double a = 0.0;
#pragma unroll
  for (int c = 0; c < NCOMP; ++c) a += (double)v[c].re * (double)v[c].re + (double)v[c].im * (double)v[c].im;
  return a;
}

// ---------------------------------------------------------------------------
// Epilogues.  apply() runs per site after the hop is computed; finalize()
// runs on thread 0 of the last block with the grid totals of red[].
// ---------------------------------------------------------------------------
template <int P, int NCOMP>
struct EpiStore {
  using R = typename Store<P>::R;
  static constexpr int NRED = 0;
  SpinorView<P> out;
  EpiStore(const EpiArgs& a, int64_t V) : out{(typename Store<P>::elem*)a.out.data, a.out.norm, V} {}
  __device__ bool active() const { return true; }
  __device__ void apply(const Lat& lat, int64_t s, const SpinorView<P>& in, const cpx<R> (&hop)[NCOMP],
                        double*) const {
    cpx<R> c[NCOMP], res[NCOMP];
    in.template load<NCOMP>(s, c);
    wilson_finish<NCOMP, R>(lat, c, hop, res);
    out.template store<NCOMP>(s, res);
  }
  __device__ void finalize(const double*) const {}
};

template <int P, int NCOMP>
struct EpiAxpy {
  using R = typename Store<P>::R;
  static constexpr int NRED = 0;
  SpinorView<P> out, w;
  double sigma;
  EpiAxpy(const EpiArgs& a, int64_t V)
      : out{(typename Store<P>::elem*)a.out.data, a.out.norm, V},
        w{(typename Store<P>::elem*)a.aux0.data, a.aux0.norm, V},
        sigma(a.sigma) {}
  __device__ bool active() const { return true; }
  __device__ void apply(const Lat& lat, int64_t s, const SpinorView<P>& in, const cpx<R> (&hop)[NCOMP],
                        double*) const {
    cpx<R> c[NCOMP], res[NCOMP], wv[NCOMP];
    in.template load<NCOMP>(s, c);
    w.template load<NCOMP>(s, wv);
    wilson_finish<NCOMP, R>(lat, c, hop, res);
    const R sg = (R)sigma;
#pragma unroll
    for (int k = 0; k < NCOMP; ++k) res[k] = mk<R>(res[k].re + sg * wv[k].re, res[k].im + sg * wv[k].im);
    out.template store<NCOMP>(s, res);
  }
  __device__ void finalize(const double*) const {}
};

// CG-M step 1: t = Q p, reduce |t|^2 and |p|^2 (<p,(Q^2+s0)p> = |Qp|^2 + s0|p|^2).
template <int P, int NCOMP>
struct EpiCgmK1 {
  using R = typename Store<P>::R;
  static constexpr int NRED = 2;
  SpinorView<P> out;
  SolverState* st;
  EpiCgmK1(const EpiArgs& a, int64_t V) : out{(typename Store<P>::elem*)a.out.data, a.out.norm, V}, st(a.st) {}
  // plain CG-M enters at PH_K1, the fast variant at PH_KP (after p = r + beta p)
  __device__ bool active() const { return st->phase == PH_K1 || st->phase == PH_KP; }
  __device__ void apply(const Lat& lat, int64_t s, const SpinorView<P>& in, const cpx<R> (&hop)[NCOMP],
                        double* red) const {
    cpx<R> c[NCOMP], res[NCOMP];
    in.template load<NCOMP>(s, c);
    wilson_finish<NCOMP, R>(lat, c, hop, res);
    out.template store<NCOMP>(s, res);
    red[0] += norm2<NCOMP, R>(res);
    red[1] += norm2<NCOMP, R>(c);
  }
  __device__ void finalize(const double* sums) const { cgm_fin_a(st, sums[0], sums[1]); }
};

// CG-M step 2: Ap = Q t + s0 p; r <- r - alpha Ap; reduce |r|^2.
template <int P, int NCOMP>
struct EpiCgmK2 {
  using R = typename Store<P>::R;
  static constexpr int NRED = 1;
  SpinorView<P> p, r;
  SolverState* st;
  EpiCgmK2(const EpiArgs& a, int64_t V)
      : p{(typename Store<P>::elem*)a.aux0.data, a.aux0.norm, V},
        r{(typename Store<P>::elem*)a.aux1.data, a.aux1.norm, V},
        st(a.st) {}
  __device__ bool active() const { return st->phase == PH_K2; }
  __device__ void apply(const Lat& lat, int64_t s, const SpinorView<P>& in, const cpx<R> (&hop)[NCOMP],
                        double* red) const {
    cpx<R> c[NCOMP], res[NCOMP], pv[NCOMP], rv[NCOMP];
    in.template load<NCOMP>(s, c);
    p.template load<NCOMP>(s, pv);
    r.template load<NCOMP>(s, rv);
    wilson_finish<NCOMP, R>(lat, c, hop, res);
    const R s0 = (R)st->sigma0;
    const R al = (R)st->alpha;
#pragma unroll
    for (int k = 0; k < NCOMP; ++k) {
      const R apr = res[k].re + s0 * pv[k].re, api = res[k].im + s0 * pv[k].im;
      rv[k] = mk<R>(rv[k].re - al * apr, rv[k].im - al * api);
    }
    r.template store<NCOMP>(s, rv);
    red[0] += norm2<NCOMP, R>(rv);
  }
  __device__ void finalize(const double* sums) const { cgm_fin_b(st, sums[0]); }
};

// Fast CG-M step 2: residuals live in a ring of nslot fields; iteration k reads
// r^k from slot k % nslot and writes r^{k+1} to slot (k+1) % nslot, so the
// last block's worth of residuals stays available for the deferred shift
// updates.  r^{k+1} = r^k - alpha (Q t + s0 p); reduce |r^{k+1}|^2.
template <int P, int NCOMP>
struct EpiCgmK2F {
  using R = typename Store<P>::R;
  using E = typename Store<P>::elem;
  static constexpr int NRED = 1;
  SpinorView<P> p;
  E* ring;
  float* ring_norm;
  int64_t slot_elems;
  int nslot;
  SolverState* st;
  EpiCgmK2F(const EpiArgs& a, int64_t V)
      : p{(E*)a.aux0.data, a.aux0.norm, V},
        ring((E*)a.ring.data),
        ring_norm(a.ring.norm),
        slot_elems(a.slot_elems),
        nslot(a.nslot),
        st(a.st) {}
  __device__ bool active() const { return st->phase == PH_K2; }
  __device__ void apply(const Lat& lat, int64_t s, const SpinorView<P>& in, const cpx<R> (&hop)[NCOMP],
                        double* red) const {
    const int k = st->iter;
    const int si = k % nslot, so = (k + 1) % nslot;
    const SpinorView<P> rin{ring + si * slot_elems, ring_norm ? ring_norm + si * lat.V : nullptr, lat.V};
    const SpinorView<P> rout{ring + so * slot_elems, ring_norm ? ring_norm + so * lat.V : nullptr, lat.V};
    cpx<R> c[NCOMP], res[NCOMP], pv[NCOMP], rv[NCOMP];
    in.template load<NCOMP>(s, c);
    p.template load<NCOMP>(s, pv);
    rin.template load<NCOMP>(s, rv);
    wilson_finish<NCOMP, R>(lat, c, hop, res);
    const R s0 = (R)st->sigma0;
    const R al = (R)st->alpha;
#pragma unroll
    for (int q = 0; q < NCOMP; ++q) {
      const R apr = res[q].re + s0 * pv[q].re, api = res[q].im + s0 * pv[q].im;
      rv[q] = mk<R>(rv[q].re - al * apr, rv[q].im - al * api);
    }
    rout.template store<NCOMP>(s, rv);
    red[0] += norm2<NCOMP, R>(rv);
  }
  __device__ void finalize(const double* sums) const { cgm_fin_b_fast(st, sums[0]); }
};

// MR step 1: t = Op r (phase gated).
template <int P, int NCOMP>
struct EpiMrK1 {
  using R = typename Store<P>::R;
  static constexpr int NRED = 0;
  SpinorView<P> out;
  SolverState* st;
  EpiMrK1(const EpiArgs& a, int64_t V) : out{(typename Store<P>::elem*)a.out.data, a.out.norm, V}, st(a.st) {}
  __device__ bool active() const { return st->phase == PH_K1; }
  __device__ void apply(const Lat& lat, int64_t s, const SpinorView<P>& in, const cpx<R> (&hop)[NCOMP],
                        double*) const {
    cpx<R> c[NCOMP], res[NCOMP];
    in.template load<NCOMP>(s, c);
    wilson_finish<NCOMP, R>(lat, c, hop, res);
    out.template store<NCOMP>(s, res);
  }
  __device__ void finalize(const double*) const { st->phase = PH_K2; }
};

// MR step 2: p = Op t + s_N r; reduce <p,r> (complex) and <p,p>.
template <int P, int NCOMP>
struct EpiMrK2 {
  using R = typename Store<P>::R;
  static constexpr int NRED = 3;
  SpinorView<P> out, r;
  SolverState* st;
  EpiMrK2(const EpiArgs& a, int64_t V)
      : out{(typename Store<P>::elem*)a.out.data, a.out.norm, V},
        r{(typename Store<P>::elem*)a.aux0.data, a.aux0.norm, V},
        st(a.st) {}
  __device__ bool active() const { return st->phase == PH_K1; }
  __device__ void apply(const Lat& lat, int64_t s, const SpinorView<P>& in, const cpx<R> (&hop)[NCOMP],
                        double* red) const {
    cpx<R> c[NCOMP], res[NCOMP], rv[NCOMP];
    in.template load<NCOMP>(s, c);
    r.template load<NCOMP>(s, rv);
    wilson_finish<NCOMP, R>(lat, c, hop, res);
    const R sg = (R)st->shift[st->nshift - 1];
#pragma unroll
    for (int k = 0; k < NCOMP; ++k) {
      res[k] = mk<R>(res[k].re + sg * rv[k].re, res[k].im + sg * rv[k].im);
      // <p, r> = sum conj(p) r
      red[0] += (double)res[k].re * rv[k].re + (double)res[k].im * rv[k].im;
      red[1] += (double)res[k].re * rv[k].im - (double)res[k].im * rv[k].re;
      red[2] += (double)res[k].re * res[k].re + (double)res[k].im * res[k].im;
    }
    out.template store<NCOMP>(s, res);
  }
  __device__ void finalize(const double* sums) const { mr_fin(st, sums[0], sums[1], sums[2]); }
};

// ---------------------------------------------------------------------------
// The kernel
// ---------------------------------------------------------------------------
template <int D, int NC, bool ADJ, bool PROJ, int P, class Epi>
__global__ __launch_bounds__(kBlock) void k_wilson(Lat lat, SiteMap map, SpinorView<P> in,
                                                   LinkView<typename Store<P>::R> lk,
                                                   LinkView<typename Store<P>::R> lkr, Epi epi, RedCtl rc) {
  using R = typename Store<P>::R;
  constexpr int NCOMP = Gam<D>::NS * (ADJ ? NC * NC : NC);
  if (!epi.active()) return;
  constexpr int NR = Epi::NRED > 0 ? Epi::NRED : 1;
  double red[NR];
#pragma unroll
  for (int r = 0; r < NR; ++r) red[r] = 0.0;
  int64_t s;
  int x[D];
  if (site_of_thread<D>(lat, map, s, x)) {
    cpx<R> hop[NCOMP];
    wilson_hop_x<D, NC, ADJ, PROJ, R>(lat, s, x, in, lk, lkr, hop);
    epi.apply(lat, s, in, hop, red);
  }
  if constexpr (Epi::NRED > 0) block_partials<NR>(red, rc.partials);
}

// ---------------------------------------------------------------------------
// Dispatch: runtime (kind, ndim, nc, rep, prec, mode) -> template instance.
// ---------------------------------------------------------------------------
template <int D, int NC, bool ADJ, int P, template <int, int> class EpiT>
static int launch_one(const OpDesc& d, bool proj, Field in, const EpiArgs& a, RedCtl rc, hipStream_t stream) {
  using R = typename Store<P>::R;
  constexpr int NCOMP = Gam<D>::NS * (ADJ ? NC * NC : NC);
  using Epi = EpiT<P, NCOMP>;
  const int64_t V = d.lat.V;
  const void* l = (P == P64) ? d.links64 : d.links32;
  const void* lr = (P == P64) ? (d.linksr64 ? d.linksr64 : d.links64) : (d.linksr32 ? d.linksr32 : d.links32);
  if (!l) {
    set_error(P == P64 ? "operator has no complex128 links" : "operator has no complex64 links (call with a float32 copy)");
    return LYM_ERR_ARG;
  }
  LinkView<R> lk{(const cpx<R>*)l, V};
  LinkView<R> lkr{(const cpx<R>*)lr, V};
  SpinorView<P> sv{(typename Store<P>::elem*)in.data, in.norm, V};
  Epi epi(a, V);
  const dim3 grid(d.wblocks), block(d.wthreads);
  if (proj) {
    hipLaunchKernelGGL((k_wilson<D, NC, ADJ, true, P, Epi>), grid, block, 0, stream, d.lat, d.map, sv, lk, lkr, epi,
                       rc);
  } else {
    if constexpr (P == P64 && std::is_same<Epi, EpiStore<P, NCOMP>>::value) {
      hipLaunchKernelGGL((k_wilson<D, NC, ADJ, false, P, Epi>), grid, block, 0, stream, d.lat, d.map, sv, lk, lkr,
                         epi, rc);
    } else {
      set_error("simple (full-projector) mode is only built for the plain complex128 apply");
      return LYM_ERR_UNSUPPORTED;
    }
  }
  LYM_HIP(hipGetLastError());
  return LYM_OK;
}

template <int P, template <int, int> class EpiT>
static int launch_geom(const OpDesc& d, bool proj, Field in, const EpiArgs& a, RedCtl rc, hipStream_t st) {
  const int key = d.ndim * 100 + d.nc * 10 + d.adj;
  if constexpr (P == P16) {
    // half-precision storage is built for the fundamental representation only
    switch (key) {
      case 420: return launch_one<4, 2, false, P, EpiT>(d, proj, in, a, rc, st);
      case 430: return launch_one<4, 3, false, P, EpiT>(d, proj, in, a, rc, st);
      case 220: return launch_one<2, 2, false, P, EpiT>(d, proj, in, a, rc, st);
      case 230: return launch_one<2, 3, false, P, EpiT>(d, proj, in, a, rc, st);
      default:
        set_error("half-precision storage: fundamental representation only");
        return LYM_ERR_UNSUPPORTED;
    }
  }
  switch (key) {
    case 420: return launch_one<4, 2, false, P, EpiT>(d, proj, in, a, rc, st);
    case 430: return launch_one<4, 3, false, P, EpiT>(d, proj, in, a, rc, st);
    case 220: return launch_one<2, 2, false, P, EpiT>(d, proj, in, a, rc, st);
    case 230: return launch_one<2, 3, false, P, EpiT>(d, proj, in, a, rc, st);
    case 421: return launch_one<4, 2, true, P, EpiT>(d, proj, in, a, rc, st);
    case 431: return launch_one<4, 3, true, P, EpiT>(d, proj, in, a, rc, st);
    case 221: return launch_one<2, 2, true, P, EpiT>(d, proj, in, a, rc, st);
    case 231: return launch_one<2, 3, true, P, EpiT>(d, proj, in, a, rc, st);
    default:
      set_error("unsupported (ndim, nc, representation)");
      return LYM_ERR_UNSUPPORTED;
  }
}

template <template <int, int> class EpiT>
static int launch_prec(int prec, const OpDesc& d, bool proj, Field in, const EpiArgs& a, RedCtl rc, hipStream_t st) {
  switch (prec) {
    case P64: return launch_geom<P64, EpiT>(d, proj, in, a, rc, st);
    case P32: return launch_geom<P32, EpiT>(d, proj, in, a, rc, st);
    case P16: return launch_geom<P16, EpiT>(d, proj, in, a, rc, st);
    default:
      set_error("unknown precision");
      return LYM_ERR_ARG;
  }
}

int wilson_launch(EpiKind kind, const OpDesc& d, int prec, bool proj, Field in, const EpiArgs& a, RedCtl rc,
                  hipStream_t stream) {
  switch (kind) {
    case EPI_STORE: return launch_prec<EpiStore>(prec, d, proj, in, a, rc, stream);
    case EPI_AXPY: return launch_prec<EpiAxpy>(prec, d, proj, in, a, rc, stream);
    case EPI_CGM_K1: return launch_prec<EpiCgmK1>(prec, d, proj, in, a, rc, stream);
    case EPI_CGM_K2: return launch_prec<EpiCgmK2>(prec, d, proj, in, a, rc, stream);
    case EPI_CGM_K2F: return launch_prec<EpiCgmK2F>(prec, d, proj, in, a, rc, stream);
    case EPI_MR_K1:
    case EPI_MR_K2:
      if (prec != P64) {
        set_error("the multishift MR solver runs in complex128 only");
        return LYM_ERR_UNSUPPORTED;
      }
      if (kind == EPI_MR_K1) return launch_geom<P64, EpiMrK1>(d, proj, in, a, rc, stream);
      return launch_geom<P64, EpiMrK2>(d, proj, in, a, rc, stream);
  }
  set_error("unknown epilogue");
  return LYM_ERR_ARG;
}

}  // namespace lym
