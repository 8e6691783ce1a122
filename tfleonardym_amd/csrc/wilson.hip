// Wilson hopping kernels for gfx950: one site per lane, site-coalesced
// component planes, XCD-aware block schedule, fused epilogues that store,
// combine and reduce (FP64) in the same pass.  A launch may cover a batch of
// right-hand sides: fields of RHS i are offset by i * rhs_stride elements and
// solver scalars by one SolverState per RHS.
#include <hip/hip_runtime.h>

#include <type_traits>
#include <utility>

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

template <int P>
__host__ __device__ __forceinline__ SpinorView<P> field_view(const Field& f, int64_t V) {
  return SpinorView<P>{(typename Store<P>::elem*)f.data, f.norm, V};
}

// view of RHS `rhs` of a batched field (stride in complex elements; norms per V)
template <int P>
__device__ __forceinline__ SpinorView<P> rhs_view(SpinorView<P> v, int rhs, int64_t stride) {
  if (rhs) {
    if (v.data) v.data += rhs * stride;
    if (v.norm) v.norm += (int64_t)rhs * v.V;
  }
  return v;
}

// Input view of RHS `rhs`: rhs_view, unless the epilogue picks the input
// itself (in_view: a slot of a per-RHS ring chosen by that RHS's state)
template <class Epi, int P>
__device__ __forceinline__ SpinorView<P> epi_in_view(const Epi& e, const SpinorView<P>& in, int rhs,
                                                     int64_t stride) {
  if constexpr (requires { e.in_view(in, rhs); })
    return e.in_view(in, rhs);
  else
    return rhs_view(in, rhs, stride);
}

// ---------------------------------------------------------------------------
// Epilogues.  at(rhs) returns the epilogue of one RHS; apply() runs per site
// after the hop is computed and adds into red[] (FP64 block partials).  `in`
// supplies the center value of the site through load(s): the global field in
// the plain kernel, the LDS tile in the tiled one.
// ---------------------------------------------------------------------------
template <int P, int NCOMP>
struct EpiStore {
  using R = typename Store<P>::R;
  static constexpr int NRED = 0;
  SpinorView<P> out;
  int64_t stride;
  EpiStore(const EpiArgs& a, int64_t V) : out(field_view<P>(a.out, V)), stride(a.rhs_stride) {}
  __device__ EpiStore at(int rhs) const {
    EpiStore e = *this;
    e.out = rhs_view(out, rhs, stride);
    return e;
  }
  __device__ bool active() const { return true; }
  template <class CV>
  __device__ void apply(const Lat& lat, int64_t s, const CV& in, const cpx<R> (&hop)[NCOMP],
                        double*) const {
    cpx<R> c[NCOMP], res[NCOMP];
    in.template load<NCOMP>(s, c);
    wilson_finish<NCOMP, R>(lat, c, hop, res);
    out.template store<NCOMP>(s, res);
  }
};

template <int P, int NCOMP>
struct EpiAxpy {
  using R = typename Store<P>::R;
  static constexpr int NRED = 0;
  SpinorView<P> out, w;
  double sigma;
  int64_t stride;
  EpiAxpy(const EpiArgs& a, int64_t V)
      : out(field_view<P>(a.out, V)), w(field_view<P>(a.aux0, V)), sigma(a.sigma), stride(a.rhs_stride) {}
  __device__ EpiAxpy at(int rhs) const {
    EpiAxpy e = *this;
    e.out = rhs_view(out, rhs, stride);
    e.w = rhs_view(w, rhs, stride);
    return e;
  }
  __device__ bool active() const { return true; }
  template <class CV>
  __device__ void apply(const Lat& lat, int64_t s, const CV& in, const cpx<R> (&hop)[NCOMP],
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
};

// Store / axpy gated on a running solve (BiCGStab: launches enqueued after
// the stop test cost only the launch).
template <int P, int NCOMP>
struct EpiGStore : EpiStore<P, NCOMP> {
  SolverState* st;
  EpiGStore(const EpiArgs& a, int64_t V) : EpiStore<P, NCOMP>(a, V), st(a.st) {}
  __device__ EpiGStore at(int) const { return *this; }
  __device__ bool active() const { return st->phase == PH_K1; }
};
template <int P, int NCOMP>
struct EpiGAxpy : EpiAxpy<P, NCOMP> {
  SolverState* st;
  EpiGAxpy(const EpiArgs& a, int64_t V) : EpiAxpy<P, NCOMP>(a, V), st(a.st) {}
  __device__ EpiGAxpy at(int) const { return *this; }
  __device__ bool active() const { return st->phase == PH_K1; }
};

// CG step 1: t = Q p, reduce |t|^2 and |p|^2 (<p,(Q^2+s0)p> = |Qp|^2 + s0|p|^2).
template <int P, int NCOMP>
struct EpiCgmK1 {
  using R = typename Store<P>::R;
  static constexpr int NRED = 2;
  SpinorView<P> out;
  SolverState* st;
  int64_t stride;
  EpiCgmK1(const EpiArgs& a, int64_t V) : out(field_view<P>(a.out, V)), st(a.st), stride(a.rhs_stride) {}
  __device__ EpiCgmK1 at(int rhs) const {
    EpiCgmK1 e = *this;
    e.out = rhs_view(out, rhs, stride);
    e.st = st + rhs;
    return e;
  }
  // plain CG-M enters at PH_K1, the fast variant at PH_KP (after p = r + beta p)
  __device__ bool active() const { return st->phase == PH_K1 || st->phase == PH_KP; }
  template <class CV>
  __device__ void apply(const Lat& lat, int64_t s, const CV& in, const cpx<R> (&hop)[NCOMP],
                        double* red) const {
    cpx<R> c[NCOMP], res[NCOMP];
    in.template load<NCOMP>(s, c);
    wilson_finish<NCOMP, R>(lat, c, hop, res);
    out.template store<NCOMP>(s, res);
    red[0] += norm2<NCOMP, R>(res);
    red[1] += norm2<NCOMP, R>(c);
  }
};

// Plain CG-M step 2: Ap = Q t + s0 p; r <- r - alpha Ap; reduce |r|^2.
template <int P, int NCOMP>
struct EpiCgmK2 {
  using R = typename Store<P>::R;
  static constexpr int NRED = 1;
  SpinorView<P> p, r;
  SolverState* st;
  int64_t stride;
  EpiCgmK2(const EpiArgs& a, int64_t V)
      : p(field_view<P>(a.aux0, V)), r(field_view<P>(a.aux1, V)), st(a.st), stride(a.rhs_stride) {}
  __device__ EpiCgmK2 at(int rhs) const {
    EpiCgmK2 e = *this;
    e.p = rhs_view(p, rhs, stride);
    e.r = rhs_view(r, rhs, stride);
    e.st = st + rhs;
    return e;
  }
  __device__ bool active() const { return st->phase == PH_K2; }
  template <class CV>
  __device__ void apply(const Lat& lat, int64_t s, const CV& in, const cpx<R> (&hop)[NCOMP],
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
};

// Fast CG-M step 2: residuals live in a ring of nslot fields; iteration k reads
// r^k from slot k % nslot and writes r^{k+1} to slot (k+1) % nslot, so the
// last block's worth of residuals stays available for the deferred shift
// updates.  r^{k+1} = r^k - alpha (Q t + s0 p); reduce |r^{k+1}|^2.  16-bit
// rings keep one norm per site and slot: slot q of RHS i at
// norm + (i * nslot + q) * V.
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
  int64_t stride, ring_stride;
  EpiCgmK2F(const EpiArgs& a, int64_t V)
      : p(field_view<P>(a.aux0, V)),
        ring((E*)a.ring.data),
        ring_norm(a.ring.norm),
        slot_elems(a.slot_elems),
        nslot(a.nslot),
        st(a.st),
        stride(a.rhs_stride),
        ring_stride(a.ring_rhs_stride) {}
  __device__ EpiCgmK2F at(int rhs) const {
    EpiCgmK2F e = *this;
    e.p = rhs_view(p, rhs, stride);
    e.ring = ring + rhs * ring_stride;
    if (ring_norm) e.ring_norm = ring_norm + (int64_t)rhs * nslot * p.V;
    e.st = st + rhs;
    return e;
  }
  __device__ bool active() const { return st->phase == PH_K2; }
  template <class CV>
  __device__ void apply(const Lat& lat, int64_t s, const CV& in, const cpx<R> (&hop)[NCOMP],
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
};

// MR step 1: t = Op r (phase gated).
template <int P, int NCOMP>
struct EpiMrK1 {
  using R = typename Store<P>::R;
  static constexpr int NRED = 0;
  SpinorView<P> out;
  SolverState* st;
  EpiMrK1(const EpiArgs& a, int64_t V) : out(field_view<P>(a.out, V)), st(a.st) {}
  __device__ EpiMrK1 at(int) const { return *this; }
  __device__ bool active() const { return st->phase == PH_K1; }
  template <class CV>
  __device__ void apply(const Lat& lat, int64_t s, const CV& in, const cpx<R> (&hop)[NCOMP],
                        double*) const {
    cpx<R> c[NCOMP], res[NCOMP];
    in.template load<NCOMP>(s, c);
    wilson_finish<NCOMP, R>(lat, c, hop, res);
    out.template store<NCOMP>(s, res);
  }
};

// MR step 2: p = Op t + s_N r; reduce <p,r> (complex) and <p,p>.
template <int P, int NCOMP>
struct EpiMrK2 {
  using R = typename Store<P>::R;
  static constexpr int NRED = 3;
  SpinorView<P> out, r;
  SolverState* st;
  EpiMrK2(const EpiArgs& a, int64_t V) : out(field_view<P>(a.out, V)), r(field_view<P>(a.aux0, V)), st(a.st) {}
  __device__ EpiMrK2 at(int) const { return *this; }
  __device__ bool active() const { return st->phase == PH_K1; }
  template <class CV>
  __device__ void apply(const Lat& lat, int64_t s, const CV& in, const cpx<R> (&hop)[NCOMP],
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
};

// The direction ring of the mixed CG: p_k (the direction of step k) of RHS i
// in slot k % nslot at ring + i * ring_stride + (k % nslot) * slot_elems,
// 16-bit norms at norm + (i * nslot + k % nslot) * V.  Keeping a block of
// directions lets the solution update sum_k alpha_k p_k be deferred and
// formed from the very directions the iteration used.
template <int P>
struct RingRef {
  typename Store<P>::elem* ring;
  float* norm;
  int64_t slot_elems, ring_stride;
  int nslot;
  RingRef(const EpiArgs& a)
      : ring((typename Store<P>::elem*)a.ring.data),
        norm(a.ring.norm),
        slot_elems(a.slot_elems),
        ring_stride(a.ring_rhs_stride),
        nslot(a.nslot) {}
  __device__ RingRef at(int rhs, int64_t V) const {
    RingRef r = *this;
    r.ring = ring + rhs * ring_stride;
    if (norm) r.norm = norm + (int64_t)rhs * nslot * V;
    return r;
  }
  __device__ SpinorView<P> slot(int k, int64_t V) const {
    const int q = k % nslot;
    return SpinorView<P>{ring + q * slot_elems, norm ? norm + (int64_t)q * V : nullptr, V};
  }
};

// Mixed CG step 1 (low precision, gate KP): t = Q p_k with p_k the ring slot
// of this RHS's iteration k (the launch's input field is the ring); reduce
// |t|^2, |p|^2.
template <int P, int NCOMP>
struct EpiCgK1R {
  using R = typename Store<P>::R;
  static constexpr int NRED = 2;
  SpinorView<P> out;
  SolverState* st;
  int64_t stride, slot_elems, ring_stride;
  int nslot;
  EpiCgK1R(const EpiArgs& a, int64_t V)
      : out(field_view<P>(a.out, V)), st(a.st), stride(a.rhs_stride), slot_elems(a.slot_elems),
        ring_stride(a.ring_rhs_stride), nslot(a.nslot) {}
  __device__ EpiCgK1R at(int rhs) const {
    EpiCgK1R e = *this;
    e.out = rhs_view(out, rhs, stride);
    e.st = st + rhs;
    return e;
  }
  __device__ bool active() const { return st->phase == PH_KP; }
  __device__ SpinorView<P> in_view(const SpinorView<P>& in, int rhs) const {
    const int q = st->iter % nslot;
    SpinorView<P> v = in;
    v.data += rhs * ring_stride + q * slot_elems;
    if (v.norm) v.norm += ((int64_t)rhs * nslot + q) * in.V;
    return v;
  }
  template <class CV>
  __device__ void apply(const Lat& lat, int64_t s, const CV& in, const cpx<R> (&hop)[NCOMP],
                        double* red) const {
    cpx<R> c[NCOMP], res[NCOMP];
    in.template load<NCOMP>(s, c);
    wilson_finish<NCOMP, R>(lat, c, hop, res);
    out.template store<NCOMP>(s, res);
    red[0] += norm2<NCOMP, R>(res);
    red[1] += norm2<NCOMP, R>(c);
  }
};

// Mixed CG step 2 (low precision): Ap = Q t + s0 p_k; r <- r - alpha Ap;
// reduce |r|^2.  p_k (ring slot k) is read only for a non-zero shift.
template <int P, int NCOMP>
struct EpiCgK2X {
  using R = typename Store<P>::R;
  static constexpr int NRED = 1;
  RingRef<P> pr;
  SpinorView<P> r;
  SolverState* st;
  int64_t stride;
  EpiCgK2X(const EpiArgs& a, int64_t V)
      : pr(a), r(field_view<P>(a.aux1, V)), st(a.st), stride(a.rhs_stride) {}
  __device__ EpiCgK2X at(int rhs) const {
    EpiCgK2X e = *this;
    e.pr = pr.at(rhs, r.V);
    e.r = rhs_view(r, rhs, stride);
    e.st = st + rhs;
    return e;
  }
  __device__ bool active() const { return st->phase == PH_K2; }
  template <class CV>
  __device__ void apply(const Lat& lat, int64_t s, const CV& in, const cpx<R> (&hop)[NCOMP],
                        double* red) const {
    cpx<R> c[NCOMP], res[NCOMP], v[NCOMP];
    in.template load<NCOMP>(s, c);
    wilson_finish<NCOMP, R>(lat, c, hop, res);
    const R s0 = (R)st->sigma0;
    const R al = (R)st->alpha;
    if (s0 != R(0)) {
      cpx<R> pv[NCOMP];
      pr.slot(st->iter, lat.V).template load<NCOMP>(s, pv);
#pragma unroll
      for (int k = 0; k < NCOMP; ++k) res[k] = mk<R>(res[k].re + s0 * pv[k].re, res[k].im + s0 * pv[k].im);
    }
    r.template load<NCOMP>(s, v);
#pragma unroll
    for (int k = 0; k < NCOMP; ++k) v[k] = mk<R>(v[k].re - al * res[k].re, v[k].im - al * res[k].im);
    r.template store<NCOMP>(s, v);
    red[0] += norm2<NCOMP, R>(v);
  }
};

// Reliable update, first half (complex128, phase gated): t = Q x_hi.
template <int P, int NCOMP>
struct EpiRu1 {
  using R = typename Store<P>::R;
  static constexpr int NRED = 0;
  SpinorView<P> out;
  SolverState* st;
  int64_t stride;
  EpiRu1(const EpiArgs& a, int64_t V) : out(field_view<P>(a.out, V)), st(a.st), stride(a.rhs_stride) {}
  __device__ EpiRu1 at(int rhs) const {
    EpiRu1 e = *this;
    e.out = rhs_view(out, rhs, stride);
    e.st = st + rhs;
    return e;
  }
  __device__ bool active() const { return st->phase == PH_RU; }
  template <class CV>
  __device__ void apply(const Lat& lat, int64_t s, const CV& in, const cpx<R> (&hop)[NCOMP],
                        double*) const {
    cpx<R> c[NCOMP], res[NCOMP];
    in.template load<NCOMP>(s, c);
    wilson_finish<NCOMP, R>(lat, c, hop, res);
    out.template store<NCOMP>(s, res);
  }
};

// Reliable update, second half (complex128): r = b - (Q t + s0 x_hi), the
// true residual; reduce |r|^2 in FP64 and store r at the low precision PL
// (complex64 or half + per-site norm) as the new iterated residual; reduce
// <r, p> with the last direction (ring slot iter - 1).  b of RHS i is bs[i]
// (caller's arrays).
template <int P, int NCOMP, int PL>
struct EpiRu2Base {
  using R = typename Store<P>::R;
  static constexpr int NRED = 3;  // |r|^2, <r, p> (re, im)
  SpinorView<P> x, b;
  SpinorView<PL> rlo;
  RingRef<PL> pr;
  const Field* bs;
  SolverState* st;
  int64_t stride;
  EpiRu2Base(const EpiArgs& a, int64_t V)
      : x(field_view<P>(a.aux0, V)), b(field_view<P>(Field{}, V)), rlo(field_view<PL>(a.out, V)), pr(a),
        bs(a.bs), st(a.st), stride(a.rhs_stride) {}
  __device__ EpiRu2Base at(int rhs) const {
    EpiRu2Base e = *this;
    e.x = rhs_view(x, rhs, stride);
    e.rlo = rhs_view(rlo, rhs, stride);
    e.pr = pr.at(rhs, x.V);
    e.b.data = (typename Store<P>::elem*)bs[rhs].data;
    e.st = st + rhs;
    return e;
  }
  __device__ bool active() const { return st->phase == PH_RU; }
  template <class CV>
  __device__ void apply(const Lat& lat, int64_t s, const CV& in, const cpx<R> (&hop)[NCOMP],
                        double* red) const {
    cpx<R> c[NCOMP], res[NCOMP], xv[NCOMP];
    in.template load<NCOMP>(s, c);
    x.template load<NCOMP>(s, xv);
    wilson_finish<NCOMP, R>(lat, c, hop, res);
    const R s0 = (R)st->sigma0;
    b.template load<NCOMP>(s, c);
    cpx<float> lo[NCOMP], pv[NCOMP];
    pr.slot(st->iter + pr.nslot - 1, lat.V).template load<NCOMP>(s, pv);
#pragma unroll
    for (int k = 0; k < NCOMP; ++k) {
      const R rr = c[k].re - (res[k].re + s0 * xv[k].re);
      const R ri = c[k].im - (res[k].im + s0 * xv[k].im);
      red[0] += (double)rr * rr + (double)ri * ri;
      red[1] += (double)rr * pv[k].re + (double)ri * pv[k].im;  // <r, p> = sum conj(r) p
      red[2] += (double)rr * pv[k].im - (double)ri * pv[k].re;
      lo[k] = mk<float>((float)rr, (float)ri);
    }
    rlo.template store<NCOMP>(s, lo);
  }
};
template <int P, int NCOMP>
using EpiRu2 = EpiRu2Base<P, NCOMP, P32>;
template <int P, int NCOMP>
using EpiRu2H = EpiRu2Base<P, NCOMP, P16>;

// ---------------------------------------------------------------------------
// The kernel.  gs = site blocks; the grid has gs * nrhs blocks.
// ---------------------------------------------------------------------------
template <int D, int NC, bool ADJ, bool PROJ, int P, class Epi>
__global__ __launch_bounds__(kBlock) void k_wilson(Lat lat, SiteMap map, SpinorView<P> in, int64_t in_stride,
                                                   LinkView<typename Store<P>::R> lk,
                                                   LinkView<typename Store<P>::R> lkr, Epi epi, RedCtl rc, int nrhs,
                                                   int gs) {
  using R = typename Store<P>::R;
  constexpr int NCOMP = Gam<D>::NS * (ADJ ? NC * NC : NC);
  int sblk = blockIdx.x, rhs = 0;
  if (nrhs > 1) batch_block(nrhs, sblk, rhs);
  const Epi e = epi.at(rhs);
  if (!e.active()) return;
  const SpinorView<P> inr = epi_in_view(e, in, rhs, in_stride);
  constexpr int NR = Epi::NRED > 0 ? Epi::NRED : 1;
  double red[NR];
#pragma unroll
  for (int r = 0; r < NR; ++r) red[r] = 0.0;
  int64_t s;
  int x[D];
  if (site_of_thread<D>(lat, map, sblk, gs, s, x)) {
    cpx<R> hop[NCOMP];
    wilson_hop_x<D, NC, ADJ, PROJ, R>(lat, s, x, inr, lk, lkr, hop);
    e.apply(lat, s, inr, hop, red);
  }
  if constexpr (Epi::NRED > 0) block_partials<NR>(red, rc.partials + (int64_t)rhs * NR * gs, sblk, gs);
}

// ---------------------------------------------------------------------------
// LDS-tiled kernel (4D, complex128/complex64; fundamental and adjoint
// spinors, the adjoint with one hop in flight): a block owns a tile of
// a0 x a1 x k x T sites (T = L_3, whole lines), one site per thread, thread
// id = ((t0 * a1 + t1) * k + t2) * T + t3.  The block first copies the input
// spinor of its tile into LDS (one coalesced read per site), then every hop
// whose neighbour lies inside the tile reads LDS; only hops that leave the
// tile read global memory.  Global spinor reads per site drop from 1 + 2D to
// 1 + sum_{mu<3} 2/a_mu (a direction the tile spans completely costs none).
// ---------------------------------------------------------------------------
// 16-bit storage keeps the tile raw in LDS: the packed integer units
// [NCOMP / G][nsite] (4G bytes each, as in global memory), then the per-site
// norms [nsite]; 52 B per site at NCOMP = 12 against 96 decoded.
template <int P, int NCOMP>
__host__ __device__ constexpr size_t tile_lds_bytes(int nsite) {
  if constexpr (P == P16)
    return (size_t)nsite * (NCOMP * 4 + 4);
  else
    return (size_t)nsite * NCOMP * sizeof(typename Store<P>::lds_elem);
}

template <int P>
struct TileView {
  using R = typename Store<P>::R;
  using E = typename Store<P>::lds_elem;
  SpinorView<P> g;  // the field in global memory (hops leaving the tile)
  const E* lds;     // [NCOMP][nsite] decoded (complex128 / complex64), raw for 16-bit
  int nsite, self;  // tile size, this thread's tile index
  int t[4];         // tile coordinates of this thread
  int a[4];         // tile extents
  int str[4];       // tile index strides
  bool full[4];     // the tile spans the whole lattice in mu (wrap inside)

  // 16-bit storage: the integers of tile site idx as floats; returns its norm
  template <int NCOMP>
  __device__ __forceinline__ R raw16(int idx, cpx<R> (&v)[NCOMP]) const {
    using U = typename Pack16<NCOMP>::unit;
    constexpr int G = Pack16<NCOMP>::G;
    const U* lu = reinterpret_cast<const U*>(lds);
#pragma unroll
    for (int q = 0; q < NCOMP / G; ++q) {
      const U u = lu[q * nsite + idx];
#pragma unroll
      for (int gg = 0; gg < G; ++gg) v[q * G + gg] = mk<R>((R)u.v[2 * gg], (R)u.v[2 * gg + 1]);
    }
    return reinterpret_cast<const float*>(lu + (NCOMP / G) * nsite)[idx];
  }
  template <int NCOMP>
  __device__ __forceinline__ void tile_value(int idx, cpx<R> (&v)[NCOMP]) const {
    if constexpr (P == P16) {
      const R n = raw16<NCOMP>(idx, v);
#pragma unroll
      for (int c = 0; c < NCOMP; ++c) v[c] = mk<R>(v[c].re * n, v[c].im * n);
    } else {
#pragma unroll
      for (int c = 0; c < NCOMP; ++c) {
        const E e = lds[c * nsite + idx];
        v[c] = mk<R>(e.x, e.y);
      }
    }
  }
  // center value: this thread's own site
  template <int NCOMP>
  __device__ __forceinline__ void load(int64_t, cpx<R> (&v)[NCOMP]) const {
    tile_value<NCOMP>(self, v);
  }
  // neighbour in direction mu (dir 0: +mu, 1: -mu); nb is its global site
  template <int NCOMP>
  __device__ __forceinline__ void load_hop(int mu, int dir, int64_t nb, cpx<R> (&v)[NCOMP]) const {
    int tc = t[mu] + (dir == 0 ? 1 : -1);
    bool inside = true;
    if (tc < 0) {
      tc += a[mu];
      inside = full[mu];
    } else if (tc >= a[mu]) {
      tc -= a[mu];
      inside = full[mu];
    }
    if (inside) {
      tile_value<NCOMP>(self + (tc - t[mu]) * str[mu], v);
    } else {
      g.template load<NCOMP>(nb, v);
    }
  }
  // 16-bit storage: a neighbour comes as integers and its norm (value =
  // v * n), from the raw LDS tile or from global memory
  static constexpr bool kScaled = SpinorView<P>::kScaled;
  template <int NCOMP>
  __device__ __forceinline__ void load_hop_scaled(int mu, int dir, int64_t nb, cpx<R> (&v)[NCOMP], R& n) const {
    int tc = t[mu] + (dir == 0 ? 1 : -1);
    bool inside = true;
    if (tc < 0) {
      tc += a[mu];
      inside = full[mu];
    } else if (tc >= a[mu]) {
      tc -= a[mu];
      inside = full[mu];
    }
    if (inside) {
      const int idx = self + (tc - t[mu]) * str[mu];
      if constexpr (P == P16) {
        n = raw16<NCOMP>(idx, v);
      } else {
        tile_value<NCOMP>(idx, v);
        n = R(1);
      }
    } else {
      g.template load_raw<NCOMP>(nb, v, n);
    }
  }
};

template <int NC, bool ADJ, int P, class Epi, int HIF>
__global__ __launch_bounds__(ADJ ? kTileAdjThreads
                                 : HIF == 1 ? (P == P64 ? kTileMaxThreads64 : kTileMaxThreads)
                                            : (P == P64 ? kTileHifThreads64 : kTileHifThreads32),
                             ADJ ? 1 : HIF == 1 ? 3 : 2) void k_wilson_tile(Lat lat, TileMap tm, SpinorView<P> in,
                                                                 int64_t in_stride,
                                                                 LinkView<typename Store<P>::R> lk,
                                                                 LinkView<typename Store<P>::R> lkr, Epi epi,
                                                                 RedCtl rc, int nrhs) {
  using R = typename Store<P>::R;
  using E = typename Store<P>::lds_elem;
  constexpr int NCOMP = Gam<4>::NS * (ADJ ? NC * NC : NC);
  extern __shared__ __align__(16) unsigned char tile_smem[];
  int tblk = blockIdx.x, rhs = 0;
  if (nrhs > 1) batch_block(nrhs, tblk, rhs);
  const Epi e = epi.at(rhs);
  if (!e.active()) return;  // uniform per block (state of one RHS)
  const SpinorView<P> inr = epi_in_view(e, in, rhs, in_stride);
  // tile origin: XCD slab, then x0 fastest, x2, x1 (walk 1: x0, x1, x2)
  const int xcd = tblk % kNumXcd;
  const int j = tblk / kNumXcd;
  const int i0 = j % tm.n0;
  const int r = j / tm.n0;
  const int i2 = tm.walk ? r / tm.n1 : r % tm.n2;
  const int i1 = tm.walk ? r % tm.n1 : r / tm.n2;
  const int tid = (int)threadIdx.x;
  int q = tid / tm.T;
  const int t3 = tid - q * tm.T;
  const int t2 = q % tm.k;
  q /= tm.k;
  const int t1 = q % tm.a1;
  const int t0 = q / tm.a1;
  int x[4];
  x[0] = xcd * tm.slab + i0 * tm.a0 + t0;
  x[1] = i1 * tm.a1 + t1;
  x[2] = i2 * tm.k + t2;
  x[3] = t3;
  const int64_t s = (((int64_t)x[0] * lat.L[1] + x[1]) * lat.L[2] + x[2]) * tm.T + x[3];
  E* lds = reinterpret_cast<E*>(tile_smem);
  if constexpr (P == P16) {
    using U = typename Pack16<NCOMP>::unit;
    constexpr int G = Pack16<NCOMP>::G;
    U u[NCOMP / G];
    float n;
    inr.template load_packed<NCOMP>(s, u, n);
    U* lu = reinterpret_cast<U*>(tile_smem);
#pragma unroll
    for (int k = 0; k < NCOMP / G; ++k) lu[k * tm.nsite + tid] = u[k];
    reinterpret_cast<float*>(lu + (NCOMP / G) * tm.nsite)[tid] = n;
  } else {
    cpx<R> c[NCOMP];
    inr.template load<NCOMP>(s, c);
#pragma unroll
    for (int k = 0; k < NCOMP; ++k) {
      E v;
      v.x = c[k].re;
      v.y = c[k].im;
      lds[k * tm.nsite + tid] = v;
    }
  }
  __syncthreads();
  TileView<P> view;
  view.g = inr;
  view.lds = lds;
  view.nsite = tm.nsite;
  view.self = tid;
  view.t[0] = t0;
  view.t[1] = t1;
  view.t[2] = t2;
  view.t[3] = t3;
  view.a[0] = tm.a0;
  view.a[1] = tm.a1;
  view.a[2] = tm.k;
  view.a[3] = tm.T;
  view.str[3] = 1;
  view.str[2] = tm.T;
  view.str[1] = tm.k * tm.T;
  view.str[0] = tm.a1 * tm.k * tm.T;
  view.full[0] = tm.a0 == lat.L[0];
  view.full[1] = tm.a1 == lat.L[1];
  view.full[2] = tm.k == lat.L[2];
  view.full[3] = true;
  constexpr int NR = Epi::NRED > 0 ? Epi::NRED : 1;
  double red[NR];
#pragma unroll
  for (int i = 0; i < NR; ++i) red[i] = 0.0;
  cpx<R> hop[NCOMP];
  wilson_hop_x<4, NC, ADJ, true, R, TileView<P>, HIF>(lat, s, x, view, lk, lkr, hop);
  e.apply(lat, s, view, hop, red);
  if constexpr (Epi::NRED > 0)
    block_partials<NR, kTileMaxThreads / 64>(red, rc.partials + (int64_t)rhs * NR * tm.blocks, tblk, tm.blocks);
}

// ---------------------------------------------------------------------------
// Multi-RHS LDS-tiled kernel (4D fundamental, complex64): a block owns a tile
// and a group of kMrRhs consecutive right-hand sides, stages all of their
// spinors in LDS, and applies every link it loads to each of them, so link
// traffic per RHS drops by the group size.  Per RHS the arithmetic is the
// single-RHS kernel's (same order), hence the same result.  Blocks of one
// tile's groups run back to back on one XCD.
// ---------------------------------------------------------------------------
template <class T, int N>
struct Arr {
  T v[N];
};

template <int NC, int P, class Epi, size_t... I>
__device__ __forceinline__ Arr<Epi, sizeof...(I)> group_epis(const Epi& epi, int rhs0, int nrhs,
                                                             std::index_sequence<I...>) {
  return Arr<Epi, sizeof...(I)>{{epi.at(min(rhs0 + (int)I, nrhs - 1))...}};
}

template <int NC, int P, class Epi>
__global__ __launch_bounds__(kTileMaxThreadsMr, 2) void k_wilson_tile_mr(Lat lat, TileMap tm, SpinorView<P> in,
                                                                       int64_t in_stride,
                                                                       LinkView<typename Store<P>::R> lk, Epi epi,
                                                                       RedCtl rc, int nrhs, int ngroups) {
  using R = typename Store<P>::R;
  using E = typename Store<P>::lds_elem;
  constexpr int NCOMP = Gam<4>::NS * NC;
  constexpr int RB = kMrRhs;
  extern __shared__ __align__(16) unsigned char tile_smem[];
  const int b = blockIdx.x;
  const int bx = b % kNumXcd;
  const int jj = b / kNumXcd;
  const int grp = jj % ngroups;
  const int tblk = (jj / ngroups) * kNumXcd + bx;
  const int rhs0 = grp * RB;
  const Arr<Epi, RB> e = group_epis<NC, P>(epi, rhs0, nrhs, std::make_index_sequence<RB>{});
  bool act[RB];
  bool any = false;
#pragma unroll
  for (int r = 0; r < RB; ++r) {
    act[r] = (rhs0 + r < nrhs) && e.v[r].active();  // uniform per block
    any = any || act[r];
  }
  if (!any) return;
  // tile origin and thread coordinates as in k_wilson_tile
  const int xcd = tblk % kNumXcd;
  const int j = tblk / kNumXcd;
  const int i0 = j % tm.n0;
  const int rr = j / tm.n0;
  const int i2 = tm.walk ? rr / tm.n1 : rr % tm.n2;
  const int i1 = tm.walk ? rr % tm.n1 : rr / tm.n2;
  const int tid = (int)threadIdx.x;
  int q = tid / tm.T;
  const int t3 = tid - q * tm.T;
  const int t2 = q % tm.k;
  q /= tm.k;
  const int t1 = q % tm.a1;
  const int t0 = q / tm.a1;
  int x[4];
  x[0] = xcd * tm.slab + i0 * tm.a0 + t0;
  x[1] = i1 * tm.a1 + t1;
  x[2] = i2 * tm.k + t2;
  x[3] = t3;
  const int64_t s = (((int64_t)x[0] * lat.L[1] + x[1]) * lat.L[2] + x[2]) * tm.T + x[3];
  E* lds = reinterpret_cast<E*>(tile_smem);
  TileView<P> view[RB];
#pragma unroll
  for (int r = 0; r < RB; ++r) {
    view[r].g = epi_in_view(e.v[r], in, min(rhs0 + r, nrhs - 1), in_stride);
    view[r].lds = lds + (int64_t)r * NCOMP * tm.nsite;
    view[r].nsite = tm.nsite;
    view[r].self = tid;
    view[r].t[0] = t0;
    view[r].t[1] = t1;
    view[r].t[2] = t2;
    view[r].t[3] = t3;
    view[r].a[0] = tm.a0;
    view[r].a[1] = tm.a1;
    view[r].a[2] = tm.k;
    view[r].a[3] = tm.T;
    view[r].str[3] = 1;
    view[r].str[2] = tm.T;
    view[r].str[1] = tm.k * tm.T;
    view[r].str[0] = tm.a1 * tm.k * tm.T;
    view[r].full[0] = tm.a0 == lat.L[0];
    view[r].full[1] = tm.a1 == lat.L[1];
    view[r].full[2] = tm.k == lat.L[2];
    view[r].full[3] = true;
    if (act[r]) {
      cpx<R> c[NCOMP];
      view[r].g.template load<NCOMP>(s, c);
#pragma unroll
      for (int k = 0; k < NCOMP; ++k) {
        E v;
        v.x = c[k].re;
        v.y = c[k].im;
        lds[((int64_t)r * NCOMP + k) * tm.nsite + tid] = v;
      }
    }
  }
  __syncthreads();
  cpx<R> acc[RB][NCOMP];
#pragma unroll
  for (int r = 0; r < RB; ++r)
#pragma unroll
    for (int c = 0; c < NCOMP; ++c) acc[r][c] = czero<R>();
#pragma unroll
  for (int h = 0; h < 8; ++h) {
    const int mu = h >> 1, dir = h & 1;
    const int L = lat.L[mu];
    const int64_t st = lat.stride[mu];
    int64_t nb = dir == 0 ? ((x[mu] == L - 1) ? s - (int64_t)(L - 1) * st : s + st)
                          : ((x[mu] == 0) ? s + (int64_t)(L - 1) * st : s - st);
    int64_t ls = dir == 0 ? s : nb;
#pragma unroll
    for (int r = 0; r < RB; ++r) hop_commit<NCOMP, R>(acc[r]);
    hop_dep<R>(nb, ls, acc[0][0].re);
    cpx<R> u[NC][NC];
    lk.template load<NC>(mu, ls, u);
#pragma unroll
    for (int r = 0; r < RB; ++r) {
      if (act[r]) {
        cpx<R> v[NCOMP];
        view[r].template load_hop<NCOMP>(mu, dir, nb, v);
        wilson_hop_apply<4, NC, false, true, R>(lat, mu, dir, x[mu], v, u, u, acc[r]);
      }
    }
  }
  constexpr int NR = Epi::NRED > 0 ? Epi::NRED : 1;
#pragma unroll
  for (int r = 0; r < RB; ++r) {
    if (!act[r]) continue;
    double red[NR];
#pragma unroll
    for (int i = 0; i < NR; ++i) red[i] = 0.0;
    e.v[r].apply(lat, s, view[r], acc[r], red);
    if constexpr (Epi::NRED > 0) {
      block_partials<NR, kTileMaxThreadsMr / 64>(red, rc.partials + (int64_t)(rhs0 + r) * NR * tm.blocks, tblk,
                                               tm.blocks);
      __syncthreads();  // the reduction's LDS is reused by the next RHS
    }
  }
}

// ---------------------------------------------------------------------------
// Dispatch: runtime (kind, ndim, nc, rep, prec, mode) -> template instance.
// ---------------------------------------------------------------------------
template <int D, int NC, bool ADJ, int P, template <int, int> class EpiT>
static int launch_one(const OpDesc& d, bool proj, Field in, const EpiArgs& a, RedCtl rc, hipStream_t stream,
                      int nrhs) {
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
  const int recon = (!ADJ && d.link_su) ? 1 : 0;
  LinkView<R> lk{(const cpx<R>*)l, V, recon};
  LinkView<R> lkr{(const cpx<R>*)lr, V, recon};
  SpinorView<P> sv{(typename Store<P>::elem*)in.data, in.norm, V};
  Epi epi(a, V);
  if constexpr (D == 4 && !ADJ && P == P32) {
    if (proj && wilson_mr(d, P, nrhs)) {
      const TileMap& tm = d.tile_mr;
      const int ngroups = (nrhs + kMrRhs - 1) / kMrRhs;
      const size_t lds = (size_t)kMrRhs * NCOMP * tm.nsite * sizeof(typename Store<P>::lds_elem);
      const dim3 grid((unsigned)tm.blocks * (unsigned)ngroups), block(tm.nsite);
      hipLaunchKernelGGL((k_wilson_tile_mr<NC, P, Epi>), grid, block, lds, stream, d.lat, tm, sv, a.rhs_stride, lk,
                         epi, rc, nrhs, ngroups);
      LYM_HIP(hipGetLastError());
      return LYM_OK;
    }
  }
  if constexpr (D == 4 && ADJ) {
    // adjoint: the same tile kernel with one hop in flight (register budget)
    if (proj && wilson_tiled(d, P)) {
      const TileMap& tm = d.tile[P == P64 ? 0 : 1];
      const size_t lds = (size_t)NCOMP * tm.nsite * sizeof(typename Store<P>::lds_elem);
      const dim3 grid((unsigned)tm.blocks * (unsigned)nrhs), block(tm.nsite);
      if (tm.nsite > kTileAdjThreads) {
        set_error("adjoint LDS tile larger than its block limit");
        return LYM_ERR_ARG;
      }
      if (lds > 64 * 1024) {
        static std::atomic<uint32_t> raised;
        const int rc = raise_lds_limit((const void*)k_wilson_tile<NC, ADJ, P, Epi, 1>, (int)kTileMaxLds, raised);
        if (rc) return rc;
      }
      hipLaunchKernelGGL((k_wilson_tile<NC, ADJ, P, Epi, 1>), grid, block, lds, stream, d.lat, tm, sv, a.rhs_stride,
                         lk, lkr, epi, rc, nrhs);
      LYM_HIP(hipGetLastError());
      return LYM_OK;
    }
  }
  if constexpr (D == 4 && !ADJ) {
    if (proj && wilson_tiled(d, P)) {
      const TileMap& tm = d.tile[P == P64 ? 0 : 1];
      const size_t lds = tile_lds_bytes<P, NCOMP>(tm.nsite);
      const dim3 grid((unsigned)tm.blocks * (unsigned)nrhs), block(tm.nsite);
      // hops in flight: 4 is built for complex64 only (complex128 would
      // spill); more than one needs blocks of at most 512 (complex128: 320)
      // threads
      int hif = (P != P32 && tm.hif == 4) ? 2 : tm.hif;
      if (tm.nsite > tile_hif_threads(P)) hif = 1;
      if (lds > 64 * 1024) {
        // above the default dynamic-LDS limit: raise it once per kernel and device
        static std::atomic<uint32_t> raised[3];
        const int hi = hif == 4 ? 2 : hif - 1;
        const void* fn = hif == 4   ? (const void*)k_wilson_tile<NC, ADJ, P, Epi, (P == P32 ? 4 : 1)>
                         : hif == 2 ? (const void*)k_wilson_tile<NC, ADJ, P, Epi, 2>
                                    : (const void*)k_wilson_tile<NC, ADJ, P, Epi, 1>;
        const int rc = raise_lds_limit(fn, (int)kTileMaxLds, raised[hi]);
        if (rc) return rc;
      }
      if (hif == 4) {
        if constexpr (P == P32)
          hipLaunchKernelGGL((k_wilson_tile<NC, ADJ, P, Epi, 4>), grid, block, lds, stream, d.lat, tm, sv,
                             a.rhs_stride, lk, lkr, epi, rc, nrhs);
      } else if (hif == 2) {
        hipLaunchKernelGGL((k_wilson_tile<NC, ADJ, P, Epi, 2>), grid, block, lds, stream, d.lat, tm, sv, a.rhs_stride,
                           lk, lkr, epi, rc, nrhs);
      } else {
        hipLaunchKernelGGL((k_wilson_tile<NC, ADJ, P, Epi, 1>), grid, block, lds, stream, d.lat, tm, sv, a.rhs_stride,
                           lk, lkr, epi, rc, nrhs);
      }
      LYM_HIP(hipGetLastError());
      return LYM_OK;
    }
  }
  const dim3 grid((unsigned)d.wblocks * (unsigned)nrhs), block(d.wthreads);
  if (proj) {
    hipLaunchKernelGGL((k_wilson<D, NC, ADJ, true, P, Epi>), grid, block, 0, stream, d.lat, d.map, sv, a.rhs_stride,
                       lk, lkr, epi, rc, nrhs, d.wblocks);
  } else {
    if constexpr (P == P64 && std::is_same<Epi, EpiStore<P, NCOMP>>::value) {
      hipLaunchKernelGGL((k_wilson<D, NC, ADJ, false, P, Epi>), grid, block, 0, stream, d.lat, d.map, sv,
                         a.rhs_stride, lk, lkr, epi, rc, nrhs, d.wblocks);
    } else {
      set_error("simple (full-projector) mode is only built for the plain complex128 apply");
      return LYM_ERR_UNSUPPORTED;
    }
  }
  LYM_HIP(hipGetLastError());
  return LYM_OK;
}

template <int P, template <int, int> class EpiT>
static int launch_geom(const OpDesc& d, bool proj, Field in, const EpiArgs& a, RedCtl rc, hipStream_t st,
                       int nrhs) {
  const int key = d.ndim * 100 + d.nc * 10 + d.adj;
  if constexpr (P == P16) {
    // half-precision storage is built for the fundamental representation only
    switch (key) {
      case 420: return launch_one<4, 2, false, P, EpiT>(d, proj, in, a, rc, st, nrhs);
      case 430: return launch_one<4, 3, false, P, EpiT>(d, proj, in, a, rc, st, nrhs);
      case 220: return launch_one<2, 2, false, P, EpiT>(d, proj, in, a, rc, st, nrhs);
      case 230: return launch_one<2, 3, false, P, EpiT>(d, proj, in, a, rc, st, nrhs);
      default:
        set_error("half-precision storage: fundamental representation only");
        return LYM_ERR_UNSUPPORTED;
    }
  } else {
    switch (key) {
      case 420: return launch_one<4, 2, false, P, EpiT>(d, proj, in, a, rc, st, nrhs);
      case 430: return launch_one<4, 3, false, P, EpiT>(d, proj, in, a, rc, st, nrhs);
      case 220: return launch_one<2, 2, false, P, EpiT>(d, proj, in, a, rc, st, nrhs);
      case 230: return launch_one<2, 3, false, P, EpiT>(d, proj, in, a, rc, st, nrhs);
      case 421: return launch_one<4, 2, true, P, EpiT>(d, proj, in, a, rc, st, nrhs);
      case 431: return launch_one<4, 3, true, P, EpiT>(d, proj, in, a, rc, st, nrhs);
      case 221: return launch_one<2, 2, true, P, EpiT>(d, proj, in, a, rc, st, nrhs);
      case 231: return launch_one<2, 3, true, P, EpiT>(d, proj, in, a, rc, st, nrhs);
      default:
        set_error("unsupported (ndim, nc, representation)");
        return LYM_ERR_UNSUPPORTED;
    }
  }
}

template <template <int, int> class EpiT>
static int launch_prec(int prec, const OpDesc& d, bool proj, Field in, const EpiArgs& a, RedCtl rc, hipStream_t st,
                       int nrhs) {
  switch (prec) {
    case P64: return launch_geom<P64, EpiT>(d, proj, in, a, rc, st, nrhs);
    case P32: return launch_geom<P32, EpiT>(d, proj, in, a, rc, st, nrhs);
    case P16: return launch_geom<P16, EpiT>(d, proj, in, a, rc, st, nrhs);
    default:
      set_error("unknown precision");
      return LYM_ERR_ARG;
  }
}

int wilson_launch(EpiKind kind, const OpDesc& d, int prec, bool proj, Field in, const EpiArgs& a, RedCtl rc,
                  hipStream_t stream, int nrhs) {
  if (nrhs < 1) {
    set_error("batch size must be >= 1");
    return LYM_ERR_ARG;
  }
  switch (kind) {
    case EPI_STORE: return launch_prec<EpiStore>(prec, d, proj, in, a, rc, stream, nrhs);
    case EPI_AXPY: return launch_prec<EpiAxpy>(prec, d, proj, in, a, rc, stream, nrhs);
    case EPI_CGM_K1: return launch_prec<EpiCgmK1>(prec, d, proj, in, a, rc, stream, nrhs);
    case EPI_CGM_K2: return launch_prec<EpiCgmK2>(prec, d, proj, in, a, rc, stream, nrhs);
    case EPI_CGM_K2F: return launch_prec<EpiCgmK2F>(prec, d, proj, in, a, rc, stream, nrhs);
    case EPI_CG_K2X: return launch_prec<EpiCgK2X>(prec, d, proj, in, a, rc, stream, nrhs);
    case EPI_CG_K1R:
      if (prec == P64) {
        set_error("the mixed CG's direction ring is complex64 or 16-bit");
        return LYM_ERR_UNSUPPORTED;
      }
      if (prec == P16) return launch_geom<P16, EpiCgK1R>(d, proj, in, a, rc, stream, nrhs);
      return launch_geom<P32, EpiCgK1R>(d, proj, in, a, rc, stream, nrhs);
    case EPI_RU1:
    case EPI_RU2:
    case EPI_RU2H:
      if (prec != P64) {
        set_error("the reliable update runs in complex128");
        return LYM_ERR_UNSUPPORTED;
      }
      if (kind == EPI_RU1) return launch_geom<P64, EpiRu1>(d, proj, in, a, rc, stream, nrhs);
      if (kind == EPI_RU2H) return launch_geom<P64, EpiRu2H>(d, proj, in, a, rc, stream, nrhs);
      return launch_geom<P64, EpiRu2>(d, proj, in, a, rc, stream, nrhs);
    case EPI_GSTORE:
    case EPI_GAXPY:
      if (prec != P64 || nrhs != 1) {
        set_error("BiCGStab runs in complex128, one RHS per call");
        return LYM_ERR_UNSUPPORTED;
      }
      if (kind == EPI_GSTORE) return launch_geom<P64, EpiGStore>(d, proj, in, a, rc, stream, 1);
      return launch_geom<P64, EpiGAxpy>(d, proj, in, a, rc, stream, 1);
    case EPI_MR_K1:
    case EPI_MR_K2:
      if (prec != P64 || nrhs != 1) {
        set_error("the multishift MR solver runs in complex128, one RHS per call");
        return LYM_ERR_UNSUPPORTED;
      }
      if (kind == EPI_MR_K1) return launch_geom<P64, EpiMrK1>(d, proj, in, a, rc, stream, 1);
      return launch_geom<P64, EpiMrK2>(d, proj, in, a, rc, stream, 1);
  }
  set_error("unknown epilogue");
  return LYM_ERR_ARG;
}

}  // namespace lym
