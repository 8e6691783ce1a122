// Krylov solvers on the device.
//
// Every iteration is a fixed launch sequence (two Wilson kernels with fused
// epilogues + one per-site vector kernel); all scalars live in SolverState
// and are advanced by the last block of each reducing kernel, so the host
// only enqueues chunks of iterations and polls one pinned flag per chunk.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <type_traits>
#include <vector>

#include "lym_fin.h"
#include "lym_internal.h"

namespace lym {

template <int N>
using IC = std::integral_constant<int, N>;

template <class F>
static int with_ncomp(int ncomp, F&& f) {
  switch (ncomp) {
    case 4: return f(IC<4>{});
    case 6: return f(IC<6>{});
    case 8: return f(IC<8>{});
    case 12: return f(IC<12>{});
    case 16: return f(IC<16>{});
    case 18: return f(IC<18>{});
    case 36: return f(IC<36>{});
  }
  set_error("unsupported number of spinor components per site");
  return LYM_ERR_UNSUPPORTED;
}

template <class F>
static int with_prec_ncomp(int prec, int ncomp, F&& f) {
  switch (prec) {
    case P64: return with_ncomp(ncomp, [&](auto N) { return f(IC<P64>{}, N); });
    case P32: return with_ncomp(ncomp, [&](auto N) { return f(IC<P32>{}, N); });
    case P16:
      return with_ncomp(ncomp, [&](auto N) {
        if constexpr (decltype(N)::value <= 12) return f(IC<P16>{}, N);
        set_error("half-precision storage: fundamental representation only");
        return (int)LYM_ERR_UNSUPPORTED;
      });
  }
  set_error("unknown precision");
  return LYM_ERR_ARG;
}

template <int P>
__device__ __forceinline__ SpinorView<P> view(const Field& f, int64_t V) {
  return SpinorView<P>{(typename Store<P>::elem*)f.data, f.norm, V};
}

// ---------------------------------------------------------------------------
// CG-M vector kernels
// ---------------------------------------------------------------------------
template <int P, int NCOMP>
__global__ __launch_bounds__(kBlock) void k_cgm_init(SolverState* st, Field b, Field r, const Field* xs,
                                                     const Field* ps, int nshift, int64_t V, RedCtl rc) {
  using R = typename Store<P>::R;
  const int64_t s = xcd_block() * kBlock + threadIdx.x;
  double red[1] = {0.0};
  if (s < V) {
    cpx<R> v[NCOMP], z[NCOMP];
    view<P>(b, V).template load<NCOMP>(s, v);
    view<P>(r, V).template store<NCOMP>(s, v);
#pragma unroll
    for (int c = 0; c < NCOMP; ++c) {
      z[c] = czero<R>();
      red[0] += (double)v[c].re * v[c].re + (double)v[c].im * v[c].im;
    }
    for (int k = 0; k < nshift; ++k) {
      view<P>(ps[k], V).template store<NCOMP>(s, v);
      view<P>(xs[k], V).template store<NCOMP>(s, z);
    }
  }
  double sums[1];
  if (block_reduce_last<1>(red, rc, sums) && threadIdx.x == 0) {
    st->bb = sums[0];
    st->rr = sums[0];
    for (int k = 0; k < st->nshift; ++k) st->tol2_s[k] *= sums[0];
    if (sums[0] > 0.0) {
      st->phase = PH_K1;
    } else {
      st->status = ST_CONVERGED;
      st->phase = PH_DONE;
    }
  }
}

// x_s += alpha_s p_s ; p_s = zeta_s r + beta_s p_s   for every shift updated this iteration
template <int P, int NCOMP>
__global__ __launch_bounds__(kBlock) void k_cgm_k3(SolverState* st, Field r, const Field* xs, const Field* ps,
                                                   int nshift, int64_t V, unsigned* ticket) {
  using R = typename Store<P>::R;
  if (st->phase != PH_K3) return;
  const int64_t s = xcd_block() * kBlock + threadIdx.x;
  if (s < V) {
    cpx<R> rv[NCOMP];
    view<P>(r, V).template load<NCOMP>(s, rv);
    for (int k = 0; k < nshift; ++k) {
      if (!st->upd[k]) continue;
      const R a = (R)st->alpha_s[k], z = (R)st->zeta[k], bt = (R)st->beta_s[k];
      const SpinorView<P> xv = view<P>(xs[k], V), pv = view<P>(ps[k], V);
      cpx<R> x[NCOMP], p[NCOMP];
      xv.template load<NCOMP>(s, x);
      pv.template load<NCOMP>(s, p);
#pragma unroll
      for (int c = 0; c < NCOMP; ++c) {
        x[c] = mk<R>(x[c].re + a * p[c].re, x[c].im + a * p[c].im);
        p[c] = mk<R>(z * rv[c].re + bt * p[c].re, z * rv[c].im + bt * p[c].im);
      }
      xv.template store<NCOMP>(s, x);
      pv.template store<NCOMP>(s, p);
    }
  }
  if (block_is_last(ticket) && threadIdx.x == 0) st->phase = st->final_k3 ? PH_DONE : PH_K1;
}

// ---------------------------------------------------------------------------
// MR vector kernels (complex128)
// ---------------------------------------------------------------------------
template <int NCOMP>
__global__ __launch_bounds__(kBlock) void k_mr_init(Field b, Field r, const Field* xs, int first, int nshift,
                                                    int64_t V) {
  const int64_t s = xcd_block() * kBlock + threadIdx.x;
  if (s >= V) return;
  cpx<double> v[NCOMP], z[NCOMP];
  view<P64>(b, V).template load<NCOMP>(s, v);
  view<P64>(r, V).template store<NCOMP>(s, v);
#pragma unroll
  for (int c = 0; c < NCOMP; ++c) z[c] = czero<double>();
  for (int k = first; k < nshift; ++k) view<P64>(xs[k], V).template store<NCOMP>(s, z);
}

// x_i += c_i r ; r -= alpha p ; reduce |r|^2  (dirac_solver.py:37-47)
template <int NCOMP>
__global__ __launch_bounds__(kBlock) void k_mr_k3(SolverState* st, Field r, Field p, const Field* xs, int first,
                                                  int nshift, int64_t V, RedCtl rc) {
  if (st->phase != PH_K3) return;
  const int64_t s = xcd_block() * kBlock + threadIdx.x;
  double red[1] = {0.0};
  if (s < V) {
    cpx<double> rv[NCOMP], pv[NCOMP];
    view<P64>(r, V).template load<NCOMP>(s, rv);
    view<P64>(p, V).template load<NCOMP>(s, pv);
    for (int k = first; k < nshift; ++k) {
      const cpx<double> c = mk<double>(st->c_re[k], st->c_im[k]);
      const SpinorView<P64> xv = view<P64>(xs[k], V);
      cpx<double> x[NCOMP];
      xv.template load<NCOMP>(s, x);
#pragma unroll
      for (int i = 0; i < NCOMP; ++i) cfma(x[i], c, rv[i]);
      xv.template store<NCOMP>(s, x);
    }
    const cpx<double> a = mk<double>(st->mr_alpha_re, st->mr_alpha_im);
#pragma unroll
    for (int i = 0; i < NCOMP; ++i) {
      rv[i] -= cmul(a, pv[i]);
      red[0] += rv[i].re * rv[i].re + rv[i].im * rv[i].im;
    }
    view<P64>(r, V).template store<NCOMP>(s, rv);
  }
  double sums[1];
  if (block_reduce_last<1>(red, rc, sums) && threadIdx.x == 0) mr_fin_rr(st, sums[0]);
}

// ---------------------------------------------------------------------------
// Fast CG-M kernels: residual ring + deferred shift updates
// ---------------------------------------------------------------------------
// ring slot 0 = b, live p = 0, x_s = 0, p_s = b (FP64); reduce |b|^2.
template <int NCOMP>
__global__ __launch_bounds__(kBlock) void k_fast_init(SolverState* st, Field b, Field r0, Field plive,
                                                      const Field* xs, const Field* ps, int nshift, int64_t V,
                                                      RedCtl rc) {
  const int64_t s = xcd_block() * kBlock + threadIdx.x;
  double red[1] = {0.0};
  if (s < V) {
    cpx<double> v[NCOMP], z[NCOMP];
    view<P64>(b, V).template load<NCOMP>(s, v);
    view<P64>(r0, V).template store<NCOMP>(s, v);
#pragma unroll
    for (int c = 0; c < NCOMP; ++c) {
      z[c] = czero<double>();
      red[0] += v[c].re * v[c].re + v[c].im * v[c].im;
    }
    view<P64>(plive, V).template store<NCOMP>(s, z);
    for (int k = 0; k < nshift; ++k) {
      view<P64>(ps[k], V).template store<NCOMP>(s, v);
      view<P64>(xs[k], V).template store<NCOMP>(s, z);
    }
  }
  double sums[1];
  if (block_reduce_last<1>(red, rc, sums) && threadIdx.x == 0) {
    st->bb = sums[0];
    st->rr = sums[0];
    for (int k = 0; k < st->nshift; ++k) st->tol2_s[k] *= sums[0];
    st->switch_rr *= sums[0];
    if (sums[0] > 0.0) {
      st->phase = PH_KP;
    } else {
      st->status = ST_CONVERGED;
      st->phase = PH_DONE;
    }
  }
}

template <int P>
__device__ __forceinline__ SpinorView<P> ring_view(const Field& ring, int64_t slot_elems, int slot, int64_t V) {
  using E = typename Store<P>::elem;
  return SpinorView<P>{(E*)ring.data + slot * slot_elems, ring.norm ? ring.norm + slot * V : nullptr, V};
}

// p = r^k + beta_{k-1} p   (r^k in ring slot k % nslot)
template <int P, int NCOMP>
__global__ __launch_bounds__(kBlock) void k_fast_kp(SolverState* st, Field ring, int64_t slot_elems, int nslot,
                                                    Field p, int64_t V, unsigned* ticket) {
  using R = typename Store<P>::R;
  if (st->phase != PH_KP) return;
  const int64_t s = xcd_block() * kBlock + threadIdx.x;
  if (s < V) {
    const SpinorView<P> rv = ring_view<P>(ring, slot_elems, st->iter % nslot, V);
    const SpinorView<P> pv = view<P>(p, V);
    const R b = (R)st->beta_old;
    cpx<R> r[NCOMP], q[NCOMP];
    rv.template load<NCOMP>(s, r);
    pv.template load<NCOMP>(s, q);
#pragma unroll
    for (int c = 0; c < NCOMP; ++c) q[c] = mk<R>(r[c].re + b * q[c].re, r[c].im + b * q[c].im);
    pv.template store<NCOMP>(s, q);
  }
  if (block_is_last(ticket) && threadIdx.x == 0) st->phase = PH_K1;
}

// Deferred shift updates of one block: for every shift, apply its recorded
// steps j < n_upd[s]:  x_s += a_j p_s ;  p_s = z_j r^{k0+j+1} + b_j p_s
// (x_s, p_s in FP64; residuals read once per block from the ring).
template <int PR, int NCOMP, int MB>
__global__ __launch_bounds__(kBlock) void k_fast_k3(SolverState* st, Field ring, int64_t slot_elems, int nslot,
                                                    const Field* xs, const Field* ps, int nshift, int64_t V,
                                                    unsigned* ticket) {
  using E = typename Store<PR>::elem;
  if (st->phase != PH_K3) return;
  const int64_t s = xcd_block() * kBlock + threadIdx.x;
  const int nb = st->blk_j;
  const int k0 = st->iter - nb;
  if (s < V) {
    const uint32_t fb = (uint32_t)(NCOMP * V * 16);
    const uint32_t rb = (uint32_t)(NCOMP * V * (int64_t)sizeof(E));
    float rn[MB];
#pragma unroll
    for (int j = 0; j < MB; ++j) {
      rn[j] = 1.f;
      if constexpr (PR == P16) {
        if (j < nb) {
          const int slot = (k0 + j + 1) % nslot;
          rn[j] = ld<float>(ring.norm + slot * V, (uint32_t)(V * 4), 0, V, s);
        }
      }
    }
#pragma unroll 1
    for (int c = 0; c < NCOMP; ++c) {
      cpx<double> r[MB];
#pragma unroll
      for (int j = 0; j < MB; ++j) {
        r[j] = czero<double>();
        if (j < nb) {
          const int slot = (k0 + j + 1) % nslot;
          const E e = ld<E>((const E*)ring.data + slot * slot_elems, rb, c, V, s);
          if constexpr (PR == P16) {
            const float2 f = __half22float2(e);
            r[j] = mk<double>((double)(f.x * rn[j]), (double)(f.y * rn[j]));
          } else {
            r[j] = mk<double>((double)e.x, (double)e.y);
          }
        }
      }
      for (int k = 0; k < nshift; ++k) {
        const int n = st->n_upd[k];
        if (n == 0) continue;
        double2 x = ld<double2>((const double2*)xs[k].data, fb, c, V, s);
        double2 p = ld<double2>((const double2*)ps[k].data, fb, c, V, s);
#pragma unroll
        for (int j = 0; j < MB; ++j) {
          if (j < n) {
            const double a = st->h_alpha[j][k], z = st->h_zeta[j][k], b = st->h_beta[j][k];
            x.x += a * p.x;
            x.y += a * p.y;
            p.x = z * r[j].re + b * p.x;
            p.y = z * r[j].im + b * p.y;
          }
        }
        put<double2>((double2*)xs[k].data, fb, c, V, s, x);
        put<double2>((double2*)ps[k].data, fb, c, V, s, p);
      }
    }
  }
  if (block_is_last(ticket) && threadIdx.x == 0) {
    st->blk_j = 0;
    for (int k = 0; k < st->nshift; ++k) st->n_upd[k] = 0;
    st->phase = st->final_k3 ? PH_DONE : PH_KP;
  }
}

// Precision switch at a block boundary: low-precision copies of the current
// residual (ring slot iter % nslot) and of the live direction.
template <int PL, int NCOMP>
__global__ __launch_bounds__(kBlock) void k_fast_switch(SolverState* st, Field ring64, Field ringlo,
                                                        int64_t slot_elems, int nslot, Field p64, Field plo,
                                                        int64_t V) {
  const int64_t s = xcd_block() * kBlock + threadIdx.x;
  if (s < V) {
    const int slot = st->iter % nslot;
    cpx<double> r[NCOMP], p[NCOMP];
    ring_view<P64>(ring64, slot_elems, slot, V).template load<NCOMP>(s, r);
    view<P64>(p64, V).template load<NCOMP>(s, p);
    using R = typename Store<PL>::R;
    cpx<R> rl[NCOMP], pl[NCOMP];
#pragma unroll
    for (int c = 0; c < NCOMP; ++c) {
      rl[c] = mk<R>((R)r[c].re, (R)r[c].im);
      pl[c] = mk<R>((R)p[c].re, (R)p[c].im);
    }
    ring_view<PL>(ringlo, slot_elems, slot, V).template store<NCOMP>(s, rl);
    view<PL>(plo, V).template store<NCOMP>(s, pl);
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) st->precision_phase = 1;
}

// ---------------------------------------------------------------------------
// flat FP64 helpers for true residuals and defect correction
// ---------------------------------------------------------------------------
// r = b - ax ; out[0] = |r|^2, out[1] = |b|^2
__global__ __launch_bounds__(kBlock) void k_residual64(int64_t n, const double2* __restrict__ b,
                                                       const double2* __restrict__ ax, double2* __restrict__ r,
                                                       RedCtl rc, double* out) {
  double v[2] = {0.0, 0.0};
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock) {
    const double2 bv = b[i], a = ax[i];
    double2 d;
    d.x = bv.x - a.x;
    d.y = bv.y - a.y;
    if (r) r[i] = d;
    v[0] += d.x * d.x + d.y * d.y;
    v[1] += bv.x * bv.x + bv.y * bv.y;
  }
  double sums[2];
  if (block_reduce_last<2>(v, rc, sums) && threadIdx.x == 0) {
    out[0] = sums[0];
    out[1] = sums[1];
  }
}

// x64 += e32
__global__ __launch_bounds__(kBlock) void k_add32to64(int64_t n, const float2* __restrict__ e, double2* __restrict__ x) {
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock) {
    const float2 a = e[i];
    double2 b = x[i];
    b.x += (double)a.x;
    b.y += (double)a.y;
    x[i] = b;
  }
}

static int red_grid_flat(int64_t n) {
  int64_t g = (n + kBlock * 4 - 1) / (kBlock * 4);
  if (g > 2048) g = 2048;
  if (g < 1) g = 1;
  return (int)g;
}

// ---------------------------------------------------------------------------
// workspace layout
// ---------------------------------------------------------------------------
struct Carver {
  char* base;
  size_t off = 0;
  void* take(size_t bytes) {
    off = (off + 255) & ~(size_t)255;
    void* p = base ? (void*)(base + off) : nullptr;
    off += bytes;
    return p;
  }
};

static size_t field_bytes(const OpDesc& d, int prec) {
  const int64_t n = field_elems(d);
  const size_t eb = prec == P64 ? 16 : prec == P32 ? 8 : 4;
  return (size_t)n * eb;
}

static Field take_field(Carver& c, const OpDesc& d, int prec) {
  Field f;
  f.data = c.take(field_bytes(d, prec));
  if (prec == P16) f.norm = (float*)c.take((size_t)d.lat.V * sizeof(float));
  return f;
}

struct Parts {
  SolverState* st;
  unsigned* ticket;
  double* partials;
  double* results;
  Field* xs_dev;
  Field* ps_dev;
  Field r64, t64, u64, v64;
  Field b_lo, r_lo, t_lo;  // low-precision copies (mixed solves)
  Field x_lo[kMaxShifts];
  Field p[kMaxShifts];     // search directions at the solve precision
  // fast CG-M (mixed): residual rings, live base direction, FP64 shift directions
  Field ring64, ring32;    // kFastBlock + 1 contiguous slots each
  int64_t slot_elems = 0;
  Field plive64, plive32;
  Field ps64[kMaxShifts];
};

constexpr int kFastBlock = 8;  // iterations per deferred shift-update block

// prec: precision of the Krylov vectors (P64 for FP64 solves, P32/P16 mixed)
static size_t carve(const OpDesc& d, int prec, int nshift, char* base, Parts* pt) {
  Carver c{base};
  Parts tmp;
  Parts& P = pt ? *pt : tmp;
  P.st = (SolverState*)c.take(sizeof(SolverState));
  P.ticket = (unsigned*)c.take(64);
  P.partials = (double*)c.take(sizeof(double) * kMaxRed * (size_t)max_red_blocks(d.lat.V));
  P.results = (double*)c.take(sizeof(double) * 16);
  P.xs_dev = (Field*)c.take(sizeof(Field) * kMaxShifts);
  P.ps_dev = (Field*)c.take(sizeof(Field) * kMaxShifts);
  P.r64 = take_field(c, d, P64);
  P.t64 = take_field(c, d, P64);
  P.u64 = take_field(c, d, P64);
  P.v64 = take_field(c, d, P64);
  if (prec == P64) {
    for (int k = 0; k < nshift; ++k) P.p[k] = take_field(c, d, P64);
  } else {
    // single-shift defect correction buffers
    P.b_lo = take_field(c, d, P32);
    P.r_lo = take_field(c, d, P32);
    P.t_lo = take_field(c, d, P32);
    P.x_lo[0] = take_field(c, d, P32);
    P.p[0] = take_field(c, d, P32);
    // fast CG-M
    P.slot_elems = field_elems(d);
    P.ring64.data = c.take(field_bytes(d, P64) * (kFastBlock + 1));
    P.ring32.data = c.take(field_bytes(d, P32) * (kFastBlock + 1));
    P.plive64 = take_field(c, d, P64);
    P.plive32 = take_field(c, d, P32);
    for (int k = 0; k < nshift; ++k) P.ps64[k] = take_field(c, d, P64);
  }
  return c.off + 256;
}

size_t solver_workspace_bytes(const OpDesc& d, int prec, int nshift) { return carve(d, prec, nshift, nullptr, nullptr); }

// ---------------------------------------------------------------------------
// host loop helpers
// ---------------------------------------------------------------------------
struct HostSync {
  int* pinned = nullptr;  // 2 slots x (phase, iter)
  hipEvent_t ev[2] = {nullptr, nullptr};
  int device = -1;
  ~HostSync() {}
};
static thread_local HostSync g_sync;

static int ensure_sync() {
  int dev = 0;
  LYM_HIP(hipGetDevice(&dev));
  if (g_sync.pinned && g_sync.device == dev) return LYM_OK;
  if (g_sync.pinned) {
    (void)hipHostFree(g_sync.pinned);
    (void)hipEventDestroy(g_sync.ev[0]);
    (void)hipEventDestroy(g_sync.ev[1]);
    g_sync.pinned = nullptr;
  }
  LYM_HIP(hipHostMalloc((void**)&g_sync.pinned, 16 * sizeof(int), hipHostMallocDefault));
  LYM_HIP(hipEventCreateWithFlags(&g_sync.ev[0], hipEventDisableTiming));
  LYM_HIP(hipEventCreateWithFlags(&g_sync.ev[1], hipEventDisableTiming));
  g_sync.device = dev;
  return LYM_OK;
}

// Enqueue `enqueue_iter` chunk by chunk, polling st->phase; two chunks in flight.
template <class F>
static int run_until_done(SolverState* st_dev, int max_iter, hipStream_t stream, F&& enqueue_iter) {
  int rc = ensure_sync();
  if (rc) return rc;
  int chunk = 4;
  const int64_t max_chunks = (int64_t)max_iter / 4 + 8;
  int64_t issued = 0, polled = 0;
  auto issue = [&](int slot) -> int {
    for (int i = 0; i < chunk; ++i) {
      const int e = enqueue_iter();
      if (e) return e;
    }
    LYM_HIP(hipMemcpyAsync(&g_sync.pinned[2 * slot], st_dev, 2 * sizeof(int), hipMemcpyDeviceToHost, stream));
    LYM_HIP(hipEventRecord(g_sync.ev[slot], stream));
    ++issued;
    if (chunk < 32) chunk *= 2;
    return LYM_OK;
  };
  if ((rc = issue(0))) return rc;
  if ((rc = issue(1))) return rc;
  while (true) {
    const int slot = (int)(polled & 1);
    LYM_HIP(hipEventSynchronize(g_sync.ev[slot]));
    ++polled;
    if (g_sync.pinned[2 * slot] == PH_DONE) break;
    if (issued > max_chunks) {
      set_error("solver did not reach its stop state (internal error)");
      return LYM_ERR_HIP;
    }
    if ((rc = issue(slot))) return rc;
  }
  LYM_HIP(hipStreamSynchronize(stream));
  return LYM_OK;
}

static RedCtl redctl(const Parts& P) { return RedCtl{P.partials, P.ticket}; }

// FP64 true residual of one shift: out = b - (Op^2 + sigma) x, returns |out|^2/|b|^2.
static int true_residual(const OpDesc& d, const Parts& P, Field b, Field x, double sigma, Field out, double* rel2,
                         hipStream_t st) {
  EpiArgs a;
  a.out = P.t64;
  int rc = wilson_launch(EPI_STORE, d, P64, true, x, a, redctl(P), st);
  if (rc) return rc;
  EpiArgs a2;
  a2.out = P.u64;
  a2.aux0 = x;
  a2.sigma = sigma;
  rc = wilson_launch(EPI_AXPY, d, P64, true, P.t64, a2, redctl(P), st);
  if (rc) return rc;
  const int64_t n = field_elems(d);
  hipLaunchKernelGGL(k_residual64, dim3(red_grid_flat(n)), dim3(kBlock), 0, st, n, (const double2*)b.data,
                     (const double2*)P.u64.data, (double2*)out.data, redctl(P), P.results);
  LYM_HIP(hipGetLastError());
  double h[2];
  LYM_HIP(hipMemcpyAsync(h, P.results, sizeof(h), hipMemcpyDeviceToHost, st));
  LYM_HIP(hipStreamSynchronize(st));
  *rel2 = h[1] > 0 ? h[0] / h[1] : 0.0;
  return LYM_OK;
}

// Core CG-M at Krylov precision `prec`: b, xs[], r, t, p[] all at `prec`.
static int cgm_core(const OpDesc& d, int prec, const Parts& P, Field b, const Field* xs, Field r, Field t,
                    int nshift, const double* shifts, const double* tol_rel, int max_iter, hipStream_t st,
                    SolverState* out_state) {
  SolverState h;
  std::memset(&h, 0, sizeof(h));
  int base = 0;
  for (int k = 1; k < nshift; ++k)
    if (shifts[k] < shifts[base]) base = k;
  h.phase = PH_K1;
  h.max_iter = max_iter;
  h.nshift = nshift;
  h.base = base;
  h.sigma0 = shifts[base];
  h.alpha_old = 1.0;
  h.beta_old = 0.0;
  for (int k = 0; k < nshift; ++k) {
    h.shift[k] = shifts[k];
    h.delta[k] = shifts[k] - shifts[base];
    h.zeta[k] = h.zeta_old[k] = 1.0;
    h.active[k] = 1;
    h.tol2_s[k] = tol_rel[k] * tol_rel[k];
    h.conv_iter[k] = -1;
  }
  Field hx[kMaxShifts], hp[kMaxShifts];
  for (int k = 0; k < nshift; ++k) {
    hx[k] = xs[k];
    hp[k] = P.p[k];
  }
  LYM_HIP(hipMemcpyAsync(P.st, &h, sizeof(h), hipMemcpyHostToDevice, st));
  LYM_HIP(hipMemcpyAsync(P.xs_dev, hx, sizeof(Field) * nshift, hipMemcpyHostToDevice, st));
  LYM_HIP(hipMemcpyAsync(P.ps_dev, hp, sizeof(Field) * nshift, hipMemcpyHostToDevice, st));
  const int64_t V = d.lat.V;
  const int grid = grid_for(V);
  int rc = with_prec_ncomp(prec, d.ncomp, [&](auto PP, auto N) -> int {
    constexpr int PV = decltype(PP)::value, NC = decltype(N)::value;
    hipLaunchKernelGGL((k_cgm_init<PV, NC>), dim3(grid), dim3(kBlock), 0, st, P.st, b, r, P.xs_dev, P.ps_dev, nshift,
                       V, redctl(P));
    LYM_HIP(hipGetLastError());
    return LYM_OK;
  });
  if (rc) return rc;
  const Field pbase = P.p[base];
  auto iter = [&]() -> int {
    EpiArgs a1;
    a1.out = t;
    a1.st = P.st;
    int e = wilson_launch(EPI_CGM_K1, d, prec, true, pbase, a1, redctl(P), st);
    if (e) return e;
    EpiArgs a2;
    a2.aux0 = pbase;
    a2.aux1 = r;
    a2.st = P.st;
    e = wilson_launch(EPI_CGM_K2, d, prec, true, t, a2, redctl(P), st);
    if (e) return e;
    return with_prec_ncomp(prec, d.ncomp, [&](auto PP, auto N) -> int {
      constexpr int PV = decltype(PP)::value, NC = decltype(N)::value;
      hipLaunchKernelGGL((k_cgm_k3<PV, NC>), dim3(grid), dim3(kBlock), 0, st, P.st, r, P.xs_dev, P.ps_dev, nshift, V,
                         P.ticket);
      LYM_HIP(hipGetLastError());
      return LYM_OK;
    });
  };
  rc = run_until_done(P.st, max_iter, st, iter);
  if (rc) return rc;
  LYM_HIP(hipMemcpyAsync(out_state, P.st, sizeof(SolverState), hipMemcpyDeviceToHost, st));
  LYM_HIP(hipStreamSynchronize(st));
  return LYM_OK;
}

// Fast multishift CG: blocks of kFastBlock iterations
//   [ p = r + beta p ; t = Q p (|t|^2,|p|^2) ; r' = r - alpha (Q t + s0 p) (|r'|^2) ] x m
//   then one deferred update of every shift's x_s, p_s (FP64) from the ring.
// Krylov vectors are FP64 until want_switch (base rr <= switch_tol^2 |b|^2),
// then `lowprec`.  xs are the caller's FP64 solutions.
static int cgm_fast(const OpDesc& d, const Parts& P, Field b, void* const* xs, int nshift, const double* shifts,
                    const double* tol_rel, int max_iter, double switch_tol, int lowprec, hipStream_t st,
                    SolverState* out_state) {
  SolverState h;
  std::memset(&h, 0, sizeof(h));
  int base = 0;
  for (int k = 1; k < nshift; ++k)
    if (shifts[k] < shifts[base]) base = k;
  h.phase = PH_KP;
  h.max_iter = max_iter;
  h.nshift = nshift;
  h.base = base;
  h.sigma0 = shifts[base];
  h.alpha_old = 1.0;
  h.beta_old = 0.0;
  h.mblk = kFastBlock;
  h.switch_rr = switch_tol * switch_tol;
  for (int k = 0; k < nshift; ++k) {
    h.shift[k] = shifts[k];
    h.delta[k] = shifts[k] - shifts[base];
    h.zeta[k] = h.zeta_old[k] = 1.0;
    h.active[k] = 1;
    h.tol2_s[k] = tol_rel[k] * tol_rel[k];
    h.conv_iter[k] = -1;
  }
  Field hx[kMaxShifts], hp[kMaxShifts];
  for (int k = 0; k < nshift; ++k) {
    hx[k].data = xs[k];
    hp[k] = P.ps64[k];
  }
  LYM_HIP(hipMemcpyAsync(P.st, &h, sizeof(h), hipMemcpyHostToDevice, st));
  LYM_HIP(hipMemcpyAsync(P.xs_dev, hx, sizeof(Field) * nshift, hipMemcpyHostToDevice, st));
  LYM_HIP(hipMemcpyAsync(P.ps_dev, hp, sizeof(Field) * nshift, hipMemcpyHostToDevice, st));
  const int64_t V = d.lat.V;
  const int grid = grid_for(V);
  const int nslot = kFastBlock + 1;
  int rc = with_ncomp(d.ncomp, [&](auto N) -> int {
    constexpr int NC = decltype(N)::value;
    hipLaunchKernelGGL((k_fast_init<NC>), dim3(grid), dim3(kBlock), 0, st, P.st, b, P.ring64, P.plive64, P.xs_dev,
                       P.ps_dev, nshift, V, redctl(P));
    LYM_HIP(hipGetLastError());
    return LYM_OK;
  });
  if (rc) return rc;

  int prec = P64;
  auto enqueue_block = [&]() -> int {
    const Field ring = prec == P64 ? P.ring64 : P.ring32;
    const Field pl = prec == P64 ? P.plive64 : P.plive32;
    const Field t = prec == P64 ? P.t64 : P.t_lo;
    for (int j = 0; j < kFastBlock; ++j) {
      int e = with_prec_ncomp(prec, d.ncomp, [&](auto PP, auto N) -> int {
        constexpr int PV = decltype(PP)::value, NC = decltype(N)::value;
        hipLaunchKernelGGL((k_fast_kp<PV, NC>), dim3(grid), dim3(kBlock), 0, st, P.st, ring, P.slot_elems, nslot, pl,
                           V, P.ticket);
        LYM_HIP(hipGetLastError());
        return LYM_OK;
      });
      if (e) return e;
      EpiArgs a1;
      a1.out = t;
      a1.st = P.st;
      e = wilson_launch(EPI_CGM_K1, d, prec, true, pl, a1, redctl(P), st);
      if (e) return e;
      EpiArgs a2;
      a2.aux0 = pl;
      a2.st = P.st;
      a2.ring = ring;
      a2.slot_elems = P.slot_elems;
      a2.nslot = nslot;
      e = wilson_launch(EPI_CGM_K2F, d, prec, true, t, a2, redctl(P), st);
      if (e) return e;
    }
    return with_prec_ncomp(prec, d.ncomp, [&](auto PP, auto N) -> int {
      constexpr int PV = decltype(PP)::value, NC = decltype(N)::value;
      hipLaunchKernelGGL((k_fast_k3<PV, NC, kFastBlock>), dim3(grid), dim3(kBlock), 0, st, P.st, ring, P.slot_elems,
                         nslot, P.xs_dev, P.ps_dev, nshift, V, P.ticket);
      LYM_HIP(hipGetLastError());
      return LYM_OK;
    });
  };
  auto enqueue_switch = [&]() -> int {
    return with_ncomp(d.ncomp, [&](auto N) -> int {
      constexpr int NC = decltype(N)::value;
      if (lowprec == P32) {
        hipLaunchKernelGGL((k_fast_switch<P32, NC>), dim3(grid), dim3(kBlock), 0, st, P.st, P.ring64, P.ring32,
                           P.slot_elems, nslot, P.plive64, P.plive32, V);
      } else {
        set_error("fast CG-M: unsupported low precision");
        return (int)LYM_ERR_UNSUPPORTED;
      }
      LYM_HIP(hipGetLastError());
      return LYM_OK;
    });
  };

  rc = ensure_sync();
  if (rc) return rc;
  const int64_t max_blocks = (int64_t)max_iter / kFastBlock + 8;
  int64_t issued = 0, polled = 0;
  auto issue = [&](int slot) -> int {
    int e = enqueue_block();
    if (e) return e;
    LYM_HIP(hipMemcpyAsync(&g_sync.pinned[4 * slot], P.st, 4 * sizeof(int), hipMemcpyDeviceToHost, st));
    LYM_HIP(hipEventRecord(g_sync.ev[slot], st));
    ++issued;
    return LYM_OK;
  };
  if ((rc = issue(0))) return rc;
  if ((rc = issue(1))) return rc;
  while (true) {
    const int slot = (int)(polled & 1);
    LYM_HIP(hipEventSynchronize(g_sync.ev[slot]));
    ++polled;
    const int* s4 = &g_sync.pinned[4 * slot];
    if (s4[0] == PH_DONE) break;
    if (issued > max_blocks) {
      set_error("fast CG-M did not reach its stop state (internal error)");
      return LYM_ERR_HIP;
    }
    if (s4[2] && prec == P64 && lowprec != P64) {
      if ((rc = enqueue_switch())) return rc;
      prec = lowprec;
    }
    if ((rc = issue(slot))) return rc;
  }
  LYM_HIP(hipStreamSynchronize(st));
  LYM_HIP(hipMemcpyAsync(out_state, P.st, sizeof(SolverState), hipMemcpyDeviceToHost, st));
  LYM_HIP(hipStreamSynchronize(st));
  return LYM_OK;
}

static void fill_stats(lym_solve_stats* s, const SolverState& h, int nshift) {
  if (!s) return;
  s->status = h.status;
  s->iterations = h.iter;
  s->nshift = nshift;
  s->final_rr = h.rr;
  for (int k = 0; k < nshift && k < 64; ++k) s->conv_iter[k] = h.conv_iter[k];
}

int solve_cgm(const OpDesc& d, int prec, int nshift, const double* shifts, const void* b, void* const* xs, double tol,
              int max_iter, void* ws, size_t ws_bytes, lym_solve_stats* stats, hipStream_t st) {
  const auto t0 = std::chrono::steady_clock::now();
  if (!d.herm) {
    set_error("CG needs a Hermitian operator: build the Wilson operator with hermitian=True");
    return LYM_ERR_ARG;
  }
  if (nshift < 1 || nshift > kMaxShifts) {
    set_error("number of shifts must be in [1, 64]");
    return LYM_ERR_ARG;
  }
  for (int k = 0; k < nshift; ++k)
    if (!(shifts[k] >= 0.0)) {
      set_error("CG-M needs non-negative shifts");
      return LYM_ERR_ARG;
    }
  const int inner = (prec == P64) ? P64 : prec;
  const size_t need = solver_workspace_bytes(d, inner, nshift);
  if (!ws || ws_bytes < need) {
    set_error("solver workspace too small: need " + std::to_string(need) + " bytes");
    return LYM_ERR_ARG;
  }
  Parts P;
  char* base = (char*)(((uintptr_t)ws + 255) & ~(uintptr_t)255);
  carve(d, inner, nshift, base, &P);
  const int64_t n = field_elems(d);
  Field bf;
  bf.data = const_cast<void*>(b);
  std::vector<double> tol_rel(nshift, tol);
  SolverState h;
  int rc = 0;
  int restarts = 0;
  std::vector<double> rel(nshift, 0.0);
  if (inner == P64) {
    // iterate slightly below the target so the recomputed true residual meets it
    for (auto& v : tol_rel) v = 0.5 * tol;
    Field xf[kMaxShifts];
    for (int k = 0; k < nshift; ++k) xf[k].data = xs[k];
    rc = cgm_core(d, P64, P, bf, xf, P.r64, P.t64, nshift, shifts, tol_rel.data(), max_iter, st, &h);
    if (rc) return rc;
  } else {
    // mixed: FP64 Krylov vectors until the base residual is below 1e-4 of |b|,
    // low precision after; shift solutions and directions accumulate in FP64
    // with deferred block updates.  FP64 true residuals are checked below.
    if (inner == P16) {
      set_error("half-precision CG-M is not built yet");
      return LYM_ERR_UNSUPPORTED;
    }
    for (auto& v : tol_rel) v = 0.5 * tol;
    rc = cgm_fast(d, P, bf, xs, nshift, shifts, tol_rel.data(), max_iter, 1e-4, inner, st, &h);
    if (rc) return rc;
  }
  // FP64 true residuals, and defect correction where needed
  const int max_refine = 12;
  for (int k = 0; k < nshift; ++k) {
    Field xk;
    xk.data = xs[k];
    double r2 = 0;
    rc = true_residual(d, P, bf, xk, shifts[k], P.r64, &r2, st);
    if (rc) return rc;
    int cycles = 0;
    while (std::sqrt(r2) > tol && cycles < max_refine && h.status != ST_BREAKDOWN) {
      // solve (A + s_k) e = r at the inner precision to a relative tol_in
      const double relnow = std::sqrt(r2);
      double tol_in = (inner == P64) ? 0.5 * tol / relnow : std::max(0.5 * tol / relnow, 1e-5);
      if (tol_in > 0.1) tol_in = 0.1;
      SolverState hi;
      if (inner == P64) {
        // FP64 refinement (only when the FP64 recurrence drifted): solve for
        // the correction e (u64) from the residual copied to v64
        rc = blas_convert(P64, P.r64.data, P64, P.v64.data, n, st);
        if (rc) return rc;
        double sh = shifts[k];
        Field xe[1] = {P.u64};
        rc = cgm_core(d, P64, P, P.v64, xe, P.r64, P.t64, 1, &sh, &tol_in, max_iter, st, &hi);
        if (rc) return rc;
        rc = blas_axpby(P64, n, 1.0, 0.0, P.u64.data, 1.0, 0.0, xs[k], st);
        if (rc) return rc;
      } else {
        rc = blas_convert(P64, P.r64.data, P32, P.b_lo.data, n, st);
        if (rc) return rc;
        double sh = shifts[k];
        Field xe[1] = {P.x_lo[0]};
        rc = cgm_core(d, P32, P, P.b_lo, xe, P.r_lo, P.t_lo, 1, &sh, &tol_in, max_iter, st, &hi);
        if (rc) return rc;
        hipLaunchKernelGGL(k_add32to64, dim3(2048), dim3(kBlock), 0, st, n, (const float2*)P.x_lo[0].data,
                           (double2*)xs[k]);
        LYM_HIP(hipGetLastError());
      }
      ++restarts;
      ++cycles;
      rc = true_residual(d, P, bf, xk, shifts[k], P.r64, &r2, st);
      if (rc) return rc;
    }
    rel[k] = std::sqrt(r2);
  }
  fill_stats(stats, h, nshift);
  int status = h.status;
  for (int k = 0; k < nshift; ++k)
    if (rel[k] > tol && status == ST_CONVERGED) status = ST_MAXITER;
  if (stats) {
    stats->status = status;
    stats->restarts = restarts;
    for (int k = 0; k < nshift && k < 64; ++k) stats->true_rel_resid[k] = rel[k];
    stats->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  }
  return status == ST_CONVERGED ? LYM_OK : (status == ST_BREAKDOWN ? LYM_BREAKDOWN : LYM_NOT_CONVERGED);
}

int solve_mr(const OpDesc& d, int nshift, const double* shifts, int first, const void* b, void* const* xs,
             int max_steps, double precision, double omega, void* ws, size_t ws_bytes, lym_solve_stats* stats,
             hipStream_t st) {
  const auto t0 = std::chrono::steady_clock::now();
  if (nshift < 1 || nshift > kMaxShifts || first < 0 || first >= nshift) {
    set_error("bad shift list for the multishift MR solver");
    return LYM_ERR_ARG;
  }
  const size_t need = solver_workspace_bytes(d, P64, nshift);
  if (!ws || ws_bytes < need) {
    set_error("solver workspace too small: need " + std::to_string(need) + " bytes");
    return LYM_ERR_ARG;
  }
  Parts P;
  char* base = (char*)(((uintptr_t)ws + 255) & ~(uintptr_t)255);
  carve(d, P64, nshift, base, &P);
  SolverState h;
  std::memset(&h, 0, sizeof(h));
  h.phase = PH_K1;
  h.max_iter = max_steps;
  h.nshift = nshift;
  h.base = first;
  h.omega = omega;
  h.tol2 = precision;
  for (int k = 0; k < nshift; ++k) {
    h.shift[k] = shifts[k];
    h.f_re[k] = 1.0;
    h.f_im[k] = 0.0;
  }
  Field hx[kMaxShifts];
  for (int k = 0; k < nshift; ++k) hx[k].data = xs[k];
  LYM_HIP(hipMemcpyAsync(P.st, &h, sizeof(h), hipMemcpyHostToDevice, st));
  LYM_HIP(hipMemcpyAsync(P.xs_dev, hx, sizeof(Field) * nshift, hipMemcpyHostToDevice, st));
  const int64_t V = d.lat.V;
  const int grid = grid_for(V);
  Field bf;
  bf.data = const_cast<void*>(b);
  int rc = with_ncomp(d.ncomp, [&](auto N) -> int {
    constexpr int NC = decltype(N)::value;
    hipLaunchKernelGGL((k_mr_init<NC>), dim3(grid), dim3(kBlock), 0, st, bf, P.r64, P.xs_dev, first, nshift, V);
    LYM_HIP(hipGetLastError());
    return LYM_OK;
  });
  if (rc) return rc;
  auto iter = [&]() -> int {
    EpiArgs a1;
    a1.out = P.t64;
    a1.st = P.st;
    int e = wilson_launch(EPI_MR_K1, d, P64, true, P.r64, a1, redctl(P), st);
    if (e) return e;
    EpiArgs a2;
    a2.out = P.u64;
    a2.aux0 = P.r64;
    a2.st = P.st;
    e = wilson_launch(EPI_MR_K2, d, P64, true, P.t64, a2, redctl(P), st);
    if (e) return e;
    return with_ncomp(d.ncomp, [&](auto N) -> int {
      constexpr int NC = decltype(N)::value;
      hipLaunchKernelGGL((k_mr_k3<NC>), dim3(grid), dim3(kBlock), 0, st, P.st, P.r64, P.u64, P.xs_dev, first, nshift,
                         V, redctl(P));
      LYM_HIP(hipGetLastError());
      return LYM_OK;
    });
  };
  rc = run_until_done(P.st, max_steps, st, iter);
  if (rc) return rc;
  LYM_HIP(hipMemcpyAsync(&h, P.st, sizeof(SolverState), hipMemcpyDeviceToHost, st));
  LYM_HIP(hipStreamSynchronize(st));
  fill_stats(stats, h, nshift);
  if (stats) stats->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  return h.status == ST_CONVERGED ? LYM_OK : (h.status == ST_BREAKDOWN ? LYM_BREAKDOWN : LYM_NOT_CONVERGED);
}

}  // namespace lym
