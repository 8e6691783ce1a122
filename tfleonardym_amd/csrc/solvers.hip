// Krylov solvers on the device.
//
// Every iteration is a fixed launch sequence: Wilson kernels with fused
// epilogues, flat element-parallel vector kernels, and one-block finalizer
// kernels that reduce the block partials in a fixed order and advance the
// FP64 scalars in SolverState.  The host only enqueues chunks of iterations
// and polls four ints of pinned memory per chunk.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <tuple>
#include <type_traits>
#include <vector>

#include "lym_fin.h"
#include "lym_internal.h"

namespace lym {

template <int N>
using IC = std::integral_constant<int, N>;

template <int P>
using Elem = typename Store<P>::elem;

// ---------------------------------------------------------------------------
// grid sizes of flat kernels (depend only on n: reductions are deterministic)
// ---------------------------------------------------------------------------
static int flat_grid(int64_t n) {
  int64_t g = (n + kBlock - 1) / kBlock;
  if (g > 16384) g = 16384;
  if (g < 1) g = 1;
  return (int)g;
}
static int red_grid_flat(int64_t n) {
  int64_t g = (n + kBlock * 4 - 1) / (kBlock * 4);
  if (g > 2048) g = 2048;
  if (g < 1) g = 1;
  return (int)g;
}

#define LYM_GRID_STRIDE(i, n) \
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < (n); i += (int64_t)gridDim.x * kBlock)

template <typename E>
__device__ __forceinline__ double2 to64(E e) {
  double2 d;
  d.x = (double)e.x;
  d.y = (double)e.y;
  return d;
}
template <typename E>
__device__ __forceinline__ E from64(double2 d) {
  E e;
  e.x = (decltype(e.x))d.x;
  e.y = (decltype(e.y))d.y;
  return e;
}

// ---------------------------------------------------------------------------
// finalizer kernel
// ---------------------------------------------------------------------------
// Batch bookkeeping the host polls: RHS that finished, RHS ready for the
// precision switch (wanting it or finished).
struct BatchCtl {
  int ndone;
  int nswitch;
  int nrhs;
  int pad;
};

// block i of the grid finalizes RHS i: state st + i, partials + i * nr * G
__global__ __launch_bounds__(kBlock) void k_finalize(SolverState* st, const double* partials, int G, int kind, int nr,
                                                     double* out, BatchCtl* ctl) {
  if (st) st += blockIdx.x;
  if (partials) partials += (int64_t)blockIdx.x * nr * G;
  if (out) out += (int64_t)blockIdx.x * nr;
  bool run = true;
  if (st) {
    const int ph = st->phase;
    switch (kind) {
      case FK_CGM_A: run = (ph == PH_K1 || ph == PH_KP); break;
      case FK_CGM_B:
      case FK_CGM_BF: run = (ph == PH_K2); break;
      case FK_CGM_K3ADV:
      case FK_FAST_K3ADV: run = (ph == PH_K3); break;
      case FK_MR: run = (ph == PH_K1); break;
      case FK_MR_RR: run = (ph == PH_K3); break;
      case FK_CG_A: run = (ph == PH_KP); break;
      case FK_CG_B: run = (ph == PH_K2); break;
      case FK_CG_RU: run = (ph == PH_RU); break;
      case FK_BICG_INIT:
      case FK_BICG_A:
      case FK_BICG_W:
      case FK_BICG_R: run = (ph == PH_K1); break;
      default: run = true;
    }
  }
  // the partial sums and the per-shift inputs are read before the phase test
  // resolves, so their loads overlap the load of the phase (a gated-out
  // launch reads them for nothing); nothing is written before the test
  double sums[4] = {0.0, 0.0, 0.0, 0.0};
  if (nr == 1) {
    double s[1];
    reduce_partials<1>(partials, G, s);
    sums[0] = s[0];
  } else if (nr == 2) {
    double s[2];
    reduce_partials<2>(partials, G, s);
    sums[0] = s[0];
    sums[1] = s[1];
  } else if (nr >= 3) {
    double s[3];
    reduce_partials<3>(partials, G, s);
    sums[0] = s[0];
    sums[1] = s[1];
    sums[2] = s[2];
  }
  __shared__ ShiftIn shifts;
  if (kind == FK_CGM_A || kind == FK_CG_A || kind == FK_CGM_B || kind == FK_CGM_BF) stage_shifts(st, shifts);
  if (!run) return;
  // kinds whose per-shift recurrences run on every thread of the block
  switch (kind) {
    case FK_INIT:
      if ((int)threadIdx.x < st->nshift) st->tol2_s[threadIdx.x] *= sums[0];
      break;

    case FK_FAST_K3ADV:
      if ((int)threadIdx.x < st->nshift) st->n_upd[threadIdx.x] = 0;
      break;
    default: break;
  }
  if (threadIdx.x != 0) return;
  switch (kind) {
    case FK_SUM:
      for (int r = 0; r < nr; ++r) out[r] = sums[r];
      break;
    case FK_INIT:
      st->bb = sums[0];
      st->rr = sums[0];
      st->switch_rr *= sums[0];
      st->switch2_rr *= sums[0];
      st->tol2_true *= sums[0];
      st->rr_max = sums[0];
      if (!(sums[0] > 0.0)) {
        // a zero source is solved by x = 0; a non-finite one is a breakdown
        st->status = isfinite(sums[0]) ? ST_CONVERGED : ST_BREAKDOWN;
        st->phase = PH_DONE;
        if (ctl) {  // a zero source is finished (and needs no precision switch)
          atomicAdd(&ctl->ndone, 1);
          st->switch_counted = 1;
          atomicAdd(&ctl->nswitch, 1);
        }
      }
      break;
    case FK_CGM_A:
      cgm_fin_a(st, shifts, sums[0], sums[1]);
      // a breakdown ends the solve here, not at a block flush: count it
      if (ctl && st->phase == PH_DONE) {
        atomicAdd(&ctl->ndone, 1);
        if (!st->switch_counted) {
          st->switch_counted = 1;
          atomicAdd(&ctl->nswitch, 1);
        }
      }
      break;
    case FK_CGM_B: cgm_fin_b(st, shifts, sums[0]); break;
    case FK_CGM_BF:
      cgm_fin_b_fast(st, shifts, sums[0]);
      if (ctl && st->want_switch && !st->switch_counted) {
        st->switch_counted = 1;
        atomicAdd(&ctl->nswitch, 1);
      }
      break;
    case FK_CGM_K3ADV: st->phase = st->final_k3 ? PH_DONE : PH_K1; break;
    case FK_FAST_K3ADV:
      st->blk_j = 0;
      st->phase = st->final_k3 ? PH_DONE : PH_KP;
      if (ctl && st->phase == PH_DONE) {
        atomicAdd(&ctl->ndone, 1);
        if (!st->switch_counted) {
          st->switch_counted = 1;
          atomicAdd(&ctl->nswitch, 1);
        }
      }
      break;
    case FK_MR: mr_fin(st, sums[0], sums[1], sums[2]); break;
    case FK_MR_RR: mr_fin_rr(st, sums[0]); break;
    case FK_BICG_INIT: bicg_fin_init(st, sums[0], sums[1], sums[2]); break;
    case FK_BICG_A: bicg_fin_a(st, sums[0], sums[1]); break;
    case FK_BICG_W: bicg_fin_w(st, sums[0], sums[1], sums[2]); break;
    case FK_BICG_R: bicg_fin_r(st, sums[0], sums[1], sums[2]); break;
    case FK_CG_A:
    case FK_CG_B:
    case FK_CG_RU:
      if (kind == FK_CG_A) {
        if (st->blk_j >= st->mblk) st->blk_j = 0;  // k_rel_kp folded the full block
        cgm_fin_a(st, shifts, sums[0], sums[1]);
      }
      else if (kind == FK_CG_B) cg_fin_b_rel(st, sums[0]);
      else if (kind == FK_CG_RU) cg_fin_ru(st, sums[0], sums[1], sums[2]);
      if (ctl && st->phase == PH_DONE) atomicAdd(&ctl->ndone, 1);
      break;
  }
}

static int launch_fin(FinKind kind, SolverState* st, const double* partials, int G, int nr, double* out,
                      hipStream_t stream, int nb, BatchCtl* ctl) {
  hipLaunchKernelGGL(k_finalize, dim3(nb), dim3(kBlock), 0, stream, st, partials, G, (int)kind, nr, out, ctl);
  LYM_HIP(hipGetLastError());
  return LYM_OK;
}

int launch_finalize(FinKind kind, SolverState* st, const double* partials, int G, int nr, double* out,
                    hipStream_t stream, int nb) {
  return launch_fin(kind, st, partials, G, nr, out, stream, nb, nullptr);
}

// ---------------------------------------------------------------------------
// flat vector kernels: element i of every field (component planes are
// contiguous, so element i of a spinor is plane i / V, site i % V)
// ---------------------------------------------------------------------------
// r = b ; p_s = b ; x_s = 0 ; partial |b|^2
template <int P>
__global__ __launch_bounds__(kBlock) void k_cgm_init(int64_t n, const Elem<P>* __restrict__ b, Elem<P>* __restrict__ r,
                                                     const Field* xs, const Field* ps, int nshift, double* partials) {
  using E = Elem<P>;
  double red[1] = {0.0};
  LYM_GRID_STRIDE(i, n) {
    const E v = b[i];
    r[i] = v;
    red[0] += (double)v.x * v.x + (double)v.y * v.y;
    E z;
    z.x = 0;
    z.y = 0;
    for (int k = 0; k < nshift; ++k) {
      ((E*)ps[k].data)[i] = v;
      ((E*)xs[k].data)[i] = z;
    }
  }
  block_partials<1>(red, partials);
}

// x_s += alpha_s p_s ; p_s = zeta_s r + beta_s p_s   for every shift updated this iteration
template <int P>
__global__ __launch_bounds__(kBlock) void k_cgm_k3(const SolverState* st, int64_t n, const Elem<P>* __restrict__ r,
                                                   const Field* xs, const Field* ps, int nshift) {
  using E = Elem<P>;
  using R = typename Store<P>::R;
  if (st->phase != PH_K3) return;
  LYM_GRID_STRIDE(i, n) {
    const E rv = r[i];
    for (int k = 0; k < nshift; ++k) {
      if (!st->upd[k]) continue;
      const R a = (R)st->alpha_s[k], z = (R)st->zeta[k], bt = (R)st->beta_s[k];
      E* xp = (E*)xs[k].data;
      E* pp = (E*)ps[k].data;
      E x = xp[i], p = pp[i];
      x.x += a * p.x;
      x.y += a * p.y;
      p.x = z * rv.x + bt * p.x;
      p.y = z * rv.y + bt * p.y;
      xp[i] = x;
      pp[i] = p;
    }
  }
}

// fast CG-M init (grid.y = RHS): ring slot 0 = b, live p = 0, x_s = 0,
// p_s = b (FP64); partial |b|^2
__global__ __launch_bounds__(kBlock) void k_fast_init(int64_t n, const Field* bs, double2* __restrict__ ring64,
                                                      int64_t ring_stride, double2* __restrict__ plive64,
                                                      const Field* xs, const Field* ps, int nshift, double* partials) {
  const int rhs = blockIdx.y;
  const double2* __restrict__ b = (const double2*)bs[rhs].data;
  double2* __restrict__ r0 = ring64 + rhs * ring_stride;
  double2* __restrict__ plive = plive64 + rhs * n;
  xs += rhs * nshift;
  ps += rhs * nshift;
  partials += (int64_t)rhs * gridDim.x;
  double red[1] = {0.0};
  LYM_GRID_STRIDE(i, n) {
    const double2 v = b[i];
    r0[i] = v;
    red[0] += v.x * v.x + v.y * v.y;
    const double2 z = make_double2(0.0, 0.0);
    plive[i] = z;
    for (int k = 0; k < nshift; ++k) {
      ((double2*)ps[k].data)[i] = v;
      ((double2*)xs[k].data)[i] = z;
    }
  }
  block_partials<1>(red, partials);
}

// p = r^k + beta_{k-1} p   (r^k in ring slot k % nslot)
template <int P>
__global__ __launch_bounds__(kBlock) void k_fast_kp(const SolverState* st, int64_t n, const Elem<P>* __restrict__ ring,
                                                    int64_t ring_stride, int nslot, Elem<P>* __restrict__ p) {
  using E = Elem<P>;
  using R = typename Store<P>::R;
  const int rhs = blockIdx.y;
  st += rhs;
  if (st->phase != PH_KP) return;
  ring += rhs * ring_stride;
  p += rhs * n;
  const E* r = ring + (int64_t)(st->iter % nslot) * n;
  const R b = (R)st->beta_old;
  LYM_GRID_STRIDE(i, n) {
    const E rv = r[i];
    E q = p[i];
    q.x = rv.x + b * q.x;
    q.y = rv.y + b * q.y;
    p[i] = q;
  }
}

// The same for 16-bit storage (per-site norms): one thread per site.
// Ring slot q of RHS i: data + (i * nslot + q) * NCOMP * V, norms + (i * nslot + q) * V.
template <int NCOMP>
__global__ __launch_bounds__(kBlock) void k_fast_kp16(const SolverState* st, int64_t V, Field ring, int nslot,
                                                      Field p) {
  const int rhs = blockIdx.y;
  st += rhs;
  if (st->phase != PH_KP) return;
  const int64_t slot = (int64_t)rhs * nslot + st->iter % nslot;
  const SpinorView<P16> rv{(short2*)ring.data + slot * NCOMP * V, ring.norm + slot * V, V};
  const SpinorView<P16> pv{(short2*)p.data + (int64_t)rhs * NCOMP * V, p.norm + (int64_t)rhs * V, V};
  const float b = (float)st->beta_old;
  for (int64_t s = (int64_t)blockIdx.x * kBlock + threadIdx.x; s < V; s += (int64_t)gridDim.x * kBlock) {
    cpx<float> a[NCOMP], q[NCOMP];
    rv.template load<NCOMP>(s, a);
    pv.template load<NCOMP>(s, q);
#pragma unroll
    for (int c = 0; c < NCOMP; ++c) q[c] = mk<float>(fmad(b, q[c].re, a[c].re), fmad(b, q[c].im, a[c].im));
    pv.template store<NCOMP>(s, q);
  }
}

// Deferred shift updates of one block: every active shift applies its
// recorded steps j < n_upd[s] (x_s += a_j p_s ; p_s = z_j r^{k0+j+1} + b_j p_s,
// x_s and p_s in FP64) in the folded form of fold_shift_block:
//   x_s += A_s p_s + sum_j d_sj r_j ,  p_s = B_s p_s + sum_j e_sj r_j .
// One thread per element (component plane c, site) holds x_s, p_s of a
// group of up to kK3Group shifts in registers and streams the block's
// residuals through them once (complex128, or complex64 / 16-bit widened to
// FP64, a 16-bit ring's site norm multiplying the integers); grid.z runs the
// groups, each reading the residuals again.  Per block the kernel moves the
// nb residuals once per group and 4 FP64 vectors per active shift, so the
// shift traffic per iteration falls as 1/nb with the block length, which the
// ring's HBM footprint (nb + 1 slots) bounds rather than registers.  The
// workgroup first folds its shifts' recorded steps into LDS (one thread per
// shift, the recurrence of fold_shift_block), as (d, e) pairs step-major so
// one 16-byte LDS read serves a shift's two FMAs per component.
// Work item m covers site 64 * (m / 64 / ncomp) + m % 64 of plane
// (m / 64) % ncomp: consecutive waves take the planes of the same 64 sites,
// so the per-site norms of a 16-bit ring are re-read from cache.
constexpr int kK3Block64 = 64;
constexpr int kK3BlockLo = 128;
constexpr int kK3Group = 12;   // shifts per thread: x_s, p_s take 8 VGPRs each
constexpr int kK3Unroll = 2;   // residuals in flight per thread

// groups of at most kK3Group shifts, as even as possible
__host__ __device__ inline int k3_groups(int nact) { return (nact + kK3Group - 1) / kK3Group; }

template <int PR>
__global__ __launch_bounds__(kBlock) void k_fast_k3(const SolverState* st, int ncomp, int64_t V, Field ring,
                                                    int64_t ring_stride, int nslot, const Field* xs, const Field* ps,
                                                    int nshift) {
  constexpr int G = kK3Group, U = kK3Unroll;
  using E = Elem<PR>;
  // LDS (dynamic, k3_lds_bytes): (d, e) [step][group slot], (A, B) per
  // group slot, the active shifts' indices and step counts
  extern __shared__ __align__(16) unsigned char k3_smem[];
  const int mbmax = nslot - 1;
  double2* sde = reinterpret_cast<double2*>(k3_smem);
  double2* sab = sde + (size_t)mbmax * G;
  int* sk = reinterpret_cast<int*>(sab + G);
  int* sm = sk + nshift;
  __shared__ int snact;
  __shared__ double2* sxp[2 * G];
  const int rhs = blockIdx.y;
  st += rhs;
  if (st->phase != PH_K3) return;
  if (threadIdx.x == 0) {
    int na = 0;
    for (int k = 0; k < nshift; ++k)
      if (st->n_upd[k] > 0) {
        sk[na] = k;
        sm[na] = st->n_upd[k];
        ++na;
      }
    snact = na;
  }
  __syncthreads();
  const int nact = snact;
  const int ngrp = k3_groups(nact);
  if ((int)blockIdx.z >= ngrp) return;
  const int per = (nact + ngrp - 1) / ngrp;
  const int g0 = (int)blockIdx.z * per;
  const int ng = min(per, nact - g0);
  const int nb = st->blk_j;
  if ((int)threadIdx.x < ng) {
    const int g = (int)threadIdx.x;
    const int k = sk[g0 + g], mk = sm[g0 + g];
    double A, B;
    fold_shift_block(st->h_alpha[k], st->h_zeta[k], st->h_beta[k], mk, sde + g, G, A, B);
    for (int j = mk; j < nb; ++j) sde[j * G + g] = make_double2(0.0, 0.0);
    sab[g] = make_double2(A, B);
    sxp[g] = (double2*)xs[rhs * nshift + k].data;
    sxp[G + g] = (double2*)ps[rhs * nshift + k].data;
  }
  __syncthreads();
  const E* rg = (const E*)ring.data + rhs * ring_stride;
  const float* rn = ring.norm ? ring.norm + (int64_t)rhs * nslot * V : nullptr;
  // the group's field pointers, read from LDS at every element (held in
  // scalar registers they spill)
  double2** xp = sxp;
  double2** pp = sxp + G;
  const int k0 = st->iter - nb;
  const int64_t n = (int64_t)ncomp * V;
  const int64_t nwork = ((V + 63) / 64) * ncomp * 64;
  LYM_GRID_STRIDE(m, nwork) {
    const int64_t grp = m >> 6;
    const int c = (int)(grp % ncomp);
    const int64_t site = (grp / ncomp) * 64 + (m & 63);
    if (site >= V) continue;
    const int64_t i = (int64_t)c * V + site;
    const int64_t ie = PR == P16 ? p16_index(ncomp, c, V, site) : i;
    double2 x[G], p[G];
#pragma unroll
    for (int g = 0; g < G; ++g)
      if (g < ng) {
        x[g] = xp[g][i];
        p[g] = pp[g][i];
      }
#pragma unroll
    for (int g = 0; g < G; ++g)
      if (g < ng) {
        const double2 ab = sab[g];
        x[g].x = fmad(ab.x, p[g].x, x[g].x);
        x[g].y = fmad(ab.x, p[g].y, x[g].y);
        p[g].x *= ab.y;
        p[g].y *= ab.y;
      }
    for (int j0 = 0; j0 < nb; j0 += U) {
      double rr[U], ri[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        rr[u] = 0.0;
        ri[u] = 0.0;
        if (j0 + u < nb) {
          const int q = (k0 + j0 + u + 1) % nslot;
          const E e = rg[(int64_t)q * n + ie];
          if constexpr (PR == P16) {
            const float sc = rn[(int64_t)q * V + site];
            rr[u] = (double)((float)e.x * sc);
            ri[u] = (double)((float)e.y * sc);
          } else {
            rr[u] = (double)e.x;
            ri[u] = (double)e.y;
          }
        }
      }
#pragma unroll
      for (int u = 0; u < U; ++u) {
        if (j0 + u < nb) {
          const double2* cj = sde + (j0 + u) * G;
#pragma unroll
          for (int g = 0; g < G; ++g)
            if (g < ng) {
              const double2 de = cj[g];
              x[g].x = fmad(de.x, rr[u], x[g].x);
              x[g].y = fmad(de.x, ri[u], x[g].y);
              p[g].x = fmad(de.y, rr[u], p[g].x);
              p[g].y = fmad(de.y, ri[u], p[g].y);
            }
        }
      }
    }
#pragma unroll
    for (int g = 0; g < G; ++g)
      if (g < ng) {
        xp[g][i] = x[g];
        pp[g][i] = p[g];
      }
  }
}

// dynamic LDS of k_fast_k3 for a ring of `nslot` slots and `nshift` shifts
static size_t k3_lds_bytes(int nslot, int nshift) {
  return (size_t)((nslot - 1) * kK3Group + kK3Group) * sizeof(double2) + (size_t)2 * nshift * sizeof(int);
}

// Shared state transition of a precision switch (block (0, rhs), thread 0):
// the next phase, its block length and threshold; the batch tally restarts
// from the RHS that have finished.
__device__ inline void fast_switch_state(SolverState* st, BatchCtl* ctl, int mblk_lo) {
  if (blockIdx.x != 0 || threadIdx.x != 0) return;
  st->precision_phase += 1;
  st->mblk = mblk_lo;
  st->want_switch = 0;
  st->switch_counted = st->phase == PH_DONE ? 1 : 0;
  st->switch_rr = st->switch2_rr;
  st->switch2_rr = 0.0;
  if (blockIdx.y == 0) ctl->nswitch = ctl->ndone;
}

// precision switch, before the new ring is written: every RHS still in a
// block (PH_KP) has it flushed by the shift update that follows
__global__ void k_fast_arm_flush(SolverState* st) {
  if (threadIdx.x != 0) return;
  st += blockIdx.y;
  if (st->phase == PH_KP) st->phase = PH_K3;
}

// precision switch: the current residual (ring slot iter % nslot) copied to
// the old phase's t field, n elements per RHS; the rings of the phases share
// memory, so the new ring's slot may overlap the old one's
template <typename E>
__global__ __launch_bounds__(kBlock) void k_fast_stage(const SolverState* st, int64_t n, const E* __restrict__ ring,
                                                       int64_t ring_stride, int nslot, E* __restrict__ t) {
  const int rhs = blockIdx.y;
  st += rhs;
  ring += rhs * ring_stride + (int64_t)(st->iter % nslot) * n;
  t += rhs * n;
  LYM_GRID_STRIDE(i, n) t[i] = ring[i];
}

// precision switch FP64 -> complex64 after the flush: the staged residual
// r64 (n elements per RHS) goes to low ring slot iter % nslot_lo and the live
// direction is converted; later blocks run mblk_lo iterations
template <int PL>
__global__ __launch_bounds__(kBlock) void k_fast_switch(SolverState* st, BatchCtl* ctl, int64_t n,
                                                        const double2* __restrict__ r64, Elem<PL>* __restrict__ ringlo,
                                                        int64_t ring_stride_lo, int nslot_lo,
                                                        const double2* __restrict__ p64, Elem<PL>* __restrict__ plo,
                                                        int mblk_lo) {
  const int rhs = blockIdx.y;
  st += rhs;
  r64 += rhs * n;
  ringlo += rhs * ring_stride_lo;
  p64 += rhs * n;
  plo += rhs * n;
  const int64_t offlo = (int64_t)(st->iter % nslot_lo) * n;
  LYM_GRID_STRIDE(i, n) {
    ringlo[offlo + i] = from64<Elem<PL>>(r64[i]);
    plo[i] = from64<Elem<PL>>(p64[i]);
  }
  fast_switch_state(st, ctl, mblk_lo);
}

// precision switch complex64 -> 16-bit storage after the flush (one thread
// per site): the staged residual r32 (NCOMP * V elements per RHS) goes to
// 16-bit ring slot iter % nslot16
template <int NCOMP>
__global__ __launch_bounds__(kBlock) void k_fast_switch16(SolverState* st, BatchCtl* ctl, int64_t V,
                                                          const float2* __restrict__ r32, Field ring16, int nslot16,
                                                          const float2* __restrict__ p32, Field p16, int mblk_lo) {
  const int rhs = blockIdx.y;
  st += rhs;
  const int64_t n = (int64_t)NCOMP * V;
  const int64_t q16 = (int64_t)rhs * nslot16 + st->iter % nslot16;
  const SpinorView<P32> ri{const_cast<float2*>(r32) + rhs * n, nullptr, V};
  const SpinorView<P32> pi{const_cast<float2*>(p32) + rhs * n, nullptr, V};
  const SpinorView<P16> ro{(short2*)ring16.data + q16 * n, ring16.norm + q16 * V, V};
  const SpinorView<P16> po{(short2*)p16.data + rhs * n, p16.norm + (int64_t)rhs * V, V};
  for (int64_t s = (int64_t)blockIdx.x * kBlock + threadIdx.x; s < V; s += (int64_t)gridDim.x * kBlock) {
    cpx<float> v[NCOMP];
    ri.template load<NCOMP>(s, v);
    ro.template store<NCOMP>(s, v);
    pi.template load<NCOMP>(s, v);
    po.template store<NCOMP>(s, v);
  }
  fast_switch_state(st, ctl, mblk_lo);
}

// ---- mixed CG with reliable updates: site kernels (grid.y = RHS) -----------
// A thread owns whole sites.  r, p and t = Q p are stored at precision PV
// (complex64, or 16-bit fixed point with per-site norms); the solution is
// accumulated in complex128 (a 16-bit running solution under complex64 r, p
// was measured: point sources on 24^3x48 then need 2.8x the iterations).
// Fields of RHS i start at data + i * NCOMP * V (and norm + i * V).
#define LYM_SITE_STRIDE(s, V) \
  for (int64_t s = (int64_t)blockIdx.x * kBlock + threadIdx.x; s < (V); s += (int64_t)gridDim.x * kBlock)

template <int PV, int NCOMP>
__device__ __forceinline__ SpinorView<PV> vec_view(const Field& f, int64_t V, int rhs) {
  return SpinorView<PV>{(typename Store<PV>::elem*)f.data + (int64_t)rhs * NCOMP * V,
                        f.norm ? f.norm + (int64_t)rhs * V : nullptr, V};
}

// Steps per deferred solution block of the mixed CG, and the slots of its
// direction ring: p_k lives in slot k % kRelBlock.  A full block's fold reads
// p_{k0}..p_{k0+15} at a site before the direction update of step k0+16
// overwrites slot k0 % 16 there (same thread), so the ring needs no spare slot.
constexpr int kRelBlock = 16;
static_assert(kRelBlock <= kMaxBlock && kRelBlock <= kBlock, "mixed CG block exceeds the recorded steps");

// slot q of RHS rhs in the mixed CG's direction ring (nslot slots per RHS)
template <int PV, int NCOMP>
__device__ __forceinline__ SpinorView<PV> ring_view(const Field& f, int64_t V, int rhs, int nslot, int q) {
  const int64_t k = (int64_t)rhs * nslot + q;
  return SpinorView<PV>{(typename Store<PV>::elem*)f.data + k * NCOMP * V, f.norm ? f.norm + k * V : nullptr, V};
}

// init: r = b, p_{-1} = 0 (ring slot nslot - 1), x_hi = 0; partial |b|^2 (FP64)
template <int PV, int NCOMP>
__global__ __launch_bounds__(kBlock) void k_rel_init(int64_t V, const Field* bs, Field r, Field ring, int nslot,
                                                     double2* __restrict__ x64, double* partials) {
  const int rhs = blockIdx.y;
  const SpinorView<P64> b{(double2*)bs[rhs].data, nullptr, V};
  const SpinorView<PV> rv = vec_view<PV, NCOMP>(r, V, rhs), pv = ring_view<PV, NCOMP>(ring, V, rhs, nslot, nslot - 1);
  const SpinorView<P64> xh{x64 + (int64_t)rhs * NCOMP * V, nullptr, V};
  partials += (int64_t)rhs * gridDim.x;
  double red[1] = {0.0};
  LYM_SITE_STRIDE(s, V) {
    cpx<double> v[NCOMP];
    b.template load<NCOMP>(s, v);
    cpx<float> lo[NCOMP], z[NCOMP];
    cpx<double> z64[NCOMP];
#pragma unroll
    for (int c = 0; c < NCOMP; ++c) {
      red[0] += v[c].re * v[c].re + v[c].im * v[c].im;
      lo[c] = mk<float>((float)v[c].re, (float)v[c].im);
      z[c] = czero<float>();
      z64[c] = czero<double>();
    }
    rv.template store<NCOMP>(s, lo);
    pv.template store<NCOMP>(s, z);
    xh.template store<NCOMP>(s, z64);
  }
  block_partials<1>(red, partials);
}

// Deferred solution update of the mixed CG: a block of nb steps k0..k0+nb-1
// (k0 = iter - nb) adds sum_j alpha_j p_{k0+j} to x_hi, with the stored
// directions of the ring (the ones the iteration used), accumulated in FP64.
__device__ inline void rel_fold_alphas(const SolverState* st, int nb, double* sa) {
  if ((int)threadIdx.x < nb) sa[threadIdx.x] = st->h_alpha[0][threadIdx.x];
  __syncthreads();
}

template <int PV, int NCOMP>
__device__ __forceinline__ void rel_fold_site(int64_t s, int64_t V, const Field& ring, int nslot, int rhs, int k0,
                                              int nb, const double* sa, const SpinorView<P64>& xh) {
  cpx<double> w[NCOMP];
  cpx<float> v[NCOMP];
  xh.template load<NCOMP>(s, w);
  for (int j = 0; j < nb; ++j) {
    ring_view<PV, NCOMP>(ring, V, rhs, nslot, (k0 + j) % nslot).template load<NCOMP>(s, v);
    const double a = sa[j];
#pragma unroll
    for (int c = 0; c < NCOMP; ++c)
      w[c] = mk<double>(fma(a, (double)v[c].re, w[c].re), fma(a, (double)v[c].im, w[c].im));
  }
  xh.template store<NCOMP>(s, w);
}

// direction update (gate KP): p_k = kp_a r + beta p_{k-1} into ring slot
// k % nslot (k = iter).  When the block is full (rel_fold_due) its solution
// update is applied first and a new block starts at step k (k_finalize's
// FK_CG_A clears the block counter).
template <int PV, int NCOMP>
__global__ __launch_bounds__(kBlock) void k_rel_kp(const SolverState* st, int64_t V, Field r, Field ring, int nslot,
                                                   double2* __restrict__ x64) {
  const int rhs = blockIdx.y;
  st += rhs;
  if (st->phase != PH_KP) return;
  const int nb = st->blk_j, k = st->iter;
  const bool fold = rel_fold_due(st);
  __shared__ double sa[kRelBlock];
  if (fold) rel_fold_alphas(st, nb, sa);
  const SpinorView<PV> rv = vec_view<PV, NCOMP>(r, V, rhs),
                       po = ring_view<PV, NCOMP>(ring, V, rhs, nslot, (k + nslot - 1) % nslot),
                       pn = ring_view<PV, NCOMP>(ring, V, rhs, nslot, k % nslot);
  const SpinorView<P64> xh{x64 + (int64_t)rhs * NCOMP * V, nullptr, V};
  const float b = (float)st->beta_old;
  const float ar = (float)st->kp_a_re, ai = (float)st->kp_a_im;
  LYM_SITE_STRIDE(s, V) {
    if (fold) rel_fold_site<PV, NCOMP>(s, V, ring, nslot, rhs, k - nb, nb, sa, xh);
    cpx<float> a[NCOMP], q[NCOMP];
    po.template load<NCOMP>(s, q);
    rv.template load<NCOMP>(s, a);
#pragma unroll
    for (int c = 0; c < NCOMP; ++c)
      q[c] = mk<float>(ar * a[c].re - ai * a[c].im + b * q[c].re, ar * a[c].im + ai * a[c].re + b * q[c].im);
    pn.template store<NCOMP>(s, q);
  }
}

// reliable update, accumulation (gate RU, rel_fold_due): the solution update
// of the steps since the block started goes into x_hi (k_finalize's FK_CG_RU
// clears the block counter)
template <int PV, int NCOMP>
__global__ __launch_bounds__(kBlock) void k_ru_fold(const SolverState* st, int64_t V, Field ring, int nslot,
                                                    double2* __restrict__ x64) {
  const int rhs = blockIdx.y;
  st += rhs;
  if (st->phase != PH_RU || !rel_fold_due(st)) return;
  const int nb = st->blk_j, k = st->iter;
  __shared__ double sa[kRelBlock];
  rel_fold_alphas(st, nb, sa);
  const SpinorView<P64> xh{x64 + (int64_t)rhs * NCOMP * V, nullptr, V};
  LYM_SITE_STRIDE(s, V) rel_fold_site<PV, NCOMP>(s, V, ring, nslot, rhs, k - nb, nb, sa, xh);
}

// result: xs[rhs] = x_hi
template <int NCOMP>
__global__ __launch_bounds__(kBlock) void k_rel_out(int64_t V, const double2* __restrict__ x64, const Field* xs) {
  const int rhs = blockIdx.y;
  const SpinorView<P64> xh{const_cast<double2*>(x64) + (int64_t)rhs * NCOMP * V, nullptr, V};
  const SpinorView<P64> out{(double2*)xs[rhs].data, nullptr, V};
  LYM_SITE_STRIDE(s, V) {
    cpx<double> w[NCOMP];
    xh.template load<NCOMP>(s, w);
    out.template store<NCOMP>(s, w);
  }
}

// ---- MR (complex128) -------------------------------------------------------
__global__ __launch_bounds__(kBlock) void k_mr_init(int64_t n, const double2* __restrict__ b, double2* __restrict__ r,
                                                    const Field* xs, int first, int nshift) {
  LYM_GRID_STRIDE(i, n) {
    r[i] = b[i];
    for (int k = first; k < nshift; ++k) ((double2*)xs[k].data)[i] = make_double2(0.0, 0.0);
  }
}

// x_i += c_i r ; r -= alpha p ; partial |r|^2   (dirac_solver.py:37-47)
__global__ __launch_bounds__(kBlock) void k_mr_k3(const SolverState* st, int64_t n, double2* __restrict__ r,
                                                  const double2* __restrict__ p, const Field* xs, int first,
                                                  int nshift, double* partials) {
  if (st->phase != PH_K3) return;
  double red[1] = {0.0};
  const double ar = st->mr_alpha_re, ai = st->mr_alpha_im;
  LYM_GRID_STRIDE(i, n) {
    double2 rv = r[i];
    for (int k = first; k < nshift; ++k) {
      const double cr = st->c_re[k], ci = st->c_im[k];
      double2* xp = (double2*)xs[k].data;
      double2 x = xp[i];
      x.x += cr * rv.x - ci * rv.y;
      x.y += cr * rv.y + ci * rv.x;
      xp[i] = x;
    }
    const double2 pv = p[i];
    rv.x -= ar * pv.x - ai * pv.y;
    rv.y -= ar * pv.y + ai * pv.x;
    r[i] = rv;
    red[0] += rv.x * rv.x + rv.y * rv.y;
  }
  block_partials<1>(red, partials);
}

// ---- BiCGStab (complex128, dirac_solver.py:56-110) --------------------------
// Every kernel runs only while the solve is running (phase PH_K1).
__device__ __forceinline__ double2 cmul2(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

// r = b - A x0 (init holds A x0), r_hat = x0 + A x0, p = nu = 0;
// partials |r|^2, <r_hat, r> (re, im)
__global__ __launch_bounds__(kBlock) void k_bicg_init(const SolverState* st, int64_t n, const double2* __restrict__ b,
                                                      const double2* __restrict__ x, const double2* __restrict__ init,
                                                      double2* __restrict__ r, double2* __restrict__ rh,
                                                      double2* __restrict__ p, double2* __restrict__ nu,
                                                      double* partials) {
  double v[3] = {0.0, 0.0, 0.0};
  if (st->phase == PH_K1) {
    LYM_GRID_STRIDE(i, n) {
      const double2 a = init[i], bv = b[i], xv = x[i];
      const double2 rv = make_double2(bv.x - a.x, bv.y - a.y);
      const double2 hv = make_double2(xv.x + a.x, xv.y + a.y);
      r[i] = rv;
      rh[i] = hv;
      p[i] = make_double2(0.0, 0.0);
      nu[i] = make_double2(0.0, 0.0);
      v[0] += rv.x * rv.x + rv.y * rv.y;
      v[1] += hv.x * rv.x + hv.y * rv.y;  // <r_hat, r> = sum conj(r_hat) r
      v[2] += hv.x * rv.y - hv.y * rv.x;
    }
  }
  block_partials<3>(v, partials);
}

// p = r + beta (p - omega nu)
__global__ __launch_bounds__(kBlock) void k_bicg_p(const SolverState* st, int64_t n, const double2* __restrict__ r,
                                                   const double2* __restrict__ nu, double2* __restrict__ p) {
  if (st->phase != PH_K1) return;
  const double2 be = make_double2(st->bi_beta[0], st->bi_beta[1]);
  const double2 om = make_double2(st->bi_omega[0], st->bi_omega[1]);
  LYM_GRID_STRIDE(i, n) {
    const double2 on = cmul2(om, nu[i]);
    const double2 pv = p[i];
    const double2 q = cmul2(be, make_double2(pv.x - on.x, pv.y - on.y));
    const double2 rv = r[i];
    p[i] = make_double2(rv.x + q.x, rv.y + q.y);
  }
}

// partials <a, c> (re, im) = sum conj(a) c, and |a|^2
__global__ __launch_bounds__(kBlock) void k_bicg_dots(const SolverState* st, int64_t n, const double2* __restrict__ a,
                                                      const double2* __restrict__ c, double* partials) {
  double v[3] = {0.0, 0.0, 0.0};
  if (st->phase == PH_K1) {
    LYM_GRID_STRIDE(i, n) {
      const double2 av = a[i], cv = c[i];
      v[0] += av.x * cv.x + av.y * cv.y;
      v[1] += av.x * cv.y - av.y * cv.x;
      v[2] += av.x * av.x + av.y * av.y;
    }
  }
  block_partials<3>(v, partials);
}

// s = r - alpha nu
__global__ __launch_bounds__(kBlock) void k_bicg_s(const SolverState* st, int64_t n, const double2* __restrict__ r,
                                                   const double2* __restrict__ nu, double2* __restrict__ s) {
  if (st->phase != PH_K1) return;
  const double2 al = make_double2(st->bi_alpha[0], st->bi_alpha[1]);
  LYM_GRID_STRIDE(i, n) {
    const double2 an = cmul2(al, nu[i]);
    const double2 rv = r[i];
    s[i] = make_double2(rv.x - an.x, rv.y - an.y);
  }
}

// x += alpha p + omega s ; r = s - omega t ; partials |r|^2, <r_hat, r>
__global__ __launch_bounds__(kBlock) void k_bicg_xr(const SolverState* st, int64_t n, double2* __restrict__ x,
                                                    const double2* __restrict__ p, const double2* __restrict__ s,
                                                    const double2* __restrict__ t, const double2* __restrict__ rh,
                                                    double2* __restrict__ r, double* partials) {
  double v[3] = {0.0, 0.0, 0.0};
  if (st->phase == PH_K1) {
    const double2 al = make_double2(st->bi_alpha[0], st->bi_alpha[1]);
    const double2 om = make_double2(st->bi_omega[0], st->bi_omega[1]);
    LYM_GRID_STRIDE(i, n) {
      const double2 sv = s[i];
      const double2 ap = cmul2(al, p[i]), os = cmul2(om, sv), ot = cmul2(om, t[i]);
      double2 xv = x[i];
      xv.x += ap.x + os.x;
      xv.y += ap.y + os.y;
      x[i] = xv;
      const double2 rv = make_double2(sv.x - ot.x, sv.y - ot.y);
      r[i] = rv;
      const double2 hv = rh[i];
      v[0] += rv.x * rv.x + rv.y * rv.y;
      v[1] += hv.x * rv.x + hv.y * rv.y;
      v[2] += hv.x * rv.y - hv.y * rv.x;
    }
  }
  block_partials<3>(v, partials);
}

// ---- FP64 helpers for true residuals and defect correction -------------------
// r = b - ax ; partials |r|^2, |b|^2
__global__ __launch_bounds__(kBlock) void k_residual64(int64_t n, const double2* __restrict__ b,
                                                       const double2* __restrict__ ax, double2* __restrict__ r,
                                                       double* partials) {
  double v[2] = {0.0, 0.0};
  LYM_GRID_STRIDE(i, n) {
    const double2 bv = b[i], a = ax[i];
    double2 d;
    d.x = bv.x - a.x;
    d.y = bv.y - a.y;
    if (r) r[i] = d;
    v[0] += d.x * d.x + d.y * d.y;
    v[1] += bv.x * bv.x + bv.y * bv.y;
  }
  block_partials<2>(v, partials);
}

// x64 += e32
__global__ __launch_bounds__(kBlock) void k_add32to64(int64_t n, const float2* __restrict__ e, double2* __restrict__ x) {
  LYM_GRID_STRIDE(i, n) {
    const float2 a = e[i];
    double2 b = x[i];
    b.x += (double)a.x;
    b.y += (double)a.y;
    x[i] = b;
  }
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

// bytes per site of one spinor at storage precision prec
static double vec_bytes_site(const OpDesc& d, int prec) {
  const double ws = prec == P64 ? 8.0 : prec == P32 ? 4.0 : 2.0;
  return d.ncomp * 2.0 * ws + (prec == P16 ? 4.0 : 0.0);
}

static Field take_field(Carver& c, const OpDesc& d, int prec) {
  Field f;
  f.data = c.take(field_bytes(d, prec));
  if (prec == P16) f.norm = (float*)c.take((size_t)d.lat.V * sizeof(float));
  return f;
}

// Iterations per deferred shift-update block (CgmTuning::mblk_64 and
// mblk_lo, at most kMaxBlock; each ring has one slot more than its block is
// long).  A block reads its residuals once per shift group and every active
// shift's FP64 x_s, p_s once (4 vectors per shift), so the shift traffic per
// iteration falls as 1/m; the ring's footprint is m + 1 fields per RHS.

// Thresholds of the fast multishift CG (relative to |b|): the base residual
// below which complex64 Krylov vectors (switch32) and 16-bit storage
// (switch16) take over, and the iterated tolerance as a fraction of the
// target (the FP64 true residual decides at the end).  A rounding error of
// relative size e made while the base residual is r shows up in the shifted
// true residuals at about e r, summed over the iterations of the phase; a
// shift that still misses the target gets FP64 defect correction.  Measured
// on MI355X (solves/s; every shift to 1e-10 after correction):
//   switch32, switch16   32^3x64, 24 shifts   24^3x48, 8 sources x 12 shifts
//   3e-5, 1e-7           1.099                5.22
//   1e-4, 1e-7           1.125                5.45
//   1e-4, 1e-6           1.123                5.53
//   3e-4, 1e-7           1.059                5.57
//   3e-4, 1e-6           1.173                5.72
//   1e-3, 1e-6           0.866 (8 corrections, 519 it)   5.77
// Iterations run per host chunk (one graph replay, one poll of the batch
// state) are fewer than a block: a precision switch waits for the chunk in
// flight, so the chunk length bounds how far a phase overruns its threshold.
// Blocks are whole numbers of chunks; a switch flushes the block in progress.
// LYM_CGM_TUNE="switch32,switch16,iter_factor[,mblk_64,mblk_lo]" overrides
// the thresholds and block lengths (tuning runs).
struct CgmTuning {
  double switch32 = 3e-4;
  double switch16 = 1e-6;
  double iter_factor = 0.5;
  int mblk_lo = kK3BlockLo;  // block length of the complex64 / 16-bit phases
  int mblk_64 = kK3Block64;  // block length of the FP64 phase
  int chunk_lo = 32;         // iterations per chunk, complex64 / 16-bit phases
  int chunk_64 = 16;         // iterations per chunk, FP64 phase
};
static CgmTuning cgm_tuning() {
  CgmTuning t;
  const char* v = std::getenv("LYM_CGM_TUNE");
  if (v && *v) {
    double a = 0, b = 0, c = 0;
    int m64 = 0, mlo = 0;
    const int got = std::sscanf(v, "%lf,%lf,%lf,%d,%d", &a, &b, &c, &m64, &mlo);
    if (got >= 3 && a > 0 && b > 0 && c > 0 && c <= 1) {
      t.switch32 = a;
      t.switch16 = b;
      t.iter_factor = c;
    }
    if (got == 5 && m64 >= t.chunk_64 && m64 <= kMaxBlock && m64 % t.chunk_64 == 0 && mlo >= t.chunk_lo &&
        mlo <= kMaxBlock && mlo % t.chunk_lo == 0) {
      t.mblk_64 = m64;
      t.mblk_lo = mlo;
    }
  }
  return t;
}

struct Parts {
  SolverState* st;  // nrhs states
  BatchCtl* ctl;
  double* partials;  // nrhs x kMaxRed x max_red_blocks
  double* results;
  Field* xs_dev;  // nrhs x nshift
  Field* ps_dev;  // nrhs x nshift
  Field* bs_dev;  // nrhs sources
  int nrhs = 1;
  Field r64, t64, u64, v64;  // single-field FP64 temporaries (true residuals, refinement)
  Field b_lo, r_lo, t_lo;    // low-precision single-shift defect correction
  Field x_lo[kMaxShifts];
  Field p[kMaxShifts];  // plain CG-M search directions
  // fast CG-M (mixed), every field batched over nrhs RHS:
  Field ring64, ring32, ring16;   // nrhs x nslot64 / nslot_lo slots
  int nslot64 = 0, nslot_lo = 0;
  Field plive64, plive32, plive16;  // live base direction
  Field tb64, tb32, tb16;           // t = Q p
  Field ps64;               // nrhs x nshift FP64 shift directions
  // mixed CG with reliable updates (one shift), every field nrhs x batched:
  // residual, direction ring (rl_nslot slots), Q p (complex64 or 16-bit
  // storage)
  Field rl_r, rl_ring, rl_t;
  int rl_nslot = 0;
  Field rl_x64, rl_t64;  // complex128 accumulated solution, Q x
};

// prec: P64 = plain FP64 solves (and MR); otherwise the fast mixed solver
static size_t carve(const OpDesc& d, int prec, int nshift, int nrhs, char* base, Parts* pt) {
  Carver c{base};
  Parts tmp;
  Parts& P = pt ? *pt : tmp;
  P.nrhs = nrhs;
  P.st = (SolverState*)c.take(sizeof(SolverState) * nrhs);
  P.ctl = (BatchCtl*)c.take(sizeof(BatchCtl));
  P.partials = (double*)c.take(sizeof(double) * kMaxRed * (size_t)max_red_blocks(d.lat.V) * nrhs);
  P.results = (double*)c.take(sizeof(double) * 16);
  P.xs_dev = (Field*)c.take(sizeof(Field) * kMaxShifts * nrhs);
  P.ps_dev = (Field*)c.take(sizeof(Field) * kMaxShifts * nrhs);
  P.bs_dev = (Field*)c.take(sizeof(Field) * nrhs);
  P.r64 = take_field(c, d, P64);
  P.t64 = take_field(c, d, P64);
  P.u64 = take_field(c, d, P64);
  P.v64 = take_field(c, d, P64);
  if (prec == P64) {
    for (int k = 0; k < nshift; ++k) P.p[k] = take_field(c, d, P64);
  } else {
    P.b_lo = take_field(c, d, P32);
    P.r_lo = take_field(c, d, P32);
    P.t_lo = take_field(c, d, P32);
    P.x_lo[0] = take_field(c, d, P32);
    P.p[0] = take_field(c, d, P32);
    if (nshift == 1) {
      // one shift: CG with reliable updates, Krylov vectors in complex64 or
      // half storage (per-site norms, nrhs x V floats)
      const int pl = prec == P16 ? P16 : P32;
      auto take_batch = [&](int q, int nf) {
        Field f;
        f.data = c.take(field_bytes(d, q) * nf * nrhs);
        if (q == P16) f.norm = (float*)c.take((size_t)d.lat.V * sizeof(float) * nf * nrhs);
        return f;
      };
      P.rl_nslot = kRelBlock;
      P.rl_r = take_batch(pl, 1);
      P.rl_ring = take_batch(pl, P.rl_nslot);
      P.rl_t = take_batch(pl, 1);
      // the solution is accumulated in complex128 only (each block's update
      // is folded into it); a 16-bit running solution would perturb it by
      // ~3e-5 relative per site, which the next reliable update sees as a
      // residual of ~3e-5 cond(A)
      P.rl_x64 = take_batch(P64, 1);
      P.rl_t64 = take_batch(P64, 1);
    } else {
      const CgmTuning tun = cgm_tuning();
      P.nslot64 = tun.mblk_64 + 1;
      P.nslot_lo = tun.mblk_lo + 1;
      // the rings of the three phases are used one after the other and share
      // one allocation (a precision switch stages the current residual in the
      // old phase's t field before writing it to the new ring)
      const size_t nbv = (size_t)d.lat.V * sizeof(float);
      const size_t b64 = field_bytes(d, P64) * P.nslot64 * nrhs;
      const size_t b32 = field_bytes(d, P32) * P.nslot_lo * nrhs;
      const size_t b16d = field_bytes(d, P16) * P.nslot_lo * nrhs;
      const size_t b16 = prec == P16 ? b16d + nbv * P.nslot_lo * nrhs : 0;
      char* ring = (char*)c.take(std::max(b64, std::max(b32, b16)));
      P.ring64.data = ring;
      P.ring32.data = ring;
      P.plive64.data = c.take(field_bytes(d, P64) * nrhs);
      P.plive32.data = c.take(field_bytes(d, P32) * nrhs);
      P.tb64.data = c.take(field_bytes(d, P64) * nrhs);
      P.tb32.data = c.take(field_bytes(d, P32) * nrhs);
      P.ps64.data = c.take(field_bytes(d, P64) * nshift * nrhs);
      if (prec == P16) {
        // third phase: 16-bit storage with per-site norms
        P.ring16.data = ring;
        P.ring16.norm = ring ? (float*)(ring + b16d) : nullptr;
        P.plive16.data = c.take(field_bytes(d, P16) * nrhs);
        P.plive16.norm = (float*)c.take(nbv * nrhs);
        P.tb16.data = c.take(field_bytes(d, P16) * nrhs);
        P.tb16.norm = (float*)c.take(nbv * nrhs);
      }
    }
  }
  return c.off + 256;
}

size_t solver_workspace_bytes(const OpDesc& d, int prec, int nshift, int nrhs) {
  return carve(d, prec, nshift, nrhs, nullptr, nullptr);
}

static RedCtl redctl(const Parts& P) { return RedCtl{P.partials, nullptr}; }

// ---------------------------------------------------------------------------
// host polling: pinned (phase, iter, want_switch, status) per in-flight chunk
// ---------------------------------------------------------------------------
struct HostSync {
  int* pinned = nullptr;
  hipEvent_t ev[2] = {nullptr, nullptr};
  int device = -1;
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

// ---------------------------------------------------------------------------
// chunk graphs: a chunk is a fixed launch sequence whose behaviour is driven
// by device state, so it is captured once into a HIP graph and replayed with
// one hipGraphLaunch per chunk instead of tens of kernel launches.  The key
// holds everything the launches bake in (operator description, workspace,
// solver kind, precision, batch).  Per host thread: at most kGraphCache
// executables; the least recently used one not launched by the current solve
// is dropped (a solve ends with a stream synchronisation, so earlier solves'
// graphs have finished).  LYM_NO_GRAPH=1 in the environment launches the
// kernels directly.
// ---------------------------------------------------------------------------
constexpr int kGraphCache = 8;

struct KeyBuf {
  std::string s;
  template <class T>
  KeyBuf& add(const T& v) {
    s.append(reinterpret_cast<const char*>(&v), sizeof(T));
    return *this;
  }
};

static std::string op_key(const OpDesc& d) {
  KeyBuf k;
  k.add(d.ndim).add(d.nc).add(d.adj).add(d.herm).add(d.ncomp);
  for (int mu = 0; mu < 4; ++mu) k.add(d.lat.L[mu]);
  k.add(d.lat.V).add(d.lat.kappa).add(d.lat.bc_re).add(d.lat.bc_im).add(d.lat.has_bc);
  k.add(d.map.tiled).add(d.map.k).add(d.wblocks).add(d.wthreads).add(d.use_tile).add(d.use_mr).add(d.link_su);
  for (const TileMap* t : {&d.tile[0], &d.tile[1], &d.tile_mr})
    k.add(t->on).add(t->a0).add(t->a1).add(t->k).add(t->hif);
  k.add(d.links64).add(d.linksr64).add(d.links32).add(d.linksr32);
  return k.s;
}

struct GraphEntry {
  std::string key;
  hipGraphExec_t exec = nullptr;
  uint64_t used = 0;   // tick of the last launch
  uint64_t solve = 0;  // solve that last launched it
};
struct GraphCache {
  std::vector<GraphEntry> e;
  hipStream_t cap = nullptr;
  int device = -1;
  uint64_t tick = 0, solve = 0;
};
static thread_local GraphCache g_graphs;

static bool kt_on();
static bool graphs_enabled() {
  const char* v = std::getenv("LYM_NO_GRAPH");
  return !(v && v[0] == '1') && !kt_on();
}

// a new solve: graphs launched before it have completed
static void graph_new_solve() { ++g_graphs.solve; }

// Launch `chunk` on `st` through its cached graph (captured on first use).
template <class F>
static int launch_chunk(const std::string& key, hipStream_t st, F&& chunk) {
  if (!graphs_enabled()) return chunk(st);
  GraphCache& G = g_graphs;
  int dev = 0;
  LYM_HIP(hipGetDevice(&dev));
  if (G.device != dev) {  // executables and the capture stream belong to one device
    for (auto& x : G.e) (void)hipGraphExecDestroy(x.exec);
    G.e.clear();
    if (G.cap) (void)hipStreamDestroy(G.cap);
    G.cap = nullptr;
    G.device = dev;
  }
  GraphEntry* hit = nullptr;
  for (auto& x : G.e)
    if (x.key == key) hit = &x;
  if (!hit) {
    if (!G.cap) LYM_HIP(hipStreamCreateWithFlags(&G.cap, hipStreamNonBlocking));
    LYM_HIP(hipStreamBeginCapture(G.cap, hipStreamCaptureModeThreadLocal));
    const int rc = chunk(G.cap);
    hipGraph_t g = nullptr;
    const hipError_t ec = hipStreamEndCapture(G.cap, &g);
    if (rc || ec != hipSuccess) {
      if (g) (void)hipGraphDestroy(g);
      return rc ? rc : hip_fail(ec, "hipStreamEndCapture");
    }
    hipGraphExec_t ex = nullptr;
    const hipError_t ei = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (ei != hipSuccess) return hip_fail(ei, "hipGraphInstantiate");
    if ((int)G.e.size() >= kGraphCache) {
      int victim = -1;
      for (int i = 0; i < (int)G.e.size(); ++i)
        if (G.e[i].solve != G.solve && (victim < 0 || G.e[i].used < G.e[victim].used)) victim = i;
      if (victim >= 0) {
        (void)hipGraphExecDestroy(G.e[victim].exec);
        G.e.erase(G.e.begin() + victim);
      }
    }
    G.e.push_back(GraphEntry{key, ex, 0, 0});
    hit = &G.e.back();
  }
  hit->used = ++G.tick;
  hit->solve = G.solve;
  LYM_HIP(hipGraphLaunch(hit->exec, st));
  return LYM_OK;
}

// Enqueue `chunk` (a fixed launch sequence), keeping two chunks in flight,
// until the device reports PH_DONE.  `on_poll(status4)` may enqueue extra
// work between chunks (e.g. a precision switch); `max_chunks` guards the loop.
template <class F, class G>
static int run_chunks(SolverState* st_dev, int64_t max_chunks, hipStream_t stream, const std::string& key, F&& chunk,
                      G&& on_poll) {
  int rc = ensure_sync();
  if (rc) return rc;
  int64_t issued = 0, polled = 0;
  auto issue = [&](int slot) -> int {
    const int e = launch_chunk(key, stream, chunk);
    if (e) return e;
    LYM_HIP(hipMemcpyAsync(&g_sync.pinned[4 * slot], st_dev, 4 * sizeof(int), hipMemcpyDeviceToHost, stream));
    LYM_HIP(hipEventRecord(g_sync.ev[slot], stream));
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
    if (issued > max_chunks) {
      set_error("solver did not reach its stop state (internal error)");
      return LYM_ERR_HIP;
    }
    if ((rc = on_poll(s4))) return rc;
    if ((rc = issue(slot))) return rc;
  }
  LYM_HIP(hipStreamSynchronize(stream));
  return LYM_OK;
}

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
  const int g = red_grid_flat(n);
  hipLaunchKernelGGL(k_residual64, dim3(g), dim3(kBlock), 0, st, n, (const double2*)b.data,
                     (const double2*)P.u64.data, (double2*)out.data, P.partials);
  LYM_HIP(hipGetLastError());
  if ((rc = launch_finalize(FK_SUM, nullptr, P.partials, g, 2, P.results, st))) return rc;
  double h[2];
  LYM_HIP(hipMemcpyAsync(h, P.results, sizeof(h), hipMemcpyDeviceToHost, st));
  LYM_HIP(hipStreamSynchronize(st));
  *rel2 = h[1] > 0 ? h[0] / h[1] : 0.0;
  return LYM_OK;
}

static void init_cgm_state(SolverState& h, int nshift, const double* shifts, const double* tol_rel, int max_iter) {
  std::memset(&h, 0, sizeof(h));
  int base = 0;
  for (int k = 1; k < nshift; ++k)
    if (shifts[k] < shifts[base]) base = k;
  h.max_iter = max_iter;
  h.nshift = nshift;
  h.base = base;
  h.sigma0 = shifts[base];
  h.alpha_old = 1.0;
  h.kp_a_re = 1.0;
  h.beta_old = 0.0;
  for (int k = 0; k < nshift; ++k) {
    h.shift[k] = shifts[k];
    h.delta[k] = shifts[k] - shifts[base];
    h.zeta[k] = h.zeta_old[k] = 1.0;
    h.active[k] = 1;
    h.tol2_s[k] = tol_rel[k] * tol_rel[k];
    h.conv_iter[k] = -1;
  }
}

// Plain CG-M at Krylov precision `prec` (P64 or P32): b, xs[], r, t, p[] all at `prec`.
static int cgm_core(const OpDesc& d, int prec, const Parts& P, Field b, const Field* xs, Field r, Field t,
                    int nshift, const double* shifts, const double* tol_rel, int max_iter, hipStream_t st,
                    SolverState* out_state) {
  graph_new_solve();
  SolverState h;
  init_cgm_state(h, nshift, shifts, tol_rel, max_iter);
  h.phase = PH_K1;
  Field hx[kMaxShifts], hp[kMaxShifts];
  for (int k = 0; k < nshift; ++k) {
    hx[k] = xs[k];
    hp[k] = P.p[k];
  }
  LYM_HIP(hipMemcpyAsync(P.st, &h, sizeof(h), hipMemcpyHostToDevice, st));
  LYM_HIP(hipMemcpyAsync(P.xs_dev, hx, sizeof(Field) * nshift, hipMemcpyHostToDevice, st));
  LYM_HIP(hipMemcpyAsync(P.ps_dev, hp, sizeof(Field) * nshift, hipMemcpyHostToDevice, st));
  const int64_t n = field_elems(d);
  const int gw = wilson_blocks(d, prec);  // Wilson grid: number of partials its reductions write
  const int gr = red_grid_flat(n);
  const int gf = flat_grid(n);
  if (prec == P64)
    hipLaunchKernelGGL(k_cgm_init<P64>, dim3(gr), dim3(kBlock), 0, st, n, (const double2*)b.data, (double2*)r.data,
                       P.xs_dev, P.ps_dev, nshift, P.partials);
  else
    hipLaunchKernelGGL(k_cgm_init<P32>, dim3(gr), dim3(kBlock), 0, st, n, (const float2*)b.data, (float2*)r.data,
                       P.xs_dev, P.ps_dev, nshift, P.partials);
  LYM_HIP(hipGetLastError());
  int rc = launch_finalize(FK_INIT, P.st, P.partials, gr, 1, nullptr, st);
  if (rc) return rc;
  const Field pbase = P.p[h.base];
  const int per_chunk = 8;
  auto chunk = [&](hipStream_t st) -> int {
    for (int j = 0; j < per_chunk; ++j) {
      EpiArgs a1;
      a1.out = t;
      a1.st = P.st;
      int e = wilson_launch(EPI_CGM_K1, d, prec, true, pbase, a1, redctl(P), st);
      if (!e) e = launch_finalize(FK_CGM_A, P.st, P.partials, gw, 2, nullptr, st);
      if (e) return e;
      EpiArgs a2;
      a2.aux0 = pbase;
      a2.aux1 = r;
      a2.st = P.st;
      e = wilson_launch(EPI_CGM_K2, d, prec, true, t, a2, redctl(P), st);
      if (!e) e = launch_finalize(FK_CGM_B, P.st, P.partials, gw, 1, nullptr, st);
      if (e) return e;
      if (prec == P64)
        hipLaunchKernelGGL(k_cgm_k3<P64>, dim3(gf), dim3(kBlock), 0, st, P.st, n, (const double2*)r.data, P.xs_dev,
                           P.ps_dev, nshift);
      else
        hipLaunchKernelGGL(k_cgm_k3<P32>, dim3(gf), dim3(kBlock), 0, st, P.st, n, (const float2*)r.data, P.xs_dev,
                           P.ps_dev, nshift);
      LYM_HIP(hipGetLastError());
      if ((e = launch_finalize(FK_CGM_K3ADV, P.st, nullptr, 0, 0, nullptr, st))) return e;
    }
    return LYM_OK;
  };
  KeyBuf key;
  key.add('C').add(prec).add(nshift).add(P.st).add(P.xs_dev).add(r.data).add(t.data).add(pbase.data);
  rc = run_chunks(P.st, (int64_t)max_iter / per_chunk + 8, st, op_key(d) + key.s, chunk,
                  [](const int*) { return (int)LYM_OK; });
  if (rc) return rc;
  LYM_HIP(hipMemcpyAsync(out_state, P.st, sizeof(SolverState), hipMemcpyDeviceToHost, st));
  LYM_HIP(hipStreamSynchronize(st));
  return LYM_OK;
}

// ---------------------------------------------------------------------------
// Kernel timing of the fast multishift CG (lym_kernel_timing): when enabled
// on this host thread, every Wilson, direction-update and shift-update launch
// of the solve is bracketed by HIP events on the solve's stream (the chunks
// are then launched directly, not replayed from graphs), and the per-class
// durations and algorithmic HBM bytes are added up when the solve ends.
// Algorithmic bytes per launch: each link (the rows the kernel reads) and
// each input / output vector once; the shift update's bytes follow from the
// block's residuals and the shifts still active in it (conv_iter).
// ---------------------------------------------------------------------------
struct KClass {
  std::string name;
  int64_t launches = 0;
  double us = 0.0, bytes = 0.0;
};
struct KRec {
  int cls;
  hipEvent_t e0, e1;
  double bytes;    // known bytes, or -1: shift update of chunk `chunk`
  int chunk;       // chunk index (shift updates)
};
struct KTimer {
  bool on = false;
  std::vector<KClass> cls;
  std::vector<KRec> recs;
  std::vector<hipEvent_t> pool;
  size_t used = 0;
};
static thread_local KTimer g_kt;
static bool kt_on() { return g_kt.on; }

static int kt_class(const std::string& name) {
  for (size_t i = 0; i < g_kt.cls.size(); ++i)
    if (g_kt.cls[i].name == name) return (int)i;
  g_kt.cls.push_back(KClass{name});
  return (int)g_kt.cls.size() - 1;
}
static int kt_event(hipEvent_t* e) {
  if (g_kt.used == g_kt.pool.size()) {
    hipEvent_t x;
    LYM_HIP(hipEventCreate(&x));
    g_kt.pool.push_back(x);
  }
  *e = g_kt.pool[g_kt.used++];
  return LYM_OK;
}
// time launch `f` on stream st as class `name` with `bytes` (or -1 + chunk)
template <class F>
static int kt_launch(hipStream_t st, const char* name, double bytes, int chunk, F&& f) {
  if (!g_kt.on) return f();
  KRec r{kt_class(name), nullptr, nullptr, bytes, chunk};
  int rc = kt_event(&r.e0);
  if (!rc) rc = kt_event(&r.e1);
  if (rc) return rc;
  LYM_HIP(hipEventRecord(r.e0, st));
  rc = f();
  LYM_HIP(hipEventRecord(r.e1, st));
  g_kt.recs.push_back(r);
  return rc;
}
// the solve has ended (stream synchronised): add durations and bytes.
// chunk_k0[c] / chunk_nb[c]: first iteration and length of chunk c;
// conv[s]: iteration shift s converged at (<0: never); shift_bytes(c) the
// bytes of chunk c's shift update.
template <class G>
static int kt_collect(G&& shift_bytes) {
  for (const KRec& r : g_kt.recs) {
    float ms = 0.f;
    LYM_HIP(hipEventElapsedTime(&ms, r.e0, r.e1));
    KClass& c = g_kt.cls[r.cls];
    const double b = r.bytes >= 0 ? r.bytes : shift_bytes(r.chunk);
    if (r.bytes >= 0 || b > 0) c.launches += 1;  // a shift update with no block due is not counted
    c.us += 1e3 * ms;
    c.bytes += b;
  }
  g_kt.recs.clear();
  g_kt.used = 0;
  return LYM_OK;
}

int kernel_timing(int enable) {
  g_kt.on = enable != 0;
  if (enable) {
    g_kt.cls.clear();
    g_kt.recs.clear();
    g_kt.used = 0;
  }
  return LYM_OK;
}
int kernel_timing_read(lym_kernel_time* out, int max, int* count) {
  const int n = (int)g_kt.cls.size();
  if (count) *count = n;
  for (int i = 0; i < n && i < max; ++i) {
    std::memset(&out[i], 0, sizeof(out[i]));
    std::snprintf(out[i].name, sizeof(out[i].name), "%s", g_kt.cls[i].name.c_str());
    out[i].launches = (int32_t)g_kt.cls[i].launches;
    out[i].total_us = g_kt.cls[i].us;
    out[i].bytes = g_kt.cls[i].bytes;
  }
  return LYM_OK;
}

// kernel-class suffix of a storage precision
static const char* pname(int p) { return p == P64 ? "f64" : p == P32 ? "f32" : "f16"; }

// algorithmic bytes of one Wilson launch over nrhs right-hand sides: the
// links once (D directions x rows x nc complex; one read serves every RHS of
// the launch) + `nvec` spinors per RHS (complex numbers of 2 * ws bytes, plus
// a 4-byte norm per site for 16-bit storage), per site
static double wilson_bytes(const OpDesc& d, int prec, int nvec, int nrhs) {
  const double wl = prec == P64 ? 8.0 : 4.0;
  const double rows = (!d.adj && d.link_su) ? d.nc - 1 : d.nc;
  const double links = (double)d.ndim * rows * d.nc * 2.0 * wl * (d.adj ? 2.0 : 1.0);
  return (double)d.lat.V * (links + (double)nrhs * nvec * vec_bytes_site(d, prec));
}

// Fast multishift CG over a batch of nrhs right-hand sides (one shift list).
// Blocks of m iterations (CgmTuning::mblk_64 in the FP64 phase, mblk_lo after):
//   [ p = r + beta p ; t = Q p (|t|^2,|p|^2) ; r' = r - alpha (Q t + s0 p) (|r'|^2) ] x m
//   then one deferred update of every shift's x_s, p_s (FP64) from the ring.
// Krylov vectors (r, p, t) are FP64 until every RHS has its base residual
// below switch_tol |b| (or has finished), then complex64; with lowprec = P16
// they move on to 16-bit storage (per-site norms, FP32 arithmetic) once every
// base residual is below switch2_tol |b|.  A rounding error of a phase enters
// the shifted solutions at the size of the residual it is made at, so the
// thresholds keep each phase's contribution to the true residual below the
// target.  xs[r * nshift + s] are the caller's FP64 solutions.
static int batched_tile(const OpDesc& d, const Parts& P, int pl, int nrhs, Field in, Field out, hipStream_t st,
                        OpDesc& db);

static int cgm_fast(const OpDesc& d, const Parts& P, const void* const* bs, void* const* xs, int nrhs, int nshift,
                    const double* shifts, const double* tol_rel, int max_iter, const CgmTuning& tun, int lowprec,
                    hipStream_t st, std::vector<SolverState>& out_states) {
  const double switch_tol = tun.switch32, switch2_tol = tun.switch16;
  const int mlo = tun.mblk_lo, m64 = tun.mblk_64;
  graph_new_solve();
  SolverState h;
  init_cgm_state(h, nshift, shifts, tol_rel, max_iter);
  h.phase = PH_KP;
  h.mblk = m64;
  h.switch_rr = switch_tol * switch_tol;
  h.switch2_rr = lowprec == P16 ? switch2_tol * switch2_tol : 0.0;
  std::vector<SolverState> hs(nrhs, h);
  std::vector<Field> hx((size_t)nrhs * nshift), hp((size_t)nrhs * nshift), hb(nrhs);
  const int64_t n = field_elems(d);
  const int64_t V = d.lat.V;
  for (int r = 0; r < nrhs; ++r) {
    hb[r].data = const_cast<void*>(bs[r]);
    for (int k = 0; k < nshift; ++k) {
      hx[(size_t)r * nshift + k].data = xs[(size_t)r * nshift + k];
      hp[(size_t)r * nshift + k].data = (char*)P.ps64.data + ((int64_t)r * nshift + k) * n * 16;
    }
  }
  BatchCtl ctl{0, 0, nrhs, 0};
  LYM_HIP(hipMemcpyAsync(P.st, hs.data(), sizeof(SolverState) * nrhs, hipMemcpyHostToDevice, st));
  LYM_HIP(hipMemcpyAsync(P.ctl, &ctl, sizeof(ctl), hipMemcpyHostToDevice, st));
  LYM_HIP(hipMemcpyAsync(P.xs_dev, hx.data(), sizeof(Field) * hx.size(), hipMemcpyHostToDevice, st));
  LYM_HIP(hipMemcpyAsync(P.ps_dev, hp.data(), sizeof(Field) * hp.size(), hipMemcpyHostToDevice, st));
  LYM_HIP(hipMemcpyAsync(P.bs_dev, hb.data(), sizeof(Field) * nrhs, hipMemcpyHostToDevice, st));
  const int gr = red_grid_flat(n);
  const int gf = flat_grid(n);
  const int gs = flat_grid(V);
  const int nslot64 = P.nslot64, nslot32 = P.nslot_lo, nslot16 = P.nslot_lo;
  if (m64 + 1 > nslot64 || mlo + 1 > nslot32) {
    set_error("fast CG-M: block length changed since the workspace was sized");
    return LYM_ERR_ARG;
  }
  const int64_t ring_stride64 = (int64_t)nslot64 * n, ring_stride32 = (int64_t)nslot32 * n;
  const int64_t ring_stride16 = (int64_t)nslot16 * n;
  hipLaunchKernelGGL(k_fast_init, dim3(gr, nrhs), dim3(kBlock), 0, st, n, P.bs_dev, (double2*)P.ring64.data,
                     ring_stride64, (double2*)P.plive64.data, P.xs_dev, P.ps_dev, nshift, P.partials);
  LYM_HIP(hipGetLastError());
  int rc = launch_fin(FK_INIT, P.st, P.partials, gr, 1, nullptr, st, nrhs, P.ctl);
  if (rc) return rc;
  if (lowprec == P16 && d.ncomp != 4 && d.ncomp != 6 && d.ncomp != 8 && d.ncomp != 12) {
    set_error("16-bit multishift storage: fundamental representation only");
    return LYM_ERR_UNSUPPORTED;
  }
  // site kernels of the 16-bit phase, instantiated per component count
  auto by_ncomp = [&](auto f) -> int {
    switch (d.ncomp) {
      case 4: f(IC<4>{}); break;
      case 6: f(IC<6>{}); break;
      case 8: f(IC<8>{}); break;
      case 12: f(IC<12>{}); break;
      default:
        set_error("16-bit multishift storage: unsupported number of spinor components");
        return LYM_ERR_UNSUPPORTED;
    }
    LYM_HIP(hipGetLastError());
    return LYM_OK;
  };

  // Wilson tiles of the batch, per precision (tuned on the first batched
  // solve of this shape: t = Q p timed on the phase's own fields, which hold
  // nothing yet except the FP64 direction)
  OpDesc d64 = d, d32 = d, d16 = d;
  if ((rc = batched_tile(d, P, P64, nrhs, P.plive64, P.tb64, st, d64))) return rc;
  if ((rc = batched_tile(d, P, P32, nrhs, P.plive32, P.tb32, st, d32))) return rc;
  if (lowprec == P16 && (rc = batched_tile(d, P, P16, nrhs, P.plive16, P.tb16, st, d16))) return rc;
  int prec = P64;
  // for the kernel timing (no graphs then, so the chunk body runs at every
  // issue): chunks issued so far (precision and length), and every shift
  // update launch in issue order (chunk c's end: (0, c); the flush before
  // chunk c at a precision switch: (1, c))
  std::vector<std::pair<int, int>> chunk_log, k3_log;
  // deferred shift update of the current block (a no-op for RHS not in PH_K3)
  auto shift_update = [&](hipStream_t st, int kind) -> int {
    const Field ring = prec == P64 ? P.ring64 : prec == P32 ? P.ring32 : P.ring16;
    const int nslot = prec == P64 ? nslot64 : prec == P32 ? nslot32 : nslot16;
    const int64_t ring_stride = prec == P64 ? ring_stride64 : prec == P32 ? ring_stride32 : ring_stride16;
    const int ei = (int)k3_log.size();
    k3_log.emplace_back(kind, (int)chunk_log.size() - (kind == 0 ? 1 : 0));
    const dim3 grid(gf, nrhs, k3_groups(nshift));
    const size_t lds = k3_lds_bytes(nslot, nshift);
    int e = kt_launch(st, (std::string("shift_update_") + pname(prec)).c_str(), -1.0, ei, [&]() {
      if (prec == P64)
        hipLaunchKernelGGL(k_fast_k3<P64>, grid, dim3(kBlock), lds, st, P.st, d.ncomp, V, ring, ring_stride, nslot,
                           P.xs_dev, P.ps_dev, nshift);
      else if (prec == P32)
        hipLaunchKernelGGL(k_fast_k3<P32>, grid, dim3(kBlock), lds, st, P.st, d.ncomp, V, ring, ring_stride, nslot,
                           P.xs_dev, P.ps_dev, nshift);
      else
        hipLaunchKernelGGL(k_fast_k3<P16>, grid, dim3(kBlock), lds, st, P.st, d.ncomp, V, ring, ring_stride, nslot,
                           P.xs_dev, P.ps_dev, nshift);
      LYM_HIP(hipGetLastError());
      return (int)LYM_OK;
    });
    if (e) return e;
    return launch_fin(FK_FAST_K3ADV, P.st, nullptr, 0, 0, nullptr, st, nrhs, P.ctl);
  };
  auto chunk = [&](hipStream_t st) -> int {
    const Field ring = prec == P64 ? P.ring64 : prec == P32 ? P.ring32 : P.ring16;
    const Field pl = prec == P64 ? P.plive64 : prec == P32 ? P.plive32 : P.plive16;
    const Field t = prec == P64 ? P.tb64 : prec == P32 ? P.tb32 : P.tb16;
    const OpDesc& dp = prec == P64 ? d64 : prec == P32 ? d32 : d16;
    const int gw = wilson_blocks(dp, prec, nrhs);  // Wilson partials per RHS at this precision
    const int cl = prec == P64 ? tun.chunk_64 : tun.chunk_lo;
    const int nslot = prec == P64 ? nslot64 : prec == P32 ? nslot32 : nslot16;
    const int64_t ring_stride = prec == P64 ? ring_stride64 : prec == P32 ? ring_stride32 : ring_stride16;
    chunk_log.emplace_back(prec, cl);
    const double sites = (double)V * nrhs;
    const std::string sfx = pname(prec);
    for (int j = 0; j < cl; ++j) {
      int e = kt_launch(st, ("direction_update_" + sfx).c_str(), 3.0 * vec_bytes_site(d, prec) * sites, -1, [&]() {
        if (prec == P64)
          hipLaunchKernelGGL(k_fast_kp<P64>, dim3(gf, nrhs), dim3(kBlock), 0, st, P.st, n, (const double2*)ring.data,
                             ring_stride, nslot, (double2*)pl.data);
        else if (prec == P32)
          hipLaunchKernelGGL(k_fast_kp<P32>, dim3(gf, nrhs), dim3(kBlock), 0, st, P.st, n, (const float2*)ring.data,
                             ring_stride, nslot, (float2*)pl.data);
        else {
          int e2 = by_ncomp([&](auto nc) {
            hipLaunchKernelGGL(k_fast_kp16<decltype(nc)::value>, dim3(gs, nrhs), dim3(kBlock), 0, st, P.st, V, ring,
                               nslot, pl);
          });
          if (e2) return e2;
        }
        LYM_HIP(hipGetLastError());
        return (int)LYM_OK;
      });
      if (e) return e;
      EpiArgs a1;
      a1.out = t;
      a1.st = P.st;
      a1.rhs_stride = n;
      e = kt_launch(st, ("wilson_cgm_k1_" + sfx).c_str(), wilson_bytes(d, prec, 2, nrhs), -1,
                    [&]() { return wilson_launch(EPI_CGM_K1, dp, prec, true, pl, a1, redctl(P), st, nrhs); });
      if (!e)
        e = kt_launch(st, "finalize", 0.0, -1,
                      [&]() { return launch_fin(FK_CGM_A, P.st, P.partials, gw, 2, nullptr, st, nrhs, P.ctl); });
      if (e) return e;
      EpiArgs a2;
      a2.aux0 = pl;
      a2.st = P.st;
      a2.ring = ring;
      a2.slot_elems = n;
      a2.nslot = nslot;
      a2.rhs_stride = n;
      a2.ring_rhs_stride = ring_stride;
      e = kt_launch(st, ("wilson_cgm_k2f_" + sfx).c_str(), wilson_bytes(d, prec, 4, nrhs), -1,
                    [&]() { return wilson_launch(EPI_CGM_K2F, dp, prec, true, t, a2, redctl(P), st, nrhs); });
      if (!e)
        e = kt_launch(st, "finalize", 0.0, -1,
                      [&]() { return launch_fin(FK_CGM_BF, P.st, P.partials, gw, 1, nullptr, st, nrhs, P.ctl); });
      if (e) return e;
    }
    // the block's shift update runs at the chunk end that completes it (or
    // where the solve stops); elsewhere the launch finds no RHS in PH_K3
    return shift_update(st, 0);
  };
  // poll the batch control block: (ndone, nswitch, nrhs, -).  Each in-flight
  // chunk remembers the precision it ran at: a tally read from a chunk issued
  // before the last switch is stale and never triggers another one.
  int rce = ensure_sync();
  if (rce) return rce;
  const int64_t max_chunks = (int64_t)max_iter / 8 + 8;
  int64_t issued = 0, polled = 0;
  int slot_prec[2] = {P64, P64};
  KeyBuf kb;
  kb.add('F').add(nrhs).add(nshift).add(P.st).add(P.xs_dev).add(P.ring64.data);
  const std::string key64 = op_key(d64) + kb.s + 'd', key32 = op_key(d32) + kb.s + 's';
  const std::string key16 = op_key(d16) + kb.s + 'h';
  auto issue = [&](int slot) -> int {
    const int e = launch_chunk(prec == P64 ? key64 : prec == P32 ? key32 : key16, st, chunk);
    if (e) return e;
    LYM_HIP(hipMemcpyAsync(&g_sync.pinned[4 * slot], P.ctl, 4 * sizeof(int), hipMemcpyDeviceToHost, st));
    LYM_HIP(hipEventRecord(g_sync.ev[slot], st));
    slot_prec[slot] = prec;
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
    if (s4[0] >= nrhs) break;
    if (issued > max_chunks) {
      set_error("fast CG-M did not reach its stop state (internal error)");
      return LYM_ERR_HIP;
    }
    if (s4[1] >= nrhs && slot_prec[slot] == prec && prec != lowprec) {
      // the block in progress is flushed at the old precision; the new ring
      // shares the old ring's memory, so the current residual is staged in
      // the old phase's t = Q p (recomputed before it is read again)
      hipLaunchKernelGGL(k_fast_arm_flush, dim3(1, nrhs), dim3(64), 0, st, P.st);
      LYM_HIP(hipGetLastError());
      if ((rc = shift_update(st, 1))) return rc;
      if (prec == P64) {
        hipLaunchKernelGGL(k_fast_stage<double2>, dim3(gf, nrhs), dim3(kBlock), 0, st, P.st, n,
                           (const double2*)P.ring64.data, ring_stride64, nslot64, (double2*)P.tb64.data);
        hipLaunchKernelGGL(k_fast_switch<P32>, dim3(gf, nrhs), dim3(kBlock), 0, st, P.st, P.ctl, n,
                           (const double2*)P.tb64.data, (float2*)P.ring32.data, ring_stride32, nslot32,
                           (const double2*)P.plive64.data, (float2*)P.plive32.data, mlo);
        LYM_HIP(hipGetLastError());
        prec = P32;
      } else {
        hipLaunchKernelGGL(k_fast_stage<float2>, dim3(gf, nrhs), dim3(kBlock), 0, st, P.st, n,
                           (const float2*)P.ring32.data, ring_stride32, nslot32, (float2*)P.tb32.data);
        LYM_HIP(hipGetLastError());
        int e = by_ncomp([&](auto nc) {
          hipLaunchKernelGGL(k_fast_switch16<decltype(nc)::value>, dim3(gs, nrhs), dim3(kBlock), 0, st, P.st, P.ctl,
                             V, (const float2*)P.tb32.data, P.ring16, nslot16, (const float2*)P.plive32.data,
                             P.plive16, mlo);
        });
        if (e) return e;
        prec = P16;
      }
    }
    if ((rc = issue(slot))) return rc;
  }
  LYM_HIP(hipStreamSynchronize(st));
  out_states.resize(nrhs);
  LYM_HIP(hipMemcpyAsync(out_states.data(), P.st, sizeof(SolverState) * nrhs, hipMemcpyDeviceToHost, st));
  LYM_HIP(hipStreamSynchronize(st));
  if (g_kt.on) {
    // bytes of every shift update launch, replaying the block boundaries of
    // each RHS: chunk c covers iterations k0[c] .. k0[c + 1]; a phase's blocks
    // end every mblk iterations from its first one, where the RHS stops, and
    // at the flush before a precision switch.  A block of nb steps reads its
    // residuals once per shift group and 4 FP64 vectors of every shift
    // active at its start.
    std::vector<int> k0(chunk_log.size() + 1, 0);
    for (size_t c = 0; c < chunk_log.size(); ++c) k0[c + 1] = k0[c] + chunk_log[c].second;
    std::vector<double> ev_bytes(k3_log.size(), 0.0);
    for (int r = 0; r < nrhs; ++r) {
      const SolverState& h = out_states[r];
      int f = 0, sp = 0, mb = m64;
      for (size_t e = 0; e < k3_log.size(); ++e) {
        const int kind = k3_log[e].first, c = k3_log[e].second;
        int nb = 0, pr = P64;
        if (kind == 0) {
          pr = chunk_log[c].first;
          const int end = k0[c + 1];
          if (f < h.iter && h.iter <= end) nb = h.iter - f;
          else if (f < h.iter && (end - sp) % mb == 0) nb = end - f;
        } else {
          pr = c > 0 ? chunk_log[c - 1].first : P64;
          if (h.iter > k0[c]) nb = k0[c] - f;
          sp = k0[c];
          mb = mlo;
        }
        if (nb <= 0) continue;
        int act = 0;
        for (int k = 0; k < nshift; ++k)
          if (h.conv_iter[k] < 0 || h.conv_iter[k] > f) ++act;
        ev_bytes[e] += (double)V * ((double)k3_groups(act) * nb * vec_bytes_site(d, pr) +
                                    act * 4.0 * vec_bytes_site(d, P64));
        f += nb;
      }
    }
    rc = kt_collect([&](int e) { return ev_bytes[e]; });
    if (rc) return rc;
  }
  return LYM_OK;
}

// One-shift CG on Q^2 + sigma over a batch of nrhs sources with complex64
// Krylov vectors and FP64 reliable updates (Sleijpen & van der Vorst; the
// scheme of Clark et al., arXiv:0911.3191): the low-precision recurrence runs
// until its residual has dropped by 100x since the last update (or passes the
// stopping test), then the solution update since the last one is added to
// x_hi and r = b - (Q^2 + sigma) x_hi is formed in FP64 and replace the iterated residual, p is kept.  Every RHS advances
// in the same launches; one that has stopped skips them.  xs[r] (FP64) get
// the solutions.  Iteration (each step one launch, RHS batched):
//   p_k = r + beta p_{k-1} (ring slot k) ; t = Q p_k (|t|^2, |p|^2) ; alpha ;
//   r -= alpha (Q t + sigma p_k) (|r|^2) ; beta / update test ;
//   after every second iteration, for the RHS that asked for it:
//   [x_hi += sum alpha_j p_j ; t = Q x_hi ; r = b - (Q t + sigma x_hi) (|r|^2, <r,p>) ; test]
// The solution update sum_j alpha_j p_j over a block of up to kRelBlock steps
// is deferred and formed from the ring of stored directions (k_rel_kp,
// k_ru_fold), which saves the per-step read and write of a running solution.
// (Rebuilding the block's directions from its residuals instead, as the
// multishift solver does, was measured here: with 16-bit p the rebuilt and
// the stored directions drift apart by the rounding of each store, every
// reliable update pays for the gap, and the 28-source batch at 24^3x48 took
// 1933 iterations instead of 729.)
constexpr int kRelChunk = 8;  // iterations per host chunk

// Tile of the batched Wilson kernels of the reliable-update CG.  The shape
// that is fastest for one RHS is not the fastest for a batch (measured at
// 24^3x48 with 28 RHS: a 192-site tile with two hops in flight beats the
// one-RHS choice by 5%), so the first batched solve on a lattice times its
// step-1 kernel (t = Q p with |t|^2, |p|^2, on the solve's own fields, right
// after the initialisation) over every candidate tile and hop grouping, and
// the choice holds for the process, per lattice, Nc, storage precision and
// batch size.  LYM_NO_BATCH_TUNE=1 keeps the one-RHS tile.
struct BatchTuneKey {
  int nc, pl, nrhs, L[4];
  bool operator<(const BatchTuneKey& o) const {
    return std::tie(nc, pl, nrhs, L[0], L[1], L[2], L[3]) < std::tie(o.nc, o.pl, o.nrhs, o.L[0], o.L[1], o.L[2], o.L[3]);
  }
};
static std::mutex g_btune_mu;
static std::map<BatchTuneKey, TileMap> g_btuned;

static bool batch_tune_enabled() {
  const char* v = std::getenv("LYM_NO_BATCH_TUNE");
  return !(v && v[0] == '1');
}

// db = d with the batched tile of storage precision pl in d.tile[] (unchanged
// when the lattice is not tiled, the multi-RHS kernel runs the batch, tuning
// is off, or no RHS is active).  The timed launches read `in` and leave
// `out` (batched fields of precision pl) and the partials overwritten; the
// states are only read.
static int batched_tile(const OpDesc& d, const Parts& P, int pl, int nrhs, Field in, Field out, hipStream_t st,
                        OpDesc& db) {
  db = d;
  if (nrhs < 2 || !wilson_tiled(d, pl) || wilson_mr(d, pl, nrhs) || d.adj || d.ndim != 4 || !batch_tune_enabled())
    return LYM_OK;
  const int slot = pl == P64 ? 0 : 1;
  const BatchTuneKey key{d.nc, pl, nrhs, {d.lat.L[0], d.lat.L[1], d.lat.L[2], d.lat.L[3]}};
  {
    std::lock_guard<std::mutex> lock(g_btune_mu);
    auto it = g_btuned.find(key);
    if (it != g_btuned.end()) {
      db.tile[slot] = it->second;
      return LYM_OK;
    }
  }
  BatchCtl c{};
  LYM_HIP(hipMemcpyAsync(&c, P.ctl, sizeof(c), hipMemcpyDeviceToHost, st));
  LYM_HIP(hipStreamSynchronize(st));
  if (c.ndone >= nrhs) return LYM_OK;  // every source is zero: nothing to time
  // LDS bytes per site and component: complex128, or complex64 (16-bit
  // storage is decoded to complex64 in the tile)
  const int64_t es = (int64_t)d.ncomp * (pl == P64 ? 16 : 8);
  std::vector<TileMap> cands;
  for (const TileMap& t : tile_candidates(d, es, tile_max_threads(0, pl), kTileMaxLds)) {
    for (int hif : {1, 2}) {
      if (hif > 1 && t.nsite > tile_hif_threads(pl)) continue;
      for (int walk : {0, 1}) {
        if (walk && (t.n1 == 1 || t.n2 == 1)) continue;  // the same order
        TileMap x = t;
        x.hif = hif;
        x.walk = walk;
        cands.push_back(x);
      }
    }
  }
  if (cands.empty()) return LYM_OK;
  hipEvent_t e0, e1;
  LYM_HIP(hipEventCreate(&e0));
  LYM_HIP(hipEventCreate(&e1));
  EpiArgs a1;
  a1.out = out;
  a1.st = P.st;
  a1.rhs_stride = field_elems(d);
  int rc = LYM_OK;
  double best_us = 1e30;
  TileMap best = d.tile[slot];
  const int reps = 3;
  for (size_t ci = 0; ci < cands.size() && !rc; ++ci) {
    OpDesc dc = d;
    dc.tile[slot] = cands[ci];
    for (int i = 0; i < 2 && !rc; ++i) rc = wilson_launch(EPI_CGM_K1, dc, pl, true, in, a1, redctl(P), st, nrhs);
    double us = 1e30;
    for (int trial = 0; trial < 2 && !rc; ++trial) {
      if (hipEventRecord(e0, st) != hipSuccess) rc = hip_fail(hipErrorUnknown, "hipEventRecord");
      for (int i = 0; i < reps && !rc; ++i) rc = wilson_launch(EPI_CGM_K1, dc, pl, true, in, a1, redctl(P), st, nrhs);
      if (!rc && hipEventRecord(e1, st) != hipSuccess) rc = hip_fail(hipErrorUnknown, "hipEventRecord");
      float ms = 0.f;
      if (!rc && hipEventSynchronize(e1) != hipSuccess) rc = hip_fail(hipErrorUnknown, "hipEventSynchronize");
      if (!rc && hipEventElapsedTime(&ms, e0, e1) != hipSuccess) rc = hip_fail(hipErrorUnknown, "hipEventElapsedTime");
      if (!rc) us = std::min(us, 1e3 * ms / reps);
    }
    if (!rc && us < best_us) {
      best_us = us;
      best = cands[ci];
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (rc) return rc;
  {
    std::lock_guard<std::mutex> lock(g_btune_mu);
    g_btuned[key] = best;
  }
  db.tile[slot] = best;
  return LYM_OK;
}

static int cg_reliable(const OpDesc& d, const Parts& P, int pl, const void* const* bs, void* const* xs, int nrhs,
                       double sigma, double tol, int max_iter, hipStream_t st, std::vector<SolverState>& out_states) {
  graph_new_solve();
  // site kernels instantiated per storage precision of r, p and component count
  auto dispatch = [&](auto f) -> int {
    auto by_nc = [&](auto pv) -> int {
      switch (d.ncomp) {
        case 4: f(pv, IC<4>{}); break;
        case 6: f(pv, IC<6>{}); break;
        case 8: f(pv, IC<8>{}); break;
        case 12: f(pv, IC<12>{}); break;
        case 16: f(pv, IC<16>{}); break;
        case 18: f(pv, IC<18>{}); break;
        case 36: f(pv, IC<36>{}); break;
        default:
          set_error("mixed CG: unsupported number of spinor components");
          return LYM_ERR_UNSUPPORTED;
      }
      LYM_HIP(hipGetLastError());
      return LYM_OK;
    };
    return pl == P16 ? by_nc(IC<P16>{}) : by_nc(IC<P32>{});
  };
  const int64_t V = d.lat.V;
  const int gs = flat_grid(V);
  SolverState h;
  const double tol_iter = 0.5 * tol;  // iterate below the target; the true residual decides
  init_cgm_state(h, 1, &sigma, &tol_iter, max_iter);
  h.phase = PH_KP;
  h.precision_phase = pl == P16 ? 2 : 1;
  // a reliable update after every 100x drop of the iterated residual (delta =
  // 0.01).  Measured against 10x (delta = 0.1) on the 28-source batches: the
  // same iteration counts (729 / 255 against 732 / 257 at 24^3x48 / 16^3x32)
  // and 5-8% less time; the solve still ends on the FP64 true residual.
  h.delta2 = 1e-4;
  h.tol2_true = tol * tol;
  h.mblk = kRelBlock;
  std::vector<SolverState> hs(nrhs, h);
  std::vector<Field> hb(nrhs), hx(nrhs);
  for (int r = 0; r < nrhs; ++r) {
    hb[r].data = const_cast<void*>(bs[r]);
    hx[r].data = xs[r];
  }
  BatchCtl ctl{0, 0, nrhs, 0};
  LYM_HIP(hipMemcpyAsync(P.st, hs.data(), sizeof(SolverState) * nrhs, hipMemcpyHostToDevice, st));
  LYM_HIP(hipMemcpyAsync(P.ctl, &ctl, sizeof(ctl), hipMemcpyHostToDevice, st));
  LYM_HIP(hipMemcpyAsync(P.bs_dev, hb.data(), sizeof(Field) * nrhs, hipMemcpyHostToDevice, st));
  LYM_HIP(hipMemcpyAsync(P.xs_dev, hx.data(), sizeof(Field) * nrhs, hipMemcpyHostToDevice, st));
  const int64_t n = field_elems(d);
  const int gr = red_grid_flat(n);
  int rc = dispatch([&](auto pv, auto nc) {
    hipLaunchKernelGGL((k_rel_init<decltype(pv)::value, decltype(nc)::value>), dim3(gr, nrhs), dim3(kBlock), 0, st,
                       V, P.bs_dev, P.rl_r, P.rl_ring, P.rl_nslot, (double2*)P.rl_x64.data, P.partials);
  });
  if (rc) return rc;
  rc = launch_fin(FK_INIT, P.st, P.partials, gr, 1, nullptr, st, nrhs, P.ctl);
  if (rc) return rc;
  OpDesc db;
  // the tile is timed on r (set to b), a plain batched field
  if ((rc = batched_tile(d, P, pl, nrhs, P.rl_r, P.rl_t, st, db))) return rc;
  const int gwl = wilson_blocks(db, pl, nrhs), gw64 = wilson_blocks(db, P64, nrhs);
  // the update accumulation is mostly gated off: a small grid-stride grid
  const int gsr = std::min(gs, 1024);
  static_assert(kRelChunk % 2 == 0, "the last iteration of a chunk must carry the update kernels");
  // kernel timing: class suffix and the direction update's bytes (r and
  // p_{k-1} read, p_k written at the storage precision; the block fold that
  // one launch in kRelBlock carries per RHS is not counted)
  const std::string sfx = pname(pl);
  const double kp_bytes = (double)V * nrhs * 3.0 * vec_bytes_site(d, pl);
  const int64_t ring_stride = n * P.rl_nslot;
  auto chunk = [&](hipStream_t st) -> int {
    for (int j = 0; j < kRelChunk; ++j) {
      // p_k = r + beta p_{k-1} (and the fold of a full block)
      int e = kt_launch(st, ("cg_direction_update_" + sfx).c_str(), kp_bytes, -1, [&]() {
        return dispatch([&](auto pv, auto nc) {
          hipLaunchKernelGGL((k_rel_kp<decltype(pv)::value, decltype(nc)::value>), dim3(gs, nrhs), dim3(kBlock), 0,
                             st, P.st, V, P.rl_r, P.rl_ring, P.rl_nslot, (double2*)P.rl_x64.data);
        });
      });
      if (e) return e;
      EpiArgs a1;
      a1.out = P.rl_t;
      a1.slot_elems = n;
      a1.nslot = P.rl_nslot;
      a1.ring_rhs_stride = ring_stride;
      a1.st = P.st;
      a1.rhs_stride = n;
      e = kt_launch(st, ("wilson_cg_k1_" + sfx).c_str(), wilson_bytes(d, pl, 2, nrhs), -1,
                    [&]() { return wilson_launch(EPI_CG_K1R, db, pl, true, P.rl_ring, a1, redctl(P), st, nrhs); });
      if (!e)
        e = kt_launch(st, "finalize", 0.0, -1,
                      [&]() { return launch_fin(FK_CG_A, P.st, P.partials, gwl, 2, nullptr, st, nrhs, P.ctl); });
      if (e) return e;
      EpiArgs a2;
      a2.aux1 = P.rl_r;
      a2.ring = P.rl_ring;
      a2.slot_elems = n;
      a2.nslot = P.rl_nslot;
      a2.ring_rhs_stride = ring_stride;
      a2.st = P.st;
      a2.rhs_stride = n;
      e = kt_launch(st, ("wilson_cg_k2x_" + sfx).c_str(), wilson_bytes(d, pl, sigma != 0.0 ? 4 : 3, nrhs), -1,
                    [&]() { return wilson_launch(EPI_CG_K2X, db, pl, true, P.rl_t, a2, redctl(P), st, nrhs); });
      if (!e)
        e = kt_launch(st, "finalize", 0.0, -1,
                      [&]() { return launch_fin(FK_CG_B, P.st, P.partials, gwl, 1, nullptr, st, nrhs, P.ctl); });
      if (e) return e;
      // reliable update after every second iteration (an RHS that asks for
      // one after an even iteration skips the next): every kernel below
      // returns at once for RHS not in PH_RU, so their cost is the launch
      if (j % 2 == 0) continue;
      // gated: only the RHS in PH_RU do work, so no bytes are attributed
      e = kt_launch(st, "reliable_update", 0.0, -1, [&]() {
        int e2 = dispatch([&](auto pv, auto nc) {
          hipLaunchKernelGGL((k_ru_fold<decltype(pv)::value, decltype(nc)::value>), dim3(gsr, nrhs), dim3(kBlock),
                             0, st, P.st, V, P.rl_ring, P.rl_nslot, (double2*)P.rl_x64.data);
        });
        if (e2) return e2;
        EpiArgs a3;
        a3.out = P.rl_t64;
        a3.st = P.st;
        a3.rhs_stride = n;
        e2 = wilson_launch(EPI_RU1, db, P64, true, P.rl_x64, a3, RedCtl{nullptr, nullptr}, st, nrhs);
        if (e2) return e2;
        EpiArgs a4;
        a4.out = P.rl_r;
        a4.ring = P.rl_ring;
        a4.slot_elems = n;
        a4.nslot = P.rl_nslot;
        a4.ring_rhs_stride = ring_stride;
        a4.aux0 = P.rl_x64;
        a4.bs = P.bs_dev;
        a4.st = P.st;
        a4.rhs_stride = n;
        e2 = wilson_launch(pl == P16 ? EPI_RU2H : EPI_RU2, db, P64, true, P.rl_t64, a4, redctl(P), st, nrhs);
        if (!e2) e2 = launch_fin(FK_CG_RU, P.st, P.partials, gw64, 3, nullptr, st, nrhs, P.ctl);
        return e2;
      });
      if (e) return e;
    }
    return LYM_OK;
  };
  rc = ensure_sync();
  if (rc) return rc;
  // an update takes an iteration slot without advancing iter: at most one per iteration
  const int64_t max_chunks = 2 * ((int64_t)max_iter + kRelChunk) / kRelChunk + 8;
  int64_t issued = 0, polled = 0;
  KeyBuf kb;
  kb.add('R').add(pl).add(nrhs).add(P.st).add(P.xs_dev).add(P.rl_ring.data);
  const std::string key = op_key(db) + kb.s;
  auto issue = [&](int slot) -> int {
    const int e = launch_chunk(key, st, chunk);
    if (e) return e;
    LYM_HIP(hipMemcpyAsync(&g_sync.pinned[4 * slot], P.ctl, 4 * sizeof(int), hipMemcpyDeviceToHost, st));
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
    if (g_sync.pinned[4 * slot] >= nrhs) break;
    if (issued > max_chunks) {
      set_error("mixed CG did not reach its stop state (internal error)");
      return LYM_ERR_HIP;
    }
    if ((rc = issue(slot))) return rc;
  }
  rc = dispatch([&](auto, auto nc) {
    hipLaunchKernelGGL((k_rel_out<decltype(nc)::value>), dim3(gs, nrhs), dim3(kBlock), 0, st, V,
                       (const double2*)P.rl_x64.data, P.xs_dev);
  });
  if (rc) return rc;
  out_states.resize(nrhs);
  LYM_HIP(hipMemcpyAsync(out_states.data(), P.st, sizeof(SolverState) * nrhs, hipMemcpyDeviceToHost, st));
  LYM_HIP(hipStreamSynchronize(st));
  if (g_kt.on && (rc = kt_collect([](int) { return 0.0; }))) return rc;
  return LYM_OK;
}

static void fill_stats(lym_solve_stats* s, const SolverState& h, int nshift) {
  if (!s) return;
  for (int k = 0; k < 64; ++k) {
    s->initial_rel_resid[k] = 0.0;
    s->refine_cycles[k] = 0;
  }
  s->refine_iterations = 0;
  s->precision_phase = h.precision_phase;
  s->status = h.status;
  s->iterations = h.iter;
  s->nshift = nshift;
  s->final_rr = h.rr;
  for (int k = 0; k < nshift && k < 64; ++k) s->conv_iter[k] = h.conv_iter[k];
}

// FP64 true residual of every (rhs, shift) and defect correction where one
// misses the target.  xs[r * nshift + k]; per (rhs, shift) entry r * nshift + k:
// rel0 = true relative residual before correction, rel = after, cycles =
// correction solves; iters[r] = iterations of RHS r's correction solves.
struct RefineOut {
  std::vector<double> rel0, rel;
  std::vector<int> cycles;
  std::vector<int> iters;  // per RHS
};
static int verify_and_refine(const OpDesc& d, const Parts& P, int prec, int nrhs, int nshift, const double* shifts,
                             const void* const* bs, void* const* xs, double tol, int max_iter,
                             const std::vector<SolverState>& hs, hipStream_t st, RefineOut& ro) {
  const int64_t n = field_elems(d);
  const int max_refine = 12;
  ro.rel.assign((size_t)nrhs * nshift, 0.0);
  ro.rel0.assign((size_t)nrhs * nshift, 0.0);
  ro.cycles.assign((size_t)nrhs * nshift, 0);
  ro.iters.assign(nrhs, 0);
  for (int r = 0; r < nrhs; ++r) {
    // a right-hand side whose solve broke down is reported, not refined
    const bool breakdown = hs[r].status == ST_BREAKDOWN;
    Field bf;
    bf.data = const_cast<void*>(bs[r]);
    for (int k = 0; k < nshift; ++k) {
      void* xk_ptr = xs[(size_t)r * nshift + k];
      Field xk;
      xk.data = xk_ptr;
      double r2 = 0;
      int rc = true_residual(d, P, bf, xk, shifts[k], P.r64, &r2, st);
      if (rc) return rc;
      ro.rel0[(size_t)r * nshift + k] = std::sqrt(r2);
      int cycles = 0;
      while (std::sqrt(r2) > tol && cycles < max_refine && !breakdown) {
        const double relnow = std::sqrt(r2);
        double tol_in = (prec == P64) ? 0.5 * tol / relnow : std::max(0.5 * tol / relnow, 1e-5);
        if (tol_in > 0.1) tol_in = 0.1;
        SolverState hi;
        double sh = shifts[k];
        if (prec == P64) {
          // FP64 correction e (u64) from the residual copied to v64
          rc = blas_convert(P64, P.r64.data, P64, P.v64.data, n, st);
          if (rc) return rc;
          Field xe[1] = {P.u64};
          rc = cgm_core(d, P64, P, P.v64, xe, P.r64, P.t64, 1, &sh, &tol_in, max_iter, st, &hi);
          if (rc) return rc;
          rc = blas_axpby(P64, n, 1.0, 0.0, P.u64.data, 1.0, 0.0, xk_ptr, st);
          if (rc) return rc;
        } else {
          rc = blas_convert(P64, P.r64.data, P32, P.b_lo.data, n, st);
          if (rc) return rc;
          Field xe[1] = {P.x_lo[0]};
          rc = cgm_core(d, P32, P, P.b_lo, xe, P.r_lo, P.t_lo, 1, &sh, &tol_in, max_iter, st, &hi);
          if (rc) return rc;
          hipLaunchKernelGGL(k_add32to64, dim3(flat_grid(n)), dim3(kBlock), 0, st, n, (const float2*)P.x_lo[0].data,
                             (double2*)xk_ptr);
          LYM_HIP(hipGetLastError());
        }
        ro.iters[r] += hi.iter;
        ++cycles;
        rc = true_residual(d, P, bf, xk, shifts[k], P.r64, &r2, st);
        if (rc) return rc;
      }
      ro.cycles[(size_t)r * nshift + k] = cycles;
      ro.rel[(size_t)r * nshift + k] = std::sqrt(r2);
    }
  }
  return LYM_OK;
}

// Multishift CG for nrhs right-hand sides sharing one shift list.
//   prec = P64: the plain FP64 CG-M, one RHS after the other.
//   prec = P32: the fast mixed solver, all RHS in the same launches.
// stats: nrhs entries (or NULL).
int solve_cgm_batched(const OpDesc& d, int prec, int nrhs, int nshift, const double* shifts, const void* const* bs,
                      void* const* xs, double tol, int max_iter, void* ws, size_t ws_bytes, lym_solve_stats* stats,
                      hipStream_t st) {
  const auto t0 = std::chrono::steady_clock::now();
  if (!d.herm) {
    set_error("CG needs a Hermitian operator: build the Wilson operator with hermitian=True");
    return LYM_ERR_ARG;
  }
  if (nshift < 1 || nshift > kMaxShifts) {
    set_error("number of shifts must be in [1, 64]");
    return LYM_ERR_ARG;
  }
  if (nrhs < 1) {
    set_error("number of right-hand sides must be >= 1");
    return LYM_ERR_ARG;
  }
  for (int k = 0; k < nshift; ++k)
    if (!(shifts[k] >= 0.0)) {
      set_error("CG-M needs non-negative shifts");
      return LYM_ERR_ARG;
    }
  if (prec == P16 && d.adj) {
    set_error("half-precision storage is built for the fundamental representation");
    return LYM_ERR_UNSUPPORTED;
  }
  const int nbatch = prec == P64 ? 1 : nrhs;  // plain FP64 runs one RHS at a time
  const size_t need = solver_workspace_bytes(d, prec, nshift, nbatch);
  if (!ws || ws_bytes < need) {
    set_error("solver workspace too small: need " + std::to_string(need) + " bytes");
    return LYM_ERR_ARG;
  }
  Parts P;
  char* base = (char*)(((uintptr_t)ws + 255) & ~(uintptr_t)255);
  carve(d, prec, nshift, nbatch, base, &P);
  // iterate slightly below the target so the recomputed true residual meets it
  const CgmTuning tun = cgm_tuning();
  std::vector<double> tol_rel(nshift, tun.iter_factor * tol);
  std::vector<SolverState> hs;
  int rc = 0;
  if (prec == P64) {
    hs.resize(nrhs);
    for (int r = 0; r < nrhs; ++r) {
      Field bf;
      bf.data = const_cast<void*>(bs[r]);
      Field xf[kMaxShifts];
      for (int k = 0; k < nshift; ++k) xf[k].data = xs[(size_t)r * nshift + k];
      rc = cgm_core(d, P64, P, bf, xf, P.r64, P.t64, nshift, shifts, tol_rel.data(), max_iter, st, &hs[r]);
      if (rc) return rc;
    }
  } else {
    // mixed: FP64 Krylov vectors until every base residual is below
    // switch32 |b|, complex64 after (and 16-bit storage below switch16 |b|
    // for LYM_PREC_F16); shift solutions and directions accumulate in FP64
    // through deferred block updates.  FP64 true residuals are checked below.
    // One shift: complex64 or 16-bit CG with FP64 reliable updates instead.
    if (nshift == 1)
      rc = cg_reliable(d, P, prec == P16 ? P16 : P32, bs, xs, nrhs, shifts[0], tol, max_iter, st, hs);
    else
      rc = cgm_fast(d, P, bs, xs, nrhs, nshift, shifts, tol_rel.data(), max_iter, tun, prec == P16 ? P16 : P32, st,
                    hs);
    if (rc) return rc;
  }
  RefineOut ro;
  rc = verify_and_refine(d, P, prec, nrhs, nshift, shifts, bs, xs, tol, max_iter, hs, st, ro);
  if (rc) return rc;
  int worst = LYM_OK;
  const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  for (int r = 0; r < nrhs; ++r) {
    int status = hs[r].status;
    int cycles = 0;
    for (int k = 0; k < nshift; ++k) {
      if (ro.rel[(size_t)r * nshift + k] > tol && status == ST_CONVERGED) status = ST_MAXITER;
      cycles += ro.cycles[(size_t)r * nshift + k];
    }
    if (stats) {
      lym_solve_stats* s = &stats[r];
      fill_stats(s, hs[r], nshift);
      s->status = status;
      s->restarts = cycles + hs[r].n_ru;  // defect-correction solves + reliable updates of this RHS
      for (int k = 0; k < nshift && k < 64; ++k) {
        s->true_rel_resid[k] = ro.rel[(size_t)r * nshift + k];
        s->initial_rel_resid[k] = ro.rel0[(size_t)r * nshift + k];
        s->refine_cycles[k] = ro.cycles[(size_t)r * nshift + k];
      }
      s->refine_iterations = ro.iters[r];
      s->seconds = secs;
    }
    const int code = status == ST_CONVERGED ? LYM_OK : (status == ST_BREAKDOWN ? LYM_BREAKDOWN : LYM_NOT_CONVERGED);
    if (code > worst) worst = code;
  }
  return worst;
}

int solve_cgm(const OpDesc& d, int prec, int nshift, const double* shifts, const void* b, void* const* xs, double tol,
              int max_iter, void* ws, size_t ws_bytes, lym_solve_stats* stats, hipStream_t st) {
  const void* bs[1] = {b};
  return solve_cgm_batched(d, prec, 1, nshift, shifts, bs, xs, tol, max_iter, ws, ws_bytes, stats, st);
}

// ---------------------------------------------------------------------------
// BiCGStab (dirac_solver.py:56-110) on A = Op + shift or A = Op^2 + shift
// (square), Op = D or Q, complex128, device-resident: every scalar lives in
// the SolverState, an iteration is a fixed launch sequence (chunks of 8
// iterations, HIP graphs), and the host polls four ints per chunk.
// Workspace: state, partials, and r, r_hat, p, nu, s, t, tmp.
// ---------------------------------------------------------------------------
struct BicgParts {
  SolverState* st;
  double* partials;
  Field r, rh, p, nu, s, t, tmp;
};
static size_t bicg_carve(const OpDesc& d, char* base, BicgParts* bp) {
  Carver c{base};
  BicgParts tmp;
  BicgParts& B = bp ? *bp : tmp;
  B.st = (SolverState*)c.take(sizeof(SolverState));
  B.partials = (double*)c.take(sizeof(double) * kMaxRed * (size_t)max_red_blocks(d.lat.V));
  for (Field* f : {&B.r, &B.rh, &B.p, &B.nu, &B.s, &B.t, &B.tmp}) *f = take_field(c, d, P64);
  return c.off + 256;
}

size_t bicgstab_workspace_bytes(const OpDesc& d) { return bicg_carve(d, nullptr, nullptr); }

int solve_bicgstab(const OpDesc& d, int square, double shift, const void* b, const void* x0, void* x, int max_steps,
                   double precision, void* ws, size_t ws_bytes, lym_solve_stats* stats, hipStream_t st) {
  const auto t0 = std::chrono::steady_clock::now();
  graph_new_solve();
  const size_t need = bicgstab_workspace_bytes(d);
  if (!ws || ws_bytes < need) {
    set_error("BiCGStab workspace too small: need " + std::to_string(need) + " bytes");
    return LYM_ERR_ARG;
  }
  BicgParts B;
  char* base = (char*)(((uintptr_t)ws + 255) & ~(uintptr_t)255);
  bicg_carve(d, base, &B);
  const int64_t n = field_elems(d);
  SolverState h;
  std::memset(&h, 0, sizeof(h));
  h.phase = PH_K1;
  h.max_iter = max_steps;
  h.tol2 = precision;
  h.bi_shift = shift;
  LYM_HIP(hipMemcpyAsync(B.st, &h, sizeof(h), hipMemcpyHostToDevice, st));
  LYM_HIP(hipMemcpyAsync(x, x0 ? x0 : b, (size_t)n * 16, hipMemcpyDeviceToDevice, st));
  const RedCtl red{B.partials, nullptr};
  const RedCtl nored{nullptr, nullptr};
  // out = A in (tmp: Op in when squared); gated on a running solve
  auto applyA = [&](Field in, Field out, hipStream_t s) -> int {
    Field src = in;
    if (square) {
      EpiArgs a;
      a.out = B.tmp;
      a.st = B.st;
      const int e = wilson_launch(EPI_GSTORE, d, P64, true, in, a, nored, s);
      if (e) return e;
      src = B.tmp;
    }
    EpiArgs a2;
    a2.out = out;
    a2.aux0 = in;
    a2.sigma = shift;
    a2.st = B.st;
    return wilson_launch(EPI_GAXPY, d, P64, true, src, a2, nored, s);
  };
  const int gr = red_grid_flat(n);
  const int gf = flat_grid(n);
  Field xf;
  xf.data = x;
  int rc = applyA(xf, B.nu, st);
  if (rc) return rc;
  hipLaunchKernelGGL(k_bicg_init, dim3(gr), dim3(kBlock), 0, st, B.st, n, (const double2*)b, (const double2*)x,
                     (const double2*)B.nu.data, (double2*)B.r.data, (double2*)B.rh.data, (double2*)B.p.data,
                     (double2*)B.nu.data, B.partials);
  LYM_HIP(hipGetLastError());
  if ((rc = launch_finalize(FK_BICG_INIT, B.st, B.partials, gr, 3, nullptr, st))) return rc;
  const int per_chunk = 8;
  auto chunk = [&](hipStream_t s) -> int {
    for (int j = 0; j < per_chunk; ++j) {
      hipLaunchKernelGGL(k_bicg_p, dim3(gf), dim3(kBlock), 0, s, B.st, n, (const double2*)B.r.data,
                         (const double2*)B.nu.data, (double2*)B.p.data);
      LYM_HIP(hipGetLastError());
      int e = applyA(B.p, B.nu, s);
      if (e) return e;
      hipLaunchKernelGGL(k_bicg_dots, dim3(gr), dim3(kBlock), 0, s, B.st, n, (const double2*)B.rh.data,
                         (const double2*)B.nu.data, B.partials);
      LYM_HIP(hipGetLastError());
      if ((e = launch_finalize(FK_BICG_A, B.st, B.partials, gr, 3, nullptr, s))) return e;
      hipLaunchKernelGGL(k_bicg_s, dim3(gf), dim3(kBlock), 0, s, B.st, n, (const double2*)B.r.data,
                         (const double2*)B.nu.data, (double2*)B.s.data);
      LYM_HIP(hipGetLastError());
      if ((e = applyA(B.s, B.t, s))) return e;
      hipLaunchKernelGGL(k_bicg_dots, dim3(gr), dim3(kBlock), 0, s, B.st, n, (const double2*)B.t.data,
                         (const double2*)B.s.data, B.partials);
      LYM_HIP(hipGetLastError());
      if ((e = launch_finalize(FK_BICG_W, B.st, B.partials, gr, 3, nullptr, s))) return e;
      hipLaunchKernelGGL(k_bicg_xr, dim3(gr), dim3(kBlock), 0, s, B.st, n, (double2*)x, (const double2*)B.p.data,
                         (const double2*)B.s.data, (const double2*)B.t.data, (const double2*)B.rh.data,
                         (double2*)B.r.data, B.partials);
      LYM_HIP(hipGetLastError());
      if ((e = launch_finalize(FK_BICG_R, B.st, B.partials, gr, 3, nullptr, s))) return e;
    }
    return LYM_OK;
  };
  KeyBuf key;
  key.add('B').add(square).add(shift).add(B.st).add(x).add(b);
  rc = run_chunks(B.st, (int64_t)max_steps / per_chunk + 8, st, op_key(d) + key.s, chunk,
                  [](const int*) { return (int)LYM_OK; });
  if (rc) return rc;
  LYM_HIP(hipMemcpyAsync(&h, B.st, sizeof(SolverState), hipMemcpyDeviceToHost, st));
  LYM_HIP(hipStreamSynchronize(st));
  fill_stats(stats, h, 1);
  if (stats) stats->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  return h.status == ST_CONVERGED ? LYM_OK : (h.status == ST_BREAKDOWN ? LYM_BREAKDOWN : LYM_NOT_CONVERGED);
}

int solve_mr(const OpDesc& d, int nshift, const double* shifts, int first, const void* b, void* const* xs,
             int max_steps, double precision, double omega, void* ws, size_t ws_bytes, lym_solve_stats* stats,
             hipStream_t st) {
  const auto t0 = std::chrono::steady_clock::now();
  graph_new_solve();
  if (nshift < 1 || nshift > kMaxShifts || first < 0 || first >= nshift) {
    set_error("bad shift list for the multishift MR solver");
    return LYM_ERR_ARG;
  }
  const size_t need = solver_workspace_bytes(d, P64, nshift, 1);
  if (!ws || ws_bytes < need) {
    set_error("solver workspace too small: need " + std::to_string(need) + " bytes");
    return LYM_ERR_ARG;
  }
  Parts P;
  char* base = (char*)(((uintptr_t)ws + 255) & ~(uintptr_t)255);
  carve(d, P64, nshift, 1, base, &P);
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
  const int64_t n = field_elems(d);
  const int gw = wilson_blocks(d, P64);  // Wilson grid: number of partials its reductions write
  const int gr = red_grid_flat(n);
  const int gf = flat_grid(n);
  hipLaunchKernelGGL(k_mr_init, dim3(gf), dim3(kBlock), 0, st, n, (const double2*)b, (double2*)P.r64.data, P.xs_dev,
                     first, nshift);
  LYM_HIP(hipGetLastError());
  const int per_chunk = 8;
  auto chunk = [&](hipStream_t st) -> int {
    for (int j = 0; j < per_chunk; ++j) {
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
      if (!e) e = launch_finalize(FK_MR, P.st, P.partials, gw, 3, nullptr, st);
      if (e) return e;
      hipLaunchKernelGGL(k_mr_k3, dim3(gr), dim3(kBlock), 0, st, P.st, n, (double2*)P.r64.data,
                         (const double2*)P.u64.data, P.xs_dev, first, nshift, P.partials);
      LYM_HIP(hipGetLastError());
      if ((e = launch_finalize(FK_MR_RR, P.st, P.partials, gr, 1, nullptr, st))) return e;
    }
    return LYM_OK;
  };
  KeyBuf key;
  key.add('M').add(nshift).add(P.st).add(P.xs_dev);
  int rc = run_chunks(P.st, (int64_t)max_steps / per_chunk + 8, st, op_key(d) + key.s, chunk,
                      [](const int*) { return (int)LYM_OK; });
  if (rc) return rc;
  LYM_HIP(hipMemcpyAsync(&h, P.st, sizeof(SolverState), hipMemcpyDeviceToHost, st));
  LYM_HIP(hipStreamSynchronize(st));
  fill_stats(stats, h, nshift);
  if (stats) stats->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  return h.status == ST_CONVERGED ? LYM_OK : (h.status == ST_BREAKDOWN ? LYM_BREAKDOWN : LYM_NOT_CONVERGED);
}

}  // namespace lym
