// Scalar recurrences of the solvers, run by thread 0 of the one-block
// finalizer kernel after it has summed the partials of the preceding
// reducing kernel.
#pragma once
#include <hip/hip_runtime.h>
#include <math.h>

#include "lym_state.h"

namespace lym {

// Per-shift inputs of the multishift recurrences.  The finalizer block
// stages them in LDS, one thread per shift, before thread 0 runs the serial
// recurrence: thread 0 then reads LDS instead of walking a dependent chain
// of global loads per shift (stores to the state keep those loads from
// being overlapped), and it still makes every write to the state itself.
struct ShiftIn {
  double zeta[kMaxShifts], zeta_old[kMaxShifts], zeta_new[kMaxShifts], alpha_s[kMaxShifts];
  double delta[kMaxShifts], tol2_s[kMaxShifts];
  int active[kMaxShifts];
};

// A block of deferred multishift steps folded into one update.  Steps
// j < mk of the CG-M shift recurrence (r_j the base residual of step j)
//   x += a_j p ;  p = z_j r_j + b_j p
// amount to
//   x <- x + A p + sum_j d_j r_j ,   p <- B p + sum_j e_j r_j
// with, for j < mk,
//   d_j = z_j S_j ,  e_j = z_j prod_{l > j} b_l ,
//   S_j = sum_{i > j} a_i prod_{j < l < i} b_l   (S_{mk-1} = 0, S_{j-1} = a_j + b_j S_j),
//   A = S_{-1} = sum_i a_i prod_{l < i} b_l ,  B = prod_l b_l .
// The residuals of a CG recurrence are mutually orthogonal and p before the
// block is orthogonal to all of them, so the terms of each sum are
// orthogonal and no cancellation enters the FP64 accumulation.  Writes
// de[j * stride] = (d_j, e_j) for j < mk.
__host__ __device__ inline void fold_shift_block(const double* a, const double* z, const double* b, int mk,
                                                 double2* de, int stride, double& A, double& B) {
  double S = 0.0, Gp = 1.0;
  for (int j = mk - 1; j >= 0; --j) {
    de[j * stride] = make_double2(z[j] * S, z[j] * Gp);
    S = fma(b[j], S, a[j]);
    Gp *= b[j];
  }
  A = S;
  B = Gp;
}

// Every thread of the block calls this (one barrier); kMaxShifts <= block size.
__device__ inline void stage_shifts(const SolverState* st, ShiftIn& in) {
  const int s = (int)threadIdx.x;
  if (s < st->nshift) {
    in.zeta[s] = st->zeta[s];
    in.zeta_old[s] = st->zeta_old[s];
    in.zeta_new[s] = st->zeta_new[s];
    in.alpha_s[s] = st->alpha_s[s];
    in.delta[s] = st->delta[s];
    in.tol2_s[s] = st->tol2_s[s];
    in.active[s] = st->active[s];
  }
  __syncthreads();
}

// CG-M after <p, (A+s0) p> is known: step length alpha and the shifted
// zeta_{k+1}, alpha^s_k (Jegerlehner hep-lat/9612014):
//   zeta_{k+1} = zeta_k zeta_{k-1} a_{k-1} /
//                ( a_k b_{k-1} (zeta_{k-1} - zeta_k) + zeta_{k-1} a_{k-1} (1 + d_s a_k) )
//   alpha^s_k  = a_k zeta_{k+1} / zeta_k
__device__ inline void cgm_fin_a(SolverState* st, const ShiftIn& in, double tt, double pp) {
  const double pAp = tt + st->sigma0 * pp;
  if (!(pAp > 0.0) || !isfinite(pAp)) {
    st->status = ST_BREAKDOWN;
    st->phase = PH_DONE;
    return;
  }
  const double alpha = st->rr / pAp;
  st->pAp = pAp;
  st->alpha = alpha;
  const double ao = st->alpha_old, bo = st->beta_old;
  for (int s = 0; s < st->nshift; ++s) {
    if (!in.active[s]) continue;
    const double z = in.zeta[s], zo = in.zeta_old[s];
    const double den = alpha * bo * (zo - z) + zo * ao * (1.0 + in.delta[s] * alpha);
    const double zn = z * zo * ao / den;
    st->zeta_new[s] = zn;
    st->alpha_s[s] = alpha * zn / z;
  }
  st->phase = PH_K2;
}

// Convergence of the shifts after <r_{k+1}, r_{k+1}> (the zetas already
// rolled: zeta_s = in.zeta_new[s]): |zeta_s|^2 rr <= tol2_s, the base last
// (it carries the common Krylov space: it stops only after all others);
// status and final_k3 of the solve.
__device__ inline void cgm_converged(SolverState* st, const ShiftIn& in, double rrn) {
  int nact = 0;
  for (int s = 0; s < st->nshift; ++s) {
    if (!in.active[s] || s == st->base) continue;
    const double zn = in.zeta_new[s];
    if (zn * zn * rrn <= in.tol2_s[s]) {
      st->active[s] = 0;
      st->conv_iter[s] = st->iter;
    } else {
      ++nact;
    }
  }
  if (in.active[st->base]) {
    if (nact == 0 && rrn <= in.tol2_s[st->base]) {
      st->active[st->base] = 0;
      st->conv_iter[st->base] = st->iter;
    } else {
      ++nact;
    }
  }
  st->nactive = nact;
  if (!isfinite(rrn)) {
    st->status = ST_BREAKDOWN;
    st->final_k3 = 1;
  } else if (nact == 0) {
    st->status = ST_CONVERGED;
    st->final_k3 = 1;
  } else if (st->iter >= st->max_iter) {
    st->status = ST_MAXITER;
    st->final_k3 = 1;
  }
}

// CG-M after <r_{k+1}, r_{k+1}> is known: beta, beta^s, roll the recurrence,
// per-shift convergence, stop decision.
__device__ inline void cgm_fin_b(SolverState* st, const ShiftIn& in, double rrn) {
  const double beta = rrn / st->rr;
  for (int s = 0; s < st->nshift; ++s) {
    st->upd[s] = in.active[s];
    if (!in.active[s]) continue;
    const double z = in.zeta[s], zn = in.zeta_new[s];
    const double q = zn / z;
    st->beta_s[s] = beta * q * q;
    st->zeta_old[s] = z;
    st->zeta[s] = zn;
  }
  st->alpha_old = st->alpha;
  st->beta_old = beta;
  st->rr = rrn;
  st->iter += 1;
  cgm_converged(st, in, rrn);
  st->phase = PH_K3;
}

// Fast CG-M after <r_{k+1}, r_{k+1}>: as cgm_fin_b, but the shift updates of
// this iteration are recorded as step blk_j of the current block instead of
// being applied now; the block is flushed (PH_K3) when full or when the solve
// stops, otherwise the next iteration starts with p = r + beta p (PH_KP).
__device__ inline void cgm_fin_b_fast(SolverState* st, const ShiftIn& in, double rrn) {
  const double beta = rrn / st->rr;
  const int j = st->blk_j;
  for (int s = 0; s < st->nshift; ++s) {
    if (!in.active[s]) continue;
    const double z = in.zeta[s], zn = in.zeta_new[s];
    const double q = zn / z;
    st->h_alpha[s][j] = in.alpha_s[s];
    st->h_zeta[s][j] = zn;
    st->h_beta[s][j] = beta * q * q;
    st->n_upd[s] = j + 1;
    st->zeta_old[s] = z;
    st->zeta[s] = zn;
  }
  st->alpha_old = st->alpha;
  st->beta_old = beta;
  st->rr = rrn;
  st->iter += 1;
  st->blk_j = j + 1;
  cgm_converged(st, in, rrn);
  if (rrn <= st->switch_rr) st->want_switch = 1;
  st->phase = (st->final_k3 || st->blk_j >= st->mblk) ? PH_K3 : PH_KP;
}

// Mixed CG after the low-precision <r_{k+1}, r_{k+1}>: beta = rrn / rr and
// p = r + beta p next (PH_KP), unless a reliable update is due: the iterated
// residual fell by delta below the true residual of the last update, or
// passed the stopping test, or the iteration budget is spent (PH_RU; rr keeps
// <r_k, r_k> so the update can form beta from the true residual).  rr_max is
// the true residual of the last update (|b|^2 before the first).
__device__ inline void cg_fin_b_rel(SolverState* st, double rrn) {
  // step blk_j of the current solution block: x_hi += alpha p is deferred to
  // the block's fold (rel_fold_due); beta is recorded below when the
  // recurrence continues
  const int j = st->blk_j;
  st->h_alpha[0][j] = st->alpha;
  st->h_beta[0][j] = 0.0;
  st->blk_j = j + 1;
  st->iter += 1;
  if (!isfinite(rrn)) {
    st->status = ST_BREAKDOWN;
    st->phase = PH_DONE;
    return;
  }
  const bool stop = rrn <= st->tol2_s[0];
  if (st->iter >= st->max_iter) st->status = ST_MAXITER;
  if (stop || rrn <= st->delta2 * st->rr_max || st->status == ST_MAXITER) {
    st->ru_final = stop ? 1 : 0;
    st->phase = PH_RU;
    return;
  }
  st->beta_old = rrn / st->rr;
  st->h_beta[0][j] = st->beta_old;
  st->kp_a_re = 1.0;
  st->kp_a_im = 0.0;
  st->rr = rrn;
  st->phase = PH_KP;
}

// The mixed CG's deferred solution update runs (and its block restarts)
// before a reliable update, or when the block is full while iterating.
__host__ __device__ inline bool rel_fold_due(const SolverState* st) {
  return (st->phase == PH_RU && st->blk_j > 0) || (st->phase == PH_KP && st->blk_j >= st->mblk);
}

// After the FP64 true residual of a reliable update: converged when it meets
// the target, stopped at the iteration budget, else the low-precision
// recurrence restarts from it with p = r_true + beta (p - c r_true),
// beta = rr_true / rr_k, c = <r_true, p> / rr_true: the kept direction is
// made orthogonal to the new residual again (as in Clark et al.,
// arXiv:0911.3191), which low-precision storage of p otherwise erodes.
__device__ inline void cg_fin_ru(SolverState* st, double rr_true, double rp_re, double rp_im) {
  st->n_ru += 1;
  st->blk_j = 0;  // the block was folded into x_hi before the update
  if (!isfinite(rr_true)) {
    st->status = ST_BREAKDOWN;
    st->phase = PH_DONE;
    return;
  }
  if (rr_true <= st->tol2_true) {
    st->status = ST_CONVERGED;
    st->rr = rr_true;
    st->phase = PH_DONE;
    return;
  }
  if (st->status == ST_MAXITER) {
    st->rr = rr_true;
    st->phase = PH_DONE;
    return;
  }
  const double beta = rr_true / st->rr;
  st->beta_old = beta;
  st->kp_a_re = 1.0 - beta * rp_re / rr_true;
  st->kp_a_im = -beta * rp_im / rr_true;
  st->rr = rr_true;
  st->rr_max = rr_true;
  st->phase = PH_KP;
}

// Multishift MR (dirac_solver.py:35-43): alpha = omega <p,r>/<p,p>;
// f_i <- f_i / (1 + (s_i - s_N) alpha); x update coefficients c_i = alpha f_i.
__device__ inline void mr_fin(SolverState* st, double pr_re, double pr_im, double pp) {
  if (!(pp > 0.0) || !isfinite(pp)) {
    st->status = ST_BREAKDOWN;
    st->phase = PH_DONE;
    return;
  }
  const double ar = st->omega * pr_re / pp, ai = st->omega * pr_im / pp;
  st->mr_alpha_re = ar;
  st->mr_alpha_im = ai;
  const int n = st->nshift;
  const double sN = st->shift[n - 1];
  for (int i = st->base; i < n - 1; ++i) {
    // d = 1 + (s_i - s_N) alpha ; f /= d
    const double ds = st->shift[i] - sN;
    const double dr = 1.0 + ds * ar, di = ds * ai;
    const double den = dr * dr + di * di;
    const double fr = st->f_re[i], fi = st->f_im[i];
    const double nr = (fr * dr + fi * di) / den, ni = (fi * dr - fr * di) / den;
    st->f_re[i] = nr;
    st->f_im[i] = ni;
    st->c_re[i] = ar * nr - ai * ni;
    st->c_im[i] = ar * ni + ai * nr;
  }
  st->c_re[n - 1] = ar;
  st->c_im[n - 1] = ai;
  st->phase = PH_K3;
}

// After r <- r - alpha p: stop when |<r,r>| < precision (dirac_solver.py:47-50).
__device__ inline void mr_fin_rr(SolverState* st, double rr) {
  st->rr = rr;
  st->iter += 1;
  if (!isfinite(rr)) {
    st->status = ST_BREAKDOWN;
    st->phase = PH_DONE;
  } else if (fabs(rr) < st->tol2) {
    st->status = ST_CONVERGED;
    st->phase = PH_DONE;
  } else if (st->iter >= st->max_iter) {
    st->status = ST_MAXITER;
    st->phase = PH_DONE;
  } else {
    st->phase = PH_K1;
  }
}

// ---- BiCGStab (dirac_solver.py:56-110), complex scalars (re, im) -----------
__device__ inline void cdiv(const double* a, const double* b, double* out) {
  const double d = b[0] * b[0] + b[1] * b[1];
  const double re = (a[0] * b[0] + a[1] * b[1]) / d, im = (a[1] * b[0] - a[0] * b[1]) / d;
  out[0] = re;
  out[1] = im;
}
__device__ inline void cmulv(const double* a, const double* b, double* out) {
  const double re = a[0] * b[0] - a[1] * b[1], im = a[0] * b[1] + a[1] * b[0];
  out[0] = re;
  out[1] = im;
}
// beta = (rho_next / rho) (alpha / omega)   (dirac_solver.py:80)
__device__ inline bool bicg_beta(SolverState* st) {
  if (st->bi_rho[0] == 0.0 && st->bi_rho[1] == 0.0) return false;
  if (st->bi_omega[0] == 0.0 && st->bi_omega[1] == 0.0) return false;
  double q1[2], q2[2];
  cdiv(st->bi_rho_next, st->bi_rho, q1);
  cdiv(st->bi_alpha, st->bi_omega, q2);
  cmulv(q1, q2, st->bi_beta);
  return isfinite(st->bi_beta[0]) && isfinite(st->bi_beta[1]);
}
__device__ inline void bicg_stop(SolverState* st, int status) {
  st->status = status;
  st->phase = PH_DONE;
}
// after the initial residual: alpha = omega = rho = 1 (dirac_solver.py:71-76),
// rho_next = <r_hat, r> of the first iteration.  A zero residual (b = A x0
// exactly) is solved at once.
__device__ inline void bicg_fin_init(SolverState* st, double rr, double re, double im) {
  st->bi_rho[0] = st->bi_alpha[0] = st->bi_omega[0] = 1.0;
  st->bi_rho[1] = st->bi_alpha[1] = st->bi_omega[1] = 0.0;
  st->bi_rho_next[0] = re;
  st->bi_rho_next[1] = im;
  st->rr = rr;
  st->bb = rr;
  if (!isfinite(rr)) return bicg_stop(st, ST_BREAKDOWN);
  if (rr == 0.0) return bicg_stop(st, ST_CONVERGED);
  if (!bicg_beta(st)) return bicg_stop(st, ST_BREAKDOWN);
}
// alpha = rho_next / <r_hat, nu>   (dirac_solver.py:88)
__device__ inline void bicg_fin_a(SolverState* st, double re, double im) {
  const double g[2] = {re, im};
  if ((re == 0.0 && im == 0.0) || !isfinite(re) || !isfinite(im)) return bicg_stop(st, ST_BREAKDOWN);
  cdiv(st->bi_rho_next, g, st->bi_alpha);
}
// omega = <t, s> / <t, t>   (dirac_solver.py:91); t = 0 (then s = 0) gives 0
__device__ inline void bicg_fin_w(SolverState* st, double re, double im, double tt) {
  if (!isfinite(tt) || !isfinite(re) || !isfinite(im)) return bicg_stop(st, ST_BREAKDOWN);
  st->bi_omega[0] = tt > 0.0 ? re / tt : 0.0;
  st->bi_omega[1] = tt > 0.0 ? im / tt : 0.0;
}
// after r = s - omega t: stop when |<r, r>| < precision (absolute,
// dirac_solver.py:96-103); else rho = rho_next, rho_next = <r_hat, r>, beta
__device__ inline void bicg_fin_r(SolverState* st, double rr, double re, double im) {
  st->iter += 1;
  st->rr = rr;
  if (!isfinite(rr)) return bicg_stop(st, ST_BREAKDOWN);
  if (fabs(rr) < st->tol2) return bicg_stop(st, ST_CONVERGED);
  if (st->iter >= st->max_iter) return bicg_stop(st, ST_MAXITER);
  st->bi_rho[0] = st->bi_rho_next[0];
  st->bi_rho[1] = st->bi_rho_next[1];
  st->bi_rho_next[0] = re;
  st->bi_rho_next[1] = im;
  if (!bicg_beta(st)) return bicg_stop(st, ST_BREAKDOWN);
}

}  // namespace lym
