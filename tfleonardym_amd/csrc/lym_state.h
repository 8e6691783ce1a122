// Device-resident solver state.  All Krylov scalars live here in FP64 and are
// updated by the last block of each reducing kernel, so an iteration is a
// fixed launch sequence with no host round trip.  `phase` is a small state
// machine: a kernel whose phase does not match returns at once, so iterations
// launched after convergence cost only their launch.
#pragma once
#include <stdint.h>

namespace lym {

constexpr int kMaxShifts = 64;
constexpr int kMaxBlock = 128;  // deferred multishift updates: iterations per block

enum Phase : int {
  PH_K1 = 0,    // t = Q p
  PH_K2 = 1,    // Ap, residual update
  PH_K3 = 2,    // per-shift vector updates (plain) / deferred block flush (fast)
  PH_DONE = 3,
  PH_KP = 4,    // fast path: p = r + beta p
  PH_RU = 5,    // mixed CG: reliable update (FP64 x += x_lo, true residual)
};

enum SolveStatus : int { ST_RUNNING = 0, ST_CONVERGED = 1, ST_MAXITER = 2, ST_BREAKDOWN = 3 };

struct SolverState {
  // the first four ints are what the host polls between chunks
  int phase;
  int iter;
  int want_switch;  // fast CG-M: base residual below the precision-switch threshold
  int status;
  int max_iter;
  int nshift;
  int base;      // index of the base (smallest) shift
  int nactive;
  int final_k3;  // K3 of the last iteration must still run
  double rr;     // <r, r> (base residual, iterated)
  double bb;     // <b, b>
  double tol2;   // stopping threshold on rr (absolute)
  double alpha, alpha_old, beta, beta_old;
  double sigma0;  // base shift
  double omega;   // MR over-relaxation
  double pAp;
  double mr_alpha_re, mr_alpha_im;
  double zeta[kMaxShifts], zeta_old[kMaxShifts], zeta_new[kMaxShifts];
  double alpha_s[kMaxShifts], beta_s[kMaxShifts], delta[kMaxShifts];
  double shift[kMaxShifts];
  double tol2_s[kMaxShifts];  // per-shift threshold on zeta^2 rr
  // MR multishift (complex): f_i and the x update coefficient alpha*f_i
  double f_re[kMaxShifts], f_im[kMaxShifts];
  double c_re[kMaxShifts], c_im[kMaxShifts];
  int upd[kMaxShifts];     // update x_s / p_s in this iteration's K3
  int active[kMaxShifts];  // still iterating
  int conv_iter[kMaxShifts];
  // ---- fast CG-M: deferred shift updates over blocks of `mblk` iterations.
  // Step j of the current block used r^{k0+j+1} (ring slot (k0+j+1) % (mblk+1))
  // and the scalars below; n_upd[s] = steps recorded for shift s.  The
  // history is stored shift-major ([shift][step]) so the update kernel reads
  // one shift's steps as consecutive scalars.
  int mblk;
  int blk_j;         // steps recorded in the current block
  int precision_phase;  // 0 = FP64 Krylov vectors, 1 = complex64, 2 = 16-bit
  int switch_counted;   // this RHS has been counted in the batch's switch tally
  double switch_rr;  // absolute threshold on rr for want_switch (next phase)
  double switch2_rr; // threshold of the phase after that (0: none)
  int n_upd[kMaxShifts];
  double h_alpha[kMaxShifts][kMaxBlock];
  double h_zeta[kMaxShifts][kMaxBlock];
  double h_beta[kMaxShifts][kMaxBlock];
  // ---- mixed-precision CG with reliable updates (one shift): the low-
  // precision recurrence runs until rr < delta2 * rr_max (rr_max = the true
  // residual |r|^2 of the last reliable update, or |b|^2 before the first),
  // then x_hi += x_lo and r = b - (A + s) x_hi in FP64 replace the iterated
  // residual.
  double rr_max;
  double delta2;
  double tol2_true;  // target on the true rr (absolute after the init scaling)
  // next direction p = kp_a r + beta_old p; kp_a = 1 except right after a
  // reliable update, where it also removes the component of p along the new
  // residual (kp_a = 1 - beta <r,p>/<r,r>)
  double kp_a_re, kp_a_im;
  int ru_final;   // the pending update follows the iterated stopping test
  int n_ru;       // reliable updates done
  // ---- BiCGStab (dirac_solver.py:56-110): complex scalars (re, im)
  double bi_rho[2], bi_rho_next[2], bi_alpha[2], bi_omega[2], bi_beta[2];
  double bi_shift;  // A = Op (+ Op) + shift
};

}  // namespace lym
