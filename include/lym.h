/*
 * lym.h -- C ABI of liblym.so, the MI355X (gfx950) implementation of the
 * TfLeonardYM fermion-matrix hot path: the Wilson-Dirac operator D / Q = g5 D,
 * D^dag D, the reference's multishift solver, and (new) CG / multishift CG.
 *
 * Conventions shared by every entry point
 *   - All field pointers are DEVICE pointers.  The library borrows them for
 *     the duration of the call and keeps nothing.
 *   - Layouts are the reference's (site index last, last lattice coordinate
 *     fastest, lattice.py:92-114):
 *       links               (ndim, nc, nc, V) complex
 *       fundamental spinor  (Ns, nc, V) complex          Ns = 2 (2D) or 4 (4D)
 *       adjoint spinor      (Ns, nc, nc, V) complex matrices
 *   - Precision LYM_PREC_F64 = complex128, LYM_PREC_F32 = complex64.
 *   - `stream` is a hipStream_t (NULL = the default stream).  Calls are
 *     asynchronous on that stream unless stated otherwise.
 *   - Return value: 0 OK, <0 error (see lym_last_error()), >0 solver status
 *     (not converged / breakdown).
 */
#ifndef LYM_H
#define LYM_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define LYM_ABI_VERSION 6

enum {
  LYM_OK = 0,
  LYM_ERR_ARG = -1,
  LYM_ERR_HIP = -2,
  LYM_ERR_NOMEM = -3,
  LYM_ERR_UNSUPPORTED = -4,
  LYM_NOT_CONVERGED = 1,
  LYM_BREAKDOWN = 2
};

enum { LYM_PREC_F64 = 0, LYM_PREC_F32 = 1, LYM_PREC_F16 = 2 };
enum { LYM_REP_FUNDAMENTAL = 0, LYM_REP_ADJOINT = 1 };
enum { LYM_MODE_PROJECTED = 0, LYM_MODE_SIMPLE = 1 };

/* Operator description (replaces the state DiracWilsonOperator builds in
 * dirac.py:254-378).  `links` are the gauge links as stored by the caller;
 * the antiperiodic boundary factor (boundary_conditions.py:23-46) is applied
 * on the fly to U_{ndim-1}(x) at x_{ndim-1} = L-1 when (bc_re, bc_im) != (1,0).
 * Adjoint representation: the hop is U_left X U_right^dag (dirac.py:416-426);
 * links_right = NULL means U_right = links without the boundary factor. */
typedef struct lym_wilson {
  int32_t ndim;           /* 2 or 4 (dirac.py:299-300) */
  int32_t nc;             /* 2 or 3 */
  int32_t representation; /* LYM_REP_* */
  int32_t hermitian;      /* 1: Q = g5 D (dirac.py:649-662) */
  int32_t dims[4];        /* lattice extents, direction order (--geometry) */
  double kappa;           /* D = 1 + kappa * hopping (dirac.py:458-523) */
  double bc_re, bc_im;    /* boundary factor, (1,0) = periodic */
  const void* links;            /* complex128 */
  const void* links_right;      /* complex128 or NULL (adjoint only) */
  const void* links_f32;        /* complex64 copy of links (lym_pack_links_f32 layout), or NULL */
  const void* links_right_f32;  /* the same of links_right, or NULL */
  /* LYM_LINKS_SU: every link is special unitary (checked by the caller with
   * lym_link_deviation); fundamental-representation kernels then read the
   * first nc-1 rows of each link and rebuild the last one (SU(3): row 2 =
   * conj(row 0 x row 1); SU(2): row 1 = (-conj(u01), conj(u00))).
   * LYM_LINKS_FULL: every entry is read (any complex matrices). */
  int32_t link_format;
  int32_t reserved;
} lym_wilson;

#define LYM_LINKS_FULL 0
#define LYM_LINKS_SU 1

typedef struct lym_solve_stats {
  int32_t status;      /* 1 converged, 2 max iterations, 3 breakdown */
  int32_t iterations;  /* iterations of the (base) recurrence */
  int32_t nshift;
  int32_t restarts;    /* reliable updates + defect-correction solves of this RHS */
  double final_rr;     /* last iterated <r,r> of the base system */
  double seconds;      /* wall time of the call (host clock, synchronised) */
  int32_t conv_iter[64];
  double true_rel_resid[64];    /* ||b - (A+s)x|| / ||b||, FP64, when computed */
  double initial_rel_resid[64]; /* the same before any defect correction */
  int32_t refine_cycles[64];    /* defect-correction solves of each shift */
  int32_t refine_iterations;    /* iterations of those solves (not in `iterations`) */
  int32_t precision_phase;      /* fast CG-M: last Krylov storage, 0 FP64, 1 complex64, 2 16-bit */
} lym_solve_stats;

/* ---- runtime --------------------------------------------------------- */
int lym_abi_version(void);
/* Select `device` for this host thread's subsequent calls; no-op if current. */
int lym_init(int device);
/* Message of the last error on this host thread ("" if none). */
const char* lym_last_error(void);

/* ---- operator ---------------------------------------------------------
 * out = D in, or Q in when op->hermitian.
 * replaces DiracWilsonOperator.apply (dirac.py:270-277), i.e.
 * multiply_in_fundamental_representation (dirac.py:458-501),
 * simple_multiply_in_fundamental_representation (dirac.py:503-523),
 * multiply_in_adjoint_representation (dirac.py:404-456),
 * simple_multiply_in_adjoint_representation (dirac.py:380-402).
 * mode: LYM_MODE_PROJECTED (half spinors) or LYM_MODE_SIMPLE (full
 * projectors; F64 only).  `in` and `out` must not overlap. */
int lym_wilson_apply(const lym_wilson* op, int prec, int mode, const void* in, void* out, void* stream);

/* out = Op(Op(in)) + shift * in, using `tmp` (same shape) as scratch.
 * replaces SquareOperator.apply (dirac.py:36-37) composed with
 * ShiftedOperator.apply (dirac.py:24-25). */
int lym_wilson_apply_square(const lym_wilson* op, int prec, double shift, const void* in, void* out,
                            void* tmp, void* stream);

/* ---- BLAS-1 over n complex numbers ------------------------------------
 * lym_dot: result[0] + i result[1] = sum conj(x) y, accumulated in FP64,
 * deterministic; synchronises `stream` and writes `result` on the host.
 * replaces algebra_utils.dot (algebra_utils.py:5-11) + mth_global_sum
 * (lattice.py:375-376). `scratch` >= lym_reduce_scratch_bytes(n) bytes. */
size_t lym_reduce_scratch_bytes(int64_t n);
int lym_dot(int prec, int64_t n, const void* x, const void* y, double* result, void* scratch, void* stream);
/* lym_dot without a scratch argument and without allocation: no counterpart in
 * the reference; for callers that cannot call hipMalloc (code under HIP graph
 * capture, tight serving loops, harnesses that forbid allocation).
 * result_dev[0] + i result_dev[1] = sum conj(x) y in FP64 with lym_dot's
 * kernels and summation order (bitwise the same value), written to DEVICE
 * memory, asynchronously on `stream` (no synchronisation, no host copy).
 * Scratch: a static 32 KiB device array per GPU (2048 block partials; any n,
 * the reduction grid is capped there), held until the process ends; its
 * device address is looked up by lym_init or the first call on a device.
 * Calls that overlap on one GPU share that array and race. */
int lym_dot_noalloc(int prec, int64_t n, const void* x, const void* y, double* result_dev, void* stream);
/* y = a x + b y (complex scalars); replaces the solver updates written as
 * tensor arithmetic in dirac_solver.py:33-45 and dirac_solver.py:83-97. */
int lym_axpby(int prec, int64_t n, double a_re, double a_im, const void* x, double b_re, double b_im, void* y,
              void* stream);
/* Negate components with spin >= Ns/2 of `n_per_spin * ns` complex numbers;
 * replaces apply_hermiticity / apply_gammad (dirac.py:649-662, 686-696). */
int lym_gamma5(int prec, int ns, int64_t n_per_spin, const void* in, void* out, void* stream);
/* out = a x + b y over `n_per_spin * ns` complex128 numbers, with complex
 * coefficients that differ between spins < ns/2 and the others:
 * coef = {a_up re, im, a_lo re, im, b_up re, im, b_lo re, im}.  Replaces the
 * per-spin combination of Overlap.apply (dirac.py:120-141). */
int lym_spin_axpby(int ns, int64_t n_per_spin, const double* coef, const void* x, const void* y, void* out,
                   void* stream);
/* Precision conversion of n complex numbers (F64 <-> F32). No reference
 * counterpart: the reference runs in complex128 only. */
int lym_convert(int prec_in, const void* in, int prec_out, void* out, int64_t n, void* stream);
/* complex64 copy of the complex128 links (ndim, nc, nc, volume) in the layout
 * the kernels read (lym_wilson.links_f32): per direction mu, U = ceil(nc^2/2)
 * planes of `volume` 16-byte units; unit q of (mu, site) holds the row-major
 * matrix elements 2q and 2q+1 as two complex64 numbers (zero padding after
 * the last element), so a kernel lane reads a link in 16-byte loads.
 * lym_links_f32_bytes gives the size (0 for bad arguments).  No reference
 * counterpart (the reference keeps complex128 links only, lattice.py). */
size_t lym_links_f32_bytes(int32_t ndim, int32_t nc, int64_t volume);
int lym_pack_links_f32(int32_t ndim, int32_t nc, int64_t volume, const void* links, void* links_f32, void* stream);
/* dev_out[0] (device double) = max over links and entries of
 * |stored last row - row rebuilt from the others| for op->links (F64) or
 * op->links_f32 (F32), the test that allows LYM_LINKS_SU.  No reference
 * counterpart (the reference always reads the full matrix); the boundary
 * conditions materialised by dirac.py:351-378 leave links that fail it. */
int lym_link_deviation(const lym_wilson* op, int prec, double* dev_out, void* stream);

/* ---- fermion force ------------------------------------------------------
 * out (+)= sum_p weights[p] * [ L(xs[p], ys[p]) + symmetric * L(ys[p], xs[p]) ]
 * with L(X, Y) = DiracWilsonOperator.lie_derivative(X, Y) (dirac.py:525-565):
 * a (D, nc, nc, V) complex128 field kappa * sum_k T^k l_k(X, Y) on the
 * operator's links (boundary factor included).  symmetric = 1 with
 * ys[p] = Op xs[p] and weights = the rational alphas is the sum
 * n_flavor.force builds (fermion_action.py:64-93).  Spinors are complex128 in
 * the operator's layout: (Ns, nc, V), or (Ns, nc, nc, V) matrices for the
 * adjoint representation (the reference's conversion to x^a = 2 tr(T^a X)
 * and its real adjoint links are folded in; the result is identical).
 * accumulate = 0 overwrites out (npairs = 0 then zeroes it), 1 adds to it.
 * xs, ys, weights: host arrays of npairs entries (device pointers in xs/ys);
 * no entry may alias out.  F64 only.  Asynchronous on `stream`. */
int lym_lie_derivative(const lym_wilson* op, int npairs, const void* const* xs, const void* const* ys,
                       const double* weights, int symmetric, int accumulate, void* out, void* stream);

/* ---- gauge-field files --------------------------------------------------
 * Unpack a gauge field stored as U* (the conjugate of U) with two columns
 * per SU(3) link in XDR floats (big-endian IEEE single) in the order
 * (site, mu, row, col < 2, re/im), the words read_config reads
 * (lattice.py:10-70): conjugate, normalise column 0, orthogonalise column 1
 * against it and normalise, column 2 = conj(col0 x col1), written to `links`
 * ((ndim, 3, 3, volume) complex128).  raw: device copy of the
 * volume * ndim * 12 words exactly as in the file.  Asynchronous. */
int lym_unpack_ustar(const void* raw, int64_t volume, int ndim, void* links, void* stream);
/* out = X (X^dag X)^-1/2 / det(...)^(1/nc) for every link X of `in`
 * ((ndim, nc, nc, volume) complex128; nc 2 or 3): the unitary polar factor
 * divided by the principal nc-th root of its determinant, as
 * unitary_sqrt_projection (smearing.py:18-31) builds the random links of
 * hotstart (hmc.py:16-31).  in and out may be the same.  Asynchronous. */
int lym_unitary_projection(int ndim, int nc, int64_t volume, const void* in, void* out, void* stream);

/* ---- measurements -------------------------------------------------------
 * out[z] (host, L_dir doubles) = sum over the nvec complex128 spinors v
 * (op's layout) and the sites x with x_dir = z of
 *   sum_{s,c} |v_sc(x)|^2                    (fundamental)
 *   sum_{s,c,d} conj(v_scd(x)) v_sdc(x)      (adjoint matrices; real)
 * the per-slice sums pion_correlator forms from its propagators
 * (fermion_measurements.py:73-84).  scratch: device, >= (V + L_dir) doubles.
 * Deterministic (fixed summation order).  Synchronous. */
int lym_slice_sums(const lym_wilson* op, int dir, int nvec, const void* const* vecs, void* scratch,
                   size_t scratch_bytes, double* out, void* stream);

/* ---- kernel autotuning (no reference counterpart) ----------------------
 * The LDS-tiled Wilson kernel (4D fundamental lattices with L0 divisible
 * by 8) works on tiles of a0 x a1 x k x L3 sites with `hif` hops' loads in
 * flight per wave.  lym_tune_wilson times D (projected form) for every
 * admissible tile and hop grouping, and the untiled kernel, on the caller's
 * fields `in` / `out` (op's layout, precision `prec`), `reps` launches
 * each, and records the fastest for this lattice, Nc and precision in a
 * process-wide table that every later call on such an operator uses.
 * Synchronous.  lym_tile_info reports the choice in effect (tuned, forced
 * by LYM_TILE / LYM_HIF, or the built-in rule); a0 = 0 means untiled. */
typedef struct lym_tile_choice {
  int32_t a0, a1, k, hif;
  int32_t nsite;        /* sites (threads) per block */
  int32_t tuned;        /* 1: from lym_tune_wilson */
  int32_t walk;         /* order of an XCD's tiles: 0 = x0 fastest, then x2,
                           x1; 1 = x0, then x1, x2 (tuned with the shape) */
  int32_t reserved;
  double us;            /* measured time per launch (lym_tune_wilson only) */
} lym_tile_choice;
int lym_tune_wilson(const lym_wilson* op, int prec, int reps, const void* in, void* out, lym_tile_choice* best,
                    void* stream);
int lym_tile_info(const lym_wilson* op, int prec, lym_tile_choice* out);

/* ---- kernel timing (no reference counterpart; measurement only) --------
 * lym_kernel_timing(1) clears the counters and starts timing on the calling
 * host thread: every multishift CG solve with several shifts and
 * LYM_PREC_F32 / LYM_PREC_F16 (lym_cgm, lym_cgm_batched) then brackets each
 * of its Wilson, direction-update, shift-update and finalizer launches with
 * HIP events on the solve's stream (launching directly instead of replaying
 * HIP graphs) and, when the solve ends, adds per kernel class the number of
 * launches, the summed event durations and the summed algorithmic HBM bytes
 * (each link row read and each vector read or written once; a shift
 * update's bytes count its block's residuals and the shifts active in it).
 * lym_kernel_timing(0) stops timing.  lym_kernel_timing_read copies up to
 * `max` classes to `out` and the number of classes to *count. */
typedef struct lym_kernel_time {
  char name[48];
  int32_t launches;
  int32_t reserved;
  double total_us;
  double bytes;
} lym_kernel_time;
int lym_kernel_timing(int enable);
int lym_kernel_timing_read(lym_kernel_time* out, int max, int* count);

/* ---- solvers ---------------------------------------------------------- */
/* Device workspace (bytes) for the solvers below. */
size_t lym_solver_workspace_bytes(const lym_wilson* op, int prec, int nshift);

/* Multishift minimal-residual solver exactly as multishift_solver.solve
 * (dirac_solver.py:13-53) for A = Op^2 (SquareOperator): x_i = (A + s_i)^-1 b
 * for the non-negative shifts; base shift = shifts[nshift-1]; stops when
 * |<r,r>| < precision (absolute, iterated).  Leading negative shifts are
 * NOT handled here (the host binding routes them to BiCGStab as
 * dirac_solver.py:26-30 does); `first` is the index of the first shift
 * solved here.  F64 only.  Synchronous. */
int lym_mr_multishift(const lym_wilson* op, int nshift, const double* shifts, int first, const void* b,
                      void* const* xs, int max_steps, double precision, double omega, void* workspace,
                      size_t workspace_bytes, lym_solve_stats* stats, void* stream);

/* BiCGStab exactly as biconjugate_gradient.solve (dirac_solver.py:56-110) on
 * A = Op + shift (square = 0) or A = Op^2 + shift (square = 1), Op = D, or
 * Q when op->hermitian: x0 = initial guess (NULL: b), r = b - A x0, shadow
 * vector r_hat = x0 + A x0 (dirac_solver.py:67-70), alpha = omega = rho = 1,
 * p = nu = 0, the standard updates, stop when |<r,r>| < precision
 * (absolute, iterated, dirac_solver.py:96-103) or after max_steps.  The
 * reference's r = x0 - A x0 (correct only for x0 = b) is not reproduced.
 * Replaces the reference's solver for Propagator.apply (dirac.py:670-672) and
 * for the leading negative pole of the multishift solver
 * (dirac_solver.py:26-30).  F64; device-resident (no host round trip per
 * iteration); x must not alias b.  stats->iterations = iterations done.
 * Synchronous. */
size_t lym_bicgstab_workspace_bytes(const lym_wilson* op);
int lym_bicgstab(const lym_wilson* op, int square, double shift, const void* b, const void* x0, void* x,
                 int max_steps, double precision, void* workspace, size_t workspace_bytes, lym_solve_stats* stats,
                 void* stream);

/* Multishift conjugate gradient (Jegerlehner, hep-lat/9612014) on
 * A = Q^2 + s_i (requires op->hermitian): x_i = (Q^2 + s_i)^-1 b, all shifts
 * >= 0, any order (the base is the smallest).  Stops each shift when its
 * iterated relative residual <= tol; then recomputes the FP64 true relative
 * residual of every shift into stats->true_rel_resid.
 *   prec = LYM_PREC_F64: FP64 throughout.
 *   prec = LYM_PREC_F32, several shifts: FP64 Krylov vectors until the base
 *          residual is below 3e-4 |b|, complex64 after; shift solutions and
 *          directions accumulate in FP64 (deferred block updates over 64
 *          FP64 / 128 low-precision iterations, whose residuals the workspace
 *          keeps: about 1 KB per complex number of the field); FP64 defect
 *          correction of any shift whose true residual misses tol.
 *   prec = LYM_PREC_F16, several shifts, fundamental representation: as
 *          LYM_PREC_F32, then 16-bit storage of the Krylov vectors (int16
 *          fixed point scaled by a per-site float norm, 4 bytes per complex
 *          number; FP32 arithmetic) below 1e-6 |b|.
 *   prec = LYM_PREC_F32, one shift: complex64 CG with FP64 reliable updates
 *          (true residual recomputed in FP64 whenever the iterated one has
 *          dropped 100x, and at the end); the solution accumulates in FP64
 *          in blocks of 16 steps from the last 16 directions, which the
 *          workspace keeps (16 fields per source at the storage precision).
 *   prec = LYM_PREC_F16, one shift, fundamental representation: the same
 *          with 16-bit storage of r, p and Q p; arithmetic in FP32.
 * b and xs are complex128.  Synchronous.  New entry point (no counterpart in
 * the reference, whose multishift solver is the MR above); it serves
 * multishift_solver.solve / RationalApproximation.apply (dirac.py:52-59). */
int lym_cgm(const lym_wilson* op, int prec, int nshift, const double* shifts, const void* b, void* const* xs,
            double tol, int max_iter, void* workspace, size_t workspace_bytes, lym_solve_stats* stats,
            void* stream);

/* Multishift CG for nrhs right-hand sides sharing one shift list:
 * xs[r * nshift + s] = (Q^2 + shifts[s])^-1 bs[r].  LYM_PREC_F32 runs all RHS
 * in the same launches (one Wilson launch covers the batch, links are read
 * once per site block and reused from L2 by the other RHS); LYM_PREC_F64 is
 * the plain FP64 CG-M, one RHS after the other.  stats: nrhs entries or NULL.
 * Workspace: lym_cgm_batched_workspace_bytes.  Synchronous.  New entry point
 * (no counterpart in the reference, which solves one source per call,
 * e.g. fermion_measurements.py:57-68, 101-108). */
size_t lym_cgm_batched_workspace_bytes(const lym_wilson* op, int prec, int nshift, int nrhs);
int lym_cgm_batched(const lym_wilson* op, int prec, int nrhs, int nshift, const double* shifts, const void* const* bs,
                    void* const* xs, double tol, int max_iter, void* workspace, size_t workspace_bytes,
                    lym_solve_stats* stats, void* stream);

#ifdef __cplusplus
}
#endif
#endif /* LYM_H */
