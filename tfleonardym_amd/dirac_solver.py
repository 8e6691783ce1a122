"""Krylov solvers, drop-in for the reference's dirac_solver.py (13-110).

``multishift_solver(max_steps, precision, omega=0.95)`` keeps the reference's
signature and, by default, its algorithm: the over-relaxed multishift minimal
residual recurrence with base shift ``shifts[-1]`` and the stop rule
``|<r,r>| < precision`` (absolute, iterated); leading negative shifts go to
BiCGStab first (dirac_solver.py:26-30).  ``algorithm="cg"`` selects the
multishift conjugate gradient (CG-M) with a true relative residual target
``tolerance`` (default 1e-10), in FP64 or mixed precision.

When the operator is ``SquareOperator(DiracWilsonOperator)`` the whole solve
runs in liblym (device-resident scalars, no per-iteration host round trip).
Any other operator runs the same recurrences from Python on device vectors
(operator.apply + liblym BLAS-1).
"""
import ctypes

import numpy as np
import torch

from . import _lib
from . import algebra_utils
from . import dirac
from .tensors import to_device

_PREC = {"double": _lib.PREC_F64, "f64": _lib.PREC_F64, "mixed": _lib.PREC_F32, "f32": _lib.PREC_F32,
         "half": _lib.PREC_F16, "f16": _lib.PREC_F16}

_workspaces = {}


def _workspace(nbytes, device):
    """Cached device workspace (grows, never shrinks) per device and stream."""
    key = (device.index, torch.cuda.current_stream(device).cuda_stream)
    t = _workspaces.get(key)
    if t is None or t.numel() < nbytes:
        t = torch.zeros(int(nbytes) + 256, dtype=torch.uint8, device=device)
        _workspaces[key] = t
    return t


def _status_name(rc):
    return {0: "converged", 1: "max_iterations", 2: "breakdown"}.get(rc, str(rc))


class multishift_solver:
    def __init__(self, max_steps, precision, omega=0.95, algorithm="mr", tolerance=1e-10,
                 precision_mode="double", verbose=True):
        self.max_steps = max_steps
        self.precision = precision
        self.omega = omega
        if algorithm not in ("mr", "cg"):
            raise RuntimeError(f"unknown multishift algorithm {algorithm}")
        self.algorithm = algorithm
        self.tolerance = tolerance
        if precision_mode not in _PREC:
            raise RuntimeError(f"unknown precision mode {precision_mode}")
        self.precision_mode = precision_mode
        self.verbose = verbose
        self.last_stats = None
        self.last_status = None

    def solve(self, operator, source, shifts):
        shifts = [float(np.real(s)) for s in shifts]
        src = to_device(source)
        solutions = [torch.zeros_like(src) for _ in shifts]
        i0 = 0
        while i0 < len(shifts) and shifts[i0] < 0:
            solver = biconjugate_gradient(self.max_steps, self.precision, verbose=self.verbose)
            solutions[i0] = solver.solve(dirac.ShiftedOperator(operator, shifts[i0]), src)
            i0 += 1
        if i0 == len(shifts):
            return solutions
        w = dirac.wilson_of_square(operator)
        if self.algorithm == "mr":
            if w is not None:
                self._mr_device(w, src, shifts, i0, solutions)
            else:
                self._mr_generic(operator, src, shifts, i0, solutions)
        else:
            if any(s < 0 for s in shifts[i0:]):
                raise RuntimeError("CG-M needs non-negative shifts after the leading negative ones")
            if w is not None and w.hermitian:
                self._cgm_device(w, src, shifts, i0, solutions)
            else:
                self._cgm_generic(operator, src, shifts, i0, solutions)
        return solutions

    def solve_batch(self, operator, sources, shifts):
        """Multishift CG for several sources sharing one shift list, in the same
        device launches (``algorithm="cg"``; new, the reference solves one
        source per call).  Returns one list of solutions per source, in
        ``shifts`` order; ``last_stats`` becomes a list (one dict per source)."""
        w = dirac.wilson_of_square(operator)
        if w is None or not w.hermitian:
            return [self.solve(operator, s, shifts) for s in sources]
        shifts = [float(np.real(s)) for s in shifts]
        if any(s < 0 for s in shifts):
            raise RuntimeError("batched CG-M needs non-negative shifts")
        srcs = [to_device(s) for s in sources]
        nrhs, n = len(srcs), len(shifts)
        lib = _lib.lib()
        prec = _PREC[self.precision_mode]
        if prec != _lib.PREC_F64:
            w.ensure_f32()
        desc = w.desc
        nbytes = lib.lym_cgm_batched_workspace_bytes(ctypes.byref(desc), prec, n, nrhs)
        ws = _workspace(nbytes, srcs[0].device)
        sols = [[torch.empty_like(srcs[0]) for _ in shifts] for _ in srcs]
        sh = (ctypes.c_double * n)(*shifts)
        bs = (ctypes.c_void_p * nrhs)(*[s.data_ptr() for s in srcs])
        xs = (ctypes.c_void_p * (nrhs * n))(*[x.data_ptr() for row in sols for x in row])
        st = (_lib.LymSolveStats * nrhs)()
        rc = _lib.check(lib.lym_cgm_batched(ctypes.byref(desc), prec, nrhs, n, sh, bs, xs, float(self.tolerance),
                                            int(self.max_steps), _lib.ptr(ws), ws.numel(), st,
                                            _lib.stream_handle(srcs[0].device)), "cgm_batched")
        self.last_stats = [st[i].as_dict(n) for i in range(nrhs)]
        self.last_status = _status_name(rc)
        return sols

    # ---- device paths ------------------------------------------------------
    def _mr_device(self, w, src, shifts, i0, solutions):
        n = len(shifts)
        lib = _lib.lib()
        desc = w.desc
        nbytes = lib.lym_solver_workspace_bytes(ctypes.byref(desc), _lib.PREC_F64, n)
        ws = _workspace(nbytes, src.device)
        sh = (ctypes.c_double * n)(*shifts)
        xs = (ctypes.c_void_p * n)(*[t.data_ptr() for t in solutions])
        st = _lib.LymSolveStats()
        rc = _lib.check(lib.lym_mr_multishift(ctypes.byref(desc), n, sh, i0, _lib.ptr(src), xs, int(self.max_steps),
                                              float(self.precision), float(self.omega), _lib.ptr(ws), ws.numel(),
                                              ctypes.byref(st), _lib.stream_handle(src.device)), "mr_multishift")
        self.last_stats = st.as_dict(n)
        self.last_status = _status_name(rc)
        if self.verbose:
            if rc == 0:
                print(f"MMMRMultishiftsolver: convergence after {st.iterations - 1} steps")
            else:
                print(f"MMMRMultishiftsolver: failed to find convergence after {st.iterations - 1} steps")

    def _cgm_device(self, w, src, shifts, i0, solutions):
        sub = shifts[i0:]
        n = len(sub)
        lib = _lib.lib()
        prec = _PREC[self.precision_mode]
        if prec != _lib.PREC_F64:
            w.ensure_f32()
        desc = w.desc
        nbytes = lib.lym_solver_workspace_bytes(ctypes.byref(desc), prec, n)
        ws = _workspace(nbytes, src.device)
        sh = (ctypes.c_double * n)(*sub)
        xs = (ctypes.c_void_p * n)(*[t.data_ptr() for t in solutions[i0:]])
        st = _lib.LymSolveStats()
        rc = _lib.check(lib.lym_cgm(ctypes.byref(desc), prec, n, sh, _lib.ptr(src), xs, float(self.tolerance),
                                    int(self.max_steps), _lib.ptr(ws), ws.numel(), ctypes.byref(st),
                                    _lib.stream_handle(src.device)), "cgm")
        self.last_stats = st.as_dict(n)
        self.last_status = _status_name(rc)
        if self.verbose:
            print(f"MultishiftCG: {self.last_status} after {st.iterations} iterations, "
                  f"max true residual {max(self.last_stats['true_rel_resid']):.3e}")

    # ---- generic paths (any Operator; device vectors) -------------------------
    def _mr_generic(self, operator, src, shifts, i0, solutions):
        """dirac_solver.py:32-53 on device vectors."""
        n = len(shifts)
        f = [1.0 + 0j for _ in shifts]
        r = src.clone()
        base = shifts[-1]
        for step in range(self.max_steps):
            p = operator.apply(r)
            algebra_utils.axpby(base, r, 1.0, p)
            alpha = self.omega * algebra_utils.dot(p, r) / algebra_utils.dot(p, p)
            algebra_utils.axpby(alpha, r, 1.0, solutions[-1])
            for i in range(i0, n - 1):
                f[i] = f[i] / (1 + (shifts[i] - base) * alpha)
                algebra_utils.axpby(alpha * f[i], r, 1.0, solutions[i])
            algebra_utils.axpby(-alpha, p, 1.0, r)
            if abs(algebra_utils.dot(r, r)) < self.precision:
                self.last_status = "converged"
                if self.verbose:
                    print(f"MMMRMultishiftsolver: convergence after {step} steps")
                return
        self.last_status = "max_iterations"
        if self.verbose:
            print(f"MMMRMultishiftsolver: failed to find convergence after {step} steps")

    def _cgm_generic(self, operator, src, shifts, i0, solutions):
        """Multishift CG on an arbitrary Hermitian positive operator."""
        sub = shifts[i0:]
        n = len(sub)
        base = int(np.argmin(sub))
        s0 = sub[base]
        xs = solutions[i0:]
        ps = [src.clone() for _ in sub]
        r = src.clone()
        zeta = np.ones(n)
        zeta_old = np.ones(n)
        active = np.ones(n, dtype=bool)
        a_old, b_old = 1.0, 0.0
        rr = algebra_utils.dot(r, r).real
        bb = rr
        it = 0
        tol2 = self.tolerance ** 2
        while active.any() and it < self.max_steps:
            p = ps[base]
            Ap = operator.apply(p)
            algebra_utils.axpby(s0, p, 1.0, Ap)
            a = rr / algebra_utils.dot(p, Ap).real
            zn = np.ones(n)
            a_s = np.zeros(n)
            for s in range(n):
                if active[s]:
                    den = a * b_old * (zeta_old[s] - zeta[s]) + zeta_old[s] * a_old * (1 + (sub[s] - s0) * a)
                    zn[s] = zeta[s] * zeta_old[s] * a_old / den
                    a_s[s] = a * zn[s] / zeta[s]
            algebra_utils.axpby(-a, Ap, 1.0, r)
            rr_new = algebra_utils.dot(r, r).real
            b = rr_new / rr
            for s in range(n):
                if active[s]:
                    algebra_utils.axpby(a_s[s], ps[s], 1.0, xs[s])
                    algebra_utils.axpby(zn[s], r, b * (zn[s] / zeta[s]) ** 2, ps[s])
            zeta_old, zeta = zeta, np.where(active, zn, zeta)
            a_old, b_old, rr = a, b, rr_new
            it += 1
            for s in range(n):
                if s != base and active[s] and zeta[s] ** 2 * rr <= 0.25 * tol2 * bb:
                    active[s] = False
            others = np.delete(active, base)
            if active[base] and not others.any() and rr <= 0.25 * tol2 * bb:
                active[base] = False
        self.last_status = "converged" if not active.any() else "max_iterations"


def _wilson_form(operator):
    """(W, square, shift) when operator is W, SquareOperator(W) or either
    shifted by a real number, W a DiracWilsonOperator; else None."""
    shift = 0.0
    op = operator
    if isinstance(op, dirac.ShiftedOperator):
        sh = complex(op.shift)
        if sh.imag != 0.0:
            return None
        shift = sh.real
        op = op.operator_to_shift
    w = dirac.wilson_of_square(op)
    if w is not None:
        return w, 1, shift
    if isinstance(op, dirac.DiracWilsonOperator):
        return op, 0, shift
    return None


class biconjugate_gradient:
    """BiCGStab as in dirac_solver.py:56-110 (shadow vector r_hat = x0 + A x0,
    stop on |<r,r>| < precision).  On a Wilson operator, its square or either
    shifted (the reference's uses: Propagator, dirac.py:670-672, and the
    negative pole of the multishift solver, dirac_solver.py:26-30) the whole
    solve runs in liblym (lym_bicgstab: device-resident scalars, no host round
    trip per iteration); any other operator runs the same recurrence from
    Python on device vectors.  ``last_steps`` is the reference's step index."""

    def __init__(self, max_steps, precision, verbose=True):
        self.max_steps = max_steps
        self.precision = precision
        self.verbose = verbose
        self.last_status = None
        self.last_steps = None
        self.last_stats = None

    def solve(self, operator, source, initial_guess=None):
        b = to_device(source)
        form = _wilson_form(operator)
        if form is not None and b.dtype == torch.complex128:
            return self._solve_device(form, b, initial_guess)
        x = b.clone() if initial_guess is None else to_device(initial_guess).clone()
        init = operator.apply(x)
        r = algebra_utils.lincomb(1.0, b, -1.0, init)
        r_hat = algebra_utils.lincomb(1.0, x, 1.0, init)
        alpha = omega = rho = 1.0 + 0j
        p = torch.zeros_like(b)
        nu = torch.zeros_like(b)
        step = 0
        for step in range(self.max_steps):
            rho_next = algebra_utils.dot(r_hat, r)
            beta = (rho_next / rho) * (alpha / omega)
            algebra_utils.axpby(-omega, nu, 1.0, p)
            algebra_utils.axpby(1.0, r, beta, p)
            nu = operator.apply(p)
            alpha = rho_next / algebra_utils.dot(r_hat, nu)
            s = algebra_utils.lincomb(1.0, r, -alpha, nu)
            t = operator.apply(s)
            omega = algebra_utils.dot(t, s) / algebra_utils.dot(t, t)
            algebra_utils.axpby(alpha, p, 1.0, x)
            algebra_utils.axpby(omega, s, 1.0, x)
            r = algebra_utils.lincomb(1.0, s, -omega, t)
            if abs(algebra_utils.dot(r, r)) < self.precision:
                self.last_status, self.last_steps = "converged", step
                if self.verbose:
                    print(f"BiconjugateGradient: convergence after {step} steps")
                return x
            rho = rho_next
        self.last_status, self.last_steps = "max_iterations", step
        if self.verbose:
            print(f"BiconjugateGradient: failed to find convergence after {step} steps")
        return x

    def _solve_device(self, form, b, initial_guess):
        w, square, shift = form
        lib = _lib.lib()
        desc = w.desc
        x0 = None if initial_guess is None else to_device(initial_guess).to(torch.complex128).contiguous()
        nbytes = lib.lym_bicgstab_workspace_bytes(ctypes.byref(desc))
        if nbytes == 0:
            raise RuntimeError(lib.lym_last_error().decode())
        ws = _workspace(nbytes, b.device)
        x = torch.empty_like(b)
        st = _lib.LymSolveStats()
        rc = _lib.check(lib.lym_bicgstab(ctypes.byref(desc), int(square), float(shift), _lib.ptr(b),
                                         _lib.ptr(x0) if x0 is not None else None, _lib.ptr(x), int(self.max_steps),
                                         float(self.precision), _lib.ptr(ws), ws.numel(), ctypes.byref(st),
                                         _lib.stream_handle(b.device)), "bicgstab")
        self.last_stats = st.as_dict(1)
        step = max(0, int(st.iterations) - 1)
        self.last_steps = step
        self.last_status = _status_name(rc)
        if self.verbose:
            if rc == 0:
                print(f"BiconjugateGradient: convergence after {step} steps")
            else:
                print(f"BiconjugateGradient: failed to find convergence after {step} steps")
        return x


class conjugate_gradient:
    """CG on a Hermitian positive operator, true relative residual target
    ``tolerance``.  ``operator`` is SquareOperator(W) or
    ShiftedOperator(SquareOperator(W), s) with W hermitian (device path), or
    any Hermitian positive Operator (generic path).  New: the reference's
    single-shift solver is BiCGStab."""

    def __init__(self, max_steps, tolerance=1e-10, precision_mode="double", verbose=False):
        self.max_steps = max_steps
        self.tolerance = tolerance
        self.precision_mode = precision_mode
        self.verbose = verbose
        self.last_stats = None

    def solve(self, operator, source, initial_guess=None):
        if initial_guess is not None:
            raise RuntimeError("conjugate_gradient starts from zero")
        shift = 0.0
        op = operator
        if isinstance(op, dirac.ShiftedOperator):
            shift = float(op.shift)
            op = op.operator_to_shift
        ms = multishift_solver(self.max_steps, 0.0, algorithm="cg", tolerance=self.tolerance,
                               precision_mode=self.precision_mode, verbose=self.verbose)
        x = ms.solve(op, source, [shift])[0]
        self.last_stats = ms.last_stats
        return x
