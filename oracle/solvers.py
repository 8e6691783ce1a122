"""Oracle: Krylov solvers (test infrastructure only).

``mr_multishift`` and ``bicgstab`` restate dirac_solver.py:13-110 step by
step (complex128 scalars, ``dot(a, b) = sum conj(a) b`` as algebra_utils.py:5-11,
stopping on the absolute squared iterated residual ``|<r,r>| < precision``).
``cg`` and ``cgm`` are the textbook conjugate gradient and its multi-shift
form (Jegerlehner, hep-lat/9612014) used for the 1e-10 true-residual solves.

``apply`` is any callable v -> A v on NumPy arrays.
"""
import numpy as np


def dot(a, b):
    """sum conj(a) * b over all indices (algebra_utils.py:5-11)."""
    if a.shape != b.shape:
        raise RuntimeError("Incompatible vector shapes!")
    return np.vdot(a, b)


def bicgstab(apply, source, max_steps, precision, initial_guess=None):
    """dirac_solver.py:56-110 with its shadow vector r_hat = x0 + A x0.

    The initial residual is b - A x0 (the reference writes x0 - A x0, which is
    the same thing for its default x0 = b)."""
    x = np.array(source if initial_guess is None else initial_guess, dtype=np.complex128, copy=True)
    init = apply(x)
    r = source - init
    r_hat = x + init
    alpha = 1.0 + 0j
    omega = 1.0 + 0j
    rho = 1.0 + 0j
    p = np.zeros_like(source)
    nu = np.zeros_like(source)
    steps = max_steps
    for step in range(max_steps):
        rho_next = dot(r_hat, r)
        beta = (rho_next / rho) * (alpha / omega)
        p = r + beta * (p - omega * nu)
        nu = apply(p)
        alpha = rho_next / dot(r_hat, nu)
        s = r - alpha * nu
        t = apply(s)
        omega = dot(t, s) / dot(t, t)
        x = x + alpha * p + omega * s
        r = s - omega * t
        if abs(dot(r, r)) < precision:
            steps = step
            return x, step, True
        rho = rho_next
    return x, steps - 1, False


def mr_multishift(apply, source, shifts, max_steps, precision, omega=0.95):
    """Multishift minimal residual with over-relaxation (dirac_solver.py:13-53).

    Base shift is ``shifts[-1]``; leading negative shifts are solved by
    BiCGStab on A + shift (dirac_solver.py:26-30).  Returns (solutions, step,
    converged)."""
    n = len(shifts)
    xs = [np.zeros_like(source) for _ in shifts]
    f = [1.0 + 0j for _ in shifts]
    r = np.array(source, dtype=np.complex128, copy=True)
    i0 = 0
    while shifts[i0] < 0:
        sh = shifts[i0]
        xs[i0], _, _ = bicgstab(lambda v: apply(v) + sh * v, source, max_steps, precision)
        i0 += 1
    base = shifts[-1]
    for step in range(max_steps):
        p = apply(r) + base * r
        alpha = omega * dot(p, r) / dot(p, p)
        xs[-1] = xs[-1] + alpha * r
        for i in range(i0, n - 1):
            f[i] = f[i] / (1 + (shifts[i] - base) * alpha)
            xs[i] = xs[i] + alpha * f[i] * r
        r = r - alpha * p
        if abs(dot(r, r)) < precision:
            return xs, step, True
    return xs, max_steps - 1, False


def cg(apply, b, shift=0.0, tol=1e-10, max_iter=100000, x0=None):
    """Conjugate gradient on (A + shift) x = b; stops on ||r||/||b|| <= tol
    (iterated residual).  Returns (x, iterations)."""
    x = np.zeros_like(b) if x0 is None else np.array(x0, copy=True)
    r = b - (apply(x) + shift * x) if x0 is not None else np.array(b, copy=True)
    p = r.copy()
    rr = np.vdot(r, r).real
    bb = np.vdot(b, b).real
    it = 0
    while rr > tol * tol * bb and it < max_iter:
        Ap = apply(p) + shift * p
        a = rr / np.vdot(p, Ap).real
        x = x + a * p
        r = r - a * Ap
        rr_new = np.vdot(r, r).real
        p = r + (rr_new / rr) * p
        rr = rr_new
        it += 1
    return x, it


def cgm(apply, b, shifts, tol=1e-10, max_iter=100000):
    """Multi-shift CG: solves (A + shifts[s]) x_s = b for all s at once.

    Base system = smallest shift.  Shifted residuals are zeta_s * r; a shift
    stops updating when |zeta_s| ||r|| <= tol ||b||.  Returns (xs, iterations).
    """
    shifts = [float(s) for s in shifts]
    n = len(shifts)
    base = int(np.argmin(shifts))
    s0 = shifts[base]
    delta = [s - s0 for s in shifts]
    xs = [np.zeros_like(b) for _ in shifts]
    ps = [np.array(b, copy=True) for _ in shifts]
    r = np.array(b, copy=True)
    zeta = np.ones(n)
    zeta_old = np.ones(n)
    active = np.ones(n, dtype=bool)
    alpha_old = 1.0
    beta_old = 0.0
    rr = np.vdot(r, r).real
    bb = rr
    it = 0
    while active.any() and it < max_iter:
        p = ps[base]
        Ap = apply(p) + s0 * p
        alpha = rr / np.vdot(p, Ap).real
        zeta_new = np.ones(n)
        alpha_s = np.zeros(n)
        for s in range(n):
            if not active[s]:
                continue
            den = (alpha * beta_old * (zeta_old[s] - zeta[s])
                   + zeta_old[s] * alpha_old * (1.0 + delta[s] * alpha))
            zeta_new[s] = zeta[s] * zeta_old[s] * alpha_old / den
            alpha_s[s] = alpha * zeta_new[s] / zeta[s]
        r = r - alpha * Ap
        rr_new = np.vdot(r, r).real
        beta = rr_new / rr
        for s in range(n):
            if not active[s]:
                continue
            beta_s = beta * (zeta_new[s] / zeta[s]) ** 2
            xs[s] = xs[s] + alpha_s[s] * ps[s]
            ps[s] = zeta_new[s] * r + beta_s * ps[s]
        zeta_old, zeta = zeta, zeta_new
        alpha_old, beta_old, rr = alpha, beta, rr_new
        it += 1
        for s in range(n):
            if active[s] and zeta[s] ** 2 * rr <= tol * tol * bb:
                active[s] = False
    return xs, it


def true_residual(apply, b, x, shift):
    """||b - (A + shift) x|| / ||b|| in float64."""
    r = b - (apply(x) + shift * x)
    return float(np.linalg.norm(r) / np.linalg.norm(b))
