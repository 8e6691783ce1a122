"""Torch-CPU complex128 restatement of the reference's Wilson operator and of
one multishift CG iteration (test and measurement infrastructure only: the
CPU baseline of the benchmark; nothing on the product path imports it).

The reference runs TensorFlow-eager ops on the CPU; TensorFlow is not
installed here, so the baseline restates the same op sequence with torch
tensor ops (which, like TF's, run multithreaded on the host cores):

    per direction mu (dirac.py:458-501, projected fundamental form):
      project   h+ = (1 - g_mu)[:Ns/2] psi,  h- = (1 + g_mu)[:Ns/2] psi
      gather    h+(x + mu), h-(x - mu)                (translate = tf.gather)
      colour    U_mu(x) h+(x+mu),  U_mu(x-mu)^dag h-(x-mu)        (einsum)
      rebuild   the lower spin rows from the upper ones (LU back-projector)
      add       out += kappa (hop+ + hop-)
    then g5 if hermitian (dirac.py:649-662).

The back-projector is built once (the reference re-runs scipy.linalg.lu on
every call, dirac.py:485-486; that host cost is left out).  Down links
U_mu(x - mu) are precomputed as the reference does (dirac.py:372-378).
Pinned against the NumPy oracle (oracle/wilson.py) in
tests/test_oracle_torch_cpu.py.
"""
import numpy as np
import torch

from . import lattice as olat
from . import wilson as ow


class TorchWilson:
    """Fundamental-representation Wilson D / Q = g5 D in complex128 on the CPU."""

    def __init__(self, links, geometry, kappa, hermitian=True):
        links = np.asarray(links, dtype=np.complex128)
        self.D = links.shape[0]
        self.nc = links.shape[1]
        self.kappa = float(kappa)
        self.hermitian = hermitian
        g, _ = ow.gammas(self.D)
        self.ns = g.shape[1]
        half = self.ns // 2
        eye = np.eye(self.ns)
        tables = olat.lookup_tables(geometry)
        self.sup = [torch.from_numpy(np.ascontiguousarray(tables[0][mu])).long() for mu in range(self.D)]
        self.sdn = [torch.from_numpy(np.ascontiguousarray(tables[1][mu])).long() for mu in range(self.D)]
        self.U = torch.from_numpy(links)
        self.Udn = torch.stack([self.U[mu][:, :, self.sdn[mu]] for mu in range(self.D)])
        self.proj_p, self.proj_m, self.back_p, self.back_m = [], [], [], []
        for mu in range(self.D):
            plus = eye - g[mu]
            minus = eye + g[mu]
            # full projected spinor = B @ (its upper Ns/2 rows): B = P pinv(P[:half])
            bp = plus @ np.linalg.pinv(plus[:half])
            bm = minus @ np.linalg.pinv(minus[:half])
            self.proj_p.append(torch.from_numpy(plus[:half].astype(np.complex128)))
            self.proj_m.append(torch.from_numpy(minus[:half].astype(np.complex128)))
            self.back_p.append(torch.from_numpy(bp.astype(np.complex128)))
            self.back_m.append(torch.from_numpy(bm.astype(np.complex128)))

    def apply(self, psi):
        out = psi.clone()
        for mu in range(self.D):
            hp = torch.einsum("mn,njc->mjc", self.proj_p[mu], psi)
            hp = torch.index_select(hp, 2, self.sup[mu])
            hp = torch.einsum("ijc,mjc->mic", self.U[mu], hp)
            hm = torch.einsum("mn,njc->mjc", self.proj_m[mu], psi)
            hm = torch.index_select(hm, 2, self.sdn[mu])
            hm = torch.einsum("jic,mjc->mic", self.Udn[mu].conj(), hm)
            hop = torch.einsum("mn,njc->mjc", self.back_p[mu], hp) + torch.einsum("mn,njc->mjc", self.back_m[mu], hm)
            out = out + self.kappa * hop
        if self.hermitian:
            out[self.ns // 2:] = -out[self.ns // 2:]
        return out

    def apply_square(self, psi):
        return self.apply(self.apply(psi))


def cgm_iterations(op, b, shifts, iters):
    """`iters` iterations of multishift CG (Jegerlehner's recurrences, the
    algorithm of the GPU solver) on Q^2 + shifts[s], every shift active:
    one Q^2 application, the base residual update and x_s, p_s updates of
    every shift per iteration.  Returns the base residual norm (so the work
    cannot be skipped)."""
    shifts = [float(s) for s in shifts]
    n = len(shifts)
    base = int(np.argmin(shifts))
    s0 = shifts[base]
    xs = [torch.zeros_like(b) for _ in shifts]
    ps = [b.clone() for _ in shifts]
    r = b.clone()
    zeta = np.ones(n)
    zeta_old = np.ones(n)
    alpha_old, beta_old = 1.0, 0.0
    rr = torch.vdot(r.flatten(), r.flatten()).real.item()
    for _ in range(iters):
        p = ps[base]
        ap = op.apply_square(p) + s0 * p
        alpha = rr / torch.vdot(p.flatten(), ap.flatten()).real.item()
        zn = np.ones(n)
        als = np.zeros(n)
        for s in range(n):
            den = alpha * beta_old * (zeta_old[s] - zeta[s]) + zeta_old[s] * alpha_old * (1.0 + (shifts[s] - s0) * alpha)
            zn[s] = zeta[s] * zeta_old[s] * alpha_old / den
            als[s] = alpha * zn[s] / zeta[s]
        r = r - alpha * ap
        rr_new = torch.vdot(r.flatten(), r.flatten()).real.item()
        beta = rr_new / rr
        for s in range(n):
            bs = beta * (zn[s] / zeta[s]) ** 2
            xs[s].add_(ps[s], alpha=als[s])
            ps[s].mul_(bs).add_(r, alpha=zn[s])
        zeta_old, zeta = zeta, zn
        alpha_old, beta_old, rr = alpha, beta, rr_new
    return float(np.sqrt(rr))
