"""Wilson-Dirac operator and operator algebra, drop-in for the reference's
dirac.py (dirac.py:10-73, 253-378, 380-523, 649-662).

Same class names, constructor signatures, attributes and mode methods; fields
are torch tensors on the HIP device in the reference layout (spinors
(Ns, Nc, V), adjoint spinors (Ns, Nc, Nc, V), links (D, Nc, Nc, V),
complex128).  Every application runs the liblym Wilson kernel; nothing here
computes the operator on the host.
"""
import ctypes

import numpy as np
import torch

from . import _lib
from . import algebra_utils
from . import boundary_conditions as bcmod
from . import lie_generators
from .lattice import Configuration
from .tensors import to_device

_GAMMA2 = np.array([[[0, 1], [1, 0]], [[0, -1j], [1j, 0]]], dtype=np.complex128)
_GAMMA4 = np.array([[[0, 0, 0, -1j], [0, 0, -1j, 0], [0, 1j, 0, 0], [1j, 0, 0, 0]],
                    [[0, 0, 0, -1], [0, 0, 1, 0], [0, 1, 0, 0], [-1, 0, 0, 0]],
                    [[0, 0, -1j, 0], [0, 0, 0, 1j], [1j, 0, 0, 0], [0, -1j, 0, 0]],
                    [[0, 0, 1, 0], [0, 0, 0, 1], [1, 0, 0, 0], [0, 1, 0, 0]]], dtype=np.complex128)


class Operator:
    """Operator protocol: .apply(v), .set_gauge_configuration(cfg), .hermitian
    (dirac.py:10-12)."""

    def __init__(self, hermitian=False):
        self.hermitian = hermitian


class ShiftedOperator(Operator):
    """A + shift (dirac.py:15-25)."""

    def __init__(self, operator_to_shift, shift):
        super().__init__(operator_to_shift.hermitian)
        self.operator_to_shift = operator_to_shift
        self.shift = shift

    def set_gauge_configuration(self, cfg):
        self.operator_to_shift.set_gauge_configuration(cfg)

    def apply(self, vector):
        w = wilson_of_square(self.operator_to_shift)
        if w is not None:
            return w.apply_square(vector, self.shift)
        out = self.operator_to_shift.apply(vector)
        v = to_device(vector, device=out.device)
        return algebra_utils.axpby(self.shift, v, 1.0, out)


class SquareOperator(Operator):
    """A = Op(Op(v)) (dirac.py:28-37); D^dag D when Op = Q = g5 D."""

    def __init__(self, operator_to_square):
        super().__init__(operator_to_square.hermitian)
        self.operator_to_square = operator_to_square

    def set_gauge_configuration(self, cfg):
        self.operator_to_square.set_gauge_configuration(cfg)

    def apply(self, vector):
        if isinstance(self.operator_to_square, DiracWilsonOperator):
            return self.operator_to_square.apply_square(vector, 0.0)
        return self.operator_to_square.apply(self.operator_to_square.apply(vector))


def wilson_of_square(op):
    """The DiracWilsonOperator W if op is SquareOperator(W), else None."""
    if isinstance(op, SquareOperator) and isinstance(op.operator_to_square, DiracWilsonOperator):
        return op.operator_to_square
    return None


class RationalApproximation(Operator):
    """r(A) v = shift v + sum_i alpha_i (A + beta_i)^-1 v through one multishift
    solve (dirac.py:40-73)."""

    def __init__(self, operator, alphas, betas, shift, solver):
        super().__init__(operator.hermitian)
        self.operator = operator
        self.alphas = alphas
        self.betas = betas
        self.shift = shift
        self.solver = solver

    def set_gauge_configuration(self, cfg):
        self.operator.set_gauge_configuration(cfg)

    def apply(self, vector):
        solutions = self.solver.solve(self.operator, vector, self.betas)
        v = to_device(vector)
        result = torch.zeros_like(v)
        algebra_utils.axpby(self.shift, v, 0.0, result)
        for solution, alpha in zip(solutions, self.alphas):
            algebra_utils.axpby(complex(alpha), solution, 1.0, result)
        return result

    def evaluate(self, x):
        result = self.shift
        for alpha, beta in zip(self.alphas, self.betas):
            result += alpha / (x + beta)
        return result

    def test_inverter(self, vector):
        """Prints <A x_i - v, A x_i - v> for every pole as the reference does
        (it omits beta_i, so it is zero only for beta_i = 0); returns them."""
        solutions = self.solver.solve(self.operator, vector, self.betas)
        v = to_device(vector)
        out = []
        for solution in solutions:
            d = self.operator.apply(solution)
            algebra_utils.axpby(-1.0, v, 1.0, d)
            val = algebra_utils.dot(d, d)
            print("It must be zero", val)
            out.append(val)
        return out


class PolynomialApproximation(Operator):
    """prod_i scaling (A - root_i) v (dirac.py:76-94), on the device kernels."""

    def __init__(self, operator, roots, scaling):
        super().__init__(operator.hermitian)
        self.roots = roots
        self.scaling = scaling
        self.operator = operator

    def apply(self, vector):
        output = to_device(vector).clone()
        for root in self.roots:
            a = self.operator.apply(output)
            algebra_utils.axpby(-complex(root) * self.scaling, output, complex(self.scaling), a)
            output = a
        return output

    def set_gauge_configuration(self, cfg):
        self.operator.set_gauge_configuration(cfg)

    def order(self):
        return len(self.roots)


class Overlap(Operator):
    """Overlap operator built on a Wilson operator and a polynomial
    approximation P of (Op^2)^-1/2 (dirac.py:97-172):
        hermitian:     O v = (1+m)/2 g5 v + (1-m)/2 Op P(Op^2) v
        non-hermitian: O v = (1+m)/2 v + (1-m)/2 g5 Op P(Op^2) v
    The polynomial and Op run in the device kernels; the final per-spin
    combination is one lym_spin_axpby launch."""

    def __init__(self, dirac_operator, square_root_approximation, mass, hermitian=False):
        super().__init__(hermitian)
        self.square_root_approximation = square_root_approximation
        self.dirac_operator = dirac_operator
        self.mass = mass
        self._copy_shape()
        self.representation = self.dirac_operator.representation

    def _copy_shape(self):
        self.number_of_dimensions = self.dirac_operator.number_of_dimensions
        self.number_of_colors = self.dirac_operator.number_of_colors
        self.volume = self.dirac_operator.volume
        self.spinor_dimension = self.dirac_operator.spinor_dimension

    def set_gauge_configuration(self, cfg):
        self.square_root_approximation.set_gauge_configuration(cfg)
        self.dirac_operator.set_gauge_configuration(cfg)
        self._copy_shape()

    def apply(self, vector):
        v = to_device(vector)
        output = self.dirac_operator.apply(self.square_root_approximation.apply(v))
        a = (1 + self.mass) / 2
        b = (1 - self.mass) / 2
        if self.hermitian:
            coef = (a, -a, b, b)      # upper: a v + b out, lower: -a v + b out
        else:
            coef = (a, a, b, -b)      # upper: a v + b out, lower: a v - b out
        c = [complex(z) for z in coef]
        arr = (ctypes.c_double * 8)(*[p for z in c for p in (z.real, z.imag)])
        res = torch.empty_like(v)
        ns = v.shape[0]
        _lib.check(_lib.lib().lym_spin_axpby(ns, v.numel() // ns, arr, _lib.ptr(v), _lib.ptr(output), _lib.ptr(res),
                                             _lib.stream_handle(v.device)), "spin_axpby")
        return res

    def _poly_step(self, v, root):
        """scaling (A v - root v) with A = the polynomial's operator."""
        s = complex(self.square_root_approximation.scaling)
        out = self.square_root_approximation.operator.apply(v)
        return algebra_utils.axpby(-s * complex(root), v, s, out)

    def lie_derivative(self, X, Y):
        """dirac.py:143-172: (1-m)/2 [L(X, P Y) + s sum_i (L(Op l_i, r_{i+1})
        + L(l_i, Op r_{i+1}))], l_0 = Op X, l_{i+1} = s (A - conj(root_i)) l_i,
        r_i = the polynomial factors after root i applied to Y; all terms in
        one force sum."""
        roots = list(self.square_root_approximation.roots)
        scaling = complex(self.square_root_approximation.scaling)
        factor = 0.5 - self.mass / 2.
        x = to_device(X)
        right = [to_device(Y, device=x.device)]
        for root in reversed(roots):
            right.append(self._poly_step(right[-1], root))
        right = list(reversed(right))
        left = [self.dirac_operator.apply(x)]
        for root in roots:
            left.append(self._poly_step(left[-1], np.conj(complex(root))))
        pairs = [(x, right[0])]
        weights = [factor]
        D = self.dirac_operator
        for i in range(len(roots)):
            pairs.append((D.apply(left[i]), right[i + 1]))
            pairs.append((left[i], D.apply(right[i + 1])))
            weights += [scaling * factor, scaling * factor]
        if isinstance(D, DiracWilsonOperator):
            return D.force_sum(pairs, weights, symmetric=False)
        force = None
        for (a, b), w in zip(pairs, weights):
            term = D.lie_derivative(a, b)
            if force is None:
                # x and y of lym_axpby must not alias: scale into a fresh tensor
                force = algebra_utils.axpby(w, term, 0.0, torch.zeros_like(term))
            else:
                force = algebra_utils.axpby(w, term, 1.0, force)
        return force


class DiracWilsonOperator(Operator):
    """Wilson-Dirac operator on SU(N) links in 2 or 4 dimensions
    (dirac.py:253-662).  ``apply`` is D, or Q = g5 D when hermitian."""

    def __init__(self, cfg, kappa, hermitian=False, mode="projected", representation="adjoint",
                 boundary_condition_provider=None):
        super().__init__(hermitian)
        self.kappa = kappa
        self.mode = mode
        self.representation = representation
        if mode not in ("projected", "simple"):
            raise RuntimeError(f"Unknown dirac multiplication mode {mode}")
        if representation not in ("fundamental", "adjoint"):
            raise RuntimeError(f"Unknown representation {representation}")
        if (mode, representation) == ("projected", "adjoint"):
            self.apply = self.multiply_in_adjoint_representation
        if (mode, representation) == ("simple", "adjoint"):
            self.apply = self.simple_multiply_in_adjoint_representation
        if (mode, representation) == ("projected", "fundamental"):
            self.apply = self.multiply_in_fundamental_representation
        if (mode, representation) == ("simple", "fundamental"):
            self.apply = self.simple_multiply_in_fundamental_representation
        self.boundary_condition_provider = boundary_condition_provider
        self.set_gauge_configuration(cfg)
        self.set_gamma_matrices()

    # ---- set-up -----------------------------------------------------------
    def set_gamma_matrices(self):
        """Chiral gammas, g5, projectors and LU-derived back-projectors as host
        arrays (dirac.py:283-349); the kernels hard-code the same basis."""
        from scipy.linalg import lu
        if self.number_of_dimensions == 2:
            self.gamma = _GAMMA2.copy()
            self.gammad_diagonal = np.array([1, -1], dtype=np.complex128)
            self.spinor_dimension = 2
        elif self.number_of_dimensions == 4:
            self.gamma = _GAMMA4.copy()
            self.gammad_diagonal = np.array([1, 1, -1, -1], dtype=np.complex128)
            self.spinor_dimension = 4
        else:
            raise RuntimeError(f"Number of dimensions {self.number_of_dimensions} not supported for fermions!")
        ns = self.spinor_dimension
        self.identity = np.eye(ns, dtype=np.complex128)
        self.gammad = np.diag(self.gammad_diagonal)
        self.plus_projectors, self.minus_projectors = [], []
        self.back_plus_projectors, self.back_minus_projectors = [], []
        for mu in range(self.number_of_dimensions):
            pp = self.identity - self.gamma[mu]
            pm = self.identity + self.gamma[mu]
            if self.hermitian:
                pp = self.gammad @ pp
                pm = self.gammad @ pm
            self.plus_projectors.append(pp)
            self.minus_projectors.append(pm)
            for P, dest in ((pp, self.back_plus_projectors), (pm, self.back_minus_projectors)):
                pl, _ = lu(P, permute_l=True)
                b = np.zeros((ns, ns // 2), dtype=np.complex128)
                for n in range(ns // 2):
                    b[n, n] = self.kappa
                b[ns // 2:] = self.kappa * pl[ns // 2:, :ns // 2]
                dest.append(b)
        self.plus_projectors = np.array(self.plus_projectors)
        self.minus_projectors = np.array(self.minus_projectors)
        self.back_plus_projectors = np.array(self.back_plus_projectors)
        self.back_minus_projectors = np.array(self.back_minus_projectors)

    def set_gauge_configuration(self, cfg):
        """Upload the links (dirac.py:351-378).  The standard antiperiodic
        provider becomes an in-kernel boundary factor; any other provider is
        applied once here and its result kept resident."""
        gf = to_device(cfg.gauge_field)
        if gf.dim() != 4 or gf.shape[1] != gf.shape[2]:
            raise RuntimeError("gauge field must have shape (D, Nc, Nc, V)")
        self.number_of_dimensions = int(gf.shape[0])
        self.number_of_colors = int(gf.shape[1])
        self.volume = int(gf.shape[-1])
        if self.number_of_dimensions not in (2, 4):
            raise RuntimeError(f"Number of dimensions {self.number_of_dimensions} not supported for fermions!")
        geometry = tuple(int(L) for L in cfg.geometry)
        if len(geometry) != self.number_of_dimensions or int(np.prod(geometry)) != self.volume:
            raise RuntimeError(f"geometry {geometry} does not match gauge field shape {tuple(gf.shape)}")
        self.geometry = geometry
        self.original_cfg = cfg
        self._links = gf
        self._links_right = None
        self._bc = 1.0 + 0j
        prov = self.boundary_condition_provider
        if prov is None:
            pass
        elif prov is bcmod.apply_antiboundary_conditions_in_t_direction:
            g, factor = bcmod.current_factor()
            if g != geometry:
                raise RuntimeError(f"boundary field was set for geometry {g}, operator has {geometry}")
            self._bc = factor
        else:
            signed = to_device(prov(gf))
            if tuple(signed.shape) != tuple(gf.shape):
                raise RuntimeError("boundary_condition_provider changed the gauge field shape")
            self._links = signed
            if self.representation == "adjoint":
                self._links_right = gf
        self._links32 = None
        self._links_right32 = None
        self._adjoint_links = None
        self._desc_fund_adj = None
        self._desc = self._make_desc(self._links, self._links_right, self.number_of_colors,
                                     _lib.REP_ADJOINT if self.representation == "adjoint" else _lib.REP_FUNDAMENTAL)

    def _make_desc(self, links, links_right, nc, rep):
        d = _lib.LymWilson()
        d.ndim = self.number_of_dimensions
        d.nc = nc
        d.representation = rep
        d.hermitian = 1 if self.hermitian else 0
        for mu, L in enumerate(self.geometry):
            d.dims[mu] = L
        d.kappa = float(np.real(self.kappa))
        if np.imag(self.kappa) != 0:
            raise RuntimeError("complex kappa is not supported")
        d.bc_re = float(self._bc.real)
        d.bc_im = float(self._bc.imag)
        d.links = links.data_ptr()
        d.links_right = links_right.data_ptr() if links_right is not None else None
        # special-unitary links: the kernels read nc-1 rows and rebuild the last
        if rep == _lib.REP_FUNDAMENTAL:
            d.link_format = _lib.su_link_format(d, links.device)
        return d

    def autotune(self, reps=20, precisions=("f64", "f32")):
        """Time every LDS tile shape and hop grouping of the Wilson kernel (and
        the untiled kernel) for this operator's lattice and keep the fastest
        for the rest of the process (lym_tune_wilson).  Returns the choices.
        Later solves use the tuned kernel; results stay within rounding of
        the untuned ones (block sums of the reductions regroup)."""
        desc = self.ensure_f32()
        shape = self.spinor_shape
        g = torch.Generator(device="cpu").manual_seed(12345)
        out = {}
        for name in precisions:
            prec, dt = (_lib.PREC_F64, torch.complex128) if name == "f64" else (_lib.PREC_F32, torch.complex64)
            x = torch.randn(shape, dtype=dt, generator=g).to(self._links.device)
            y = torch.empty_like(x)
            choice = _lib.LymTileChoice()
            _lib.check(_lib.lib().lym_tune_wilson(ctypes.byref(desc), prec, int(reps), _lib.ptr(x), _lib.ptr(y),
                                                  ctypes.byref(choice), _lib.stream_handle(self._links.device)),
                       "autotune")
            out[name] = choice.as_dict()
        return out

    def tile_info(self, precision="f32"):
        """The Wilson-kernel tile in effect for this operator (a0 = 0: untiled)."""
        prec = {"f64": _lib.PREC_F64, "f32": _lib.PREC_F32, "f16": _lib.PREC_F16}[precision]
        choice = _lib.LymTileChoice()
        _lib.check(_lib.lib().lym_tile_info(ctypes.byref(self._desc), prec, ctypes.byref(choice)), "tile info")
        return choice.as_dict()

    def ensure_f32(self):
        """complex64 copy of the links for the mixed-precision solvers."""
        if self._links32 is None:
            self._links32 = _lib.pack_links_f32(self._links)
            self._desc.links_f32 = self._links32.data_ptr()
            if self._links_right is not None:
                self._links_right32 = _lib.pack_links_f32(self._links_right)
                self._desc.links_right_f32 = self._links_right32.data_ptr()
        return self._desc

    @property
    def desc(self):
        return self._desc

    @property
    def cfg(self):
        """The configuration after the boundary provider (dirac.py:366-370)."""
        if self.boundary_condition_provider is None:
            return self.original_cfg
        return self.boundary_condition_provider(self.original_cfg)

    @property
    def fermion_gauge_field(self):
        """U (fundamental) or V_ab = 2 tr(T^a U T^b U^dag) (adjoint), after the
        boundary provider (dirac.py:356-370)."""
        if self.representation == "fundamental":
            return self.cfg.gauge_field
        return self.adjoint_links(with_bc=True)

    def adjoint_links(self, with_bc=True):
        if self._adjoint_links is None:
            T = torch.from_numpy(lie_generators.generators(self.number_of_colors, "fundamental")).to(self._links.device)
            U = to_device(self.original_cfg.gauge_field)
            V = 2.0 * torch.einsum("ijk,mklc,nlo,mjoc->minc", T, U, T, torch.conj(U))
            self._adjoint_links = V.real.to(torch.complex128).contiguous()
        V = self._adjoint_links
        if with_bc and self.boundary_condition_provider is not None:
            return to_device(self.boundary_condition_provider(V))
        return V

    @property
    def spinor_shape(self):
        ns = 2 if self.number_of_dimensions == 2 else 4
        nc = self.number_of_colors
        if self.representation == "adjoint":
            return (ns, nc, nc, self.volume)
        return (ns, nc, self.volume)

    # ---- application --------------------------------------------------------
    def _run(self, desc, v, mode, shape):
        x = to_device(v)
        if tuple(x.shape) != tuple(shape):
            raise RuntimeError(f"vector shape {tuple(x.shape)} does not match operator shape {tuple(shape)}")
        out = torch.empty_like(x)
        _lib.check(_lib.lib().lym_wilson_apply(ctypes.byref(desc), _lib.PREC_F64, mode, _lib.ptr(x),
                                               _lib.ptr(out), _lib.stream_handle(x.device)), "wilson_apply")
        return out

    def _fund_desc(self):
        """Descriptor of the fundamental-form kernel: U, or for an adjoint
        operator the real adjoint links V_ab acting on x^a (tests.py:85-90)."""
        if self.representation == "fundamental":
            return self._desc, self.number_of_colors
        if self._desc_fund_adj is None:
            # the standard antiperiodic provider stays an in-kernel factor;
            # any other provider is applied to V here
            custom = self.boundary_condition_provider is not None and self._bc == 1
            V = self.adjoint_links(with_bc=custom)
            n = V.shape[1]
            if n not in (2, 3):
                raise RuntimeError(f"fundamental-form kernel with {n}x{n} adjoint links is not built (Nc in 2, 3)")
            self._adj_v = V
            self._desc_fund_adj = self._make_desc(V, None, n, _lib.REP_FUNDAMENTAL)
        return self._desc_fund_adj, self._adj_v.shape[1]

    def multiply_in_fundamental_representation(self, input_vector):
        """dirac.py:458-501 (half-spinor projection)."""
        d, n = self._fund_desc()
        ns = 2 if self.number_of_dimensions == 2 else 4
        return self._run(d, input_vector, _lib.MODE_PROJECTED, (ns, n, self.volume))

    def simple_multiply_in_fundamental_representation(self, input_vector):
        """dirac.py:503-523 (full projectors)."""
        d, n = self._fund_desc()
        ns = 2 if self.number_of_dimensions == 2 else 4
        return self._run(d, input_vector, _lib.MODE_SIMPLE, (ns, n, self.volume))

    def multiply_in_adjoint_representation(self, input_vector):
        """dirac.py:404-456: hop = U X U^dag on matrix-valued spinors."""
        self._need_adjoint()
        return self._run(self._desc, input_vector, _lib.MODE_PROJECTED, self.spinor_shape)

    def simple_multiply_in_adjoint_representation(self, input_vector):
        """dirac.py:380-402."""
        self._need_adjoint()
        return self._run(self._desc, input_vector, _lib.MODE_SIMPLE, self.spinor_shape)

    def _need_adjoint(self):
        if self.representation != "adjoint":
            raise RuntimeError("adjoint multiplication on an operator built in the fundamental representation")

    def apply_hermiticity(self, input_vector):
        """g5 when hermitian, else a copy (dirac.py:649-662)."""
        x = to_device(input_vector)
        out = torch.empty_like(x)
        if not self.hermitian:
            out.copy_(x)
            return out
        ns = x.shape[0]
        _lib.check(_lib.lib().lym_gamma5(_lib.PREC_F64, ns, x.numel() // ns, _lib.ptr(x), _lib.ptr(out),
                                         _lib.stream_handle(x.device)), "gamma5")
        return out

    # ---- fermion force -------------------------------------------------------
    def lie_derivative(self, X, Y):
        """dirac.py:525-565: the (D, Nc, Nc, V) field kappa sum_k T^k l_k(X, Y),
        the derivative of X^dag Op Y along U_mu(x) -> exp(i eps T^k) U_mu(x)
        (the X^dag term; n_flavor symmetrises).  Adjoint spinors are the
        (Ns, Nc, Nc, V) matrices; the result is in fundamental generators."""
        return self.force_sum([(X, Y)], [1.0], symmetric=False)

    def force_sum(self, pairs, weights, symmetric=True, out=None):
        """sum_p w_p [L(x_p, y_p) + L(y_p, x_p)] (symmetric) or sum_p w_p L(x_p, y_p)
        over (x_p, y_p) in `pairs`, in one force launch per 32 pairs
        (lym_lie_derivative; complex weights: one more for the imaginary
        parts); added to `out` when given, else a new tensor.  This is the
        sum n_flavor.force forms (fermion_action.py:86-91)."""
        shape = tuple(self.spinor_shape)
        xs, ys = [], []
        for x, y in pairs:
            xd = to_device(x)
            yd = to_device(y, device=xd.device)
            if tuple(xd.shape) != shape or tuple(yd.shape) != shape:
                raise RuntimeError(f"force vectors must have shape {shape}, got {tuple(xd.shape)}, {tuple(yd.shape)}")
            xs.append(xd)
            ys.append(yd)
        if len(weights) != len(xs):
            raise RuntimeError("one weight per pair")
        cw = [complex(a) for a in weights]
        if any(c.imag != 0 for c in cw):
            # L is real-linear: sum w L = sum Re(w) L + i sum Im(w) L
            re = self.force_sum(list(zip(xs, ys)), [c.real for c in cw], symmetric, out)
            im = self.force_sum(list(zip(xs, ys)), [c.imag for c in cw], symmetric)
            return algebra_utils.axpby(1j, im, 1.0, re)
        w = [c.real for c in cw]
        fshape = (self.number_of_dimensions, self.number_of_colors, self.number_of_colors, self.volume)
        dev = self._links.device
        if out is None:
            out = torch.empty(fshape, dtype=torch.complex128, device=dev)
            accumulate = 0
        else:
            if tuple(out.shape) != fshape or out.dtype != torch.complex128 or not out.is_contiguous():
                raise RuntimeError(f"force output must be a contiguous complex128 tensor of shape {fshape}")
            accumulate = 1
        n = len(xs)
        xp = (ctypes.c_void_p * max(n, 1))(*[t.data_ptr() for t in xs])
        yp = (ctypes.c_void_p * max(n, 1))(*[t.data_ptr() for t in ys])
        wp = (ctypes.c_double * max(n, 1))(*w)
        _lib.check(_lib.lib().lym_lie_derivative(ctypes.byref(self._desc), n, xp, yp, wp, 1 if symmetric else 0,
                                                 accumulate, _lib.ptr(out), _lib.stream_handle(dev)),
                   "lie_derivative")
        return out

    def apply_square(self, vector, shift=0.0):
        """Op(Op(v)) + shift v in two fused launches (SquareOperator /
        ShiftedOperator fast path)."""
        x = to_device(vector)
        if tuple(x.shape) != tuple(self.spinor_shape):
            raise RuntimeError(f"vector shape {tuple(x.shape)} does not match operator shape {self.spinor_shape}")
        out = torch.empty_like(x)
        tmp = torch.empty_like(x)
        _lib.check(_lib.lib().lym_wilson_apply_square(ctypes.byref(self._desc), _lib.PREC_F64, float(shift),
                                                      _lib.ptr(x), _lib.ptr(out), _lib.ptr(tmp),
                                                      _lib.stream_handle(x.device)), "wilson_apply_square")
        return out


class Propagator:
    """D^-1 via the square operator: g5 D (A^-1 src); for an Overlap operator
    g5 (sol - O_{m=0} sol) (dirac.py:665-696)."""

    def __init__(self, operator, inverter):
        self.operator = operator
        self.inverter = inverter

    def apply(self, source):
        square_solution = self.inverter.solve(SquareOperator(self.operator), source)
        solution = self.operator.apply(square_solution)
        if not isinstance(self.operator, Overlap):
            return self.apply_gammad(solution)
        # overlap: subtract the massless operator (dirac.py:675-683)
        mass = self.operator.mass
        self.operator.mass = 0
        try:
            massless = self.operator.apply(solution)
        finally:
            self.operator.mass = mass
        return self.apply_gammad(algebra_utils.axpby(-1.0, massless, 1.0, solution))

    def apply_gammad(self, input_vector):
        x = to_device(input_vector)
        out = torch.empty_like(x)
        ns = x.shape[0]
        _lib.check(_lib.lib().lym_gamma5(_lib.PREC_F64, ns, x.numel() // ns, _lib.ptr(x), _lib.ptr(out),
                                         _lib.stream_handle(x.device)), "gamma5")
        return out


_lib.load()
