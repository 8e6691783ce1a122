"""GPU parity of the fermion force (lym_lie_derivative through
DiracWilsonOperator.lie_derivative / force_sum and fermion_action.n_flavor)
against the oracle (dirac.py:525-565, fermion_action.py:7-93) on identical
seeded inputs.  Bar: FP64 relative error <= 1e-13 for the force kernel;
n_flavor results, which include solves to a relative residual of 1e-12,
<= 1e-9."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

from oracle import force as ofo
from oracle import wilson as ow
from tfleonardym_amd import fields

TOL64 = 1e-13


def rel(a, b):
    a = a.detach().cpu().numpy() if isinstance(a, torch.Tensor) else a
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def make_op(U, geometry, kappa, herm, rep, bc, custom=False):
    from tfleonardym_amd import boundary_conditions, dirac, lattice
    cfg = lattice.Configuration(gauge_field=U, geometry=geometry, colors=U.shape[1],
                                number_of_dimensions=len(geometry))
    prov = None
    if bc is not None:
        boundary_conditions.set_antiperiodic_field_in_t_direction(geometry, factor=bc)
        prov = boundary_conditions.apply_antiboundary_conditions_in_t_direction
        if custom:
            std = prov

            def prov(g):  # not the standard provider object: links are signed on upload
                return std(g)
    return dirac.DiracWilsonOperator(cfg, kappa, herm, representation=rep, boundary_condition_provider=prov)


def _vecs(geometry, nc, rep, seed, n=2):
    D = len(geometry)
    ns = 2 if D == 2 else 4
    V = fields.volume(geometry)
    rng = np.random.default_rng(seed)
    shp = (ns, nc, V) if rep == "fundamental" else (ns, nc, nc, V)
    return [rng.standard_normal(shp) + 1j * rng.standard_normal(shp) for _ in range(n)]


CASES = []
for geometry in [(4, 6), (8, 8), (2, 4, 2, 4), (4, 4, 4, 8)]:
    for nc in (2, 3):
        for rep in ("fundamental", "adjoint"):
            CASES.append((geometry, nc, rep))


@pytest.mark.parametrize("geometry,nc,rep", CASES)
@pytest.mark.parametrize("herm", [False, True])
@pytest.mark.parametrize("bc", [None, -1.0])
def test_lie_derivative_matches_oracle(geometry, nc, rep, herm, bc):
    print(f"case force geometry={geometry} nc={nc} rep={rep} herm={herm} bc={bc}", flush=True)
    U = fields.gauge_field(geometry, nc, seed=41, kind="haar")
    X, Y = _vecs(geometry, nc, rep, 3)
    kappa = 0.1357
    wo = ow.WilsonOracle(U, geometry, kappa, herm, rep, bc_factor=bc)
    op = make_op(U, geometry, kappa, herm, rep, bc)
    got = op.lie_derivative(X, Y)
    assert tuple(got.shape) == (len(geometry), nc, nc, fields.volume(geometry))
    assert rel(got, ofo.lie_derivative(wo, X, Y)) < TOL64


@pytest.mark.parametrize("geometry", [(8, 2, 8, 64), (8, 4, 8, 8), (16, 2, 4, 16)])
@pytest.mark.parametrize("nc", [2, 3])
@pytest.mark.parametrize("herm", [True, False])
@pytest.mark.parametrize("lds", ["0", "1"])
def test_force_tiled_schedules_match_oracle(geometry, nc, herm, lds, monkeypatch):
    """Lattices on the XCD-tiled schedule: the unstaged kernel and the opt-in
    LDS-staged one (LYM_FORCE_LDS=1, hermitian; tile faces in x2 read
    globally when the block does not span x2, in-tile wrap when it does),
    3 symmetrised weighted pairs."""
    print(f"case force tiled geometry={geometry} nc={nc} herm={herm} lds={lds}", flush=True)
    monkeypatch.setenv("LYM_FORCE_LDS", lds)
    U = fields.gauge_field(geometry, nc, seed=47, kind="haar")
    wo = ow.WilsonOracle(U, geometry, 0.125, herm, "fundamental", bc_factor=-1.0)
    op = make_op(U, geometry, 0.125, herm, "fundamental", -1.0)
    vs = _vecs(geometry, nc, "fundamental", 6, n=4)
    pairs = [(vs[0], vs[1]), (vs[2], vs[3]), (vs[1], vs[2])]
    weights = [0.4, -1.1, 2.3]
    want = sum(w * (ofo.lie_derivative(wo, a, b) + ofo.lie_derivative(wo, b, a)) for (a, b), w in zip(pairs, weights))
    assert rel(op.force_sum(pairs, weights, symmetric=True), want) < TOL64
    want1 = ofo.lie_derivative(wo, vs[0], vs[3])
    assert rel(op.lie_derivative(vs[0], vs[3]), want1) < TOL64


@pytest.mark.parametrize("rep,nc", [("fundamental", 3), ("adjoint", 2), ("adjoint", 3)])
def test_lie_derivative_custom_boundary_provider(rep, nc):
    """A provider other than the standard one: links signed on upload (and
    for the adjoint the unsigned right links kept)."""
    geometry = (4, 2, 2, 6)
    print(f"case force custom provider rep={rep} nc={nc}", flush=True)
    U = fields.gauge_field(geometry, nc, seed=42, kind="haar")
    X, Y = _vecs(geometry, nc, rep, 4)
    wo = ow.WilsonOracle(U, geometry, 0.12, True, rep, bc_factor=-1.0)
    op = make_op(U, geometry, 0.12, True, rep, -1.0, custom=True)
    assert rel(op.lie_derivative(X, Y), ofo.lie_derivative(wo, X, Y)) < TOL64


@pytest.mark.parametrize("geometry,nc", [((8, 8, 8, 16), 3), ((16, 4, 4, 8), 2)])
def test_force_sum_many_pairs(geometry, nc):
    """40 weighted symmetrised pairs (two launches of at most 32), added to
    a prior field, against the oracle sum; and zero pairs."""
    print(f"case force_sum geometry={geometry} nc={nc}", flush=True)
    U = fields.gauge_field(geometry, nc, seed=43, kind="near_identity")
    wo = ow.WilsonOracle(U, geometry, 0.13, True, "fundamental", bc_factor=-1.0)
    op = make_op(U, geometry, 0.13, True, "fundamental", -1.0)
    base = _vecs(geometry, nc, "fundamental", 5, n=6)
    rng = np.random.default_rng(9)
    pairs, weights, ij = [], [], []
    for p in range(40):
        i, j = (int(v) for v in rng.integers(0, 6, 2))
        pairs.append((base[i], base[j]))
        ij.append((i, j))
        weights.append(float(rng.uniform(-2, 2)))
    L = {(i, j): ofo.lie_derivative(wo, base[i], base[j]) for i in range(6) for j in range(6)}
    prior_np = np.random.default_rng(2).standard_normal((len(geometry), nc, nc, fields.volume(geometry))) * (1 + 1j)
    want = prior_np.copy()
    for (i, j), w in zip(ij, weights):
        want = want + w * (L[(i, j)] + L[(j, i)])
    out = torch.from_numpy(prior_np).to("cuda")
    got = op.force_sum(pairs, weights, symmetric=True, out=out)
    assert got.data_ptr() == out.data_ptr()
    assert rel(got, want) < TOL64
    zero = op.force_sum([], [], symmetric=True)
    torch.cuda.synchronize()
    assert not bool(zero.abs().max().item())


def test_force_shape_errors_and_complex_weights():
    geometry = (4, 6)
    U = fields.gauge_field(geometry, 2, seed=1, kind="haar")
    op = make_op(U, geometry, 0.1, True, "fundamental", None)
    X, Y = _vecs(geometry, 2, "fundamental", 1)
    with pytest.raises(RuntimeError):
        op.lie_derivative(X[:, :, :-1], Y[:, :, :-1])
    with pytest.raises(RuntimeError):
        op.force_sum([(X, Y)], [1.0, 2.0])
    # complex weights: sum Re(w) L + i sum Im(w) L
    L = op.lie_derivative(X, Y).cpu().numpy()
    got = op.force_sum([(X, Y)], [0.5 + 2.0j], symmetric=False).cpu().numpy()
    assert np.abs(got - (0.5 + 2.0j) * L).max() < 1e-13 * np.abs(L).max()


def test_force_full_size_properties():
    """24^3 x 48 SU(3) (the production lattice): the symmetrised force of
    y = Q x is anti-Hermitian and traceless at every link (its generator
    coefficients are imaginary), and a two-pair launch equals two
    accumulated single-pair launches up to rounding."""
    geometry = (24, 24, 24, 48)
    print("case force full size 24^3x48", flush=True)
    U = fields.gauge_field(geometry, 3, seed=44, kind="near_identity")
    op = make_op(U, geometry, 0.1335, True, "fundamental", -1.0)
    g = torch.Generator(device="cpu").manual_seed(3)
    shp = op.spinor_shape
    x1 = torch.randn(shp, dtype=torch.complex128, generator=g).to("cuda")
    x2 = torch.randn(shp, dtype=torch.complex128, generator=g).to("cuda")
    y1, y2 = op.apply(x1), op.apply(x2)
    F = op.force_sum([(x1, y1), (x2, y2)], [0.6, 1.7], symmetric=True)
    Fh = F + F.conj().transpose(1, 2)
    assert float(Fh.abs().max()) < 1e-12 * float(F.abs().max())
    tr = F.diagonal(dim1=1, dim2=2).sum(-1)
    assert float(tr.abs().max()) < 1e-12 * float(F.abs().max())
    F1 = op.force_sum([(x1, y1)], [0.6], symmetric=True)
    op.force_sum([(x2, y2)], [1.7], symmetric=True, out=F1)
    assert float((F1 - F).abs().max()) < 1e-13 * float(F.abs().max())


def _n_flavor(U, geometry, rep, kappa, approx_tables, tol):
    from tfleonardym_amd import approximations as ap
    from tfleonardym_amd import dirac, dirac_solver, fermion_action
    op = make_op(U, geometry, kappa, True, rep, -1.0)
    solver = dirac_solver.multishift_solver(100000, 1e-30, algorithm="cg", tolerance=tol, verbose=False)
    ras = []
    for t in approx_tables:
        al, be = ap.split(t)
        ras.append(dirac.RationalApproximation(dirac.SquareOperator(op), al, be, 0.0, solver))
    return op, fermion_action.n_flavor(ras, ras, ras, solver, op)


@pytest.mark.parametrize("geometry,nc,rep", [((4, 4, 4, 8), 3, "fundamental"), ((8, 8), 2, "adjoint"),
                                             ((2, 2, 4, 4), 2, "adjoint")])
def test_n_flavor_force_matches_oracle(geometry, nc, rep):
    """n_flavor.force (multishift CG, Q x_i, one force launch over all poles
    of both pseudofermions) against the oracle's solves and Lie derivatives."""
    from tfleonardym_amd import approximations as ap
    print(f"case n_flavor geometry={geometry} nc={nc} rep={rep}", flush=True)
    U = fields.gauge_field(geometry, nc, seed=45, kind="near_identity")
    tables = [ap.force_level1_rational_approximation_coefficients] * 2
    op, nf = _n_flavor(U, geometry, rep, 0.125, tables, 1e-12)
    D = len(geometry)
    ns = 2 if D == 2 else 4
    phis = [fields.gaussian_spinor(geometry, ns, nc, seed=s, representation=rep) for s in (1, 2)]
    nf.pseudofermions = [torch.from_numpy(p).to("cuda") for p in phis]
    F = nf.force()
    wo = ow.WilsonOracle(U, geometry, 0.125, True, rep, bc_factor=-1.0)
    approx = [tuple(ap.split(t)) + (0.0,) for t in tables]
    want = ofo.force(wo, phis, approx, tol=1e-13)
    assert rel(F, want) < 1e-9
    e_want = ofo.energy(wo, phis, approx, tol=1e-13)
    assert abs(nf.energy() - e_want) < 1e-9 * abs(e_want)
    assert abs(nf.number_of_flavors() - ofo.number_of_flavors(approx)) < 1e-14


def test_n_flavor_energy_gradient():
    """tests.py:147-256 on the device: finite differences of n_flavor.energy
    under U_mu(x0) -> exp(i e T^a) U_mu(x0) against the force it returns,
    dS/de = -2i tr(T^a F_mu(x0))."""
    from tfleonardym_amd import approximations as ap
    from tfleonardym_amd import lattice
    geometry = (4, 2, 2, 4)
    nc = 3
    print("case n_flavor energy gradient", flush=True)
    U = fields.gauge_field(geometry, nc, seed=46, kind="near_identity")
    tables = [ap.force_level1_rational_approximation_coefficients]
    op, nf = _n_flavor(U, geometry, "fundamental", 0.13, tables, 1e-13)
    nf.initialize_pseudofermions([fields.gaussian_spinor(geometry, 4, nc, seed=7)])
    F = nf.force().cpu().numpy()
    T = ow.generators(nc)
    eps = 1e-5
    for mu, site, a in [(3, 63, 2), (0, 5, 7), (1, 17, 4)]:
        vals = []
        for e in (eps, -eps):
            cfg = lattice.Configuration(gauge_field=ofo.perturbed_links(U, mu, site, T[a], e), geometry=geometry,
                                        colors=nc, number_of_dimensions=4)
            nf.set_gauge_configuration(cfg)
            vals.append(nf.energy())
        fd = (vals[0] - vals[1]) / (2 * eps)
        want = -2j * np.trace(T[a] @ F[mu, :, :, site])
        assert abs(fd - want.real) < 1e-6 * max(1.0, abs(want)), (mu, site, a, fd, want)
