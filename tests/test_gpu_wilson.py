"""GPU parity of the Wilson operator (liblym through the drop-in Python API)
against the oracle on identical seeded inputs.  Bar: FP64 relative error
<= 1e-13 (norm-wise); FP32 <= 2e-6."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

from oracle import wilson as ow
from tfleonardym_amd import fields

TOL64 = 1e-13
TOL32 = 2e-6


def rel(a, b):
    a = a.detach().cpu().numpy() if isinstance(a, torch.Tensor) else a
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def make_op(U, geometry, kappa, herm, rep, bc, mode="projected"):
    from tfleonardym_amd import boundary_conditions, dirac, lattice
    cfg = lattice.Configuration(gauge_field=U, geometry=geometry, colors=U.shape[1],
                                number_of_dimensions=len(geometry))
    prov = None
    if bc is not None:
        boundary_conditions.set_antiperiodic_field_in_t_direction(geometry, factor=bc)
        prov = boundary_conditions.apply_antiboundary_conditions_in_t_direction
    return dirac.DiracWilsonOperator(cfg, kappa, herm, mode=mode, representation=rep,
                                     boundary_condition_provider=prov)


CASES = []
for geometry in [(4, 6), (8, 8), (2, 4, 2, 4), (4, 4, 4, 8)]:
    for nc in (2, 3):
        for rep in ("fundamental", "adjoint"):
            CASES.append((geometry, nc, rep))


@pytest.mark.parametrize("geometry,nc,rep", CASES)
@pytest.mark.parametrize("herm", [False, True])
@pytest.mark.parametrize("bc", [None, -1.0])
@pytest.mark.parametrize("mode", ["projected", "simple"])
def test_wilson_matches_oracle(geometry, nc, rep, herm, bc, mode):
    print(f"case geometry={geometry} nc={nc} rep={rep} herm={herm} bc={bc} mode={mode}", flush=True)
    D = len(geometry)
    ns = 2 if D == 2 else 4
    U = fields.gauge_field(geometry, nc, seed=101, kind="haar")
    psi = fields.gaussian_spinor(geometry, ns, nc, seed=7, representation=rep)
    kappa = 0.1357
    want = ow.WilsonOracle(U, geometry, kappa, herm, rep, bc_factor=bc, mode=mode).apply(psi)
    op = make_op(U, geometry, kappa, herm, rep, bc, mode)
    got = op.apply(psi)
    assert rel(got, want) < TOL64


@pytest.mark.parametrize("geometry", [(8, 8, 8, 8), (8, 4, 4, 16), (16, 8, 8, 16)])
@pytest.mark.parametrize("nc,rep", [(3, "fundamental"), (2, "adjoint"), (2, "fundamental")])
@pytest.mark.parametrize("bc", [None, -1.0])
def test_tiled_schedule_parity(geometry, nc, rep, bc):
    """Lattices where the Wilson kernels use the XCD-tiled block schedule."""
    print(f"case tiled geometry={geometry} nc={nc} rep={rep} bc={bc}", flush=True)
    U = fields.gauge_field(geometry, nc, seed=33, kind="haar")
    psi = fields.gaussian_spinor(geometry, 4, nc, seed=34, representation=rep)
    want = ow.WilsonOracle(U, geometry, 0.125, True, rep, bc_factor=bc).apply(psi)
    op = make_op(U, geometry, 0.125, True, rep, bc)
    assert rel(op.apply(psi), want) < TOL64


def _tile_shapes(geometry, es, ncomp, max_sites=None):
    """Every LDS tile (a0, a1, k) the library admits on `geometry`: a0 divides
    L0/8, a1 | L1, k | L2, a0*a1*k*L3 a multiple of 64 and <= 1024 sites
    (640 for complex128, 256 for adjoint spinors), tile spinor <= 150 KiB."""
    L0, L1, L2, T = geometry
    slab = L0 // 8
    if max_sites is None:
        max_sites = 640 if es == 16 else 1024
    out = []
    for a0 in range(1, slab + 1):
        for a1 in range(1, L1 + 1):
            for k in range(1, L2 + 1):
                ns = a0 * a1 * k * T
                if slab % a0 or L1 % a1 or L2 % k or ns % 64 or ns > max_sites or ns * ncomp * es > 150 * 1024:
                    continue
                out.append((a0, a1, k))
    return out


@pytest.mark.parametrize("prec", ["f64", "f32"])
@pytest.mark.parametrize("geometry", [(16, 8, 8, 16), (8, 4, 6, 32)])
def test_lds_tile_shapes(prec, geometry, monkeypatch):
    """Every admissible LDS tile shape (forced through LYM_TILE) gives the
    oracle's D, and the same result as the untiled kernel (LYM_NO_TILE=1) to
    rounding: partial tiles, full-extent wrap, tile edges in every direction.
    The other walk order over the tiles gives the identical result."""
    import ctypes
    from tfleonardym_amd import _lib
    print(f"case lds tiles prec={prec} geometry={geometry}", flush=True)
    U = fields.gauge_field(geometry, 3, seed=41, kind="haar")
    psi = fields.gaussian_spinor(geometry, 4, 3, seed=42)
    op = make_op(U, geometry, 0.127, True, "fundamental", -1.0)
    desc = op.ensure_f32()
    want = ow.WilsonOracle(U, geometry, 0.127, True, "fundamental", bc_factor=-1.0).apply(psi)
    pc = _lib.PREC_F64 if prec == "f64" else _lib.PREC_F32
    dt = torch.complex128 if prec == "f64" else torch.complex64
    x = torch.from_numpy(psi).to("cuda").to(dt)

    def run():
        out = torch.zeros_like(x)
        _lib.check(_lib.lib().lym_wilson_apply(ctypes.byref(desc), pc, 0, _lib.ptr(x), _lib.ptr(out),
                                               _lib.stream_handle()))
        return out

    monkeypatch.setenv("LYM_NO_TILE", "1")
    plain = run()
    monkeypatch.setenv("LYM_NO_TILE", "0")
    shapes = _tile_shapes(geometry, 16 if prec == "f64" else 8, 12)
    assert shapes
    tol = TOL64 if prec == "f64" else TOL32
    for i, (a0, a1, k) in enumerate(shapes):
        monkeypatch.setenv("LYM_TILE", f"{a0},{a1},{k}")
        # hops in flight per wave: every value on the first and last shapes
        for hif in ((1, 2, 4) if i in (0, len(shapes) - 1) else (1,)):
            monkeypatch.setenv("LYM_HIF", str(hif))
            got = run()
            assert rel(got.to(torch.complex128), want) < tol, (a0, a1, k, hif)
            d = (got - plain).abs().max().item() / plain.abs().max().item()
            assert d < (1e-15 if prec == "f64" else 1e-6), (a0, a1, k, hif, d)
            if hif == 1:
                walk0 = got
        monkeypatch.delenv("LYM_HIF")
        monkeypatch.setenv("LYM_TILE", f"{a0},{a1},{k},1")
        assert torch.equal(run(), walk0), (a0, a1, k, "walk 1")


@pytest.mark.parametrize("prec", ["f64", "f32"])
@pytest.mark.parametrize("nc", [2, 3])
@pytest.mark.parametrize("herm,bc", [(True, -1.0), (False, None)])
def test_lds_tile_shapes_adjoint(prec, nc, herm, bc, monkeypatch):
    """Adjoint representation (matrix spinors U X U^dag) on the LDS-tiled
    kernel: every admissible tile shape (at most 256 sites, one hop in
    flight) gives the oracle's D and the untiled kernel's result to rounding."""
    import ctypes
    from tfleonardym_amd import _lib
    geometry = (16, 4, 4, 16)
    print(f"case adjoint lds tiles prec={prec} nc={nc} herm={herm} bc={bc}", flush=True)
    U = fields.gauge_field(geometry, nc, seed=51, kind="haar")
    psi = fields.gaussian_spinor(geometry, 4, nc, seed=52, representation="adjoint")
    op = make_op(U, geometry, 0.131, herm, "adjoint", bc)
    desc = op.ensure_f32()
    want = ow.WilsonOracle(U, geometry, 0.131, herm, "adjoint", bc_factor=bc).apply(psi)
    pc = _lib.PREC_F64 if prec == "f64" else _lib.PREC_F32
    dt = torch.complex128 if prec == "f64" else torch.complex64
    x = torch.from_numpy(psi).to("cuda").to(dt)

    def run():
        out = torch.zeros_like(x)
        _lib.check(_lib.lib().lym_wilson_apply(ctypes.byref(desc), pc, 0, _lib.ptr(x), _lib.ptr(out),
                                               _lib.stream_handle()))
        return out

    monkeypatch.setenv("LYM_NO_TILE", "1")
    plain = run()
    monkeypatch.setenv("LYM_NO_TILE", "0")
    info = _lib.LymTileChoice()
    _lib.check(_lib.lib().lym_tile_info(ctypes.byref(desc), pc, ctypes.byref(info)))
    assert info.a0 > 0 and info.nsite <= 256, info.as_dict()  # the default runs a tile
    es = 16 if prec == "f64" else 8
    shapes = _tile_shapes(geometry, es, 4 * nc * nc, max_sites=256)
    assert shapes
    tol = TOL64 if prec == "f64" else TOL32
    for a0, a1, k in shapes:
        monkeypatch.setenv("LYM_TILE", f"{a0},{a1},{k}")
        got = run()
        assert rel(got.to(torch.complex128), want) < tol, (a0, a1, k)
        d = (got - plain).abs().max().item() / plain.abs().max().item()
        assert d < (1e-15 if prec == "f64" else 1e-6), (a0, a1, k, d)


def test_autotune_records_and_keeps_parity():
    """lym_tune_wilson times every tile / hop grouping and the untiled kernel,
    records the fastest for this lattice; later applications use it (tile
    info says tuned) and still give the oracle's D, in both precisions.
    (A lattice no other test uses: the tuned table lives for the process.)"""
    import ctypes
    from tfleonardym_amd import _lib
    geometry = (16, 4, 4, 16)
    U = fields.gauge_field(geometry, 3, seed=91, kind="haar")
    psi = fields.gaussian_spinor(geometry, 4, 3, seed=92)
    op = make_op(U, geometry, 0.121, True, "fundamental", -1.0)
    want = ow.WilsonOracle(U, geometry, 0.121, True, "fundamental", bc_factor=-1.0).apply(psi)
    assert op.tile_info("f32")["tuned"] == 0
    choices = op.autotune(reps=3)
    for name, tol in (("f64", TOL64), ("f32", TOL32)):
        c = choices[name]
        assert c["tuned"] == 1 and c["us"] > 0
        info = op.tile_info(name)
        assert info["tuned"] == 1 and (info["a0"], info["a1"], info["k"]) == (c["a0"], c["a1"], c["k"])
        if c["a0"]:
            assert geometry[1] % c["a1"] == 0 and geometry[2] % c["k"] == 0 and c["nsite"] % 64 == 0
        dt = torch.complex128 if name == "f64" else torch.complex64
        x = torch.from_numpy(psi).to("cuda").to(dt)
        y = torch.empty_like(x)
        prec = _lib.PREC_F64 if name == "f64" else _lib.PREC_F32
        _lib.check(_lib.lib().lym_wilson_apply(ctypes.byref(op.desc), prec, 0, _lib.ptr(x), _lib.ptr(y),
                                               _lib.stream_handle()))
        assert rel(y.to(torch.complex128), want) < tol, (name, c)


@pytest.mark.parametrize("nc", [2, 3])
def test_su_link_rebuild(nc, monkeypatch):
    """SU(N) links are flagged LINKS_SU and the kernels rebuild each link's
    last row: D matches the oracle (which reads all entries) and the
    full-read kernels (LYM_FULL_LINKS=1) to FP64 rounding.  Links that are not
    SU(N) (random complex matrices; -U for SU(3), det = -1) are flagged
    LINKS_FULL and still match the oracle."""
    from tfleonardym_amd import _lib
    print(f"case su link rebuild nc={nc}", flush=True)
    geometry = (8, 4, 4, 8)
    U = fields.gauge_field(geometry, nc, seed=51, kind="haar")
    psi = fields.gaussian_spinor(geometry, 4, nc, seed=52)
    want = ow.WilsonOracle(U, geometry, 0.13, True, "fundamental", bc_factor=-1.0).apply(psi)
    op = make_op(U, geometry, 0.13, True, "fundamental", -1.0)
    assert op.desc.link_format == _lib.LINKS_SU
    got = op.apply(psi)
    assert rel(got, want) < TOL64
    monkeypatch.setenv("LYM_FULL_LINKS", "1")
    op_full = make_op(U, geometry, 0.13, True, "fundamental", -1.0)
    assert op_full.desc.link_format == _lib.LINKS_FULL
    assert rel(got, op_full.apply(psi).cpu().numpy()) < 1e-14
    monkeypatch.delenv("LYM_FULL_LINKS")
    rng = np.random.default_rng(5)
    bad = [rng.standard_normal(U.shape) + 1j * rng.standard_normal(U.shape)]
    if nc == 3:
        bad.append(-U)
    for W in bad:
        op_w = make_op(W, geometry, 0.13, True, "fundamental", -1.0)
        assert op_w.desc.link_format == _lib.LINKS_FULL
        want_w = ow.WilsonOracle(W, geometry, 0.13, True, "fundamental", bc_factor=-1.0).apply(psi)
        assert rel(op_w.apply(psi), want_w) < TOL64


def test_custom_provider_is_materialised():
    """A provider that is not the standard one is applied once to the links."""
    from tfleonardym_amd import boundary_conditions, dirac, lattice
    geometry = (4, 4, 4, 4)
    U = fields.gauge_field(geometry, 3, seed=5, kind="haar")
    boundary_conditions.set_antiperiodic_field_in_t_direction(geometry)
    cfg = lattice.Configuration(U, geometry, 3, 4)
    for rep in ("fundamental", "adjoint"):
        op = dirac.DiracWilsonOperator(cfg, 0.12, True, representation=rep,
                                       boundary_condition_provider=lambda x: boundary_conditions.
                                       apply_antiboundary_conditions_in_t_direction(x))
        psi = fields.gaussian_spinor(geometry, 4, 3, seed=2, representation=rep)
        want = ow.WilsonOracle(U, geometry, 0.12, True, rep, bc_factor=-1.0).apply(psi)
        assert rel(op.apply(psi), want) < TOL64


def test_adjoint_operator_fundamental_form_with_adjoint_links():
    """tests.py:83-90 on the device: an adjoint operator's fundamental-form
    method runs with V_ab links on x^a = 2 tr(T^a X)."""
    geometry = (8, 8)
    U = fields.gauge_field(geometry, 2, seed=9, kind="haar")
    op = make_op(U, geometry, 0.1, True, "adjoint", -1.0)
    X = fields.gaussian_spinor(geometry, 2, 2, seed=3, representation="adjoint")
    t1 = op.simple_multiply_in_adjoint_representation(X).cpu().numpy()
    t2 = op.multiply_in_adjoint_representation(X).cpu().numpy()
    assert rel(t1, t2) < TOL64
    x = ow.to_adjoint_vector(X)
    t3 = op.multiply_in_fundamental_representation(x).cpu().numpy()
    t4 = ow.to_adjoint_vector(t1)
    assert rel(t3, t4) < 1e-12


@pytest.mark.parametrize("geometry,nc", [((8, 8), 2), ((4, 4, 4, 8), 3), ((4, 4, 4, 4), 2)])
def test_square_and_shift(geometry, nc):
    from tfleonardym_amd import dirac
    ns = 2 if len(geometry) == 2 else 4
    U = fields.gauge_field(geometry, nc, seed=1, kind="haar")
    psi = fields.gaussian_spinor(geometry, ns, nc, seed=8)
    oop = ow.WilsonOracle(U, geometry, 0.13, True, "fundamental", bc_factor=-1.0)
    op = make_op(U, geometry, 0.13, True, "fundamental", -1.0)
    assert rel(dirac.SquareOperator(op).apply(psi), oop.apply_square(psi)) < TOL64
    sh = dirac.ShiftedOperator(dirac.SquareOperator(op), 0.37)
    assert rel(sh.apply(psi), oop.apply_square(psi, 0.37)) < TOL64


def test_float32_apply():
    import ctypes
    from tfleonardym_amd import _lib
    geometry = (4, 4, 4, 8)
    U = fields.gauge_field(geometry, 3, seed=3, kind="near_identity", eps=0.3)
    psi = fields.gaussian_spinor(geometry, 4, 3, seed=1)
    op = make_op(U, geometry, 0.13, True, "fundamental", -1.0)
    desc = op.ensure_f32()
    x = torch.from_numpy(psi).to("cuda").to(torch.complex64)
    out = torch.empty_like(x)
    _lib.check(_lib.lib().lym_wilson_apply(ctypes.byref(desc), _lib.PREC_F32, 0, _lib.ptr(x), _lib.ptr(out),
                                           _lib.stream_handle()))
    want = ow.WilsonOracle(U, geometry, 0.13, True, "fundamental", bc_factor=-1.0).apply(psi)
    assert rel(out.to(torch.complex128), want) < TOL32


@pytest.mark.parametrize("geometry", [(16, 16, 16, 32), (24, 24, 24, 48)])
def test_float32_apply_full_size(geometry):
    """complex64 D on the benchmark lattices (rebuilt link rows, antiperiodic
    t) against the complex128 D on the same input: the largest component
    difference stays within 4e-7 of the largest component, the rounding of
    complex64 arithmetic on a 9-term sum."""
    import ctypes
    from tfleonardym_amd import _lib
    print(f"case f32 vs f64 D geometry={geometry}", flush=True)
    U = fields.gauge_field(geometry, 3, seed=20250, kind="near_identity", eps=0.25)
    op = make_op(U, geometry, 0.1335, True, "fundamental", -1.0)
    desc = op.ensure_f32()
    g = torch.Generator(device="cpu").manual_seed(3)
    x64 = torch.randn(op.spinor_shape, dtype=torch.complex128, generator=g).cuda()
    y64 = torch.empty_like(x64)
    _lib.check(_lib.lib().lym_wilson_apply(ctypes.byref(desc), _lib.PREC_F64, 0, _lib.ptr(x64), _lib.ptr(y64),
                                           _lib.stream_handle()))
    x32 = x64.to(torch.complex64)
    y32 = torch.empty_like(x32)
    _lib.check(_lib.lib().lym_wilson_apply(ctypes.byref(desc), _lib.PREC_F32, 0, _lib.ptr(x32), _lib.ptr(y32),
                                           _lib.stream_handle()))
    err = (y32.to(torch.complex128) - y64).abs().max().item() / y64.abs().max().item()
    assert np.isfinite(err) and err < 4e-7, err


def test_c2_size_parity_and_hermiticity():
    """16^3 x 32 SU(3) (the multishift workload's lattice): FP64 parity and
    <v2, Q v1> = <Q v2, v1> at full size."""
    from tfleonardym_amd import algebra_utils
    geometry = (16, 16, 16, 32)
    U = fields.gauge_field(geometry, 3, seed=2024, kind="near_identity", eps=0.3)
    v1 = fields.gaussian_spinor(geometry, 4, 3, seed=11)
    v2 = fields.gaussian_spinor(geometry, 4, 3, seed=12)
    op = make_op(U, geometry, 0.13, True, "fundamental", None)
    q1 = op.apply(v1)
    want = ow.WilsonOracle(U, geometry, 0.13, True, "fundamental").apply(v1)
    assert rel(q1, want) < TOL64
    q2 = op.apply(v2)
    a = algebra_utils.dot(torch.from_numpy(v2).cuda(), q1)
    b = algebra_utils.dot(q2, torch.from_numpy(v1).cuda())
    assert abs(a - b) < 1e-11 * abs(a)
    # deterministic: same bits on a second application
    assert torch.equal(op.apply(v1), q1)


def test_blas():
    from tfleonardym_amd import algebra_utils
    rng = np.random.default_rng(0)
    for n in (1, 7, 1000, 123457):
        a = rng.standard_normal(n) + 1j * rng.standard_normal(n)
        b = rng.standard_normal(n) + 1j * rng.standard_normal(n)
        ta = torch.from_numpy(a).cuda()
        tb = torch.from_numpy(b).cuda()
        d = algebra_utils.dot(ta, tb)
        assert abs(d - np.vdot(a, b)) < 1e-12 * max(1.0, abs(np.vdot(a, b)))
        assert algebra_utils.dot(ta, tb) == d
        y = tb.clone()
        algebra_utils.axpby(0.3 - 0.2j, ta, 1.5 + 0.1j, y)
        assert np.allclose(y.cpu().numpy(), (0.3 - 0.2j) * a + (1.5 + 0.1j) * b, rtol=1e-14, atol=1e-14)
    with pytest.raises(RuntimeError):
        algebra_utils.dot(torch.zeros(3, dtype=torch.complex128).cuda(), torch.zeros(4, dtype=torch.complex128).cuda())


def test_gamma5_and_errors():
    from tfleonardym_amd import dirac, lattice
    geometry = (4, 4)
    U = fields.gauge_field(geometry, 2, seed=1, kind="haar")
    op = make_op(U, geometry, 0.1, True, "fundamental", None)
    psi = fields.gaussian_spinor(geometry, 2, 2, seed=1)
    g = op.apply_hermiticity(psi).cpu().numpy()
    assert np.array_equal(g, ow.gamma5(psi))
    with pytest.raises(RuntimeError):
        dirac.DiracWilsonOperator(lattice.Configuration(U, geometry, 2, 2), 0.1, mode="bogus")
    U3 = fields.gauge_field((2, 2, 2), 2, seed=1, kind="haar")
    with pytest.raises(RuntimeError):
        dirac.DiracWilsonOperator(lattice.Configuration(U3, (2, 2, 2), 2, 3), 0.1)
    with pytest.raises(RuntimeError):
        op.apply(np.zeros((2, 2, 15), dtype=np.complex128))


@pytest.mark.parametrize("ndim,nc", [(4, 3), (4, 2), (2, 3), (2, 2)])
def test_pack_links_f32_layout(ndim, nc):
    """lym_pack_links_f32 against the layout it documents: per direction,
    ceil(nc^2 / 2) planes of 16-byte units; unit q of a site holds the
    row-major elements 2q and 2q + 1 as complex64, zero after the last."""
    from tfleonardym_amd import _lib
    dev = _lib.require_device()
    rng = np.random.default_rng(100 + 10 * ndim + nc)
    V = 96
    U = rng.standard_normal((ndim, nc, nc, V)) + 1j * rng.standard_normal((ndim, nc, nc, V))
    packed = _lib.pack_links_f32(torch.from_numpy(U).to(dev)).cpu().numpy()
    qu = (nc * nc + 1) // 2
    units = packed.reshape(ndim, qu, V, 4)
    flat = U.reshape(ndim, nc * nc, V).astype(np.complex64)
    for q in range(qu):
        for half, e in ((0, 2 * q), (1, 2 * q + 1)):
            want = flat[:, e, :] if e < nc * nc else np.zeros((ndim, V), np.complex64)
            assert np.array_equal(units[:, q, :, 2 * half], want.real)
            assert np.array_equal(units[:, q, :, 2 * half + 1], want.imag)
