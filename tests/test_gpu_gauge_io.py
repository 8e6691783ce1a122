"""GPU parity of the gauge-field file formats (lattice.py:10-70,
TfLeonardYMNDim.py:57-103): read_config (host read, device unpack and SU(3)
reconstruction) against the oracle reader, the Wilson operator on the loaded
field, and the save / load round trip with the reference's file naming."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

from oracle import gauge_io as og
from oracle import wilson as ow
from tfleonardym_amd import fields


@pytest.mark.parametrize("geometry,kind", [((4, 4, 4, 8), "haar"), ((2, 6), "skew"), ((8, 8, 8, 16), "haar")])
def test_read_config_matches_oracle(tmp_path, geometry, kind):
    from tfleonardym_amd import gauge_io
    print(f"case read_config geometry={geometry} kind={kind}", flush=True)
    D = len(geometry)
    V = fields.volume(geometry)
    if kind == "haar":
        U = fields.gauge_field(geometry, 3, seed=7, kind="haar")
    else:
        rng = np.random.default_rng(3)
        U = rng.standard_normal((D, 3, 3, V)) + 1j * rng.standard_normal((D, 3, 3, V))
    path = str(tmp_path / "cfg.xdr")
    og.write_ustar(path, U, geometry, params=(5.5, 0.14, 2.0))
    want, _, _ = og.read_config(path, geometry)
    cfg = gauge_io.read_config(path, geometry, verbose=False)
    assert cfg.number_of_dimensions == D and tuple(cfg.geometry) == tuple(geometry) and cfg.colors == 3
    got = cfg.gauge_field.cpu().numpy()
    assert np.abs(got - want).max() < 1e-14


def test_wilson_on_loaded_field(tmp_path):
    from tfleonardym_amd import dirac, gauge_io
    geometry = (4, 2, 4, 6)
    U = fields.gauge_field(geometry, 3, seed=8, kind="near_identity")
    path = str(tmp_path / "w.xdr")
    og.write_ustar(path, U, geometry)
    want_links, _, _ = og.read_config(path, geometry)
    cfg = gauge_io.read_config(path, geometry, verbose=False)
    op = dirac.DiracWilsonOperator(cfg, 0.13, True, representation="fundamental")
    psi = fields.gaussian_spinor(geometry, 4, 3, seed=2)
    want = ow.WilsonOracle(want_links, geometry, 0.13, True, "fundamental").apply(psi)
    got = op.apply(psi).cpu().numpy()
    assert np.linalg.norm(got - want) / np.linalg.norm(want) < 1e-13


def test_read_config_errors(tmp_path):
    from tfleonardym_amd import gauge_io
    path = str(tmp_path / "short.xdr")
    open(path, "wb").write(b"\0" * 100)
    with pytest.raises(RuntimeError):
        gauge_io.read_config(path, (2, 2, 2, 2), verbose=False)
    with pytest.raises(RuntimeError):
        gauge_io.read_config(path, (2, 2), Nc=2, verbose=False)


def test_save_load_round_trip(tmp_path):
    from tfleonardym_amd import gauge_io, lattice
    geometry = (4, 4, 2, 6)
    folder = str(tmp_path) + "/"
    fields_saved = []
    for n in (1, 2, 3):
        U = fields.gauge_field(geometry, 3, seed=n, kind="haar")
        fields_saved.append(U)
        cfg = lattice.Configuration(torch.from_numpy(U).cuda(), geometry, 3, 4)
        gauge_io.save_configuration(cfg, 5.6, 0.1335, n, folder)
    cfg, n = gauge_io.load_configuration(geometry, 5.6, 0.1335, None, folder)
    assert n == 3 and np.array_equal(cfg.gauge_field.cpu().numpy(), fields_saved[2])
    cfg, n = gauge_io.load_configuration(geometry, 5.6, 0.1335, 2, folder)
    assert n == 2 and np.array_equal(cfg.gauge_field.cpu().numpy(), fields_saved[1])
    assert tuple(cfg.geometry) == geometry and cfg.colors == 3 and cfg.number_of_dimensions == 4
