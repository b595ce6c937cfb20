"""GPU checks of the work moved off the host around the 2D time loop: real-space
fields, frame recording, and the device initial conditions (the reference's
shear layer; the Philox random-phase field)."""
import numpy as np
import pytest

import oracle as O
import tara_amd as T
from conftest import rel_l2

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("tag,d", [("asrun", 3.0), ("intent", 3.0 / 128.0)])
def test_device_shear_layer_vs_reference(golden, tag, d):
    """The device shear layer equals the reference's IC (src/main2csh.f95:84-110),
    and 2000 AB2 steps from it reproduce the reference trajectory."""
    print(f"device IC {tag}", flush=True)
    s = T.CshSolver(64, 64, 0.01, 1e-4, T.SCHEME_AB2)
    s.ic_shear(2.0, 2.0, 25.0, 2.0, d)
    assert rel_l2(s.download(), golden[f"c1_{tag}_w0"]) < 1e-14
    s.advance(2000)
    assert rel_l2(s.download(), golden[f"c1_{tag}_w2000"]) < 1e-10
    s.close()


@pytest.mark.parametrize("nx,ny,seed,k0", [(64, 64, 1, 4.0), (256, 128, 3, 6.5), (4096, 4096, 1, 4.0)])
def test_device_random_phase_vs_oracle(nx, ny, seed, k0):
    print(f"device random IC {nx}x{ny}", flush=True)
    s = T.CshSolver(nx, ny, 1e-4, 2e-4, T.SCHEME_RK4)
    s.ic_random(seed, k0, 1.0)
    got = s.download()
    s.close()
    assert rel_l2(got, O.ic_random(nx, ny, seed, k0, 1.0)) < 1e-14


def test_fields_vs_oracle():
    n = 128
    w = O.ic_random(n, n, 5, 4.0, 1.0)
    s = T.CshSolver(n, n, 1e-3, 1e-3, T.SCHEME_RK4)
    s.upload(w)
    om, ux, uy = s.fields()
    s.close()
    kx = np.arange(n // 2 + 1)[:, None]
    ky = np.fft.fftfreq(n, 1.0 / n)[None, :]
    k2 = kx * kx + ky * ky
    psi = np.where(k2 > 0, w / np.where(k2 > 0, k2, 1.0), 0)
    assert rel_l2(om, O.c2r(w, n) / n**2) < 1e-14
    assert rel_l2(ux, O.c2r(1j * ky * psi, n) / n**2) < 1e-14
    assert rel_l2(uy, O.c2r(-1j * kx * psi, n) / n**2) < 1e-14


def test_record_matches_advance_download():
    n = 64
    w = O.ic_random(n, n, 9, 4.0, 1.0)
    a = T.CshSolver(n, n, 1e-3, 1e-3, T.SCHEME_RK4)
    b = T.CshSolver(n, n, 1e-3, 1e-3, T.SCHEME_RK4)
    a.upload(w)
    b.upload(w)
    frames = a.record(3, 5)
    for f in frames:
        b.advance(5)
        assert np.array_equal(f, b.download())
    a.close()
    b.close()
