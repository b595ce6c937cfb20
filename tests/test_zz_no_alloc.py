"""The allocation-free derive2csh (tara_derive2csh_noalloc): device arrays in,
device arrays out, every intermediate in the library's static per-device
scratch.  Its results equal tara_derive2csh's (host arrays, allocated plan)
bit for bit, on consecutive calls whose second input is smaller (the scratch
keeps the first call's contents, and the second call's tables replace the
first's), and a grid beyond the scratch is refused before anything runs."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

import oracle as O  # noqa: E402
import tara_amd as T  # noqa: E402

pytestmark = pytest.mark.gpu


def _field(nx, ny, seed):
    rng = np.random.default_rng(seed)
    return O.r2c(np.asfortranarray(rng.standard_normal((nx, ny))))


def _on_device(w):
    # (Nh, Ny) Fortran order = Ny rows of Nh complex values, kx fastest
    return torch.from_numpy(np.ascontiguousarray(w.T)).to("cuda")


def _noalloc(w, nx, ny, nu):
    d_in = _on_device(w)
    d_out = torch.full_like(d_in, complex(np.nan, np.nan))
    stream = torch.cuda.current_stream().cuda_stream
    T.derive2csh_noalloc(nx, ny, nx // 2 + 1, nu, d_in.data_ptr(), d_out.data_ptr(), stream)
    torch.cuda.synchronize()
    return np.asfortranarray(d_out.cpu().numpy().T)


@pytest.mark.parametrize("sizes", [[(512, 256), (128, 64)], [(1024, 1024), (256, 512), (64, 32)]])
def test_noalloc_equals_derive2csh(sizes):
    for k, (nx, ny) in enumerate(sizes):
        print(f"noalloc derive2csh {nx}x{ny} (call {k})", flush=True)
        w = _field(nx, ny, 10 * nx + ny)
        got = _noalloc(w, nx, ny, 1e-3)
        want = T.derive2csh(nx, ny, nx // 2 + 1, 1e-3, w)
        assert np.array_equal(got, want), (nx, ny)


def test_noalloc_repeat_same_grid():
    """Two calls on the same grid (tables kept) agree with the drop-in too."""
    for seed in (1, 2):
        w = _field(256, 256, seed)
        assert np.array_equal(_noalloc(w, 256, 256, 5e-3), T.derive2csh(256, 256, 129, 5e-3, w))


def test_noalloc_rejects_oversize_grid():
    with pytest.raises(T.TaraError) as ei:
        T.derive2csh_noalloc(2048, 2048, 1025, 1e-3, 1, 1, 0)
    assert "unsupported" in str(ei.value)
