"""Snapshot and data formats (include/tara_hip.h, "snapshots and data"):
FP64 checkpoints, the reference's text snapshot and the FP32/BF16 trajectory-pair
store.  Writers call libtara_hip.so; readers are plain numpy (memory-mapped for
the pair store), so training code can stream pairs without the library."""
import ctypes as C

import numpy as np

from . import _check, _dp, _p, _real, _spec, lib

_CKPT = np.dtype([("magic", "S8"), ("version", "<u4"), ("kind", "<u4"), ("nx", "<u4"), ("ny", "<u4"),
                  ("nz", "<u4"), ("nfields", "<u4"), ("step", "<i8"), ("time", "<f8"), ("reserved", "V16")])
_PAIR = np.dtype([("magic", "S8"), ("version", "<u4"), ("dtype", "<u4"), ("nx", "<u4"), ("ny", "<u4"),
                  ("npairs", "<u8"), ("dt_pair", "<f8"), ("nu", "<f8"), ("reserved", "V16")])
FP32, BF16 = 1, 2


def _bind():
    L = lib()
    if not hasattr(L, "_snapio_bound"):
        L.tara_ckpt_write.argtypes = [C.c_char_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_long, C.c_double,
                                      C.c_void_p]
        L.tara_fort_write.argtypes = [C.c_char_p, C.c_int, C.c_int, C.c_double, C.c_double, _dp, _dp, _dp]
        L.tara_pairs_create.argtypes = [C.c_char_p, C.c_int, C.c_int, C.c_int, C.c_double, C.c_double,
                                        C.POINTER(C.c_void_p)]
        L.tara_pairs_append.argtypes = [C.c_void_p, _dp, _dp]
        L.tara_pairs_close.argtypes = [C.c_void_p]
        L._snapio_bound = True
    return L


def write_checkpoint(path, fields, nx, step=0, time=0.0):
    """fields: list of (Nh, Ny) or (Nh, Ny, Nz) complex half spectra."""
    fs = [_spec(f) for f in fields]
    ny = fs[0].shape[1]
    nz = fs[0].shape[2] if fs[0].ndim == 3 else 1
    ptrs = (C.c_void_p * len(fs))(*[f.ctypes.data for f in fs])
    _check(_bind().tara_ckpt_write(str(path).encode(), nx, ny, nz, len(fs), step, time, ptrs))


def read_checkpoint(path):
    """Returns (fields, header dict)."""
    h = np.fromfile(path, dtype=_CKPT, count=1)[0]
    if h["magic"] != b"TARACKPT" or h["version"] != 1:
        raise ValueError(f"{path} is not a checkpoint")
    nx, ny, nz, nf = (int(h[k]) for k in ("nx", "ny", "nz", "nfields"))
    shape = (nx // 2 + 1, ny) if nz == 1 else (nx // 2 + 1, ny, nz)
    n = int(np.prod(shape))
    data = np.fromfile(path, dtype="<c16", offset=_CKPT.itemsize, count=n * nf)
    fields = [np.asfortranarray(data[k * n:(k + 1) * n].reshape(shape[::-1]).T) for k in range(nf)]
    return fields, {"nx": nx, "ny": ny, "nz": nz, "step": int(h["step"]), "time": float(h["time"])}


def write_fort(path, omega, ux, uy, lx, ly):
    om, a, b = _real(omega), _real(ux), _real(uy)
    nx, ny = om.shape
    _check(_bind().tara_fort_write(str(path).encode(), nx, ny, lx, ly, _p(om), _p(a), _p(b)))


class PairWriter:
    def __init__(self, path, nx, ny, dtype=BF16, dt_pair=0.0, nu=0.0):
        h = C.c_void_p()
        _check(_bind().tara_pairs_create(str(path).encode(), nx, ny, dtype, dt_pair, nu, C.byref(h)))
        self._h = h

    def append(self, omega_in, omega_out):
        _check(_bind().tara_pairs_append(self._h, _p(_real(omega_in)), _p(_real(omega_out))))

    def close(self):
        if self._h:
            _check(_bind().tara_pairs_close(self._h))
            self._h = None


def read_pairs(path):
    """Returns (inputs, outputs) as float32 arrays of shape (npairs, Nx, Ny)
    (memory-mapped for FP32 files) and the header dict."""
    h = np.fromfile(path, dtype=_PAIR, count=1)[0]
    if h["magic"] != b"TARAPAIR" or h["version"] != 1:
        raise ValueError(f"{path} is not a pair store")
    nx, ny, n, dt = int(h["nx"]), int(h["ny"]), int(h["npairs"]), int(h["dtype"])
    raw = np.memmap(path, dtype="<f4" if dt == FP32 else "<u2", mode="r", offset=_PAIR.itemsize,
                    shape=(n, 2, ny, nx))
    if dt == BF16:
        raw = (np.asarray(raw).astype(np.uint32) << 16).view(np.float32)
    arr = np.asarray(raw).transpose(0, 1, 3, 2)  # (npairs, 2, Nx, Ny)
    return arr[:, 0], arr[:, 1], {"nx": nx, "ny": ny, "npairs": n, "dt_pair": float(h["dt_pair"]),
                                  "nu": float(h["nu"])}
