"""The matrix-core GEMM (tara_gemm) against numpy on the same (bf16-rounded)
operands: shapes with ragged edges, transposes through strides, two-level
batches, the three epilogues.  Device memory comes from torch."""
import numpy as np
import pytest

import tara_amd as T

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def _bf16_round(x):
    return torch.from_numpy(np.ascontiguousarray(x, dtype=np.float32)).to(torch.bfloat16).to(torch.float32).numpy()


@pytest.mark.parametrize("dtype", [T.GEMM_BF16, T.GEMM_F32])
@pytest.mark.parametrize("m,n,k,ta,tb", [(64, 64, 32, False, False), (100, 70, 45, False, True),
                                         (256, 128, 96, True, False), (33, 300, 513, True, True)])
def test_gemm_vs_numpy(dtype, m, n, k, ta, tb):
    print(f"gemm dtype={dtype} {m}x{n}x{k} ta={ta} tb={tb}", flush=True)
    rng = np.random.default_rng(m * 7 + n * 3 + k)
    nb1, nb2 = 3, 2
    A = rng.standard_normal((nb2, nb1, m, k)).astype(np.float32)
    B = rng.standard_normal((nb2, nb1, k, n)).astype(np.float32)
    if dtype == T.GEMM_BF16:
        A, B = _bf16_round(A), _bf16_round(B)
    tdt = torch.bfloat16 if dtype == T.GEMM_BF16 else torch.float32
    # storage: A as (z2, z1, m, k) or transposed (z2, z1, k, m); B as (.., k, n) or (.., n, k)
    As = np.swapaxes(A, 2, 3) if ta else A
    Bs = np.swapaxes(B, 2, 3) if tb else B
    dA = torch.from_numpy(np.ascontiguousarray(As)).to(tdt).cuda()
    dB = torch.from_numpy(np.ascontiguousarray(Bs)).to(tdt).cuda()
    sam, sak = (1, m) if ta else (k, 1)
    sbk, sbn = (1, k) if tb else (n, 1)
    want = np.einsum("yzmk,yzkn->yzmn", A.astype(np.float64), B.astype(np.float64))
    for epi in (T.GEMM_STORE_F32, T.GEMM_ACCUM_F32, T.GEMM_STORE_BF16):
        out_dt = torch.bfloat16 if epi == T.GEMM_STORE_BF16 else torch.float32
        base = torch.ones((nb2, nb1, m, n), dtype=out_dt).cuda()
        T.gemm(dtype, epi, m, n, k, dA.data_ptr(), sam, sak, dB.data_ptr(), sbk, sbn, base.data_ptr(), n, 1,
               nb1, m * k, k * n, m * n, nb2, nb1 * m * k, nb1 * k * n, nb1 * m * n, alpha=0.5)
        got = base.float().cpu().numpy().astype(np.float64)
        ref = 0.5 * want + (1.0 if epi == T.GEMM_ACCUM_F32 else 0.0)
        tol = 1e-2 if epi == T.GEMM_STORE_BF16 else 1e-5
        err = np.linalg.norm(got - ref) / np.linalg.norm(ref)
        assert err < tol, (epi, err)
