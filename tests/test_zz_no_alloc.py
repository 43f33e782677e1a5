"""mfx_gemm_noalloc (the entry point that allocates nothing) against mfx_gemm,
through the C ABI: the same kernels on the same inputs, so results must be
bit-identical -- on two calls in a row, the second with a smaller input (the
fixed array then still holds the first call's larger panels), for every
precision it takes, with transposed operands; a request whose packed operands
exceed the array is refused before anything is launched, leaving C untouched.
"""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

DEV = "cuda"


def _seeded(h, rows, cols, mid, pad=0):
    base = torch.empty((rows, cols + pad), device=DEV)
    x = base[:, :cols]
    h.fill_uniform(x, 21, mid)
    return x


@pytest.mark.parametrize("prec", ["split-bf16", "bf16", "fp32"])
@pytest.mark.parametrize("ta,tb", [(False, False), (True, True)])
def test_noalloc_matches_gemm_on_two_calls(handle, prec, ta, tb):
    for step, (m, n, k) in enumerate(((1000, 700, 1500), (300, 260, 129))):
        print(f"no-alloc {prec} ta={ta} tb={tb} call {step}: m={m} n={n} k={k}", flush=True)
        A = _seeded(handle, k if ta else m, m if ta else k, 0, pad=4)
        B = _seeded(handle, n if tb else k, k if tb else n, 1)
        C0 = _seeded(handle, m, n, 2, pad=8)
        want = C0.clone()
        got = C0.clone()
        handle.gemm(A, B, want, alpha=1.5, beta=1.2, prec=prec, transA=ta, transB=tb)
        handle.gemm_noalloc(A, B, got, alpha=1.5, beta=1.2, prec=prec, transA=ta, transB=tb)
        torch.cuda.synchronize()
        assert torch.equal(got, want), f"call {step}: {(got - want).abs().max().item()}"
        # and it is the product: FP64 on the CPU within the precision's tolerance
        ref = (1.5 * (A.double().cpu().T if ta else A.double().cpu()) @ (B.double().cpu().T if tb else B.double().cpu())
               + 1.2 * C0.double().cpu())
        err = (torch.linalg.norm(got.double().cpu() - ref) / torch.linalg.norm(ref)).item()
        assert err < {"fp32": 2e-6, "split-bf16": 1e-5, "bf16": 1e-2}[prec]


def test_noalloc_refuses_beyond_the_array(handle, mfx):
    # split BF16 packs about 4 (m + n) k bytes: 4 * 8192 * 2048 = 64 MiB of A alone
    m, n, k = 8192, 256, 2048
    A = torch.zeros((m, k), device=DEV)
    B = torch.zeros((k, n), device=DEV)
    C = torch.full((m, n), 3.0, device=DEV)
    with pytest.raises(mfx.MfxError) as e:
        handle.gemm_noalloc(A, B, C, alpha=1.0, beta=0.0, prec="split-bf16")
    assert e.value.status == 2 and "scratch" in str(e.value)
    with pytest.raises(mfx.MfxError) as e:
        handle.gemm_noalloc(A[:64, :64], B[:64, :64], C[:64, :64], prec="fp8-block")
    assert e.value.status == 2
    torch.cuda.synchronize()
    assert torch.all(C == 3.0).item()
    # the same shape at half the rows fits: 4 * (4096 + 256) * 2048 < 64 MiB
    C2 = torch.full((4096, n), 3.0, device=DEV)
    handle.gemm_noalloc(A[:4096], B, C2, prec="split-bf16")
    torch.cuda.synchronize()
    assert torch.all(C2 == 0.0).item()
