"""Roofline calibrators on the GPU: the MFMA burn loop (mfx_mfma_burn) and
the HBM copy (mfx_hbm_copy), plus their benchmark programs.

Bars: every wave's burn checksum equals the exact closed form NACC * I * K *
a * b (powers of two, so exact in FP32) and the reported FLOP count equals
waves * 8 * I * 2 * 32 * 32 * K; the copy is byte-exact for ragged sizes; bad
arguments are rejected on the host before any launch."""
import pytest
import torch

from test_gpu_benchmarks import bench

pytestmark = pytest.mark.gpu
DEV = "cuda"

# K of each precision's MFMA and the per-MFMA product a * b * K of the
# operands the burn uses (0.125 * 0.125 * 2, 0.125 * 0.125 * 16, 1 * 1 * 64).
BURN = {"fp32": (2, 1 / 32), "bf16": (16, 1 / 4), "split-bf16": (16, 1 / 4), "fp8-block": (64, 64.0)}


@pytest.mark.parametrize("prec", list(BURN))
@pytest.mark.parametrize("iters", [1, 37, 256])
@pytest.mark.parametrize("random_operands", [False, True])
def test_mfma_burn_checksum_and_flop(handle, prec, iters, random_operands):
    waves = handle.mfma_burn_waves()
    assert waves == 8 * torch.cuda.get_device_properties(0).multi_processor_count
    out = torch.full((waves,), float("nan"), device=DEV)
    if random_operands and prec.startswith("fp8"):
        with pytest.raises(RuntimeError, match="FP32 and BF16 only"):
            handle.mfma_burn(out, iters, prec=prec, random_operands=True)
        return
    flop, expected = handle.mfma_burn(out, iters, prec=prec, random_operands=random_operands)
    k, per = BURN[prec]
    # random: 4 cancelling MFMAs per iteration and one constant one, so every
    # accumulator ends at exactly one constant product
    mfmas = 4 * iters + 1 if random_operands else iters
    assert expected == 8 * (1 if random_operands else iters) * per
    assert flop == waves * 8 * mfmas * 2 * 32 * 32 * k
    torch.cuda.synchronize()
    assert torch.all(out == expected).item(), out.unique()


def test_mfma_burn_rejects_bad_arguments(handle):
    waves = handle.mfma_burn_waves()
    out = torch.empty(waves, device=DEV)
    with pytest.raises(RuntimeError):
        handle.mfma_burn(out, 0)
    with pytest.raises(RuntimeError):
        handle.mfma_burn(out, (1 << 20) + 1)
    with pytest.raises(RuntimeError):
        handle.mfma_burn(out[: waves - 1], 4)


@pytest.mark.parametrize("nbytes", [16, 4096 * 16, (1 << 24) + 13, 3 * (1 << 20) + 7])
def test_hbm_copy_is_byte_exact(handle, nbytes):
    g = torch.Generator(device=DEV).manual_seed(nbytes)
    src = torch.randint(0, 256, (nbytes,), dtype=torch.uint8, device=DEV, generator=g)
    dst = torch.zeros(nbytes + 64, dtype=torch.uint8, device=DEV)
    handle.hbm_copy(src, dst)
    torch.cuda.synchronize()
    assert torch.equal(dst[:nbytes], src)
    assert torch.count_nonzero(dst[nbytes:]).item() == 0


def test_hbm_copy_rejects_misaligned(handle):
    src = torch.zeros(1024, dtype=torch.uint8, device=DEV)
    dst = torch.zeros(1024, dtype=torch.uint8, device=DEV)
    with pytest.raises(RuntimeError):
        handle.hbm_copy(src[4:], dst)


@pytest.mark.parametrize("operands", ["constant", "random"])
@pytest.mark.parametrize("prec,name", [("fp32", "MicroBench_MFMA_fp32"), ("bf16", "MicroBench_MFMA_bf16"),
                                       ("split-bf16", "MicroBench_MFMA_bf16"), ("fp8-block", "MicroBench_MFMA_fp8block")])
def test_mfma_peak_program(prec, name, operands):
    r, res = bench("mfma_peak", "--size=512", "--num-runs=2", "--queue-profiling", f"--precision={prec}",
                   f"--operands={operands}")
    if operands == "random":
        name += "_random"
        if prec.startswith("fp8"):
            assert r.returncode != 0 and "FP32 and BF16 only" in r.stderr + r.stdout
            return
    assert r.returncode == 0, r.stderr
    assert res["_names"] == [name]
    assert res["Verification"] == "PASS"
    assert res["throughput-metric"].endswith("[FLOP]")


def test_dram_copy_program():
    r, res = bench("hbm_copy", "--size=255", "--num-runs=2", "--queue-profiling")
    assert r.returncode == 0, r.stderr
    assert res["_names"] == ["MicroBench_DRAM_float_1", "MicroBench_DRAM_double_1"]
    assert r.stdout.count("Verification: PASS") == 2
    assert res["throughput-metric"].endswith("[GiB]")
