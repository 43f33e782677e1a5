"""Persistent grouped launch, chains, hipGraph replay and the autotuner (GPU,
through the C ABI).  Bars: relative Frobenius error vs FP64 per problem with
the precision's tolerance; bit equality with the single-problem path when both
use the same tile configuration; every graph replay equal to the eager result."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu
DEV = "cuda"
TOL = {"fp32": 2e-6, "split-bf16": 1e-5, "bf16": 1e-2}


def rand_problems(mfx, rng, count, lo=16, hi=600, transposes=True):
    """Uneven problems (M, N, K multiples of 16 drawn from rng) with device
    operands: (A, B, C, transA, transB, matrix ids for the seeded fill)."""
    keep = []
    for i in range(count):
        m, n, k = (int(rng.integers(lo // 16, hi // 16 + 1)) * 16 for _ in range(3))
        ta = bool(transposes and rng.integers(2))
        tb = bool(transposes and rng.integers(2))
        A = torch.empty((k, m) if ta else (m, k), device=DEV)
        B = torch.empty((n, k) if tb else (k, n), device=DEV)
        C = torch.empty((m, n), device=DEV)
        keep.append((A, B, C, ta, tb, (3 * i, 3 * i + 1, 3 * i + 2)))
    return keep


def fill(handle, keep, seed):
    for A, B, C, ta, tb, ids in keep:
        handle.fill_uniform(A, seed, ids[0])
        handle.fill_uniform(B, seed, ids[1])
        handle.fill_uniform(C, seed, ids[2])


def fp64(A, B, C0, alpha, beta, ta, tb):
    a = A.double().T if ta else A.double()
    b = B.double().T if tb else B.double()
    return alpha * (a @ b) + beta * C0.double()


def rel(x, r):
    return (torch.linalg.norm(x.double() - r) / torch.linalg.norm(r)).item()


@pytest.mark.parametrize("prec", ["fp32", "split-bf16", "bf16"])
def test_grouped_uneven_problems_vs_fp64(handle, mfx, prec):
    rng = np.random.default_rng(1)
    keep = rand_problems(mfx, rng, 40)
    fill(handle, keep, 5)
    C0 = [k[2].clone() for k in keep]
    probs = [mfx.problem(A, B, C, alpha=1.5, beta=1.2, transA=ta, transB=tb) for A, B, C, ta, tb, _ in keep]
    handle.grouped(probs, prec=prec)
    handle.grouped_status()
    for (A, B, C, ta, tb, _), c0 in zip(keep, C0):
        err = rel(C, fp64(A, B, c0, 1.5, 1.2, ta, tb))
        assert err < TOL[prec], err


@pytest.mark.parametrize("prec", ["fp32", "split-bf16"])
@pytest.mark.parametrize("config", [0, 1, 2, 3, 4, 5, 6])
def test_grouped_bitwise_equals_single_launch(handle, mfx, prec, config):
    rng = np.random.default_rng(2 + config)
    keep = rand_problems(mfx, rng, 12, hi=520)
    fill(handle, keep, 6)
    C0 = [k[2].clone() for k in keep]
    handle.force_config(prec, config)
    handle.force_splits(1)  # the grouped kernels run each tile whole: compare with unsplit single launches
    try:
        probs = [mfx.problem(A, B, C, alpha=0.5, beta=-1.0, transA=ta, transB=tb) for A, B, C, ta, tb, _ in keep]
        handle.grouped(probs, prec=prec)
        for (A, B, C, ta, tb, _), c0 in zip(keep, C0):
            single = c0.clone()
            handle.gemm(A, B, single, alpha=0.5, beta=-1.0, prec=prec, transA=ta, transB=tb)
            torch.cuda.synchronize()
            assert torch.equal(C, single)
    finally:
        handle.force_config(prec, -1)
        handle.force_splits(-1)


def chain_fp64(mats):
    level = [m.double() for m in mats]
    while len(level) > 1:
        up = [level[i] @ level[i + 1] for i in range(0, len(level) - 1, 2)]
        if len(level) % 2:
            up.append(level[-1])
        level = up
    return level[0]


@pytest.mark.parametrize("prec", ["fp32", "split-bf16"])
@pytest.mark.parametrize("count,n", [(8, 256), (4, 320), (5, 192), (2, 128)])
def test_chain(handle, prec, count, n):
    mats = [torch.empty((n, n), device=DEV) for _ in range(count)]
    for i, m in enumerate(mats):
        handle.fill_uniform(m, 11, i)
    scratch = [torch.empty((n, n), device=DEV) for _ in range(max(0, count - 2))]
    out = torch.empty((n, n), device=DEV)
    handle.chain(mats, scratch, out, prec=prec)
    handle.grouped_status()
    err = rel(out, chain_fp64(mats))
    assert err < TOL[prec] * (count - 1) ** 0.5, err


def test_chain_identity_is_exact(handle):
    n = 200
    eye = torch.eye(n, device=DEV)
    mats = [eye.clone() for _ in range(4)]
    scratch = [torch.empty((n, n), device=DEV) for _ in range(2)]
    out = torch.full((n, n), float("nan"), device=DEV)
    handle.chain(mats, scratch, out, prec="split-bf16")
    assert torch.equal(out, eye)


@pytest.mark.parametrize("prec", ["fp32", "split-bf16"])
def test_grouped_2mm_3mm_dependencies(handle, mfx, prec):
    """A 2mm chain (T = A B; D = 1.5 T C + 1.2 D) and a 3mm diamond in one launch."""
    n = 384
    X = [torch.empty((n, n), device=DEV) for _ in range(8)]
    for i, x in enumerate(X):
        handle.fill_uniform(x, 12, i)
    A, B, C, D, P, Q, R, S = X
    T = torch.empty((n, n), device=DEV)
    E = torch.empty((n, n), device=DEV)
    F = torch.empty((n, n), device=DEV)
    G = torch.empty((n, n), device=DEV)
    D0 = D.clone()
    probs = [mfx.problem(A, B, T), mfx.problem(P, Q, E), mfx.problem(R, S, F),
             mfx.problem(T, C, D, alpha=1.5, beta=1.2), mfx.problem(E, F, G)]
    deps = [(3, 0), (4, 1), (4, 2)]
    handle.grouped(probs, deps, prec=prec)
    handle.grouped_status()
    want_d = 1.5 * (A.double() @ B.double()) @ C.double() + 1.2 * D0.double()
    want_g = (P.double() @ Q.double()) @ (R.double() @ S.double())
    assert rel(D, want_d) < TOL[prec] * 2 ** 0.5
    assert rel(G, want_g) < TOL[prec] * 2 ** 0.5


def test_dependency_order_is_validated(handle, mfx):
    n = 64
    A = torch.empty((n, n), device=DEV)
    probs = [mfx.problem(A, A, torch.empty((n, n), device=DEV)) for _ in range(2)]
    with pytest.raises(mfx.MfxError) as e:
        handle.grouped(probs, [(0, 1)], prec="fp32")
    assert e.value.status == 1
    with pytest.raises(mfx.MfxError):
        handle.grouped(probs, [(0, 5)], prec="fp32")


@pytest.mark.parametrize("prec", ["fp32", "split-bf16"])
def test_graph_replays_match_eager(handle, mfx, prec):
    rng = np.random.default_rng(3)
    keep = rand_problems(mfx, rng, 24, hi=400)
    fill(handle, keep, 7)
    C0 = [k[2].clone() for k in keep]
    probs = [mfx.problem(A, B, C, alpha=1.0, beta=0.5, transA=ta, transB=tb) for A, B, C, ta, tb, _ in keep]
    handle.grouped(probs, prec=prec)
    eager = [k[2].clone() for k in keep]
    graph = handle.graph_create(probs, prec=prec)
    try:
        for replay in range(3):
            for k, c0 in zip(keep, C0):
                k[2].copy_(c0)  # eager work between replays
            torch.cuda.synchronize()
            graph.launch()
            graph.status()
            for k, e in zip(keep, eager):
                assert torch.equal(k[2], e), replay
            # unrelated eager library work between replays
            x = torch.empty((128, 128), device=DEV)
            handle.fill_uniform(x, 1, 1)
            handle.gemm(x, x, x.clone(), prec=prec)
            torch.cuda.synchronize()
    finally:
        graph.close()


def test_autotune_single_and_table_roundtrip(handle, mfx, tmp_path):
    best = handle.autotune("split-bf16", 512, 512, 512)
    assert 0 <= best < mfx.config_count("split-bf16")
    path = str(tmp_path / "tune.json")
    lib = mfx.load()
    assert lib.mfx_autotune_save(path.encode()) == 0
    assert lib.mfx_autotune_load(path.encode()) == 0
    text = open(path).read()
    assert "mfx-autotune-2" in text and '"1:9:9:9": %d' % best in text
    # a tuned gemm still matches FP64
    A = torch.empty((512, 512), device=DEV)
    B = torch.empty((512, 512), device=DEV)
    handle.fill_uniform(A, 3, 0)
    handle.fill_uniform(B, 3, 1)
    C = torch.empty((512, 512), device=DEV)
    handle.gemm(A, B, C, prec="split-bf16")
    assert rel(C, A.double() @ B.double()) < 1e-5


def test_autotune_grouped(handle, mfx):
    rng = np.random.default_rng(4)
    keep = rand_problems(mfx, rng, 30, lo=64, hi=1024, transposes=False)
    fill(handle, keep, 8)
    probs = [mfx.problem(A, B, C) for A, B, C, ta, tb, _ in keep]
    p = handle.autotune_grouped(probs, prec="split-bf16")
    assert 0 <= p < mfx.grouped_policy_count()
    handle.grouped(probs, prec="split-bf16")
    for A, B, C, ta, tb, _ in keep:
        assert rel(C, A.double() @ B.double()) < 1e-5


@pytest.mark.parametrize("prec", ["fp32", "split-bf16", "bf16"])
def test_every_tile_policy(handle, mfx, prec):
    """Each grouped tile policy (pairs of tile configurations) against FP64,
    and bit-equal to single-problem launches on the tile each problem got."""
    rng = np.random.default_rng(11)
    keep = rand_problems(mfx, rng, 24, lo=16, hi=700)
    handle.force_splits(1)  # grouped tiles run whole: compare with unsplit single launches
    try:
        for pol in range(mfx.grouped_policy_count()):
            big, small, bound, per_level = mfx.grouped_policy_info(pol)
            fill(handle, keep, 6)
            C0 = [k[2].clone() for k in keep]
            probs = [mfx.problem(A, B, C, alpha=0.5, beta=2.0, transA=ta, transB=tb) for A, B, C, ta, tb, _ in keep]
            handle.force_grouped_policy(pol)
            handle.grouped(probs, prec=prec)
            handle.grouped_status()
            for (A, B, C, ta, tb, _), c0 in zip(keep, C0):
                err = rel(C, fp64(A, B, c0, 0.5, 2.0, ta, tb))
                assert err < TOL[prec], (pol, err)
            if big != small:
                got = [k[2].clone() for k in keep]
                fill(handle, keep, 6)
                for (A, B, C, ta, tb, _), g in zip(keep, got):
                    m, n = C.shape
                    lg, sm = big, small
                    if prec == "fp32":  # no grouped FP32 256x256 tile: 256x128 runs in its place
                        lg, sm = (3 if lg == 0 else lg), (3 if sm == 0 else sm)
                    bm, bn = mfx.config_info(prec, lg)[:2]
                    cfg = sm if -(-m // bm) * bm * -(-n // bn) * bn > bound * m * n else lg
                    handle.force_config(prec, cfg)
                    handle.gemm(A, B, C, alpha=0.5, beta=2.0, prec=prec, transA=ta, transB=tb)
                    handle.force_config(prec, -1)
                    assert torch.equal(C, g), (pol, cfg, m, n)
    finally:
        handle.force_grouped_policy(-1)
        handle.force_config(prec, -1)
        handle.force_splits(-1)


@pytest.mark.parametrize("prec", ["fp32", "split-bf16", "bf16"])
def test_grouped_odd_shapes_scalar_loads(handle, mfx, prec):
    """Extents that are not multiples of 4 force the bounds-checked scalar
    loads; K = 0 leaves beta * C."""
    shapes = [(33, 17, 45, False, False), (130, 257, 7, True, False), (5, 300, 129, False, True),
              (257, 255, 253, True, True), (64, 48, 0, False, False)]
    keep, probs = [], []
    for i, (m, n, k, ta, tb) in enumerate(shapes):
        A = torch.empty((k, m) if ta else (m, k), device=DEV)
        B = torch.empty((n, k) if tb else (k, n), device=DEV)
        C = torch.empty((m, n), device=DEV)
        for j, x in enumerate((A, B, C)):
            if x.numel():
                handle.fill_uniform(x, 9, 3 * i + j)
        keep.append((A, B, C, ta, tb, C.clone()))
        probs.append(mfx.problem(A, B, C, alpha=1.25, beta=-0.5, transA=ta, transB=tb))
    handle.grouped(probs, prec=prec)
    handle.grouped_status()
    for A, B, C, ta, tb, c0 in keep:
        want = fp64(A, B, c0, 1.25, -0.5, ta, tb)
        assert rel(C, want) < TOL[prec], (C.shape, A.shape)


@pytest.mark.parametrize("prec", ["fp32", "split-bf16"])
def test_chain_every_tile_policy(handle, mfx, prec):
    """Dependent levels under every policy, both one persistent launch with
    in-kernel waits and one launch per level."""
    n, count = 320, 8
    mats = [torch.empty((n, n), device=DEV) for _ in range(count)]
    for j, x in enumerate(mats):
        handle.fill_uniform(x, 13, j)
    want = chain_fp64(mats)
    try:
        for pol in range(mfx.grouped_policy_count()):
            scratch = [torch.empty((n, n), device=DEV) for _ in range(count - 2)]
            out = torch.empty((n, n), device=DEV)
            handle.force_grouped_policy(pol)
            handle.chain(mats, scratch, out, prec=prec)
            handle.grouped_status()
            assert rel(out, want) < TOL[prec] * 3, pol
    finally:
        handle.force_grouped_policy(-1)


@pytest.mark.parametrize("policy", [25, 29, 33])
@pytest.mark.parametrize("prec", ["split-bf16", "bf16"])
def test_direct_tiles_edges(handle, mfx, prec, policy):
    """Policies whose tiles read FP32 operands directly (no packing): ragged
    M / N (rows past the operand read its last row again), K = 16 and K = 0,
    padded leading dimensions, every transpose, plus problems the direct
    tile must not take (K % 16 != 0, a row-major operand whose row count is
    not a multiple of 4, a base address off 16 bytes), which fall back to
    packing -- all against FP64, and a consumer reading a producer's C."""
    shapes = [(300, 260, 64, False, False), (36, 44, 16, True, False), (260, 100, 48, False, True),
              (100, 36, 32, True, True), (64, 72, 0, False, False), (70, 66, 40, False, False),
              (33, 40, 24, False, False), (40, 35, 48, True, True), (37, 40, 32, True, False)]
    keep, probs = [], []
    for i, (m, n, k, ta, tb) in enumerate(shapes):
        pad = 4 * (i % 3)
        A = torch.empty(((k, m + pad) if ta else (m, k + pad)), device=DEV)[:, : (m if ta else k)]
        B = torch.empty(((n, k + pad) if tb else (k, n + pad)), device=DEV)[:, : (k if tb else n)]
        C = torch.empty((m, n), device=DEV)
        for j, x in enumerate((A, B, C)):
            if x.numel():
                handle.fill_uniform(x, 17, 3 * i + j)
        keep.append((A, B, C, ta, tb, C.clone()))
        probs.append(mfx.problem(A, B, C, alpha=0.75, beta=1.5, transA=ta, transB=tb))
    # a view that starts one float into its storage: not 16-byte aligned
    base = torch.empty((129, 97), device=DEV)
    handle.fill_uniform(base, 17, 90)
    A = base.view(-1)[1:1 + 128 * 96].view(128, 96)
    B = torch.empty((96, 64), device=DEV)
    handle.fill_uniform(B, 17, 91)
    C = torch.empty((128, 64), device=DEV)
    handle.fill_uniform(C, 17, 92)
    keep.append((A, B, C, False, False, C.clone()))
    probs.append(mfx.problem(A, B, C, alpha=0.75, beta=1.5))
    # consumer of problem 0's output: E = C0 (300 x 260) * F (260 x 48)
    F = torch.empty((260, 48), device=DEV)
    handle.fill_uniform(F, 17, 93)
    E = torch.empty((300, 48), device=DEV)
    probs.append(mfx.problem(keep[0][2], F, E))
    try:
        handle.force_grouped_policy(policy)
        handle.grouped(probs, deps=[(len(probs) - 1, 0)], prec=prec)
        handle.grouped_status()
    finally:
        handle.force_grouped_policy(-1)
    for A, B, C, ta, tb, c0 in keep:
        assert rel(C, fp64(A, B, c0, 0.75, 1.5, ta, tb)) < TOL[prec], (tuple(C.shape), ta, tb)
    assert rel(E, keep[0][2].double() @ F.double()) < TOL[prec]


@pytest.mark.parametrize("prec", ["fp32", "split-bf16", "bf16"])
def test_k0_consumer_waits_for_its_producer(handle, mfx, prec):
    """A K = 0 problem whose C is its producer's output (C = beta * C, no
    product): under every policy it must run after the producer, so the
    result is beta times the producer's product -- in-kernel waits included
    (its tiles have nothing to pack, but still wait)."""
    m, n, k = 300, 200, 96
    A = torch.empty((m, k), device=DEV)
    B = torch.empty((k, n), device=DEV)
    handle.fill_uniform(A, 31, 0)
    handle.fill_uniform(B, 31, 1)
    want = 0.5 * (A.double() @ B.double())
    try:
        for pol in range(mfx.grouped_policy_count()):
            X = torch.full((m, n), float("nan"), device=DEV)
            probs = [mfx.problem(A, B, X, alpha=1.0, beta=0.0),
                     mfx.problem(torch.zeros((m, 0), device=DEV), torch.zeros((0, n), device=DEV), X, beta=0.5)]
            handle.force_grouped_policy(pol)
            handle.grouped(probs, deps=[(1, 0)], prec=prec)
            handle.grouped_status()
            assert rel(X, want) < TOL[prec], pol
    finally:
        handle.force_grouped_policy(-1)
