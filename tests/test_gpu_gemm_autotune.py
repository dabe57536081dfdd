"""GEMM tile autotuner (tnh_gemm_autotune): each tuned size class keeps the size rule's
tile unless a timed candidate is at least 3% faster, the choice is reused, small or batched
products keep the size rule, and results are bitwise identical to the size rule's tile
(tiles differ only in the output block a workgroup computes, not in the order of the sum
over k)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

NCAND = {"f64": 3, "f32": 3, "bf16x3": 2}


def randn(rng, shape):
    return np.asfortranarray(rng.standard_normal(shape))


def run_gemm(ctx, prec, ops, A, B, C0, M, N, K):
    from tnhip import Tensor, ops as T

    dt = np.float64 if prec == "f64" else np.float32
    dA, dB, dC = (Tensor.from_numpy(x, dt, ctx=ctx) for x in (A, B, C0))
    T.gemm(ops[0], ops[1], M, N, K, 0.75, dA, A.shape[0], dB, B.shape[0], -0.5, dC, M, prec)
    return dC.numpy()


@pytest.mark.parametrize("prec", ["f64", "f32", "bf16x3"])
@pytest.mark.parametrize("ops", ["NN", "NT", "TN"])
def test_autotuned_tile_matches_rule_bitwise(ctx, rng, prec, ops):
    M, N, K = 1000, 900, 600  # 1.08 GFLOP: above the tuning threshold
    A = randn(rng, (K, M) if ops[0] == "T" else (M, K))
    B = randn(rng, (N, K) if ops[1] == "T" else (K, N))
    C0 = randn(rng, (M, N))
    try:
        ctx.gemm_autotune(False)
        ctx.gemm_autotune_reset()
        rule = run_gemm(ctx, prec, ops, A, B, C0, M, N, K)
        assert ctx.gemm_tuning() == []  # the size rule tunes nothing
        ctx.gemm_autotune(True)
        tuned = run_gemm(ctx, prec, ops, A, B, C0, M, N, K)
        table = ctx.gemm_tuning()
        assert len(table) == 1
        e = table[0]
        assert (e["mode"], e["trans"], e["M"], e["N"], e["K"]) == (prec, ops, M, N, K if prec != "bf16x3" else 608)
        assert len(e["ms"]) == NCAND[prec] and all(t > 0 for t in e["ms"])
        shapes = {"f64": ["64x64", "128x64", "128x128"], "f32": ["64x64", "128x128", "128x64"],
                  "bf16x3": ["64x64", "128x128"]}[prec]
        # the rule's tile is kept unless the fastest candidate beats it by at least 3%
        # (the report rounds to 1e-4 ms)
        ti, ri = shapes.index(e["tile"]), shapes.index(e["rule"])
        if ti == ri:
            assert min(e["ms"]) >= 0.97 * e["ms"][ri] - 2e-4
        else:
            assert e["ms"][ti] == min(e["ms"]) and e["ms"][ti] <= 0.97 * e["ms"][ri] + 2e-4
        again = run_gemm(ctx, prec, ops, A, B, C0, M, N, K)
        assert len(ctx.gemm_tuning()) == 1  # reused, not re-timed
        assert np.array_equal(tuned, rule)
        assert np.array_equal(again, rule)
        # a product of the same size class (M = 992 and 1000 share one) reuses its tile
        M2 = 992
        A2 = A[:, :M2] if ops[0] == "T" else A[:M2, :]
        run_gemm(ctx, prec, ops, A2, B, C0[:M2, :], M2, N, K)
        assert len(ctx.gemm_tuning()) == 1
    finally:
        ctx.gemm_autotune(True)
        ctx.gemm_autotune_reset()


def test_small_products_keep_the_rule(ctx, rng):
    try:
        ctx.gemm_autotune(True)
        ctx.gemm_autotune_reset()
        for (M, N, K) in [(300, 260, 129), (90, 4000, 4000), (4000, 90, 4000)]:
            A, B, C0 = randn(rng, (M, K)), randn(rng, (K, N)), randn(rng, (M, N))
            out = run_gemm(ctx, "f64", "NN", A, B, C0, M, N, K)
            ref = 0.75 * A @ B - 0.5 * C0
            assert np.linalg.norm(out - ref) <= 1e-13 * np.linalg.norm(ref)
        assert ctx.gemm_tuning() == []
    finally:
        ctx.gemm_autotune_reset()


@pytest.mark.parametrize("prec", ["f64", "f32", "bf16x3"])
def test_matvec_autotuned_equals_rule(ctx, rng, prec):
    """The H_eff matvec at chi = 256 (its two large products are tuned) is bitwise the
    same with tuned tiles as with the size rule's tiles."""
    from tnhip import Tensor, ops as T
    from test_gpu_contract import heisenberg_w, random_env

    chi = 256
    W = heisenberg_w()
    L, R = random_env(rng, chi, 5), random_env(rng, chi, 5)
    x = randn(rng, (chi, 2, 2, chi))
    dt = np.float64 if prec == "f64" else np.float32
    dL, dR = Tensor.from_numpy(L, dt, ctx=ctx), Tensor.from_numpy(R, dt, ctx=ctx)
    dW, dx = Tensor.from_numpy(W, ctx=ctx), Tensor.from_numpy(x, ctx=ctx)
    try:
        ctx.gemm_autotune(False)
        rule = T.heff_apply(dL, dW, dW, dR, dx, precision=prec).numpy()
        ctx.gemm_autotune(True)
        ctx.gemm_autotune_reset()
        tuned = T.heff_apply(dL, dW, dW, dR, dx, precision=prec).numpy()
        assert len(ctx.gemm_tuning()) >= 1
        assert np.array_equal(tuned, rule)
    finally:
        ctx.gemm_autotune(True)
        ctx.gemm_autotune_reset()


def test_tuning_report_truncation_and_bad_mode(ctx, rng):
    """tnh_gemm_tuning_report writes at most len bytes (NUL-terminated) and returns the full
    size; tnh_gemm_autotune rejects modes other than 0, 1, 2."""
    import ctypes as C

    import tnhip
    from tnhip._lib import lib

    try:
        ctx.gemm_autotune(True)
        ctx.gemm_autotune_reset()
        M, N, K = 1000, 900, 600
        run_gemm(ctx, "f32", "NN", randn(rng, (M, K)), randn(rng, (K, N)), randn(rng, (M, N)), M, N, K)
        need = C.c_int64()
        assert lib.tnh_gemm_tuning_report(ctx.h, None, 0, C.byref(need)) == 0
        full = C.create_string_buffer(int(need.value))
        assert lib.tnh_gemm_tuning_report(ctx.h, full, int(need.value), C.byref(need)) == 0
        assert len(full.value) + 1 == need.value and full.value.startswith(b"f32 NN M=1000 N=900 K=600")
        short = C.create_string_buffer(8)
        assert lib.tnh_gemm_tuning_report(ctx.h, short, 8, C.byref(need)) == 0
        assert short.value == full.value[:7] and need.value == len(full.value) + 1
        with pytest.raises(tnhip.TnhError):
            ctx.call("tnh_gemm_autotune", 3)
    finally:
        ctx.gemm_autotune(True)
        ctx.gemm_autotune_reset()
