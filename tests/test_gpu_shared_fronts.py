"""Shared fronts of the factorization.

A front whose subtree assembles no phase-coupled slot has a real A(k) that is the same
for every harmonic at one (operator, rpm, shift); each such class of problems factors
those fronts once (by its first problem) and the other members receive the result.
The results must equal those of every problem factoring every front (TMHIP_SHARE=0)
bit for bit, and the factorization's work must fall.
"""
import os

import numpy as np
import pytest

from common import engine_mesh, generated, oracle_model

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    from turbomodal_amd import capi
    c = capi.Context()
    yield c
    c.close()


def _solve(ctx, probs, share):
    from turbomodal_amd import capi
    old = os.environ.pop("TMHIP_SHARE", None)
    try:
        if not share:
            os.environ["TMHIP_SHARE"] = "0"
        ctx.set_profiling(True)
        ctx.reset_kernel_stats()
        out = capi.solve_batch(ctx, probs)
        stats = ctx.kernel_stats()
        ctx.set_profiling(False)
    finally:
        os.environ.pop("TMHIP_SHARE", None)
        if old is not None:
            os.environ["TMHIP_SHARE"] = old
    return out, stats


@pytest.mark.parametrize("case,rpms,ks,method", [
    ("flat_disk_c1", (0.0, 5000.0, 12000.0), None, "block"),
    ("flat_disk_c1", (0.0, 5000.0), None, "lanczos"),
    ("bladed_c2", (3000.0, 10263.0), (0, 1, 5, 9, 13, 17, 18), "block"),
])
def test_shared_fronts_bit_identical(ctx, case, rpms, ks, method):
    from turbomodal_amd import capi, meshgen
    X, T, S, N = generated(case)
    m = oracle_model(X, T, S, N)
    em = engine_mesh(ctx, m)
    op = capi.EngineOperator(em, **meshgen.STEEL, prestress=True)
    ks = ks if ks is not None else range(N // 2 + 1)
    meth = capi.TM_METHOD_BLOCK if method == "block" else capi.TM_METHOD_LANCZOS
    probs = [dict(op=op, rpm=r, harmonic=k, num_modes=10, method=meth) for r in rpms for k in ks]
    print(f"[shared] {case} {method}: {len(probs)} problems", flush=True)
    shared, s1 = _solve(ctx, probs, True)
    alone, s0 = _solve(ctx, probs, False)
    for a, b in zip(shared, alone):
        assert a.status == 0 and b.status == 0
        assert np.array_equal(a.frequencies, b.frequencies)
        assert np.array_equal(a.eigenvalues, b.eigenvalues)
        assert np.array_equal(a.mode_shapes, b.mode_shapes)
    # the complex harmonics' factorizations did less work
    f1, f0 = s1["factor"]["alg_flops"], s0["factor"]["alg_flops"]
    print(f"[shared] {case} {method}: factor flops {f1:.3e} shared, {f0:.3e} every front", flush=True)
    assert f1 < 0.95 * f0
    op.close()
    em.close()
