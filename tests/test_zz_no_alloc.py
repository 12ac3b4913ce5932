"""The allocation-free inertia entry point (tm_operator_inertia_noalloc, factor workspace in
the library's static device scratch) against tm_operator_inertia (workspace from the
context's pool): the same kernels on the same inputs, so the outputs are identical.  Two
calls in a row, the second on a smaller mesh, so the second runs on what the first left in
the scratch; a mesh whose workspace exceeds the scratch is refused."""
import numpy as np
import pytest

from common import engine_mesh, fixture_mesh, generated, oracle_model

pytestmark = pytest.mark.gpu


def _operator(ctx, which):
    from turbomodal_amd import capi
    if which == "c1":
        X, T, S, N = generated("flat_disk_c1")
    elif which == "c2":
        X, T, S, N = generated("bladed_c2")
    else:
        X, T, S, N = fixture_mesh(which)
    m = oracle_model(X, T, S, N)
    em = engine_mesh(ctx, m)
    return em, capi.EngineOperator(em, 200e9, 0.3, 7800.0, True)


def test_noalloc_inertia_equals_pooled():
    from turbomodal_amd import capi
    ctx = capi.Context()
    try:
        ops = [_operator(ctx, "c1"), _operator(ctx, "wedge_sector.msh")]
        for (em, op), (rpm, k, shift) in zip(ops, ((6000.0, 3, 2.0e8), (3000.0, 1, 1.0e9))):
            print(em.info()["front_bytes_fp64"], flush=True)
            a = op.inertia(rpm, k, shift)
            b = op.inertia(rpm, k, shift, noalloc=True)
            assert a == b, (a, b)
            assert a[0] > 0  # shifts inside the spectrum: some eigenvalues below
            # the shift at 1 is below the spectrum: no negative pivot, in either entry point
            assert op.inertia(rpm, k, 1.0, noalloc=True)[0] == 0
        # the 45k-DOF sector's workspace exceeds the scratch: refused, nothing run
        em2, op2 = _operator(ctx, "c2")
        with pytest.raises(ValueError):
            op2.inertia(0.0, 0, 1.0, noalloc=True)
        for em, op in ops + [(em2, op2)]:
            op.close()
            em.close()
    finally:
        ctx.close()
