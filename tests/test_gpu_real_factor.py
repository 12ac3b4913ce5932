"""The real factor of the real harmonics (k = 0 and N/2) against the complex one.

At k = 0 and N/2 the phase is exactly +-1 and K(k) - sigma M(k) is real symmetric (the
reference's real path, src/cyclic_solver.cpp:673-723, 1019-1043): the block method then
factors it in real FP64 (8-byte entries, one real MFMA product per block instead of
four).  TMHIP_REAL_FACTOR=0 runs the same problems through the complex factor.  The
two factors differ at rounding level only (the complex arithmetic adds exact zeros; the
real kernels may fuse multiplies and adds where the complex ones round in between), and
the factor only preconditions the Krylov method, whose Rayleigh-Ritz
runs on the exact K(k), M(k): the results agree to well inside the oracle's bar (relative
frequency difference <= 1e-9 against the oracle's 1e-8), and both meet the oracle's bar.
"""
import os

import numpy as np
import pytest

from common import engine_mesh, generated, oracle_model
from oracle import cyclic

pytestmark = pytest.mark.gpu


def _solve(ctx, probs, real):
    from turbomodal_amd import capi
    old = os.environ.get("TMHIP_REAL_FACTOR")
    os.environ["TMHIP_REAL_FACTOR"] = "1" if real else "0"
    try:
        return capi.solve_batch(ctx, probs)
    finally:
        if old is None:
            del os.environ["TMHIP_REAL_FACTOR"]
        else:
            os.environ["TMHIP_REAL_FACTOR"] = old


@pytest.mark.parametrize("case", ["flat_disk_c1", "bladed_c2"])
def test_real_factor_matches_complex_and_oracle(case):
    from turbomodal_amd import capi
    ctx = capi.Context()
    X, T, S, N = generated(case)
    m = oracle_model(X, T, S, N)
    em = engine_mesh(ctx, m)
    op = capi.EngineOperator(em, 200e9, 0.3, 7800.0, True)
    rpms = (0.0, 9000.0) if case == "flat_disk_c1" else (10263.0,)
    nm = 5 if case == "flat_disk_c1" else 10
    probs = [dict(op=op, rpm=r, harmonic=k, num_modes=nm) for r in rpms for k in (0, N // 2)]
    # with a complex harmonic in the same call: it keeps the complex factor
    probs.append(dict(op=op, rpm=rpms[-1], harmonic=1, num_modes=nm))
    re = _solve(ctx, probs, True)
    cx = _solve(ctx, probs, False)
    worst = 0.0
    for a, b in zip(re, cx):
        assert a.status == b.status == 0, (a.rpm, a.harmonic)
        d = np.max(np.abs(a.frequencies / b.frequencies - 1))
        worst = max(worst, d)
        assert d <= 1e-9, (a.rpm, a.harmonic, d)
        for q in range(len(a.frequencies)):
            assert cyclic.compute_mac(a.mode_shapes[:, q], b.mode_shapes[:, q])[0, 0] >= 1 - 1e-9
        assert np.all(a.residuals <= np.maximum(4e-8, 4 * b.residuals))
    # the complex harmonic is untouched: bit-identical either way
    assert np.array_equal(re[-1].frequencies, cx[-1].frequencies)
    assert np.array_equal(re[-1].mode_shapes, cx[-1].mode_shapes)
    if case == "flat_disk_c1":
        for o in re[:-1]:
            ops = cyclic.sector_operators(m, 200e9, 0.3, 7800.0, o.rpm)
            ref = cyclic.solve_harmonic(m, ops, o.harmonic, nm, o.rpm, "tight")
            assert np.max(np.abs(o.frequencies / ref.frequencies - 1)) <= 1e-8, (o.rpm, o.harmonic)
            for q in range(nm):
                assert cyclic.compute_mac(o.mode_shapes[:, q], ref.mode_shapes[:, q])[0, 0] >= 0.999
    print(f"\n{case}: real vs complex factor, largest relative frequency difference {worst:.2e}")
    op.close()
    em.close()
    ctx.close()
