"""Captured launches (hipGraph replay) of the block method against direct launches.

The batched factorization and the block method's triangular sweeps are captured
once per slot group and argument set and replayed for every later wave and block
step (TMHIP_GRAPHS=0 launches them directly).  A sweep of 4 rpm points x 19
harmonics on the flat disk runs three waves per slot group (rpm continuation), so
every captured factorization is replayed with the program's own launches, copies
and synchronisations in between, and every block step replays the sweep graph;
each replay's results must be bit-identical to the direct launches'.  The same for
the complex-FP32 factor (TM_PREC_MIXED) and the Coriolis pencil (2n sweeps).
"""
import os

import numpy as np
import pytest

from common import engine_mesh, generated, oracle_model
from oracle import cyclic

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def disk():
    from turbomodal_amd import capi
    ctx = capi.Context()
    X, T, S, N = generated("flat_disk_c1")
    m = oracle_model(X, T, S, N)
    em = engine_mesh(ctx, m)
    op = capi.EngineOperator(em, 200e9, 0.3, 7800.0, True)
    yield ctx, m, op
    op.close()
    em.close()
    ctx.close()


def _solve(ctx, probs, graphs):
    from turbomodal_amd import capi
    old = os.environ.get("TMHIP_GRAPHS")
    os.environ["TMHIP_GRAPHS"] = "1" if graphs else "0"
    try:
        return capi.solve_batch(ctx, probs)
    finally:
        if old is None:
            del os.environ["TMHIP_GRAPHS"]
        else:
            os.environ["TMHIP_GRAPHS"] = old


def _same(a, b):
    for x, y in zip(a, b):
        assert x.status == y.status == 0, (x.rpm, x.harmonic)
        assert np.array_equal(x.frequencies, y.frequencies), (x.rpm, x.harmonic)
        assert np.array_equal(x.residuals, y.residuals), (x.rpm, x.harmonic)
        assert np.array_equal(x.mode_shapes, y.mode_shapes), (x.rpm, x.harmonic)


@pytest.mark.parametrize("variant", ["fp64", "mixed", "coriolis"])
def test_replayed_equals_direct(disk, variant):
    from turbomodal_amd import capi
    ctx, m, op = disk
    N = m.num_sectors
    if variant == "coriolis":
        probs = [dict(op=op, rpm=6000.0, harmonic=k, num_modes=6, coriolis=True) for k in range(1, N // 2 + 1)]
    else:
        prec = capi.TM_PREC_MIXED if variant == "mixed" else capi.TM_PREC_FP64
        probs = [dict(op=op, rpm=float(r), harmonic=k, num_modes=5, precision=prec, warm_start=True)
                 for r in (0.0, 4000.0, 8000.0, 12000.0) for k in range(N // 2 + 1)]
    direct = _solve(ctx, probs, False)
    # twice with graphs: the second call replays what the first one captured where
    # the slot groups' buffers come back at the same addresses
    g1 = _solve(ctx, probs, True)
    g2 = _solve(ctx, probs, True)
    _same(direct, g1)
    _same(direct, g2)
    # and the replayed path is the oracle's at one spot check
    o = g2[3]
    ops = cyclic.sector_operators(m, 200e9, 0.3, 7800.0, o.rpm)
    if variant == "coriolis":
        ref = cyclic.solve_harmonic_qep(m, ops, o.harmonic, 6, o.rpm)
    else:
        ref = cyclic.solve_harmonic(m, ops, o.harmonic, 5, o.rpm, "tight")
    assert np.max(np.abs(o.frequencies / ref.frequencies - 1)) <= 1e-8
