"""Continuous-batching server (tm_server_*): results are the same alone and in a mix of
uneven requests that arrive while others are on the GPU, and they match the batch path
and the oracle.  Bars as in test_gpu_parity: frequency 1e-8, MAC 0.999."""
import threading
import time

import numpy as np
import pytest

from common import engine_mesh, generated, oracle_model
from oracle import cyclic

pytestmark = pytest.mark.gpu

STEEL = (200e9, 0.3, 7800.0)
TITANIUM = (114e9, 0.34, 4430.0)


@pytest.fixture(scope="module")
def setup():
    from turbomodal_amd import capi
    ctx = capi.Context()
    models, meshes, ops = {}, {}, {}
    for N in (36, 24):
        X, T, S, _ = generated("flat_disk_c1", N)
        m = oracle_model(X, T, S, N)
        models[N] = m
        meshes[N] = engine_mesh(ctx, m)
        for name, mat in (("steel", STEEL), ("ti", TITANIUM)):
            ops[(N, name)] = capi.EngineOperator(meshes[N], *mat, True)
    yield capi, ctx, models, ops
    for o in ops.values():
        o.close()
    for m in meshes.values():
        m.close()
    ctx.close()


def _target(ops):
    return dict(op=ops[(36, "steel")], mat="steel", rpm=7000.0, harmonic=5, num_modes=10)


def test_alone_equals_mixed_and_batch(setup):
    capi, ctx, models, ops = setup
    tgt = _target(ops)
    with capi.Server(ctx, max_slots=16) as srv:
        alone = srv.wait(srv.submit(tgt))
    rng = np.random.default_rng(7)
    mix = []
    for i in range(24):
        N = (36, 24)[i % 2]
        mat = ("steel", "ti")[(i // 2) % 2]
        mix.append(dict(op=ops[(N, mat)], mat=mat, rpm=float(rng.uniform(0, 15000)),
                        harmonic=int(rng.integers(0, N // 2 + 1)), num_modes=int(rng.choice([5, 10, 20]))))
    with capi.Server(ctx, max_slots=16) as srv:
        early = [srv.submit(p) for p in mix[:12]]
        time.sleep(0.3)  # the first requests are on the GPU now
        tt = srv.submit(tgt)
        late = [srv.submit(p) for p in mix[12:]]
        outs = {t: srv.wait(t) for t in early + late}
        mixed = srv.wait(tt)
    assert alone.status == 0 and mixed.status == 0
    assert np.array_equal(alone.frequencies, mixed.frequencies)
    assert np.array_equal(alone.mode_shapes, mixed.mode_shapes)
    assert np.array_equal(alone.residuals, mixed.residuals)
    # the batch path gives the same result
    batch = capi.solve_batch(ctx, [tgt])[0]
    assert np.array_equal(batch.frequencies, alone.frequencies)
    assert np.array_equal(batch.mode_shapes, alone.mode_shapes)
    # every request finished; a sample against the oracle
    assert all(o.status == 0 for o in outs.values())
    for p, t in list(zip(mix, early + late))[:4] + [(tgt, None)]:
        o = outs[t] if t is not None else mixed
        N = p["op"].mesh.num_sectors
        m = models[N]
        E, nu, rho = STEEL if p["mat"] == "steel" else TITANIUM
        ref_ops = cyclic.sector_operators(m, E, nu, rho, p["rpm"])
        ref = cyclic.solve_harmonic(m, ref_ops, p["harmonic"], p["num_modes"], p["rpm"], "tight")
        assert len(o.frequencies) == len(ref.frequencies)
        assert np.max(np.abs(o.frequencies / ref.frequencies - 1)) <= 1e-8
        for q in range(len(o.frequencies)):
            assert cyclic.compute_mac(o.mode_shapes[:, q], ref.mode_shapes[:, q])[0, 0] >= 0.999


def test_concurrent_submitters_and_timeout(setup):
    capi, ctx, models, ops = setup
    with capi.Server(ctx, max_slots=8) as srv:
        tickets, lock = [], threading.Lock()

        def sub(seed):
            r = np.random.default_rng(seed)
            for _ in range(4):
                t = srv.submit(dict(op=ops[(24, "steel")], rpm=float(r.uniform(0, 9000)),
                                    harmonic=int(r.integers(0, 13)), num_modes=5))
                with lock:
                    tickets.append(t)

        th = [threading.Thread(target=sub, args=(s,)) for s in range(3)]
        for x in th:
            x.start()
        for x in th:
            x.join()
        r0 = srv.wait(tickets[-1], timeout=0.0)  # may or may not be done yet
        assert r0 is None or r0.status == 0
        res = [srv.wait(t) for t in tickets[:-1]]
        assert all(r is not None and r.status == 0 for r in res)
    with pytest.raises(ValueError):
        with capi.Server(ctx) as srv:
            srv.submit(dict(op=ops[(24, "steel")], rpm=0.0, harmonic=13, num_modes=5))
