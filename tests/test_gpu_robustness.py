"""Failure handling and capability parity on the GPU.

* A shift placed exactly on an eigenvalue: the LDL^H flags a near-singular pivot, the
  problem is solved again alone with the shift moved (the reference falls back from LDL^H
  to LU, src/hermitian_lanczos.cpp:27-55, and retries failed harmonics,
  src/cyclic_solver.cpp:1334-1342), and the result is the oracle's.
* Grounded springs (ELASTIC_SUPPORT, src/cyclic_solver.cpp:193-212, 747-751) through the
  C ABI and through the drop-in ConstraintGroup, against the oracle with the same springs.
* Sylvester inertia of K(k) - sigma M(k) against the dense spectrum.
"""
import numpy as np
import pytest
import scipy.linalg as sla

from common import engine_mesh, fixture_mesh, generated, oracle_model
from oracle import cyclic

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    from turbomodal_amd import capi
    c = capi.Context()
    yield c
    c.close()


def _dense(op, rpm, k, n):
    from common import exported
    K = exported(op, rpm, k, 0, n).toarray()
    M = exported(op, rpm, k, 1, n).toarray()
    return sla.eigh(0.5 * (K + K.conj().T), 0.5 * (M + M.conj().T), eigvals_only=True)


@pytest.mark.parametrize("case", ["wedge", "c1"])
def test_shift_on_an_eigenvalue(ctx, case):
    # The LDL^H of K(k) - lambda_3 M(k) has a near-singular pivot: the inertia call flags
    # it, the solve at that shift is flagged and solved again at a moved shift, and it
    # returns the oracle's lowest pairs.  wedge: < 256 free DOFs, the Lanczos method;
    # c1: the block method.
    from turbomodal_amd import capi
    if case == "wedge":
        X, T, S, N = fixture_mesh("wedge_sector.msh")
    else:
        X, T, S, N = generated("flat_disk_c1")
    m = oracle_model(X, T, S, N)
    em = engine_mesh(ctx, m)
    op = capi.EngineOperator(em, 200e9, 0.3, 7800.0, True)
    rpm = 3000.0
    ops = cyclic.sector_operators(m, 200e9, 0.3, 7800.0, rpm)
    for k in (0, 1):
        ref = cyclic.solve_harmonic(m, ops, k, 4, rpm, "tight")
        sigma = float(ref.eigenvalues[min(2, len(ref.eigenvalues) - 1)])
        neg, ratio, st = op.inertia(rpm, k, sigma)
        assert st == 2 and ratio <= 1e-10, (case, k, st, ratio)  # near-singular, not zero
        assert neg in (len(ref.eigenvalues[ref.eigenvalues < sigma]), 1 + len(ref.eigenvalues[ref.eigenvalues < sigma]))
        o = capi.solve_batch(ctx, [dict(op=op, rpm=rpm, harmonic=k, num_modes=4, shift=sigma)])[0]
        assert o.status == 0, (case, k)
        assert o.factor_retries >= 1 and o.pivot_flags == 0 and o.shift_used < sigma, \
            (o.factor_retries, o.pivot_flags, o.shift_used)
        assert np.max(np.abs(o.frequencies / ref.frequencies - 1)) <= 1e-8, (case, k, o.frequencies, ref.frequencies)
        for q in range(len(o.frequencies)):
            assert cyclic.compute_mac(o.mode_shapes[:, q], ref.mode_shapes[:, q])[0, 0] >= 0.999
        # a shift below the spectrum is clean and is not moved
        o2 = capi.solve_batch(ctx, [dict(op=op, rpm=rpm, harmonic=k, num_modes=4)])[0]
        assert o2.status == 0 and o2.factor_retries == 0 and o2.pivot_flags == 0 and o2.shift_used == 1.0
    op.close()
    em.close()


def test_inertia_counts_eigenvalues_below_the_shift(ctx):
    from turbomodal_amd import capi
    X, T, S, N = generated("flat_disk_c1")
    m = oracle_model(X, T, S, N)
    em = engine_mesh(ctx, m)
    op = capi.EngineOperator(em, 200e9, 0.3, 7800.0, True)
    n = em.info()["n_free"]
    for rpm, k in ((0.0, 0), (8000.0, 5), (8000.0, 18)):
        w = _dense(op, rpm, k, n)
        for j in (0, 3, 11):
            sigma = 0.5 * (w[j] + w[j + 1]) if j else 0.5 * w[0]
            neg, ratio, st = op.inertia(rpm, k, sigma)
            assert neg == (j + 1 if j else 0), (rpm, k, j, neg)
            assert st == 0 and ratio > 1e-10
    op.close()
    em.close()


def _springs(X, m):
    k = np.zeros(3 * len(X))
    for i in sorted(set(range(0, len(X), 3)) | set(m.left) | set(m.right)):
        k[3 * i:3 * i + 3] = (3e8, 1e8, 2e8)
    return k


def test_springs_match_oracle(ctx):
    from turbomodal_amd import capi
    X, T, S, N = fixture_mesh("wedge_sector.msh")
    m = oracle_model(X, T, S, N)
    em = engine_mesh(ctx, m)
    op = capi.EngineOperator(em, 200e9, 0.3, 7800.0, True)
    ks = _springs(X, m)
    op.add_springs(ks)
    rpm = 3000.0
    ops = cyclic.sector_operators(m, 200e9, 0.3, 7800.0, rpm, springs=ks)
    n = em.info()["n_free"]
    import scipy.sparse as sp
    for k in (0, 1, N // 2):
        r, c, v = op.export(rpm, k, 0)
        A = sp.csc_matrix((v, (r, c)), shape=(n, n)).toarray()
        B = cyclic.harmonic_matrix(ops.Kc, ops.Kp, k, N).toarray()
        assert np.abs(A - B).max() <= 1e-13 * np.abs(B).max(), k
    outs = capi.solve_batch(ctx, [dict(op=op, rpm=rpm, harmonic=k, num_modes=5) for k in range(N // 2 + 1)])
    for o in outs:
        assert o.status == 0
        ref = cyclic.solve_harmonic(m, ops, o.harmonic, 5, rpm, "tight")
        assert np.max(np.abs(o.frequencies / ref.frequencies - 1)) <= 1e-8, o.harmonic
        for q in range(len(o.frequencies)):
            assert cyclic.compute_mac(o.mode_shapes[:, q], ref.mode_shapes[:, q])[0, 0] >= 0.999
    op.close()
    em.close()


def test_springs_through_the_dropin_constraint_groups():
    import turbomodal_amd as tm
    from common import GOLDEN
    import os
    mesh = tm.load_mesh(os.path.join(GOLDEN, "wedge_sector.msh"), 24)
    mat = tm.Material(200e9, 0.3, 7800.0)
    X, T, S, N = fixture_mesh("wedge_sector.msh")
    m = oracle_model(X, T, S, N)
    ks = _springs(X, m)
    hub = tm.ConstraintGroup()
    hub.name = "hub"
    hub.type = tm.BCType.FIXED
    hub.node_ids = list(S["hub_constraint"])
    spring = tm.ConstraintGroup()
    spring.name = "springs"
    spring.type = tm.BCType.ELASTIC_SUPPORT
    spring.node_ids = [int(i) for i in range(len(X)) if ks[3 * i] > 0]
    spring.spring_stiffness = [3e8, 1e8, 2e8]
    solver = tm.CyclicSymmetrySolver(mesh, mat, [hub, spring])
    res = solver.solve_at_rpm(3000.0, 5, [0, 3, 12])
    ops = cyclic.sector_operators(m, 200e9, 0.3, 7800.0, 3000.0, springs=ks)
    for r in res:
        ref = cyclic.solve_harmonic(m, ops, r.harmonic_index, 5, 3000.0, "tight")
        assert np.max(np.abs(np.asarray(r.frequencies) / ref.frequencies - 1)) <= 1e-8


def test_many_modes_block_and_coriolis(ctx):
    # more wanted pairs than the default bases hold: the block method takes a longer basis
    # (k = 0: 60 pairs, basis 160; k = N/2: 100 pairs, basis 240), and so does the Coriolis
    # pencil (31 modes: basis 112), against the oracle (the reference requests
    # min(6 M, 2 n - 1) QEP eigenvalues, src/modal_solver.cpp:222-229, and keeps M)
    from turbomodal_amd import capi
    X, T, S, N = generated("flat_disk_c1")
    m = oracle_model(X, T, S, N)
    em = engine_mesh(ctx, m)
    op = capi.EngineOperator(em, 200e9, 0.3, 7800.0, True)
    rpm = 6000.0
    ops = cyclic.sector_operators(m, 200e9, 0.3, 7800.0, rpm)
    outs = capi.solve_batch(ctx, [dict(op=op, rpm=rpm, harmonic=0, num_modes=60),
                                  dict(op=op, rpm=rpm, harmonic=N // 2, num_modes=100)])
    for o, nm in zip(outs, (60, 100)):
        assert o.status == 0 and o.converged and len(o.frequencies) == nm
        ref = cyclic.solve_harmonic(m, ops, o.harmonic, nm, rpm, "tight")
        assert np.max(np.abs(o.frequencies / ref.frequencies - 1)) <= 1e-8, o.harmonic
        neg, ratio, st = op.inertia(rpm, o.harmonic, float(o.eigenvalues[nm - 1]) * (1 + 1e-7))
        assert neg == nm
    op.close()
    em.close()
    X, T, S, N = fixture_mesh("blade_sector.msh")
    m = oracle_model(X, T, S, N)
    em = engine_mesh(ctx, m)
    op = capi.EngineOperator(em, 200e9, 0.3, 7800.0, True)
    rpm = 20000.0
    ops = cyclic.sector_operators(m, 200e9, 0.3, 7800.0, rpm)
    q = capi.solve_batch(ctx, [dict(op=op, rpm=rpm, harmonic=2, num_modes=31, coriolis=True)])[0]
    assert q.status == 0 and len(q.frequencies) == 31
    ref = cyclic.solve_harmonic_qep(m, ops, 2, 31, rpm)
    assert np.max(np.abs(q.frequencies / ref.frequencies - 1)) <= 1e-8
    assert np.array_equal(q.whirl, ref.whirl_direction)
    op.close()
    em.close()
