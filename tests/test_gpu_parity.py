"""GPU parity: the HIP engine (through the C ABI) against the CPU oracle.

Bars (from the workload definition): relative frequency error <= 1e-8,
MAC >= 0.999 per mode, ||K phi - lambda M phi|| / ||K phi|| <= 1e-8, all
against the FP64 oracle solved to a relative Ritz residual of 1e-12.
Assembled matrices are compared entry-wise (rounding-level differences only:
the engine sums contributions in a different, fixed order).
"""
import numpy as np
import pytest
import scipy.sparse as sp

from common import engine_mesh, fixture_mesh, generated, oracle_model
from oracle import cyclic

pytestmark = pytest.mark.gpu

FREQ_TOL = 1e-8
MAC_TOL = 0.999
RES_TOL = 1e-8
EPS = np.finfo(float).eps


def free_vector(m, phi):
    """Full-sector mode shape -> the reference's free-DOF vector (interior DOFs,
    then left DOFs, constrained removed)."""
    red = np.array(m.interior_dofs + m.left_dofs)
    return phi[red[cyclic.reduced_free_map(m)]]


def residual_and_floor(K, M, x, lam):
    """||K x - lam M x|| / ||K x|| and its FP64 floor eps || |K||x| + lam |M||x| || / ||K x||:
    the residual of the exact eigenvector rounded to FP64 is of the size of the floor, so
    the floor, not 1e-8, bounds what any FP64 solver can return when lambda_max/lambda is
    large (thin blades reach lambda_max/lambda_1 ~ 5e8)."""
    Kx = K @ x
    r = Kx - lam * (M @ x)
    nk = np.linalg.norm(Kx)
    fl = EPS * np.linalg.norm(abs(K) @ np.abs(x) + lam * (abs(M) @ np.abs(x))) / nk
    return np.linalg.norm(r) / nk, fl


@pytest.fixture(scope="module")
def ctx():
    from turbomodal_amd import capi
    c = capi.Context()
    yield c
    c.close()


def _case(name):
    if name == "wedge":
        X, T, S, N = fixture_mesh("wedge_sector.msh")
        return X, T, S, N, 3000.0
    if name == "blade_fixture":
        X, T, S, N = fixture_mesh("blade_sector.msh")
        return X, T, S, N, 8000.0
    X, T, S, N = generated("flat_disk_c1")
    return X, T, S, N, 5000.0


@pytest.mark.parametrize("case", ["wedge", "blade_fixture", "c1"])
def test_assembly_and_static_prestress(ctx, case):
    from turbomodal_amd import capi
    X, T, S, N, rpm = _case(case)
    m = oracle_model(X, T, S, N)
    em = engine_mesh(ctx, m)
    op = capi.EngineOperator(em, 200e9, 0.3, 7800.0, True)
    ops = cyclic.sector_operators(m, 200e9, 0.3, 7800.0, rpm)
    for k in (0, 1, N // 2):
        for which, (Ac, Ap) in ((0, (ops.Kc, ops.Kp)), (1, (ops.Mc, ops.Mp))):
            r, c, v = op.export(rpm, k, which)
            A = sp.csc_matrix((v, (r, c)), shape=Ac.shape).toarray()
            B = cyclic.harmonic_matrix(Ac, Ap, k, N).toarray()
            assert np.abs(A - B).max() <= 1e-13 * np.abs(B).max(), (case, k, which)
    # two backward-stable direct solves (supernodal LDL^T vs SuperLU) agree to
    # kappa(K) * eps; kappa of the thin blade sector is ~1e8
    w2 = (rpm * 2 * np.pi / 60) ** 2
    u = op.static_u()
    assert np.linalg.norm(w2 * u - ops.u_static) <= 1e-8 * np.linalg.norm(ops.u_static)
    op.close()
    em.close()


def _check_against_oracle(m, ops, outs, num_modes, rpm):
    for o in outs:
        assert o.status == 0, o.harmonic
        ref = cyclic.solve_harmonic(m, ops, o.harmonic, num_modes, rpm, "tight")
        assert len(o.frequencies) == len(ref.frequencies)
        assert np.array_equal(o.whirl, ref.whirl_direction)
        err = np.max(np.abs(o.frequencies / ref.frequencies - 1))
        assert err <= FREQ_TOL, (o.harmonic, err)
        for q in range(len(o.frequencies)):
            mac = cyclic.compute_mac(o.mode_shapes[:, q], ref.mode_shapes[:, q])[0, 0]
            assert mac >= MAC_TOL, (o.harmonic, q, mac)
        # residual recomputed here from the returned full-sector shapes (FP64 K(k), M(k))
        for q in range(0, len(o.frequencies), 2 if o.harmonic not in (0, m.num_sectors // 2) else 1):
            x = free_vector(m, o.mode_shapes[:, q])
            lam = (2 * np.pi * o.frequencies[q]) ** 2
            res, floor = residual_and_floor(ref.K, ref.M, x, lam)
            assert res <= max(RES_TOL, 4 * floor), (o.harmonic, q, res, floor)
            # the engine's own residual (free space, before expansion) agrees up to rounding
            assert abs(res - o.residuals[q // (2 if o.harmonic not in (0, m.num_sectors // 2) else 1)]) <= max(
                0.5 * res, 4 * floor, 1e-9)
        assert np.allclose(o.modal_mass, m.num_sectors, rtol=1e-9)


@pytest.mark.parametrize("method", ["lanczos", "block"])
@pytest.mark.parametrize("case", ["wedge", "blade_fixture", "c1"])
def test_all_harmonics_match_oracle(ctx, case, method):
    # wedge (< 256 free DOFs) runs the Lanczos method under either setting
    from turbomodal_amd import capi
    X, T, S, N, rpm = _case(case)
    m = oracle_model(X, T, S, N)
    em = engine_mesh(ctx, m)
    op = capi.EngineOperator(em, 200e9, 0.3, 7800.0, True)
    ops = cyclic.sector_operators(m, 200e9, 0.3, 7800.0, rpm)
    nm = 5
    meth = capi.TM_METHOD_BLOCK if method == "block" else capi.TM_METHOD_LANCZOS
    outs = capi.solve_batch(ctx, [dict(op=op, rpm=rpm, harmonic=k, num_modes=nm, method=meth)
                                  for k in range(N // 2 + 1)])
    _check_against_oracle(m, ops, outs, nm, rpm)
    op.close()
    em.close()


def test_reference_policy_and_even_mode_count(ctx):
    # the reference's own convergence test; even and odd num_modes duplication
    from turbomodal_amd import capi
    X, T, S, N, rpm = _case("c1")
    m = oracle_model(X, T, S, N)
    em = engine_mesh(ctx, m)
    op = capi.EngineOperator(em, 200e9, 0.3, 7800.0, True)
    ops = cyclic.sector_operators(m, 200e9, 0.3, 7800.0, rpm)
    outs = capi.solve_batch(ctx, [dict(op=op, rpm=rpm, harmonic=k, num_modes=6, conv_policy=capi.TM_CONV_REFERENCE)
                                  for k in (0, 3, 18)])
    _check_against_oracle(m, ops, outs, 6, rpm)
    op.close()
    em.close()


@pytest.mark.parametrize("method", ["lanczos", "block"])
def test_batch_composition_does_not_change_results(ctx, method):
    # a problem's arithmetic must not depend on what else shares its batch
    from turbomodal_amd import capi
    X, T, S, N, rpm = _case("c1")
    m = oracle_model(X, T, S, N)
    em = engine_mesh(ctx, m)
    op = capi.EngineOperator(em, 200e9, 0.3, 7800.0, True)
    meth = capi.TM_METHOD_BLOCK if method == "block" else capi.TM_METHOD_LANCZOS
    alone = capi.solve_batch(ctx, [dict(op=op, rpm=rpm, harmonic=7, num_modes=5, method=meth)])[0]
    mixed = capi.solve_batch(ctx, [dict(op=op, rpm=r, harmonic=k, num_modes=nm, method=meth)
                                   for r, k, nm in ((0.0, 0, 10), (rpm, 7, 5), (9000.0, 18, 3), (rpm, 2, 5))])[1]
    assert np.array_equal(alone.frequencies, mixed.frequencies)
    assert np.array_equal(alone.mode_shapes, mixed.mode_shapes)
    op.close()
    em.close()


def test_slot_groups_side_by_side_equal_serial(ctx):
    # a call's slot groups (block sizes; a large run in two) run side by side on their own
    # streams, or one after another with TMHIP_SERIAL_GROUPS=1: bit-identical results
    import os
    from turbomodal_amd import capi
    X, T, S, N, _ = _case("c1")
    m = oracle_model(X, T, S, N)
    em = engine_mesh(ctx, m)
    op = capi.EngineOperator(em, 200e9, 0.3, 7800.0, True)
    probs = [dict(op=op, rpm=r, harmonic=k, num_modes=5) for r in (0.0, 5000.0, 10000.0) for k in range(N // 2 + 1)]
    probs += [dict(op=op, rpm=7500.0, harmonic=k, num_modes=10) for k in range(N // 2 + 1)]
    old = os.environ.pop("TMHIP_SERIAL_GROUPS", None)
    try:
        side = capi.solve_batch(ctx, probs)
        os.environ["TMHIP_SERIAL_GROUPS"] = "1"
        serial = capi.solve_batch(ctx, probs)
    finally:
        os.environ.pop("TMHIP_SERIAL_GROUPS", None)
        if old is not None:
            os.environ["TMHIP_SERIAL_GROUPS"] = old
    for a, b in zip(side, serial):
        assert np.array_equal(a.frequencies, b.frequencies)
        assert np.array_equal(a.mode_shapes, b.mode_shapes)
    op.close()
    em.close()


def test_memory_reserve_sizes_batches_not_results(ctx):
    # memory_reserve_fraction (src/python_bindings.cpp:360) keeps device memory out of batch
    # sizing; a reserve of 1 leaves one problem per batch (the reference's max(1, budget / per),
    # src/cyclic_solver.cpp:1263), and the results stay bit-identical to full batches
    from turbomodal_amd import capi
    X, T, S, N, rpm = _case("c1")
    m = oracle_model(X, T, S, N)
    em = engine_mesh(ctx, m)
    op = capi.EngineOperator(em, 200e9, 0.3, 7800.0, True)
    probs = [dict(op=op, rpm=r, harmonic=k, num_modes=nm, method=meth)
             for r, k, nm in ((0.0, 0, 10), (rpm, 7, 5), (9000.0, 18, 3), (rpm, 2, 5))
             for meth in (capi.TM_METHOD_BLOCK, capi.TM_METHOD_LANCZOS)]
    for bad in (-0.1, 1.5, float("nan")):
        with pytest.raises(ValueError):
            ctx.set_memory_reserve(bad)
    try:
        full = capi.solve_batch(ctx, probs)
        ctx.set_memory_reserve(1.0)
        single = capi.solve_batch(ctx, probs)
    finally:
        ctx.set_memory_reserve(0.15)
    for a, b in zip(full, single):
        assert np.array_equal(a.frequencies, b.frequencies)
        assert np.array_equal(a.mode_shapes, b.mode_shapes)
    op.close()
    em.close()


def test_rpm_sweep_matches_independent_solves(ctx):
    # one operator at unit omega serves every rpm: compare against oracles built per rpm
    from turbomodal_amd import capi
    X, T, S, N, _ = _case("c1")
    m = oracle_model(X, T, S, N)
    em = engine_mesh(ctx, m)
    op = capi.EngineOperator(em, 200e9, 0.3, 7800.0, True)
    for rpm in (0.0, 7500.0, 15000.0):
        ops = cyclic.sector_operators(m, 200e9, 0.3, 7800.0, rpm)
        outs = capi.solve_batch(ctx, [dict(op=op, rpm=rpm, harmonic=k, num_modes=5) for k in (0, 5, 18)])
        _check_against_oracle(m, ops, outs, 5, rpm)
    op.close()
    em.close()


def test_bladed_c2_block_vs_oracle(ctx):
    # the 45k-DOF bladed sector of the rpm-sweep benchmark, ND 0 (10 modes, pb = 16) and
    # ND 1 (5 unique pairs, pb = 8) at 10263 rpm, against the oracle
    from turbomodal_amd import capi
    X, T, S, N = generated("bladed_c2")
    m = oracle_model(X, T, S, N)
    em = engine_mesh(ctx, m)
    op = capi.EngineOperator(em, 200e9, 0.3, 7800.0, True)
    rpm = 10263.0
    ops = cyclic.sector_operators(m, 200e9, 0.3, 7800.0, rpm)
    outs = capi.solve_batch(ctx, [dict(op=op, rpm=rpm, harmonic=k, num_modes=10) for k in (0, 1)])
    _check_against_oracle(m, ops, outs, 10, rpm)
    op.close()
    em.close()


def test_free_hub_min_frequency(ctx):
    # no constraints: rigid-body modes filtered by min_frequency, no prestress
    from turbomodal_amd import capi
    X, T, S, N, _ = _case("wedge")
    m = oracle_model(X, T, S, N, hub="free")
    em = engine_mesh(ctx, m)
    op = capi.EngineOperator(em, 200e9, 0.3, 7800.0, False)
    outs = capi.solve_batch(ctx, [dict(op=op, rpm=0.0, harmonic=k, num_modes=4, min_frequency=10.0)
                                  for k in (0, 1, 2)])
    ops = cyclic.sector_operators(m, 200e9, 0.3, 7800.0, 0.0)
    for o in outs:
        assert o.status == 0
        assert np.all(o.frequencies > 10.0)
        # oracle: all eigenpairs of the free system, keep f > 10 Hz
        Kk = cyclic.harmonic_matrix(ops.Kc, ops.Kp, o.harmonic, N).toarray()
        Mk = cyclic.harmonic_matrix(ops.Mc, ops.Mp, o.harmonic, N).toarray()
        import scipy.linalg as sla
        w = sla.eigh(0.5 * (Kk + Kk.conj().T), 0.5 * (Mk + Mk.conj().T), eigvals_only=True)
        f = np.sqrt(np.maximum(w, 0)) / (2 * np.pi)
        f = f[f > 10.0]
        uniq = o.frequencies[::2] if o.harmonic not in (0, N // 2) else o.frequencies
        assert np.allclose(uniq, f[:len(uniq)], rtol=1e-8)
    op.close()
    em.close()


def test_error_paths(ctx):
    from turbomodal_amd import capi
    X, T, S, N, _ = _case("wedge")
    m = oracle_model(X, T, S, N)
    em = engine_mesh(ctx, m)
    with pytest.raises(ValueError):
        capi.EngineOperator(em, -1.0, 0.3, 7800.0, True)
    with pytest.raises(ValueError):
        capi.EngineOperator(em, 200e9, 0.5, 7800.0, True)
    op = capi.EngineOperator(em, 200e9, 0.3, 7800.0, True)
    with pytest.raises(ValueError):
        capi.solve_batch(ctx, [dict(op=op, rpm=0.0, harmonic=N, num_modes=5)])
    # a right node without a matched partner is a runtime error (src/cyclic_solver.cpp:303-307)
    cls, partner, cm = capi.mesh_arrays_from_model(m.nodes, m.elements, m.left, m.right, m.pairs, m.constrained)
    partner = partner.copy()
    partner[m.pairs[0][1]] = -1
    with pytest.raises(capi.EngineError):
        capi.EngineMesh(ctx, m.nodes, m.elements, cls, partner, cm, N)
    op.close()
    em.close()


def test_dropin_api_solve_and_sweep():
    import turbomodal_amd as tm
    from common import GOLDEN
    import os
    path = os.path.join(GOLDEN, "wedge_sector.msh")
    mesh = tm.load_mesh(path, 24)
    mat = tm.Material(200e9, 0.3, 7800.0)
    res = tm.solve(mesh, mat, rpm=2000.0, num_modes=4, verbose=0, harmonic_indices=[0, 3, 12])
    assert [r.harmonic_index for r in res] == [0, 3, 12]
    X, T, S, N = fixture_mesh("wedge_sector.msh")
    m = oracle_model(X, T, S, N)
    ops = cyclic.sector_operators(m, 200e9, 0.3, 7800.0, 2000.0)
    for r in res:
        ref = cyclic.solve_harmonic(m, ops, r.harmonic_index, 4, 2000.0, "tight")
        assert np.max(np.abs(np.asarray(r.frequencies) / ref.frequencies - 1)) <= FREQ_TOL
        assert np.asarray(r.mode_shapes).shape == (3 * len(X), len(ref.frequencies))
        assert np.allclose(r.modal_mass_kg, 24.0, rtol=1e-9)
    sweep = tm.rpm_sweep(mesh, mat, [0.0, 2000.0], num_modes=4, harmonic_indices=[0, 3, 12])
    assert len(sweep) == 2 and [r.rpm for r in sweep[1]] == [2000.0] * 3
    for a, b in zip(sweep[1], res):
        assert np.allclose(a.frequencies, b.frequencies, rtol=1e-12)
    # stationary frame and npz round trip of served results
    cd = tm.campbell_data(sweep, 24)
    assert cd["frequencies"].shape[:2] == (2, 3)
    import tempfile
    with tempfile.TemporaryDirectory() as d:
        p = os.path.join(d, "r.npz")
        tm.save_results(p, sweep)
        data = tm.load_results(p)
        assert data["is_sweep"] and data["mode_shapes"].shape[:2] == (2, 3)
        assert np.allclose(data["frequencies"][1, 1, :4], sweep[1][1].frequencies)
