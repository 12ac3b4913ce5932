"""TM_PREC_MIXED: the supernodal LDL^H in complex FP32 (Schur tiles on
v_mfma_f32_16x16x4 with FP32 accumulation, factor stored in FP32) and every shift-invert
solve of the block method refined in FP64 against the FP64 K(k), M(k) (two refinement
steps), with the FP64 Rayleigh-Ritz step and residual checks unchanged.  On the flat disk
(4.9k DOF, lambda_max / lambda_1 about 1e7, so the FP32 factor's error kappa eps_32 stays
below 1) the results meet the FP64 oracle's bars; requests the mixed path does not serve
are refused."""
import numpy as np
import pytest

from common import engine_mesh, fixture_mesh, generated, oracle_model
from oracle import cyclic

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    from turbomodal_amd import capi
    c = capi.Context()
    yield c
    c.close()


def test_mixed_flat_disk_matches_oracle(ctx):
    from turbomodal_amd import capi
    X, T, S, N = generated("flat_disk_c1")
    m = oracle_model(X, T, S, N)
    em = engine_mesh(ctx, m)
    op = capi.EngineOperator(em, 200e9, 0.3, 7800.0, True)
    rpm = 6000.0
    ops = cyclic.sector_operators(m, 200e9, 0.3, 7800.0, rpm)
    ks = (0, 5, 18)
    mixed = capi.solve_batch(ctx, [dict(op=op, rpm=rpm, harmonic=k, num_modes=10, precision=capi.TM_PREC_MIXED)
                                   for k in ks])
    fp64 = capi.solve_batch(ctx, [dict(op=op, rpm=rpm, harmonic=k, num_modes=10) for k in ks])
    for o, f in zip(mixed, fp64):
        assert o.status == 0 and o.converged, o.harmonic
        ref = cyclic.solve_harmonic(m, ops, o.harmonic, 10, rpm, "tight")
        err = np.max(np.abs(o.frequencies / ref.frequencies - 1))
        assert err <= 1e-8, (o.harmonic, err)
        assert np.max(np.abs(o.frequencies / f.frequencies - 1)) <= 1e-8
        for q in range(len(o.frequencies)):
            assert cyclic.compute_mac(o.mode_shapes[:, q], ref.mode_shapes[:, q])[0, 0] >= 0.999
        assert np.all(o.residuals <= 4e-8)
        assert np.allclose(o.modal_mass, N, rtol=1e-9)
    # the FP32 factor holds half the bytes
    assert op.footprint(10, 5, precision=capi.TM_PREC_MIXED) < op.footprint(10, 5)
    op.close()
    em.close()


def test_mixed_refuses_what_it_does_not_serve(ctx):
    from turbomodal_amd import capi
    X, T, S, N = generated("flat_disk_c1")
    m = oracle_model(X, T, S, N)
    em = engine_mesh(ctx, m)
    op = capi.EngineOperator(em, 200e9, 0.3, 7800.0, True)
    for extra in (dict(method=capi.TM_METHOD_LANCZOS), dict(coriolis=True, rpm=6000.0),
                  dict(conv_policy=capi.TM_CONV_REFERENCE)):
        p = dict(op=op, rpm=0.0, harmonic=2, num_modes=6, precision=capi.TM_PREC_MIXED)
        p.update(extra)
        with pytest.raises(ValueError):
            capi.solve_batch(ctx, [p])
    with pytest.raises(ValueError):
        capi.solve_batch(ctx, [dict(op=op, rpm=0.0, harmonic=2, num_modes=6, precision=7)])
    op.close()
    em.close()
    X, T, S, N = fixture_mesh("wedge_sector.msh")  # below 256 free DOFs: no block method
    m = oracle_model(X, T, S, N)
    em = engine_mesh(ctx, m)
    op = capi.EngineOperator(em, 200e9, 0.3, 7800.0, True)
    with pytest.raises(ValueError):
        capi.solve_batch(ctx, [dict(op=op, rpm=0.0, harmonic=1, num_modes=4, precision=capi.TM_PREC_MIXED)])
    op.close()
    em.close()


def test_mixed_bladed_sector_matches_oracle(ctx):
    # the 45k-DOF bladed sector (thin blade: lambda_max / lambda_1 ~ 5e8, so kappa eps_32 > 1
    # for its lowest modes): the refinement still brings every solve to the FP64 level;
    # rpm 0 x ND 0, 9, 18 against the oracle's answers (tests/golden/c4_sample.npz)
    from common import load_golden, subset_mac
    from turbomodal_amd import capi, meshgen
    gold = load_golden("c4_sample.npz")
    X, T, S, N = generated("bladed_c2")
    m = oracle_model(X, T, S, N)
    em = engine_mesh(ctx, m)
    op = capi.EngineOperator(em, **meshgen.STEEL, prestress=True)
    outs = capi.solve_batch(ctx, [dict(op=op, rpm=0.0, harmonic=k, num_modes=10, precision=capi.TM_PREC_MIXED)
                                  for k in (0, 9, 18)])
    sub = gold["subset"]
    for o in outs:
        assert o.status == 0 and o.converged, o.harmonic
        key = f"r0_k{o.harmonic}"
        assert np.max(np.abs(o.frequencies / gold[key + "_freq"] - 1)) <= 1e-8, o.harmonic
        for q in range(10):
            assert subset_mac(o.mode_shapes[sub, q], gold[key + "_shape"][:, q]) >= 0.999
        assert np.all(o.residuals <= 4e-8)
    op.close()
    em.close()
