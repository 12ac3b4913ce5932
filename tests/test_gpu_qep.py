"""Coriolis QEP on the GPU (tm_problem.coriolis) against the oracle's Lancaster-pencil
solve: frequencies |w| / 2 pi to 1e-8, whirl = sign(w), MAC >= 0.999, the QEP residual
||(K + w D - w^2 M) phi|| / ||K phi|| <= max(1e-8, 4 x its FP64 floor) recomputed here and
equal to the reported residual, k = 0 on the Hermitian path."""
import numpy as np
import pytest

from common import engine_mesh, fixture_mesh, generated, oracle_model
from oracle import cyclic

pytestmark = pytest.mark.gpu


def free_vector(m, phi):
    red = np.array(m.interior_dofs + m.left_dofs)
    return phi[red[cyclic.reduced_free_map(m)]]


@pytest.fixture(scope="module")
def ctx():
    from turbomodal_amd import capi
    c = capi.Context()
    yield c
    c.close()


def _check(m, ops, outs, nm, rpm):
    N = m.num_sectors
    for o in outs:
        assert o.status == 0, o.harmonic
        if o.harmonic == 0:
            ref = cyclic.solve_harmonic(m, ops, 0, nm, rpm, "tight")
            assert np.max(np.abs(o.frequencies / ref.frequencies - 1)) <= 1e-8
            continue
        ref = cyclic.solve_harmonic_qep(m, ops, o.harmonic, nm, rpm)
        assert len(o.frequencies) == len(ref.frequencies) == nm
        assert np.max(np.abs(o.frequencies / ref.frequencies - 1)) <= 1e-8, o.harmonic
        Kk, Mk, Dk = ref.K, ref.M, ref.D
        real_h = N % 2 == 0 and o.harmonic == N // 2
        for grp in _tie_groups(ref.frequencies):
            # a real harmonic (k = N/2) has exact +-omega pairs: the order inside a
            # tie is rounding, so whirl is compared as a set and shapes by whirl
            if real_h and grp[-1] == nm - 1 and len(grp) % 2:
                continue  # a pair cut by the truncation: either member is right
            assert sorted(o.whirl[grp]) == sorted(ref.whirl_direction[grp]), (o.harmonic, o.whirl, ref.whirl_direction)
            for q in grp:
                qr = [j for j in grp if ref.whirl_direction[j] == o.whirl[q]]
                qr = q if q in qr else qr[0]
                _check_mode(m, N, o, ref, q, qr, Kk, Mk, Dk)


def _tie_groups(f, rel=1e-9):
    groups, cur = [], [0]
    for q in range(1, len(f)):
        if abs(f[q] - f[cur[-1]]) <= rel * f[q]:
            cur.append(q)
        else:
            groups.append(cur)
            cur = [q]
    groups.append(cur)
    return groups


def _check_mode(m, N, o, ref, q, qr, Kk, Mk, Dk):
    mac = cyclic.compute_mac(o.mode_shapes[:, q], ref.mode_shapes[:, qr])[0, 0]
    assert mac >= 0.999, (o.harmonic, q, mac)
    phi = free_vector(m, o.mode_shapes[:, q])
    w = o.eigenvalues[q]
    assert abs(abs(w) / (2 * np.pi) - o.frequencies[q]) <= 1e-12 * o.frequencies[q]
    assert np.sign(w) == o.whirl[q]
    r = (Kk @ phi) + w * (Dk @ phi) - w * w * (Mk @ phi)
    kphi = np.linalg.norm(Kk @ phi)
    # FP64 rounding floor of this residual: eps || |K||phi| + |w||D||phi| + w^2 |M||phi| ||
    floor = np.finfo(float).eps * np.linalg.norm(
        abs(Kk) @ abs(phi) + abs(w) * (abs(Dk) @ abs(phi)) + w * w * (abs(Mk) @ abs(phi))) / kphi
    assert np.linalg.norm(r) <= max(1e-8, 4 * floor) * kphi, (o.harmonic, q, np.linalg.norm(r) / kphi)
    assert abs(o.residuals[q] - np.linalg.norm(r) / kphi) <= max(0.5 * o.residuals[q], 4 * floor, 1e-12)
    # the mode is normalised as the oracle's (z^H L2 z = 1): same modal mass N Re(a^H M a)
    ref_phi = free_vector(m, ref.mode_shapes[:, qr])
    mm_ref = N * np.real(ref_phi.conj() @ (Mk @ ref_phi))
    assert abs(o.modal_mass[q] - N * np.real(phi.conj() @ (Mk @ phi))) <= 1e-10 * mm_ref
    assert abs(o.modal_mass[q] / mm_ref - 1) <= 1e-6, (o.harmonic, q)


def test_wedge_all_harmonics(ctx):
    from turbomodal_amd import capi
    X, T, S, N = fixture_mesh("wedge_sector.msh")
    m = oracle_model(X, T, S, N)
    rpm = 30000.0
    em = engine_mesh(ctx, m)
    op = capi.EngineOperator(em, 200e9, 0.3, 7800.0, True)
    ops = cyclic.sector_operators(m, 200e9, 0.3, 7800.0, rpm)
    outs = capi.solve_batch(ctx, [dict(op=op, rpm=rpm, harmonic=k, num_modes=6, coriolis=True)
                                  for k in range(N // 2 + 1)])
    _check(m, ops, outs, 6, rpm)
    op.close()
    em.close()


def test_flat_disk_c1(ctx):
    from turbomodal_amd import capi
    X, T, S, N = generated("flat_disk_c1")
    m = oracle_model(X, T, S, N)
    rpm = 12000.0
    em = engine_mesh(ctx, m)
    op = capi.EngineOperator(em, 200e9, 0.3, 7800.0, True)
    ops = cyclic.sector_operators(m, 200e9, 0.3, 7800.0, rpm)
    outs = capi.solve_batch(ctx, [dict(op=op, rpm=rpm, harmonic=k, num_modes=10, coriolis=True) for k in (0, 1, 6, 18)])
    _check(m, ops, outs, 10, rpm)
    op.close()
    em.close()


def test_dropin_include_coriolis():
    import turbomodal_amd as tm
    from turbomodal_amd import meshgen
    sm = meshgen.config_mesh("flat_disk_c1")
    mesh = meshgen.to_core_mesh(sm, tm._core)
    mat = tm.Material(200e9, 0.3, 7800.0)
    res = tm.solve(mesh, mat, rpm=9000.0, num_modes=6, harmonic_indices=[0, 2, 18], include_coriolis=True, verbose=0)
    assert [r.harmonic_index for r in res] == [0, 2, 18]
    assert np.all(res[0].whirl_direction == 0)
    for r in res[1:]:
        assert set(np.unique(r.whirl_direction)) <= {-1, 1}
        assert np.all(np.diff(r.frequencies) >= -1e-9)
