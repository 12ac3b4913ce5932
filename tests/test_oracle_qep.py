"""Oracle pinning for the Coriolis QEP (src/modal_solver.cpp:149-312,
src/cyclic_solver.cpp:858-1005): D(k) Hermitian (tests/test_cyclic.cpp:678-731), the
selected pairs solve (K + w D - w^2 M) phi = 0, a second linearisation (first companion
form, non-Hermitian eig) gives the same eigenvalues, the rotating-frame limit, and a
full-annulus brute force: the cyclic QEP eigenvalues of every harmonic k >= 1 (and their
mirror -w for 0 < k < N/2, the harmonic N - k) are eigenvalues of the assembled 360-degree
gyroscopic model with the same prestress."""
import math

import numpy as np
import pytest
import scipy.linalg as sla
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from common import fixture_mesh, oracle_model
from oracle import cyclic, fem

RPM = 30000.0


@pytest.fixture(scope="module")
def wedge():
    X, T, S, N = fixture_mesh("wedge_sector.msh")
    m = oracle_model(X, T, S, N)
    return X, T, S, N, m, cyclic.sector_operators(m, 200e9, 0.3, 7800.0, RPM)


def test_coriolis_matrix_hermitian_and_direct_projection(wedge):
    X, T, S, N, m, ops = wedge
    omega = RPM * 2 * math.pi / 60
    _, G = fem.assemble_rotating(m.nodes, m.elements, 7800.0, omega, np.array([0.0, 0.0, 1.0]))
    red = np.array(m.interior_dofs + m.left_dofs)
    fm = red[cyclic.reduced_free_map(m)]
    for k in (1, 5, N // 2):
        D = cyclic.coriolis_matrix(ops, k, N).toarray()
        assert np.abs(D - D.conj().T).max() <= 1e-12 * np.abs(D).max()
        # the phase split equals i 2 omega T(k)^H G T(k) on the free DOFs
        Tk = cyclic.cyclic_transformation(m, k)
        Gk = (Tk.conj().T @ G @ Tk).toarray()
        free = cyclic.reduced_free_map(m)
        Dd = 2j * omega * Gk[np.ix_(free, free)]
        assert np.abs(D - Dd).max() <= 1e-12 * np.abs(Dd).max()
    # G is real skew-symmetric before the cyclic projection
    assert np.abs((G + G.T).toarray()).max() <= 1e-12 * np.abs(G.toarray()).max()


def test_selected_pairs_solve_the_qep_and_companion_form(wedge):
    X, T, S, N, m, ops = wedge
    for k in (1, 4, N // 2):
        Kk = cyclic.harmonic_matrix(ops.Kc, ops.Kp, k, N).toarray()
        Mk = cyclic.harmonic_matrix(ops.Mc, ops.Mp, k, N).toarray()
        Dk = cyclic.coriolis_matrix(ops, k, N).toarray()
        r = cyclic.solve_harmonic_qep(m, ops, k, 6, RPM)
        assert len(r.frequencies) == 6
        assert np.all(np.diff(r.frequencies) >= -1e-9)
        for w, phi in zip(r.eigenvalues, r.reduced_vectors.T):
            Q = Kk + w * Dk - w * w * Mk
            assert np.linalg.norm(Q @ phi) <= 1e-8 * np.linalg.norm(Kk @ phi)
        # first companion linearisation: [[0, I], [K, D]] - w [[I, 0], [0, M]]
        n = Kk.shape[0]
        A = np.block([[np.zeros((n, n)), np.eye(n)], [Kk, Dk]])
        B = np.block([[np.eye(n), np.zeros((n, n))], [np.zeros((n, n)), Mk]])
        wc = sla.eig(A, B, right=False)
        wc = wc[np.isfinite(wc)].real
        # the non-Hermitian companion pencil is less well conditioned for LAPACK's
        # QZ: agreement to 1e-6 (the residual check above is the strict one)
        for w in r.eigenvalues:
            assert np.min(np.abs(wc - w)) <= 1e-6 * abs(w)


def test_slow_rotation_limit(wedge):
    # at small omega the QEP eigenvalues split symmetrically around +-sqrt(lambda_GEP)
    X, T, S, N, m, _ = wedge
    ops = cyclic.sector_operators(m, 200e9, 0.3, 7800.0, 10.0)
    for k in (1, 3):
        q = cyclic.solve_harmonic_qep(m, ops, k, 4, 10.0)
        g = cyclic.solve_harmonic(m, ops, k, 4, 10.0, "tight")
        for f in q.frequencies:
            assert np.min(np.abs(g.frequencies - f)) / f < 1e-3


def test_full_annulus_gyroscopic_brute_force(wedge):
    X, T, S, N, m, ops = wedge
    nn = len(X)
    r2l = {r: l for l, r in m.pairs}
    gid = -np.ones((N, nn), int)
    cnt = 0
    for s in range(N):
        for i in range(nn):
            if i not in r2l:
                gid[s, i] = cnt
                cnt += 1
    for s in range(N):
        for i, l in r2l.items():
            gid[s, i] = gid[(s + 1) % N, l]
    coords = np.zeros((cnt, 3))
    for s in range(N):
        a = 2 * math.pi * s / N
        R = np.array([[math.cos(a), -math.sin(a), 0], [math.sin(a), math.cos(a), 0], [0, 0, 1]])
        for i in range(nn):
            coords[gid[s, i]] = R @ X[i]
    els = np.concatenate([gid[s][T] for s in range(N)])
    omega = RPM * 2 * math.pi / 60
    ax = np.array([0.0, 0.0, 1.0])
    K, M = fem.assemble_KM(coords, els, 200e9, 0.3, 7800.0)
    Kw, G = fem.assemble_rotating(coords, els, 7800.0, omega, ax)
    F = fem.assemble_centrifugal(coords, els, 7800.0, omega, ax)
    hub = {gid[s, i] for s in range(N) for i in S["hub_constraint"]}
    free = np.array([d for d in range(3 * cnt) if d // 3 not in hub])
    u = np.zeros(3 * cnt)
    u[free] = spla.spsolve(K.tocsc()[free][:, free], F[free])
    Ks = fem.assemble_stress_stiffening(coords, els, 200e9, 0.3, u)
    Ke = (K - Kw + Ks).toarray()[np.ix_(free, free)]
    Mf = M.toarray()[np.ix_(free, free)]
    Df = 2j * omega * G.toarray()[np.ix_(free, free)]
    n = len(free)
    L1 = np.block([[Df, Ke], [Ke, np.zeros((n, n))]])
    L2 = np.block([[Mf, np.zeros((n, n))], [np.zeros((n, n)), Ke]])
    w_full = sla.eigh(0.5 * (L1 + L1.conj().T), 0.5 * (L2 + L2.conj().T), eigvals_only=True)
    for k in range(1, N // 2 + 1):
        r = cyclic.solve_harmonic_qep(m, ops, k, 6, RPM)
        for w in r.eigenvalues:
            assert np.min(np.abs(w_full - w)) <= 1e-8 * abs(w), (k, w)
            if k < N // 2:
                assert np.min(np.abs(w_full + w)) <= 1e-8 * abs(w), (k, -w)


def test_sparse_path_equals_dense(wedge):
    # the large-system path (eigsh shift-invert at 0 with L1^-1 through a factor of K(k))
    # returns the dense path's pairs
    X, T, S, N, m, ops = wedge
    for k in (1, 7):
        d = cyclic.solve_harmonic_qep(m, ops, k, 6, RPM)
        s = cyclic.solve_harmonic_qep(m, ops, k, 6, RPM, dense_max=0)
        assert np.allclose(s.eigenvalues, d.eigenvalues, rtol=1e-10)
        assert np.array_equal(s.whirl_direction, d.whirl_direction)
        for q in range(6):
            assert cyclic.compute_mac(s.mode_shapes[:, q], d.mode_shapes[:, q])[0, 0] >= 1 - 1e-9
