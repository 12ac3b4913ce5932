"""Pins the oracle's element layer (oracle/fem.py) independently of its own formulas.

Checks come from the reference's tests (tests/test_element.cpp,
tests/test_material.cpp, tests/test_rotating.cpp) and from mathematics the
restatement does not use: exact simplex integration of barycentric monomials
(int L1^a L2^b L3^c L4^d dV = 6V a!b!c!d!/(a+b+c+d+3)!), finite differences,
rigid-body null spaces and patch tests.
"""
import math

import numpy as np
import pytest

from oracle import fem

REF_TET = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1],
                    [.5, 0, 0], [.5, .5, 0], [0, .5, 0], [0, 0, .5], [.5, 0, .5], [0, .5, .5]], float)
EDGES = ((0, 1), (1, 2), (0, 2), (0, 3), (1, 3), (2, 3))


def random_tet(rng, scale=0.01):
    c = rng.normal(size=(4, 3)) * scale
    if np.linalg.det(np.stack([c[1] - c[0], c[2] - c[0], c[3] - c[0]])) < 0:
        c[[1, 2]] = c[[2, 1]]
    X = np.zeros((10, 3))
    X[:4] = c
    for m, (a, b) in enumerate(EDGES):
        X[4 + m] = 0.5 * (c[a] + c[b])
    return X


def test_partition_of_unity_and_nodal_values():
    # tests/test_element.cpp:55-86
    for p in [(0.25, 0.25, 0.25), (0.1, 0.1, 0.1), (0.7, 0.1, 0.05)]:
        assert abs(fem.shape_functions(*p).sum() - 1.0) < 1e-14
        assert np.allclose(fem.shape_derivatives(*p).sum(axis=0), 0.0, atol=1e-14)
    for i, xi in enumerate(REF_TET):
        N = fem.shape_functions(*xi)
        e = np.zeros(10)
        e[i] = 1.0
        assert np.allclose(N, e, atol=1e-14)
    # centroid values (tests/test_element.cpp:88-106)
    N = fem.shape_functions(0.25, 0.25, 0.25)
    assert np.allclose(N[:4], -0.125) and np.allclose(N[4:], 0.25)


def test_derivatives_finite_difference():
    # tests/test_element.cpp:108-139
    h = 1e-6
    for p in [(0.2, 0.3, 0.1), (0.05, 0.6, 0.2)]:
        d = fem.shape_derivatives(*p)
        for k in range(3):
            pp = list(p); pm = list(p)
            pp[k] += h; pm[k] -= h
            fd = (fem.shape_functions(*pp) - fem.shape_functions(*pm)) / (2 * h)
            assert np.allclose(d[:, k], fd, atol=1e-8)


def exact_monomial(a, b, c):
    """int over the unit simplex of xi^a eta^b zeta^c."""
    return math.factorial(a) * math.factorial(b) * math.factorial(c) / math.factorial(a + b + c + 3)


@pytest.mark.parametrize("deg", [0, 1, 2, 3, 4])
def test_keast14_exact_to_degree_4(deg):
    for a in range(deg + 1):
        for b in range(deg + 1 - a):
            c = deg - a - b
            q = sum(w * p[0] ** a * p[1] ** b * p[2] ** c for p, w in zip(fem.KEAST14_PTS, fem.KEAST14_W))
            assert abs(q - exact_monomial(a, b, c)) <= 2e-16 * max(1.0, 1 / exact_monomial(a, b, c)) * exact_monomial(a, b, c) * 10


def test_gauss4_exact_to_degree_2_only():
    # the centrifugal-load rule integrates degree 2 exactly and not degree 3
    for a, b, c in [(0, 0, 0), (1, 0, 0), (1, 1, 0), (2, 0, 0)]:
        q = sum(w * p[0] ** a * p[1] ** b * p[2] ** c for p, w in zip(fem.GAUSS4_PTS, fem.GAUSS4_W))
        assert abs(q - exact_monomial(a, b, c)) < 1e-15
    q = sum(w * p[0] ** 3 for p, w in zip(fem.GAUSS4_PTS, fem.GAUSS4_W))
    assert abs(q - exact_monomial(3, 0, 0)) > 1e-4
    assert abs(fem.KEAST14_W.sum() - 1 / 6) < 1e-15 and abs(fem.GAUSS4_W.sum() - 1 / 6) < 1e-16


def test_constitutive_known_values():
    # tests/test_material.cpp:90-139
    D = fem.constitutive_matrix(200e9, 0.3)
    assert abs(D[0, 0] - 269.2308e9) < 1e5 and abs(D[0, 1] - 115.3846e9) < 1e5
    assert abs(D[3, 3] - 76.9231e9) < 1e5 and abs(D[3, 3] - 200e9 / 2.6) < 1e-2
    assert D[0, 0] == D[1, 1] == D[2, 2] and D[3, 3] == D[4, 4] == D[5, 5]
    assert D[2, 5] == 0.0
    D = fem.constitutive_matrix(110e9, 0.34)
    c = 110e9 / (1.34 * 0.32)
    assert abs(D[0, 0] - c * 0.66) < 1e-2 and abs(D[0, 1] - c * 0.34) < 1e-2
    with pytest.raises(ValueError):
        fem.validate_material(-1, 0.3, 7800)
    with pytest.raises(ValueError):
        fem.validate_material(200e9, 0.5, 7800)
    with pytest.raises(ValueError):
        fem.validate_material(200e9, 0.3, 0)


def barycentric_mass(V):
    """Exact TET10 scalar mass int N_a N_b dV from barycentric monomials."""
    # N_a as polynomials in (L1..L4): dict exponent-tuple -> coefficient
    def lin(i):
        e = [0, 0, 0, 0]; e[i] = 1
        return {tuple(e): 1.0}

    def mul(p, q):
        r = {}
        for ea, ca in p.items():
            for eb, cb in q.items():
                e = tuple(x + y for x, y in zip(ea, eb))
                r[e] = r.get(e, 0.0) + ca * cb
        return r

    def add(p, q, s=1.0):
        r = dict(p)
        for e, c in q.items():
            r[e] = r.get(e, 0.0) + s * c
        return r

    N = []
    for i in range(4):
        N.append(add({k: 2 * v for k, v in mul(lin(i), lin(i)).items()}, lin(i), -1.0))
    for a, b in EDGES:
        N.append({k: 4 * v for k, v in mul(lin(a), lin(b)).items()})

    def integ(p):
        return sum(c * 6 * V * np.prod([math.factorial(x) for x in e]) / math.factorial(sum(e) + 3)
                   for e, c in p.items())

    return np.array([[integ(mul(N[i], N[j])) for j in range(10)] for i in range(10)])


def test_mass_equals_exact_consistent_mass():
    rng = np.random.default_rng(3)
    for _ in range(5):
        X = random_tet(rng)
        V = np.linalg.det(np.stack([X[1] - X[0], X[2] - X[0], X[3] - X[0]])) / 6
        m = fem.element_scalar_mass(X[None], 1.0)[0]
        ex = barycentric_mass(V)
        assert np.allclose(m, ex, rtol=1e-13, atol=1e-13 * abs(ex).max())
        # the familiar 1/420 table: corner diag 6, corner-corner 1, corner-mid -4/-6, mid 32/16/8
        t = ex / (V / 420)
        assert abs(t[0, 0] - 6) < 1e-9 and abs(t[0, 1] - 1) < 1e-9
        assert abs(t[0, 4] + 4) < 1e-9 and abs(t[0, 5] + 6) < 1e-9
        assert abs(t[4, 4] - 32) < 1e-9 and abs(t[4, 5] - 16) < 1e-9 and abs(t[4, 9] - 8) < 1e-9
        Me = fem.element_mass(X[None], 7800.0)[0]
        # Me sum over one component = rho V (tests/test_element.cpp:279-340)
        assert abs(Me[0::3, 0::3].sum() - 7800.0 * V) < 1e-12 * 7800 * V
        assert np.all(Me[0::3, 1::3] == 0.0)


def rigid_modes(X):
    R = []
    for d in range(3):
        v = np.zeros(30); v[d::3] = 1.0; R.append(v)
    for axis in range(3):
        w = np.zeros(3); w[axis] = 1.0
        R.append(np.cross(w, X).ravel())
    return np.array(R).T


def test_stiffness_invariants():
    rng = np.random.default_rng(5)
    D = fem.constitutive_matrix(200e9, 0.3)
    for _ in range(4):
        X = random_tet(rng)
        K = fem.element_stiffness(X[None], D)[0]
        assert np.allclose(K, K.T, atol=1e-9 * abs(K).max())
        R = rigid_modes(X)
        assert np.abs(K @ R).max() < 1e-8 * abs(K).max()
        w = np.linalg.eigvalsh(K)
        assert w.min() > -1e-8 * w.max()
        assert np.sum(w > 1e-8 * w.max()) == 24  # 30 - 6 rigid modes
        K2 = fem.element_stiffness(X[None], fem.constitutive_matrix(400e9, 0.3))[0]
        assert np.allclose(K2, 2 * K, rtol=1e-12, atol=1e-12 * abs(K).max())


def test_patch_and_internal_force():
    # constant strain from a linear field (tests/test_element.cpp:355-563) and
    # internal force of a constant stress = integral of B^T sigma (565-611)
    rng = np.random.default_rng(7)
    X = random_tet(rng)
    A = rng.normal(size=(3, 3)) * 1e-3
    u = (X @ A.T).ravel()
    eps_exact = np.array([A[0, 0], A[1, 1], A[2, 2], A[0, 1] + A[1, 0], A[1, 2] + A[2, 1], A[0, 2] + A[2, 0]])
    detJ, dNdx = fem._geometry14(X[None])
    B = fem._b_matrix(dNdx)[0]
    for g in range(14):
        assert np.allclose(B[g] @ u, eps_exact, atol=1e-12)
    D = fem.constitutive_matrix(200e9, 0.3)
    sig = fem.element_prestress(X[None], D, u[None])[0]
    assert np.allclose(sig, (D @ eps_exact)[None, :], rtol=1e-10)
    K = fem.element_stiffness(X[None], D)[0]
    V = abs(detJ[0, 0]) / 6
    # closed form of int B^T sigma for a constant sigma: int grad N_corner = 0,
    # int grad(4 L_i L_j) = V (grad L_i + grad L_j); grad L from the corner coordinates
    s = D @ eps_exact
    S = np.array([[s[0], s[3], s[5]], [s[3], s[1], s[4]], [s[5], s[4], s[2]]])
    A4 = np.hstack([np.ones((4, 1)), X[:4]])
    gradL = np.linalg.inv(A4)[1:, :].T          # row k = grad L_k
    f_int = np.zeros((10, 3))
    for m, (i, j) in enumerate(EDGES):
        f_int[4 + m] = V * S @ (gradL[i] + gradL[j])
    f_int = f_int.ravel()
    assert np.allclose(K @ u, f_int, rtol=1e-9, atol=1e-9 * abs(f_int).max())


def test_rotating_effects():
    rng = np.random.default_rng(11)
    X = random_tet(rng) + np.array([0.2, 0.05, 0.0])
    ax = np.array([0, 0, 1.0])
    G = fem.element_gyroscopic(X[None], 7800.0, ax)[0]
    assert np.allclose(G, -G.T, atol=1e-15 * abs(G).max() + 1e-30)
    K1 = fem.element_spin_softening(X[None], 7800.0, 100.0, ax)[0]
    K2 = fem.element_spin_softening(X[None], 7800.0, 200.0, ax)[0]
    assert np.allclose(K2, 4 * K1, rtol=1e-14) and np.allclose(K1, K1.T)
    F1 = fem.element_centrifugal_load(X[None], 7800.0, 100.0, ax)[0]
    F2 = fem.element_centrifugal_load(X[None], 7800.0, 200.0, ax)[0]
    assert np.allclose(F2, 4 * F1, rtol=1e-14)
    # radial: total force along the centroid direction, no axial component
    tot = F1.reshape(10, 3).sum(axis=0)
    cen = X[:4].mean(axis=0)
    assert abs(tot[2]) < 1e-12 * np.linalg.norm(tot)
    assert np.dot(tot[:2], cen[:2]) > 0.99 * np.linalg.norm(tot[:2]) * np.linalg.norm(cen[:2])
    # total centrifugal force of a degree-1 field (r_perp linear) is exact: rho w^2 V centroid
    V = np.linalg.det(np.stack([X[1] - X[0], X[2] - X[0], X[3] - X[0]])) / 6
    assert np.allclose(tot[:2], 7800.0 * 100.0 ** 2 * V * cen[:2], rtol=1e-12)


def test_stress_stiffening():
    rng = np.random.default_rng(13)
    X = random_tet(rng)
    zero = np.zeros((1, 14, 6))
    assert np.all(fem.element_stress_stiffening(X[None], zero) == 0.0)
    s = np.array([1e6, -2e6, 3e5, 4e5, -1e5, 2e5])
    Ks = fem.element_stress_stiffening(X[None], np.broadcast_to(s, (1, 14, 6)).copy())[0]
    assert np.allclose(Ks, Ks.T, atol=1e-12 * abs(Ks).max())
    # constant stress closed form: block (a,b) = (int grad N_a^T S grad N_b) I3 (Keast-14 exact: degree 2)
    detJ, dNdx = fem._geometry14(X[None])
    S = np.array([[s[0], s[3], s[5]], [s[3], s[1], s[4]], [s[5], s[4], s[2]]])
    blk = np.einsum("gai,ij,gbj,g->ab", dNdx[0], S, dNdx[0], detJ[0] * fem.KEAST14_W)
    assert np.allclose(Ks[0::3, 0::3], blk, rtol=1e-12, atol=1e-12 * abs(blk).max())
    assert np.all(Ks[0::3, 1::3] == 0.0)


def test_assembly_single_element_and_total_mass():
    # tests/test_assembler.cpp:96-122, 213-270
    rng = np.random.default_rng(17)
    X = random_tet(rng)
    el = np.arange(10)[None]
    K, M = fem.assemble_KM(X, el, 200e9, 0.3, 7800.0)
    Ke = fem.element_stiffness(X[None], fem.constitutive_matrix(200e9, 0.3))[0]
    tol = 1e-15 * abs(Ke).max()
    Kd = K.toarray()
    keep = np.abs(Ke) > tol
    assert np.allclose(Kd[keep], Ke[keep], rtol=0, atol=0)
    assert np.all(Kd[~keep] == 0.0)
    V = np.linalg.det(np.stack([X[1] - X[0], X[2] - X[0], X[3] - X[0]])) / 6
    assert abs(M.toarray()[0::3, 0::3].sum() - 7800 * V) < 1e-12 * 7800 * V
