"""Oracle: TET10 element integrals and triplet assembly (numpy restatement).

TEST INFRASTRUCTURE ONLY.  This module is the CPU checker for the HIP
element/assembly kernels.  Only ``tests/``, ``__graft_entry__.smoke()`` and
the ``cpu_baseline`` leg of ``bench.py`` may import it; the product never
does.

Every function restates the reference algorithm it cites (file:line in the
turbomodal sources).  Arrays are vectorised over elements; the per-element
arithmetic is the reference's: 14-point Keast quadrature for Ke, Me, Ks, Kw
and G, the 4-point rule for the centrifugal load, |detJ| for Ke/Me and the
signed detJ for the rotating-effect integrals.
"""
from __future__ import annotations

import numpy as np
import scipy.sparse as sp

# --- quadrature (src/element.cpp:13-69) ---------------------------------
G4_A = 0.5854101966249685
G4_B = 0.1381966011250105
GAUSS4_PTS = np.array([[G4_B, G4_B, G4_B], [G4_A, G4_B, G4_B],
                       [G4_B, G4_A, G4_B], [G4_B, G4_B, G4_A]])
GAUSS4_W = np.full(4, 1.0 / 24.0)

_K2A, _K2B = 0.698419704324386603, 0.100526765225204467
_K3A, _K3B = 0.0568813795204234229, 0.314372873493192195
KEAST14_PTS = np.array([
    [0.5, 0.5, 0.0], [0.5, 0.0, 0.5], [0.5, 0.0, 0.0],
    [0.0, 0.5, 0.5], [0.0, 0.5, 0.0], [0.0, 0.0, 0.5],
    [_K2B, _K2B, _K2B], [_K2A, _K2B, _K2B], [_K2B, _K2A, _K2B], [_K2B, _K2B, _K2A],
    [_K3B, _K3B, _K3B], [_K3A, _K3B, _K3B], [_K3B, _K3A, _K3B], [_K3B, _K3B, _K3A],
])
KEAST14_W = np.array([0.00317460317460317503] * 6 + [0.01476497079049678] * 4
                     + [0.02213979111426512] * 4)


def shape_functions(xi, eta, zeta):
    """N (10,) at a natural point (src/element.cpp:90-112)."""
    L1, L2, L3, L4 = 1.0 - xi - eta - zeta, xi, eta, zeta
    return np.array([L1 * (2 * L1 - 1), L2 * (2 * L2 - 1), L3 * (2 * L3 - 1),
                     L4 * (2 * L4 - 1), 4 * L1 * L2, 4 * L2 * L3, 4 * L1 * L3,
                     4 * L1 * L4, 4 * L2 * L4, 4 * L3 * L4])


def shape_derivatives(xi, eta, zeta):
    """dN/d(xi,eta,zeta) (10,3) (src/element.cpp:114-201)."""
    L1, L2, L3, L4 = 1.0 - xi - eta - zeta, xi, eta, zeta
    d = np.zeros((10, 3))
    d[0] = 1.0 - 4.0 * L1
    d[1, 0] = 4.0 * L2 - 1.0
    d[2, 1] = 4.0 * L3 - 1.0
    d[3, 2] = 4.0 * L4 - 1.0
    d[4] = [4.0 * (L1 - L2), -4.0 * L2, -4.0 * L2]
    d[5] = [4.0 * L3, 4.0 * L2, 0.0]
    d[6] = [-4.0 * L3, 4.0 * (L1 - L3), -4.0 * L3]
    d[7] = [-4.0 * L4, -4.0 * L4, 4.0 * (L1 - L4)]
    d[8] = [4.0 * L4, 0.0, 4.0 * L2]
    d[9] = [0.0, 4.0 * L4, 4.0 * L3]
    return d


N14 = np.array([shape_functions(*p) for p in KEAST14_PTS])       # (14,10)
DN14 = np.array([shape_derivatives(*p) for p in KEAST14_PTS])    # (14,10,3)
N4 = np.array([shape_functions(*p) for p in GAUSS4_PTS])
DN4 = np.array([shape_derivatives(*p) for p in GAUSS4_PTS])


def constitutive_matrix(E, nu):
    """Isotropic D, Voigt {xx,yy,zz,xy,yz,xz}, engineering shear
    (src/material.cpp:24-31)."""
    c = E / ((1.0 + nu) * (1.0 - 2.0 * nu))
    D = np.zeros((6, 6))
    D[:3, :3] = c * nu
    D[0, 0] = D[1, 1] = D[2, 2] = c * (1.0 - nu)
    D[3, 3] = D[4, 4] = D[5, 5] = c * (1.0 - 2.0 * nu) / 2.0
    return D


def validate_material(E, nu, rho):
    """src/material.cpp:15-22."""
    if E <= 0.0:
        raise ValueError("Young's modulus E must be positive")
    if nu <= 0.0 or nu >= 0.5:
        raise ValueError("Poisson's ratio nu must be in (0, 0.5)")
    if rho <= 0.0:
        raise ValueError("Density rho must be positive")


def _jac(X, dN):
    """J = dNdxi^T X per element and point. X (E,10,3), dN (G,10,3) -> (E,G,3,3)."""
    return np.einsum("gai,eaj->egij", dN, X)


def _b_matrix(dNdx):
    """B (…,6,30) from physical derivatives (…,10,3) (src/element.cpp:213-258)."""
    sh = dNdx.shape[:-2]
    B = np.zeros(sh + (6, 30))
    gx, gy, gz = dNdx[..., 0], dNdx[..., 1], dNdx[..., 2]
    B[..., 0, 0::3] = gx
    B[..., 1, 1::3] = gy
    B[..., 2, 2::3] = gz
    B[..., 3, 0::3] = gy
    B[..., 3, 1::3] = gx
    B[..., 4, 1::3] = gz
    B[..., 4, 2::3] = gy
    B[..., 5, 0::3] = gz
    B[..., 5, 2::3] = gx
    return B


def _geometry14(X):
    J = _jac(X, DN14)
    detJ = np.linalg.det(J)
    Jinv = np.linalg.inv(J)
    dNdx = np.einsum("gak,egjk->egaj", DN14, Jinv)     # dNdxi @ Jinv^T
    return detJ, dNdx


def element_stiffness(X, D):
    """Ke = sum_14 B^T D B |detJ| w  (src/element.cpp:292-335). X (E,10,3)."""
    detJ, dNdx = _geometry14(X)
    B = _b_matrix(dNdx)
    f = np.abs(detJ) * KEAST14_W[None, :]
    return np.einsum("egki,kl,eglj,eg->eij", B, D, B, f, optimize=True)


def element_scalar_mass(X, rho, signed=False):
    """m_ab = rho sum_14 N_a N_b detJ w; |detJ| unless ``signed``.

    Me = m (x) I3 (src/element.cpp:337-371); the spin-softening and
    gyroscopic integrals use the same scalar with the signed detJ
    (src/rotating_effects.cpp:140-228).
    """
    J = _jac(X, DN14)
    detJ = np.linalg.det(J)
    dj = detJ if signed else np.abs(detJ)
    return rho * np.einsum("ga,gb,eg,g->eab", N14, N14, dj, KEAST14_W, optimize=True)


def element_mass(X, rho):
    """Me (E,30,30) = sum_14 rho N_mat^T N_mat |detJ| w (src/element.cpp:337-371)."""
    m = element_scalar_mass(X, rho)
    return np.einsum("eab,cd->eacbd", m, np.eye(3)).reshape(-1, 30, 30)


def element_spin_softening(X, rho, omega, axis):
    """Kw = rho sum_14 N^T omega^2 (I - a a^T) N detJ w (src/rotating_effects.cpp:140-179)."""
    a = np.asarray(axis, float) / np.linalg.norm(axis)
    Om = omega * omega * (np.eye(3) - np.outer(a, a))
    m = element_scalar_mass(X, rho, signed=True)
    return np.einsum("eab,cd->eacbd", m, Om).reshape(-1, 30, 30)


def element_gyroscopic(X, rho, axis):
    """G = rho sum_14 N^T [a x] N detJ w (src/rotating_effects.cpp:181-228)."""
    a = np.asarray(axis, float) / np.linalg.norm(axis)
    S = np.array([[0, -a[2], a[1]], [a[2], 0, -a[0]], [-a[1], a[0], 0]])
    m = element_scalar_mass(X, rho, signed=True)
    return np.einsum("eab,cd->eacbd", m, S).reshape(-1, 30, 30)


def element_centrifugal_load(X, rho, omega, axis):
    """Fe (E,30), 4-point rule, signed detJ (src/rotating_effects.cpp:5-51)."""
    a = np.asarray(axis, float) / np.linalg.norm(axis)
    J = _jac(X, DN4)
    detJ = np.linalg.det(J)
    pos = np.einsum("ga,eaj->egj", N4, X)
    rperp = pos - np.einsum("egj,j->eg", pos, a)[..., None] * a
    f = rho * omega * omega * np.einsum("ga,egc,eg,g->eac", N4, rperp, detJ, GAUSS4_W)
    return f.reshape(-1, 30)


def element_prestress(X, D, u_e):
    """sigma at the 14 points = D B(gp) u_e (src/assembler.cpp:102-110). -> (E,14,6)."""
    detJ, dNdx = _geometry14(X)
    B = _b_matrix(dNdx)
    return np.einsum("kl,eglj,ej->egk", D, B, u_e, optimize=True)


def element_stress_stiffening(X, sigma):
    """Ks = sum_14 B_NL^T S B_NL detJ w, S = blkdiag(s,s,s) (src/rotating_effects.cpp:53-138)."""
    detJ, dNdx = _geometry14(X)
    E = X.shape[0]
    BNL = np.zeros((E, 14, 9, 30))
    for c in range(3):
        for d in range(3):
            BNL[:, :, 3 * c + d, c::3] = dNdx[..., d]
    s = sigma
    sm = np.empty(s.shape[:2] + (3, 3))
    sm[..., 0, 0], sm[..., 1, 1], sm[..., 2, 2] = s[..., 0], s[..., 1], s[..., 2]
    sm[..., 0, 1] = sm[..., 1, 0] = s[..., 3]
    sm[..., 1, 2] = sm[..., 2, 1] = s[..., 4]
    sm[..., 0, 2] = sm[..., 2, 0] = s[..., 5]
    S = np.zeros(s.shape[:2] + (9, 9))
    for c in range(3):
        S[..., 3 * c:3 * c + 3, 3 * c:3 * c + 3] = sm
    f = detJ * KEAST14_W[None, :]
    return np.einsum("egki,egkl,eglj,eg->eij", BNL, S, BNL, f, optimize=True)


def dof_map(elements):
    """dof_map(3n+c) = 3*node+c (src/assembler.cpp:21-29)."""
    el = np.asarray(elements, dtype=np.int64)
    return (3 * el[:, :, None] + np.arange(3)[None, None, :]).reshape(el.shape[0], 30)


def assemble_matrix(Ke, elements, ndof):
    """Triplets with |v| > 1e-15 max|Ke| per element, summed into CSC
    (src/assembler.cpp:274-289, 65-68)."""
    dm = dof_map(elements)
    tol = 1e-15 * np.abs(Ke).reshape(Ke.shape[0], -1).max(axis=1)
    keep = np.abs(Ke) > tol[:, None, None]
    rows = np.broadcast_to(dm[:, :, None], Ke.shape)[keep]
    cols = np.broadcast_to(dm[:, None, :], Ke.shape)[keep]
    return sp.csc_matrix((Ke[keep], (rows, cols)), shape=(ndof, ndof))


def assemble_vector(Fe, elements, ndof):
    """F(dof_map(i)) += Fe(i) (src/assembler.cpp:253-262)."""
    dm = dof_map(elements)
    F = np.zeros(ndof)
    np.add.at(F, dm.ravel(), Fe.ravel())
    return F


def _chunks(E, size=4096):
    for s in range(0, E, size):
        yield slice(s, min(E, s + size))


def assemble_KM(nodes, elements, E, nu, rho):
    """GlobalAssembler::assemble (src/assembler.cpp:11-69)."""
    D = constitutive_matrix(E, nu)
    ndof = 3 * nodes.shape[0]
    K = sp.csc_matrix((ndof, ndof))
    M = sp.csc_matrix((ndof, ndof))
    for s in _chunks(elements.shape[0]):
        X = nodes[elements[s]]
        K = K + assemble_matrix(element_stiffness(X, D), elements[s], ndof)
        M = M + assemble_matrix(element_mass(X, rho), elements[s], ndof)
    return K.tocsc(), M.tocsc()


def assemble_rotating(nodes, elements, rho, omega, axis):
    """GlobalAssembler::assemble_rotating_effects -> (K_omega, G) (src/assembler.cpp:157-218)."""
    ndof = 3 * nodes.shape[0]
    Kw = sp.csc_matrix((ndof, ndof))
    G = sp.csc_matrix((ndof, ndof))
    for s in _chunks(elements.shape[0]):
        X = nodes[elements[s]]
        Kw = Kw + assemble_matrix(element_spin_softening(X, rho, omega, axis), elements[s], ndof)
        G = G + assemble_matrix(element_gyroscopic(X, rho, axis), elements[s], ndof)
    return Kw.tocsc(), G.tocsc()


def assemble_centrifugal(nodes, elements, rho, omega, axis):
    """GlobalAssembler::assemble_centrifugal_load (src/assembler.cpp:220-272)."""
    ndof = 3 * nodes.shape[0]
    F = np.zeros(ndof)
    for s in _chunks(elements.shape[0]):
        X = nodes[elements[s]]
        F += assemble_vector(element_centrifugal_load(X, rho, omega, axis), elements[s], ndof)
    return F


def assemble_stress_stiffening(nodes, elements, E, nu, u):
    """GlobalAssembler::assemble_stress_stiffening (src/assembler.cpp:71-155)."""
    D = constitutive_matrix(E, nu)
    ndof = 3 * nodes.shape[0]
    Ks = sp.csc_matrix((ndof, ndof))
    dm = dof_map(elements)
    for s in _chunks(elements.shape[0]):
        X = nodes[elements[s]]
        sig = element_prestress(X, D, u[dm[s]])
        Ks = Ks + assemble_matrix(element_stress_stiffening(X, sig), elements[s], ndof)
    return Ks.tocsc()
