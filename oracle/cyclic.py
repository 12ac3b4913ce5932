"""Oracle: cyclic-symmetry reduction and shift-invert Lanczos (numpy/scipy).

TEST INFRASTRUCTURE ONLY.  This module is the CPU checker for the HIP
cyclic modal solve.  Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import it; the product never does.

It restates, step by step, the reference path behind ``tm.solve``:
boundary matching (``src/mesh.cpp:260-351``), DOF classes
(``src/cyclic_solver.cpp:214-238``), T(k) (``240-330``), the static
centrifugal solve (``456-558``), the phase-split projections
(``560-723``), ``solve_at_rpm``/``solve_harmonic`` (``725-1357``) and the
Hermitian shift-invert Lanczos with implicit restart
(``src/hermitian_lanczos.cpp:12-460``).  Sparse LU (SuperLU via scipy)
stands in for Eigen's SimplicialLDLT; both are exact direct solvers, so only
rounding differs.  Spectra (k = 0, N/2 in the reference) is absent from the
reference tree; those harmonics use the same Lanczos on the real matrices.

Two modes:
* ``policy="reference"`` - the reference's nev/ncv/sigma/tolerance policy and
  its (effectively absolute) convergence test;
* ``policy="tight"`` - the same algorithm iterated until every wanted Ritz
  pair has relative shift-invert residual <= 1e-12; this is the FP64 oracle
  that accuracy targets are measured against.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np
import scipy.linalg as sla
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from . import fem

PI = math.pi


# --------------------------------------------------------------------------
# mesh side
# --------------------------------------------------------------------------
def axis_planes(rotation_axis):
    """(c1, c2, axial) component indices (src/mesh.cpp:272-283)."""
    return {0: (1, 2, 0), 1: (0, 2, 1)}.get(rotation_axis, (0, 1, 2))


def match_boundary_nodes(nodes, left, right, num_sectors, rotation_axis=2):
    """Greedy nearest match of each left node to an unused right node after
    rotating right nodes by -alpha; tolerance max(1e-6, 1e-4 r)
    (src/mesh.cpp:260-351).  Returns (left, right, pairs, free)."""
    alpha = 2.0 * PI / num_sectors
    ca, sa = math.cos(-alpha), math.sin(-alpha)
    c1, c2, cax = axis_planes(rotation_axis)
    used = np.zeros(len(right), dtype=bool)
    R = nodes[np.asarray(right, dtype=np.int64)]
    r1 = ca * R[:, c1] - sa * R[:, c2]
    r2 = sa * R[:, c1] + ca * R[:, c2]
    rax = R[:, cax]
    new_l, new_r, pairs, free = [], [], [], []
    for lid in left:
        l1, l2, lax = nodes[lid, c1], nodes[lid, c2], nodes[lid, cax]
        d = np.sqrt((l1 - r1) ** 2 + (l2 - r2) ** 2 + (lax - rax) ** 2)
        d[used] = np.inf
        j = int(np.argmin(d)) if len(d) else -1
        tol = max(1e-6, 1e-4 * math.sqrt(l1 * l1 + l2 * l2))
        if j < 0 or not np.isfinite(d[j]) or d[j] > tol:
            free.append(int(lid))
            continue
        used[j] = True
        new_l.append(int(lid))
        new_r.append(int(right[j]))
        pairs.append((int(lid), int(right[j])))
    free += [int(right[j]) for j in range(len(right)) if not used[j]]
    return new_l, new_r, pairs, sorted(set(free))


@dataclass
class CyclicModel:
    nodes: np.ndarray
    elements: np.ndarray
    num_sectors: int
    left: list
    right: list
    pairs: list
    constrained: list           # full DOF ids
    rotation_axis: int = 2
    interior_dofs: list = field(default_factory=list)
    left_dofs: list = field(default_factory=list)
    right_dofs: list = field(default_factory=list)

    @property
    def ndof(self):
        return 3 * self.nodes.shape[0]


def build_model(nodes, elements, node_sets, num_sectors, rotation_axis=2, hub="fixed"):
    """Mesh::load_from_arrays + CyclicSymmetrySolver ctor
    (src/mesh.cpp:208-258, src/cyclic_solver.cpp:88-238)."""
    nodes = np.asarray(nodes, float)
    elements = np.asarray(elements, np.int64)
    left0 = sorted(int(i) for i in node_sets["left_boundary"])
    right0 = sorted(int(i) for i in node_sets["right_boundary"])
    left, right, pairs, _free = match_boundary_nodes(nodes, left0, right0, num_sectors,
                                                      rotation_axis)
    constrained = []
    if hub == "fixed" and "hub_constraint" in node_sets:
        for n in sorted(set(int(i) for i in node_sets["hub_constraint"])):
            constrained += [3 * n, 3 * n + 1, 3 * n + 2]
    m = CyclicModel(nodes, elements, num_sectors, left, right, pairs,
                    sorted(set(constrained)), rotation_axis)
    ls, rs = set(left), set(right)
    for i in range(nodes.shape[0]):
        for d in range(3):
            dof = 3 * i + d
            if i in rs:
                m.right_dofs.append(dof)
            elif i in ls:
                m.left_dofs.append(dof)
            else:
                m.interior_dofs.append(dof)
    return m


def cyclic_transformation(m: CyclicModel, k):
    """T(k) (3 n_nodes x n_reduced) complex (src/cyclic_solver.cpp:240-330)."""
    alpha = 2.0 * PI / m.num_sectors
    phase = complex(math.cos(k * alpha), math.sin(k * alpha))
    red = m.interior_dofs + m.left_dofs
    f2r = {d: i for i, d in enumerate(red)}
    rows, cols, vals = [], [], []
    for d in red:
        rows.append(d); cols.append(f2r[d]); vals.append(1.0 + 0j)
    r2l = {r: l for l, r in m.pairs}
    ca, sa = math.cos(alpha), math.sin(alpha)
    c1, c2, ax = axis_planes(m.rotation_axis)
    seen = set()
    for rd in m.right_dofs:
        rn = rd // 3
        if rn in seen:
            continue
        seen.add(rn)
        if rn not in r2l:
            raise RuntimeError(f"Right boundary node {rn} has no matched left partner")
        ln = r2l[rn]
        b = 3 * rn
        l1, l2, la = f2r[3 * ln + c1], f2r[3 * ln + c2], f2r[3 * ln + ax]
        rows += [b + c1, b + c1, b + c2, b + c2, b + ax]
        cols += [l1, l2, l1, l2, la]
        vals += [phase * ca, phase * (-sa), phase * sa, phase * ca, phase]
    return sp.csc_matrix((np.array(vals), (rows, cols)), shape=(m.ndof, len(red)))


def reduced_free_map(m: CyclicModel):
    """free_reduced_map: reduced indices that are not constrained
    (src/cyclic_solver.cpp:573-584, 633-642)."""
    red = m.interior_dofs + m.left_dofs
    cs = set(m.constrained)
    return np.array([i for i, d in enumerate(red) if d not in cs], dtype=np.int64)


def _partition(A, right_set):
    """same: row-rightness == col-rightness; nr: row not right, col right
    (src/cyclic_solver.cpp:606-626)."""
    A = A.tocoo()
    rr = np.isin(A.row, right_set)
    cr = np.isin(A.col, right_set)
    same = rr == cr
    nr = (~rr) & cr
    n = A.shape[0]
    As = sp.csc_matrix((A.data[same], (A.row[same], A.col[same])), shape=(n, n))
    Anr = sp.csc_matrix((A.data[nr], (A.row[nr], A.col[nr])), shape=(n, n))
    return As, Anr


def project_phase_split(m: CyclicModel, A):
    """(A_c, A_p) restricted to free DOFs: A(k) = A_c + P A_p + conj(P) A_p^H
    (src/cyclic_solver.cpp:560-671, 942-947)."""
    T0 = cyclic_transformation(m, 0)
    right = np.array(m.right_dofs, dtype=np.int64)
    As, Anr = _partition(sp.csc_matrix(A, dtype=complex), right)
    Ac = (T0.conj().T @ As @ T0).tocsc()
    Ap = (T0.conj().T @ Anr @ T0).tocsc()
    fm = reduced_free_map(m)
    return Ac[fm][:, fm].tocsc(), Ap[fm][:, fm].tocsc()


def harmonic_matrix(Ac, Ap, k, N):
    alpha = 2.0 * PI / N
    P = complex(math.cos(k * alpha), math.sin(k * alpha))
    return (Ac + P * Ap + np.conj(P) * Ap.conj().T).tocsc()


def static_centrifugal(m: CyclicModel, K, F):
    """u_static: k=0 cyclic BCs, constrained DOFs removed, sparse LU
    (src/cyclic_solver.cpp:456-558)."""
    T0 = cyclic_transformation(m, 0).real
    T0 = sp.csc_matrix(T0)
    T0.data[np.abs(T0.data) <= 1e-15] = 0.0
    T0.eliminate_zeros()
    Kc = (T0.T @ K @ T0).tocsc()
    Fc = T0.T @ F
    fm = reduced_free_map(m)
    if fm.size == 0:
        return np.zeros(m.ndof)
    Kf = Kc[fm][:, fm].tocsc()
    uf = spla.splu(Kf, permc_spec="MMD_AT_PLUS_A").solve(Fc[fm])
    uc = np.zeros(T0.shape[1])
    uc[fm] = uf
    return T0 @ uc


# --------------------------------------------------------------------------
# tridiagonal eigen + Lanczos (src/hermitian_lanczos.cpp)
# --------------------------------------------------------------------------
def _tridiag_eig(alpha, beta):
    n = len(alpha)
    if n == 1:
        return np.array([alpha[0]]), np.ones((1, 1))
    return sla.eigh_tridiagonal(alpha, beta[: n - 1])


def implicit_restart(V, MV, alpha, beta, nev, ncv, Mop):
    """Exact-shift implicit QR restart (src/hermitian_lanczos.cpp:114-260)."""
    ritz, _ = _tridiag_eig(alpha[:ncv], beta[:ncv])
    shifts = np.sort(ritz)[: ncv - nev]
    Q = np.eye(ncv)
    for mu in shifts:
        x = alpha[0] - mu
        z = beta[0]
        for i in range(ncv - 1):
            r = math.hypot(x, z)
            if r < 1e-300:
                c, s = 1.0, 0.0
            else:
                c, s = x / r, z / r
            if i > 0:
                beta[i - 1] = r
            a0, a1, b0 = alpha[i], alpha[i + 1], beta[i]
            alpha[i] = c * c * a0 + 2.0 * c * s * b0 + s * s * a1
            alpha[i + 1] = s * s * a0 - 2.0 * c * s * b0 + c * c * a1
            beta[i] = c * s * (a1 - a0) + (c * c - s * s) * b0
            if i + 2 < ncv:
                bulge = s * beta[i + 1]
                beta[i + 1] = c * beta[i + 1]
                x, z = beta[i], bulge
            q0 = Q[:, i].copy()
            q1 = Q[:, i + 1].copy()
            Q[:, i] = c * q0 + s * q1
            Q[:, i + 1] = -s * q0 + c * q1
    beta_tri = beta[nev - 1]
    beta_dangling = beta[ncv - 1]
    f = None
    if nev < ncv:
        f = beta_tri * (V[:, :ncv] @ Q[:, nev]) + beta_dangling * Q[ncv - 1, nev - 1] * V[:, ncv]
    V[:, :nev] = V[:, :ncv] @ Q[:, :nev]
    if nev < ncv:
        Mf = Mop(f)
        bn = math.sqrt(abs(np.real(np.vdot(f, Mf))))
        if bn > 1e-14:
            V[:, nev] = f / bn
            MV[:, nev] = Mf / bn
            beta[nev - 1] = bn
        else:
            beta[nev - 1] = 0.0
    for j in range(nev):
        MV[:, j] = Mop(V[:, j])


@dataclass
class LanczosResult:
    eigenvalues: np.ndarray
    eigenvectors: np.ndarray
    nconv: int
    converged: bool
    restarts: int = 0
    message: str = ""


def lanczos(K, M, nev, ncv=0, shift=1.0, tol=1e-8, max_iterations=1000,
            policy="reference", rel_tol=1e-12, seed=42):
    """HermitianLanczosEigenSolver::solve (src/hermitian_lanczos.cpp:264-460)."""
    n = K.shape[0]
    nev = min(nev, n - 1)
    if nev <= 0:
        return LanczosResult(np.zeros(0), np.zeros((n, 0)), 0, False)
    if n <= 200:
        Kd = K.toarray(); Md = M.toarray()
        Kd = 0.5 * (Kd + Kd.conj().T); Md = 0.5 * (Md + Md.conj().T)
        w, X = sla.eigh(Kd, Md)
        o = np.argsort(w)[:nev]
        return LanczosResult(w[o], X[:, o], nev, True, 0, "Dense solver")
    if ncv <= 0:
        ncv = min(max(4 * nev + 1, 40), n)
    ncv = min(ncv, n)
    A = (K - shift * M).tocsc()
    cplx = np.iscomplexobj(A.data) and np.any(A.data.imag != 0)
    lu_ = spla.splu(A if cplx else sp.csc_matrix(A.real), permc_spec="MMD_AT_PLUS_A")

    class _Solve:
        @staticmethod
        def solve(b):
            if cplx:
                return lu_.solve(b)
            return lu_.solve(np.ascontiguousarray(b.real)) + 1j * lu_.solve(np.ascontiguousarray(b.imag))
    lu = _Solve
    Mop = lambda x: M @ x
    V = np.zeros((n, ncv + 1), complex)
    MV = np.zeros((n, ncv + 1), complex)
    alpha = np.zeros(ncv)
    beta = np.zeros(ncv)
    rng = np.random.default_rng(seed)
    v0 = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    Mv0 = Mop(v0)
    mn = math.sqrt(np.real(np.vdot(v0, Mv0)))
    V[:, 0] = v0 / mn
    MV[:, 0] = Mv0 / mn
    for restart in range(max_iterations + 1):
        j_start = 0 if restart == 0 else nev
        for j in range(j_start, ncv):
            w = lu.solve(MV[:, j])
            alpha[j] = np.real(np.vdot(MV[:, j], w))
            w = w - alpha[j] * V[:, j]
            if j > 0:
                w = w - beta[j - 1] * V[:, j - 1]
            for kk in range(j + 1):
                h = np.vdot(MV[:, kk], w)
                w = w - h * V[:, kk]
            Mw = Mop(w)
            bj = math.sqrt(abs(np.real(np.vdot(w, Mw))))
            if bj < 1e-14:
                r2 = np.random.default_rng(seed + restart * 100 + j)
                w = r2.standard_normal(n) + 1j * r2.standard_normal(n)
                for kk in range(j + 1):
                    w = w - np.vdot(MV[:, kk], w) * V[:, kk]
                Mw = Mop(w)
                bj = math.sqrt(abs(np.real(np.vdot(w, Mw))))
                if bj < 1e-14:
                    ncv = j + 1
                    break
            beta[j] = bj
            V[:, j + 1] = w / bj
            MV[:, j + 1] = Mw / bj
        act = ncv
        theta, S = _tridiag_eig(alpha[:act], beta[:act])
        lam = np.where(np.abs(theta) > 1e-15, shift + 1.0 / np.where(theta == 0, 1, theta), 1e30)
        idx = np.argsort(lam, kind="stable")
        bl = beta[act - 1]
        nconv = 0
        for i in range(min(nev, act)):
            ii = idx[i]
            res = abs(bl * S[act - 1, ii])
            if policy == "reference":
                ok = res < tol * max(1.0, abs(theta[ii]))
            else:
                ok = res <= rel_tol * abs(theta[ii])
            nconv += int(ok)
        if nconv >= nev or restart == max_iterations:
            ne = min(nconv if nconv > 0 else nev, act, nev)
            X = V[:, :act] @ S[:, idx[:ne]]
            return LanczosResult(lam[idx[:ne]], X, ne, nconv >= nev, restart,
                                 "Converged" if nconv >= nev else "Not converged")
        implicit_restart(V, MV, alpha, beta, nev, act, Mop)
    raise RuntimeError("unreachable")


# --------------------------------------------------------------------------
# solve_at_rpm (src/cyclic_solver.cpp:725-1357)
# --------------------------------------------------------------------------
@dataclass
class HarmonicResult:
    harmonic_index: int
    rpm: float
    frequencies: np.ndarray
    mode_shapes: np.ndarray          # (3 n_nodes, M) complex, full sector DOFs
    whirl_direction: np.ndarray
    converged: bool
    eigenvalues: np.ndarray = None   # reduced-space lambda before duplication
    reduced_vectors: np.ndarray = None
    K: object = None                 # free-space K(k), M(k) used
    M: object = None


@dataclass
class SectorOperators:
    """Free-space phase-split operators of one (mesh, material, rpm)."""
    Kc: object
    Kp: object
    Mc: object
    Mp: object
    K_eff_full: object = None
    M_full: object = None
    u_static: np.ndarray = None
    Gc: object = None                # gyroscopic G (omega-free) phase split (cyclic_solver.cpp:858-899)
    Gp: object = None
    omega: float = 0.0


def sector_operators(m: CyclicModel, E, nu, rho, rpm, springs=None):
    """Steps 1-3 of solve_at_rpm (src/cyclic_solver.cpp:735-818).  springs: optional
    3 n_nodes grounded spring stiffnesses (ELASTIC_SUPPORT, build_spring_stiffness,
    src/cyclic_solver.cpp:193-212) added to K before the rotating terms (step 1b,
    747-751); the static prestress solve keeps the base K (static_centrifugal uses
    assembler_.K(), 456-470)."""
    omega = rpm * 2.0 * PI / 60.0
    K, M = fem.assemble_KM(m.nodes, m.elements, E, nu, rho)
    K_eff = K
    if springs is not None and np.any(np.asarray(springs) > 0):
        K_eff = (K + sp.diags(np.maximum(np.asarray(springs, float), 0.0))).tocsc()
    u = None
    Gc = Gp = None
    if abs(omega) > 0.0:
        ax = np.eye(3)[m.rotation_axis]
        Kw, G = fem.assemble_rotating(m.nodes, m.elements, rho, omega, ax)
        Gc, Gp = project_phase_split(m, G)
        K_eff = K_eff - Kw
        if len(m.constrained) > 0:
            F = fem.assemble_centrifugal(m.nodes, m.elements, rho, omega, ax)
            u = static_centrifugal(m, K, F)
            Ks = fem.assemble_stress_stiffening(m.nodes, m.elements, E, nu, u)
            K_eff = K_eff + Ks
    Kc, Kp = project_phase_split(m, K_eff)
    Mc, Mp = project_phase_split(m, M)
    return SectorOperators(Kc, Kp, Mc, Mp, K_eff.tocsc(), M, u, Gc, Gp, omega)


def expand_modes(m: CyclicModel, X, k):
    """zeros at constrained DOFs -> T0 x -> right rows * P (src/cyclic_solver.cpp:1185-1211)."""
    fm = reduced_free_map(m)
    T0 = cyclic_transformation(m, 0)
    nred = T0.shape[1]
    Xr = np.zeros((nred, X.shape[1]), complex)
    Xr[fm] = X
    full = T0 @ Xr
    alpha = 2.0 * PI / m.num_sectors
    P = complex(math.cos(k * alpha), math.sin(k * alpha))
    full[np.array(m.right_dofs, dtype=np.int64)] *= P
    return full


def solve_harmonic(m: CyclicModel, ops: SectorOperators, k, num_modes, rpm,
                   policy="reference"):
    """solve_harmonic (src/cyclic_solver.cpp:937-1224), GEP branch."""
    N = m.num_sectors
    max_k = N // 2
    n_free = ops.Kc.shape[0]
    if n_free <= 1:
        return None
    Kk = harmonic_matrix(ops.Kc, ops.Kp, k, N)
    Mk = harmonic_matrix(ops.Mc, ops.Mp, k, N)
    degenerate = k > 0 and not (N % 2 == 0 and k == max_k)
    nev_target = (num_modes + 1) // 2 if degenerate else num_modes
    nev = min(nev_target, n_free - 1)
    ncv = min(max(4 * nev + 1, 40), n_free)
    real_case = (k == 0) or (N % 2 == 0 and k == max_k)
    if real_case:
        Kk = sp.csc_matrix(Kk.real); Mk = sp.csc_matrix(Mk.real)
    res = lanczos(Kk, Mk, nev, ncv, 1.0, 1e-8, 1000, policy=policy)
    if res.nconv == 0:
        return None
    lam = res.eigenvalues[:nev_target]
    X = res.eigenvectors[:, :nev_target]
    freqs = np.where(lam > 0, np.sqrt(np.maximum(lam, 0)) / (2 * PI), 0.0)
    whirl = np.zeros(len(freqs), np.int32)
    vecs = X
    if degenerate:
        nexp = min(2 * len(freqs), num_modes)
        f2, v2, w2 = [], [], []
        for q in range(nexp):
            src = q // 2
            f2.append(freqs[src])
            if q % 2 == 0:
                v2.append(X[:, src]); w2.append(1)
            else:
                v2.append(np.conj(X[:, src])); w2.append(-1)
        freqs = np.array(f2); vecs = np.array(v2).T; whirl = np.array(w2, np.int32)
    shapes = expand_modes(m, vecs, k)
    return HarmonicResult(k, rpm, freqs, shapes, whirl, res.converged, lam, X, Kk, Mk)


# --------------------------------------------------------------------------
# Coriolis: Lancaster-linearised QEP (src/modal_solver.cpp:149-312,
# src/cyclic_solver.cpp:986-1005)
# --------------------------------------------------------------------------
def lancaster_pencil(Kk, Mk, Dk):
    """L1 = [[D, K], [K, 0]], L2 = [[M, 0], [0, K]] (build_lancaster_L1/L2,
    src/modal_solver.cpp:151-198): L1 z = w L2 z with z = [w phi; phi] is
    (K + w D - w^2 M) phi = 0."""
    L1 = sp.bmat([[Dk, Kk], [Kk, None]], format="csc")
    L2 = sp.bmat([[Mk, None], [None, Kk]], format="csc")
    return L1, L2


def coriolis_matrix(ops: SectorOperators, k, N):
    """D = i 2 omega G(k) (src/cyclic_solver.cpp:986-994) with G(k) = T(k)^H G T(k).
    G is skew-symmetric, so the right-row coupling block of the phase split is
    -Gp^T: G(k) = Gc + P Gp - conj(P) Gp^T, skew-Hermitian, and D Hermitian (the
    property tests/test_cyclic.cpp:678-731 checks on the direct projection).  The
    reference's production code adds +conj(P) Gp^H, the form of the symmetric K and
    M, which leaves its D non-Hermitian; this follows the projection instead."""
    alpha = 2.0 * PI / N
    P = complex(math.cos(k * alpha), math.sin(k * alpha))
    return (2j * ops.omega) * (ops.Gc + P * ops.Gp - np.conj(P) * ops.Gp.T).tocsc()


def select_qep(omega, vectors, num_modes, n):
    """Selection of solve_lancaster_qep (src/modal_solver.cpp:258-307): drop |w| < 1e-6,
    sort by |w|, keep up to num_modes forward (w > 0) and num_modes backward (w < 0),
    re-sort by |w|; frequency |w| / 2 pi, whirl sign(w), mode = head(n) of z."""
    keep = [i for i in range(len(omega)) if abs(omega[i]) >= 1e-6]
    keep.sort(key=lambda i: abs(omega[i]))
    sel, nf, nb = [], 0, 0
    for i in keep:
        if omega[i] > 0 and nf < num_modes:
            sel.append(i); nf += 1
        elif omega[i] < 0 and nb < num_modes:
            sel.append(i); nb += 1
        if nf >= num_modes and nb >= num_modes:
            break
    sel.sort(key=lambda i: abs(omega[i]))
    freqs = np.array([abs(omega[i]) / (2 * PI) for i in sel])
    whirl = np.array([1 if omega[i] > 0 else -1 for i in sel], np.int32)
    shapes = vectors[:n, sel] if len(sel) else np.zeros((n, 0), complex)
    return freqs, whirl, shapes, np.array([omega[i] for i in sel])


def lancaster_inverse(Kk, Dk):
    """L1^-1 as a LinearOperator through one sparse factor of K(k):
    L1 [a; b] = [c; d] gives K a = d and D a + K b = c, so a = K^-1 d and
    b = K^-1 (c - D a)."""
    n = Kk.shape[0]
    lu_ = spla.splu(sp.csc_matrix(Kk), permc_spec="MMD_AT_PLUS_A")
    Dc = sp.csr_matrix(Dk)

    def mv(z):
        z = np.asarray(z).reshape(-1)
        a = lu_.solve(np.ascontiguousarray(z[n:], dtype=complex))
        b = lu_.solve(np.ascontiguousarray(z[:n] - Dc @ a, dtype=complex))
        return np.concatenate([a, b])

    return spla.LinearOperator((2 * n, 2 * n), matvec=mv, dtype=complex)


def solve_qep(Kk, Mk, Dk, num_modes, dense_max=400):
    """The forward/backward gyroscopic modes the reference's selection keeps: the
    eigenvalues w of the Lancaster pencil with the smallest |w| of each sign.
    Dense generalised eigh of (L1, L2) for 2n <= dense_max (the reference's dense
    path covers 2n <= 200, src/modal_solver.cpp:224-226, and returns every
    eigenvalue); above that, scipy eigsh with shift-invert at 0 (L1^-1 applied
    through a factor of K(k), lancaster_inverse), which finds exactly the
    eigenvalues of smallest magnitude, to ARPACK's tolerance tol=1e-13.  (The
    reference's Lanczos for large systems uses sigma = sqrt(||K||_F / ||M||_F) with
    ascending-lambda selection, whose result depends on its iteration history:
    parity unpinned there.)"""
    n = Kk.shape[0]
    L1, L2 = lancaster_pencil(Kk, Mk, Dk)
    if 2 * n <= dense_max:
        A = L1.toarray(); B = L2.toarray()
        A = 0.5 * (A + A.conj().T); B = 0.5 * (B + B.conj().T)
        w, Z = sla.eigh(A, B)
    else:
        nev = min(2 * num_modes + 6, 2 * n - 2)
        w, Z = spla.eigsh(L1, k=nev, M=L2, sigma=0.0, which="LM", OPinv=lancaster_inverse(Kk, Dk),
                          tol=1e-13)
    return select_qep(w, Z, num_modes, n) + (L1, L2)


def solve_harmonic_qep(m: CyclicModel, ops: SectorOperators, k, num_modes, rpm, dense_max=400):
    """QEP branch of solve_harmonic for k > 0 with Coriolis (src/cyclic_solver.cpp:986-1005,
    1058-1096): no duplication; truncated to num_modes."""
    N = m.num_sectors
    Kk = harmonic_matrix(ops.Kc, ops.Kp, k, N)
    Mk = harmonic_matrix(ops.Mc, ops.Mp, k, N)
    Dk = coriolis_matrix(ops, k, N)
    freqs, whirl, heads, om, L1, L2 = solve_qep(Kk, Mk, Dk, num_modes, dense_max)
    freqs, whirl, heads, om = freqs[:num_modes], whirl[:num_modes], heads[:, :num_modes], om[:num_modes]
    shapes = expand_modes(m, heads, k)
    r = HarmonicResult(k, rpm, freqs, shapes, whirl, True, om, heads, Kk, Mk)
    r.D = Dk
    return r


def solve_at_rpm(m: CyclicModel, E, nu, rho, rpm, num_modes, harmonics=None,
                 policy="reference", ops=None):
    ops = ops or sector_operators(m, E, nu, rho, rpm)
    max_k = m.num_sectors // 2
    ks = list(range(max_k + 1)) if not harmonics else [k for k in harmonics if 0 <= k <= max_k]
    out = []
    for k in ks:
        r = solve_harmonic(m, ops, k, num_modes, rpm, policy)
        if r is not None:
            out.append(r)
    return out


def modal_mass(M_sector, shapes, num_sectors):
    """N * Re(phi^H M phi) per mode (python/turbomodal/solver.py:91-103)."""
    Mphi = M_sector @ shapes
    return num_sectors * np.real(np.sum(shapes.conj() * Mphi, axis=0))


def compute_mac(phi_a, phi_b):
    """|a^H b|^2 / (|a|^2 |b|^2) (python/turbomodal/mac.py:13-40)."""
    a = np.atleast_2d(np.asarray(phi_a, dtype=np.complex128))
    b = np.atleast_2d(np.asarray(phi_b, dtype=np.complex128))
    na = np.real(np.sum(a * np.conj(a), axis=1))
    nb = np.real(np.sum(b * np.conj(b), axis=1))
    cross = np.abs(a @ b.conj().T) ** 2
    return cross / np.maximum(np.outer(na, nb), 1e-300)
