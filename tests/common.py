"""Shared helpers for the tests: oracle models and engine meshes built from
the same inputs."""
from __future__ import annotations

import os

import numpy as np

from oracle import cyclic, gmsh
from turbomodal_amd import meshgen

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def fixture_mesh(name):
    """(nodes, elements, node_sets, num_sectors) of a reference fixture."""
    X, T, S = gmsh.load_msh(os.path.join(GOLDEN, name))
    n = {"wedge_sector.msh": 24, "blade_sector.msh": 24, "leissa_disk_sector.msh": 24}[name]
    return X, T, S, n


def oracle_model(nodes, elements, node_sets, num_sectors, hub="fixed"):
    return cyclic.build_model(nodes, elements, node_sets, num_sectors, 2, hub)


def engine_mesh(ctx, model):
    from turbomodal_amd import capi
    cls, partner, cm = capi.mesh_arrays_from_model(model.nodes, model.elements, model.left,
                                                    model.right, model.pairs, model.constrained)
    return capi.EngineMesh(ctx, model.nodes, model.elements, cls, partner, cm, model.num_sectors, 2)


def generated(name, num_sectors=None):
    sm = meshgen.config_mesh(name, num_sectors)
    return sm.nodes, sm.elements, sm.node_sets, sm.num_sectors


def rel_err(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    return np.max(np.abs(a - b) / np.maximum(np.abs(b), 1e-300))


def load_golden(name):
    """Oracle answers written by tests/golden/make_golden.py (plain arrays, no pickle)."""
    with np.load(os.path.join(GOLDEN, name), allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def subset_mac(a, b):
    """MAC |a^H b|^2 / (|a|^2 |b|^2) of two vectors (on a DOF subset)."""
    a = np.asarray(a, complex)
    b = np.asarray(b, complex)
    return abs(np.vdot(a, b)) ** 2 / (np.vdot(a, a).real * np.vdot(b, b).real)


def free_vector(m, phi):
    """Full-sector mode shape -> the reference's free-DOF vector (interior DOFs, then
    left DOFs, constrained removed)."""
    red = np.array(m.interior_dofs + m.left_dofs)
    return phi[red[cyclic.reduced_free_map(m)]]


def exported(op, rpm, k, which, n):
    """The engine's assembled K_eff(k) (0), M(k) (1) or G(k) (2) as a scipy CSR matrix."""
    import scipy.sparse as sp
    r, c, v = op.export(rpm, k, which)
    return sp.csr_matrix((v, (r, c)), shape=(n, n))


def residual_floor(K, M, x, lam):
    """||K x - lam M x|| / ||K x|| and its FP64 rounding floor
    eps || |K||x| + lam |M||x| || / ||K x||."""
    Kx = K @ x
    nk = np.linalg.norm(Kx)
    r = Kx - lam * (M @ x)
    fl = np.finfo(float).eps * np.linalg.norm(abs(K) @ np.abs(x) + abs(lam) * (abs(M) @ np.abs(x))) / nk
    return np.linalg.norm(r) / nk, fl
