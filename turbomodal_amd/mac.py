"""Modal Assurance Criterion (python/turbomodal/mac.py:13-56)."""
from __future__ import annotations

import numpy as np


def compute_mac(phi_a, phi_b):
    """MAC[i, j] = |phi_a[i]^H phi_b[j]|^2 / (||phi_a[i]||^2 ||phi_b[j]||^2)."""
    a = np.atleast_2d(np.asarray(phi_a, dtype=np.complex128))
    b = np.atleast_2d(np.asarray(phi_b, dtype=np.complex128))
    na = np.real(np.sum(a * np.conj(a), axis=1))
    nb = np.real(np.sum(b * np.conj(b), axis=1))
    cross = np.abs(a @ b.conj().T) ** 2
    return cross / np.maximum(np.outer(na, nb), 1e-300)


def compute_auto_mac(phi):
    return compute_mac(phi, phi)


def subspace_mac(phi_a, phi_b):
    """Cosines^2 of the principal angles between two mode subspaces
    (columns = modes); equals the MAC for one-dimensional subspaces and is
    the meaningful comparison inside clusters of (near-)equal frequencies."""
    qa, _ = np.linalg.qr(np.asarray(phi_a, dtype=np.complex128))
    qb, _ = np.linalg.qr(np.asarray(phi_b, dtype=np.complex128))
    s = np.linalg.svd(qa.conj().T @ qb, compute_uv=False)
    return np.clip(s, 0.0, 1.0) ** 2
