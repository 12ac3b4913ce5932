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
