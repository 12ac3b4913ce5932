"""Host-side logic of the drop-in package (no GPU): the _core object model,
gmsh parsing and cyclic matching against the oracle, constrained DOFs,
stationary-frame known answers (tests/test_cyclic.cpp:451-676), Campbell
data and the npz result layout, the mesh generators' sizes, and the C ABI
library's exports."""
import os
import re

import numpy as np
import pytest

from common import GOLDEN, fixture_mesh, oracle_model

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_library_exports_every_declared_symbol():
    import ctypes
    from turbomodal_amd import capi
    hdr = open(os.path.join(ROOT, "include", "tmhip.h")).read()
    declared = set(re.findall(r"\b(tm_[a-z_]+)\s*\(", hdr))
    assert declared == set(capi.EXPORTS)
    lib = ctypes.CDLL(capi.LIB_PATH)
    for name in declared:
        assert hasattr(lib, name), name


def test_material_validation_and_temperature():
    import turbomodal_amd as tm
    with pytest.raises(ValueError):
        tm.Material(-1.0, 0.3, 7800)
    with pytest.raises(ValueError):
        tm.Material(200e9, 0.5, 7800)
    with pytest.raises(ValueError):
        tm.Material(200e9, 0.3, 0.0)
    m = tm.Material(200e9, 0.3, 7800, 293.15, -5e7)
    m2 = m.at_temperature(393.15)
    assert abs(m2.E - (200e9 - 5e9)) < 1e-3
    D = np.asarray(tm.Material(200e9, 0.3, 7850).constitutive_matrix())
    assert abs(D[0, 0] - 269.2308e9) < 1e5 and abs(D[3, 3] - 76.9231e9) < 1e5


@pytest.mark.parametrize("name", ["wedge_sector.msh", "blade_sector.msh", "leissa_disk_sector.msh"])
def test_gmsh_loader_and_matching_agree_with_oracle(name):
    import turbomodal_amd as tm
    mesh = tm.load_mesh(os.path.join(GOLDEN, name), 24)
    X, T, S, N = fixture_mesh(name)
    assert np.allclose(np.asarray(mesh.nodes), X)
    assert np.array_equal(np.asarray(mesh.elements), T)
    m = oracle_model(X, T, S, N)
    assert sorted(mesh.matched_pairs) == sorted(m.pairs)
    assert sorted(mesh.left_boundary) == sorted(m.left)
    sets = {ns.name: sorted(ns.node_ids) for ns in mesh.node_sets}
    for k, v in S.items():
        assert sets[k] == list(v)


def test_constrained_dofs():
    import turbomodal_amd as tm
    mesh = tm.load_mesh(os.path.join(GOLDEN, "wedge_sector.msh"), 24)
    mat = tm.Material(200e9, 0.3, 7800)
    s = tm.CyclicSymmetrySolver(mesh, mat)
    hub = mesh.find_node_set("hub_constraint").node_ids
    assert s.build_constrained_dofs() == sorted(3 * n + c for n in hub for c in range(3))
    cg = tm.ConstraintGroup()
    cg.node_ids = [1, 2]
    cg.type = tm.BCType.DISPLACEMENT
    cg.constrained_components = [False, True, False]
    s2 = tm.CyclicSymmetrySolver(mesh, mat, [cg])
    assert s2.build_constrained_dofs() == [4, 7]
    assert tm.CyclicSymmetrySolver(mesh, mat, tm.FluidConfig(), False).build_constrained_dofs() == []


def _result(k, rpm, freqs, whirl):
    import turbomodal_amd as tm
    r = tm.ModalResult()
    r.harmonic_index = k
    r.rpm = rpm
    r.frequencies = list(freqs)
    r.whirl_direction = list(whirl)
    return r


def test_stationary_frame_known_answers():
    import turbomodal_amd as tm
    S = tm.CyclicSymmetrySolver
    # kinematic split (tests/test_cyclic.cpp:451-477): 3000 rpm, k=2 -> 100 Hz shift
    sf = S.compute_stationary_frame(_result(2, 3000.0, [300.0, 1000.0], [0, 0]), 24)
    assert np.allclose(sf.frequencies, [200.0, 400.0, 900.0, 1100.0], atol=0.1)
    sf = S.compute_stationary_frame(_result(0, 3000.0, [100.0, 200.0, 300.0], [0, 0, 0]), 24)
    assert np.allclose(sf.frequencies, [100, 200, 300]) and np.all(np.asarray(sf.whirl_direction) == 0)
    sf = S.compute_stationary_frame(_result(12, 3000.0, [500.0, 1000.0], [0, 0]), 24)
    assert np.allclose(sf.frequencies, [500, 1000])
    sf = S.compute_stationary_frame(_result(3, 0.0, [500.0, 1000.0], [0, 0]), 24)
    assert np.allclose(sf.frequencies, [500, 1000])
    # labelled modes (tests/test_cyclic.cpp:644-676): per-mode shift by k Omega
    sf = S.compute_stationary_frame(_result(1, 3000.0, [200.0, 250.0, 650.0, 700.0], [-1, 1, -1, 1]), 24)
    assert np.allclose(sf.frequencies, [150.0, 300.0, 600.0, 750.0], atol=1e-6)


def test_campbell_and_npz_roundtrip(tmp_path):
    import turbomodal_amd as tm
    from turbomodal_amd.solver import EnrichedModalResult
    rows = []
    for rpm in (0.0, 3000.0):
        row = []
        for k in (0, 1, 12):
            w = [0, 0] if k in (0, 12) else [1, -1]
            r = _result(k, rpm, [100.0 + k, 100.0 + k], w)
            r.mode_shapes = np.ones((6, 2), complex) * (k + 1)
            row.append(EnrichedModalResult(_result=r, modal_mass_kg=np.array([24.0, 24.0])))
        rows.append(row)
    cd = tm.campbell_data(rows, 24)
    assert cd["frequencies"].shape == (2, 3, 4)
    # k=1 at 3000 rpm: FW 101 + 50, BW |101 - 50|
    assert np.allclose(np.sort(cd["frequencies"][1, 1, :2]), [51.0, 151.0])
    p = str(tmp_path / "sweep.npz")
    tm.save_results(p, rows)
    d = tm.load_results(p)
    assert d["is_sweep"] is True
    assert d["frequencies"].shape == (2, 3, 2) and d["mode_shapes"].shape == (2, 3, 2, 6)
    assert np.allclose(d["modal_mass_kg"], 24.0) and np.allclose(d["mode_shapes"][0, 2], 13.0)


def test_generator_sizes():
    from turbomodal_amd import meshgen
    for name, nodes, elems in (("flat_disk_c1", 1625, 864), ("bladed_c2", 15105, 9216)):
        sm = meshgen.config_mesh(name)
        assert sm.num_nodes == nodes and sm.num_elements == elems
        p = sm.nodes[sm.elements[:, :4]]
        vol = np.einsum("ti,ti->t", np.cross(p[:, 1] - p[:, 0], p[:, 2] - p[:, 0]), p[:, 3] - p[:, 0])
        assert vol.min() > 0
        m = oracle_model(sm.nodes, sm.elements, sm.node_sets, sm.num_sectors)
        assert len(m.pairs) == len(sm.node_sets["left_boundary"])


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_grid_matching_equals_greedy_scan(seed):
    """The grid search of match_boundary_nodes returns the reference's greedy
    nearest-unused pairing (src/mesh.cpp:289-333), restated as a full scan in the
    oracle: jittered partners, partners beyond the tolerance, right nodes with no
    partner, two right nodes at the same distance from a left node, and a left
    node whose nearest right node is taken by an earlier left node."""
    import turbomodal_amd as tm
    from oracle import cyclic
    rng = np.random.default_rng(seed)
    N = 24
    a = 2 * np.pi / N
    nl = 300
    r = rng.uniform(0.1, 0.3, nl)
    z = rng.uniform(0.0, 0.05, nl)
    left = np.stack([r, np.zeros(nl), z], 1)
    rot = np.array([[np.cos(a), -np.sin(a), 0], [np.sin(a), np.cos(a), 0], [0, 0, 1]])
    right = left @ rot.T
    right += rng.normal(0, 1e-6, right.shape) * (rng.uniform(0, 1, (nl, 1)) < 0.5)  # within tolerance
    right[:10] += 1e-2                                                            # beyond tolerance
    extra = rng.uniform(0.1, 0.3, (20, 1)) * np.array([[np.cos(a), np.sin(a), 0]]) + [0, 0, 0.07]
    twin = right[50] + np.array([0.0, 0.0, 2e-6])                                 # tie candidate
    right = np.vstack([right, extra, twin, right[60] + 1e-7])                     # a near-duplicate
    X = np.vstack([left, right, np.zeros((10, 3))])
    nL, nR = len(left), len(right)
    lid = list(range(nL))
    rid = list(range(nL, nL + nR))
    perm = rng.permutation(nR)
    rid = [rid[p] for p in perm]
    T = np.tile(np.arange(nL + nR, nL + nR + 10, dtype=np.int32), (1, 1))
    sets = []
    for name, ids in (("left_boundary", lid), ("right_boundary", rid)):
        ns = tm.NodeSet()
        ns.name = name
        ns.node_ids = ids
        sets.append(ns)
    mesh = tm.Mesh()
    mesh.load_from_arrays(X, T, sets, N, 2)
    l2, r2, pairs, free = cyclic.match_boundary_nodes(X, lid, rid, N)
    assert [tuple(p) for p in mesh.matched_pairs] == [tuple(p) for p in pairs]
    assert list(mesh.left_boundary) == l2 and list(mesh.right_boundary) == r2
    assert sorted(mesh.free_boundary) == free


def test_reference_package_name_resolves_to_the_engine():
    # `import turbomodal as tm` code keeps working (python/turbomodal/__init__.py:1-60)
    import turbomodal as tm
    import turbomodal_amd
    from turbomodal import _core
    from turbomodal.io import load_mesh
    from turbomodal.solver import rpm_sweep
    assert tm.solve is turbomodal_amd.solve and rpm_sweep is turbomodal_amd.rpm_sweep
    assert _core.CyclicSymmetrySolver is turbomodal_amd.CyclicSymmetrySolver
    m = load_mesh(os.path.join(GOLDEN, "wedge_sector.msh"), 24)
    assert m.num_sectors == 24 and len(m.matched_pairs) > 0
    with pytest.raises(ValueError):
        load_mesh(os.path.join(GOLDEN, "wedge_sector.msh"), 24, file_format="nastran")
