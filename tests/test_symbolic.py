"""Symbolic analysis on the host (turbomodal_amd/csrc/symbolic.cpp): the nested
dissection, the supernodal tree and the front structures, checked by
tests/native/symcheck.cpp against quantities it re-derives from the K(k) block
pattern itself.  The checker shares the analysis code, not its bookkeeping:
* the supernodes' pivot ranges tile the free nodes in postorder;
* every separator separates (no pattern edge between the two children's subtrees);
* every front's structure equals symbolic elimination's (the later nodes adjacent
  to its pivots or in its children's structures), so the multifrontal factor has
  exactly the fill of the chosen ordering;
* nnz(L) and the factor flops recomputed from those structures agree.
The fill and flop bounds on the generated C2 bladed sector guard the ordering's
quality (minimum degree gives 20.2 M on the same scalar pattern)."""
import json
import os
import shutil
import subprocess

import numpy as np
import pytest

from common import GOLDEN, fixture_mesh, generated, oracle_model

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")


@pytest.fixture(scope="module")
def symcheck(tmp_path_factory):
    if not (os.path.exists(HIPCC) or shutil.which(HIPCC)):
        pytest.skip("hipcc not available to build the host checker")
    out = str(tmp_path_factory.mktemp("symcheck") / "symcheck")
    csrc = os.path.join(ROOT, "turbomodal_amd", "csrc")
    subprocess.run([HIPCC, "-O2", "-std=c++17", "-x", "hip", "--offload-arch=gfx950", "-I", csrc,
                    "-I", os.path.join(ROOT, "include"), os.path.join(ROOT, "tests", "native", "symcheck.cpp"),
                    os.path.join(csrc, "symbolic.cpp"), "-o", out], check=True)
    return out


def run_check(exe, tmp_path, nodes, elements, node_sets, num_sectors):
    from turbomodal_amd import capi
    model = oracle_model(nodes, elements, node_sets, num_sectors)
    cls, partner, cm = capi.mesh_arrays_from_model(model.nodes, model.elements, model.left, model.right,
                                                    model.pairs, model.constrained)
    path = str(tmp_path / "mesh.bin")
    with open(path, "wb") as f:
        np.array([len(model.nodes), len(model.elements), num_sectors, 2], np.int32).tofile(f)
        np.asarray(model.nodes, np.float64).tofile(f)
        np.asarray(model.elements, np.int32).tofile(f)
        cls.astype(np.int32).tofile(f)
        partner.astype(np.int32).tofile(f)
        cm.astype(np.uint8).tofile(f)
    res = subprocess.run([exe, path], check=True, capture_output=True, text=True)
    return json.loads(res.stdout)


def assert_consistent(r):
    assert r["tile_ok"] and r["tree_ok"] and r["root_ok"]
    assert r["shared_ok"]  # the shared fronts re-derived from the node ranges
    assert r["n_shared"] < r["nsn"]  # the phase-coupled slots lie below the root
    assert r["sep_ok"]  # (a fixture smaller than one leaf has no separator to check)
    assert r["struct_ok"]
    assert r["nnz_L"] == r["nnz_L_recomputed"]
    assert r["flops"] == pytest.approx(r["flops_recomputed"], rel=1e-12)
    assert r["n_free"] <= 3 * r["nf"]  # partially constrained nodes keep pinned rows


@pytest.mark.parametrize("name", ["wedge_sector.msh", "blade_sector.msh", "leissa_disk_sector.msh"])
def test_reference_fixtures(symcheck, tmp_path, name):
    r = run_check(symcheck, tmp_path, *fixture_mesh(name))
    assert_consistent(r)


def test_flat_disk_c1(symcheck, tmp_path):
    r = run_check(symcheck, tmp_path, *generated("flat_disk_c1"))
    assert_consistent(r)
    assert r["nf"] == 1440 and r["nsn"] > 1 and r["sep_edges_checked"] > 0


def test_bladed_sector_c2_fill_and_flops(symcheck, tmp_path):
    r = run_check(symcheck, tmp_path, *generated("bladed_c2"))
    assert_consistent(r)
    assert r["n_free"] == 43344 and r["sep_edges_checked"] > 0
    # the ordering's quality: fill within 10 % of minimum degree's, flops as measured
    assert r["nnz_L"] <= 22_200_000
    assert r["flops"] <= 7.0e10
    # most fronts lie in subtrees away from the phase seam: a third of the flops is shared
    # by every harmonic at one rpm (measured 517 of 605 fronts, 2.29e10 flops)
    assert r["n_shared"] > r["nsn"] // 2
    assert 0.25 * r["flops"] < r["flops_shared"] < 0.5 * r["flops"]
