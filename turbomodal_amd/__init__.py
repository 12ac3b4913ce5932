"""MI355X-native cyclic-symmetry modal solve with turbomodal's API.

``solve`` / ``rpm_sweep`` and the ``_core`` classes keep the reference's
signatures (python/turbomodal/solver.py:237-628, src/python_bindings.cpp);
every numeric stage runs in the gfx950 HIP engine ``libtmhip.so``.
"""
from turbomodal_amd._core import (  # noqa: F401
    BCType,
    ConstraintGroup,
    CyclicSymmetrySolver,
    FluidConfig,
    FluidType,
    Material,
    Mesh,
    ModalResult,
    NodeSet,
    SolverConfig,
    SolverProgress,
    SolverStatus,
    StationaryFrameResult,
)
from turbomodal_amd.mac import compute_auto_mac, compute_mac  # noqa: F401
from turbomodal_amd.solver import (  # noqa: F401
    BoundaryCondition,
    EnrichedModalResult,
    campbell_data,
    compute_modal_mass,
    load_results,
    rpm_sweep,
    save_results,
    solve,
)


def load_mesh(path: str, num_sectors: int, rotation_axis: int = 2) -> Mesh:
    """gmsh MSH 2.x sector mesh -> Mesh with cyclic boundaries matched."""
    m = Mesh()
    m.load_from_gmsh(path)
    m.num_sectors = int(num_sectors)
    m.rotation_axis = int(rotation_axis)
    m.identify_cyclic_boundaries()
    m.match_boundary_nodes()
    return m
