"""The reference's package name (``import turbomodal as tm``) over the MI355X engine.

Existing code written against turbomodal's cyclic modal solve keeps its imports:
``turbomodal.solve``, ``turbomodal.rpm_sweep``, ``turbomodal.io.load_mesh``,
``turbomodal._core.CyclicSymmetrySolver`` ... resolve to ``turbomodal_amd``
(python/turbomodal/__init__.py:1-60 names them).  Only the cyclic modal solve path
and its result formats are provided; the reference's other subsystems (CAD import,
fluid coupling, forced response, mistuning, visualisation) are not part of this
package and their names are absent.
"""
import sys as _sys

import turbomodal_amd as _impl
from turbomodal_amd import _core, mac, solver  # noqa: F401
from turbomodal_amd import (  # noqa: F401
    BCType,
    BoundaryCondition,
    ConstraintGroup,
    CyclicSymmetrySolver,
    EnrichedModalResult,
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
    campbell_data,
    compute_auto_mac,
    compute_mac,
    compute_modal_mass,
    load_results,
    rpm_sweep,
    save_results,
    solve,
)
from turbomodal.io import load_mesh  # noqa: F401,E402

# submodules under the reference's names
_sys.modules[__name__ + "._core"] = _core
_sys.modules[__name__ + ".mac"] = mac
_sys.modules[__name__ + ".solver"] = solver
__version__ = getattr(_impl, "__version__", "mi355x")
