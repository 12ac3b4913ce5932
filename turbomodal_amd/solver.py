"""High-level drop-in API: ``solve``, ``rpm_sweep`` and the result helpers.

Signatures and result objects follow python/turbomodal/solver.py (solve:
237-375, rpm_sweep: 499-628, campbell_data: 631-744, save_results /
load_results: 751-909).  The difference is where the work runs: every
harmonic of every rpm point is one problem of a batch on the MI355X engine
(shared symbolic analysis, one assembly at unit omega), and the modal mass
``N * Re(phi^H M phi)`` comes back from the device with the modes instead of
a CPU re-assembly of the sector mass matrix (solver.py:91-114).
"""
from __future__ import annotations

import logging
import sys
import time
from dataclasses import dataclass, field
from typing import Sequence

import numpy as np

from turbomodal_amd._core import (
    BCType,
    ConstraintGroup,
    CyclicSymmetrySolver,
    FluidConfig,
    Material,
    Mesh,
    ModalResult,
    SolverConfig,
)

logger = logging.getLogger(__name__)

SILENT = 0
PROGRESS = 1
DETAILED = 2


@dataclass
class EnrichedModalResult:
    """ModalResult plus the full-annulus modal mass in kg.

    All ModalResult attributes (harmonic_index, frequencies, mode_shapes,
    whirl_direction, converged, rpm) are reachable on this object.
    """

    _result: ModalResult = field(repr=False)
    modal_mass_kg: np.ndarray

    def __getattr__(self, name):
        return getattr(self._result, name)


@dataclass
class BoundaryCondition:
    """Constraint group given by explicit node ids (python/turbomodal/solver.py:117-155).

    Plane-based node selection belongs to the reference's mesh-editing side,
    which this package does not carry; pass ``node_ids``.
    """

    name: str
    type: str
    plane_point: np.ndarray | None = None
    plane_normal: np.ndarray | None = None
    constrained_components: tuple = (True, True, True)
    tolerance: float = 1e-6
    node_ids: list | None = None
    selection_radius: float | None = None
    spring_stiffness: np.ndarray | None = None
    cylinder_axis: np.ndarray | None = None
    cylinder_origin: np.ndarray | None = None
    surface_normal: np.ndarray | None = None


_BC_TYPES = {
    "fixed": BCType.FIXED,
    "displacement": BCType.DISPLACEMENT,
    "frictionless": BCType.FRICTIONLESS,
    "elastic_support": BCType.ELASTIC_SUPPORT,
    "cylindrical": BCType.CYLINDRICAL,
}


def _build_constraint_groups(mesh: Mesh, bcs: list[BoundaryCondition]) -> list[ConstraintGroup]:
    """python/turbomodal/solver.py:158-210 (explicit node ids only)."""
    groups = []
    valid = ", ".join(f"'{k}'" for k in _BC_TYPES)
    for bc in bcs:
        t = _BC_TYPES.get(bc.type.lower())
        if t is None:
            raise ValueError(f"Unknown BC type '{bc.type}'. Must be one of {valid}.")
        if bc.node_ids is None:
            raise ValueError(f"BC '{bc.name}': node_ids are required (plane selection is not provided)")
        cg = ConstraintGroup()
        cg.name = bc.name
        cg.type = t
        cg.node_ids = [int(i) for i in bc.node_ids]
        cg.constrained_components = [bool(c) for c in bc.constrained_components]
        if t == BCType.FRICTIONLESS:
            if bc.surface_normal is None:
                raise ValueError(f"BC '{bc.name}': frictionless needs surface_normal")
            cg.surface_normal = [float(x) for x in bc.surface_normal]
        elif t == BCType.ELASTIC_SUPPORT:
            if bc.spring_stiffness is None:
                raise ValueError(f"BC '{bc.name}': elastic_support requires spring_stiffness (3,) array.")
            cg.spring_stiffness = [float(x) for x in bc.spring_stiffness]
        elif t == BCType.CYLINDRICAL:
            ax = bc.cylinder_axis
            if ax is None:
                ax = np.zeros(3)
                ax[mesh.rotation_axis] = 1.0
            cg.cylinder_axis = [float(x) for x in ax]
            cg.cylinder_origin = [float(x) for x in (bc.cylinder_origin if bc.cylinder_origin is not None
                                                      else np.zeros(3))]
        groups.append(cg)
    return groups


def _validate_modal_results(results):
    """python/turbomodal/solver.py:218-234."""
    for r in results:
        freqs = np.asarray(r.frequencies)
        if freqs.size == 0:
            continue
        if np.any(np.isnan(freqs)):
            logger.warning("ND=%d: NaN frequencies detected — solver may not have converged", r.harmonic_index)
        if np.any(freqs < 0):
            logger.warning("ND=%d: negative frequencies detected", r.harmonic_index)
        shapes = np.asarray(r.mode_shapes)
        if shapes.size > 0 and np.any(np.linalg.norm(shapes, axis=0) < 1e-15):
            logger.warning("ND=%d: zero-norm mode shapes detected", r.harmonic_index)
    return results


def compute_modal_mass(mesh: Mesh, material: Material, results, *, temperature: float | None = None):
    """Full-annulus modal mass N * Re(phi^H M phi) per mode, one (n_modes,) array per result
    (python/turbomodal/solver.py:56-114).  The engine computes it on the device with the
    modes, from the sector mass matrix of the material the results were solved with, so
    ``material`` must be that material (the temperature law changes only E,
    src/material.cpp:33-38, and leaves M as it is); ``mesh`` is not needed."""
    del mesh, material, temperature
    return [np.asarray(getattr(r, "_result", r).modal_mass, dtype=np.float64) for r in results]


def _enrich(results):
    return [EnrichedModalResult(_result=r, modal_mass_kg=np.asarray(r.modal_mass, dtype=np.float64))
            for r in results]


def _make_solver(mesh, material, fluid, hub_constraint, boundary_conditions):
    if fluid is None:
        fluid = FluidConfig()
    if boundary_conditions is not None:
        return CyclicSymmetrySolver(mesh, material, _build_constraint_groups(mesh, boundary_conditions), fluid)
    return CyclicSymmetrySolver(mesh, material, fluid, hub_constraint == "fixed")


def _progress_bar(done, total, prefix="", elapsed=0.0, width=30, suffix=""):
    frac = done / max(total, 1)
    n = int(round(width * frac))
    return f"\r{prefix}[{'#' * n}{'.' * (width - n)}] {done}/{total} {elapsed:.1f}s{suffix}"


def solve(
    mesh: Mesh,
    material: Material,
    rpm: float = 0.0,
    num_modes: int = 20,
    fluid: FluidConfig | None = None,
    config: SolverConfig | None = None,
    verbose: int = PROGRESS,
    harmonic_indices: list[int] | None = None,
    max_threads: int = 0,
    hub_constraint: str = "fixed",
    include_coriolis: bool = False,
    min_frequency: float = 0.0,
    temperature: float | None = None,
    condition=None,
    boundary_conditions: list[BoundaryCondition] | None = None,
    show_convergence: bool = False,
    memory_reserve_fraction: float = 0.2,
) -> list[EnrichedModalResult]:
    """Cyclic-symmetry modal solve at one rpm (python/turbomodal/solver.py:237-375).

    ``config`` is accepted and, as in the reference, not used (the cyclic
    path fixes sigma = 1, nev/ncv and the tolerance itself).
    ``show_convergence`` (a live matplotlib chart) is accepted and ignored.
    """
    if condition is not None:
        rpm = condition.rpm
        if temperature is None:
            temperature = condition.temperature
    if temperature is not None:
        material = material.at_temperature(temperature)
    hi = list(harmonic_indices or [])
    if verbose >= PROGRESS:
        max_k = mesh.num_sectors // 2
        nd_str = f"ND {hi}" if hi else f"ND 0..{max_k}"
        cs = " [Coriolis]" if include_coriolis else ""
        ts = f"  T={temperature:.1f}K" if temperature is not None else ""
        print(f"Solving at {rpm:.0f} RPM{cs}{ts}  ({mesh.num_nodes()} nodes, {mesh.num_elements()} elements, "
              f"{mesh.num_sectors} sectors, {nd_str}, {num_modes} modes/ND)")
    t0 = time.perf_counter()
    solver = _make_solver(mesh, material, fluid, hub_constraint, boundary_conditions)
    progress_cb = None
    if verbose >= PROGRESS:
        def progress_cb(prog):
            sys.stdout.write(_progress_bar(prog.completed, prog.total, prefix="  ", elapsed=prog.elapsed_s))
            sys.stdout.flush()
            if prog.completed == prog.total:
                sys.stdout.write("\n")
    results = solver.solve_at_rpm(rpm, num_modes, hi, max_threads, include_coriolis, min_frequency,
                                  progress_cb, memory_reserve_fraction=memory_reserve_fraction)
    elapsed = time.perf_counter() - t0
    if verbose >= PROGRESS:
        print(f"  Solved {len(results)} harmonic indices in {elapsed:.1f}s")
    results = _enrich(_validate_modal_results(results))
    if verbose >= DETAILED:
        for r in results:
            fs = ", ".join(f"{f:.1f}" for f in r.frequencies[:4])
            ms = ", ".join(f"{m:.4e}" for m in r.modal_mass_kg[:4])
            print(f"    ND={r.harmonic_index:2d}: [{fs}] Hz (rotating frame)")
            print(f"           modal mass [kg]: [{ms}]")
    return results


def rpm_sweep(
    mesh: Mesh,
    material: Material,
    rpm_values: np.ndarray | Sequence[float],
    num_modes: int = 20,
    fluid: FluidConfig | None = None,
    verbose: int = SILENT,
    harmonic_indices: list[int] | None = None,
    max_threads: int = 0,
    hub_constraint: str = "fixed",
    include_coriolis: bool = False,
    min_frequency: float = 0.0,
    temperature: float | None = None,
    boundary_conditions: list[BoundaryCondition] | None = None,
    memory_reserve_fraction: float = 0.2,
) -> list[list[EnrichedModalResult]]:
    """Modal solves over rpm points (python/turbomodal/solver.py:499-628).

    All (rpm, harmonic) problems are submitted as one batch: one assembly at
    unit omega serves every rpm (K_eff = K + omega^2 (Ks1 - Kw1)) and all
    problems share one symbolic factorization.
    """
    if temperature is not None:
        material = material.at_temperature(temperature)
    hi = list(harmonic_indices or [])
    rpm_arr = np.asarray(rpm_values, dtype=np.float64)
    if verbose >= PROGRESS:
        print(f"RPM sweep: {len(rpm_arr)} points [{rpm_arr[0]:.0f} .. {rpm_arr[-1]:.0f}] RPM")
    t0 = time.perf_counter()
    solver = _make_solver(mesh, material, fluid, hub_constraint, boundary_conditions)
    rows = solver.solve_rpm_sweep([float(r) for r in rpm_arr], num_modes, hi, max_threads, include_coriolis,
                                  min_frequency, memory_reserve_fraction=memory_reserve_fraction)
    out = [_enrich(_validate_modal_results(row)) for row in rows]
    if verbose >= PROGRESS:
        total = time.perf_counter() - t0
        print(f"  Sweep complete in {total:.1f}s ({total / max(len(rpm_arr), 1):.2f}s/point avg)")
    return out


def campbell_data(results, num_sectors: int = 0) -> dict:
    """Stationary-frame Campbell data from an rpm sweep (python/turbomodal/solver.py:631-744)."""
    if not results:
        return {"rpm": np.array([]), "frequencies": np.array([]), "harmonic_index": np.array([]),
                "whirl_direction": np.array([])}
    n_rpm = len(results)
    harmonics = sorted({r.harmonic_index for row in results for r in row})
    harmonic_index = np.array(harmonics)
    h_to_idx = {h: i for i, h in enumerate(harmonics)}
    if num_sectors <= 0:
        max_k = int(harmonic_index[-1]) if len(harmonic_index) else 0
        num_sectors = 2 * max_k
    max_k = num_sectors // 2
    n_rot = max((len(r.frequencies) for row in results for r in row), default=0)
    n_modes = 2 * n_rot
    rpm = np.zeros(n_rpm)
    freqs = np.full((n_rpm, len(harmonics), n_modes), np.nan)
    whirl = np.zeros((n_rpm, len(harmonics), n_modes), dtype=np.int32)
    for i in range(n_rpm):
        if results[i]:
            rpm[i] = results[i][0].rpm
        omega = rpm[i] * 2.0 * np.pi / 60.0
        for r in results[i]:
            h = h_to_idx.get(r.harmonic_index)
            if h is None:
                continue
            k = r.harmonic_index
            f_rot = np.asarray(r.frequencies)
            w_rot = np.asarray(r.whirl_direction)
            nm = len(f_rot)
            # whirl labels present (QEP or degenerate duplication): per-mode shift
            if np.any(w_rot != 0) and abs(omega) > 0 and k > 0:
                kw = k * abs(omega) / (2.0 * np.pi)
                stat_f = np.abs(f_rot + w_rot * kw)
                order = np.argsort(stat_f, kind="stable")
                n_out = min(n_modes, nm)
                freqs[i, h, :n_out] = stat_f[order][:n_out]
                whirl[i, h, :n_out] = w_rot[order][:n_out]
            elif abs(omega) > 0 and 0 < k < max_k:
                kw = k * abs(omega) / (2.0 * np.pi)
                all_f = np.empty(2 * nm)
                all_w = np.empty(2 * nm, dtype=np.int32)
                all_f[0::2] = f_rot + kw
                all_f[1::2] = np.abs(f_rot - kw)
                all_w[0::2] = 1
                all_w[1::2] = -1
                order = np.argsort(all_f, kind="stable")
                n_out = min(n_modes, 2 * nm)
                freqs[i, h, :n_out] = all_f[order][:n_out]
                whirl[i, h, :n_out] = all_w[order][:n_out]
            else:
                n_out = min(n_modes, nm)
                freqs[i, h, :n_out] = f_rot[:n_out]
    return {"rpm": rpm, "frequencies": freqs, "harmonic_index": harmonic_index, "whirl_direction": whirl}


def save_results(path: str, results, *, include_mode_shapes: bool = True) -> None:
    """npz layout of python/turbomodal/solver.py:751-866."""
    if results and not isinstance(results[0], list):
        sweep, is_sweep = [results], False
    else:
        sweep, is_sweep = results, True
    n_rpm = len(sweep)
    if n_rpm == 0:
        np.savez_compressed(path, is_sweep=is_sweep)
        return
    n_h = len(sweep[0])
    n_modes = max((len(r.frequencies) for row in sweep for r in row), default=0)
    harmonics = np.array([r.harmonic_index for r in sweep[0]], dtype=np.int32)
    h_to_idx = {int(h): i for i, h in enumerate(harmonics)}
    rpms = np.zeros(n_rpm)
    f_arr = np.full((n_rpm, n_h, n_modes), np.nan)
    w_arr = np.zeros((n_rpm, n_h, n_modes), dtype=np.int32)
    m_arr = np.full((n_rpm, n_h, n_modes), np.nan)
    c_arr = np.ones((n_rpm, n_h), dtype=bool)
    n_dof = 0
    if include_mode_shapes:
        for row in sweep:
            for r in row:
                s = np.asarray(r.mode_shapes)
                if s.size:
                    n_dof = s.shape[0]
                    break
            if n_dof:
                break
    shapes = np.zeros((n_rpm, n_h, n_modes, n_dof), np.complex128) if include_mode_shapes and n_dof else None
    for i, row in enumerate(sweep):
        if row:
            rpms[i] = row[0].rpm
        for r in row:
            h = h_to_idx.get(r.harmonic_index)
            if h is None:
                continue
            nm = len(r.frequencies)
            f_arr[i, h, :nm] = np.asarray(r.frequencies)[:nm]
            w = np.asarray(r.whirl_direction)
            w_arr[i, h, :min(nm, len(w))] = w[:nm]
            c_arr[i, h] = r.converged
            if hasattr(r, "modal_mass_kg"):
                mm = np.asarray(r.modal_mass_kg)
                m_arr[i, h, :min(nm, len(mm))] = mm[:nm]
            if shapes is not None:
                s = np.asarray(r.mode_shapes)
                if s.size:
                    ns = min(nm, s.shape[1])
                    shapes[i, h, :ns, :] = s[:, :ns].T
    arrays = {"is_sweep": np.bool_(is_sweep), "rpm": rpms, "harmonic_index": harmonics, "frequencies": f_arr,
              "whirl_direction": w_arr, "modal_mass_kg": m_arr, "converged": c_arr}
    if shapes is not None:
        arrays["mode_shapes"] = shapes
    np.savez_compressed(path, **arrays)


def load_results(path: str) -> dict:
    """python/turbomodal/solver.py:869-909 (no pickles)."""
    data = dict(np.load(path, allow_pickle=False))
    if "is_sweep" in data:
        data["is_sweep"] = bool(data["is_sweep"])
    return data
