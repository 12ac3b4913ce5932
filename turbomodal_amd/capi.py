"""ctypes binding of the engine's C ABI (``include/tmhip.h``).

The library ``libtmhip.so`` is built in-tree next to this file.  Importing
this module never falls back to anything else: if the library is missing or
lacks a symbol, :func:`lib` raises.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libtmhip.so")

TM_OK = 0
TM_ERR_ARG = -1
TM_ERR_RUNTIME = -2
TM_ERR_HIP = -3
TM_ERR_NUMERIC = -4
TM_ERR_MEMORY = -5

TM_NODE_INTERIOR, TM_NODE_LEFT, TM_NODE_RIGHT = 0, 1, 2
TM_PREC_FP64, TM_PREC_MIXED = 0, 1
TM_CONV_TIGHT, TM_CONV_REFERENCE = 0, 1
TM_METHOD_LANCZOS, TM_METHOD_BLOCK = 0, 1

# Every function declared in include/tmhip.h.
EXPORTS = (
    "tm_context_create", "tm_context_destroy", "tm_last_error", "tm_set_profiling",
    "tm_kernel_stats", "tm_reset_kernel_stats", "tm_mesh_create", "tm_mesh_destroy",
    "tm_mesh_get_info", "tm_operator_create", "tm_operator_destroy", "tm_solve_batch",
    "tm_problem_footprint", "tm_device_free_bytes", "tm_operator_export",
    "tm_operator_static_u", "tm_server_create", "tm_server_submit", "tm_server_wait",
    "tm_server_destroy", "tm_set_memory_reserve", "tm_operator_add_springs", "tm_operator_inertia",
    "tm_operator_inertia_noalloc",
)


class KernelStat(C.Structure):
    _fields_ = [("name", C.c_char * 48), ("launches", C.c_int64), ("total_ms", C.c_double),
                ("alg_bytes", C.c_double), ("alg_flops", C.c_double)]


class MeshInfo(C.Structure):
    _fields_ = [("n_nodes", C.c_int32), ("n_elems", C.c_int32), ("n_reduced_nodes", C.c_int32),
                ("n_free_nodes", C.c_int32), ("n_free", C.c_int32), ("n_dof_system", C.c_int32),
                ("nnz_blocks", C.c_int64), ("nnz_L", C.c_int64), ("factor_flops", C.c_double),
                ("n_supernodes", C.c_int32), ("tree_levels", C.c_int32), ("max_front", C.c_int32),
                ("front_bytes_fp64", C.c_int64)]


class Problem(C.Structure):
    _fields_ = [("op", C.c_void_p), ("rpm", C.c_double), ("harmonic", C.c_int32),
                ("num_modes", C.c_int32), ("precision", C.c_int32), ("conv_policy", C.c_int32),
                ("shift", C.c_double), ("tol_rel", C.c_double), ("max_restarts", C.c_int32),
                ("method", C.c_int32), ("min_frequency", C.c_double), ("coriolis", C.c_int32),
                ("warm_start", C.c_int32)]


class Result(C.Structure):
    _fields_ = [("frequencies", C.POINTER(C.c_double)), ("mode_shapes", C.POINTER(C.c_double)),
                ("whirl", C.POINTER(C.c_int32)), ("eigenvalues", C.POINTER(C.c_double)),
                ("residuals", C.POINTER(C.c_double)), ("modal_mass", C.POINTER(C.c_double)),
                ("n_modes", C.c_int32),
                ("n_unique", C.c_int32), ("converged", C.c_int32), ("restarts", C.c_int32),
                ("status", C.c_int32), ("seconds", C.c_double), ("factor_retries", C.c_int32),
                ("pivot_flags", C.c_int32), ("shift_used", C.c_double)]


_lib = None


def lib():
    """Load ``libtmhip.so`` (raises if it is not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"HIP engine library not built: {LIB_PATH} (run __graft_entry__.build())")
    L = C.CDLL(LIB_PATH)
    for name in EXPORTS:
        if not hasattr(L, name):
            raise RuntimeError(f"{LIB_PATH} does not export {name}")
    vp, i32, i64, dbl = C.c_void_p, C.c_int32, C.c_int64, C.c_double
    P = C.POINTER
    L.tm_context_create.argtypes = [C.c_int, P(vp)]
    L.tm_context_destroy.argtypes = [vp]
    L.tm_last_error.argtypes = [vp]
    L.tm_last_error.restype = C.c_char_p
    L.tm_set_profiling.argtypes = [vp, C.c_int]
    L.tm_kernel_stats.argtypes = [vp, P(KernelStat), C.c_int]
    L.tm_reset_kernel_stats.argtypes = [vp]
    L.tm_mesh_create.argtypes = [vp, P(dbl), i32, P(i32), i32, P(i32), P(i32), P(C.c_uint8), i32, i32, P(vp)]
    L.tm_mesh_destroy.argtypes = [vp]
    L.tm_mesh_get_info.argtypes = [vp, P(MeshInfo)]
    L.tm_operator_create.argtypes = [vp, vp, dbl, dbl, dbl, i32, P(vp)]
    L.tm_operator_destroy.argtypes = [vp]
    L.tm_solve_batch.argtypes = [vp, P(Problem), i32, P(Result)]
    L.tm_problem_footprint.argtypes = [vp, i32, i32, i32, i32]
    L.tm_problem_footprint.restype = i64
    L.tm_device_free_bytes.argtypes = [vp]
    L.tm_set_memory_reserve.argtypes = [vp, dbl]
    L.tm_device_free_bytes.restype = i64
    L.tm_operator_export.argtypes = [vp, dbl, i32, i32, P(i64), P(i32), P(i32), P(dbl)]
    L.tm_operator_static_u.argtypes = [vp, P(dbl)]
    L.tm_server_create.argtypes = [vp, i32, dbl, P(vp)]
    L.tm_server_submit.argtypes = [vp, P(Problem), P(Result), P(C.c_uint64)]
    L.tm_server_wait.argtypes = [vp, C.c_uint64, dbl, P(i32)]
    L.tm_server_destroy.argtypes = [vp]
    L.tm_server_destroy.restype = None
    L.tm_operator_add_springs.argtypes = [vp, P(dbl)]
    L.tm_operator_inertia.argtypes = [vp, dbl, i32, dbl, P(i64), P(dbl), P(i32)]
    L.tm_operator_inertia_noalloc.argtypes = [vp, dbl, i32, dbl, P(i64), P(dbl), P(i32)]
    _lib = L
    return L


class EngineError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"[{code}] {msg}")
        self.code = code


def _ptr(a, ctype):
    return a.ctypes.data_as(C.POINTER(ctype))


class Context:
    def __init__(self, device: int = -1):
        L = lib()
        h = C.c_void_p()
        rc = L.tm_context_create(device, C.byref(h))
        if rc != TM_OK:
            raise EngineError(rc, L.tm_last_error(None).decode())
        self.h = h
        self.L = L

    def check(self, rc):
        if rc != TM_OK:
            msg = self.L.tm_last_error(self.h).decode()
            if rc == TM_ERR_ARG:
                raise ValueError(msg)
            raise EngineError(rc, msg)

    def close(self):
        if self.h:
            self.L.tm_context_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_profiling(self, on: bool):
        self.check(self.L.tm_set_profiling(self.h, 1 if on else 0))

    def kernel_stats(self):
        n = self.L.tm_kernel_stats(self.h, None, 0)
        if n < 0:
            self.check(n)
        arr = (KernelStat * max(n, 1))()
        self.L.tm_kernel_stats(self.h, arr, n)
        return {arr[i].name.decode(): dict(launches=arr[i].launches, total_ms=arr[i].total_ms,
                                           alg_bytes=arr[i].alg_bytes, alg_flops=arr[i].alg_flops)
                for i in range(n)}

    def reset_kernel_stats(self):
        self.check(self.L.tm_reset_kernel_stats(self.h))

    def free_bytes(self) -> int:
        return int(self.L.tm_device_free_bytes(self.h))

    def set_memory_reserve(self, fraction: float):
        """Share of device memory kept out of solve-batch sizing (solve_at_rpm's
        memory_reserve_fraction, src/python_bindings.cpp:360)."""
        self.check(self.L.tm_set_memory_reserve(self.h, float(fraction)))


class EngineMesh:
    def __init__(self, ctx: Context, nodes, elements, node_class, right_partner, constrained,
                 num_sectors, rotation_axis=2):
        self.ctx = ctx
        self.nodes = np.ascontiguousarray(nodes, dtype=np.float64)
        el = np.ascontiguousarray(elements, dtype=np.int32)
        cls = np.ascontiguousarray(node_class, dtype=np.int32)
        prt = np.ascontiguousarray(right_partner, dtype=np.int32)
        cm = np.ascontiguousarray(constrained, dtype=np.uint8)
        h = C.c_void_p()
        ctx.check(ctx.L.tm_mesh_create(ctx.h, _ptr(self.nodes, C.c_double), self.nodes.shape[0],
                                       _ptr(el, C.c_int32), el.shape[0], _ptr(cls, C.c_int32),
                                       _ptr(prt, C.c_int32), _ptr(cm, C.c_uint8), int(num_sectors),
                                       int(rotation_axis), C.byref(h)))
        self.h = h
        self.num_sectors = int(num_sectors)
        self.n_nodes = self.nodes.shape[0]

    def info(self) -> dict:
        mi = MeshInfo()
        self.ctx.check(self.ctx.L.tm_mesh_get_info(self.h, C.byref(mi)))
        return {f: getattr(mi, f) for f, _ in MeshInfo._fields_}

    def close(self):
        if self.h:
            self.ctx.L.tm_mesh_destroy(self.h)
            self.h = None


class EngineOperator:
    def __init__(self, mesh: EngineMesh, E, nu, rho, prestress=True):
        self.mesh = mesh
        ctx = mesh.ctx
        h = C.c_void_p()
        ctx.check(ctx.L.tm_operator_create(ctx.h, mesh.h, float(E), float(nu), float(rho),
                                           1 if prestress else 0, C.byref(h)))
        self.h = h

    def close(self):
        if self.h:
            self.mesh.ctx.L.tm_operator_destroy(self.h)
            self.h = None

    def export(self, rpm: float, harmonic: int, which: int):
        """(rows, cols, complex values) of K_eff(k) / M(k) / G(k) over the reference free DOFs."""
        ctx = self.mesh.ctx
        nnz = C.c_int64()
        ctx.check(ctx.L.tm_operator_export(self.h, rpm, harmonic, which, C.byref(nnz), None, None, None))
        n = nnz.value
        r = np.empty(n, np.int32)
        c = np.empty(n, np.int32)
        v = np.empty(2 * n, np.float64)
        ctx.check(ctx.L.tm_operator_export(self.h, rpm, harmonic, which, C.byref(nnz),
                                           _ptr(r, C.c_int32), _ptr(c, C.c_int32), _ptr(v, C.c_double)))
        return r, c, v[0::2] + 1j * v[1::2]

    def static_u(self):
        u = np.zeros(3 * self.mesh.n_nodes)
        self.mesh.ctx.check(self.mesh.ctx.L.tm_operator_static_u(self.h, _ptr(u, C.c_double)))
        return u

    def add_springs(self, spring_full):
        """Grounded springs (3 n_nodes stiffnesses, N/m) on the diagonal of K (ELASTIC_SUPPORT)."""
        k = np.ascontiguousarray(spring_full, dtype=np.float64)
        if k.shape != (3 * self.mesh.n_nodes,):
            raise ValueError("spring_full must hold 3 * n_nodes stiffnesses")
        self.mesh.ctx.check(self.mesh.ctx.L.tm_operator_add_springs(self.h, _ptr(k, C.c_double)))

    def inertia(self, rpm: float, harmonic: int, shift: float, noalloc: bool = False):
        """(negative pivots of the LDL^H of K(k) - shift M(k), smallest |d| / |a_jj|, status bits);
        noalloc: through tm_operator_inertia_noalloc (the library's static device scratch)."""
        neg, ratio, st = C.c_int64(), C.c_double(), C.c_int32()
        fn = self.mesh.ctx.L.tm_operator_inertia_noalloc if noalloc else self.mesh.ctx.L.tm_operator_inertia
        self.mesh.ctx.check(fn(self.h, float(rpm), int(harmonic), float(shift), C.byref(neg), C.byref(ratio),
                               C.byref(st)))
        return int(neg.value), float(ratio.value), int(st.value)

    def footprint(self, num_modes, harmonic, precision=TM_PREC_FP64, method=None):
        method = TM_METHOD_BLOCK if method is None else method
        return int(self.mesh.ctx.L.tm_problem_footprint(self.h, num_modes, precision, harmonic, method))


@dataclass
class SolveOutput:
    harmonic: int
    rpm: float
    frequencies: np.ndarray
    mode_shapes: np.ndarray      # (3 n_nodes, n_modes) complex
    whirl: np.ndarray
    eigenvalues: np.ndarray
    residuals: np.ndarray
    converged: bool
    restarts: int
    status: int
    seconds: float
    modal_mass: np.ndarray = None
    factor_retries: int = 0
    pivot_flags: int = 0
    shift_used: float = 0.0


def _prepare(p, want_shapes, prob, res):
    """Fill a Problem and a Result (caller-owned buffers) from a dict."""
    op = p["op"]
    prob.op = op.h
    prob.rpm = float(p.get("rpm", 0.0))
    prob.harmonic = int(p["harmonic"])
    prob.num_modes = int(p["num_modes"])
    prob.precision = int(p.get("precision", TM_PREC_FP64))
    prob.conv_policy = int(p.get("conv_policy", TM_CONV_TIGHT))
    prob.shift = float(p.get("shift", 1.0))
    prob.tol_rel = float(p.get("tol_rel", 1e-12))
    prob.max_restarts = int(p.get("max_restarts", 1000))
    prob.method = int(p.get("method", TM_METHOD_BLOCK))
    prob.min_frequency = float(p.get("min_frequency", 0.0))
    prob.coriolis = 1 if p.get("coriolis", False) else 0
    prob.warm_start = 1 if p.get("warm_start", False) else 0
    nm = prob.num_modes
    nn = op.mesh.n_nodes
    f = np.zeros(nm)
    w = np.zeros(nm, np.int32)
    ev = np.zeros(nm)
    rs = np.zeros(nm)
    ms = np.zeros(2 * 3 * nn * nm) if want_shapes else None
    mm = np.zeros(nm)
    res.modal_mass = _ptr(mm, C.c_double)
    res.frequencies = _ptr(f, C.c_double)
    res.whirl = _ptr(w, C.c_int32)
    res.eigenvalues = _ptr(ev, C.c_double)
    res.residuals = _ptr(rs, C.c_double)
    res.mode_shapes = _ptr(ms, C.c_double) if want_shapes else None
    return (f, w, ev, rs, ms, mm)


def _output(p, r, keep):
    f, w, ev, rs, ms, mm = keep
    nm = r.n_modes
    nn = p["op"].mesh.n_nodes
    shapes = None
    if ms is not None:
        # the engine writes (re, im) pairs: the buffer is already complex128, viewed without a copy
        shapes = ms.view(np.complex128)[: 3 * nn * nm].reshape(3 * nn, nm)
    return SolveOutput(int(p["harmonic"]), float(p.get("rpm", 0.0)), f[:nm].copy(), shapes,
                       w[:nm].copy(), ev[: r.n_unique].copy(), rs[: r.n_unique].copy(),
                       bool(r.converged), int(r.restarts), int(r.status), float(r.seconds),
                       mm[:nm].copy(), int(r.factor_retries), int(r.pivot_flags), float(r.shift_used))


def solve_batch(ctx: Context, problems, want_shapes=True):
    """problems: list of dicts {op, rpm, harmonic, num_modes, [precision, conv_policy, shift,
    tol_rel, max_restarts, method, min_frequency]} -> list[SolveOutput]."""
    n = len(problems)
    parr = (Problem * n)()
    rarr = (Result * n)()
    keep = [_prepare(p, want_shapes, parr[i], rarr[i]) for i, p in enumerate(problems)]
    ctx.check(ctx.L.tm_solve_batch(ctx.h, parr, n, rarr))
    return [_output(p, rarr[i], keep[i]) for i, p in enumerate(problems)]


class Server:
    """Continuous-batching server on a context (tm_server_*): submit() returns a
    ticket at once, wait() blocks until that problem's result is in."""

    def __init__(self, ctx: Context, max_slots: int = 64, hbm_fraction: float = 0.85):
        h = C.c_void_p()
        ctx.check(ctx.L.tm_server_create(ctx.h, int(max_slots), float(hbm_fraction), C.byref(h)))
        self.ctx, self.h, self._live = ctx, h, {}

    def submit(self, p, want_shapes=True) -> int:
        prob, res = Problem(), Result()
        keep = _prepare(p, want_shapes, prob, res)
        t = C.c_uint64()
        self.ctx.check(self.ctx.L.tm_server_submit(self.h, C.byref(prob), C.byref(res), C.byref(t)))
        self._live[t.value] = (p, res, keep)
        return t.value

    def wait(self, ticket: int, timeout=None):
        done = C.c_int32()
        self.ctx.check(self.ctx.L.tm_server_wait(self.h, ticket, -1.0 if timeout is None else float(timeout),
                                                 C.byref(done)))
        if not done.value:
            return None
        p, res, keep = self._live.pop(ticket)
        return _output(p, res, keep)

    def close(self):
        if self.h:
            self.ctx.L.tm_server_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


def mesh_arrays_from_model(nodes, elements, left, right, pairs, constrained_dofs):
    """node_class / right_partner / constrained mask from matched cyclic boundaries."""
    nn = len(nodes)
    cls = np.zeros(nn, np.int32)
    cls[list(left)] = TM_NODE_LEFT
    cls[list(right)] = TM_NODE_RIGHT
    partner = -np.ones(nn, np.int32)
    for l, r in pairs:
        partner[r] = l
    cm = np.zeros(3 * nn, np.uint8)
    cm[list(constrained_dofs)] = 1
    return cls, partner, cm


def engine_from_core(ctx: Context, core_mesh, constrained_dofs):
    """EngineMesh of a ``_core.Mesh`` whose cyclic boundaries are matched (the same
    inputs the drop-in solver hands the engine)."""
    cls, partner, cm = mesh_arrays_from_model(core_mesh.nodes, core_mesh.elements, core_mesh.left_boundary,
                                              core_mesh.right_boundary, core_mesh.matched_pairs, constrained_dofs)
    return EngineMesh(ctx, core_mesh.nodes, core_mesh.elements, cls, partner, cm, core_mesh.num_sectors,
                      int(getattr(core_mesh, "rotation_axis", 2)))
