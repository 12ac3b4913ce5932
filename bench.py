#!/usr/bin/env python3
"""Benchmark: ND modal solves/sec of the cyclic-symmetry modal solve on one MI355X.

Workload (default ``c4``): rpm sweep on the ~45k-DOF bladed sector (disk
(24,12,4) + twisted blade (24,8,2) Kuhn TET10 lattices, 36 sectors, steel),
20 rpm points linspace(0, 15000) x harmonic indices 0..18 = 380 complex
Hermitian eigenproblems, 10 modes each, centrifugal prestress + spin
softening.  One step = one full sweep through the drop-in
``CyclicSymmetrySolver.solve_rpm_sweep`` (results incl. full-sector mode
shapes delivered to host memory).  The mesh, its symbolic analysis and the
assembled operator are resident on the device before the timed region, as
the engine keeps them across calls.

Prints one JSON line (see the keys below).  ``--gpus N`` with N > 1 exits:
this build runs on one GPU.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "ND modal solves/sec + p99 latency, TET10 bladed sector, freq err ≤1e-8, MAC≥.999"
HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6
FP32_PEAK_TFLOPS = 157.3


def build_solver(workload, precision):
    import turbomodal_amd as tm
    from turbomodal_amd import meshgen
    name = {"c4": "bladed_c2", "c2": "bladed_c2", "c1": "flat_disk_c1", "c3": "bladed_c3"}[workload]
    sm = meshgen.config_mesh(name)
    mesh = meshgen.to_core_mesh(sm, tm._core)
    mat = tm.Material(**(meshgen.TITANIUM if workload == "c3" else meshgen.STEEL))
    solver = tm.CyclicSymmetrySolver(mesh, mat)
    return tm, sm, mesh, mat, solver


def workload_spec(workload):
    if workload == "c4":
        return np.linspace(0.0, 15000.0, 20), 10
    if workload == "c2":
        return np.array([10000.0]), 10
    if workload == "c3":
        return np.array([12000.0]), 20
    return np.array([0.0]), 5


def source_hash():
    """SHA-1 of the engine's sources (turbomodal_amd/csrc, sorted by name): a traffic
    figure measured on other sources is not reported."""
    import hashlib
    h = hashlib.sha1()
    d = os.path.join(ROOT, "turbomodal_amd", "csrc")
    for name in sorted(os.listdir(d)):
        with open(os.path.join(d, name), "rb") as f:
            h.update(name.encode() + b"\0" + f.read())
    return h.hexdigest()


def measured_traffic(args, key):
    """HBM bytes per launch of the factorization ('factor') or the triangular-sweep pair
    ('trisolve') of this workload, as the FETCH_SIZE and WRITE_SIZE counters read them
    (rocprofv3 --pmc, scaled by a plain copy kernel), recorded with their commands in
    profiles/ by separate profiled runs (the counters cannot be read inside the timed
    run).  None unless the profile was taken on these very sources."""
    if args.workload != "c4" or args.method != "block" or args.coriolis:
        return None
    path = os.path.join(ROOT, "profiles", "traffic_c4_rpm_sweep_fp64_block.md")
    vals = {}
    try:
        for line in open(path):
            if ":" in line and not line.startswith(("|", " ", "#")):
                k, v = line.split(":", 1)
                vals[k.strip()] = v.strip()
    except OSError:
        return None
    if vals.get("source_hash") != source_hash():
        return None
    v = vals.get(f"{key}_traffic_bytes_per_launch")
    return float(v) if v else None


def _oracle_solve(args):
    """One harmonic of the oracle (worker of the all-core baseline; the operators come
    from the parent through fork)."""
    from threadpoolctl import threadpool_limits
    from oracle import cyclic
    k, num_modes, rpm = args
    with threadpool_limits(limits=1):  # one BLAS thread per process (the pool exists since import)
        t = time.perf_counter()
        cyclic.solve_harmonic(_BASE["m"], _BASE["ops"], k, num_modes, rpm, "reference")
        return time.perf_counter() - t


_BASE = {}


def cpu_baseline(sm, mat, rpm, num_modes):
    """The oracle (numpy/scipy restatement of the reference algorithm with the reference's
    own nev/ncv/tolerance policy, one thread per process) on one rpm point of the workload:
    the operator build, then all N/2 + 1 harmonics over the host cores (one process per
    harmonic, as the reference runs one thread per harmonic, src/cyclic_solver.cpp:1229-1231),
    and two harmonics on one core."""
    import multiprocessing as mp
    os.environ["OMP_NUM_THREADS"] = "1"
    os.environ["OPENBLAS_NUM_THREADS"] = "1"
    from oracle import cyclic
    t0 = time.perf_counter()
    m = cyclic.build_model(sm.nodes, sm.elements, sm.node_sets, sm.num_sectors)
    ops = cyclic.sector_operators(m, mat["E"], mat["nu"], mat["rho"], rpm)
    t_ops = time.perf_counter() - t0
    _BASE.update(m=m, ops=ops)
    n_h = sm.num_sectors // 2 + 1
    workers = max(1, min(16, os.cpu_count() or 1, n_h))   # the box's CPU share for one GPU
    t1 = time.perf_counter()
    with mp.get_context("fork").Pool(workers) as pool:
        per = pool.map(_oracle_solve, [(k, num_modes, rpm) for k in range(n_h)], chunksize=1)
    t_all = time.perf_counter() - t1
    t2 = time.perf_counter()
    for k in (0, 1):
        _oracle_solve((k, num_modes, rpm))
    t_one = (time.perf_counter() - t2) / 2
    return {"t_ops": t_ops, "t_all": t_all, "workers": workers, "t_one": t_one, "n_h": n_h,
            "per_harmonic_s": per}


def host_description():
    """Host CPU of this run: logical CPUs and the model name (lscpu)."""
    import subprocess
    model = None
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for line in out.splitlines():
            if line.startswith("Model name:"):
                model = line.split(":", 1)[1].strip()
    except Exception:
        pass
    return {"nproc": os.cpu_count(), "model": model}


def run_c5(args):
    """Serving: seeded Poisson stream of uneven requests through the continuous-batching
    server; meshes, symbolic analyses and operators resident before the timed region."""
    import threading
    import turbomodal_amd as tm
    from turbomodal_amd import capi, meshgen
    core = tm._core
    rng = np.random.default_rng(args.seed)
    names = ["flat_disk_c1", "bladed_c2"] + (["bladed_c3"] if args.c5_c3 else [])
    mats = {"steel": (200e9, 0.3, 7800.0), "titanium": (114e9, 0.34, 4430.0)}
    ctx = capi.Context()
    ops, keep = [], []
    t0 = time.perf_counter()
    for name in names:
        for N in (24, 36, 72):
            sm = meshgen.config_mesh(name, N)
            cm = meshgen.to_core_mesh(sm, core)
            solver = core.CyclicSymmetrySolver(cm, core.Material(*mats["steel"]))
            em = capi.engine_from_core(ctx, cm, solver.build_constrained_dofs())
            keep += [cm, solver, em]
            for mname, mat in mats.items():
                ops.append((name, N, mname, capi.EngineOperator(em, *mat, True)))
    print(f"[bench] c5: {len(ops)} resident operators in {time.perf_counter() - t0:.1f}s", file=sys.stderr, flush=True)

    def draw(n):
        out = []
        for _ in range(n):
            name, N, mname, op = ops[int(rng.integers(len(ops)))]
            out.append(dict(op=op, rpm=float(rng.uniform(0, 15000)), harmonic=int(rng.integers(0, N // 2 + 1)),
                            num_modes=int(rng.choice([5, 10, 20])), mesh=name))
        return out

    # saturation: the whole stream at once
    reqs = draw(args.c5_requests)
    with capi.Server(ctx, max_slots=args.c5_slots) as srv:
        ts = time.perf_counter()
        tickets = [srv.submit(p, want_shapes=True) for p in reqs]
        for t in tickets:
            srv.wait(t)
        sat = len(reqs) / (time.perf_counter() - ts)
    print(f"[bench] c5 saturation {sat:.2f} solves/s", file=sys.stderr, flush=True)
    # Poisson arrivals at 80 % of saturation
    rate = 0.8 * sat
    reqs = draw(args.c5_requests)
    gaps = rng.exponential(1.0 / rate, len(reqs))
    arrival, done_at = {}, {}
    lock = threading.Lock()
    with capi.Server(ctx, max_slots=args.c5_slots) as srv:
        from concurrent.futures import ThreadPoolExecutor
        pool = ThreadPoolExecutor(max_workers=16)

        def waiter(t):
            out = srv.wait(t)
            with lock:
                done_at[t] = time.perf_counter()
            return out

        futs = []
        tstart = time.perf_counter()
        nxt = tstart
        for p, g in zip(reqs, gaps):
            nxt += g
            dt = nxt - time.perf_counter()
            if dt > 0:
                time.sleep(dt)
            t = srv.submit(p)
            arrival[t] = time.perf_counter()
            futs.append(pool.submit(waiter, t))
        results = [f.result() for f in futs]
        tend = time.perf_counter()
        pool.shutdown()
    lat = np.array([done_at[t] - arrival[t] for t in arrival])  # in arrival order
    last_arrival = max(arrival.values())
    served = len(reqs) / (tend - tstart)
    ok = sum(1 for r in results if r.status == 0)
    # the plain FP64 path: one request at a time, single-vector Lanczos, on a sample
    sample = reqs[: args.c5_plain]
    tp = time.perf_counter()
    for p in sample:
        capi.solve_batch(ctx, [dict(p, method=capi.TM_METHOD_LANCZOS)])
    plain = len(sample) / (time.perf_counter() - tp)
    out = {
        "metric": "ND modal solves/sec + p99 latency (serving)", "value": round(served, 3), "unit": "ND solves/s",
        "n_gpus": 1, "steps": 1, "warmup": 0, "ms_per_step": round(1e3 * (tend - tstart), 1),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "c128 (complex f64)",
        "data": "synthetic (seeded Kuhn TET10 lattice meshes; Poisson arrivals)",
        "config": {"workload": "c5 serving stream", "requests": len(reqs), "meshes": names, "sectors": [24, 36, 72],
                   "materials": list(mats), "arrival_rate_per_s": round(rate, 3), "seed": args.seed},
        "saturation_solves_per_s": round(sat, 3),
        "p50_latency_ms": round(1e3 * float(np.percentile(lat, 50)), 1),
        "p99_latency_ms": round(1e3 * float(np.percentile(lat, 99)), 1),
        # a stream the server keeps up with has latencies that do not grow along it
        "p50_latency_ms_by_arrival_quarter": [round(1e3 * float(np.percentile(q, 50)), 1)
                                              for q in np.array_split(lat, 4)],
        "served_rate_during_arrivals": round(sum(1 for t in arrival if done_at[t] <= last_arrival)
                                             / (last_arrival - tstart), 3),
        "solved": ok,
        "plain_fp64_one_at_a_time": {"value": round(plain, 3), "unit": "ND solves/s", "sample": len(sample),
                                     "eigensolver": "lanczos"},
        "speedup_vs_plain": round(sat / plain, 2),
    }
    for o in ops:
        o[3].close()
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="c4", choices=["c4", "c2", "c1", "c3", "c5", "hbmcal"],
                    help="hbmcal: one plain 256 MiB streaming copy kernel, the reference point that scales "
                         "the FETCH_SIZE / WRITE_SIZE counters to bytes")
    ap.add_argument("--coriolis", action="store_true",
                    help="include_coriolis=True: the gyroscopic QEP for harmonics k > 0")
    ap.add_argument("--c5-requests", type=int, default=2000)
    ap.add_argument("--c5-slots", type=int, default=64, help="slots per server group (tm_server_create max_slots)")
    ap.add_argument("--c5-plain", type=int, default=24)
    ap.add_argument("--c5-c3", action="store_true", help="include the ~330k-DOF C3 generator in the stream")
    ap.add_argument("--seed", type=int, default=2024)
    ap.add_argument("--precision", default="fp64", choices=["fp64", "mixed"],
                    help="mixed: complex-FP32 factor with FP64 refinement of every shift-invert solve")
    ap.add_argument("--method", default="block", choices=["block", "lanczos"])
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-plain", action="store_true", help="skip the plain FP64 one-at-a-time measurement")
    ap.add_argument("--serial-groups", action="store_true",
                    help="run the block-size slot groups of a sweep one after another instead of side by side "
                         "(every sweep; the kernels then run alone on the GPU, as in the roofline sweep)")
    args = ap.parse_args()
    if args.serial_groups:
        os.environ["TMHIP_SERIAL_GROUPS"] = "1"
    if args.gpus != 1:
        print(f"bench.py: this build measures one MI355X; --gpus {args.gpus} is not supported "
              "(the (rpm, harmonic) problems would shard across GPUs with no collectives)",
              file=sys.stderr)
        sys.exit(2)
    if args.workload == "c5":
        run_c5(args)
        return
    if args.workload == "hbmcal":
        import torch
        x = torch.ones(32 << 20, dtype=torch.float64, device="cuda")  # 256 MiB
        y = torch.empty_like(x)
        torch.mul(x, 1.0, out=y)  # one streaming kernel: 256 MiB read, 256 MiB written
        torch.cuda.synchronize()
        print(json.dumps({"workload": "hbmcal", "read_bytes": x.numel() * 8, "write_bytes": y.numel() * 8}))
        return

    tm, sm, mesh, mat, solver = build_solver(args.workload, args.precision)
    from turbomodal_amd import meshgen
    solver.engine_method = args.method
    solver.engine_precision = args.precision
    rpms, num_modes = workload_spec(args.workload)
    n_problems = len(rpms) * (sm.num_sectors // 2 + 1)
    core = tm._core

    print(f"[bench] workload={args.workload} problems/step={n_problems} nodes={sm.num_nodes}", file=sys.stderr,
          flush=True)
    for w in range(args.warmup):
        t = time.perf_counter()
        solver.solve_rpm_sweep(list(rpms), num_modes, include_coriolis=args.coriolis)
        print(f"[bench] warmup {w} {time.perf_counter() - t:.2f}s", file=sys.stderr, flush=True)
    info = solver.engine_info()

    core.engine_set_profiling(True)
    core.engine_reset_kernel_stats()
    lat = []
    results = None
    t0 = time.perf_counter()
    for s in range(args.steps):
        ts = time.perf_counter()
        results = solver.solve_rpm_sweep(list(rpms), num_modes, include_coriolis=args.coriolis)
        lat += [r.solve_seconds for row in results for r in row]
        print(f"[bench] step {s} {time.perf_counter() - ts:.2f}s", file=sys.stderr, flush=True)
    elapsed = time.perf_counter() - t0
    stats = core.engine_kernel_stats()
    core.engine_set_profiling(False)

    n_ok = sum(len(row) for row in results)
    ms_per_step = 1e3 * elapsed / args.steps
    value = n_problems / (elapsed / args.steps)
    max_res = max(float(np.max(r.residuals)) for row in results for r in row)

    # The timed steps run a sweep's block-size slot groups side by side on their own
    # streams, so a launch's event interval there also holds the other group's
    # kernels.  The per-kernel rates come from one more sweep, right after, with the
    # groups one after another (each kernel alone on the GPU); the same rate over
    # the timed steps is reported beside it.
    kstats = stats
    if not args.serial_groups:
        os.environ["TMHIP_SERIAL_GROUPS"] = "1"
        core.engine_set_profiling(True)
        core.engine_reset_kernel_stats()
        t = time.perf_counter()
        solver.solve_rpm_sweep(list(rpms), num_modes, include_coriolis=args.coriolis)
        print(f"[bench] roofline sweep (groups serial) {time.perf_counter() - t:.2f}s", file=sys.stderr,
              flush=True)
        kstats = core.engine_kernel_stats()
        core.engine_set_profiling(False)
        del os.environ["TMHIP_SERIAL_GROUPS"]

    def rate(st, key):
        tr = st.get(key, {})
        if not tr.get("launches"):
            return None
        avg = tr["total_ms"] / tr["launches"]
        return avg, tr["alg_bytes"] / tr["launches"], tr["alg_flops"] / tr["launches"]

    measured_on = ("the timed steps (groups serial)" if args.serial_groups else
                   "one sweep after the timed steps with the block-size groups serial")
    roofline = roofline_tri = None
    mixed = args.precision == "mixed"
    fr = rate(kstats, "factor_fp32" if mixed else "factor")
    if fr:
        avg_ms, _, flops_per = fr
        ach = flops_per / (avg_ms * 1e-3) / 1e12
        peak = FP32_PEAK_TFLOPS if mixed else FP64_PEAK_TFLOPS
        roofline = {"bound": "mfma", "kernel": "factorization (batched supernodal LDL^H: front assembly, diagonal "
                                              "blocks, panel solves, Schur tiles on v_mfma_%s_16x16x4, pivot-block "
                                              "inversion)" % ("f32" if mixed else "f64"),
                    "achieved": round(ach, 2), "peak": peak, "unit": "TFLOP/s",
                    "frac": round(ach / peak, 4), "traffic": None if mixed else measured_traffic(args, "factor"),
                    "avg_launch_ms": round(avg_ms, 3), "alg_flops_per_launch": flops_per,
                    "measured_on": measured_on}
    tr = rate(kstats, "trisolve_fp32" if mixed else "trisolve")
    if tr:
        avg_ms, bytes_per, _ = tr
        ach = bytes_per / (avg_ms * 1e-3) / 1e9
        roofline_tri = {"bound": "hbm", "kernel": "trisolve (L, D, L^H sweep pair)", "achieved": round(ach, 1),
                        "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(ach / HBM_PEAK_GBS, 4),
                        "traffic": None if mixed else measured_traffic(args, "trisolve"),
                        "avg_launch_ms": round(avg_ms, 3),
                        "alg_bytes_per_launch": bytes_per, "measured_on": measured_on}
        if not args.serial_groups and rate(stats, "trisolve_fp32" if mixed else "trisolve"):
            a2, b2, _ = rate(stats, "trisolve_fp32" if mixed else "trisolve")
            roofline_tri["frac_timed_region_groups_concurrent"] = round(b2 / (a2 * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)
    tri = stats.get("trisolve_fp32" if mixed else "trisolve", {})
    all_res = np.concatenate([np.asarray(r.residuals) for row in results for r in row])

    out = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "ND solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 1),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "c128 (complex f64)" if args.precision == "fp64" else "c64 factor (complex f32) + c128 refinement",
        "data": "synthetic (seeded Kuhn TET10 lattice meshes, %s)" % ("titanium" if args.workload == "c3" else "steel"),
        "config": {"workload": {"c4": "rpm sweep bladed sector 20 rpm x 19 ND", "c2": "bladed sector 19 ND @10000 rpm",
                                "c1": "flat disk 19 ND @0 rpm",
                                "c3": "doubled bladed sector (Ti) 19 ND @12000 rpm"}[args.workload]
                   + (" + Coriolis QEP (k > 0)" if args.coriolis else ""),
                   "n_dof": int(sm.num_dof), "n_free": int(info["n_free"]), "num_sectors": int(sm.num_sectors),
                   "num_modes": num_modes, "problems_per_step": n_problems, "precision": args.precision,
                   "eigensolver": args.method,
                   "nnz_L": int(info["nnz_L"]), "factor_flops": float(info["factor_flops"])},
        "p50_latency_ms": round(1e3 * float(np.percentile(lat, 50)), 1),
        "p99_latency_ms": round(1e3 * float(np.percentile(lat, 99)), 1),
        "solved": n_ok,
        "max_residual": max_res,
        "residuals": {"pairs": int(all_res.size), "above_1e-8": int(np.sum(all_res > 1e-8)),
                      "max": float(all_res.max())},
        "roofline": roofline,
        "roofline_trisolve": roofline_tri,
        "source_hash": source_hash(),
        "kernels": {k: {"launches": v["launches"], "ms": round(v["total_ms"], 2)} for k, v in stats.items()},
        "sweeps_per_step": tri.get("launches", 0) / args.steps,
        "restarts": {"max": int(max(r.restarts for row in results for r in row)),
                     "mean": round(float(np.mean([r.restarts for row in results for r in row])), 2),
                     "worst": sorted(((int(r.restarts), int(r.harmonic_index), float(r.rpm))
                                      for row in results for r in row), reverse=True)[:5]},
        "cpu_baseline": None,
    }

    # accuracy of the timed results against the oracle's answers on a spread sample
    # (rpm 0, 7895, 15000 x ND 0, 9, 18; tests/golden/make_golden.py)
    if args.workload == "c4" and args.method == "block" and not args.coriolis:
        with np.load(os.path.join(ROOT, "tests", "golden", "c4_sample.npz"), allow_pickle=False) as z:
            gold = {k: z[k] for k in z.files}
        sub = gold["subset"]
        errs, macs = [], []
        for ir in (0, 10, 19):
            row = {r.harmonic_index: r for r in results[ir]}
            for k in (0, 9, 18):
                r = row[k]
                f_ref = gold[f"r{ir}_k{k}_freq"]
                errs.append(float(np.max(np.abs(np.asarray(r.frequencies) / f_ref - 1))))
                sh = np.asarray(r.mode_shapes)[sub]
                g = gold[f"r{ir}_k{k}_shape"]
                for q in range(len(f_ref)):
                    a, b = sh[:, q], g[:, q]
                    macs.append(float(abs(np.vdot(a, b)) ** 2 / (np.vdot(a, a).real * np.vdot(b, b).real)))
        out["accuracy_vs_oracle"] = {"problems": len(errs), "sample": "rpm {0, 7895, 15000} x ND {0, 9, 18}",
                                     "max_freq_rel_err": max(errs), "min_mac_on_dof_subset": min(macs)}

    # the plain FP64 path: single-vector Lanczos, one problem per call, over one rpm point
    if not args.no_plain and args.workload in ("c4", "c2"):
        rpm_p = float(rpms[len(rpms) * 2 // 3])
        solver.engine_method = "lanczos"
        solver.engine_precision = "fp64"
        n_h = sm.num_sectors // 2 + 1
        tp = time.perf_counter()
        for k in range(n_h):
            solver.solve_at_rpm(rpm_p, num_modes, [k], include_coriolis=args.coriolis)
        plain = n_h / (time.perf_counter() - tp)
        solver.engine_method = args.method
        solver.engine_precision = args.precision
        out["plain_fp64_one_at_a_time"] = {"value": round(plain, 3), "unit": "ND solves/s", "problems": n_h,
                                           "sample": f"ND 0-{n_h - 1} at {rpm_p:.0f} rpm, one solve_at_rpm call each",
                                           "eigensolver": "lanczos (single vector)"}
        out["speedup_vs_plain_fp64"] = round(value / plain, 2)
        print(f"[bench] plain FP64 one at a time {plain:.2f} ND solves/s", file=sys.stderr, flush=True)

    if args.workload == "c3" or args.coriolis:  # the oracle at 330k DOF / on the QEP takes minutes per harmonic
        args.no_cpu_baseline = True
    if not args.no_cpu_baseline:
        rpm_s = float(rpms[len(rpms) * 2 // 3])
        print(f"[bench] cpu baseline at {rpm_s:.0f} rpm", file=sys.stderr, flush=True)
        mat = meshgen.TITANIUM if args.workload == "c3" else meshgen.STEEL
        cb = cpu_baseline(sm, mat, rpm_s, num_modes)
        n_h = cb["n_h"]
        rate_all = n_h / (cb["t_ops"] + cb["t_all"])
        rate_one = n_h / (cb["t_ops"] + n_h * cb["t_one"])
        out["cpu_baseline"] = {
            "value": round(rate_all, 4), "unit": "ND solves/s", "cores": cb["workers"], "kind": "port",
            "sample": (f"oracle restatement (scipy SuperLU + numpy Lanczos, the reference's policy) on the same mesh, "
                       f"one rpm point ({rpm_s:.0f} rpm): operator build {cb['t_ops']:.1f}s, then ND 0-{n_h - 1} over "
                       f"{cb['workers']} processes (one thread each) in {cb['t_all']:.1f}s; rate = {n_h} ND / "
                       f"(build + solves)"),
            "one_core": {"value": round(rate_one, 4), "unit": "ND solves/s", "cores": 1,
                         "sample": f"ND 0, 1 at {cb['t_one']:.1f}s each; rate = {n_h} ND / (build + {n_h} x that)"},
            "host": host_description()}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
