#!/usr/bin/env python3
"""Benchmark: time to a 1e-8 relative residual on the 8M-cell heterogeneous
diffusion system, and HBM GB/s of the finest-level SpMV / smoother.

Workload (one GPU): 200^3 hexahedral mesh of the unit cube, interior nodes
jittered by U[-0.2h, 0.2h], conductivity exp(2.5 z) constant in 8^3-cell
blocks, source U[0,1], T=1 at x=0 and T=0 at x=1, zero flux elsewhere,
assembled on the GPU (thermal finite-volume discretization). One step = build
the AMG hierarchy + solve to ||r||_1 / ||r0||_1 < 1e-8 (FP64 true residual)
with FP64 FGMRES(30) preconditioned by an FP32 V-cycle (BF16 coarse
off-diagonals, FP64 dense coarsest solve). Reported beside it: the plain FP64
GPU AMG (coloured Gauss-Seidel) on the same system, the cycle count of the
reference-order AMG, and the CPU oracle timed on this host.

    python bench.py [--gpus 1] [--steps 3] [--warmup 1]
"""
import argparse
import ctypes
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

SEED = 20261202
METRIC = "time-to-1e-8 rel. residual, 8M-cell diffusion; SpMV/smoother HBM GB/s"
HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def build_workload(N, torch):
    from fvm_amd.thermal import Mesh, assemble
    t0 = time.perf_counter()
    mesh = Mesh.hex(N, 0.2, SEED)
    nt = mesh.n_total
    k = torch.empty(nt, dtype=torch.float64, device="cuda:0")
    s = torch.empty(nt, dtype=torch.float64, device="cuda:0")
    x = torch.zeros(nt, dtype=torch.float64, device="cuda:0")
    mesh.lognormal_field(k, 2.5, 8, SEED + 1)
    mesh.uniform_field(s, 0.0, 1.0, SEED + 2)
    bc_t = [0, 0, 0, 1, 1, 1, 1]
    bc_v = [0.0, 1.0, 0.0, 0.0, 0.0, 0.0, 0.0]
    S = assemble(mesh, k, s, x, bc_t, bc_v)
    torch.cuda.synchronize()
    return mesh, S, time.perf_counter() - t0


def make_amg(gs_order, precision, maxit=100000):
    import fvm_amd
    a = fvm_amd.AMG(device=0)
    a.verbosity = 0
    a.gsOrdering = gs_order
    a.precision = precision
    a.relativeTolerance = 1e-8
    a.nMaxIterations = maxit
    return a


def setup_on(amg, S):
    from fvm_amd._lib import check, lib
    amg._push_params()
    check(lib().fvm_amg_setup_system(amg._h, S._h))


def level_sizes(amg):
    return [(amg.level(l)["n"], amg.level(l)["nnz"]) for l in range(amg.num_levels())]


def kernel_stats(amg, kind):
    """(launches, total ms, total algorithmic bytes) of the handle's timed
    finest-level kernels of one kind (0: FP64 residual/SpMV, 1: smoother sweep
    directions), measured with HIP events on the handle's stream."""
    from fvm_amd._lib import check, lib
    n = ctypes.c_longlong()
    ms = ctypes.c_double()
    nb = ctypes.c_double()
    check(lib().fvm_amg_kernel_timing(amg._h, kind, ctypes.byref(n), ctypes.byref(ms)))
    check(lib().fvm_amg_kernel_bytes(amg._h, kind, ctypes.byref(nb)))
    return n.value, ms.value, nb.value


def kernel_entry(st, formula):
    """Per-launch figures of kernel_stats(): average duration, average algorithmic
    bytes and their rate."""
    n, ms, nb = st
    return dict(launches=n, avg_us=1e3 * ms / n, bytes=nb / n, bytes_formula=formula,
                gbs=nb / (ms * 1e-3) / 1e9)


def kernel_source_hash():
    """First 16 hex digits of the SHA-256 of the cycle kernels' source."""
    import hashlib
    with open(os.path.join(ROOT, "fvm_amd", "csrc", "cycle.hip"), "rb") as f:
        return hashlib.sha256(f.read()).hexdigest()[:16]


def profiled_traffic(kernel, N):
    """HBM bytes per launch of `kernel` from the committed counter profile
    (profiles/hbm_traffic.json: two rocprofv3 passes, FETCH_SIZE and WRITE_SIZE,
    scaled by a plain copy's counts), used only when the profile was taken on this
    workload size with the current kernel source; otherwise (None, reason)."""
    path = os.path.join(ROOT, "profiles", "hbm_traffic.json")
    try:
        with open(path) as f:
            prof = json.load(f)
    except (OSError, ValueError):
        return None, "no counter profile"
    k = prof.get("kernels", {}).get(kernel)
    if k is None:
        return None, "kernel not in the counter profile"
    if prof.get("n") != N:
        return None, "counter profile taken on %s^3" % prof.get("n")
    if prof.get("source_sha16") != kernel_source_hash():
        return None, "counter profile stale (kernel source changed since it was taken)"
    return k["bytes_per_launch"], "profiles/hbm_traffic.json"


def host_cpu():
    """CPU model and logical core count of this host (/proc/cpuinfo, os.cpu_count)."""
    model = None
    try:
        for ln in open("/proc/cpuinfo"):
            if ln.startswith("model name"):
                model = ln.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return model, os.cpu_count()


def _median(v):
    v = sorted(v)
    return v[len(v) // 2] if len(v) % 2 else 0.5 * (v[len(v) // 2 - 1] + v[len(v) // 2])


def oracle_cycle_model(O, h, reps=1):
    """Setup time, then the solve time of 1 and of 6 cycles on that hierarchy:
    seconds per cycle = (T6 - T1) / 5, fixed cost = T1 - one cycle (initial
    residual and norm). Medians over `reps`."""
    a = O.OracleAMG(n_max_iterations=2, relative_tolerance=1e-30)
    t0 = time.perf_counter()
    a.setup(h["n_self"], h["n_total"], h["row"], h["col"], h["diag"], h["off"], h["is_boundary"])
    setup = time.perf_counter() - t0
    t1s, t6s = [], []
    for _ in range(reps):
        for k, acc in ((2, t1s), (7, t6s)):
            a._p.n_max_iterations = k
            t0 = time.perf_counter()
            a.solve(h["b"])
            acc.append(time.perf_counter() - t0)
    per = (_median(t6s) - _median(t1s)) / 5
    return setup, per, _median(t1s) - per


def cpu_validation(torch, runs=5):
    """The oracle (single thread) to convergence on the 32^3 uniform conductivity-1
    system (T = 1 / 0 on the x walls, zero flux elsewhere, relTol 1e-8): median of
    `runs` timed runs after one warm-up, setup and solve apart; and the
    extrapolation model used for the 8M-cell system (oracle_cycle_model) applied
    to the same system, compared with the measured time."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as O
    from fvm_amd.thermal import Mesh, assemble
    mesh = Mesh.hex(32, 0.0, 0)
    nt = mesh.n_total
    k = torch.ones(nt, dtype=torch.float64, device="cuda:0")
    z = torch.zeros(nt, dtype=torch.float64, device="cuda:0")
    h = assemble(mesh, k, z, z, [0, 0, 0, 1, 1, 1, 1], [0.0, 1.0, 0.0, 0.0, 0.0, 0.0, 0.0]).to_host()
    setups, solves, cycles = [], [], None
    for r in range(runs + 1):
        a = O.OracleAMG(n_max_iterations=10000, relative_tolerance=1e-8)
        t0 = time.perf_counter()
        a.setup(h["n_self"], h["n_total"], h["row"], h["col"], h["diag"], h["off"], h["is_boundary"])
        t1 = time.perf_counter()
        res = a.solve(h["b"])
        t2 = time.perf_counter()
        cycles = res["cycles"]
        if r:
            setups.append(t1 - t0)
            solves.append(t2 - t1)
    _, per, fixed = oracle_cycle_model(O, h, reps=runs)
    measured = _median(setups) + _median(solves)
    model = _median(setups) + fixed + cycles * per
    return dict(workload="32^3 uniform, k=1", runs=runs, cycles=cycles, setup_s=round(_median(setups), 5),
                solve_s=round(_median(solves), 5), measured_s=round(measured, 5), model_s=round(model, 5),
                model_error=round(model / measured - 1.0, 4))


def cpu_baseline(S, cycles_needed, torch=None):
    """Oracle (C restatement of the reference AMG, single thread) on the same
    system: the hierarchy build, then 1 and 6 cycles, extrapolated to the number of
    cycles the reference-order AMG needs for 1e-8 (setup + fixed + cycles x per
    cycle); the model is checked against full timed runs on the 32^3 system
    (cpu_validation)."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as O
    h = S.to_host()
    setup, per_cycle, fixed = oracle_cycle_model(O, h)
    model, ncpu = host_cpu()
    out = dict(value=setup + fixed + cycles_needed * per_cycle, unit="s", cores=1, kind="port",
               sample="oracle setup (%.2f s), then 1 and 6 V-cycles (%.3f s/cycle) on the same 8M-cell system, "
                      "extrapolated to the %d cycles of the reference-order AMG"
                      % (setup, per_cycle, cycles_needed),
               setup_s=setup, per_cycle_s=per_cycle, fixed_s=fixed, host_threads=ncpu, cpu_model=model)
    if torch is not None:
        out["validation"] = cpu_validation(torch)
    return out


def fib_direction(i, K):
    import math
    z = 1.0 - (2 * i + 1) / K
    phi = i * math.pi * (3.0 - math.sqrt(5.0))
    r = math.sqrt(max(0.0, 1.0 - z * z))
    return [r * math.cos(phi), r * math.sin(phi), z]


def build_batch_workload(torch, K=256, lo=1e4, hi=5e5):
    """K discrete-ordinate systems (gray phonon BTE): m^3 jittered hex cubes with
    n = round(exp(U[ln lo, ln hi])), m = round(n^(1/3)); direction i on a
    Fibonacci sphere; tau_i ~ U[0.1, 10]; e0 = 1, field 0, inflow value 1 on the
    x=0 wall (0 elsewhere). Meshes and their metrics are built once."""
    import numpy as np
    from fvm_amd.thermal import Mesh
    rng = np.random.default_rng([SEED, 4])
    n = np.round(np.exp(rng.uniform(np.log(lo), np.log(hi), K)))
    m = np.maximum(2, np.round(np.cbrt(n))).astype(int)
    tau = rng.uniform(0.1, 10.0, K)
    meshes, fields = [], []
    for i in range(K):
        mesh = Mesh.hex(int(m[i]), 0.2, SEED + 1000 + i)
        mesh.metrics_device()
        nc = mesh.n_cells
        x = torch.zeros(mesh.n_total, dtype=torch.float64, device="cuda:0")
        x[nc: nc + int(m[i]) ** 2] = 1.0          # ghosts of the x=0 wall (first boundary group)
        e0 = torch.ones(nc, dtype=torch.float64, device="cuda:0")
        meshes.append(mesh)
        fields.append((fib_direction(i, K), float(tau[i]), e0, x))
    return meshes, fields, m


def batch_params(obj, gs_order=1):
    obj.relativeTolerance = 1e-8
    obj.nMaxIterations = 500
    obj.verbosity = 0
    obj.gsOrdering = gs_order


def run_batch(args):
    """Per sweep of all K directions: assemble every system, build every
    hierarchy, solve every system to 1e-8 (FP64 AMG per system with coloured
    Gauss-Seidel, PhononModel::advance), all systems in one batch handle."""
    import numpy as np
    import torch
    import fvm_amd
    from fvm_amd.thermal import assemble_phonon
    K = args.batch
    log("building %d direction systems" % K)
    meshes, fields, m = build_batch_workload(torch, K)
    ncells = int(sum(me.n_cells for me in meshes))
    B = fvm_amd.BatchAMG()
    batch_params(B)

    def sweep():
        systems = [assemble_phonon(me, *f) for me, f in zip(meshes, fields)]
        devs = [dict(S.device_pointers(), n_self=S.n_self, n_total=S.n_total) for S in systems]
        B.setup_device(devs)
        xs = [torch.zeros(S.n_total, dtype=torch.float64, device="cuda:0") for S in systems]
        r0, rf, cyc = B.solve_device([d["b"] for d in devs], xs)
        return systems, xs, r0, rf, cyc

    for w in range(args.warmup):
        _, _, r0, rf, cyc = sweep()
        log("warmup %d: cycles min %d max %d" % (w, cyc.min(), cyc.max()))
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(args.steps):
        systems, xs, r0, rf, cyc = sweep()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    ms = 1e3 * (t1 - t0) / args.steps
    st = B.setup_stats()
    # phases of one more sweep
    torch.cuda.synchronize(); ta = time.perf_counter()
    systems = [assemble_phonon(me, *f) for me, f in zip(meshes, fields)]
    torch.cuda.synchronize(); tb = time.perf_counter()
    devs = [dict(S.device_pointers(), n_self=S.n_self, n_total=S.n_total) for S in systems]
    B.setup_device(devs)
    torch.cuda.synchronize(); tc = time.perf_counter()
    xs = [torch.zeros(S.n_total, dtype=torch.float64, device="cuda:0") for S in systems]
    B.solve_device([d["b"] for d in devs], xs)
    torch.cuda.synchronize(); td = time.perf_counter()
    log("batched sweep: %.1f ms (assemble %.1f, setup %.1f, solve %.1f ms), cycles %d..%d"
        % (ms, 1e3 * (tb - ta), 1e3 * (tc - tb), 1e3 * (td - tc), cyc.min(), cyc.max()))
    # the same systems one handle at a time (unbatched loop, as the reference calls AMG::solve)
    t_loop = None
    if not args.skip_plain:
        a = fvm_amd.AMG()
        batch_params(a)
        torch.cuda.synchronize(); tl0 = time.perf_counter()
        for S, d in zip(systems, devs):
            x = torch.zeros(S.n_total, dtype=torch.float64, device="cuda:0")
            a.setup_device(S.n_self, S.n_total, d["row"], d["col"], d["diag"], d["off"], d["is_boundary"])
            a.solve_device(d["b"], x)
        torch.cuda.synchronize(); t_loop = time.perf_counter() - tl0
        log("unbatched loop over %d systems: %.3f s" % (K, t_loop))
    cpu = None
    if not args.skip_cpu:
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        import oracle_lib as O
        order = np.argsort([me.n_cells for me in meshes])
        sample = [int(order[j]) for j in np.linspace(0, K - 1, 6).astype(int)]
        t_cpu, cells = 0.0, 0
        for i in sample:
            h = systems[i].to_host()
            a = O.OracleAMG(relative_tolerance=1e-8, n_max_iterations=200)
            c0 = time.perf_counter()
            a.setup(h["n_self"], h["n_total"], h["row"], h["col"], h["diag"], h["off"], h["is_boundary"])
            a.solve(h["b"])
            t_cpu += time.perf_counter() - c0
            cells += h["n_self"]
        cpu = dict(value=round(t_cpu * ncells / cells, 3), unit="s", cores=1, kind="port",
                   sample="oracle setup+solve of %d of the %d systems (%d cells, %.2f s), scaled by cell count "
                          "to all %d cells" % (len(sample), K, cells, t_cpu, ncells))
    out = {
        "metric": "time per sweep of %d direction systems (assemble + AMG setup + solve to 1e-8)" % K,
        "value": round(ms / 1e3, 6), "unit": "s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(ms, 3), "higher_is_better": False, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic (seeded meshes, directions and relaxation times)",
        "config": {"workload": "%d phonon-BTE directions, m^3 jittered hex cubes, m in [%d, %d], %d cells total; "
                               "upwind convection + relaxation, FP64 AMG (coloured GS) per system, batched"
                               % (K, int(m.min()), int(m.max()), ncells)},
        "cycles_min": int(cyc.min()), "cycles_max": int(cyc.max()), "cycles_mean": float(cyc.mean()),
        "assemble_s": round(tb - ta, 4), "setup_s": round(tc - tb, 4), "solve_s": round(td - tc, 4),
        "levels": st["levels"], "unbatched_loop_s": None if t_loop is None else round(t_loop, 4),
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


def apply_knobs(amg, args):
    """The launch shapes the autotuner chose in a timed run (printed in its JSON
    line under autotune.chosen); none of them changes a result."""
    for key in ("sell_block", "spmv_block", "tail_max_rows", "sell_xcd", "spmv_xcd", "cycle_graph", "nt_levels"):
        amg.set_tuning(key, getattr(args, key))


def run_kinetic(args):
    """Kinetic model (KineticModel::advance): Couette slab on an N^3 hex mesh,
    walls at x = 0 and x = 1 moving in -y / +y at temperature 1, specular symmetry on
    the other faces, Q^3 Cartesian velocities (clim 5.5, T2 1), ES-BGK (fgamma 2), the
    kinetic linear solver's defaults (AMG to 1e-15 relative, at most 100 cycles per
    direction). A step is one outer iteration: every velocity assembled, all solved
    by one batched AMG, then walls, moments, collision frequency and the BGK and
    ES-BGK Newton equilibria. CPU baseline: the oracle's assembly + AMG for a sample of
    directions, scaled to all of them, plus its per-cell update once."""
    import torch
    import fvm_amd  # noqa: F401
    from fvm_amd.kinetic import KineticModel, Quadrature
    from fvm_amd.thermal import Mesh
    N, Q = args.kinetic, args.kinetic_q
    mesh = Mesh.hex(N, 0.0, SEED)
    quad = Quadrature(Q, Q, Q, 5.5, 1.0)
    K = quad.getDirCount()
    bcs = {1: ("wall", (0.0, -0.2, 0.0), 1.0), 2: ("wall", (0.0, 0.2, 0.0), 1.0),
           3: ("symmetry",), 4: ("symmetry",), 5: ("symmetry",), 6: ("symmetry",)}
    model = KineticModel(mesh, quad, bcs, fgamma=2)
    model.set_solver(gsOrdering=1)     # coloured Gauss-Seidel (the reference-order smoother is the parity mode)
    log("kinetic: %d^3 cells, %d velocities (%d unknowns per iteration)" % (N, K, K * N ** 3))
    for w in range(args.warmup):
        t0 = time.perf_counter()
        model.advance(1, relTolerance=0.0, absTolerance=0.0)
        log("warmup %d: %.3f s, norm %.4e" % (w, time.perf_counter() - t0, model.norms[-1]))
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    model.advance(args.steps, relTolerance=0.0, absTolerance=0.0)
    torch.cuda.synchronize()
    per = (time.perf_counter() - t0) / args.steps
    log("kinetic: %.3f s per outer iteration, norms %s" % (per, ["%.4e" % v for v in model.norms]))
    cpu = None
    if not args.skip_cpu:
        cpu = kinetic_cpu_baseline(N, Q, K, bcs)
        log("cpu: %.1f s per iteration (%s)" % (cpu["value"], cpu["sample"]))
    out = {
        "metric": "time per outer iteration, ES-BGK kinetic model (all velocities assembled and solved)",
        "value": round(per, 6), "unit": "s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(1e3 * per, 3), "higher_is_better": False, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic (Couette slab, rest-state Maxwellian start)",
        "config": {"workload": "%d^3 hex Couette slab, %d^3 Cartesian velocities (%d systems, %d unknowns), ES-BGK, "
                               "AMG (coloured GS) per direction to 1e-15 / 100 cycles, batched" % (N, Q, K, K * N ** 3)},
        "norms": model.norms, "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


def kinetic_cpu_baseline(N, Q, K, bcs, sample=3):
    """The oracle on the same case: direction assembly + AMG (to 1e-15, 100 cycles)
    for `sample` directions timed and scaled to K, plus one per-cell update (walls,
    moments, collision frequency, Newton BGK over every cell) timed once."""
    import numpy as np
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import hexmesh
    import kinetic_ref as KR
    import oracle_lib as O
    mesh = hexmesh.hex_mesh(N, 0.0, SEED)
    nc = mesh["n_cells"]
    nct = nc + mesh["n_ghost"]
    g = O.fv_metrics(mesh)
    quad = KR.quadrature_cartesian(Q, Q, Q, 5.5, 1.0)
    o = dict(KR.DEFAULTS)      # fgamma 2
    st = KR.init_state(nct, quad, o)
    kind = np.zeros(7, np.int32)
    wall_u = np.zeros((7, 3)); wall_u[1] = bcs[1][1]; wall_u[2] = bcs[2][1]
    # directions across the velocity grid (a wall-normal one included)
    picks = np.linspace(0, K - 1, sample + 2).astype(int)[1:-1]
    # the systems of a working iteration: after the walls have acted once
    refl = {}
    gp = mesh["group_ptr"]
    for gi in (3, 4, 5, 6):
        f0 = gp[gi]
        refl[gi] = KR.reflection_map(g["fA"][3 * f0:3 * f0 + 3] / g["fAm"][f0], *quad[:3])
    kb = {gi: ("wall", np.array(bc[1]), bc[2]) if bc[0] == "wall" else bc for gi, bc in bcs.items()}
    KR.boundary_values(mesh, g["fA"], g["fAm"], kb, refl, st["f"], (st["density"], st["v"], st["T"]), *quad,
                       o["epsilon"])
    t0 = time.perf_counter()
    for j in picks:
        S = O.fv_kinetic(mesh, [quad[0][j], quad[1][j], quad[2][j]], st["nu"][:nc], st["feq_lin"][j][:nc], st["f"][j],
                         kind, wall_u)
        oa = O.OracleAMG(n_max_iterations=100, relative_tolerance=1e-15)
        oa.setup(S["n_self"], S["n_total"], S["row"], S["col"], S["diag"], S["off"], S["is_boundary"])
        oa.solve(S["b"])
    per_dir = (time.perf_counter() - t0) / sample
    t1 = time.perf_counter()
    macro = (st["density"], st["v"], st["T"])
    KR.boundary_values(mesh, g["fA"], g["fAm"], kb, refl, st["f"], macro, *quad, o["epsilon"])
    d, v, T, _ = KR.macroparameters(st["f"], *quad)
    KR.collision_frequency(d, T, o)
    t_all = time.perf_counter() - t1
    # the per-cell Newton fits on a sample of cells, scaled to all of them
    M = min(nct, 1024)
    t2 = time.perf_counter()
    KR.newton_bgk(d[:M], v[:M], T[:M], st["coeff"][:M], *quad, o)
    feq = KR.equilibrium(st["coeff"][:M], v[:M], *quad[:3])
    T6 = KR.macroparameters_es(st["f"][:, :M], feq, v[:M], d[:M], *quad, o["Prandtl"])
    KR.newton_esbgk(d[:M], v[:M], T6, st["coeffg"][:M], *quad, o)
    t_cells = (time.perf_counter() - t2) * nct / M
    upd = t_all + t_cells
    return {"value": per_dir * K + upd, "unit": "s", "cores": 1, "kind": "port",
            "sample": "oracle assembly + AMG of %d of the %d directions (%.3f s each), scaled; walls, moments and "
                      "collision frequency of all cells (%.2f s); the BGK and ES-BGK Newton fits of %d of the %d "
                      "cells, scaled (%.2f s)" % (sample, K, per_dir, t_all, M, nct, t_cells)}


def run_profile(args):
    """Short program for rocprofv3 (kernel trace and PMC passes): the 8M-cell
    system, `--steps` headline solves (setup, FP64 FGMRES(30) with the mixed
    FP32/BF16 AMG to 1e-8, with the given launch shapes) with the finest-level
    kernels timed live as in the timed run, three plain FP64 cycles of the
    coloured hierarchy, and one 1 GiB device copy (512 MiB read + 512 MiB
    written) that calibrates the HBM byte counters."""
    import torch
    import fvm_amd
    from fvm_amd._lib import check, lib
    N = args.n
    mesh, S, _ = build_workload(N, torch)
    n_total = S.n_total
    b = S.device_pointers()["b"]
    amg = make_amg(1, 1)
    apply_knobs(amg, args)
    fg = fvm_amd.FGMRES(amg)
    fg.restart = 30
    fg.relativeTolerance = 1e-8
    fg.nMaxIterations = 2000
    check(lib().fvm_amg_set_kernel_timing(amg._h, 1))
    for _ in range(args.steps):
        x = torch.zeros(n_total, dtype=torch.float64, device="cuda:0")
        setup_on(amg, S)
        r0, rf, it = fg.solve_device(b, x)
    torch.cuda.synchronize()
    lv = level_sizes(amg)
    kernels = {"finest_fp64_spmv_residual": kernel_entry(kernel_stats(amg, 0), "28n+12nnz"),
               "finest_fp32_gs_sweep": kernel_entry(kernel_stats(amg, 1), "20 per row + 8 per off-diagonal")}
    log("profiled %d mixed solves: n=%d nnz=%d, %d FGMRES iterations; live: %s"
        % (args.steps, lv[0][0], lv[0][1], it, json.dumps(kernels)))
    del fg, amg
    plain = make_amg(1, 0)
    apply_knobs(plain, args)
    setup_on(plain, S)
    x = torch.zeros(n_total, dtype=torch.float64, device="cuda:0")
    for _ in range(3):
        check(lib().fvm_amg_smooth_device(plain._h, ctypes.c_void_p(b), ctypes.c_void_p(x.data_ptr())))
    torch.cuda.synchronize()
    log("profiled 3 plain FP64 cycles")
    del plain
    src = torch.ones(64 << 20, dtype=torch.float64, device="cuda:0")
    dst = torch.empty_like(src)
    dst.copy_(src)
    torch.cuda.synchronize()
    log("calibration copy: %d bytes read, %d bytes written" % (src.numel() * 8, dst.numel() * 8))


def _counter_rows(path, counter):
    """Dispatches of a rocprofv3 counter-collection CSV: (dispatch id, kernel
    name, grid size, value of `counter`), in dispatch order."""
    import csv
    rows = []
    with open(path, newline="") as f:
        for r in csv.DictReader(f):
            if r.get("Counter_Name") != counter:
                continue
            rows.append((int(r["Dispatch_Id"]), r["Kernel_Name"], int(r["Grid_Size"]), float(r["Counter_Value"])))
    rows.sort()
    return rows


def write_traffic_profile(fetch_csv, write_csv, n, out_path):
    """HBM bytes per launch of the finest-level kernels from the two counter
    passes of `bench.py --profile` (FETCH_SIZE, then WRITE_SIZE, KiB per
    dispatch), each scaled by what the counter reads for the program's final
    512 MiB copy. The SpMV is its largest-grid dispatches; a smoother direction is
    the FP32 colour launches with at least n^3/4 threads (the finest level's
    colours: no coarser level has that many rows in one colour), summed per
    headline solve cycle and halved (forward + backward)."""
    fetch = _counter_rows(fetch_csv, "FETCH_SIZE")
    write = _counter_rows(write_csv, "WRITE_SIZE")
    copy_bytes = 512 << 20
    scale = {}
    for name, rows in (("FETCH_SIZE", fetch), ("WRITE_SIZE", write)):
        cp = [r for r in rows if "copyBuffer" in r[1]]
        if not cp:
            raise SystemExit("no calibration copy in %s" % name)
        scale[name] = copy_bytes / (cp[-1][3] * 1024.0)
    N3 = n ** 3

    def per_launch(pick, group=1):
        tot = []
        for rows, name in ((fetch, "FETCH_SIZE"), (write, "WRITE_SIZE")):
            sel = [r for r in rows if pick(r)]
            tot.append((len(sel), sum(r[3] for r in sel) * 1024.0 * scale[name]))
        (nf, bf), (nw, bw) = tot
        if nf != nw or nf == 0:
            raise SystemExit("dispatch counts differ between the passes (%d, %d)" % (nf, nw))
        return dict(dispatches=nf, fetch_bytes=bf / nf * group, write_bytes=bw / nf * group,
                    bytes_per_launch=(bf + bw) / nf * group)

    spmv = [r for r in fetch if "k_sell_spmv" in r[1]]
    gmax = max(r[2] for r in spmv)
    kern = {"finest_fp64_spmv_residual": per_launch(lambda r: "k_sell_spmv" in r[1] and r[2] == gmax)}
    sweeps = lambda r: "k_sell_gs<" in r[1] and "float" in r[1] and "double" not in r[1] and r[2] >= N3 // 4
    nsw = sum(1 for r in fetch if sweeps(r))
    if nsw == 0 or nsw % 3:
        raise SystemExit("finest smoother launches (%d) are not 3 per cycle (two colours expected)" % nsw)
    # 2 forward colour launches and 1 backward launch per cycle with a two-colouring
    kern["finest_fp32_gs_sweep"] = per_launch(sweeps, group=nsw / (2.0 * (nsw / 3.0)))
    kern["finest_fp32_gs_sweep"]["note"] = "per direction: (2 forward + 1 backward colour launches) / 2"
    prof = {"n": n, "source_sha16": kernel_source_hash(),
            "scale": {k: round(v, 4) for k, v in scale.items()}, "kernels": kern}
    with open(out_path, "w") as f:
        json.dump(prof, f, indent=1)
    print(json.dumps(prof, indent=1))


def thermal_system(torch, N, sigma, block, seed):
    """Jittered N^3 heat-conduction system assembled on the GPU: lognormal k
    (per cell for block 1, else constant in block^3 cells), source U[0,1],
    T=1 at x=0, T=0 at x=1, zero flux elsewhere."""
    from fvm_amd.thermal import Mesh, assemble
    mesh = Mesh.hex(N, 0.2, seed)
    nt = mesh.n_total
    k = torch.empty(nt, dtype=torch.float64, device="cuda:0")
    s = torch.empty(nt, dtype=torch.float64, device="cuda:0")
    x = torch.zeros(nt, dtype=torch.float64, device="cuda:0")
    mesh.lognormal_field(k, sigma, block, seed + 1)
    mesh.uniform_field(s, 0.0, 1.0, seed + 2)
    S = assemble(mesh, k, s, x, [0, 0, 0, 1, 1, 1, 1], [0.0, 1.0, 0.0, 0.0, 0.0, 0.0, 0.0])
    return mesh, S


def level0_graph(torch, amg):
    """Device CSR of the hierarchy's finest level (self rows, natural order)."""
    lv = amg.level(0)
    t = lambda a, dt: torch.tensor(a, dtype=dt, device="cuda:0")
    return (lv["n"], t(lv["row"], torch.int32), t(lv["col"], torch.int32), t(lv["off"], torch.float64),
            t(lv["diag"], torch.float64))


def residual_of(torch, g, x, b):
    """r = b + A x on the self rows of graph g = (n, row, col, off, diag)."""
    n, row, col, off, diag = g
    rid = torch.repeat_interleave(torch.arange(n, device="cuda:0"), (row[1:] - row[:-1]).long())
    y = diag * x[:n] + b[:n]
    return y.index_add_(0, rid, off * x[col.long()])


def run_learned(args):
    """Training pairs from plain AMG on matrices of the 2.1M-cell workload's
    distribution (at 32^3), training on the GPU, then the ablation on the
    4M-cell system: FGMRES(30) + mixed AMG with and without the learned correction."""
    import numpy as np
    import torch
    import fvm_amd
    from fvm_amd import LearnedCorrection
    rng = np.random.default_rng([SEED, 5])
    n_mat, per_mat, Nt = args.train_matrices, args.train_pairs // args.train_matrices, 32
    cycles = [int(c) for c in args.train_cycles.split(",")]
    log("generating %d residual/error pairs from plain AMG on %d matrices (%d^3, sigma 1.5 per cell)"
        % (n_mat * per_mat, n_mat, Nt))
    t0 = time.perf_counter()
    samples, keep = [], []
    for j in range(n_mat):
        mesh, S = thermal_system(torch, Nt, 1.5, 1, SEED + 100 + j)
        amg = make_amg(1, 0, maxit=10)
        setup_on(amg, S)
        g = level0_graph(torch, amg)
        d = S.device_pointers()
        zero = torch.zeros(S.n_total, dtype=torch.float64, device="cuda:0")
        for q in range(per_mat):
            # exact solution first (white noise), b = -A xs, then one or two cycles of
            # plain AMG from zero: error e = xs - x and residual r = b + A x (A e = -r)
            xs = torch.tensor(rng.normal(size=S.n_total), dtype=torch.float64, device="cuda:0")
            xs[S.n_self:] = 0
            b = torch.zeros(S.n_total, dtype=torch.float64, device="cuda:0")
            b[:S.n_self] = -residual_of(torch, g, xs, zero)
            x = torch.zeros(S.n_total, dtype=torch.float64, device="cuda:0")
            from fvm_amd._lib import check, lib
            for _ in range(cycles[q % len(cycles)]):
                check(lib().fvm_amg_smooth_device(amg._h, ctypes.c_void_p(b.data_ptr()),
                                                  ctypes.c_void_p(x.data_ptr())))
            samples.append((g, residual_of(torch, g, x, b).contiguous(), (xs[:S.n_self] - x[:S.n_self]).contiguous()))
        keep.append((mesh, S, amg))
    torch.cuda.synchronize()
    t_gen = time.perf_counter() - t0
    net = LearnedCorrection(seed=args.seed)
    t0 = time.perf_counter()
    curve = net.train(samples, steps=args.train_steps, lr=args.lr, batch=8, log=log)
    torch.cuda.synchronize()
    t_train = time.perf_counter() - t0
    del samples, keep
    # ablation on the 4M-cell system (160^3, sigma 2.0 in 8^3 blocks)
    N = args.n if args.n != 200 else 160
    log("ablation on %d^3" % N)
    mesh, S = thermal_system(torch, N, 2.0, 8, SEED + 7)
    b = S.device_pointers()["b"]
    amg = make_amg(1, 1)
    setup_on(amg, S)
    amg.autotune(reps=2)
    # held-out pairs on the ablation system itself (plain FP64 AMG, one cycle):
    # energy-norm error after the cycle, and after the cycle plus the correction
    g = level0_graph(torch, amg)
    plain = make_amg(1, 0, maxit=10)
    setup_on(plain, S)
    zero = torch.zeros(S.n_total, dtype=torch.float64, device="cuda:0")
    net.set_graph(*g)
    ratios = []
    for q in range(4):
        xs = torch.tensor(rng.normal(size=S.n_total), dtype=torch.float64, device="cuda:0")
        xs[S.n_self:] = 0
        bq = torch.zeros(S.n_total, dtype=torch.float64, device="cuda:0")
        bq[:S.n_self] = -residual_of(torch, g, xs, zero)
        x = torch.zeros(S.n_total, dtype=torch.float64, device="cuda:0")
        from fvm_amd._lib import check, lib
        check(lib().fvm_amg_smooth_device(plain._h, ctypes.c_void_p(bq.data_ptr()), ctypes.c_void_p(x.data_ptr())))
        r = residual_of(torch, g, x, bq).contiguous()
        e = (xs[:S.n_self] - x[:S.n_self]).contiguous()
        en = lambda d: float(-(d @ residual_of(torch, g, d, torch.zeros_like(d))))
        corr = torch.zeros(S.n_self, dtype=torch.float64, device="cuda:0")
        net.forward(r, corr)
        ratios.append((en(e) / en(xs[:S.n_self].contiguous()), en((e - corr).contiguous()) / en(e)))
    heldout = dict(energy_after_cycle=float(np.mean([a for a, _ in ratios])),
                   energy_after_cycle_and_correction_rel=float(np.mean([b_ for _, b_ in ratios])))
    log("held-out pairs on %d^3: %s" % (N, heldout))
    # stationary iteration x <- x + M(b + A x) with M = plain AMG cycle (+ correction)
    stat = {}
    bfull = torch.zeros(S.n_total, dtype=torch.float64, device="cuda:0")
    bfull.copy_(torch.as_tensor(S.to_host()["b"], device="cuda:0"))
    for label, use in (("without", False), ("with", True)):
        x = torch.zeros(S.n_total, dtype=torch.float64, device="cuda:0")
        r0 = float(residual_of(torch, g, x, bfull).abs().sum())
        cyc, ratio = 0, 1.0
        torch.cuda.synchronize(); t0 = time.perf_counter()
        while cyc < args.stationary_max and ratio >= 1e-8:
            check(lib().fvm_amg_smooth_device(plain._h, ctypes.c_void_p(bfull.data_ptr()), ctypes.c_void_p(x.data_ptr())))
            if use:
                r = residual_of(torch, g, x, bfull).contiguous()
                xv = x[:S.n_self]
                net.forward(r, xv, accumulate=True) if xv.is_contiguous() else None
            cyc += 1
            ratio = float(residual_of(torch, g, x, bfull).abs().sum()) / r0
        torch.cuda.synchronize()
        stat[label] = dict(cycles=cyc, ratio=ratio, seconds=round(time.perf_counter() - t0, 3))
        log("stationary AMG %s correction: %s" % (label, stat[label]))
    del plain
    res = {}
    for label, use in (("without", False), ("with", True)):
        fg = fvm_amd.FGMRES(amg)
        fg.restart = 30
        fg.relativeTolerance = 1e-8
        fg.nMaxIterations = 2000
        if use:
            fg.set_learned(net, args.stall)
        times = []
        for k in range(args.steps + 1):
            x = torch.zeros(S.n_total, dtype=torch.float64, device="cuda:0")
            torch.cuda.synchronize(); t0 = time.perf_counter()
            r0, rf, it = fg.solve_device(b, x)
            torch.cuda.synchronize(); times.append(time.perf_counter() - t0)
        res[label] = dict(iterations=it, solve_s=round(min(times[1:]), 4), ratio=rf / r0,
                          **(fg.learned_stats() if use else {}))
        log("%s learned correction: %s" % (label, res[label]))
    out = {
        "metric": "FGMRES(30)+mixed AMG solve time to 1e-8 with vs without the learned correction",
        "value": res["with"]["solve_s"], "unit": "s", "n_gpus": 1, "steps": args.steps, "warmup": 1,
        "ms_per_step": round(1e3 * res["with"]["solve_s"], 3), "higher_is_better": False, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64 outer, f32/bf16 V-cycle, bf16 network on MFMA (f32 accumulate)",
        "data": "synthetic (seeded meshes and fields)",
        "config": {"workload": "%d^3 jittered hex (n=%d), lognormal k sigma=2.0 in 8^3 blocks; network trained on "
                               "%d residual/error pairs of plain AMG (after %s cycles) on %d matrices at 32^3 (sigma 1.5 per cell), "
                               "energy-norm loss" % (N, S.n_self, n_mat * per_mat, args.train_cycles, n_mat)},
        "without": res["without"], "with": res["with"], "heldout": heldout, "stationary": stat,
        "gain": round(res["without"]["solve_s"] / res["with"]["solve_s"], 3),
        "train": {"pairs": n_mat * per_mat, "steps": args.train_steps, "lr": args.lr,
                  "loss_first": float(np.mean(curve[:10])), "loss_last": float(np.mean(curve[-10:])),
                  "generate_s": round(t_gen, 2), "train_s": round(t_train, 2)},
    }
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=200, help="cells per direction (200: the 8M-cell workload)")
    ap.add_argument("--skip-cpu", action="store_true")
    ap.add_argument("--skip-exact", action="store_true")
    ap.add_argument("--skip-plain", action="store_true")
    ap.add_argument("--profile", action="store_true",
                    help="short program for rocprofv3: a few cycles of the hot kernels and a calibration copy")
    ap.add_argument("--traffic", nargs=2, metavar=("FETCH_CSV", "WRITE_CSV"),
                    help="no GPU: write profiles/hbm_traffic.json from the two counter passes of --profile")
    ap.add_argument("--sell-block", type=int, default=256, help="--profile: smoother threads per block")
    ap.add_argument("--spmv-block", type=int, default=256, help="--profile: residual/SpMV threads per block")
    ap.add_argument("--tail-max-rows", type=int, default=40000, help="--profile: single-workgroup tail size")
    ap.add_argument("--sell-xcd", type=int, default=1, help="--profile: XCD-contiguous smoother blocks")
    ap.add_argument("--spmv-xcd", type=int, default=0, help="--profile: XCD-contiguous residual/SpMV blocks")
    ap.add_argument("--cycle-graph", type=int, default=1, help="--profile: replay levels >= 1 of the cycle from a graph")
    ap.add_argument("--nt-levels", type=int, default=2, help="--profile: finest levels read with non-temporal loads")
    ap.add_argument("--learned", action="store_true",
                    help="instead of the 8M-cell solve: train the learned correction and run its ablation")
    ap.add_argument("--train-matrices", type=int, default=20)
    ap.add_argument("--train-pairs", type=int, default=2000)
    ap.add_argument("--train-steps", type=int, default=600)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--stall", type=float, default=0.9)
    ap.add_argument("--stationary-max", type=int, default=150)
    ap.add_argument("--train-cycles", default="1,2",
                    help="plain AMG cycles before a training pair is taken (cycled through)")
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--kinetic", type=int, default=0,
                    help="instead of the 8M-cell solve: time outer iterations of the BGK kinetic model on an N^3 slab")
    ap.add_argument("--kinetic-q", type=int, default=10, help="--kinetic: velocities per axis")
    ap.add_argument("--batch", type=int, default=0,
                    help="instead of the 8M-cell solve: time a sweep of this many phonon direction systems")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path "
              "(rows would shard by mesh slabs with halo exchange, not built); run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    if args.traffic:
        write_traffic_profile(args.traffic[0], args.traffic[1], args.n,
                              os.path.join(ROOT, "profiles", "hbm_traffic.json"))
        return
    if args.kinetic:
        run_kinetic(args)
        return
    if args.batch:
        run_batch(args)
        return
    if args.profile:
        run_profile(args)
        return
    if args.learned:
        run_learned(args)
        return
    import torch
    import fvm_amd
    from fvm_amd._lib import lib

    N = args.n
    log("building %d^3 workload" % N)
    mesh, S, t_asm = build_workload(N, torch)
    n_total = S.n_total
    b = S.device_pointers()["b"]          # right-hand side stays in the system's device memory
    log("assembled: n=%d n_total=%d nnz=%d in %.2f s" % (S.n_self, n_total, S.nnz, t_asm))

    # ---------------- headline: mixed-precision AMG + FP64 FGMRES, setup every solve
    amg = make_amg(1, 1)
    fg = fvm_amd.FGMRES(amg)
    fg.restart = 30
    fg.relativeTolerance = 1e-8
    fg.nMaxIterations = 2000

    def step():
        x = torch.zeros(n_total, dtype=torch.float64, device="cuda:0")
        setup_on(amg, S)
        r0, rf, it = fg.solve_device(b, x)
        return x, r0, rf, it

    setup_on(amg, S)
    tune = amg.autotune(reps=2)                 # launch shapes only: results do not change
    log("autotune: %s" % json.dumps(tune["chosen"]))
    for w in range(args.warmup):
        x, r0, rf, it = step()
        log("warmup %d: iterations %d ratio %.3e" % (w, it, rf / r0))
    lib().fvm_amg_set_kernel_timing(amg._h, 1)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(args.steps):
        x, r0, rf, it = step()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    ms_per_step = 1e3 * (t1 - t0) / args.steps
    st = amg.setup_stats()
    levels = level_sizes(amg)
    n0, nnz0 = levels[0]
    kr = kernel_stats(amg, 0)
    ks = kernel_stats(amg, 1)
    lib().fvm_amg_set_kernel_timing(amg._h, 0)
    log("mixed: %.1f ms/solve (setup %.1f ms), %d FGMRES iterations, ratio %.3e"
        % (ms_per_step, 1e3 * st["setup_seconds"], it, rf / r0))
    x_mixed = x

    # ---------------- the reference's own Krylov loops (BCGStab.cpp, CG.cpp) with the same
    # mixed AMG as preconditioner: setup + solve to the same stop rule, once each
    krylov = {}
    for name, cls in (("bcgstab", fvm_amd.BCGStab), ("cg", fvm_amd.CG)):
        ks_ = cls(amg)
        ks_.relativeTolerance = 1e-8
        ks_.nMaxIterations = 500
        xk = torch.zeros(n_total, dtype=torch.float64, device="cuda:0")
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        setup_on(amg, S)
        r0k, rfk, itk = ks_.solve_device(b, xk)
        torch.cuda.synchronize()
        krylov[name] = {"s": round(time.perf_counter() - t0, 4), "iterations": itk, "final_ratio": rfk / r0k}
        log("%s + mixed AMG: %.1f ms (setup included), %d iterations, ratio %.3e"
            % (name, 1e3 * krylov[name]["s"], itk, rfk / r0k))

    # ---------------- plain FP64 GPU AMG (coloured GS), same system
    t_plain, cyc_plain, err, tune_plain = None, None, None, None
    kp_s = (0, 0.0, 0.0)
    if not args.skip_plain:
        plain = make_amg(1, 0)
        # its own launch shapes: the FP64 hierarchy is tuned like the mixed one (untimed)
        setup_on(plain, S)
        tune_plain = plain.autotune(reps=2)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        setup_on(plain, S)
        xp = torch.zeros(n_total, dtype=torch.float64, device="cuda:0")
        lib().fvm_amg_set_kernel_timing(plain._h, 1)
        r0p, rfp, cyc_plain = plain.solve_device(b, xp)
        torch.cuda.synchronize()
        t_plain = time.perf_counter() - t0
        kp_s = kernel_stats(plain, 1)
        lib().fvm_amg_set_kernel_timing(plain._h, 0)
        log("plain FP64 AMG: %.3f s (setup %.3f s), %d cycles, ratio %.3e"
            % (t_plain, plain.setup_stats()["setup_seconds"], cyc_plain, rfp / r0p))
        err = (x_mixed[: S.n_self] - xp[: S.n_self]).abs().max().item() / xp[: S.n_self].abs().max().item()
        del plain

    # ---------------- reference-order AMG: cycle count (the reference's own number)
    cyc_exact = None
    t_exact = None
    if not args.skip_exact:
        ex = make_amg(0, 0)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        setup_on(ex, S)
        xe = torch.zeros(n_total, dtype=torch.float64, device="cuda:0")
        r0e, rfe, cyc_exact = ex.solve_device(b, xe)
        torch.cuda.synchronize()
        t_exact = time.perf_counter() - t0
        log("exact-order FP64 AMG: %.3f s, %d cycles" % (t_exact, cyc_exact))
        del ex

    # ---------------- roofline of the finest-level kernels (algorithmic bytes counted by
    # the library per timed launch: per row the diagonal, the vectors and a 4-byte row
    # offset, per off-diagonal a 4-byte column and the value; a sweep direction counts
    # the rows of the colours it updates)
    kernels = {}
    if kr[0]:
        kernels["finest_fp64_spmv_residual"] = kernel_entry(
            kr, "residual 36n+12nnz, SpMV 28n+12nnz (FP64)")
    if ks[0]:
        kernels["finest_fp32_gs_sweep"] = kernel_entry(
            ks, "20 per updated row + 8 per its off-diagonal (FP32), per direction")
    if kp_s[0]:
        kernels["plain_finest_fp64_gs_sweep"] = kernel_entry(
            kp_s, "36 per updated row + 12 per its off-diagonal (FP64), per direction")
    # dominant kernel of the headline solve: the one with the larger total time
    dom = max([("finest_fp64_spmv_residual", kr[1]), ("finest_fp32_gs_sweep", ks[1])], key=lambda t: t[1])[0]
    kd = kernels[dom]
    traffic, traffic_note = profiled_traffic(dom, N)
    roofline = dict(bound="hbm", kernel=dom, achieved=round(kd["gbs"], 1), peak=HBM_PEAK_GBS, unit="GB/s",
                    frac=round(kd["gbs"] / HBM_PEAK_GBS, 4), traffic=traffic, traffic_note=traffic_note,
                    algorithmic_bytes=kd["bytes"], avg_us=round(kd["avg_us"], 2))

    cpu = None
    if not args.skip_cpu:
        log("cpu baseline (oracle, 1 thread)")
        cpu = cpu_baseline(S, cyc_exact if cyc_exact else (cyc_plain or 0), torch)
        log("cpu extrapolation check on 32^3: %s" % json.dumps(cpu["validation"]))
        log("cpu: %.1f s estimated time-to-1e-8" % cpu["value"])

    out = {
        "metric": METRIC, "value": round(ms_per_step / 1e3, 6), "unit": "s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": False, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64 outer residual/Krylov, f32 V-cycle, bf16 coarse off-diagonals",
        "data": "synthetic (seeded jittered mesh and fields generated on the GPU)",
        "config": {"workload": "%d^3 jittered hex (n=%d, nnz=%d), lognormal k sigma=2.5 in 8^3 blocks, "
                               "source U[0,1], Dirichlet x0/x1, Neumann y/z; FP64 FGMRES(30) + mixed AMG "
                               "V(0,1), setup included, to 1e-8 (1-norm, FP64 true residual)"
                               % (N, S.n_self, nnz0),
                   "n_cells": S.n_self, "levels": len(levels)},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "setup_ms": round(1e3 * st["setup_seconds"], 2),
        "finest_coarsening_rounds": st["finest_rounds"],
        "fgmres_iterations": it,
        "final_ratio": rf / r0,
        "reference_krylov_mixed_amg": krylov,
        "plain_fp64_gpu_amg_s": None if t_plain is None else round(t_plain, 4),
        "plain_fp64_cycles": cyc_plain,
        "speedup_vs_plain_fp64_gpu": None if t_plain is None else round(t_plain / (ms_per_step / 1e3), 3),
        "exact_order_gpu_amg_s": None if t_exact is None else round(t_exact, 4),
        "exact_order_cycles": cyc_exact,
        "rel_error_vs_fp64_amg": err,
        "assembly_s": round(t_asm, 4),
        "autotune": tune,
        "plain_fp64_autotune": None if tune_plain is None else tune_plain["chosen"],
        "kernels": kernels,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
