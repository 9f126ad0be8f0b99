"""Benchmark of the stabilized-FEM formation + memLS path on one MI355X.

One step = one FP64 interior formation (flux reconstruction, e3 over all
elements, assembly straight into the memLS matrix and Res4) on the cylinder
built from a 64x64x320 hex grid clipped to r <= 1 cm and split into Kuhn tets,
with inputs resident in HBM.  value = elements formed per second (Melem/s).

Also reported on the same mesh: the Krylov SpMV bandwidth of the scaled 3x3
momentum block and of the full 4x4 matrix, and the wall time of FP64
timesteps (3 nonlinear iterations of formation + NS solve + update each,
Kspace 100, memLS NS tolerance 1e-8, resistance outlet 1333 dyn s/cm^5).

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Melem/s formed + Krylov SpMV GB/s; mixed vs FP64 GPU timestep time at 1e-8"
FP64_PEAK_TFLOPS = 78.6
HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def build_problem(grid, cfa, M):
    import numpy as np
    t0 = time.time()
    m = M.cylinder_mesh(*grid)
    colm, rowp = cfa.genadj(m.ien, m.numnp)
    y, ac = M.poiseuille_state(m)
    ibc, bc = M.dirichlet_codes(m)
    log(f"mesh {grid}: nel={m.nel} nshg={m.numnp} nnz={rowp.size} ({time.time() - t0:.1f}s)")
    return m, colm, rowp, y, ac, ibc, bc


# flops per element of the reference formulation as written (e3 per
# quadrature point, the flux pass, the scatter and fillsparseI), counted by
# the oracle's counting build (crimson_oracle.c: cfo_flops_per_element;
# pinned by the oracle tests), and per node (qpbc)
FLOPS_PER_ELEMENT = 8996.0
FLOPS_PER_NODE = 10.0
# the formation's kernels as rocprofv3 names them (profiles/*_pmc_traffic.md)
FORMATION_KERNELS = ("node_pack", "flux_node", "e3_elem<true>", "asm_row<1,16>", "bnd_elem", "bnd_node<1>")


def pmc_kernel_name(name):
    """A rocprofv3 kernel name without its return type, namespace and
    arguments, spaces removed (``asm_row<1,16>``)."""
    n = name[5:] if name.startswith("void ") else name
    n = n.replace("(anonymous namespace)::", "")
    depth = 0
    for i, ch in enumerate(n):
        if ch == "<":
            depth += 1
        elif ch == ">":
            depth -= 1
        elif ch == "(" and depth == 0 and i > 0:
            n = n[:i]
            break
    return n.replace(" ", "")


def measure_traffic(grid, steps=5, timeout=180):
    """HBM bytes per dispatch of the formation bench's kernels, measured now:
    two child runs of this script's formation alone on the same workload
    under ``rocprofv3 --pmc`` (FETCH_SIZE, then WRITE_SIZE), each run starting
    with a plain 1 GiB device copy whose counter reading gives the bytes per
    counter unit.  Returns {kernel: bytes per dispatch} averaged over the
    last ``steps`` dispatches (the timed steps), or None when rocprofv3 is
    missing, this process already runs under a profiler, or a run leaves no
    counters."""
    import csv
    import shutil
    import subprocess
    import tempfile
    from collections import defaultdict
    exe = shutil.which("rocprofv3") or "/opt/rocm/bin/rocprofv3"
    if not os.path.exists(exe):
        return None
    # not from a process that rocprofv3 already instruments (it exports its
    # ROCPROF* settings to the application): the child profiler would
    # inherit the tool
    if any(k.startswith("ROCPROF") for k in os.environ):
        return None
    child = [sys.executable, "-u", os.path.abspath(__file__), "--grid", ",".join(str(g) for g in grid),
             "--steps", str(steps), "--warmup", "1", "--timesteps", "0", "--cpu-seconds", "0", "--mixed", "0",
             "--ensemble", "0", "--adjoint", "0", "--config1", "0", "--stretched", "0", "--ensemble-timestep", "0",
             "--reference-case", "0", "--traffic", "0", "--copy-mb", "1024"]
    env = dict(os.environ, TMPDIR="/tmp")
    both = defaultdict(lambda: [0.0, 0.0])
    for side, counter in enumerate(("FETCH_SIZE", "WRITE_SIZE")):
        d = tempfile.mkdtemp(prefix="crimson_pmc_", dir="/tmp")
        try:
            subprocess.run([exe, "--pmc", counter, "--output-format", "csv", "-d", d, "-o", "pmc", "--"] + child,
                           cwd="/tmp", env=env, stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL,
                           timeout=timeout)
            path = os.path.join(d, "pmc_counter_collection.csv")
            if not os.path.exists(path):
                return None
            per = defaultdict(lambda: defaultdict(float))  # kernel -> dispatch -> counter
            with open(path) as fh:
                for row in csv.DictReader(fh):
                    if row["Counter_Name"] == counter:
                        per[pmc_kernel_name(row["Kernel_Name"])][int(row["Dispatch_Id"])] += float(row["Counter_Value"])
            copies = per.get("__amd_rocclr_copyBuffer", {})
            cal = max(copies.values()) if copies else 0.0
            if cal <= 0.0:
                return None
            scale = float(1 << 30) / cal  # bytes per counter unit
            for k, dv in per.items():
                vals = [dv[i] for i in sorted(dv)][-steps:]
                both[k][side] = sum(vals) / len(vals) * scale
        except (OSError, subprocess.SubprocessError, ValueError, KeyError):
            return None
        finally:
            shutil.rmtree(d, ignore_errors=True)
    return {k: v[0] + v[1] for k, v in both.items()}


def pmc_traffic(grid, kernel):
    """HBM bytes per dispatch of ``kernel`` from the committed rocprofv3 PMC
    summary of this workload (FETCH_SIZE and WRITE_SIZE, calibrated against a
    plain device copy), or None when no summary exists for it: the fallback
    when the live measurement (measure_traffic) is off or unavailable."""
    path = os.path.join(ROOT, "profiles", f"formation_{grid[0]}x{grid[1]}x{grid[2]}_pmc_traffic.md")
    if not os.path.exists(path):
        return None
    for line in open(path):
        parts = line.split()
        if len(parts) == 3 and parts[0] == "traffic_bytes_per_dispatch" and parts[1] == kernel.replace(" ", ""):
            return float(parts[2])
    return None


def inflow_mean(t, seed=7, period=1.0):
    """Mean inflow velocity: 10 cm/s plus 3 seeded harmonics (+/- 30 %)."""
    from crimsonflowsolver_amd.mesh import uniform_pm1
    r = uniform_pm1(seed, 6)
    u = 10.0
    for k in range(3):
        u += 3.0 * 0.5 * (1 + r[k]) / (k + 1) * math.sin(2 * math.pi * (k + 1) * t / period + math.pi * r[3 + k])
    return u


def cpu_threads():
    """Threads of the all-core CPU baseline: OMP_NUM_THREADS when set (the
    host's CPU share), else the CPUs this process may run on."""
    v = os.environ.get("OMP_NUM_THREADS", "")
    return int(v) if v.isdigit() and int(v) > 0 else len(os.sched_getaffinity(0))


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(sample_grid, target_s=15.0, threads=1):
    """Oracle (CPU restatement) interior formation (flux pass + e3 + bc3lhs +
    fillsparseI, the two-phase form of the block loop with identical bits) on
    `threads` threads of the threaded build, on a bounded slab of the same
    cross-section; returns Melem/s and the sample's size and timing."""
    from crimsonflowsolver_amd import mesh as M
    from oracle import oracle as O
    m = M.cylinder_mesh(*sample_grid)
    colm, rowp = O.genadj(m.ien, m.numnp)
    y, ac = M.poiseuille_state(m)
    p = O.make_params()
    O.lib_mt(threads)
    plan = O.Plan(m.ien, m.numnp, mt=True)

    def once():
        plan.interior(y, ac, m.x, colm, rowp, p)

    once()  # warm-up
    n, t0 = 0, time.perf_counter()
    while True:
        once()
        n += 1
        if time.perf_counter() - t0 >= target_s:
            break
    dt = time.perf_counter() - t0
    return m.nel * n / dt / 1e6, m.nel, n, dt


def cpu_timestep(grid, R=1333.0, threads=1):
    """One timestep (3 nonlinear iterations: yAlpha, formation with the
    boundary elements, bc3Res, ElmpvsQ's resistance term, Val4 packing, memLS
    NS solve, itrCorrect, itrBC) of the oracle on `grid`: the interior
    formation, Val4 packing and memLS on `threads` threads of the threaded
    build, the boundary elements and the numpy step glue on one; the same
    steps and solver settings as the GPU timestep.
    Returns (seconds, solver iterations per solve)."""
    from crimsonflowsolver_amd import mesh as M
    from oracle import oracle as O
    import numpy as np
    m = M.cylinder_mesh(*grid)
    N = m.numnp
    colm, rowp = O.genadj(m.ien, N)
    y0, ac0 = M.poiseuille_state(m)
    ibc, _ = M.dirichlet_codes(m)
    ienb, iBCB, BCB = M.cylinder_boundary(m)
    nabi = m.nabi()
    out = np.nonzero(m.outlet & (ibc == 0))[0]
    p = O.make_params()
    O.lib_mt(threads)
    plan = O.Plan(m.ien, N, mt=True)
    faceres = np.array([0.0, R * p.alfi * p.gami * p.Delt])
    ls = O.MemLS(colm, rowp, 2, relTol=1e-8, dimKry=100, maxItr=10, relTolIn=[0.01, 0.2], maxItrIn=[1, 200],
                 mt=True)
    fixed = np.nonzero(ibc)[0]
    ls.bc_create(1, fixed + 1, 3, O.BC_TYPE_Dir, np.zeros((3, fixed.size), order="F"))
    ls.bc_create(2, out + 1, 3, O.BC_TYPE_Neu, np.asfortranarray(nabi[out].T))
    rr = m.x[:, 0] ** 2 + m.x[:, 1] ** 2
    inl = m.inlet & ~m.wall
    bc = np.zeros((N, 7), order="F")
    bc[inl, 4] = 2.0 * inflow_mean(p.Delt) * np.clip(1.0 - rr[inl], 0.0, None)
    its = []
    t0 = time.perf_counter()
    y, ac = O.itr_predict(y0, ac0, p.gami)
    y = O.itr_bc(ibc, bc, y)
    for _ in range(3):
        ya, aca = O.itr_yalpha(y0, ac0, y, ac, p.alfi, p.almi)
        res, K, P = plan.interior(ya, aca, m.x, colm, rowp, p)
        res = O.boundary_res(ienb, iBCB, BCB, ya, m.x, res)
        res = O.bc3res(ibc, res)
        res, _ = O.elmpvsq_resistance(res, ya, out, nabi[out], R)
        sol = ls.solve(np.asfortranarray(res.T), O.pack_val4(colm, rowp, K, P, mt=True), res=faceres)
        its.append((ls.stats(0)["itr"], ls.stats(1)["itr"], ls.stats(2)["itr"]))
        y, ac = O.itr_correct(y, ac, sol, p.gami, p.alfi, p.Delt)
        y = O.itr_bc(ibc, bc, y)
    return time.perf_counter() - t0, its


def gpu_timestep(cfa, M, torch, grid, R=1333.0, reps=2):
    """The GPU timestep of the main bench on the slab `grid` (the same steps
    as cpu_timestep, through the device entry points); returns the mean
    milliseconds of `reps` timesteps from the same initial state and the
    solver iterations of the last."""
    import numpy as np
    dev = torch.device("cuda:0")
    m = M.cylinder_mesh(*grid)
    N = m.numnp
    colm, rowp = cfa.genadj(m.ien, N)
    y0, ac0 = M.poiseuille_state(m)
    ibc, _ = M.dirichlet_codes(m)
    prm = cfa.form_params()
    form = cfa.ElmGMRInterior(m.ien, colm, rowp, N, iBC=ibc, params=prm)
    form.set_boundary(*M.cylinder_boundary(m))
    nabi = m.nabi()
    out = np.nonzero(m.outlet & (ibc == 0))[0]
    fixed = np.nonzero(ibc)[0]
    lhs = cfa.memLS_LHS_CREATE(N, N, rowp.size, None, colm, rowp, 2)
    cfa.memLS_BC_CREATE(lhs, 1, fixed.size, 3, cfa.BC_TYPE_Dir, fixed + 1, np.zeros((3, fixed.size), order="F"))
    cfa.memLS_BC_CREATE(lhs, 2, out.size, 3, cfa.BC_TYPE_Neu, out + 1, np.asfortranarray(nabi[out].T))
    T = lambda a: torch.from_numpy(np.ascontiguousarray(a.T)).to(dev)
    dx = T(m.x)
    faceres = np.array([0.0, R * prm.alfi * prm.gami * prm.Delt])
    rr = m.x[:, 0] ** 2 + m.x[:, 1] ** 2
    inl = m.inlet & ~m.wall
    bc = np.zeros((N, 7), order="F")
    bc[inl, 4] = 2.0 * inflow_mean(prm.Delt) * np.clip(1.0 - rr[inl], 0.0, None)
    dBC, dibc = T(bc), torch.from_numpy(ibc.astype(np.int32)).to(dev)
    dglob = torch.from_numpy(out.astype(np.int32)).to(dev)
    dnabi = torch.from_numpy(np.ascontiguousarray(nabi[out])).to(dev)
    dq = torch.zeros(1, dtype=torch.float64, device=dev)
    dres4 = torch.zeros((N, 4), dtype=torch.float64, device=dev)
    yold, acold = T(y0), T(ac0)
    times = []
    for _ in range(reps):
        ls = cfa.memLS_LS_CREATE(cfa.LS_TYPE_NS, relTol=1e-8, dimKry=100, maxItr=10,
                                 relTolIn=[0.01, 0.2], maxItrIn=[1, 200])
        yc, acc = torch.empty_like(yold), torch.empty_like(acold)
        ya, aca = torch.empty_like(yold), torch.empty_like(acold)
        its = []
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        form.predict(yold, acold, yc, acc)
        form.itr_bc(dibc, dBC, yc)
        for _ in range(3):
            form.yalpha(yold, acold, yc, acc, ya, aca)
            form.sync()
            form.to_lhs(ya, aca, dx, dres4, lhs)
            form.bc3res(dres4, res4_layout=True)
            form.elmpvsq(dglob, dnabi, R, ya, dres4, dq)
            form.sync()
            cfa.memLS_SOLVE_dev(lhs, ls, 4, dres4, res=faceres)
            its.append((ls.RI["itr"], ls.GM["itr"], ls.CG["itr"]))
            form.correct(yc, acc, dres4)
            form.itr_bc(dibc, dBC, yc)
            form.sync()
        torch.cuda.synchronize()
        times.append(time.perf_counter() - t0)
    return float(np.mean(times) * 1e3), its


def ensemble_timestep(cfa, torch, ens, meshes, states, R):
    """One timestep (3 nonlinear iterations) of every instance of the union,
    run entry point by entry point and as one captured graph
    (crimson_step_graph_create: formation, bc3Res, each instance's outlet
    term, the device-driven NS solve, update); both from the same state.
    Returns the times, the captured graph's launch count and whether the
    two agree bit for bit."""
    import numpy as np
    from crimsonflowsolver_amd.timestep import ns_solver
    dev = torch.device("cuda:0")
    T = lambda a: torch.from_numpy(np.ascontiguousarray(np.asarray(a, order="F").T)).to(dev)
    p = ens.params
    y0 = T(np.concatenate([a[0] for a in states]))
    ac0 = T(np.concatenate([a[1] for a in states]))
    x = T(ens.x)
    bch = np.zeros((ens.N, 7), order="F")
    for k, mm in enumerate(meshes):
        rr = mm.x[:, 0] ** 2 + mm.x[:, 1] ** 2
        inl = mm.inlet & ~mm.wall
        col = np.zeros(mm.numnp)
        col[inl] = 2.0 * inflow_mean(p.Delt) * np.clip(1.0 - rr[inl] / mm.radius ** 2, 0.0, None)
        bch[ens.rows(k), 4] = col
    bc = T(bch)
    ibc = torch.from_numpy(ens.iBC).to(dev)
    surfaces = [(torch.from_numpy(ens.outlets[k].astype(np.int32)).to(dev),
                 torch.from_numpy(np.ascontiguousarray(ens.nabi[k])).to(dev), float(R[k]), 0.0)
                for k in range(ens.K)]
    faceres = ens.face_res()
    f = ens.form
    # entry point by entry point
    ls_e = ns_solver()
    y, ac = torch.empty_like(y0), torch.empty_like(ac0)
    ya, aca = torch.empty_like(y0), torch.empty_like(ac0)
    r4 = torch.zeros((ens.N, 4), dtype=torch.float64, device=dev)
    q = torch.zeros(1, dtype=torch.float64, device=dev)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    f.predict(y0, ac0, y, ac)
    f.itr_bc(ibc, bc, y)
    for _ in range(3):
        f.yalpha(y0, ac0, y, ac, ya, aca)
        f.sync()
        f.to_lhs(ya, aca, x, r4, ens.lhs)
        f.bc3res(r4, res4_layout=True)
        for glob, nabi, c1, c2 in surfaces:
            f.elmpvsq_coef(glob, nabi, c1, c2, ya, r4, q)
        f.sync()
        cfa.memLS_SOLVE_dev(ens.lhs, ls_e, 4, r4, res=faceres)
        f.correct(y, ac, r4)
        f.itr_bc(ibc, bc, y)
        f.sync()
    torch.cuda.synchronize()
    ms_eager = (time.perf_counter() - t0) * 1e3
    its_e = [ls_e.RI["itr"], ls_e.GM["itr"], ls_e.CG["itr"]]
    # one captured graph
    ls_g = ns_solver()
    gy0, gac0 = y0.clone(), ac0.clone()
    gy, gac, gya, gaca = (torch.empty_like(y0) for _ in range(4))
    gr4 = torch.zeros((ens.N, 4), dtype=torch.float64, device=dev)
    t0 = time.perf_counter()
    g = f.step_graph(ens.lhs, ls_g, gy0, gac0, gy, gac, gya, gaca, x, gr4, ibc, bc, surfaces=surfaces,
                     faceres=faceres, nonlinear=3)
    torch.cuda.synchronize()
    s_capture = time.perf_counter() - t0
    t0 = time.perf_counter()
    g.launch()
    torch.cuda.synchronize()
    ms_graph = (time.perf_counter() - t0) * 1e3
    same = bool(torch.equal(gy, y) and torch.equal(gac, ac))
    its_g = [ls_g.RI["itr"], ls_g.GM["itr"], ls_g.CG["itr"]]
    out = {"nonlinear_iterations": 3, "eager_ms": ms_eager, "graph_ms": ms_graph,
           "graph_capture_s": s_capture, "graph_equals_eager_bitwise": same,
           "last_solve_iterations_eager": its_e, "last_solve_iterations_graph": its_g}
    log(f"ensemble timestep (3 nonlinear iterations, {ens.K} instances): eager {ms_eager:.0f} ms, one graph launch "
        f"{ms_graph:.0f} ms (capture {s_capture:.1f} s), bitwise equal {same}")
    del g
    return out


def implicit_adjoint(cfa, M, torch, m, nsteps, every, R=1333.0):
    """The adjoint workload: the gradient of J = <cy, y_N> + <ca, ac_N>
    (seeded random cotangents) with respect to the initial state, mu and the
    outlet resistance, through nsteps generalized-alpha timesteps of 3
    nonlinear iterations each (the 1.3M-node timestep's solver settings and
    inflow), checkpointed every `every` steps with recomputation in between
    (adjoint.StepAdjoint: transpose solves by FGMRES on the VJP kernels,
    preconditioned by memLS's NS solver on the transposed tangent).  Returns
    the timings and the solver statistics."""
    from crimsonflowsolver_amd.adjoint import StepAdjoint
    from crimsonflowsolver_amd.timestep import ResistanceCylinder
    case = ResistanceCylinder(m, R=R, inflow=inflow_mean, transpose=True)
    adj = StepAdjoint(case, tol=1e-8)
    y0, ac0 = case.state()
    gen = torch.Generator(device="cuda").manual_seed(13)
    cy = torch.randn(y0.shape, dtype=torch.float64, device="cuda", generator=gen)
    ca = torch.randn(y0.shape, dtype=torch.float64, device="cuda", generator=gen)
    ca[3] = 0.0
    # forward alone (the recurrence whose gradient is taken), then the
    # checkpointed gradient (its forward pass, recomputation and reverse sweep)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    y, ac = y0, ac0
    for n in range(nsteps):
        y, ac, _ = case.step(y, ac, (n + 1) * case.params.Delt, iters=3)
        if (n + 1) % 10 == 0:
            log(f"implicit adjoint: forward step {n + 1} of {nsteps}")
    torch.cuda.synchronize()
    t_fwd = time.perf_counter() - t0
    t0 = time.perf_counter()
    J, yb, ab, mub, rb = adj.gradient(y0, ac0, 0.0, nsteps, every, cy, ca, step_kw=dict(iters=3),
                                      progress=lambda ph, n: log(f"implicit adjoint gradient: {ph} pass "
                                                                 f"{'done' if ph == 'forward' else f'back to step {n}'}"))
    torch.cuda.synchronize()
    t_grad = time.perf_counter() - t0
    its = [s[0] for s in adj.stats]
    out = {"steps": nsteps, "checkpoint_every": every, "nonlinear_iterations": 3,
           "forward_ms_per_step": t_fwd / nsteps * 1e3, "gradient_s": t_grad,
           "gradient_ms_per_step": t_grad / nsteps * 1e3,
           "transpose_solve_iterations": [int(v) for v in its],
           "transpose_solve_worst_rel_residual": max(s[1] for s in adj.stats),
           "J": J, "dJ_dmu": mub, "dJ_dR": rb, "dJ_dy0_norm": float(torch.linalg.vector_norm(yb))}
    log(f"implicit adjoint, {nsteps} steps (checkpoints every {every}): forward {t_fwd:.1f} s, gradient "
        f"{t_grad:.1f} s, transpose solves {min(its)}-{max(its)} FGMRES iterations, dJ/dmu {mub:.4e}, "
        f"dJ/dR {rb:.4e}")
    del case, adj
    return out


def reference_case(cfa, np):
    """The reference's basic case on the device (phasta_case.PhastaCase): 5
    steps from its restart.0.1; returns the outlet flow after each step,
    the reference's values and the wall time per step."""
    from crimsonflowsolver_amd.phasta_case import PhastaCase, wall_from_solver_inp
    g = os.path.join(ROOT, "tests", "golden")
    z = np.load(os.path.join(g, "basic_case.npz"))
    b = np.load(os.path.join(g, "basic_bct.npz"))
    o = np.load(os.path.join(g, "basic_outputs.npz"))
    recs = [(b["x"][k], np.concatenate([b["u"][k], b["t"][:, None]], axis=1)) for k in range(b["x"].shape[0])]
    c = PhastaCase(z["x"], z["ien"], z["ienb"], z["iBCB"], z["BCB"], z["iBC"], z["y"], z["ac"], z["u"], z["uref"],
                   z["nodetag"], z["ndsurf"], recs, {3: (o["rcr"][0], o["rcr"][1], o["rcr"][2], o["pdist"])},
                   rho=0.00106, mu=0.004, Delt=0.001, rhoinf=0.0, wall=wall_from_solver_inp(), dirichlet_srfs=(2,))
    qs, ts, its = [], [], []
    for _ in range(5):
        t0 = time.perf_counter()
        r = c.step()
        ts.append((time.perf_counter() - t0) * 1e3)
        qs.append(r[3]["Q"])
        its.append([list(map(float, v[1:])) for v in r["iterations"]])
    out = {"case": "testbin/mainTests/basic (1072 nodes, 4157 tets, deformable wall, bct.dat, RCR outlet)",
           "outlet_flow": qs, "reference_outlet_flow": [float(v) for v in o["qhist"][:, 1]],
           "step5_abs_diff": abs(qs[-1] - 714.921079082528), "reference_test_tolerance": 1e-7,
           "ms_per_step": ts, "solver_iterations": its}
    log(f"reference basic case: outlet flow after step 5 {qs[-1]!r} (reference 714.921079082528), "
        f"{np.mean(ts[1:]):.1f} ms per step")
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--grid", default="64,64,320")
    ap.add_argument("--timesteps", type=int, default=10,
                    help="timesteps of 3 nonlinear iterations to time, FP64 and mixed (0: skip)")
    ap.add_argument("--max-nonlinear", type=int, default=60,
                    help="nonlinear iterations allowed to reach the 1e-12 residual in the converged timestep")
    ap.add_argument("--cpu-seconds", type=float, default=15.0, help="CPU baseline sample length (0: skip)")
    ap.add_argument("--spmv-reps", type=int, default=50)
    ap.add_argument("--mixed", type=int, default=1, help="also time the mixed-precision timesteps (0: skip)")
    ap.add_argument("--slab", default="64,64,20",
                    help="slab on which one timestep is timed on the GPU and with the oracle on the CPU threads "
                         "(empty: skip; needs --cpu-seconds > 0)")
    ap.add_argument("--cpu-full-timestep", type=int, default=0,
                    help="also time the first timestep of the full mesh on the CPU threads (minutes)")
    ap.add_argument("--boundary", type=int, default=1,
                    help="form the boundary elements too (inlet code 0, wall 16, outlet 2), as ElmGMR does")
    ap.add_argument("--ensemble", type=int, default=48,
                    help="instances of the batched-ensemble measurement (on 8 uneven cylinders with boundary "
                         "elements, one union graph and one union NS solve; 0: skip)")
    ap.add_argument("--stretched", type=int, default=1,
                    help="also time the FP64 and mixed timesteps on the cylinder with 5% of its elements stretched "
                         "up to 50x (FP64 tau island: aspect ratio > 20)")
    ap.add_argument("--adjoint", type=int, default=1,
                    help="also time the VJP of the residual and check it against central differences")
    ap.add_argument("--ensemble-timestep", type=int, default=1,
                    help="also time one ensemble timestep entry point by entry point and as one graph (0: skip)")
    ap.add_argument("--adjoint-steps", type=int, default=10,
                    help="timesteps of the implicit-step adjoint (0: skip; the workload's figure is 50)")
    ap.add_argument("--adjoint-every", type=int, default=5, help="its checkpoint interval (the workload's: 10)")
    ap.add_argument("--reference-case", type=int, default=1,
                    help="run the reference's basic case (5 steps) against its own output (0: skip)")
    ap.add_argument("--config1", type=int, default=1,
                    help="also run the 20x20x100 case under both solver readings (GMRESV 4x4 and NS)")
    ap.add_argument("--autotune", type=int, default=1,
                    help="time each formation launch configuration (row schedule x assembly chunk) first "
                         "and keep the fastest")
    ap.add_argument("--copy-mb", type=int, default=0,
                    help="run one plain device copy of this many MB first (calibrates PMC byte counters)")
    ap.add_argument("--to-1e8", dest="to_1e8", type=int, default=1,
                    help="also drive the first timestep of each form to 1e-8 (0: skip; the fixed-iteration "
                         "timesteps still run)")
    ap.add_argument("--traffic", type=int, default=1,
                    help="measure roofline.traffic at the end with two rocprofv3 PMC child runs (0: read the "
                         "committed profile)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: the formation/solve path here has no multi-GPU form "
              "(one GPU holds the whole mesh); run it with --gpus 1.", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch
    import crimsonflowsolver_amd as cfa
    from crimsonflowsolver_amd import mesh as M

    if cfa.device_count() < 1 or not torch.cuda.is_available():
        raise SystemExit("bench.py: no GPU visible")
    dev = torch.device("cuda:0")
    grid = tuple(int(v) for v in args.grid.split(","))
    m, colm, rowp, y, ac, ibc, bc = build_problem(grid, cfa, M)
    N, nel, nnz = m.numnp, m.nel, rowp.size
    prm = cfa.form_params()
    t0 = time.time()
    form = cfa.ElmGMRInterior(m.ien, colm, rowp, N, iBC=ibc, params=prm)
    log(f"formation setup {time.time() - t0:.1f}s")

    # faces: 1 = Dirichlet (inlet + wall, all velocity components),
    # 2 = resistance outlet (Neumann, val = NABI)
    fixed = np.nonzero(ibc)[0]
    out = np.nonzero(m.outlet & (ibc == 0))[0]
    nabi = m.nabi()
    lhs = cfa.memLS_LHS_CREATE(N, N, nnz, None, colm, rowp, 2)
    cfa.memLS_BC_CREATE(lhs, 1, fixed.size, 3, cfa.BC_TYPE_Dir, fixed + 1, np.zeros((3, fixed.size), order="F"))
    cfa.memLS_BC_CREATE(lhs, 2, out.size, 3, cfa.BC_TYPE_Neu, out + 1, np.asfortranarray(nabi[out].T))

    form32 = lhs32 = None
    if args.mixed:
        form32 = cfa.ElmGMRInterior(m.ien, colm, rowp, N, iBC=ibc,
                                    params=cfa.form_params(precision=32, aspect_fp64=20.0))
        lhs32 = cfa.memLS_LHS_CREATE(N, N, nnz, None, colm, rowp, 2)
        cfa.memLS_BC_CREATE(lhs32, 1, fixed.size, 3, cfa.BC_TYPE_Dir, fixed + 1, np.zeros((3, fixed.size), order="F"))
        cfa.memLS_BC_CREATE(lhs32, 2, out.size, 3, cfa.BC_TYPE_Neu, out + 1, np.asfortranarray(nabi[out].T))

    nelb = 0
    if args.boundary:
        t0 = time.time()
        ienb, iBCB, BCB = M.cylinder_boundary(m)
        nelb = ienb.shape[0]
        for fm in (form, form32):
            if fm is not None:
                fm.set_boundary(ienb, iBCB, BCB)
        log(f"boundary elements: {nelb} ({time.time() - t0:.1f}s)")

    T = lambda a: torch.from_numpy(np.ascontiguousarray(a.T)).to(dev)  # (n,k) Fortran -> (k,n) C
    dy, dac, dx = T(y), T(ac), T(m.x)
    dres4 = torch.zeros((N, 4), dtype=torch.float64, device=dev)       # Res4(4,N) Fortran

    if args.copy_mb > 0:
        n = args.copy_mb * (1 << 20) // 8
        src = torch.ones(n, dtype=torch.float64, device=dev)
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        log(f"calibration copy: {args.copy_mb} MiB read + {args.copy_mb} MiB written")
        del src, dst

    # ---- launch configurations (no result depends on them)
    tuned = None
    if args.autotune:
        tuned = {"fp64": form.autotune(dy, dac, dx, dres4, lhs, reps=3)}
        if args.mixed:
            tuned["fp32"] = form32.autotune(dy, dac, dx, dres4, lhs32, reps=3)
        log(f"autotune: {tuned}")

    # ---- formation throughput
    for _ in range(args.warmup):
        form.to_lhs(dy, dac, dx, dres4, lhs)
    torch.cuda.synchronize()
    tm = np.zeros(4)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        form.to_lhs(dy, dac, dx, dres4, lhs)
        tt = form.timing()
        tm += [tt["flux_ms"], tt["e3_ms"], tt["assembly_ms"], tt["total_ms"]]
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    tm /= args.steps
    ms_step = elapsed / args.steps * 1e3
    melems = nel / (ms_step * 1e-3) / 1e6
    log(f"formation: {ms_step:.3f} ms/step wall, events flux {tm[0]:.3f} e3 {tm[1]:.3f} asm {tm[2]:.3f} "
        f"total {tm[3]:.3f} ms -> {melems:.1f} Melem/s")

    mixed_form = None
    if args.mixed:
        for _ in range(args.warmup):
            form32.to_lhs(dy, dac, dx, dres4, lhs32)
        torch.cuda.synchronize()
        tm32 = np.zeros(4)
        t0 = time.perf_counter()
        for _ in range(args.steps):
            form32.to_lhs(dy, dac, dx, dres4, lhs32)
            tt = form32.timing()
            tm32 += [tt["flux_ms"], tt["e3_ms"], tt["assembly_ms"], tt["total_ms"]]
        torch.cuda.synchronize()
        ms32 = (time.perf_counter() - t0) / args.steps * 1e3
        tm32 /= args.steps
        mixed_form = {"melem_s": nel / (ms32 * 1e-3) / 1e6, "ms_per_step": ms32,
                      "phase_ms": {"flux": tm32[0], "e3_residual_and_fp32_tangent": tm32[1],
                                   "res_rows_and_fp32_assembly": tm32[2]}}
        log(f"mixed formation: {ms32:.3f} ms/step -> {mixed_form['melem_s']:.1f} Melem/s, phases {tm32}")

    # ---- Krylov SpMV (scaled, as inside the solver)
    x3 = torch.rand((N, 3), dtype=torch.float64, device=dev)
    y3 = torch.empty_like(x3)
    x4 = torch.rand((N, 4), dtype=torch.float64, device=dev)
    y4 = torch.empty_like(x4)
    ms3 = lhs.spmv_dev(3, x3, y3, max(1, args.spmv_reps))
    ms4 = lhs.spmv_dev(4, x4, y4, max(1, args.spmv_reps))
    # memLS reads each block's column as a 2-byte offset from its row when
    # every offset fits 16 bits, else as a 4-byte index
    rows = np.repeat(np.arange(N), np.diff(colm))
    cb = 2 if np.abs((np.asarray(rowp) - 1) - rows).max() <= 32767 else 4
    b3 = nnz * (72 + cb) + N * (8 + 24 + 24)
    b4 = nnz * (128 + cb) + N * (8 + 32 + 32)
    gbs3, gbs4 = b3 / (ms3 * 1e-3) / 1e9, b4 / (ms4 * 1e-3) / 1e9
    log(f"spmv 3x3: {ms3:.4f} ms {gbs3:.0f} GB/s ; 4x4: {ms4:.4f} ms {gbs4:.0f} GB/s")
    spmv32 = None
    if args.mixed:
        m3 = lhs32.spmv_dev(3, x3, y3, max(1, args.spmv_reps))
        m4 = lhs32.spmv_dev(4, x4, y4, max(1, args.spmv_reps))
        c3 = nnz * (36 + cb) + N * (4 + 12 + 12)
        c4 = nnz * (64 + cb) + N * (4 + 16 + 16)
        spmv32 = {"bsr3x3": c3 / (m3 * 1e-3) / 1e9, "bsr4x4": c4 / (m4 * 1e-3) / 1e9, "ms3": m3, "ms4": m4}
        log(f"fp32 spmv 3x3: {m3:.4f} ms {spmv32['bsr3x3']:.0f} GB/s ; 4x4: {m4:.4f} ms {spmv32['bsr4x4']:.0f} GB/s")

    # ---- ensemble (the batched-ensemble workload): K instances on 8 cylinders
    # of uneven size (lengths 4-20 cm, 40k-600k tets), viscosity
    # lognormal(ln 0.04, 0.2), outlet resistance U(800, 2000); formed as one
    # union (autotuned, replayed from a HIP graph) and solved by one NS solve
    ens_info = None
    if args.ensemble > 0:
        from crimsonflowsolver_amd.ensemble import Ensemble, uneven_cylinders
        K = args.ensemble
        meshes, shapes, mu_e, R_e = uneven_cylinders(K)
        ens = Ensemble(meshes, mu_e, R_e, boundary=True)
        st = [M.poiseuille_state(mm, seed=60 + k) for k, mm in enumerate(meshes)]
        ey = T(np.concatenate([a[0] for a in st]))
        eac = T(np.concatenate([a[1] for a in st]))
        ex = T(ens.x)
        er4 = torch.zeros((ens.N, 4), dtype=torch.float64, device=dev)
        g = ens.capture(ey, eac, ex, er4)
        enel = int(ens.elem_off[-1])
        reps = max(5, args.steps)
        for _ in range(2):
            g.launch()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            g.launch()
        torch.cuda.synchronize()
        ms_graph = (time.perf_counter() - t0) / reps * 1e3
        t0 = time.perf_counter()
        for _ in range(reps):
            ens.form.to_lhs(ey, eac, ex, er4, ens.lhs)
        torch.cuda.synchronize()
        ms_eager = (time.perf_counter() - t0) / reps * 1e3
        # one NS solve of the union (the residual of the formation above)
        g.launch()
        ens.form.bc3res(er4, res4_layout=True)
        ls_e = cfa.memLS_LS_CREATE(cfa.LS_TYPE_NS, relTol=1e-8, dimKry=100, maxItr=10,
                                   relTolIn=[0.01, 0.2], maxItrIn=[1, 200])
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        cfa.memLS_SOLVE_dev(ens.lhs, ls_e, 4, er4, res=ens.face_res())
        torch.cuda.synchronize()
        ms_solve = (time.perf_counter() - t0) * 1e3
        ens_info = {"instances": K, "cylinders": [{"length": float(mm.x[:, 2].max()), "nel": int(mm.nel)}
                                                  for mm in shapes],
                    "nel": enel, "nelb": ens.nelb, "nshg": ens.N,
                    "graph_ms": ms_graph, "eager_ms": ms_eager, "graph_melem_s": enel / (ms_graph * 1e-3) / 1e6,
                    "ns_solve_ms": ms_solve, "ns_solver_iterations": [ls_e.RI["itr"], ls_e.GM["itr"], ls_e.CG["itr"]],
                    "autotune": ens.tuned[0] if ens.tuned else None}
        log(f"ensemble of {K} on 8 cylinders ({enel} tets, {ens.nelb} boundary): graph {ms_graph:.3f} ms "
            f"({ens_info['graph_melem_s']:.0f} Melem/s), eager {ms_eager:.3f} ms, union NS solve {ms_solve:.0f} ms "
            f"(RI/GM/CG {ens_info['ns_solver_iterations']}), autotune {ens_info['autotune']}")
        del g
        if args.ensemble_timestep:
            ens_info["timestep"] = ensemble_timestep(cfa, torch, ens, meshes, st, R_e)
        del ens, ey, eac, ex, er4

    # ---- timesteps: FP64, then the mixed form from the same initial state
    ts_ms = None
    ts_info = {}
    if args.timesteps > 0:
        R = 1333.0
        faceres = np.array([0.0, R * prm.alfi * prm.gami * prm.Delt])
        dibc = torch.from_numpy(ibc.astype(np.int32)).to(dev)
        rr = (m.x[:, 0] ** 2 + m.x[:, 1] ** 2)
        inl = m.inlet & ~m.wall
        dglob = torch.from_numpy(out.astype(np.int32)).to(dev)
        dnabi = torch.from_numpy(np.ascontiguousarray(nabi[out])).to(dev)   # (nf,3) C == (3,nf) F
        y0, ac0 = dy.clone(), dac.clone()
        rfin = torch.empty((4, N), dtype=torch.float64, device=dev)  # res(N,4) Fortran

        def new_ls():
            # config 2's solver: memLS NS, Kspace 100, tolerance 1e-8
            return cfa.memLS_LS_CREATE(cfa.LS_TYPE_NS, relTol=1e-8, dimKry=100, maxItr=10,
                                       relTolIn=[0.01, 0.2], maxItrIn=[1, 200])

        def inflow_bc(t):
            dBC = T(np.zeros((N, 7), order="F"))  # (7,N) C == BC(N,7) Fortran
            prof = np.zeros(N)
            prof[inl] = 2.0 * inflow_mean(t) * np.clip(1.0 - rr[inl], 0.0, None)
            dBC[4] = torch.from_numpy(prof).to(dev)   # BC(:,5) = u_z
            return dBC

        def residual(fm, lh, ya, aca, dx, dnabi, dq):
            """Res4 of the nonlinear iteration (formation into lh, bc3Res,
            ElmpvsQ) and its FP64 norm."""
            fm.to_lhs(ya, aca, dx, dres4, lh)
            fm.bc3res(dres4, res4_layout=True)
            fm.elmpvsq(dglob, dnabi, R, ya, dres4, dq)
            fm.sync()
            return torch.linalg.vector_norm(dres4)

        def run_steps(fm, lh, label, dx=dx, fref=form, dnabi=dnabi):
            """args.timesteps timesteps of 3 nonlinear iterations each (config
            2); also returns the FP64 residual norm (after bc3Res and ElmpvsQ)
            at each formation and once more after the last update, formed by
            fref outside the timed region."""
            ls = new_ls()
            dq = torch.zeros(1, dtype=torch.float64, device=dev)
            yc, acc = y0.clone(), ac0.clone()
            yold, acold = y0.clone(), ac0.clone()
            ya, aca = torch.empty_like(yc), torch.empty_like(acc)
            times, its, rnorms = [], [], []
            for n in range(args.timesteps):
                dBC = inflow_bc((n + 1) * prm.Delt)
                torch.cuda.synchronize()
                t0 = time.perf_counter()
                fm.predict(yold, acold, yc, acc)
                fm.itr_bc(dibc, dBC, yc)
                rn = []
                for it in range(3):
                    fm.yalpha(yold, acold, yc, acc, ya, aca)
                    fm.sync()
                    rn.append(residual(fm, lh, ya, aca, dx, dnabi, dq))
                    cfa.memLS_SOLVE_dev(lh, ls, 4, dres4, res=faceres)
                    its.append((ls.RI["itr"], ls.GM["itr"], ls.CG["itr"], ls.RI["fNorm"] / ls.RI["iNorm"]))
                    fm.correct(yc, acc, dres4)
                    fm.itr_bc(dibc, dBC, yc)
                    fm.sync()
                torch.cuda.synchronize()
                times.append(time.perf_counter() - t0)
                fref.yalpha(yold, acold, yc, acc, ya, aca)
                fref.sync()
                fref.dev(ya, aca, dx, rfin, lhs=0)
                fref.bc3res(rfin)
                rq = torch.zeros(1, dtype=torch.float64, device=dev)
                rres4 = rfin.t().contiguous()
                fref.elmpvsq(dglob, dnabi, R, ya, rres4, rq)
                fref.sync()
                rn.append(torch.linalg.vector_norm(rres4))
                rnorms.append([float(v) for v in rn])
                yold.copy_(yc)
                acold.copy_(acc)
                log(f"{label} timestep {n + 1}: {times[-1] * 1e3:.1f} ms, solver (RI,GM,CG itr, rel) {its[-3:]}")
            return times, its, yc, rnorms

        def converge_step(fm, lh, label, dx=dx, dnabi=dnabi, tol=1e-12, max_it=args.max_nonlinear):
            """The first timestep driven until its FP64 nonlinear residual
            (Res4 after bc3Res and ElmpvsQ, formed in FP64 by both forms) is
            <= tol times its value at the first nonlinear iteration, each
            iteration one formation + NS solve (config 2's solver settings) +
            update (each solve's right-hand side scaled to unit norm).
            Returns (seconds elapsed when each residual was formed,
            residual norms, solver iterations, the state y after each update
            (copies kept outside the timed work))."""
            ls = new_ls()
            dq = torch.zeros(1, dtype=torch.float64, device=dev)
            yc, acc = y0.clone(), ac0.clone()
            ya, aca = torch.empty_like(yc), torch.empty_like(acc)
            dBC = inflow_bc(prm.Delt)
            rn, its, tk, ys = [], [], [], []
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            paused = 0.0
            fm.predict(y0, ac0, yc, acc)
            fm.itr_bc(dibc, dBC, yc)
            for it in range(max_it + 1):
                fm.yalpha(y0, ac0, yc, acc, ya, aca)
                fm.sync()
                rn.append(float(residual(fm, lh, ya, aca, dx, dnabi, dq)))
                tk.append(time.perf_counter() - t0 - paused)
                tp = time.perf_counter()
                ys.append(yc.clone())
                torch.cuda.synchronize()
                paused += time.perf_counter() - tp
                if rn[-1] <= tol * rn[0] or it == max_it:
                    break
                # right-hand side at unit norm: memLS's GE refuses pivots below
                # 2 eps in absolute terms, which the Gram matrix of a residual
                # driven this far down reaches
                dres4.div_(rn[-1])
                cfa.memLS_SOLVE_dev(lh, ls, 4, dres4, res=faceres)
                dres4.mul_(rn[-1])
                its.append((ls.RI["itr"], ls.GM["itr"], ls.CG["itr"], ls.RI["fNorm"] / ls.RI["iNorm"]))
                fm.correct(yc, acc, dres4)
                fm.itr_bc(dibc, dBC, yc)
                fm.sync()
            torch.cuda.synchronize()
            log(f"{label} to {tol:g}: {tk[-1] * 1e3:.0f} ms, {len(its)} nonlinear iterations, residual "
                f"{rn[-1] / rn[0]:.2e} of the first ({' '.join(f'{v / rn[0]:.1e}' for v in rn)})")
            return tk, rn, its, ys

        def rel_err(a, b):
            d = (a - b).abs()
            return {"u": float(d[:3].max() / b[:3].abs().max()), "p": float(d[3].max() / b[3].abs().max())}

        def conv_pair(fa, la, fb, lb, tag, target=1e-8, **kw):
            """Time to equal final error.  The FP64 form is driven to a
            residual of 1e-12 of the first; its last state is the FP64
            solution y*.  For each form (FP64, mixed), the time to target is
            the time at which it first holds a state whose FP64 residual is
            <= target of the first and whose velocity and pressure are each
            within target of y* (max-norm relative per field).  The speedup
            is reported only when both forms reach that point."""
            ta, ra, ia, ysa = converge_step(fa, la, f"{tag}fp64", **kw)
            tb, rb, ib, ysb = converge_step(fb, lb, f"{tag}mixed", **kw)
            ystar = ysa[-1]
            ref_ok = ra[-1] <= 1e-12 * ra[0]

            def first_hit(tk, rn, ys):
                errs = [rel_err(v, ystar) for v in ys]
                for k in range(len(rn)):
                    if rn[k] <= target * rn[0] and max(errs[k].values()) <= target:
                        return k, errs
                return None, errs

            ka, ea = first_hit(ta, ra, ysa)
            kb, eb = first_hit(tb, rb, ysb)
            ok = ref_ok and ka is not None and kb is not None
            out = {"reference": "FP64 driven to 1e-12 of the first residual",
                   "reference_reached": bool(ref_ok),
                   "fp64_ms": ta[ka] * 1e3 if ka is not None else None,
                   "mixed_ms": tb[kb] * 1e3 if kb is not None else None,
                   "fp64_nonlinear_iterations": ka, "mixed_nonlinear_iterations": kb,
                   "fp64_residual_rel": [v / ra[0] for v in ra], "mixed_residual_rel": [v / rb[0] for v in rb],
                   "fp64_err_vs_reference": ea, "mixed_err_vs_reference": eb,
                   "fp64_solver_iterations": [list(map(float, v)) for v in ia],
                   "mixed_solver_iterations": [list(map(float, v)) for v in ib],
                   "mixed_rel_err_vs_fp64": eb[kb] if kb is not None else eb[-1],
                   "fp64_rel_err_vs_reference": ea[ka] if ka is not None else ea[-1],
                   "reached_1e-8": bool(ok),
                   "mixed_speedup": (ta[ka] / tb[kb]) if ok else None}
            log(f"{tag}to {target:g}: fp64 {out['fp64_ms']} ms ({ka} it), mixed {out['mixed_ms']} ms ({kb} it), "
                f"mixed error {out['mixed_rel_err_vs_fp64']}, speedup {out['mixed_speedup']}")
            del ysa, ysb
            return out

        st64, its64, y64, rn64 = run_steps(form, lhs, "fp64")
        ts_ms = float(np.mean(st64) * 1e3)
        ts_info = {"timesteps_timed": args.timesteps, "nonlinear_iterations": 3,
                   "timestep_ms_fp64_each": [t * 1e3 for t in st64],
                   "solver_iterations": [list(map(float, v)) for v in its64],
                   "residual_norms_fp64": rn64}
        if args.mixed:
            st32, its32, y32, rn32 = run_steps(form32, lhs32, "mixed")
            err = rel_err(y32, y64)
            ms32 = float(np.mean(st32) * 1e3)
            ts_info.update({"timestep_ms_mixed": ms32, "timestep_ms_mixed_each": [t * 1e3 for t in st32],
                            "fixed_iterations_speedup": ts_ms / ms32,
                            "fixed_iterations_rel_err_vs_fp64": err,
                            "residual_norms_mixed": rn32,
                            "solver_iterations_mixed": [list(map(float, v)) for v in its32]})
            log(f"mixed: {ms32:.1f} ms/timestep vs fp64 {ts_ms:.1f} -> x{ts_ms / ms32:.2f} at 3 nonlinear "
                f"iterations (state error {err})")
            if args.to_1e8:
                conv = conv_pair(form, lhs, form32, lhs32, "")
                ts_info["to_1e-8"] = conv
                ts_info["timestep_to_1e-8_ms"] = {"fp64": conv["fp64_ms"], "mixed": conv["mixed_ms"]}
                ts_info["mixed_speedup"] = conv["mixed_speedup"]
                ts_info["mixed_rel_err_vs_fp64"] = conv["mixed_rel_err_vs_fp64"]

        if args.mixed and args.stretched:
            # the same cylinder with 5 % of its elements in seeded regions of
            # anisotropic stretch up to 50x (mesh.stretch_regions); elements
            # with aspect ratio > 20 form the FP64 tau island of the mixed form
            mst = M.stretch_regions(m)
            # aspect ratios of the elements the map moved (the others keep the
            # plain mesh's, at most 2)
            asp = M.aspect_ratios(mst.x, np.asarray(mst.ien)[M.moved_elements(m, mst)])
            dxs = T(mst.x)
            nabi_s = mst.nabi()
            dnabi_s = torch.from_numpy(np.ascontiguousarray(nabi_s[out])).to(dev)
            bnd_s = M.cylinder_boundary(mst) if args.boundary else None
            pair = []
            for prec in (64, 32):
                fs = cfa.ElmGMRInterior(m.ien, colm, rowp, N, iBC=ibc,
                                        params=cfa.form_params(precision=prec, aspect_fp64=20.0))
                if bnd_s is not None:
                    fs.set_boundary(*bnd_s)
                ls_ = cfa.memLS_LHS_CREATE(N, N, nnz, None, colm, rowp, 2)
                cfa.memLS_BC_CREATE(ls_, 1, fixed.size, 3, cfa.BC_TYPE_Dir, fixed + 1,
                                    np.zeros((3, fixed.size), order="F"))
                cfa.memLS_BC_CREATE(ls_, 2, out.size, 3, cfa.BC_TYPE_Neu, out + 1, np.asfortranarray(nabi_s[out].T))
                pair.append((fs, ls_))
            cs = conv_pair(pair[0][0], pair[0][1], pair[1][0], pair[1][1], "stretched ", dx=dxs, dnabi=dnabi_s)
            ts_info["stretched"] = {
                "elements_stretched_frac": 0.05, "max_stretch": 50.0,
                "aspect_max": float(asp.max()), "fp64_island_frac": float((asp > 20.0).sum() / m.nel),
                "to_1e-8": cs}
            del pair

    # ---- adjoint: the VJP of the residual (interior and boundary elements)
    # with respect to the nodal state and the viscosity, on this mesh, timed,
    # and checked against central differences of the residual (h = 1e-6
    # relative) along a seeded state direction and in mu
    adj = None
    if args.adjoint:
        gen = torch.Generator(device=dev).manual_seed(11)
        w = torch.randn((4, N), dtype=torch.float64, device=dev, generator=gen)
        ybar = torch.empty_like(w)
        mb = form.vjp(dy, dac, dx, w, ybar)
        torch.cuda.synchronize()
        reps = max(3, args.steps // 2)
        t0 = time.perf_counter()
        for _ in range(reps):
            mb = form.vjp(dy, dac, dx, w, ybar)
        torch.cuda.synchronize()
        ms_vjp = (time.perf_counter() - t0) / reps * 1e3
        # a smooth seeded direction (a node-by-node random one makes the
        # second differences of the residual so large that no step size
        # separates truncation from rounding), zero at the outlet, whose
        # influx term switches on where u.n < 0
        rs = np.random.default_rng(12)
        vh = np.zeros((4, N))
        for c in range(4):
            k = rs.uniform(0.5, 2.0, 3)
            ph = rs.uniform(0, 2 * np.pi, 3)
            vh[c] = (np.sin(k[0] * m.x[:, 0] + ph[0]) * np.cos(k[1] * m.x[:, 1] + ph[1]) *
                     np.sin(k[2] * m.x[:, 2] / 10.0 + ph[2]))
        vh[3] *= 100.0  # pressure scale
        vh[:, m.outlet] = 0.0
        v = torch.from_numpy(vh).to(dev)
        rp, rm = torch.empty_like(w), torch.empty_like(w)
        ad_y = float((ybar * v).sum())
        fd_ys = {}
        for hr in (1e-6, 1e-4):
            h = hr * float(dy.abs().max())
            form.dev(dy + h * v, dac, dx, rp, lhs=0)
            form.dev(dy - h * v, dac, dx, rm, lhs=0)
            torch.cuda.synchronize()
            fd_ys[hr] = float((w * (rp - rm)).sum() / (2 * h))
        fd_y = fd_ys[1e-6]
        inst = np.zeros(nel, np.uint16)
        hm = 1e-6 * prm.mu
        form.set_instances(inst, np.array([prm.mu + hm]))
        form.dev(dy, dac, dx, rp, lhs=0)
        form.set_instances(inst, np.array([prm.mu - hm]))
        form.dev(dy, dac, dx, rm, lhs=0)
        form.set_instances(inst, np.zeros(0))
        torch.cuda.synchronize()
        fd_mu = float((w * (rp - rm)).sum() / (2 * hm))
        adj = {"vjp_ms": ms_vjp, "vjp_melem_s": nel / (ms_vjp * 1e-3) / 1e6,
               "fd_check_state": {"vjp": ad_y, "central_difference": fd_y, "rel": abs(ad_y - fd_y) / abs(fd_y),
                                  "rel_h1e-4": abs(ad_y - fd_ys[1e-4]) / abs(fd_ys[1e-4])},
               "fd_check_mu": {"vjp": mb, "central_difference": fd_mu, "rel": abs(mb - fd_mu) / abs(fd_mu)}}
        if args.adjoint_steps > 0:
            adj["implicit"] = implicit_adjoint(cfa, M, torch, m, args.adjoint_steps, args.adjoint_every)
        log(f"adjoint: VJP {ms_vjp:.2f} ms ({adj['vjp_melem_s']:.0f} Melem/s); central differences: state "
            f"{adj['fd_check_state']['rel']:.2e} (h 1e-6), {adj['fd_check_state']['rel_h1e-4']:.2e} (h 1e-4), "
            f"mu {adj['fd_check_mu']['rel']:.2e} relative")

    # ---- the CPU-runnable case (20x20x100 cylinder, outlet natural): one
    # formation and one memLS solve under both readings of "GMRES dim 50, rel
    # tol 1e-6": GMRESV on the full 4x4 system, and the NS solver
    c1 = None
    if args.config1:
        m1 = M.cylinder_mesh(20, 20, 100)
        n1 = m1.numnp
        colm1, rowp1 = cfa.genadj(m1.ien, n1)
        ibc1, _ = M.dirichlet_codes(m1)
        f1 = cfa.ElmGMRInterior(m1.ien, colm1, rowp1, n1, iBC=ibc1, params=cfa.form_params())
        f1.set_boundary(*M.cylinder_boundary(m1))
        y1, a1 = M.poiseuille_state(m1)
        fixed1 = np.nonzero(ibc1)[0]
        c1 = {"workload": "20x20x100 cylinder, outlet natural", "nel": int(m1.nel), "nshg": int(n1)}
        for label, lt, kw in (("GMRESV 4x4", cfa.LS_TYPE_GMRES, dict(relTol=1e-6, dimKry=50, maxItr=10)),
                              ("NS", cfa.LS_TYPE_NS, dict(relTol=1e-6, dimKry=50, maxItr=10))):
            lh1 = cfa.memLS_LHS_CREATE(n1, n1, rowp1.size, None, colm1, rowp1, 1)
            cfa.memLS_BC_CREATE(lh1, 1, fixed1.size, 3, cfa.BC_TYPE_Dir, fixed1 + 1,
                                np.zeros((3, fixed1.size), order="F"))
            r1 = torch.zeros((n1, 4), dtype=torch.float64, device=dev)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            f1.to_lhs(T(y1), T(a1), T(m1.x), r1, lh1)
            f1.bc3res(r1, res4_layout=True)
            torch.cuda.synchronize()
            t1 = time.perf_counter()
            ls1 = cfa.memLS_LS_CREATE(lt, **kw)
            cfa.memLS_SOLVE_dev(lh1, ls1, 4, r1)
            torch.cuda.synchronize()
            t2 = time.perf_counter()
            c1[label] = {"formation_ms": (t1 - t0) * 1e3, "solve_ms": (t2 - t1) * 1e3,
                         "iterations": [ls1.RI["itr"], ls1.GM["itr"], ls1.CG["itr"]], "converged": ls1.RI["suc"]}
            log(f"20x20x100, {label}: formation {(t1 - t0) * 1e3:.1f} ms, solve {(t2 - t1) * 1e3:.1f} ms, "
                f"iterations {c1[label]['iterations']}, converged {ls1.RI['suc']}")
        del f1, lh1, r1

    # ---- the reference's own basic case (testbin/mainTests/basic, held in
    # tests/golden: 1072 nodes, deformable wall, bct.dat inflow, RCR outlet)
    # for its 5 steps on the device, against the reference's output of that
    # run: the RCR outlet's flow after step 5 (its test: 714.921079082528
    # within 1e-7) and the per-step wall time
    refcase = None
    if args.reference_case:
        refcase = reference_case(cfa, np)

    # ---- measured ceilings of this device (a plain copy, FMA chains)
    peaks = cfa.measure_peaks(1 << 31)
    log(f"measured ceilings: copy {peaks['copy_gbs']:.0f} GB/s, FP64 FMA {peaks['fp64_tflops']:.1f} TFLOP/s, "
        f"FP32 FMA {peaks['fp32_tflops']:.1f} TFLOP/s (packed FP32 not selected)")

    # ---- roofline (from the live HIP-event timings).  The formation's unit
    # of work is the element; its bound is FP64 VALU (no dense contraction,
    # arithmetic intensity above the HBM balance): achieved = the counted
    # flops of the restatement (8996 per element, 10 per node) over the
    # formation kernels' duration (flux pass, e3, assembly, boundary rows).
    # The dominant kernel (the row assembly, HBM-bound) is reported beside it
    # with the bytes its design must move.
    form_ms = tm[3]
    flops = FLOPS_PER_ELEMENT * nel + FLOPS_PER_NODE * N
    ach = flops / (form_ms * 1e-3) / 1e12
    traffic = [pmc_traffic(grid, k) for k in FORMATION_KERNELS]
    roof = {"kernel": "formation (" + " + ".join(FORMATION_KERNELS) + ")", "bound": "valu",
            "achieved": ach, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s", "frac": ach / FP64_PEAK_TFLOPS,
            "algorithmic_flops": flops, "duration_ms": form_ms,
            "traffic": sum(traffic) if all(t is not None for t in traffic) else None,
            "measured_peak": peaks["fp64_tflops"], "frac_of_measured": ach / peaks["fp64_tflops"]}
    # assembly: each incidence record read once (12 doubles per element-node)
    # with its element's nodes (16 B per element-node), the packed nodes
    # once (6 doubles/node), every block written once (16 doubles/nnz), Res4
    # (4 doubles/node) and the contribution lists (2 B per off-diagonal
    # contribution, 12 per element)
    byt = nel * 4 * (12 * 8 + 16) + N * 6 * 8 + nnz * 16 * 8 + N * 4 * 8 + nel * 12 * 2
    kname = "asm_row<1,8>" if tuned is not None and tuned["fp64"][0].split("/")[1] == "8" else "asm_row<1,16>"
    roof["dominant_kernel"] = {"kernel": kname, "bound": "hbm", "achieved": byt / (tm[2] * 1e-3) / 1e9,
                               "peak": HBM_PEAK_GBS, "unit": "GB/s",
                               "frac": byt / (tm[2] * 1e-3) / 1e9 / HBM_PEAK_GBS, "algorithmic_bytes": byt,
                               "traffic": pmc_traffic(grid, kname), "duration_ms": tm[2]}
    log(f"roofline: formation {ach:.1f} TFLOP/s = {ach / FP64_PEAK_TFLOPS:.3f} of FP64 peak "
        f"({flops / 1e9:.1f} GFLOP in {form_ms:.3f} ms)")

    cpu = None
    slab = None
    if args.cpu_seconds > 0:
        P = cpu_threads()
        sample = (grid[0], grid[1], max(4, grid[2] // 16))
        v, sn, nrep, dt = cpu_baseline(sample, args.cpu_seconds, P)
        cpu = {"value": v, "unit": "Melem/s", "cores": P, "kind": "port", "threads": P, "cpu_model": cpu_model(),
               "sample": f"oracle interior formation (flux pass + e3 + fillsparseI, OpenMP build, {P} threads over "
                         f"element blocks and CSR rows) on the {sample[0]}x{sample[1]}x{sample[2]} slab of the "
                         f"same cross-section ({sn} tets), {nrep} calls in {dt:.1f}s"}
        log(f"cpu baseline: {v:.3f} Melem/s on {P} threads ({cpu_model()})")
        if args.slab:
            sg = tuple(int(v) for v in args.slab.split(","))
            gms, gits = gpu_timestep(cfa, M, torch, sg)
            cs, cits = cpu_timestep(sg, threads=P)
            slab = {"grid": list(sg), "gpu_ms": gms, "cpu_ms": cs * 1e3, "speedup": cs * 1e3 / gms,
                    "solver_iterations_gpu": [list(map(float, v)) for v in gits],
                    "solver_iterations_cpu": [list(map(float, v)) for v in cits],
                    "cpu": f"oracle (C restatement of the formation and memLS) on {P} threads (interior formation, "
                           f"Val4 packing, memLS), boundary elements and numpy step glue on one ({cpu_model()})"}
            log(f"timestep on the {sg} slab: GPU {gms:.1f} ms, CPU {cs * 1e3:.0f} ms on {P} threads "
                f"-> x{cs * 1e3 / gms:.1f}")
        if args.cpu_full_timestep and ts_ms is not None:
            cs, cits = cpu_timestep(grid, threads=P)
            slab = dict(slab or {})
            slab["full_mesh"] = {"grid": list(grid), "gpu_ms": ts_info["timestep_ms_fp64_each"][0],
                                 "cpu_ms": cs * 1e3, "speedup": cs * 1e3 / ts_info["timestep_ms_fp64_each"][0],
                                 "solver_iterations_cpu": [list(map(float, v)) for v in cits],
                                 "threads": P}
            log(f"first timestep on the full mesh: GPU {ts_info['timestep_ms_fp64_each'][0]:.0f} ms, CPU "
                f"{cs * 1e3:.0f} ms on {P} threads -> x{cs * 1e3 / ts_info['timestep_ms_fp64_each'][0]:.1f}")

    # ---- roofline.traffic: HBM bytes per launch measured now (two PMC
    # child runs of the formation on this workload), else from the committed
    # profile of the same command
    live = None
    if args.traffic:
        t0 = time.time()
        live = measure_traffic(grid)
        log(f"traffic: {'measured' if live else 'not measured (no rocprofv3, or running under it)'} "
            f"({time.time() - t0:.0f}s)")
    if live:
        ft = [live.get(k.replace(" ", "")) for k in FORMATION_KERNELS]
        roof["traffic"] = sum(ft) if all(t is not None for t in ft) else None
        roof["traffic_by_kernel"] = {k: live.get(k.replace(" ", "")) for k in FORMATION_KERNELS}
        roof["dominant_kernel"]["traffic"] = live.get(roof["dominant_kernel"]["kernel"].replace(" ", ""))
        roof["traffic_source"] = "measured: rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE child runs of this workload"
    else:
        roof["traffic_source"] = (f"profiles/formation_{grid[0]}x{grid[1]}x{grid[2]}_pmc_traffic.md"
                                  if roof["traffic"] is not None else None)

    line = {
        "metric": METRIC, "value": melems, "unit": "Melem/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms_step, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"FP64 formation (flux pass + e3 + assembly into the memLS matrix"
                               f"{' + boundary elements' if nelb else ''}) on the "
                               f"cylinder r=1 L=10 from a {grid[0]}x{grid[1]}x{grid[2]} hex grid clipped by hex-centroid "
                               f"radius, 6 Kuhn tets per hex",
                   "nel": nel, "nelb": nelb, "nshg": N, "nnzb": nnz, "rho": 1.06, "mu": 0.04, "dt": 1e-3, "rho_inf": 0.5},
        "phase_ms": {"flux": tm[0], "e3_elem": tm[1], "assembly": tm[2], "boundary": tm[3] - tm[0] - tm[1] - tm[2],
                     "events_total": tm[3]},
        "krylov_spmv_gbs": {"bsr3x3": gbs3, "bsr4x4": gbs4, "ms3": ms3, "ms4": ms4},
        "krylov_spmv_gbs_fp32": spmv32, "mixed_formation": mixed_form,
        "timestep_ms_fp64": ts_ms, **ts_info,
        "roofline": roof, "cpu_baseline": cpu, "timestep_slab": slab,
        "autotune": None if tuned is None else {k: {"chosen": c, "assembly_ms": t} for k, (c, t) in tuned.items()},
        "ensemble": ens_info,
        "case_20x20x100": c1,
        "adjoint": adj,
        "reference_case": refcase,
        "ceilings": {"measured": peaks, "published": {"hbm_gbs": HBM_PEAK_GBS, "fp64_tflops": FP64_PEAK_TFLOPS,
                                                      "fp32_tflops": 157.3}},
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
