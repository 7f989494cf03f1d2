"""Benchmark: J2 material-point update rate of the stdBrick state
determination on one MI355X.

Workload: the 64^3 stdBrick cube (274,625 nodes, 262,144 elements,
2,097,152 Gauss points) of J2Plasticity (K=166.7e3, G=76.9e3, sigma_y=250,
H=2000), FP64.  One step is one Domain::update over the whole mesh: every
element gathers its nodal coordinates and trial displacements, every Gauss
point forms its strain and runs the J2 return map, writing plastic state,
stress and the consistent tangent.  The trial displacement is a fixed
synthetic field that leaves part of the cube plastic; inputs are resident in
HBM before the timed region.

Prints one JSON line (value = material-point updates per second) with the
roofline of the update kernel (algorithmic bytes / HIP-event time) and the
CPU oracle timed on a bounded sample of the same workload on this host.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def update_bytes(nnode, nele, npts):
    # per point: read eps_p,n (6) + xi_n (1); write eps_p,n+1 (6) + xi (1) +
    # stress (6) + the two tangent coefficients (2) = 22 doubles = 176 B;
    # per node: coordinates + trial displacement = 48 B; per element: 8 int32
    return npts * 176 + nnode * 48 + nele * 32


def synthetic_field(mesh):
    x, y, z = mesh["crds"].T
    eps0 = 2.0e-3
    mod = 1.0 + 0.6 * np.sin(2 * np.pi * x) * np.cos(np.pi * y)
    U = np.zeros((len(x), 3))
    U[:, 2] = -eps0 * z * mod
    U[:, 0] = 0.3 * eps0 * (x - 0.5) * z
    U[:, 1] = 0.3 * eps0 * (y - 0.5) * z * mod
    return U


def host_cpu():
    """CPU model and physical core count of this host (lscpu / cpuinfo)."""
    model, cores = "unknown", None
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
        import subprocess
        out = subprocess.run(["lscpu", "-p=CORE,SOCKET"], capture_output=True, text=True).stdout
        pairs = {ln for ln in out.splitlines() if ln and not ln.startswith("#")}
        cores = len(pairs) or None
    except OSError:
        pass
    return model, cores


def cpu_baseline(n_sample, seconds):
    """The oracle's update of an n_sample^3 cube with the same synthetic
    field, built here with -O3 -march=native and OpenMP over elements, timed
    on one thread (the reference's serial Domain::update) and on every CPU
    this process may run on; `value` is the all-threads rate."""
    import shutil
    import tempfile

    from tests import _oracle
    from xara_amd import meshes
    tmp = tempfile.mkdtemp(prefix="xara_cpu_")
    try:
        path = os.path.join(tmp, "liboracle_native.so")
        t0 = time.time()
        _oracle.build_native(path)
        build_s = time.time() - t0
        b = _oracle.load_native(path)
        mesh = meshes.brick_cube(n_sample)
        o = _oracle.oracle_model(3, 3, binding=b)
        meshes.build_brick_cube(o, mesh, pressure=0.0)
        o.setup()
        o.set_trial_disp(synthetic_field(mesh))
        npts = 8 * n_sample ** 3
        threads = len(os.sched_getaffinity(0))

        def rate(nt, budget):
            b.set_threads(nt)
            o.update()
            ms = o.time_op("update", 1)
            reps = max(2, int(budget * 1000.0 / max(ms, 1e-3)))
            ms = o.time_op("update", reps)
            return npts / (ms * 1e-3), reps

        one, reps1 = rate(1, seconds / 2)
        allc, repsn = rate(threads, seconds / 2)
        o.close()
    finally:
        shutil.rmtree(tmp, ignore_errors=True)
    model, phys = host_cpu()
    return dict(value=allc, unit="points/s", cores=threads, kind="port", one_thread=one,
                physical_cores=phys, cpu=model, build=_oracle.NATIVE_FLAGS, build_s=round(build_s, 1),
                sample=f"oracle update of a {n_sample}^3 stdBrick J2 cube ({npts} points) with the same synthetic "
                       f"field: {repsn} updates on {threads} threads (OpenMP over elements), {reps1} on one thread")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--n", type=int, default=64, help="cube edge in elements")
    ap.add_argument("--cpu-n", type=int, default=32, help="cube edge of the CPU sample")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--copy-calib", action="store_true",
                    help="also run the 256 MiB copy kernel (profiler counter calibration)")
    ap.add_argument("--mode", choices=["update", "newton", "quad", "frame", "force", "ensemble", "adjoint"],
                    default="update")
    ap.add_argument("--adj-frame", default="5,5,10", help="bays x, bays y, stories of the adjoint mode")
    ap.add_argument("--adj-steps", type=int, default=500)
    ap.add_argument("--ckpt", type=int, default=25)
    ap.add_argument("--ddm", action="store_true", help="also run the CPU oracle's forward-mode check")
    ap.add_argument("--models", type=int, default=4096, help="frames in the ensemble mode")
    ap.add_argument("--record-steps", type=int, default=2000, help="ground-motion record length (steps)")
    ap.add_argument("--ens-start", type=int, default=200,
                    help="ensemble steps run untimed before the timed window (the record's envelope rise)")
    ap.add_argument("--ens-steps", type=int, default=800, help="ensemble steps timed (steps 200-1000 by default)")
    ap.add_argument("--looped", type=int, default=16, help="frames timed alone for the looped comparison")
    ap.add_argument("--looped-steps", type=int, default=20)
    ap.add_argument("--no-tune", action="store_true")
    ap.add_argument("--frame", default="10,10,20", help="bays x, bays y, stories of the frame mode")
    ap.add_argument("--frame-steps", type=int, default=0,
                    help="DisplacementControl steps of the pushover (frame mode: 0 = none; force mode: 0 = 60)")
    ap.add_argument("--newton-steps", type=int, default=40)
    ap.add_argument("--system", default="GpuBlockPCG")
    ap.add_argument("--spmv", choices=["auto", "csr", "sym"], default="auto", help="SpMV layout of the PCG")
    ap.add_argument("--lanes", type=int, default=0, help="SpMV threads per row (0 = from the row length)")
    ap.add_argument("--cg-graph", type=int, choices=[0, 1], default=1,
                    help="replay PCG chunks from captured graphs (0 = launch each kernel)")
    ap.add_argument("--cg-fuse", type=int, choices=[0, 1], default=1,
                    help="alpha / beta reductions inside the PCG vector kernels on small systems")
    ap.add_argument("--cg-small", type=int, choices=[0, 1], default=1,
                    help="whole PCG solve in one workgroup for systems that fit it")
    ap.add_argument("--cg-warm", type=int, choices=[0, 1], default=1,
                    help="start a step's first solve from the previous step's solution (0 = from zero)")
    ap.add_argument("--cg-wide", type=int, choices=[0, 1], default=1,
                    help="exact tiled band-Cholesky preconditioner for banded systems above 1,024 equations")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1"}))
        sys.exit(2)

    if args.mode == "newton":
        return newton_main(args)
    if args.mode == "frame":
        return frame_main(args)
    if args.mode == "force":
        return force_main(args)
    if args.mode == "quad":
        return quad_main(args)
    if args.mode == "ensemble":
        return ensemble_main(args)
    if args.mode == "adjoint":
        return adjoint_main(args)

    from xara_amd import meshes
    from xara_amd.model import Model

    n = args.n
    print(f"bench: building {n}^3 stdBrick cube", file=sys.stderr, flush=True)
    mesh = meshes.brick_cube(n)
    m = Model(3, 3, device=0)
    meshes.build_brick_cube(m, mesh, pressure=0.0)
    t0 = time.time()
    m.setup()
    setup_s = time.time() - t0
    nnode, nele = len(mesh["node_tags"]), len(mesh["ele_tags"])
    npts = 8 * nele
    m.set_trial_disp(synthetic_field(mesh))
    print(f"bench: setup {setup_s:.1f}s neq={m.neq} nnz={m.nnz}; warmup", file=sys.stderr, flush=True)
    m.time_op("update", args.warmup)
    # timed region: exactly `steps` updates on the library stream, bracketed
    # by device synchronisation (time_op synchronises before returning)
    t0 = time.perf_counter()
    ev_ms = m.time_op("update", args.steps)
    wall = time.perf_counter() - t0
    ms_per_step = 1e3 * wall / args.steps
    value = npts * args.steps / wall
    bytes_per = update_bytes(nnode, nele, npts)
    achieved = bytes_per / (ev_ms * 1e-3) / 1e9
    # HBM bytes per launch measured with rocprofv3 FETCH_SIZE / WRITE_SIZE on
    # this workload, scaled by the copy-kernel calibration (profiles/)
    traffic = None
    prof = os.path.join(ROOT, "profiles", "brick_update_64cube.json")
    if n == 64 and os.path.exists(prof):
        traffic = json.load(open(prof))["traffic_bytes"]
    if args.copy_calib:
        m.time_op("copy", 3)
    plastic = int(np.count_nonzero(m.point_data("plastic", npts)[:, 6] > 0))
    print(f"bench: {ms_per_step:.3f} ms/step (event {ev_ms:.3f} ms), {plastic} of {npts} points plastic",
          file=sys.stderr, flush=True)
    out = {
        "metric": "material-point updates/s",
        "value": value,
        "unit": "points/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"stdBrick J2 cube {n}^3 update (Domain::update)",
                   "elements": nele, "nodes": nnode, "points": npts,
                   "plastic_points": plastic, "setup_s": round(setup_s, 2)},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "bytes_per_launch": bytes_per, "event_ms": ev_ms},
    }
    if not args.no_cpu:
        print("bench: CPU baseline", file=sys.stderr, flush=True)
        cb = cpu_baseline(args.cpu_n, args.cpu_seconds)
        out["cpu_baseline"] = cb
        out["gpu_over_cpu"] = value / cb["value"]
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out))


def gpu_system(m, args, rtol, maxit):
    """The PCG system of a GPU model with the bench's issue pattern."""
    m.system(args.system, rtol, maxit)
    if m._b.prefix == "xg_":
        m.cg_config(16, bool(args.cg_graph), bool(args.cg_fuse), bool(args.cg_small), bool(args.cg_warm),
                    bool(args.cg_wide))


def newton_pushover(m, n, steps, max_seconds=None, max_steps=None, log=True):
    """The brick-cube pushover on a set-up model: first-yield pressure p_y
    from one elastic solve at unit pressure, then `steps` LoadControl steps
    of 1.5 p_y / steps (NewtonRaphson, NormUnbalance).  Returns (Newton
    iterations, wall seconds, steps run, p_y, PCG iterations, solves) of the
    load steps; stops early after max_steps steps or max_seconds."""
    from xara_amd import meshes
    mesh = meshes.brick_cube(n)
    npts = 8 * len(mesh["ele_tags"])
    m.set_load_factor(1.0)
    B = m.form_unbalance()
    m.form_tangent(values=False)
    u = m.solve(B)
    m.incr_disp(u)
    m.update()
    sig = m.point_data("stress", npts)
    s0, s1, s2, s3, s4, s5 = sig.T
    vm = np.sqrt(0.5 * ((s0 - s1) ** 2 + (s1 - s2) ** 2 + (s2 - s0) ** 2) + 3 * (s3 ** 2 + s4 ** 2 + s5 ** 2))
    p_y = 250.0 / vm.max()
    m.revert_to_start()
    m.set_load_factor(0.0)
    m.test("NormUnbalance", 1e-6 * p_y * (1.0 / n) ** 2 * n, 25)
    m.integrator("LoadControl", 1.5 * p_y / steps)
    if log:
        print(f"bench: first-yield pressure {p_y:.4f}; running {steps} steps", file=sys.stderr, flush=True)
    c0 = m.counters()
    t0 = time.perf_counter()
    run = 0
    for st in range(steps if max_steps is None else min(steps, max_steps)):
        rc = m.analyze(1)
        run += 1
        c = m.counters()
        if log:
            print(f"bench: step {st + 1} rc={rc} newton={c['newton_iters'] - c0['newton_iters']} "
                  f"cg={c['cg_iters'] - c0['cg_iters']} t={time.perf_counter() - t0:.1f}s", file=sys.stderr,
                  flush=True)
        if rc != 0:
            break
        if max_seconds is not None and time.perf_counter() - t0 > max_seconds:
            break
    wall = time.perf_counter() - t0
    c = m.counters()
    return (c["newton_iters"] - c0["newton_iters"], wall, run, p_y, c["cg_iters"] - c0["cg_iters"],
            c["solves"] - c0["solves"])


def newton_main(args):
    """Newton iterations/s of the full pushover: cube under uniform top
    pressure ramped to 1.5x the first-yield pressure in `newton_steps`
    LoadControl steps, NewtonRaphson with NormUnbalance, GPU PCG.  The CPU
    baseline: the oracle (one core, its profile LDL^T direct solve, the
    role of the reference's default solver) on the same pushover of a 12^3
    cube for about --cpu-seconds, with the GPU on that cube beside it (a
    64^3 factorisation would need tens of GB and hours)."""
    from xara_amd import meshes
    from xara_amd.model import Model

    n = args.n
    mesh = meshes.brick_cube(n)
    m = Model(3, 3, device=0)
    meshes.build_brick_cube(m, mesh, pressure=1.0)
    gpu_system(m, args, 1e-10, 50000)
    m.spmv_config(args.spmv, args.lanes)
    t0 = time.time()
    m.setup()
    setup_s = time.time() - t0
    npts = 8 * len(mesh["ele_tags"])
    iters, wall, _, p_y, cg, solves = newton_pushover(m, n, args.newton_steps)
    spmv_ms = m.time_op("spmv", 20)
    plastic = int(np.count_nonzero(m.point_data("plastic", npts)[:, 6] > 0))
    out = {
        "metric": "Newton iterations/s", "value": iters / wall, "unit": "iterations/s", "n_gpus": 1,
        "steps": args.newton_steps, "warmup": 0, "ms_per_step": 1e3 * wall / max(1, args.newton_steps),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"stdBrick J2 cube {n}^3 pushover to 1.5 p_y, LoadControl, "
                               f"{args.system} rtol 1e-10", "newton_iters": iters,
                   "cg_iters": cg, "solves": solves,
                   "spmv": args.spmv, "spmv_lanes": args.lanes, "spmv_ms": round(spmv_ms, 4),
                   "plastic_points": plastic, "points": npts, "setup_s": round(setup_s, 2),
                   "p_y": p_y, "final_load_factor": m.load_factor},
    }
    m.close()
    if not args.no_cpu:
        from tests import _oracle
        nc = 12
        res = {}
        for name in ("cpu", "gpu"):  # the GPU runs the steps the CPU sample ran
            mm = _oracle.oracle_model(3, 3) if name == "cpu" else Model(3, 3, device=0)
            meshes.build_brick_cube(mm, meshes.brick_cube(nc), pressure=1.0)
            if name == "cpu":
                mm.system(args.system, 1e-10, 50000)
            else:
                gpu_system(mm, args, 1e-10, 50000)
            mm.setup()
            it_c, wall_c, run_c = newton_pushover(
                mm, nc, args.newton_steps, max_seconds=args.cpu_seconds if name == "cpu" else None,
                max_steps=res["cpu"][2] if name == "gpu" else None, log=False)[:3]
            res[name] = (it_c, wall_c, run_c)
            mm.close()
            print(f"bench: newton {nc}^3 {name} {it_c} iterations in {wall_c:.2f}s ({run_c} steps)",
                  file=sys.stderr, flush=True)
        cpu_rate = res["cpu"][0] / res["cpu"][1]
        gpu_rate = res["gpu"][0] / res["gpu"][1]
        out["cpu_baseline"] = {
            "value": cpu_rate, "unit": "iterations/s", "cores": 1, "kind": "port",
            "sample": f"the same pushover on a {nc}^3 cube ({res['cpu'][2]} of {args.newton_steps} load steps, "
                      f"{res['cpu'][0]} Newton iterations), oracle/liboracle.so with its profile LDL^T direct "
                      f"solve, one thread"}
        out["same_sample_gpu"] = {"cube": nc, "value": gpu_rate, "unit": "iterations/s",
                                  "gpu_over_cpu": gpu_rate / cpu_rate}
    print(json.dumps(out))


def quad_main(args):
    """Plane-strain FourNodeQuad J2 strip 64x16 (20 x 5), clamped at x = 0,
    unit vertical reference load at the free-end top node,
    DisplacementControl of that node (dof 2, 0.001 per step, 50 steps),
    Newton + NormUnbalance 1e-8.  Newton iterations/s of the GPU path and of
    the CPU oracle (one core, direct solve) on the same pushover, and the
    largest relative difference of their converged displacements."""
    from tests import _oracle
    from xara_amd import meshes
    from xara_amd.model import Model

    mesh = meshes.quad_strip(64, 16)
    nsteps = 50

    def define(m):
        meshes.build_quad_strip(m, mesh)
        m.load(mesh["tip"], 0.0, 1.0)
        m.test("NormUnbalance", 1e-8, 25)
        m.integrator("DisplacementControl", mesh["tip"], 2, 0.001)

    res = {}
    for name, m in (("gpu", Model(2, 2)), ("cpu", _oracle.oracle_model(2, 2))):
        define(m)
        if name == "gpu":
            gpu_system(m, args, 1e-12, 100000)
        m.setup()
        c0 = m.counters()
        t0 = time.perf_counter()
        traj = []
        for _ in range(nsteps):
            if m.analyze(1) != 0:
                break
            traj.append(m.trial_disp().copy())
        wall = time.perf_counter() - t0
        c = m.counters()
        res[name] = (wall, c["newton_iters"] - c0["newton_iters"], np.array(traj))
        if name == "gpu":
            solver = {"solves": c["solves"] - c0["solves"], "pcg_iters": c["cg_iters"] - c0["cg_iters"],
                      "band_bw": c["band_bw"]}
        print(f"bench: quad {name} {wall:.2f}s newton={res[name][1]}", file=sys.stderr, flush=True)
    (gw, gi, gt), (cw, ci, ct) = res["gpu"], res["cpu"]
    rel = float(np.max(np.abs(gt - ct)) / np.max(np.abs(ct))) if len(gt) == len(ct) else None
    out = {
        "metric": "Newton iterations/s", "value": gi / gw, "unit": "iterations/s", "n_gpus": 1,
        "steps": nsteps, "warmup": 0, "ms_per_step": 1e3 * gw / nsteps, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": "FourNodeQuad J2 plane-strain strip 64x16, DisplacementControl pushover to 0.05",
                   "newton_iters": gi, "system": args.system, **solver, "max_rel_diff_vs_oracle": rel},
        "cpu_baseline": {"value": ci / cw, "unit": "iterations/s", "cores": 1, "kind": "port",
                         "sample": "the whole pushover, oracle direct solve"},
    }
    print(json.dumps(out))


def frame_bytes(nsteel, nconc, nsec):
    # steel fiber: committed history (10 doubles + kon) in, trial out = 168 B;
    # concrete fiber: committed ecmin, dept, eps, sig + previous trial sig, e
    # in (48 B), trial history out (40 B) = 88 B; section: deformation,
    # resultant and 4x4 tangent out = 192 B
    return nsteel * 168 + nconc * 88 + nsec * 192


def frame_field(fr, amp=60.0):
    z = fr["crds"][:, 2]
    h = z.max()
    U = np.zeros((len(z), 6))
    U[:, 0] = amp * (z / h) * (1.5 - 0.5 * z / h)
    U[:, 1] = 0.25 * amp * (z / h) ** 2
    U[:, 2] = -2e-4 * amp * z / h
    U[:, 4] = 1.5 * amp / h * (1 - z / h)
    return U


def frame_main(args):
    """Fiber update rate of the 3D RC moment frame (dispBeamColumn3d, 5 Gauss
    points, 100 Concrete02 + 20 Steel02 fibers per section)."""
    from xara_amd import meshes
    from xara_amd.model import Model

    nbx, nby, nst = (int(v) for v in args.frame.split(","))
    fr = meshes.rc_frame_3d(nbx, nby, nst)
    m = Model(3, 6, device=0)
    meshes.build_rc_frame_3d(m, fr)
    gpu_system(m, args, 1e-12, 200000)
    t0 = time.time()
    m.setup()
    setup_s = time.time() - t0
    nele = len(fr["columns"]) + len(fr["beams_x"]) + len(fr["beams_y"])
    nsec = 5 * nele
    nsteel, nconc = 20 * nsec, 100 * nsec
    m.set_trial_disp(frame_field(fr))
    m.time_op("update", args.warmup)
    t0 = time.perf_counter()
    ev_ms = m.time_op("update", args.steps)
    wall = time.perf_counter() - t0
    nfib = nsteel + nconc
    value = nfib * args.steps / wall
    byt = frame_bytes(nsteel, nconc, nsec)
    achieved = byt / (ev_ms * 1e-3) / 1e9
    out = {
        "metric": "fiber updates/s", "value": value, "unit": "fibers/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": 1e3 * wall / args.steps, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"3D RC frame {nbx}x{nby} bays x {nst} stories, dispBeamColumn3d update",
                   "elements": nele, "sections": nsec, "fibers": nfib, "neq": m.neq, "setup_s": round(setup_s, 2)},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None, "bytes_per_launch": byt,
                     "event_ms": ev_ms},
    }
    if args.frame_steps > 0:
        m.revert_to_start()
        m.test("NormUnbalance", 1e-3, 40)
        m.integrator("DisplacementControl", fr["roof"], 1, 0.03 * fr["height"] / 300.0)
        c0 = m.counters()
        t0 = time.perf_counter()
        for st in range(args.frame_steps):
            rc = m.analyze(1)
            c = m.counters()
            print(f"bench: frame step {st + 1} rc={rc} lambda={m.load_factor:.4g} newton="
                  f"{c['newton_iters'] - c0['newton_iters']} cg={c['cg_iters'] - c0['cg_iters']} "
                  f"t={time.perf_counter() - t0:.1f}s", file=sys.stderr, flush=True)
            if rc != 0:
                break
        wall = time.perf_counter() - t0
        c = m.counters()
        out["newton"] = {"steps": args.frame_steps, "newton_iters": c["newton_iters"] - c0["newton_iters"],
                         "cg_iters": c["cg_iters"] - c0["cg_iters"], "wall_s": wall,
                         "newton_iters_per_s": (c["newton_iters"] - c0["newton_iters"]) / wall}
    if not args.no_cpu:
        # the oracle's Domain::update over a 3 x 3 x 5 frame of the same
        # sections with the same kind of synthetic field, about --cpu-seconds
        from tests import _oracle
        frs = meshes.rc_frame_3d(3, 3, 5)
        o = _oracle.oracle_model(3, 6)
        meshes.build_rc_frame_3d(o, frs)
        o.setup()
        o.set_trial_disp(frame_field(frs))
        nf = 120 * 5 * (len(frs["columns"]) + len(frs["beams_x"]) + len(frs["beams_y"]))
        o.update()
        reps, t0 = 0, time.perf_counter()
        while True:
            o.update()
            reps += 1
            if time.perf_counter() - t0 > args.cpu_seconds:
                break
        cw = time.perf_counter() - t0
        o.close()
        out["cpu_baseline"] = {"value": nf * reps / cw, "unit": "fibers/s", "cores": 1, "kind": "port",
                               "sample": f"{reps} serial updates of a 3x3-bay, 5-story frame of the same "
                                         f"sections ({nf} fibers) with the same kind of synthetic field, "
                                         f"oracle/liboracle.so, one thread"}
        out["gpu_over_cpu"] = value / (nf * reps / cw)
    print(json.dumps(out))


def force_main(args):
    """Pushover of the 3D RC moment frame built from forceBeamColumn3d
    elements (5 Lobatto sections of the same 120-fiber RC section): Newton
    iterations/s of a DisplacementControl pushover on the GPU, and the same
    pushover on a reduced frame by the CPU oracle."""
    from xara_amd import meshes
    from xara_amd.model import Model

    def pushover(m, fr, steps):
        meshes.build_rc_frame_3d(m, fr, element="forceBeamColumn")
        gpu_system(m, args, 1e-12, 200000)
        m.test("NormUnbalance", 1e-3, 40)
        m.integrator("DisplacementControl", fr["roof"], 1, 0.03 * fr["height"] / 300.0)
        t0 = time.time()
        m.setup()
        setup_s = time.time() - t0
        c0 = m.counters()
        t0 = time.perf_counter()
        for st in range(steps):
            rc = m.analyze(1)
            if rc != 0:
                raise RuntimeError(f"force-based pushover step {st + 1} failed: {rc}")
        wall = time.perf_counter() - t0
        c = m.counters()
        return wall, c["newton_iters"] - c0["newton_iters"], c["cg_iters"] - c0["cg_iters"], setup_s

    nbx, nby, nst = (int(v) for v in args.frame.split(","))
    fr = meshes.rc_frame_3d(nbx, nby, nst)
    steps = args.frame_steps if args.frame_steps > 0 else 60
    gm = Model(3, 6, device=0)
    wall, it, cg, setup_s = pushover(gm, fr, steps)
    nele = len(fr["columns"]) + len(fr["beams_x"]) + len(fr["beams_y"])
    # sections whose axial stiffness fell below 95% of the elastic value
    ks = gm.point_data("tangent", 5 * nele, 16)[:, 0]
    k_el = ks.max()
    soft = float(np.mean(ks < 0.95 * k_el))
    out = {
        "metric": "Newton iterations/s", "value": it / wall, "unit": "iterations/s", "n_gpus": 1,
        "steps": steps, "warmup": 0, "ms_per_step": 1e3 * wall / steps, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"3D RC frame {nbx}x{nby} bays x {nst} stories of forceBeamColumn3d (5 Lobatto "
                               f"sections), DisplacementControl pushover to {steps * 0.01:.2f}% roof drift",
                   "elements": nele, "sections": 5 * nele, "fibers": 120 * 5 * nele, "newton_iters": it,
                   "cg_iters": cg, "setup_s": round(setup_s, 2), "system": args.system,
                   "softened_sections": soft},
    }
    if not args.no_cpu:
        from tests import _oracle
        small = meshes.rc_frame_3d(2, 2, 4)
        cwall, cit, _, _ = pushover(_oracle.oracle_model(3, 6), small, steps)
        out["cpu_baseline"] = {"value": cit / cwall, "unit": "iterations/s", "cores": 1, "kind": "port",
                               "sample": f"the same pushover of a 2x2x4 frame ({steps} steps), oracle direct solve"}
    print(json.dumps(out))


def ensemble_main(args):
    """Batched ensemble: model-steps/s of N uneven 2D RC frames advanced
    together by Newmark + Newton under their own synthetic records, against
    the same frames analysed one at a time on the GPU (looped single-model
    path) and on one CPU core (oracle)."""
    from xara_amd import ground_motion, meshes
    from xara_amd.ensemble import Ensemble
    from xara_amd.model import Model, XaraError

    n, dt = args.models, 0.01
    stories, bays, pga = meshes.frame_population(n)
    t0 = time.time()
    rec = ground_motion.kanai_tajimi_batch(np.arange(n), args.record_steps, dt, pga)
    print(f"bench: {n} records in {time.time() - t0:.1f}s", file=sys.stderr, flush=True)

    def define(m, k):
        fr = meshes.rc_frame_2d(int(stories[k]), int(bays[k]))
        meshes.define_transient_rc_frame_2d(m, fr, rec[k], dt)
        return fr

    t0 = time.time()
    ens = Ensemble()
    nsec = np.zeros(n, np.int64)
    for k in range(n):
        m = Model(2, 3)
        fr = define(m, k)
        nsec[k] = 5 * (len(fr["columns"]) + len(fr["beams"]))
        ens.add(m)
        m.close()
    ens.setup()
    setup_s = time.time() - t0
    print(f"bench: ensemble of {n} frames set up in {setup_s:.1f}s", file=sys.stderr, flush=True)
    byt_model = np.array([frame_bytes(20 * s, 100 * s, s) for s in nsec], np.float64)
    tune = None
    if args.warmup > 0:
        ens.analyze(args.warmup, dt)
    # the record envelope rises over its first 2 s: the timed window starts
    # at step ens_start (default 200), where the motion is strong
    done = args.warmup
    while done < args.ens_start:
        k = min(100, args.ens_start - done)
        ens.analyze(k, dt)
        done += k
    print(f"bench: ensemble advanced to step {done}", file=sys.stderr, flush=True)
    if not args.no_tune:
        table = ens.tune()
        best = table[np.argmin(table[:, 3])]
        tune = {"block": int(best[0]), "grid": int(best[1]), "order": int(best[2]), "ms": float(best[3]),
                "candidates": int(len(table))}
        print(f"bench: autotune {tune}", file=sys.stderr, flush=True)
    st0, it0 = ens.status()
    alive0 = st0 > -2
    nsteps = args.ens_steps
    t0 = time.perf_counter()
    done = 0
    while done < nsteps:
        k = min(50, nsteps - done)
        stopped = ens.analyze(k, dt)
        done += k
        print(f"bench: ensemble step {args.ens_start + done} stopped={stopped} t={time.perf_counter() - t0:.1f}s",
              file=sys.stderr, flush=True)
    wall = time.perf_counter() - t0
    st1, it1 = ens.status()
    newton = (it1 - it0).astype(np.float64)
    # a frame that stopped inside the timed region counts its steps up to the failure
    model_steps = float(np.sum(alive0)) * nsteps
    value = model_steps / wall
    updates_bytes = float(np.sum(byt_model * (newton + alive0 * nsteps)))
    achieved = updates_bytes / wall / 1e9

    # looped single-model GPU path on the first frames of the same population
    L = min(args.looped, n)
    t_loop, steps_loop = 0.0, 0
    for k in range(L):
        m = Model(2, 3)
        define(m, k)
        m.system("GpuBlockPCG", 1e-14, 100000)
        m.cg_config(16, bool(args.cg_graph), bool(args.cg_fuse), bool(args.cg_small), bool(args.cg_warm),
                    bool(args.cg_wide))
        m.setup()
        for _ in range(args.warmup):
            m.analyze(1, dt)
        t1 = time.perf_counter()
        ok = 0
        try:
            for _ in range(args.looped_steps):
                m.analyze(1, dt)
                ok += 1
        except XaraError:
            pass
        t_loop += time.perf_counter() - t1
        steps_loop += ok
        m.close()
    looped = steps_loop / t_loop if t_loop > 0 else None
    print(f"bench: looped single-model path {looped:.1f} model-steps/s", file=sys.stderr, flush=True)

    out = {
        "metric": "ensemble model-steps/s", "value": value, "unit": "model-steps/s", "n_gpus": 1,
        "steps": nsteps, "warmup": args.ens_start, "ms_per_step": 1e3 * wall / nsteps,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (seeded Kanai-Tajimi records, PGA 0.1-0.6 g)",
        "config": {"workload": f"ensemble of {n} 2D RC frames, stories 3-20, bays 1-6, Newmark dt {dt}, "
                               f"steps {args.ens_start}-{args.ens_start + nsteps} of {args.record_steps}",
                   "masses": "from a 100 N/mm gravity load on the beams (tributary half bays), tonnes",
                   "models": n, "sections": int(nsec.sum()), "fibers": int(120 * nsec.sum()),
                   "equations": int(sum(3 * (int(s) * (int(b) + 1)) for s, b in zip(stories, bays))),
                   "newton_iters_per_model_step": float(newton.sum() / max(model_steps, 1.0)),
                   "stopped": int(np.sum(st1 <= -2)), "stopped_before_window": int(np.sum(~alive0)),
                   "setup_s": round(setup_s, 1), "autotune": tune},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "note": "fiber-update bytes of every model update over the whole step time"},
        "looped_single_model": {"models": L, "steps_each": args.looped_steps, "model_steps_per_s": looped,
                                "ensemble_ratio": value / looped if looped else None},
    }
    if not args.no_cpu:
        out["cpu_baseline"] = ensemble_cpu_baseline(define, min(n, 8), args.cpu_seconds, dt)
    print(json.dumps(out))


def ensemble_cpu_baseline(define, nsample, seconds, dt):
    """The oracle (one core) stepping the first frames of the population."""
    from tests import _oracle
    from xara_amd.model import XaraError

    models = []
    for k in range(nsample):
        o = _oracle.oracle_model(2, 3)
        define(o, k)
        o.setup()
        models.append(o)
    t0 = time.perf_counter()
    steps = 0
    live = list(models)
    while live and time.perf_counter() - t0 < seconds:
        nxt = []
        for o in live:
            try:
                o.analyze(1, dt)
                steps += 1
                nxt.append(o)
            except XaraError:
                pass
        live = nxt
    wall = time.perf_counter() - t0
    return {"value": steps / wall, "unit": "model-steps/s", "cores": 1, "kind": "port",
            "sample": f"first {nsample} frames of the population, {steps} Newmark steps from rest"}


def adjoint_main(args):
    """Adjoint gradients of a 3D RC frame transient: loss = sum over the
    steps of the squared roof x displacements, parameters = per-element
    Steel02 fy, E, b and Concrete02 fc.  Reports adjoint (reverse sweep with
    checkpoint recompute) / forward wall time and checks the gradient along
    a random direction against central differences of the GPU forward."""
    from xara_amd import ground_motion, meshes
    from xara_amd.adjoint import FrameTransient
    from xara_amd.model import Model

    nbx, nby, nst = (int(v) for v in args.adj_frame.split(","))
    fr = meshes.rc_frame_3d(nbx, nby, nst)
    ne = len(meshes.frame_elements(fr)[0])
    dt, N = 0.01, args.adj_steps
    pga = float(meshes.frame_population(4096)[2][0])  # member 0 of the ensemble population
    ag = ground_motion.kanai_tajimi(0, N + 5, dt, pga)

    def build(m, theta):
        meshes.define_transient_frame(m, fr, theta, ag, dt, tol=1e-12, maxit=40)
        gpu_system(m, args, 1e-12, 200000)

    roof = meshes.roof_nodes(fr)
    prob = FrameTransient(3, 6, build, roof, N, dt, ckpt=args.ckpt)
    theta = meshes.default_theta(ne)
    # plain forward (no checkpoints), for the ratio
    m = Model(3, 6)
    build(m, theta)
    m.setup()
    t0 = time.perf_counter()
    for k in range(N):
        m.analyze(1, dt)
        if (k + 1) % 100 == 0:
            print(f"bench: plain forward step {k + 1} t={time.perf_counter() - t0:.1f}s", file=sys.stderr,
                  flush=True)
    plain_s = time.perf_counter() - t0
    m.close()
    t0 = time.perf_counter()
    loss = prob.run(theta)
    fwd_s = time.perf_counter() - t0
    print(f"bench: checkpointed forward {fwd_s:.2f}s loss={loss:.6e}", file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    grad = prob.gradient()
    bwd_s = time.perf_counter() - t0
    st = prob.stats()
    print(f"bench: adjoint sweep {bwd_s:.2f}s", file=sys.stderr, flush=True)
    # gradient check along a random direction: central differences (h = 1e-6 relative)
    rng = np.random.default_rng(7)
    d = rng.standard_normal((ne, 4)) * theta
    h = 1e-6
    Lp = prob.run(theta + h * d)
    Lm = prob.run(theta - h * d)
    fd = (Lp - Lm) / (2 * h)
    adj = float(np.sum(grad * d))
    out = {
        "metric": "adjoint/forward wall-time ratio", "value": (fwd_s + bwd_s) / plain_s, "unit": "x",
        "n_gpus": 1, "steps": N, "warmup": 0, "ms_per_step": 1e3 * (fwd_s + bwd_s) / N,
        "higher_is_better": False, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (seeded Kanai-Tajimi record)",
        "config": {"workload": f"3D RC frame {nbx}x{nby} bays x {nst} stories, {N}-step Newmark transient, "
                               f"loss = sum of squared roof x displacements",
                   "elements": ne, "parameters": 4 * ne, "checkpoint_every": args.ckpt, "pga_g": pga,
                   "plain_forward_s": plain_s, "checkpointed_forward_s": fwd_s, "adjoint_sweep_s": bwd_s,
                   "adjoint_sweep_over_forward": bwd_s / plain_s, "backward_cg_iters": st["backward_cg_iters"],
                   "loss": loss},
        "gradient_check": {"direction": "random (seed 7), relative to theta", "adjoint": adj,
                           "central_fd_h1e-6": fd, "rel_err": abs(adj - fd) / max(abs(fd), 1e-300)},
    }
    if args.ddm:
        from tests import _oracle
        o = _oracle.oracle_model(3, 6)
        meshes.define_transient_frame(o, fr, theta, ag, dt, tol=1e-12, maxit=40)
        o.setup()
        t0 = time.perf_counter()
        L, dL = _oracle.ddm_transient(o, N, dt, d[:, :3], d[:, 3], roof, 1)
        ddm_s = time.perf_counter() - t0
        out["gradient_check"].update({"oracle_ddm": dL, "rel_err_ddm": abs(adj - dL) / abs(dL),
                                      "oracle_loss": L})
        out["cpu_baseline"] = {"value": ddm_s, "unit": "s per directional derivative", "cores": 1,
                               "kind": "port",
                               "sample": f"oracle forward-mode (DDM) transient along one direction, {N} steps; "
                                         f"the full gradient needs {4 * ne} such runs"}
    prob.close()
    print(json.dumps(out))


if __name__ == "__main__":
    main()
