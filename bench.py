"""Benchmark of the finite-volume hot path on one MI355X.

Workload (the headline configuration): 2D Euler four-quadrant Riemann problem
(Riemann IC2), 1024 x 1024 cells on [0,1]^2, gamma 1.4, stencil-7 mesh, FP64.
One step = one seeded +-1% perturbation U*(1+delta) of the initial condition,
evaluated as WENO5 velocity() (problem A) plus first-order exact-Rusanov
jacobian() (problem B, 4x4 blocks, 5-point stencil), the two calls issued on
two streams, all inputs resident in HBM before the timed region.  Metric:
residual+Jacobian cells/s.

Prints one JSON line (the last line of stdout).
"""
import argparse
import json
import os
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for _p in (os.path.join(ROOT, "pressio-demoapps_amd"), os.path.join(ROOT, "oracle")):
    if _p not in sys.path:
        sys.path.insert(0, _p)

HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6


class KernelEvent:
    """A HIP timing event created with hipEventDisableSystemFence: recording
    it skips the system-scope fence (a cache writeback and invalidation) that
    a default event performs, which after a launch that writes 671 MB adds
    several microseconds to the measured span.  Created on the HIP runtime
    that torch and libpda share in this process; read after a device
    synchronize."""
    FLAGS = 0x20000000  # hipEventDisableSystemFence (hip_runtime_api.h)
    _hip = None

    @classmethod
    def hip(cls):
        if cls._hip is None:
            import ctypes
            path = next((l.split()[-1] for l in open("/proc/self/maps") if "libamdhip64" in l), None)
            if path is None:
                raise RuntimeError("the HIP runtime is not loaded in this process")
            h = ctypes.CDLL(path)
            h.hipEventCreateWithFlags.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_uint]
            h.hipEventRecord.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
            h.hipEventElapsedTime.argtypes = [ctypes.POINTER(ctypes.c_float), ctypes.c_void_p, ctypes.c_void_p]
            h.hipEventDestroy.argtypes = [ctypes.c_void_p]
            cls._hip = h
        return cls._hip

    def __init__(self):
        import ctypes
        self.e = ctypes.c_void_p()
        rc = self.hip().hipEventCreateWithFlags(ctypes.byref(self.e), self.FLAGS)
        if rc != 0:
            raise RuntimeError(f"hipEventCreateWithFlags failed ({rc})")

    def record(self, stream):
        rc = self.hip().hipEventRecord(self.e, stream.cuda_stream)
        if rc != 0:
            raise RuntimeError(f"hipEventRecord failed ({rc})")

    def elapsed_time(self, end):
        """milliseconds from this event to end (both complete)"""
        import ctypes
        ms = ctypes.c_float()
        rc = self.hip().hipEventElapsedTime(ctypes.byref(ms), self.e, end.e)
        if rc != 0:
            raise RuntimeError(f"hipEventElapsedTime failed ({rc})")
        return ms.value

    def __del__(self):
        if self._hip is not None and self.e:
            self._hip.hipEventDestroy(self.e)


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def traffic_bytes(n):
    """HBM bytes per launch of the Jacobian kernel as measured by rocprofv3
    PMC counters (FETCH_SIZE + WRITE_SIZE, scaled by a plain copy of known
    size), read from the committed profile summary for this grid; None when
    no measurement for this grid exists."""
    path = os.path.join(ROOT, "profiles", "jacobian_fo_euler2d_riemann_%dx%d_pmc.json" % (n, n))
    try:
        with open(path) as f:
            return json.load(f)["traffic_bytes_per_launch"]
    except (OSError, KeyError, ValueError):
        return None


def sod1d_ssprk3():
    """Sod shock tube, 1000 cells on [-0.5, 0.5] (mesh stencil 7), WENO5,
    SSPRK3 dt 1e-4 x 2000 steps (t = 0.2) with the device stepper, one
    hipGraph per step; returns the wall time and rates."""
    import torch
    import pressiodemoapps_amd as pda
    mesh = pda.create_structured_mesh((1000,), (-0.5, 0.5), 7)
    prob = pda.create_problem(mesh, pda.Euler1d.Sod, pda.InviscidFluxReconstruction.Weno5)
    U0 = prob.initialCondition()
    U = U0.clone()
    st = pda.ExplicitStepper(prob, pda.StepScheme.SSPRK3)
    # warm-up and graph capture on the tensor that is timed (the recorded
    # graph is tied to the state's address), then back to the initial state
    st.advance(U, 1e-4, 10)
    U.copy_(U0)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    st.advance(U, 1e-4, 2000)
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    return {"cells": 1000, "steps": 2000, "wall_s": el, "steps_per_s": 2000 / el,
            "us_per_step": el / 2000 * 1e6, "finite": bool(torch.isfinite(U).all().item())}


def extra_workloads():
    """The Sod 1D explicit workload (SSPRK3 x 2000), the batched shallow-water
    workload (384 slip-wall instances, RK4 x 500) and the implicit Burgers
    workload (512^2 BDF2 x 400 steps, Newton to ||R|| <= 1e-10 with the
    previous-step, extrapolated, predictor and learned guesses).  Reported
    beside the headline line; failures are reported, not fatal."""
    out = {}
    try:
        r = sod1d_ssprk3()
        out["sod1d_ssprk3"] = r
        log(f"[bench] Sod 1D SSPRK3: {r['us_per_step']:.1f} us per step, {r['wall_s']:.3f} s for 2000 steps")
    except Exception as e:  # reported, never fatal
        out["sod1d_ssprk3"] = {"error": str(e)}
        log(f"[bench] Sod 1D failed: {e}")
    try:
        from pressiodemoapps_amd import batched
        t0 = time.perf_counter()
        batch, _, sizes = batched.build(384, seed=7)
        setup = time.perf_counter() - t0
        r = batched.run(batch, sizes, steps=500, dt=0.01, log=log)
        r["setup_s"] = setup
        out["batched_swe"] = r
        log(f"[bench] batched SWE: {r['cell_steps_per_s']:.3e} cell-steps/s, {r['wall_s']:.2f} s for 500 RK4 steps")
        del batch
    except Exception as e:  # reported, never fatal
        out["batched_swe"] = {"error": str(e)}
        log(f"[bench] batched SWE failed: {e}")
    try:
        from pressiodemoapps_amd import learned
        if os.path.exists(learned.DEFAULT_WEIGHTS):
            net, meta = learned.load(learned.DEFAULT_WEIGHTS, 512)
            r = learned.compare(net, 512, 400, seed=77, log=log)
            r["training"] = meta
            out["implicit_burgers"] = r
            log(f"[bench] implicit Burgers: speedup over the previous-step guess: learned "
                f"{r['speedup_learned_vs_previous_step']:.2f} (order {r['learned'].get('order_mean', 0):.2f} on "
                f"average; {r['speedup_learned_vs_quadratic_predictor']:.3f} x the quadratic predictor, "
                f"{r['speedup_learned_vs_best_fixed_order']:.3f} x the best fixed order, {r['best_fixed_order']}); "
                f"FP64 predictors of order 1 / 2 / 3 / 4: "
                + " / ".join(f"{r[f'speedup_predictor_order_{k}_vs_previous_step']:.2f}" for k in (1, 2, 3, 4)))
        else:
            out["implicit_burgers"] = {"error": "no trained weights at " + learned.DEFAULT_WEIGHTS}
    except Exception as e:  # reported, never fatal
        out["implicit_burgers"] = {"error": str(e)}
        log(f"[bench] implicit Burgers failed: {e}")
    return out


def cpu_baseline_child(meshdir, reps):
    """Runs in a child process that never imports torch (so the reference's
    OpenMP runtime starts with this process's OMP_* settings): the
    reference's own rightHandSide (WENO5, problem A), jacobian and
    rightHandSideAndJacobian (first order, problem B) on a state, velocity
    and Jacobian created once, as a caller of adapter_cpp.hpp:105-229 runs
    them; one warm-up call then `reps` timed calls each.  Prints one JSON
    line with the per-call times."""
    import numpy as np
    import pyoracle
    path, level = pyoracle.ref_omp_build()
    t0 = time.perf_counter()
    ra = pyoracle.RefProblem(meshdir, 2, 4, 2, 2, omp=True)
    rb = pyoracle.RefProblem(meshdir, 2, 4, 0, 2, omp=True)
    setup = time.perf_counter() - t0
    U0 = ra.initial_condition()
    U = U0 * (1.0 + 0.01 * np.random.default_rng(0).uniform(-1, 1, U0.shape))
    out = {"rhs": ra.time_calls(U, 0, reps).tolist(), "jacobian": rb.time_calls(U, 1, reps).tolist(),
           "rhs_and_jacobian": rb.time_calls(U, 2, reps).tolist(), "setup_s": setup,
           "build": os.path.basename(path), "isa": level}
    print(json.dumps(out), flush=True)


def cpu_baseline(n_cpu, reps=5):
    """The reference CPU path (oracle/_ref, OpenMP, the fastest x86-64 level
    this host runs) timed on this host's cores on the n_cpu x n_cpu Riemann
    workload: cells / (median rightHandSide + median jacobian) per
    perturbation.  Without the reference build, the C restatement (one
    thread) stands in."""
    import numpy as np
    import pyoracle
    import pressiodemoapps_amd as pda
    mesh = pda.create_structured_mesh((n_cpu, n_cpu), (0.0, 1.0, 0.0, 1.0), 7)
    cells = n_cpu * n_cpu
    if pyoracle.ref_available(omp=True):
        import subprocess
        d = tempfile.mkdtemp(prefix="pda_bench_mesh_")
        mesh.write(d)
        threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        env = dict(os.environ, OMP_NUM_THREADS=str(threads), OMP_PROC_BIND="true")
        r = subprocess.run([sys.executable, os.path.abspath(__file__), "--cpu-child", d, "--cpu-reps", str(reps)],
                           env=env, capture_output=True, text=True, timeout=600)
        if r.returncode != 0:
            raise RuntimeError("reference timing failed: " + r.stderr[-2000:])
        t = json.loads(r.stdout.strip().splitlines()[-1])
        med = {k: float(np.median(t[k])) for k in ("rhs", "jacobian", "rhs_and_jacobian")}
        per = med["rhs"] + med["jacobian"]
        return {"value": cells / per, "unit": "cells/s", "cores": threads, "kind": "reference",
                "sample": f"{n_cpu}x{n_cpu} Riemann IC2, the full workload size: median of {reps} calls "
                          f"(after a warm-up) of the reference's WENO5 rightHandSide "
                          f"{med['rhs'] * 1e3:.1f} ms + first-order jacobian {med['jacobian'] * 1e3:.1f} ms "
                          f"per perturbation, V and J created once; {threads} OpenMP threads, "
                          f"{t['isa']} build",
                "rhs_ms": med["rhs"] * 1e3, "jacobian_ms": med["jacobian"] * 1e3,
                "rhs_and_jacobian_ms": med["rhs_and_jacobian"] * 1e3,
                "rhs_cells_per_s": cells / med["rhs"], "jacobian_cells_per_s": cells / med["jacobian"],
                "setup_s": t["setup_s"], "build": t["build"]}
    orc = pyoracle.Oracle()
    ms = orc.from_mesh(mesh)
    import ctypes
    pa = pda.create_problem(mesh, pda.Euler2d.Riemann, pda.InviscidFluxReconstruction.Weno5, 2)
    U0 = np.zeros(pa.totalDofStencilMesh())
    pda.lib().pda_initial_condition_host(pa.handle, U0.ctypes.data_as(ctypes.c_void_p))
    U = U0 * (1.0 + 0.01 * np.random.default_rng(0).uniform(-1, 1, U0.shape))
    pba, pbb = orc.problem(2, 4, 2, [1.4]), orc.problem(2, 4, 0, [1.4])
    tv, tj = [], []
    for r in range(reps + 1):
        t0 = time.perf_counter()
        orc.velocity(ms, pba, U)
        t1 = time.perf_counter()
        orc.velocity_and_jacobian(ms, pbb, U)
        t2 = time.perf_counter()
        if r:
            tv.append(t1 - t0)
            tj.append(t2 - t1)
    per = float(np.median(tv) + np.median(tj))
    return {"value": cells / per, "unit": "cells/s", "cores": 1, "kind": "port",
            "sample": f"{n_cpu}x{n_cpu} Riemann IC2: median of {reps} calls of the C restatement's WENO5 "
                      f"velocity + first-order Jacobian, one thread",
            "rhs_ms": float(np.median(tv)) * 1e3, "jacobian_ms": float(np.median(tj)) * 1e3}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--n", type=int, default=1024)
    ap.add_argument("--cpu-n", type=int, default=1024)
    ap.add_argument("--cpu-reps", type=int, default=5)
    ap.add_argument("--cpu-child", default=None, help=argparse.SUPPRESS)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extra", action="store_true",
                    help="skip the batched shallow-water and implicit Burgers measurements")
    ap.add_argument("--copy-calib", action="store_true",
                    help="after timing, run one plain 256 MiB device copy (the reference "
                         "kernel for scaling rocprofv3 FETCH_SIZE/WRITE_SIZE counters)")
    args = ap.parse_args()
    if args.cpu_child:
        cpu_baseline_child(args.cpu_child, args.cpu_reps)
        return
    if args.gpus != 1:
        print(json.dumps({"error": "one GPU is measured here; --gpus must be 1 (several GPUs would "
                                   "each own a slab of the mesh with its halo rows)"}))
        sys.exit(2)

    import numpy as np
    import torch
    import pressiodemoapps_amd as pda

    assert torch.cuda.is_available(), "bench needs a HIP device"
    n = args.n
    log(f"[bench] building {n}x{n} mesh and problems")
    mesh = pda.create_structured_mesh((n, n), (0.0, 1.0, 0.0, 1.0), 7)
    probA = pda.create_problem(mesh, pda.Euler2d.Riemann, pda.InviscidFluxReconstruction.Weno5, 2)
    probB = pda.create_problem(mesh, pda.Euler2d.Riemann, pda.InviscidFluxReconstruction.FirstOrder, 2)
    cells = n * n
    m = probA.numDofPerCell()
    U0 = probA.initialCondition()
    nst = probA.totalDofStencilMesh()
    npert = min(args.steps + args.warmup, 100)
    states = []
    for i in range(npert):
        g = torch.Generator(device="cuda").manual_seed(i)
        d = torch.rand(nst, dtype=torch.float64, device="cuda", generator=g) * 2.0 - 1.0
        states.append(U0 * (1.0 + 0.01 * d))
    V = probA.createRightHandSide()
    J = probB.createJacobian()
    nnz = J.nonZeros()
    torch.cuda.synchronize()
    # set-up, outside the timed region: the velocity kernel's strip height
    # measured on this problem (the values do not depend on it)
    rps, rps_us = probA.autotune(states[0])
    log(f"[bench] velocity strip height {rps} ({rps_us:.1f} us per call while tuning)")

    # One step issues problem A's rightHandSide on one stream and problem
    # B's jacobian on another, as a caller may (the calls read the same
    # state and write different outputs; each stream keeps its own calls in
    # order): the velocity waves fill the compute units that the
    # store-bound Jacobian leaves idle.
    sV, sJ = torch.cuda.Stream(), torch.cuda.Stream()

    def step(i):
        U = states[i % npert]
        with torch.cuda.stream(sV):
            probA.rightHandSide(U, 0.0, V)
        with torch.cuda.stream(sJ):
            probB.jacobian(U, 0.0, J)

    def step_serial(i):
        U = states[i % npert]
        probA.rightHandSide(U, 0.0, V)
        probB.jacobian(U, 0.0, J)

    # autotune() times back-to-back calls on one state, which stays in cache;
    # in the steps each velocity reads a state that the previous step's
    # 671 MB Jacobian write has evicted, and shares the GPU with a Jacobian,
    # so the velocity's strip height is also timed in this loop, jointly
    # with the Jacobian's (taller Jacobian strips mean fewer Jacobian waves,
    # which leaves room for velocity waves on the SIMDs); untimed set-up:
    # two passes of 3 + 12 steps per pair, best pass kept.
    def ctx_us(hv, hj):
        probA.set_velocity_strip_height(hv)
        probB.set_jacobian_strip_height(hj)
        best = 1e300
        for _ in range(2):
            for i in range(3):
                step(i)
            torch.cuda.synchronize()
            t = time.perf_counter()
            for i in range(12):
                step(3 + i)
            torch.cuda.synchronize()
            best = min(best, (time.perf_counter() - t) / 12 * 1e6)
        return best
    vcands = sorted({8, 10, 12, 14, 16, 20, rps})
    jcands = (0, 4, 6, 8, 10)  # 0: the kernel's own rule (2 rows at 1024^2)
    ctx = {(hv, hj): ctx_us(hv, hj) for hj in jcands for hv in vcands}
    rps, jrps = min(ctx, key=ctx.get)
    probA.set_velocity_strip_height(rps)
    probB.set_jacobian_strip_height(jrps)
    for hj in jcands:
        log(f"[bench] step loop, jacobian strip height {hj or 'rule'}: "
            + ", ".join(f"{hv}: {ctx[(hv, hj)]:.0f} us" for hv in vcands))
    log(f"[bench] strip heights in the step loop: velocity {rps}, jacobian {jrps or 'rule'}")

    log(f"[bench] warmup {args.warmup}")
    for i in range(args.warmup):
        step(i)
    torch.cuda.synchronize()

    log(f"[bench] timing {args.steps} steps")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(args.steps):
        step(args.warmup + i)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0

    # the same steps with both calls on one stream, for reference
    for i in range(3):
        step_serial(i)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(args.steps):
        step_serial(args.warmup + i)
    torch.cuda.synchronize()
    wall_serial = time.perf_counter() - t0

    # the roofline kernel in its own pass, at its own strip-height rule (the
    # height tuned for the two-stream steps trades the kernel's own speed
    # for room beside it): jacobian() calls alone, each between fence-free
    # events on the stream it runs on, each on another perturbation (the
    # first call, after the synchronize, is not sampled)
    probB.set_jacobian_strip_height(0)
    s = torch.cuda.current_stream()
    nj = min(args.steps, 20)
    evj = [(KernelEvent(), KernelEvent()) for _ in range(nj)]
    probB.jacobian(states[0], 0.0, J)
    for i in range(nj):
        evj[i][0].record(s)
        probB.jacobian(states[(i + 1) % npert], 0.0, J)
        evj[i][1].record(s)
    torch.cuda.synchronize()
    t_j = sum(a.elapsed_time(b) for a, b in evj) / nj * 1e-3
    # velocity() alone, likewise
    nv = min(args.steps, 20)
    evv = [(KernelEvent(), KernelEvent()) for _ in range(nv)]
    for i in range(nv):
        evv[i][0].record(s)
        probA.rightHandSide(states[i % npert], 0.0, V)
        evv[i][1].record(s)
    torch.cuda.synchronize()
    t_v = sum(a.elapsed_time(b) for a, b in evv) / nv * 1e-3
    ms_per_step = wall / args.steps * 1e3

    # algorithmic bytes: velocity 8*m*(Nst+N); Jacobian 8*nnz + 8*m*Nst
    bytes_v = 8 * m * (cells + cells)
    bytes_j = 8 * nnz + 8 * m * cells
    ach_j = bytes_j / t_j / 1e9
    value = cells / (wall / args.steps)
    log(f"[bench] velocity {t_v*1e6:.1f} us ({bytes_v/t_v/1e9:.0f} GB/s), jacobian {t_j*1e6:.1f} us "
        f"({ach_j:.0f} GB/s), step {ms_per_step*1e3:.1f} us on two streams "
        f"({wall_serial / args.steps * 1e6:.1f} us on one), unit bytes {(bytes_v+bytes_j)/1e6:.0f} MB "
        f"-> {(bytes_v+bytes_j)/(wall/args.steps)/1e9:.0f} GB/s")

    if args.copy_calib:
        src = torch.empty(32 * 1024 * 1024, dtype=torch.float64, device="cuda").fill_(1.0)
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        log("[bench] copy calibration: 256 MiB read + 256 MiB written")

    extra = {}
    if not args.no_extra:
        extra = extra_workloads()

    cpu = None
    if not args.no_cpu:
        log("[bench] cpu baseline")
        try:
            cpu = cpu_baseline(args.cpu_n, args.cpu_reps)
        except Exception as e:  # reported, never fatal
            log(f"[bench] cpu baseline failed: {e}")
    out = {
        "metric": "residual+Jacobian cells/s",
        "value": value,
        "unit": "cells/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic: Riemann IC2 with seeded +-1% uniform perturbations",
        "config": {"workload": f"2D Euler Riemann IC2 {n}x{n}, WENO5 velocity + first-order "
                               f"jacobian per perturbation", "cells": cells, "jacobian_nnz": nnz,
                   "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": ach_j, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": ach_j / HBM_PEAK_GBS, "traffic": traffic_bytes(n),
                     "kernel": "k_jac_fo_euler2d (first-order Euler 2D Jacobian, box cells and "
                               "boundary ring in one launch; the whole jacobian() call)",
                     "bytes_per_launch": bytes_j, "avg_launch_us": t_j * 1e6},
        "unit_bytes_per_step": bytes_v + bytes_j,
        "hbm_frac_per_step": (bytes_v + bytes_j) / (wall / args.steps) / 1e9 / HBM_PEAK_GBS,
        "velocity_us": t_v * 1e6,
        "velocity_rows_per_strip": rps,
        "jacobian_rows_per_strip_in_steps": jrps or "kernel rule",
        "streams": "velocity and jacobian on two streams",
        "serial_ms_per_step": wall_serial / args.steps * 1e3,
        "serial_value": cells / (wall_serial / args.steps),
        "cpu_baseline": cpu,
    }
    out.update(extra)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
