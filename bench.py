"""Benchmark of the diffusion field solve on one MI355X.

Workload (default): the 2D multi-field case, a 1024^2 plane with ~40k Voronoi
cells in a radius-450 disc and three fields (oxygen, secreted signal,
chemokine), each solved from its initial state to an FP64 relative residual
of 1e-8. One step = the three fields solved together (one batched launch
sequence) in FP32 smoothing + FP64 defect correction. The plain FP64 solve of
the same fields is timed beside it.

value = voxel-sweeps/s of the FP32+refine solve: sum over fields of (swept
voxels x smoothing sweeps), divided by the wall time of the step. Inputs are
resident in HBM: each step resets the device values with cw_field_fill.

Prints one JSON line. --gpus N with N > 1 exits with a message: the measurement is of one GPU.
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

HBM_PEAK_GBS = 8000.0


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def optimal_omega(fc):
    s = fc.span()
    longest = max(s[1] - s[0] + 1, s[3] - s[2] + 1, s[5] - s[4] + 1)
    return 2.0 / (1.0 + math.sin(math.pi / max(longest, 2)))


def bytes_per_voxel_pass(fc, precision):
    """Algorithmic HBM bytes per swept voxel per sweep-kernel launch.
    FP64 plain: read c 8 + write c 8 + folded coefficient 8 = 24.
    FP32 smoother: read e 4 + write e 4 + defect RHS 4 = 12, plus the
    correction's lower bound 4 for a field with sinks (the floor can bind)."""
    if precision == "fp64":
        return 24.0
    has_sink = bool((fc.uptake > 0).any())
    return 16.0 if has_sink else 12.0


def copy_bandwidth(gib=1.0):
    """Attainable HBM bandwidth: a device-to-device copy of `gib` GiB (read +
    write bytes over the copy kernel's time, torch events on its stream)."""
    try:
        import torch

        if not torch.cuda.is_available():
            return None
        n = int(gib * (1 << 30)) // 4
        a = torch.empty(n, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        a.fill_(1.0)
        b.copy_(a)
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        reps = 10
        e0.record()
        for _ in range(reps):
            b.copy_(a)
        e1.record()
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / reps
        del a, b
        return 2.0 * n * 4 / (ms * 1e-3) / 1e9
    except Exception:
        return None


def cpu_model():
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def host_threads():
    """The host threads this process may use: its CPU affinity, capped by
    OMP_NUM_THREADS (the CPU share a machine grants one process)."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    cap = int(os.environ.get("OMP_NUM_THREADS", str(n)))
    return max(1, min(n, cap))


def cpu_baseline(fc, sweeps_to_tol=None, omega=None, samples=5, sample_s=1.5, threads=None,
                 tol_label="1e-8"):
    """The oracle (C restatement of pde.rs: FP64, read-old/write-new
    buffers, OpenMP over rows in chunks of 4) on this host, on the oxygen
    field of the bench workload at the bench's omega:
      - the sweep rate, median of `samples` samples of about sample_s seconds
        each after one warm-up sample, on all the host threads this process
        may use and on one thread;
      - time to the relative residual of 1e-8: the sweeps the FP64 solve needs
        (the GPU's FP64 sweep count, which the oracle reproduces bit for bit),
        timed on all threads from the initial state, then the oracle's own
        residual of the result."""
    from oracle.orf import OracleField

    threads = host_threads() if threads is None else threads
    omega = optimal_omega(fc) if omega is None else omega
    span = fc.span_voxels()

    def fresh():
        f = OracleField(fc.w, fc.h, fc.d, fc.diff, fc.decay, fc.initial)
        f.values[:] = fc.initial_values()
        f.copy_next_from_values()
        f.uptake[:] = fc.uptake
        f.production[:] = fc.production
        f.held[:] = fc.held
        return f

    def rate(nthreads):
        f = fresh()
        sp = f.free_span()
        out = []
        for k in range(samples + 1):
            n, t0 = 0, time.perf_counter()
            while time.perf_counter() - t0 < sample_s:
                for _ in range(10):
                    f.sor_sweep(omega, nthreads, sp)
                n += 10
            dt = time.perf_counter() - t0
            if k:  # the first sample is the warm-up
                out.append(span * n / dt)
        return float(sorted(out)[len(out) // 2]), out

    dims = f"{fc.w}x{fc.h}" + (f"x{fc.d}" if fc.d > 1 else "")
    all_rate, all_samples = rate(threads)
    one_rate, one_samples = rate(1)
    res = {"value": all_rate, "unit": "voxel-sweeps/s", "cores": threads, "kind": "port",
           "cpu": cpu_model(), "nproc": os.cpu_count(),
           "sample": f"FP64 red-black sweeps of the {dims} {fc.name} field ({span} swept "
                     f"voxels), median of {samples} samples of ~{sample_s:.0f} s after a "
                     f"warm-up, CPU restatement of pde.rs (the Rust binary cannot be built here)",
           "samples": all_samples,
           "single_thread": {"value": one_rate, "samples": one_samples}}
    if sweeps_to_tol:
        f = fresh()
        sp = f.free_span()
        t0 = time.perf_counter()
        for _ in range(sweeps_to_tol):
            f.sor_sweep(omega, threads, sp)
        dt = time.perf_counter() - t0
        res["time_to_" + tol_label] = {"seconds": dt, "sweeps": sweeps_to_tol, "cores": threads,
                                       "rel_residual": f.rel_residual()[0],
                                       "field": f"{fc.name}, {fc.w}x{fc.h}x{fc.d}"}
    return res


def cpu_campaign(flat, sweeps, picks=8, threads=None):
    """The oracle on a sample of the campaign: `picks` histories spread over
    the size range (sorted by edge, evenly spaced), each field swept the
    number of FP64 sweeps its GPU solve took to 1e-8, on all host threads;
    the whole campaign's CPU time is estimated from the sample's voxel-sweep
    rate and the campaign's total FP64 voxel-sweeps."""
    from cellweave_amd.campaign import omega_for
    from oracle.orf import OracleField

    threads = host_threads() if threads is None else threads
    nh = len(flat) // 2
    order = sorted(range(nh), key=lambda h: flat[2 * h].w)
    hs = [order[round(i * (nh - 1) / max(1, picks - 1))] for i in range(picks)]
    vs, secs = 0, 0.0
    for h in hs:
        for k in (2 * h, 2 * h + 1):
            fc = flat[k]
            f = OracleField(fc.w, fc.h, fc.d, fc.diff, fc.decay, fc.initial)
            f.values[:] = fc.initial_values()
            f.copy_next_from_values()
            f.uptake[:] = fc.uptake
            f.production[:] = fc.production
            f.held[:] = fc.held
            sp = f.free_span()
            omega = omega_for(fc)  # the per-field omega of the GPU solve
            t0 = time.perf_counter()
            for _ in range(sweeps[k]):
                f.sor_sweep(omega, threads, sp)
            secs += time.perf_counter() - t0
            vs += fc.span_voxels() * sweeps[k]
    total = sum(fc.span_voxels() * s for fc, s in zip(flat, sweeps))
    rate = vs / secs
    return {"value": rate, "unit": "voxel-sweeps/s", "cores": threads, "kind": "port",
            "cpu": cpu_model(), "nproc": os.cpu_count(),
            "sample": f"{picks} histories ({2 * picks} fields, edges "
                      f"{min(flat[2 * h].w for h in hs)}..{max(flat[2 * h].w for h in hs)}) "
                      "swept their GPU FP64 sweep counts to 1e-8, CPU restatement of pde.rs",
            "sample_seconds": secs,
            "campaign_seconds_estimate": total / rate}


def cpu_for_workload(name, line):
    """The CPU restatement beside a workload line, on its representative
    field (the campaign: a sample of histories)."""
    from cellweave_amd import cases

    if name == "oxygen200":
        (fc,) = cases.oxygen_disc_200()
        return cpu_baseline(fc, sweeps_to_tol=line["fp64"]["sweeps"][0], omega=1.9)
    if name == "spheroid256":
        (fc,) = cases.spheroid()
        return cpu_baseline(fc, sweeps_to_tol=line["fp64"]["sweeps"][0],
                            omega=line["config"]["omega"], samples=3)
    if name == "adjoint192":
        fc, _ = cases.adjoint_spheroid()
        return cpu_baseline(fc, sweeps_to_tol=line["forward"]["fp64_sweeps"], omega=line["omega"],
                            samples=3, tol_label="1e-13")
    if name == "campaign512":
        hist = cases.campaign_histories(count=512)
        return cpu_campaign([fc for h in hist for fc in h], line["_fp64_sweeps"])
    return None


def cpu_tissue_steps(gpu_sweeps, steps=4, threads=None):
    """The same lattice-step loop on the oracle (C restatement of pde.rs and
    coupling.rs, OpenMP over rows): the warm-up solve, then the first `steps`
    lattice steps, each timed. The sweep counts sit beside the GPU's."""
    import numpy as np

    from cellweave_amd import cases, workloads as wl
    from oracle.orf import OracleField

    if threads is None:
        threads = host_threads()
    seq, ncells = wl.tissue_steps(steps=steps)
    n = 120
    consumes = np.ones(ncells + 1, np.uint8)
    consumes[0] = 0
    f = OracleField(n, n, n, cases.d_px(2000.0), 0.0, 1.0)

    def declare(ids):
        f.apply_uptake(ids, consumes, wl.TISSUE_UPTAKE)
        f.release_all()
        f.hold_bath_at(1.0, (ids == 0).astype(np.uint8))

    declare(seq[0])
    f.relax(*wl.TISSUE_WARMUP, threads=threads)
    f.mean_per_cell(seq[0], ncells + 1)
    ts, sweeps = [], []
    for ids in seq[1:]:
        t0 = time.perf_counter()
        declare(ids)
        sweeps.append(f.relax(wl.TISSUE_TOLERANCE * wl.TISSUE_UPTAKE, wl.TISSUE_MAX_SWEEPS,
                              threads=threads))
        f.mean_per_cell(ids, ncells + 1)
        ts.append(time.perf_counter() - t0)
    t = float(np.median(ts))
    return {"value": 1.0 / t, "unit": "steps/s", "cores": threads, "kind": "port",
            "sample": f"warm-up solve, then {len(ts)} lattice steps of the {n}^3 tissue "
                      f"(median {1e3 * t:.0f} ms a step), CPU restatement of pde.rs and coupling.rs",
            "sweeps_per_step": sweeps,
            "same_sweeps_as_gpu": sweeps == list(gpu_sweeps[:len(sweeps)])}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=1024)
    ap.add_argument("--spp", type=int, default=0, help="FP32 sweeps per pass (0 = library)")
    ap.add_argument("--spp64", type=int, default=0)
    ap.add_argument("--shape32", type=int, default=-1,
                    help="FP32 smoother tile shape (-1 = library); overrides --spp")
    ap.add_argument("--inner-tol", type=float, default=0.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--time-launches", type=int, default=200)
    ap.add_argument("--workload", default="multifield1024",
                    help="multifield1024 (the default line) or oxygen200, spheroid256, campaign512, "
                         "adjoint192, coupling, tissue120")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (a batch of "
              "fields would shard across GPUs with no exchange, a single field by slabs with a "
              "halo exchange per launch)", file=sys.stderr)
        sys.exit(2)

    import numpy as np

    if args.workload != "multifield1024":
        from cellweave_amd import workloads

        line = workloads.WORKLOADS[args.workload]()
        if not args.no_cpu:
            log("cpu baseline")
            if args.workload == "tissue120":
                line["cpu_baseline"] = cpu_tissue_steps(line["sweeps_per_step"])
            else:
                line["cpu_baseline"] = cpu_for_workload(args.workload, line)
        line.pop("_fp64_sweeps", None)
        print(json.dumps(line))
        return

    from cellweave_amd import cases
    from cellweave_amd.field import (CRIT_REL_RESIDUAL, PREC_FP32_REFINE, PREC_FP64, Field,
                                     batch_relax, time_passes)

    try:
        import torch
        sync = torch.cuda.synchronize if torch.cuda.is_available() else (lambda: None)
    except Exception:  # torch is plumbing only; the solve does not need it
        sync = lambda: None  # noqa: E731

    log(f"building workload: {args.n}^2 plane, three fields")
    fcs = cases.multifield_plane(n=args.n)
    fields = []
    for fc in fcs:
        g = Field(fc.w, fc.h, fc.d, fc.diff, fc.decay, fc.initial)
        g.upload(values=fc.initial_values(), held=fc.held, uptake=fc.uptake,
                 production=fc.production)
        fields.append(g)
    omega = optimal_omega(fcs[0])
    span = [fc.span_voxels() for fc in fcs]
    tol = 1e-8

    def reset():
        for g, fc in zip(fields, fcs):
            g.fill(fc.initial)

    def shape_kw(precision):
        if precision == PREC_FP32_REFINE and args.shape32 >= 0:
            return {"tile_shape": args.shape32, "sweeps_per_pass": 0}
        return {"sweeps_per_pass": args.spp if precision == PREC_FP32_REFINE else args.spp64}

    def solve(precision, spp):
        reset()
        sync()
        t0 = time.perf_counter()
        stats, _ = batch_relax(fields, criterion=CRIT_REL_RESIDUAL, tolerance=tol,
                               max_sweeps=2_000_000, omega=omega, precision=precision,
                               inner_tolerance=args.inner_tol, **shape_kw(precision))
        sync()
        return time.perf_counter() - t0, stats

    # FP64 plain: one warm-up, then timed solves.
    log("fp64 warm-up")
    solve(PREC_FP64, args.spp64)
    t64s, st64 = [], None
    for _ in range(max(1, args.steps // 3)):
        t, st64 = solve(PREC_FP64, args.spp64)
        t64s.append(t)
    t64 = float(np.median(t64s))
    vs64 = sum(s * st["sweeps"] for s, st in zip(span, st64))
    x64 = [g.download() for g in fields]
    log(f"fp64: {t64 * 1e3:.2f} ms, sweeps {[st['sweeps'] for st in st64]}, "
        f"rel {[st['rel_residual'] for st in st64]}")

    log("fp32+refine warm-up")
    for _ in range(args.warmup):
        solve(PREC_FP32_REFINE, args.spp)
    sync()
    times, st32 = [], None
    t_start = time.perf_counter()
    for _ in range(args.steps):
        t, st32 = solve(PREC_FP32_REFINE, args.spp)
        times.append(t)
    sync()
    total = time.perf_counter() - t_start
    ms_step = 1e3 * total / args.steps
    vs32 = sum(s * st["sweeps"] for s, st in zip(span, st32))
    value = vs32 / (total / args.steps)
    x32 = [g.download() for g in fields]
    errs = {}
    for fc, a, b, st in zip(fcs, x32, x64, st32):
        scale = float(np.max(np.abs(b))) or 1.0
        errs[fc.name] = {"max_abs": float(np.max(np.abs(a - b))),
                         "rel_l2": float(np.linalg.norm(a - b) / (np.linalg.norm(b) or 1.0)),
                         "max_rel_to_peak": float(np.max(np.abs(a - b)) / scale),
                         "rel_residual": st["rel_residual"], "sweeps": st["sweeps"],
                         "outer": st["outer_iters"]}
    log(f"fp32+refine: {ms_step:.2f} ms/step, sweeps {[st['sweeps'] for st in st32]}, "
        f"outer {[st['outer_iters'] for st in st32]}")

    # Live kernel timing of the dominant kernel (the FP32 smoother launch).
    spp_used = st32[0]["sweeps_per_pass"]
    # waves and rows a wave of the register-blocked FP32 shapes (S, NW, NR)
    nw32, nr32 = {15: (16, 8), 18: (16, 8), 19: (16, 8), 20: (16, 8), 21: (8, 8), 22: (8, 8),
                  23: (8, 8), 24: (16, 8)}.get(st32[0]["tile_shape"], ("?", "?"))
    ms32 = time_passes(fields, args.time_launches, criterion=CRIT_REL_RESIDUAL, tolerance=tol,
                       omega=omega, precision=PREC_FP32_REFINE, **shape_kw(PREC_FP32_REFINE))
    bytes32 = sum(s * bytes_per_voxel_pass(fc, "fp32") for s, fc in zip(span, fcs))
    achieved = bytes32 / (ms32 * 1e-3) / 1e9
    reset()
    ms64 = time_passes(fields, args.time_launches, criterion=CRIT_REL_RESIDUAL, tolerance=tol,
                       omega=omega, precision=PREC_FP64, **shape_kw(PREC_FP64))
    spp64_used = st64[0]["sweeps_per_pass"]
    bytes64 = sum(s * 24.0 for s in span)
    achieved64 = bytes64 / (ms64 * 1e-3) / 1e9
    log(f"kernel: fp32 {ms32 * 1e3:.1f} us/launch ({achieved:.0f} GB/s alg), "
        f"fp64 {ms64 * 1e3:.1f} us/launch ({achieved64:.0f} GB/s alg)")

    cpu = None
    if not args.no_cpu:
        log("cpu baseline")
        cpu = cpu_baseline(fcs[0], sweeps_to_tol=st64[0]["sweeps"], omega=omega)

    # Committed counter summaries describe one kernel shape on one grid: they
    # are attached only when the timed smoother is that shape on that grid.
    def committed(name):
        path = os.path.join(ROOT, "profiles", name)
        if not os.path.exists(path):
            return None
        with open(path) as fh:
            d = json.load(fh)
        return d if d.get("sweeps_per_launch") == spp_used and d.get("n") == args.n else None

    traffic_json = committed("traffic_fp32_smoother_multifield_1024.json")
    valu_json = committed("valu_fp32_smoother_multifield_1024.json")

    # The solve's algorithmic bytes: each smoother launch moves
    # its per-voxel pass bytes once for s sweeps; each outer iteration adds the
    # FP64 residual (24 B a voxel) and correction (20 B); the first residual
    # comes before any smoothing.
    solve_bytes = 0.0
    for s_vox, fc, st in zip(span, fcs, st32):
        b_pass = bytes_per_voxel_pass(fc, "fp32")
        solve_bytes += s_vox * (b_pass * st["sweeps"] / st["sweeps_per_pass"] +
                                24.0 * (st["outer_iters"] + 1) + 20.0 * st["outer_iters"])
    solve_gbs = solve_bytes / (ms_step * 1e-3) / 1e9
    # The smoother's per-launch working set: every field's pass bytes at once
    # (the three launches run concurrently). Within the 256 MB Infinity Cache
    # the sweeps are bound by VALU work and the tile loads' latency
    # (SQ_INSTS_VALU at 0.84 of the roof at 1024^2); beyond it each launch
    # streams from HBM as well.
    working_set = sum(s_vox * bytes_per_voxel_pass(fc, "fp32") for s_vox, fc in zip(span, fcs))
    binding = ("VALU and tile-load latency (FP32 register-blocked smoother; working set %.0f MB, "
               "inside the 256 MB Infinity Cache)" % (working_set / 1e6) if working_set <= 256e6 else
               "HBM and VALU (FP32 register-blocked smoother; working set %.0f MB, beyond the "
               "256 MB Infinity Cache)" % (working_set / 1e6))

    out = {
        "metric": "voxel-sweeps/s & % HBM roofline, FP32+refine vs FP64 SOR at 1e-8 residual",
        "value": value,
        "unit": "voxel-sweeps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 smoothing, f64 residual/correction",
        "data": "synthetic (seeded Voronoi lattice, 1024^2 plane, ~40k cells, three fields)",
        "config": {
            "workload": f"2D multi-field {args.n}x{args.n}, ~40k cells, oxygen+signal+chemokine, "
                        "FP32+refine to relative residual 1e-8",
            "grid": [args.n, args.n], "fields": [fc.name for fc in fcs], "omega": omega,
            "tolerance": tol, "swept_voxels": span, "sweeps_per_pass": spp_used,
        },
        "roofline": {
            # the whole FP32+refine solve against the HBM roof;
            # the bench's working set (about 42 MB) stays in the 256 MB
            # Infinity Cache, and the smoother is bound by its VALU work
            # (valu below), not by HBM
            "bound": "hbm", "achieved": solve_gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": solve_gbs / HBM_PEAK_GBS,
            "traffic": traffic_json["bytes_per_launch"] if traffic_json else None,
            "basis": "algorithmic bytes of the whole solve (smoother passes at B_pass/s a "
                     "voxel-sweep, FP64 residual 24 B and correction 20 B a voxel per outer "
                     "iteration) over ms_per_step",
            "alg_bytes_per_solve": solve_bytes,
            "binding_resource": binding,
            "kernel": {
                "name": f"relax2d_reg<float,{spp_used},{nw32},{nr32},...> (FP32 smoother, {spp_used} sweeps "
                        f"a launch; one launch per field, the {len(fcs)} fields' launches "
                        "concurrent on their own streams)",
                "launch_ms": ms32, "alg_bytes_all_fields_per_launch": bytes32,
                "achieved_GBs": achieved, "frac": achieved / HBM_PEAK_GBS,
                "timing": "HIP events on each field's stream over back-to-back launches of the "
                          "three fields; the longest per-stream average interval",
                "traffic_bytes_per_launch": traffic_json["bytes_per_launch"] if traffic_json else None,
                "valu_frac": valu_json["valu_frac"] if valu_json else None,
                "sweeps_per_launch": spp_used,
                # an unblocked smoother (one sweep per HBM pass) would have to
                # stream this many GB/s to sweep as fast
                "one_sweep_per_pass_equivalent_GBs": achieved * spp_used,
            },
            "attainable_copy_GBs": copy_bandwidth(),
        },
        "cpu_baseline": cpu,
        "fp64": {
            "voxel_sweeps_per_s": vs64 / t64, "ms_per_solve": t64 * 1e3,
            "sweeps": [st["sweeps"] for st in st64], "sweeps_per_pass": spp64_used,
            "kernel_launch_ms": ms64, "kernel_alg_GBs": achieved64,
            "kernel_frac": achieved64 / HBM_PEAK_GBS,
        },
        "fp32_refine": {"ms_per_solve": ms_step, "errors_vs_fp64": errs},
        "speedup_vs_fp64": {"throughput": value / (vs64 / t64), "time_to_solution": t64 / (ms_step / 1e3)},
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
