"""The measured workloads beside bench.py's default line.

Each function runs one workload on the GPU and returns a dict in the shape of
bench.py's JSON line: voxel-sweeps/s (swept voxels x sweeps / wall time), the
FP64 solve beside the FP32+refine one, and the live per-launch timing of the
dominant kernel for the roofline.

    oxygen200:    200^2, one oxygen field, omega 1.9 (the reference's CPU-sized case)
    spheroid256:  256^3 spheroid, FP32+refine against FP64
    campaign512:  512 histories x 2 fields, batched per size class, graph replays
    adjoint192:   192^3 adjoint, gradients against central finite differences
    coupling:     the device-side gathers, holds, step and per-cell means
    tissue120:    the per-lattice-step loop (sources, warm-started relax, means) at 120^3
"""
from __future__ import annotations

import math
import time

import numpy as np

from . import campaign, cases
from .field import (CRIT_MAX_CHANGE, CRIT_REL_RESIDUAL, PREC_FP32_REFINE, PREC_FP64, Field,
                    batch_relax, time_passes)

HBM_PEAK_GBS = 8000.0


def _field(fc):
    g = Field(fc.w, fc.h, fc.d, fc.diff, fc.decay, fc.initial)
    g.upload(values=fc.initial_values(), held=fc.held, uptake=fc.uptake,
             production=fc.production)
    return g


def _line(metric_value, config, extra):
    out = {
        "metric": "voxel-sweeps/s & % HBM roofline, FP32+refine vs FP64 SOR at 1e-8 residual",
        "value": metric_value, "unit": "voxel-sweeps/s", "n_gpus": 1,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "data": "synthetic (seeded Voronoi lattice)", "config": config,
    }
    out.update(extra)
    return out


def _solve_pair(fields, fcs, omega, reset, tol=1e-8, reps=3, use_graph=0):
    span = [fc.span_voxels() for fc in fcs]
    res = {}
    for name, prec in (("fp64", PREC_FP64), ("fp32_refine", PREC_FP32_REFINE)):
        reset()
        batch_relax(fields, criterion=CRIT_REL_RESIDUAL, tolerance=tol, omega=omega,
                    precision=prec, use_graph=use_graph)  # warm-up
        ts, st = [], None
        for _ in range(reps):
            reset()
            t0 = time.perf_counter()
            st, _ = batch_relax(fields, criterion=CRIT_REL_RESIDUAL, tolerance=tol, omega=omega,
                                precision=prec, use_graph=use_graph)
            ts.append(time.perf_counter() - t0)
        t = float(np.median(ts))
        vs = sum(s * x["sweeps"] for s, x in zip(span, st))
        res[name] = {"ms_per_solve": 1e3 * t, "voxel_sweeps_per_s": vs / t,
                     "sweeps": [x["sweeps"] for x in st],
                     "rel_residual": [x["rel_residual"] for x in st],
                     "outer": [x["outer_iters"] for x in st],
                     "sweeps_per_pass": [x["sweeps_per_pass"] for x in st],
                     "tile_shape": [x["tile_shape"] for x in st],
                     "values": [g.download() for g in fields]}
    return res


def _errors(res, fcs):
    out = {}
    for fc, a, b in zip(fcs, res["fp32_refine"]["values"], res["fp64"]["values"]):
        out[fc.name] = {"max_abs": float(np.max(np.abs(a - b))),
                        "rel_l2": float(np.linalg.norm(a - b) / (np.linalg.norm(b) or 1.0))}
    return out


def _strip(res):
    return {k: {kk: vv for kk, vv in v.items() if kk != "values"} for k, v in res.items()}


def oxygen200():
    (fc,) = cases.oxygen_disc_200()
    g = _field(fc)
    res = _solve_pair([g], [fc], 1.9, lambda: g.fill(1.0))
    vs = res["fp32_refine"]["voxel_sweeps_per_s"]
    # the reference's own criterion (largest change per sweep below 1e-8, pde.rs:279-284)
    ts = []
    for _ in range(4):
        g.fill(1.0)
        t0 = time.perf_counter()
        st = g.relax_ex(omega=1.9, criterion=CRIT_MAX_CHANGE, tolerance=1e-8, max_sweeps=200_000)
        ts.append(time.perf_counter() - t0)
    ref_rule = {"criterion": "max change < 1e-8 (pde.rs:279-284)", "precision": "fp64",
                "sweeps": st["sweeps"], "ms_per_solve": 1e3 * float(np.median(ts[1:])),
                "rel_residual_at_stop": g.residual()}
    return _line(vs, {"workload": "200x200 oxygen field, ~450 cells, omega 1.9, rel residual 1e-8"},
                 {"fp64": _strip(res)["fp64"], "fp32_refine": _strip(res)["fp32_refine"],
                  "errors_vs_fp64": _errors(res, [fc]), "reference_criterion": ref_rule})


def spheroid256(n=256):
    t0 = time.perf_counter()
    (fc,) = cases.spheroid(n=n, ncells=int(round(60000 * (n / 256) ** 3)), radius=110 * n / 256,
                          necrotic=40 * n / 256)
    build = time.perf_counter() - t0
    g = _field(fc)
    omega = campaign.omega_for(fc)
    res = _solve_pair([g], [fc], omega, lambda: g.fill(1.0), reps=2)
    span = fc.span_voxels()
    g.fill(1.0)
    ms32 = time_passes([g], 40, criterion=CRIT_REL_RESIDUAL, tolerance=1e-8, omega=omega,
                       precision=PREC_FP32_REFINE)
    g.fill(1.0)
    ms64 = time_passes([g], 40, criterion=CRIT_REL_RESIDUAL, tolerance=1e-8, omega=omega,
                       precision=PREC_FP64)
    # B_pass: e 4+4, rhs 4, lower bound 4 (oxygen has sinks) a voxel, read and
    # written once a launch however many sweeps the launch runs
    b32 = span * 16.0
    b64 = span * 24.0
    s = _strip(res)
    return _line(res["fp32_refine"]["voxel_sweeps_per_s"],
                 {"workload": f"{n}^3 spheroid, ~60k cells, FP32+refine vs FP64 to 1e-8",
                  "omega": omega, "swept_voxels": span, "case_build_s": build},
                 {"fp64": s["fp64"], "fp32_refine": s["fp32_refine"],
                  "errors_vs_fp64": _errors(res, [fc]),
                  "roofline": {"bound": "hbm", "unit": "GB/s", "peak": HBM_PEAK_GBS,
                               "fp32_kernel_launch_ms": ms32,
                               "fp32_sweeps_per_launch": res["fp32_refine"]["sweeps_per_pass"][0],
                               "fp32_achieved": b32 / ms32 / 1e6,
                               "fp32_frac": b32 / ms32 / 1e6 / HBM_PEAK_GBS,
                               "fp64_kernel_launch_ms": ms64,
                               "fp64_achieved": b64 / ms64 / 1e6,
                               "fp64_frac": b64 / ms64 / 1e6 / HBM_PEAK_GBS},
                  "speedup_vs_fp64": {
                      "throughput": res["fp32_refine"]["voxel_sweeps_per_s"] /
                      res["fp64"]["voxel_sweeps_per_s"],
                      "time_to_solution": res["fp64"]["ms_per_solve"] /
                      res["fp32_refine"]["ms_per_solve"]}})


def campaign512(count=512, reps=2):
    t0 = time.perf_counter()
    hist = cases.campaign_histories(count=count)
    flat = [fc for h in hist for fc in h]
    build = time.perf_counter() - t0
    fields = campaign.make_fields(flat)
    span = [fc.span_voxels() for fc in flat]
    table = campaign.load_table()
    out, per_field = {}, {}
    # Both precisions device-driven (cw_campaign_relax: one captured graph,
    # one host wait); the host-driven FP64 batch (cw_batch_relax per size
    # class, verdicts read by the host) is timed too, and the speed-ups are
    # taken against the faster FP64 solve.
    for name, prec, dd in (("fp64", PREC_FP64, True), ("fp64_host_driven", PREC_FP64, False),
                           ("fp32_refine", PREC_FP32_REFINE, True)):
        plan = campaign.plan(flat, prec, table)
        times = []
        for rep in range(1 + reps):  # the first solve is the untimed warm-up
            for g, fc in zip(fields, flat):
                g.fill(fc.initial)
            stats, secs = campaign.solve(fields, plan, precision=prec, device_driven=dd)
            if rep:
                times.append(secs)
        secs = float(np.median(times))
        per_field[name] = [int(st["sweeps"]) for st in stats]
        vs = sum(s * st["sweeps"] for s, st in zip(span, stats))
        worst = max(st["rel_residual"] for st in stats)
        out[name] = {"seconds": secs, "seconds_each": times, "voxel_sweeps_per_s": vs / secs,
                     "total_sweeps": int(sum(st["sweeps"] for st in stats)),
                     "worst_rel_residual": worst,
                     "all_converged": all(st["status"] == 0 for st in stats)}
    best64 = min((out["fp64"], out["fp64_host_driven"]), key=lambda r: r["seconds"])
    return _line(out["fp32_refine"]["voxel_sweeps_per_s"],
                 {"workload": f"{count} histories x 2 fields (oxygen, signal), edges "
                              "logU[128,768], device-driven (one captured graph for every "
                              "size class, each with its own tile shape), per-field optimal "
                              "omega",
                  "fields": len(flat), "voxels": int(sum(fc.n for fc in flat)),
                  "case_build_s": build, "shape_table": {k: table.get(k) for k in ("fp64", "fp32")}},
                 {"fp64": out["fp64"], "fp64_host_driven": out["fp64_host_driven"],
                  "fp32_refine": out["fp32_refine"],
                  "speedup_vs_fp64": {"against": "the faster FP64 solve",
                                      "throughput": out["fp32_refine"]["voxel_sweeps_per_s"] /
                                      best64["voxel_sweeps_per_s"],
                                      "time_to_solution": best64["seconds"] /
                                      out["fp32_refine"]["seconds"]},
                  # per-field FP64 sweep counts, for a CPU timing of the same solves
                  "_fp64_sweeps": per_field["fp64"]})


def neighbour_defect(c, w, h, d):
    """q_i = sum over in-box neighbours j of (c_i - c_j): the operator of the
    steady state D q(c) + decay c = source + p - u (pde.rs:384-395) without D."""
    c3 = c.reshape(d, h, w)
    q = np.zeros_like(c3)
    for ax in range(3):
        if c3.shape[ax] < 2:
            continue
        hi = [slice(None)] * 3
        lo = [slice(None)] * 3
        hi[ax] = slice(1, None)
        lo[ax] = slice(None, -1)
        step = c3[tuple(hi)] - c3[tuple(lo)]
        q[tuple(hi)] += step
        q[tuple(lo)] -= step
    return q.reshape(-1)


class SpheroidAdjoint:
    """The adjoint workload on an n^3 Voronoi spheroid (oxygen, D = 20 px^2/s,
    no decay, box faces held at g = 1, necrotic core without uptake):
    loss 0.5 sum (c - target)^2 with the target the FP64 steady state at
    D x 1.1 and uptakes x (1 + 0.1 U[-1, 1]).

    Central differences at relative step h solve each perturbed problem in
    correction form, c = base + delta, for delta' = delta + shift with
    D q(delta') = -u - D q(base) and held delta' = shift + dg. Solved directly,
    c ~ 1 carries 1e-16 rounding while a 1e-6 change of one cell's uptake moves
    c by ~1e-10, which limits the quotient to ~1e-5; in correction form the
    rounding scales with the shift (4h) instead. The shift keeps delta' above
    the solve's floor at zero, so the projected solve returns the linear one:
    with the floor inactive c = 1 - L^-1 u / D (L the graph Laplacian with
    the faces held), so a step h in D or in one cell's rate moves c by at most
    h max(1 - c), and a step h in g moves it by exactly h."""

    def __init__(self, n=192, ncells=25000, radius=82.0, necrotic=40.0):
        self.n = n
        self.fc, self.pert = cases.adjoint_spheroid(n=n, ncells=ncells, radius=radius, necrotic=necrotic)
        self.necro = cases.necrotic_mask(n, necrotic)
        self.held = self.fc.held == 1
        self.rates = self.fc.per_cell_uptake
        self.omega = campaign.omega_for(self.fc)
        assert np.array_equal(self.uptake_of(self.rates), np.where(self.held, 0.0, self.fc.uptake))

    def uptake_of(self, rates):
        up = cases.gather(self.fc.grid, rates)
        up[self.necro | self.held] = 0.0
        return up

    def solve(self, diff, g, rates, precision=PREC_FP64, tol=1e-13):
        n = self.n
        f = Field(n, n, n, diff, 0.0, 1.0)
        f.upload(values=np.where(self.held, g, 1.0), held=self.fc.held,
                 uptake=self.uptake_of(rates))
        st = f.relax_ex(omega=self.omega, criterion=CRIT_REL_RESIDUAL, tolerance=tol,
                        precision=precision, max_sweeps=400_000)
        return f, st

    def prepare(self):
        fc = self.fc
        ft, _ = self.solve(fc.diff * self.pert["diff_scale"], 1.0,
                           self.rates * self.pert["uptake_scale"])
        self.target = ft.download()
        fb, self.base_stats = self.solve(fc.diff, 1.0, self.rates)
        self.base = fb.download()
        self.q = neighbour_defect(self.base, self.n, self.n, self.n)

    def delta(self, diff, dg, rates, shift, tol=1e-13):
        n = self.n
        r = np.where(self.held, 0.0, -self.uptake_of(rates) - diff * self.q)
        f = Field(n, n, n, diff, 0.0, shift)
        f.upload(values=np.where(self.held, shift + dg, shift), held=self.fc.held,
                 uptake=np.where(r < 0, -r, 0.0), production=np.where(r > 0, r, 0.0))
        st = f.relax_ex(omega=self.omega, criterion=CRIT_REL_RESIDUAL, tolerance=tol,
                        precision=PREC_FP64, max_sweeps=400_000)
        out = f.download()
        if out.min() <= 0.0:
            raise ValueError("the shifted correction touched the floor; raise the shift")
        return out - shift, st["sweeps"]

    def quotient(self, plus, minus, den):
        (dp, sp), (dm, sm) = plus, minus
        self.fd_sweeps += [sp, sm]
        return 0.5 * math.fsum((dp - dm) * (2.0 * (self.base - self.target) + dp + dm)) / den

    def differences(self, h=1e-6, cells=()):
        """Central differences of the loss in D, g and each listed cell's rate."""
        self.fd_sweeps = []
        D, r0 = self.fc.diff, self.rates
        s = 4.0 * h * float(np.max(1.0 - self.base))  # bounds |delta| for D and cells
        sg = 4.0 * h
        out = {"D": self.quotient(self.delta(D * (1 + h), 0.0, r0, s),
                                  self.delta(D * (1 - h), 0.0, r0, s), 2 * h * D),
               "g": self.quotient(self.delta(D, h, r0, sg), self.delta(D, -h, r0, sg), 2 * h),
               "cells": {}}
        for k in cells:
            up, dn = r0.copy(), r0.copy()
            up[k] *= 1 + h
            dn[k] *= 1 - h
            out["cells"][int(k)] = self.quotient(self.delta(D, 0.0, up, s),
                                                 self.delta(D, 0.0, dn, s), 2 * h * r0[k])
        return out


def adjoint192(n=192, ncells=25000, radius=82.0, fd_cells=32, h=1e-6, adj_tol=1e-11):
    """Forward in FP32+refinement, adjoint in FP64, gradients against central
    differences at relative step h (correction-form perturbed solves)."""
    t0 = time.perf_counter()
    case = SpheroidAdjoint(n=n, ncells=ncells, radius=radius, necrotic=40.0 * n / 192)
    build = time.perf_counter() - t0
    case.prepare()
    rates = case.rates
    # the measured path: forward FP32+refinement, adjoint FP64
    t1 = time.perf_counter()
    f, st_fwd = case.solve(case.fc.diff, 1.0, rates, precision=PREC_FP32_REFINE, tol=1e-12)
    t_fwd = time.perf_counter() - t1
    c = f.download()
    # The adjoint's right-hand side c - target is interior and small, so its
    # residual floor sits near eps * cond(A) (3e-13..1e-12 at 64^3, ~9x that
    # at 192^3); the gradients agree to 1e-13 between 1e-11 and 1e-12 stops.
    t1 = time.perf_counter()
    adj = f.adjoint(case.target, grid=case.fc.grid, ncells=rates.size, tolerance=adj_tol,
                    omega=case.omega, max_sweeps=200_000)
    t_adj = time.perf_counter() - t1
    del f
    live = np.flatnonzero(np.abs(adj["du"]) > 0)
    cells = np.random.default_rng(5005).choice(live, size=min(fd_cells, live.size), replace=False)
    t1 = time.perf_counter()
    fd = case.differences(h=h, cells=cells)
    t_fd = time.perf_counter() - t1
    rel = lambda a, b: abs(a - b) / abs(b)
    checks = {"D": {"adjoint": adj["dl_dd"], "fd": fd["D"], "rel_err": rel(adj["dl_dd"], fd["D"])},
              "g": {"adjoint": adj["dl_dg"], "fd": fd["g"], "rel_err": rel(adj["dl_dg"], fd["g"])}}
    cell_err = {k: rel(adj["du"][k], v) for k, v in fd["cells"].items()}
    checks["uptake"] = {"cells": {str(k): {"adjoint": float(adj["du"][k]), "fd": v,
                                           "rel_err": cell_err[k]} for k, v in fd["cells"].items()},
                        "max_rel_err": max(cell_err.values())}
    worst = max(checks["D"]["rel_err"], checks["g"]["rel_err"], checks["uptake"]["max_rel_err"])
    return {"workload": f"{n}^3 oxygen volume, {rates.size - 1} cells, loss 0.5 sum (c - c_target)^2, "
                        "target from D x 1.1 and uptake x (1 + 0.1 U[-1,1])",
            "omega": case.omega, "case_build_s": build, "min_c": float(c.min()),
            "forward": {"precision": "fp32_refine", "seconds": t_fwd, "sweeps": st_fwd["sweeps"],
                        "outer": st_fwd["outer_iters"], "rel_residual": st_fwd["rel_residual"],
                        "max_abs_vs_fp64": float(np.max(np.abs(c - case.base))),
                        "fp64_sweeps": case.base_stats["sweeps"]},
            "adjoint": {"precision": "fp64", "tolerance": adj_tol, "seconds": t_adj,
                        "sweeps": adj["sweeps"], "rel_residual": adj["rel_residual"],
                        "loss": adj["loss"]},
            "fd": {"h_rel": h, "form": "correction (c = base + delta, shift 4h max(1-c), 4h for g)", "solve_tol": 1e-13,
                   "seconds": t_fd, "sweeps": case.fd_sweeps},
            "checks": checks, "worst_rel_err": worst, "within_1e-6": bool(worst <= 1e-6),
            "units": "dL/dD per px^2/s (the solve's D is um^2/s divided by dx^2 = 100 um^2)"}


def _timed(fn, reps=10):
    fn()  # warm-up (allocations, cached cell lists)
    ts = []
    for _ in range(reps):
        t0 = time.perf_counter()
        fn()
        ts.append(time.perf_counter() - t0)
    return float(np.median(ts))


def coupling(reps=10):
    """The steps either side of the solve, on the device (coupling.rs:26-182,
    pde.rs:415-728), at 1024^2 (three fields' lattice) and 256^3: wall time
    per call (each call is synchronous) and HBM GB/s from each kernel's
    algorithmic bytes per voxel (stated beside it)."""
    from .field import Lattice

    out = {}
    for name, fc in (("1024^2", cases.multifield_plane(n=1024)[0]), ("256^3", cases.spheroid()[0])):
        w, h, d = fc.w, fc.h, fc.d
        n = w * h * d
        ids = fc.grid.astype(np.uint32)
        ncells = int(ids[ids != 0xFFFFFFFF].max()) + 1
        consumes = np.ones(ncells, np.uint8)
        consumes[0] = 0
        prod = np.where(np.random.default_rng(7).random(ncells) < 0.3, 2e-4, 0.0)
        lat = Lattice(w, h, d)
        f = Field(w, h, d, fc.diff, 0.0, 1.0)
        ops = {
            # name: (call, algorithmic bytes per voxel)
            "lattice_upload": (lambda: lat.upload(ids), 4),
            "apply_uptake": (lambda: f.apply_uptake(lat, consumes, 1e-3), 4 + 8),
            "apply_production": (lambda: f.apply_production(lat, prod), 4 + 8),
            "release_all": (lambda: f.release_all(), 1),
            "hold_medium_bath": (lambda: f.hold_medium_bath(lat, 1.0), None),
            "free_count": (lambda: f.free_count(), 1),
            "inflow": (lambda: f.inflow(), 8 + 1),
            "step": (lambda: f.step(0.2 * 0.25 / max(fc.diff, 1e-30)), 8 + 1 + 8 + 8 + 8),
            "mean_per_cell": (lambda: f.mean_per_cell(lat, ncells), 4 + 8),
        }
        res = {}
        for op, (fn, bpv) in ops.items():
            t = _timed(fn, reps)
            r = {"ms": 1e3 * t}
            if bpv:
                r["bytes_per_voxel"] = bpv
                r["GBs"] = n * bpv / t / 1e9
                r["frac_hbm"] = r["GBs"] / HBM_PEAK_GBS
            res[op] = r
        # the cell lists are rebuilt after every lattice upload: time that too
        def fresh_mean():
            lat.upload(ids)
            f.mean_per_cell(lat, ncells)
        res["mean_per_cell_after_upload"] = {"ms": 1e3 * _timed(fresh_mean, reps)}
        out[name] = {"voxels": n, "cells": ncells, "ops": res}
    return {"workload": "device-side coupling, holds, step and per-cell means (synchronous calls)",
            "results": out,
            "notes": "lattice_upload includes the PCIe copy of the ids; hold_medium_bath is a "
                     "closed-box scan, a tiled flood to a fixed point and a finish pass"}


TISSUE_UPTAKE = 1e-3        # per s, every living cell consumes
TISSUE_TOLERANCE = 0.01     # per-step tolerance, times the uptake (simulation.rs:85)
TISSUE_MAX_SWEEPS = 20_000  # simulation.rs:88
TISSUE_WARMUP = (1e-8, 2_000_000)  # simulation.rs:95, :98


def tissue_steps(n=120, radius=45.0, steps=10, copies=3800, seed=6000):
    """The inputs of the lattice-step workload: the first lattice and one
    lattice per step (each `copies` boundary moves from the last), all
    seeded. Returns (ids per step with the first at index 0, cell count)."""
    ids, ncells = cases.tissue_box(seed=seed, n=n, radius=radius)
    rng = np.random.default_rng(seed + 1)
    seq = [ids]
    for _ in range(steps):
        seq.append(cases.lattice_copies(seq[-1], n, n, n, copies, rng))
    return seq, ncells


def tissue120(n=120, radius=45.0, steps=10, copies=3800):
    """The reference's per-step loop on a 120^3 tissue in a bath, with every
    array on the device: a warm-up solve to 1e-8 absolute (simulation.rs:1068-1083),
    then per lattice step the lattice upload, declare_sources (uptake of the
    living cells, release, the medium bath held at 1; simulation.rs:2138-2145),
    the warm-started relax to 0.01 x uptake with at most 20,000 sweeps
    (simulation.rs:1144-1154) and the per-cell means (simulation.rs:1157).
    The lattice goes over as the voxels each step changed (cw_lattice_set);
    the whole-grid upload is timed apart, as is the values download a host
    reader of every value would add. FP64 plain with the reference's
    stopping rule: the same sweeps and bits as the CPU loop. Metric: lattice
    steps/s."""
    from .field import Lattice

    seq, ncells = tissue_steps(n, radius, steps, copies)
    # the voxels each step changed, as the lattice's own step knows them
    changes = [np.flatnonzero(a != b) for a, b in zip(seq, seq[1:])]
    consumes = np.ones(ncells + 1, np.uint8)
    consumes[0] = 0
    lat = Lattice(n, n, n)
    f = Field(n, n, n, cases.d_px(2000.0), 0.0, 1.0)

    def declare():
        f.apply_uptake(lat, consumes, TISSUE_UPTAKE)
        f.release_all()
        f.hold_medium_bath(lat, 1.0)

    t0 = time.perf_counter()
    lat.upload(seq[0])
    declare()
    warm = f.relax(*TISSUE_WARMUP)
    f.mean_per_cell(lat, ncells + 1)
    t_warm = time.perf_counter() - t0
    per_step, sweeps, dl, full = [], [], [], []
    for ids, changed in zip(seq[1:], changes):
        t0 = time.perf_counter()
        lat.set(changed, ids[changed])
        declare()
        sweeps.append(f.relax(TISSUE_TOLERANCE * TISSUE_UPTAKE, TISSUE_MAX_SWEEPS))
        means = f.mean_per_cell(lat, ncells + 1)
        per_step.append(time.perf_counter() - t0)
        t1 = time.perf_counter()
        f.download()
        dl.append(time.perf_counter() - t1)
        t1 = time.perf_counter()
        lat.upload(ids)  # what a whole-grid upload would have cost instead
        full.append(time.perf_counter() - t1)
    t = float(np.median(per_step))
    span = n * n * n
    return {
        "metric": "lattice steps/s (declare sources + warm-started relax + per-cell means)",
        "value": 1.0 / t, "unit": "steps/s", "n_gpus": 1, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (seeded Voronoi ball in a bath, seeded boundary copies per step)",
        "config": {"workload": f"{n}^3 tissue in a bath, {ncells} cells in a radius-{radius:g} "
                               f"ball, {copies} boundary copies per step, {steps} steps",
                   "tolerance": TISSUE_TOLERANCE * TISSUE_UPTAKE,
                   "max_sweeps": TISSUE_MAX_SWEEPS, "precision": "fp64, reference rule"},
        "ms_per_step": 1e3 * t, "ms_per_step_each": [1e3 * x for x in per_step],
        "sweeps_per_step": sweeps, "warmup": {"sweeps": warm, "seconds": t_warm},
        "voxel_sweeps_per_s": span * float(np.sum(sweeps)) / float(np.sum(per_step)),
        "values_download_ms": 1e3 * float(np.median(dl)),
        "lattice_update": "changed voxels only (cw_lattice_set), "
                          f"{int(np.median([c.size for c in changes]))} a step",
        "whole_grid_upload_ms": 1e3 * float(np.median(full)),
        "mean_of_means": float(np.mean(means[1:])),
    }


WORKLOADS = {"oxygen200": oxygen200, "spheroid256": spheroid256, "campaign512": campaign512,
             "adjoint192": adjoint192, "coupling": coupling, "tissue120": tissue120}
