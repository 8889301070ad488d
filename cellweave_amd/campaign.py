"""Batched campaigns: many independent histories solved together.

The reference runs a campaign's histories in parallel on CPU threads, each with
its own fields (src/campaign.rs:196-204). Here every field of every history is
solved by one device-driven launch sequence (cw_campaign_relax): fields are
grouped by size class, each class swept with its own tile shape from the
autotuner table; every field keeps its own state on the device, a small kernel
between steps moves it to its next phase (FP64 sweeps and residual tests, or
the FP32 refinement's residual, smoothing segment and correction), and one
captured hipGraph holds every class's chain of steps. The host launches the
graph and waits once.

    python -m cellweave_amd.campaign --tune     # on a GPU: writes the shape table
"""
from __future__ import annotations

import argparse
import json
from concurrent.futures import ThreadPoolExecutor
import os
import time

from .field import (CRIT_REL_RESIDUAL, PREC_FP32_REFINE, PREC_FP64, Field, autotune,
                    batch_relax, campaign_relax)

TABLE_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "autotune_gfx950.json")

# Size classes by the longest grid edge.
CLASSES = ((0, 256, "<256"), (256, 512, "256-511"), (512, 1 << 30, ">=512"))

# Plane tile shapes the device-driven batch is compiled for (cw_campaign_relax).
DEVICE_SHAPES = {"fp64": (8, 9), "fp32": (15, 18, 19, 20, 21, 22, 23, 24, 25, 26, 27)}


def size_class(n: int) -> str:
    for lo, hi, name in CLASSES:
        if lo <= n < hi:
            return name
    raise ValueError(n)


def load_table(path: str = TABLE_PATH) -> dict:
    if os.path.exists(path):
        with open(path) as fh:
            return json.load(fh)
    return {}


def make_fields(cases_):
    out = []
    for fc in cases_:
        g = Field(fc.w, fc.h, fc.d, fc.diff, fc.decay, fc.initial)
        g.upload(values=fc.initial_values(), held=fc.held, uptake=fc.uptake,
                 production=fc.production)
        out.append(g)
    return out


def omega_for(fc) -> float:
    import math

    s = fc.span()
    longest = max(s[1] - s[0] + 1, s[3] - s[2] + 1, s[5] - s[4] + 1)
    return 2.0 / (1.0 + math.sin(math.pi / max(longest, 2)))


def group(cases_):
    groups = {}
    for k, fc in enumerate(cases_):
        groups.setdefault(size_class(max(fc.w, fc.h, fc.d)), []).append(k)
    return groups


def plan(cases_, precision=PREC_FP32_REFINE, table=None):
    """The launch plan, built before any timing: the fields of each size class
    with the class's tile shape from the autotuner table."""
    table = load_table() if table is None else table
    key = "fp64" if precision == PREC_FP64 else "fp32"
    return [(cls, idx, table.get(key, {}).get(cls, -1))
            for cls, idx in sorted(group(cases_).items())]


def solve(fields, plan_, precision=PREC_FP32_REFINE, tol=1e-8, use_graph=1,
          max_sweeps=2_000_000, omega=-1.0, concurrent=False, device_driven=True):
    """Solve every field to `tol`. omega < 0 gives each field its own
    2/(1+sin(pi/L)), L its longest free-span edge. Returns (stats per field,
    seconds).

    device_driven (the default): one cw_campaign_relax call for every class,
    each with its plan's tile shape, from this thread. Otherwise one
    cw_batch_relax launch sequence per size class, verdicts read by the host,
    the classes one after another; with `concurrent` they run at once, one
    host thread each (about 2% faster; a profiled run of that form once ended
    in a host segmentation fault whose cause was not found, so it is not the
    default)."""
    stats = [None] * len(fields)
    if device_driven:
        groups = [0] * len(fields)
        shapes = []
        for g, (_cls, idx, shape) in enumerate(plan_):
            for k in idx:
                groups[k] = g
            shapes.append(shape)
        t0 = time.perf_counter()
        st, _ = campaign_relax(fields, groups, shapes, criterion=CRIT_REL_RESIDUAL, tolerance=tol,
                               max_sweeps=max_sweeps, omega=omega, precision=precision)
        return st, time.perf_counter() - t0

    def one(entry):
        cls, idx, shape = entry
        st, _ = batch_relax([fields[k] for k in idx], criterion=CRIT_REL_RESIDUAL, tolerance=tol,
                            max_sweeps=max_sweeps, omega=omega, precision=precision,
                            tile_shape=shape, use_graph=use_graph)
        for k, s in zip(idx, st):
            stats[k] = s

    t0 = time.perf_counter()
    if concurrent and len(plan_) > 1:
        with ThreadPoolExecutor(max_workers=min(len(plan_), 4)) as ex:
            for f in [ex.submit(one, e) for e in plan_]:
                f.result()
    else:
        for e in plan_:
            one(e)
    return stats, time.perf_counter() - t0


def tune(cases_, per_class=0, launches=10, confirm=6, tol=1e-8):
    """Rank the sweep shapes on each size class, for both precisions, on every
    field of the class (per_class=0), since how many workgroup rounds a launch
    takes depends on the whole batch, or on its first per_class fields.

    Two stages: cw_autotune ranks every shape by its launch time per sweep,
    then the `confirm` best are timed on the class's whole solve, as `solve`
    runs it (graph replays, per-field omega, refinement segments), and the
    fastest solve wins. The per-sweep ranking leaves out what a shape changes
    around its launches: launches per segment, graph chunks, verdicts."""
    table = {"fp64": {}, "fp32": {}, "ranking": {}, "confirmed_ms": {}}
    for cls, idx in sorted(group(cases_).items()):
        sample = [cases_[k] for k in (idx[:per_class] if per_class > 0 else idx)]
        fields = make_fields(sample)

        def reset():
            for g, fc in zip(fields, sample):
                g.fill(fc.initial)

        for key, prec in (("fp64", PREC_FP64), ("fp32", PREC_FP32_REFINE)):
            ranked = autotune(fields, launches=launches, criterion=CRIT_REL_RESIDUAL,
                              tolerance=tol, omega=-1.0, precision=prec)
            reset()
            reset()
            # the campaign runs device-driven: only the shapes compiled for it
            runnable = [r for r in ranked if r[0] in DEVICE_SHAPES[key]]
            timed = []
            for shape, _ in runnable[:max(1, confirm)]:
                best = None
                for _rep in range(2):  # the first also builds the plan and graph
                    reset()
                    _, secs = solve(fields, [(cls, list(range(len(fields))), shape)],
                                    precision=prec, tol=tol, concurrent=False)
                    best = secs if best is None else min(best, secs)
                timed.append([shape, 1e3 * best])
            timed.sort(key=lambda e: e[1])
            print(f"tune {key} {cls}: {len(fields)} fields, solve ms by shape {timed}", flush=True)
            table[key][cls] = timed[0][0]
            table["ranking"][f"{key} {cls}"] = runnable[:6]
            table["confirmed_ms"][f"{key} {cls}"] = timed
            reset()
    return table


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--tune", action="store_true")
    ap.add_argument("--count", type=int, default=512)
    args = ap.parse_args()
    from . import cases

    hist = cases.campaign_histories(count=args.count)
    flat = [fc for h in hist for fc in h]
    if args.tune:
        table = tune(flat)
        table["workload"] = f"campaign: {args.count} plane histories x 2 fields, whole size classes"
        with open(TABLE_PATH, "w") as fh:
            json.dump(table, fh, indent=1)
        print(json.dumps(table))


if __name__ == "__main__":
    main()
