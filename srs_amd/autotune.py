"""Autotuner for the sweep kernel's launch shape and cache hints.

Candidates: elements per workgroup (tile), realizations per realization
group (group, for batches), non-temporal loads of the read-once operator rows
(nt_geometry), non-temporal stores of the new state (nt_state), SoA plane
padding (plane_pad), the compact row format, the register budget and LDS
staging.  Every candidate computes the
same bits; only speed differs.  Each trial runs on a scratch context built
by the caller's factory, so the context that is finally used has taken no
steps.  Timing is the step as it runs, a captured graph of steps timed by
HIP events around the call.
"""
from __future__ import annotations

import itertools
import time
from typing import Callable

from . import STEP_GRAPH, Context

LAYOUT_KEYS = ("tile", "plane_pad", "compact", "group")
KERNEL_KEYS = ("nt_geometry", "nt_state", "occupancy", "lds")


def autotune(make: Callable[[dict], Context], tiles=(128, 256, 512), pads=(0,), compact=(1, 0), nt_geometry=(0, 1),
             nt_state=(0, 1), occupancy=(0, 6), lds=(0, 1), groups=(1,), reps: int = 20, warm: int = 3,
             log=None) -> tuple[dict, list]:
    """Returns (best configuration, all trials as (config, step_ms))."""
    trials = []
    for tile, pad, cmp, grp in itertools.product(tiles, pads, compact, groups):
        layout = {"tile": tile, "plane_pad": pad, "compact": cmp}
        if len(groups) > 1 or grp != 1:
            layout["group"] = grp
        ctx = make(layout)
        try:
            for g, s, o, ld in itertools.product(nt_geometry, nt_state, occupancy, lds):
                if ld and grp != 1:
                    continue  # the LDS-staged sweep: groups of 1 only
                ctx.tune(nt_geometry=g, nt_state=s, occupancy=o, lds=ld)
                ctx.explicit_step(warm)
                t0 = time.perf_counter()
                # ranked by the step as it runs: a captured graph of n steps
                # (with the gated sweep the gating is part of the sweep
                # launch, otherwise a launch of its own), timed by the
                # events around the call
                n = max(reps, 16)
                ctx.prepare(n, STEP_GRAPH)
                ctx.explicit_step(n, STEP_GRAPH)
                total, _ = ctx.last_timing()
                ms = total / n
                cfg = {**layout, "nt_geometry": g, "nt_state": s, "occupancy": o, "lds": ld}
                trials.append((cfg, ms))
                if log:
                    log(f"autotune {cfg}: step {ms * 1e3:.2f} us ({time.perf_counter() - t0:.2f} s)")
        finally:
            ctx.close()
    best = min(trials, key=lambda x: x[1])[0]
    return best, trials


def apply(ctx: Context, cfg: dict, layout: bool) -> Context:
    """Applies the layout keys (before any mesh is added) or the kernel keys."""
    keys = LAYOUT_KEYS if layout else KERNEL_KEYS
    ctx.tune(**{k: v for k, v in cfg.items() if k in keys})
    return ctx
