"""Autotuner of the class-binned schedules.

Sweeps the schedule parameters of `pg_asm_set_tiles` -- nodes per block of the
node-group kernel and the variant of the small-class cut kernel (elements per
wavefront, register cap) -- timing each with HIP events on the assembly stream,
and keeps the fastest.  Results are cached in a JSON file keyed by the problem
class (dimension, cohesive law, size bucket) and the mode.
"""
from __future__ import annotations

import json
import math
import os

import numpy as np

from .device import DeviceBuffer, synchronize

CACHE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tuning_cache.json")
GROUPS = {2: (8, 16, 32), 3: (2, 4, 8)}
CUT_VARIANTS = (0, 1, 2, 3)


def _key(problem, mode):
    bucket = int(round(math.log2(max(problem.n_elem, 1))))
    return f"d{problem.dim}-coh{problem.cfg.cohesive}-e2^{bucket}-{mode}"


def _load():
    try:
        with open(CACHE) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def tune(asm, U=None, reps: int = 5, use_cache: bool = True, log=None):
    """Pick (group_nodes, cut_variant) for `asm` (an AssemblerGPU in a binned mode);
    returns (choice, table of mean kernel ms per candidate)."""
    p = asm.problem
    key = _key(p, asm.mode)
    cache = _load()
    if use_cache and key in cache:
        g, c = cache[key]["choice"]
        asm.set_tiles(g, c)
        return (g, c), cache[key].get("table", {})
    if U is None:
        U = p.random_u(0)
    dU = DeviceBuffer.from_numpy(U)
    dIc = DeviceBuffer.zeros(max(p.n_cqp, 1))
    dIt = DeviceBuffer.zeros(max(p.n_cqp, 1))
    dK = DeviceBuffer.zeros(p.nnz_total)
    dP = DeviceBuffer.zeros(p.n_dof)
    ic = dIc if p.n_cqp else None
    it = dIt if p.n_cqp else None
    table = {}
    asm.set_timing(True)
    try:
        for g in GROUPS[p.dim]:
            for c in CUT_VARIANTS:
                asm.set_tiles(g, c)
                asm.assemble_device(dU, ic, it, dK, dP)  # warm
                ms = []
                for _ in range(reps):
                    asm.assemble_device(dU, ic, it, dK, dP)
                    ms.append(asm.last_timing()[0])
                table[f"{g},{c}"] = float(np.mean(ms))
                if log:
                    log(f"[tune] group_nodes={g} cut_variant={c}: {table[f'{g},{c}']:.3f} ms")
    finally:
        asm.set_timing(False)
        synchronize()
    best = min(table, key=table.get)
    g, c = (int(x) for x in best.split(","))
    asm.set_tiles(g, c)
    cache[key] = {"choice": [g, c], "table": table}
    try:
        with open(CACHE, "w") as f:
            json.dump(cache, f, indent=1, sort_keys=True)
    except OSError:
        pass
    return (g, c), table
