"""Autotuner of the class-binned schedules.

Sweeps the schedule parameters of `pg_asm_set_tiles` -- nodes per block, threads
per node and form of the node-group kernel, overlap, and the variant of the
small-class cut kernel (elements per wavefront, register cap) jointly, then the
size of the groups that run after the cut kernels -- timing the whole assembly
with HIP events, and keeps the fastest.  Every choice gives
bit-identical results.  Results are cached in a JSON file keyed by the problem
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
# (bit 7: the structured-grid nodes through the generic kernels instead)
GROUP_VARIANTS = {2: (0, 1, 2, 3, 4, 6), 3: (0, 1, 4, 5, 128, 132)}
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
    """Pick (group_nodes, cut_variant, group_variant) for `asm` (an AssemblerGPU in
    a binned mode); returns (choice, table of mean kernel ms per candidate)."""
    p = asm.problem
    key = _key(p, asm.mode)
    cache = _load()
    if use_cache and key in cache and len(cache[key]["choice"]) == 3:
        choice = tuple(cache[key]["choice"])
        asm.set_tiles(*choice)
        return choice, cache[key].get("table", {})
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

    def timed(g, c, v, which):
        asm.set_tiles(g, c, v)
        asm.assemble_device(dU, ic, it, dK, dP)  # warm
        ms = []
        for _ in range(reps):
            asm.assemble_device(dU, ic, it, dK, dP)
            ms.append(asm.last_timing()[which])
        return float(np.mean(ms))

    asm.set_timing(True)
    try:
        # node-group size x group variant x cut-kernel variant, each timed on
        # the whole assembly (the groups of nodes without cut incidences overlap
        # the cut kernel, so the choices are coupled)
        for g in GROUPS[p.dim]:
            for v in GROUP_VARIANTS[p.dim]:
                for c in CUT_VARIANTS:
                    table[f"{g},{c},{v}"] = timed(g, c, v, 0)
                    if log:
                        log(f"[tune] group_nodes={g} cut_variant={c} group_variant={v}: {table[f'{g},{c},{v}']:.3f} ms")
        g, c, v = (int(x) for x in min(table, key=table.get).split(","))
        # the groups that wait for the cut kernels may take another size
        # (group variant bits 5-6: 1 + index of the size)
        if not v & 4:
            for k, g2 in enumerate(GROUPS[p.dim]):
                if g2 == g:
                    continue
                v2 = v | ((k + 1) << 5)
                table[f"{g},{c},{v2}"] = timed(g, c, v2, 0)
                if log:
                    log(f"[tune] group_nodes={g} cut_variant={c} group_variant={v2}: {table[f'{g},{c},{v2}']:.3f} ms")
            g, c, v = (int(x) for x in min(table, key=table.get).split(","))
        # consecutive groups per block of the structured-grid kernels (group
        # variant bits 8-9)
        if asm.grid_nodes() > 0:
            for k in (1, 2, 3):
                v2 = (v & 255) | (k << 8)
                table[f"{g},{c},{v2}"] = timed(g, c, v2, 0)
                if log:
                    log(f"[tune] group_nodes={g} cut_variant={c} group_variant={v2}: {table[f'{g},{c},{v2}']:.3f} ms")
            g, c, v = (int(x) for x in min(table, key=table.get).split(","))
    finally:
        asm.set_timing(False)
        synchronize()
    asm.set_tiles(g, c, v)
    cache[key] = {"choice": [g, c, v], "table": table}
    try:
        with open(CACHE, "w") as f:
            json.dump(cache, f, indent=1, sort_keys=True)
    except OSError:
        pass
    return (g, c, v), table
