"""Benchmark: K+R assembly throughput of the 2D cohesive RVE on one MI355X.

Workload (default C3): 2000 x 2000 tri3 background mesh (8M elements), 3000 circular
fibers r ~ U[0.004, 0.008], Ortiz-Pandolfi cohesive interfaces, U ~ U[-1e-3, 1e-3].
One step = one full K (4 CSR blocks) + R assembly with inputs resident in HBM.

Prints ONE JSON line: value = elements/s of the mixed-precision class-binned
schedule (FP32 tangents, FP64 damage / sums / residual); the plain FP64 schedule,
the binned FP64 schedule, the speed-ups, K and R relative errors (mixed vs FP64),
the roofline of the dominant kernel and the CPU baseline are reported beside it.
The 3D cube (C4) and the batch of 256 RVEs (C5) are then measured the same way
and reported under their own keys (--extra '' skips them).
"""
from __future__ import annotations

import argparse
import gc
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def workload(cfg, batch, uscale):
    law = {0: "perfect bond", 1: "bilinear cohesive", 2: "Ortiz-Pandolfi cohesive"}.get(cfg.cohesive, str(cfg.cohesive))
    if batch:
        return ("C5: 256 independent tri3 RVEs, n x n with n log-uniform in [64,512], 5-150 fibers each "
                f"(area fraction <= 0.45), r~U[4h,10h], bilinear cohesive, one launch sequence, U~U[-{uscale:g},{uscale:g}]")
    if cfg.dim == 3:
        return (f"{cfg.name}: {cfg.n}^3 cubes x 6 tet4, {cfg.n_incl} spheres r~U[{cfg.r_min},{cfg.r_max}], "
                f"{law}, U~U[-{uscale:g},{uscale:g}]")
    return (f"{cfg.name}: {cfg.n}x{cfg.n} tri3 RVE, {cfg.n_incl} fibers r~U[{cfg.r_min},{cfg.r_max}], "
            f"{law}, U~U[-{uscale:g},{uscale:g}]")


def build_problem(name, device_pattern=True):
    """The workload's set-up; detection, enrichment and the pattern stage run
    on the GPU unless device_pattern is False.  Returns the problem, its
    config, the batch's configs (or None) and the set-up wall seconds (total
    and per stage)."""
    import dataclasses
    from pigfem_amd import RVEProblem, batch_config, config
    t0 = time.time()
    if name.upper() == "C5":
        cfgs = batch_config(256)
        prob = RVEProblem.batch(cfgs, device_pattern=device_pattern, device_setup=device_pattern)
        cfg = dataclasses.replace(cfgs[0], name="C5")
    else:
        cfgs = None
        cfg = config(name)
        prob = RVEProblem(cfg, device_pattern=device_pattern, device_setup=device_pattern)
    setup = {"total": time.time() - t0, "setup_on": "gpu" if device_pattern else "host"}
    setup.update({k: round(v, 4) for k, v in prob.stage_seconds().items()})
    log(f"[bench] {cfg.name} built in {setup['total']:.1f}s ({setup}): {prob.stats()}")
    return prob, cfg, cfgs, setup


def node_pair_kernels(dim, tiles, grid_nodes):
    """Names of the node-pair kernels one mixed-schedule assembly launches with
    these tiles: the structured-grid kernel (tri3 / tet4 nodes checked against
    the grid formulas) and the generic node-group kernel for the other nodes."""
    g, _, v = tiles
    tv, lists = v & 1, dim == 3 or bool(v & 2)
    sizes = (8, 16, 32) if dim == 2 else (2, 4, 8)
    rsel = (v >> 5) & 3
    g_rest = sizes[rsel - 1] if rsel else g
    overlap = not v & 4

    def bs(ng, t):
        per = (8 if dim == 2 else 16) * (2 if t else 1) * ng
        return 64 if per < 64 else (512 if per > 512 else (per + 63) // 64 * 64)
    names = []
    if grid_nodes > 0 and not v & (128 | 8):
        gs = g if overlap else g_rest
        if dim == 2:
            names.append(f"k_pair_struct<float, {gs}, {tv}, {bs(gs, tv)}>")
        else:
            # the tet4 grid's table of element tangents (the generator's cube: one spacing,
            # two materials), unless group variant bit 10 turns it off
            tab = "false" if v & 1024 else "true"
            names.append(f"k_pair_struct3<float, {gs}, {tv}, {tab}, {bs(gs, tv)}>")
    for ng in ((g, g_rest) if overlap else (g_rest,)):
        name = (f"k_pair_lists<{dim}, float, float, {ng}, {tv}, {bs(ng, tv)}>" if lists
                else f"k_pair_group<{dim}, float, {ng}, {tv}, {bs(ng, tv)}>")
        if name not in names:
            names.append(name)
    return " + ".join(names)


def measure(prob, cfg, args, full):
    """Time the schedules on one problem; full adds the FP64-cut variant and the
    host-buffer API.  Returns a dict of results (the caller frees the problem)."""
    import numpy as np
    from pigfem_amd import AssemblerGPU, DeviceBuffer, synchronize
    from pigfem_amd.autotune import tune

    asm = AssemblerGPU(prob, "mixed")
    # bilinear law: openings beyond ~1e-4 in compression overflow the contact
    # branch's sinh (alpha = 50, delta_c1 = 5e-5), so its fields are scaled down
    uscale = 1e-5 if cfg.cohesive in (1, 3) else 1e-3
    U = prob.random_u(0, uscale)
    dU = DeviceBuffer.from_numpy(U)
    dIc = DeviceBuffer.zeros(max(prob.n_cqp, 1))
    dIt = DeviceBuffer.zeros(max(prob.n_cqp, 1))
    dK = DeviceBuffer.zeros(prob.nnz_total)
    dP = DeviceBuffer.zeros(prob.n_dof)
    ic = dIc if prob.n_cqp else None
    it = dIt if prob.n_cqp else None

    binned_modes = ("binned64", "mixed", "mixed_cut64") if full else ("binned64", "mixed")
    tuned = {}
    for mode in binned_modes:
        asm.mode = mode
        if args.tiles:
            tuned[mode] = tuple(int(x) for x in args.tiles.split(","))
        else:
            tuned[mode] = tune(asm, U, use_cache=False, log=log)[0]
        log(f"[bench] {cfg.name} tuned {mode}: group_nodes, cut_variant, group_variant = {tuned[mode]}")

    def run(mode, steps):
        asm.mode = mode
        if mode in tuned:
            asm.set_tiles(*tuned[mode])
        for _ in range(steps):
            asm.assemble_device(dU, ic, it, dK, dP)

    results = {}
    for mode in ("fp64",) + binned_modes:
        run(mode, args.warmup)
        synchronize()
        t = time.perf_counter()
        run(mode, args.steps)
        synchronize()
        ms = (time.perf_counter() - t) * 1e3 / args.steps
        asm.check_error()
        results[mode] = {"ms": ms, "eps": prob.n_elem / (ms * 1e-3)}
        log(f"[bench] {cfg.name} {mode}: {ms:.3f} ms/step, {results[mode]['eps']:.4g} elements/s")

    # accuracy of the mixed schedule against the FP64 schedule on the same inputs
    def snapshot(mode):
        asm.mode = mode
        if mode in tuned:
            asm.set_tiles(*tuned[mode])
        asm.assemble_device(dU, ic, it, dK, dP)
        synchronize()
        return (dK.download(prob.nnz_total), dP.download(prob.n_dof),
                dIt.download(prob.n_cqp) if prob.n_cqp else np.zeros(0))
    K64, P64, I64 = snapshot("fp64")
    K32, P32, I32 = snapshot("mixed")
    acc = {"rel_err_K": float(np.linalg.norm(K32 - K64) / np.linalg.norm(K64)),
           "rel_err_P": float(np.linalg.norm(P32 - P64) / np.linalg.norm(P64)),
           "isv_max_abs_diff": float(np.abs(I32 - I64).max()) if prob.n_cqp else 0.0}
    log(f"[bench] {cfg.name} mixed vs fp64: rel K {acc['rel_err_K']:.3e}, rel P {acc['rel_err_P']:.3e}, "
        f"max |isv diff| {acc['isv_max_abs_diff']:.3e}")
    del K32, P32, I32
    if full:
        Kc, Pc_, _ = snapshot("mixed_cut64")
        acc["mixed_cut64_rel_err_K"] = float(np.linalg.norm(Kc - K64) / np.linalg.norm(K64))
        acc["mixed_cut64_rel_err_P"] = float(np.linalg.norm(Pc_ - P64) / np.linalg.norm(P64))
        log(f"[bench] {cfg.name} mixed_cut64 vs fp64: rel K {acc['mixed_cut64_rel_err_K']:.3e}, "
            f"rel P {acc['mixed_cut64_rel_err_P']:.3e}")
        del Kc, Pc_
    del K64, P64, I64

    # per-kernel durations with HIP events on the launch streams (mixed
    # schedule).  For the roofline the node-pair kernels run alone: the tuned
    # tiles with every node group after the cut kernels on one stream (group
    # variant bit 2) and of the tuned group size (bits 5-6 clear), so no other
    # kernel shares the GPU while they run.  The timed steps above keep the
    # tuned schedule, whose side-stream groups run alongside the cut kernels;
    # its stream times are reported beside.
    asm.mode = "mixed"

    def kernel_times(tiles):
        asm.set_tiles(*tiles)
        asm.set_timing(True)
        kt = []
        for _ in range(max(args.steps, 5)):
            asm.assemble_device(dU, ic, it, dK, dP)
            kt.append(asm.last_timing())
        asm.set_timing(False)
        return np.array(kt).mean(axis=0)
    roof_tiles = (tuned["mixed"][0], tuned["mixed"][1], (tuned["mixed"][2] | 4) & ~(3 << 5))
    ko = kernel_times(tuned["mixed"])
    kt = kernel_times(roof_tiles)
    pre_ms, pair_ms = float(kt[1]), float(kt[2])
    B_alg = asm.algorithmic_bytes()
    import ctypes
    from pigfem_amd import _ffi
    cut_bytes = ctypes.sizeof(_ffi.PgCutRec) * prob.n_cut + 16 * prob.n_cqp
    pair_bytes = B_alg - cut_bytes
    achieved = pair_bytes / (pair_ms * 1e-3) / 1e9
    whole = B_alg / (results["mixed"]["ms"] * 1e-3) / 1e9
    log(f"[bench] {cfg.name} events, tiles {roof_tiles}: permute+cut {pre_ms:.3f} ms, node-pair {pair_ms:.3f} ms "
        f"-> {achieved:.0f} GB/s (algorithmic {pair_bytes/1e9:.3f} GB); whole assembly {whole:.0f} GB/s; "
        f"tuned schedule {tuple(tuned['mixed'])}: whole {ko[0]:.3f} ms, cut {ko[1]:.3f}, side-stream groups "
        f"{ko[3]:.3f}, remaining groups {ko[4]:.3f}")

    kname = node_pair_kernels(prob.dim, roof_tiles, asm.grid_nodes())
    traffic = None
    tf = os.path.join(ROOT, "profiles", f"hbm_traffic_{cfg.name}.md")
    if os.path.exists(tf):
        for line in open(tf):
            if line.startswith("traffic_bytes_per_launch:"):
                name, val = line.split(":", 1)[1].rsplit("=", 1)
                if name.strip() == kname:
                    traffic = int(val.split()[0])

    out = {"results": results, "acc": acc, "tuned": tuned, "uscale": uscale,
           "roofline": {"bound": "hbm", "kernel": kname, "achieved": achieved, "peak": HBM_PEAK_GBS,
                        "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                        "algorithmic_bytes_per_launch": pair_bytes, "kernel_ms": pair_ms,
                        "tiles": list(roof_tiles)},
           "tuned_schedule_ms": {"whole": float(ko[0]), "cut": float(ko[1]), "side_stream_groups": float(ko[3]),
                                 "remaining_groups": float(ko[4])},
           "whole_GBps": whole, "B_alg": B_alg, "cut_kernel_ms": pre_ms, "device_bytes": asm.device_bytes(), "U": U,
           "grid_nodes": asm.grid_nodes()}

    if full:
        # host-buffer API (PCIe-inclusive: U in, K / P / ISVs out), mixed schedule
        from pigfem_amd import PMatrix, PVector
        Kh, Ph = PMatrix(prob), PVector(prob)
        ith = np.zeros(prob.n_cqp)
        asm.assemble(Kh, Ph, 0.0, ith, U, True, True)
        t = time.perf_counter()
        for _ in range(3):
            asm.assemble(Kh, Ph, 0.0, ith, U, True, True)
        out["host_ms"] = (time.perf_counter() - t) * 1e3 / 3
        log(f"[bench] {cfg.name} host-buffer API: {out['host_ms']:.2f} ms per call (PCIe-inclusive); "
            f"device memory held {out['device_bytes']/1e9:.2f} GB")
        del Kh, Ph, ith
    asm.close()
    for b in (dU, dIc, dIt, dK, dP):
        b.free()
    return out


def per_instance_rates(cfgs, tuned, uscale, batched_ms):
    """Each instance of the batch assembled on its own (own handle, same
    schedule and tiles): what a loop over 256 separate assemblies achieves."""
    import numpy as np
    from pigfem_amd import AssemblerGPU, DeviceBuffer, RVEProblem, synchronize
    inst_ms, inst_el = [], []
    for i, c in enumerate(cfgs):
        pi = RVEProblem(c)
        ai = AssemblerGPU(pi, "mixed")
        ai.set_tiles(*tuned)
        bu = [DeviceBuffer.from_numpy(pi.random_u(i, uscale)), DeviceBuffer.zeros(max(pi.n_cqp, 1)),
              DeviceBuffer.zeros(max(pi.n_cqp, 1)), DeviceBuffer.zeros(pi.nnz_total), DeviceBuffer.zeros(pi.n_dof)]
        for _ in range(2):
            ai.assemble_device(*bu)
        synchronize()
        t = time.perf_counter()
        for _ in range(5):
            ai.assemble_device(*bu)
        synchronize()
        inst_ms.append((time.perf_counter() - t) * 1e3 / 5)
        inst_el.append(pi.n_elem)
        ai.close()
        del bu, pi
    eps_i = np.array(inst_el) / (np.array(inst_ms) * 1e-3)
    log(f"[bench] separate instances: {sum(inst_ms):.2f} ms total, per-instance elements/s "
        f"min {eps_i.min():.3g} median {np.median(eps_i):.3g} max {eps_i.max():.3g}; batched launch {batched_ms:.2f} ms")
    return {
        "instances": len(cfgs), "elements_min": int(min(inst_el)), "elements_max": int(max(inst_el)),
        "elements_per_s_min": float(eps_i.min()), "elements_per_s_median": float(np.median(eps_i)),
        "elements_per_s_max": float(eps_i.max()),
        "sum_of_separate_ms": float(sum(inst_ms)),
        "separate_aggregate_elements_per_s": float(sum(inst_el) / (sum(inst_ms) * 1e-3)),
        "batched_speedup_vs_separate": float(sum(inst_ms) / batched_ms),
    }


def cpu_baselines(prob, cfg, U, args):
    import numpy as np
    from oracle.oracle import Oracle
    # size the element range for ~cpu_seconds of work from a short probe
    probe = min(prob.n_elem, 20000)
    o = Oracle(prob, 0, probe)
    t = time.perf_counter()
    o.assemble(U, np.zeros(prob.n_cqp))
    rate = probe / (time.perf_counter() - t)
    o.close()
    m = int(min(prob.n_elem, 4_000_000, max(probe, rate * args.cpu_seconds)))
    o = Oracle(prob, 0, m)
    best = 1e30
    for _ in range(2):
        t = time.perf_counter()
        o.assemble(U, np.zeros(prob.n_cqp))
        best = min(best, time.perf_counter() - t)
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            model = next((ln.split(":", 1)[1].strip() for ln in f if ln.startswith("model name")), "")
    except OSError:
        pass
    cpu = {"value": m / best, "unit": "elements/s", "cores": 1, "kind": "port", "cpu_model": model,
           "sample": f"first {m} elements (element-loop order) of the {cfg.name} RVE, K+R, FP64, "
                     f"best of 2 single-thread runs of the oracle restatement"}
    log(f"[bench] cpu baseline: {cpu['value']:.4g} elements/s over {m} elements")
    # the same sample on every host thread the run may use, element colour by
    # colour (elements of one colour share no node)
    nt = max(1, int(os.environ.get("OMP_NUM_THREADS", "0")) or (os.cpu_count() or 1))
    o.assemble(U, np.zeros(prob.n_cqp), threads=nt)  # colouring + warm-up
    best_t = 1e30
    for _ in range(3):
        t = time.perf_counter()
        o.assemble(U, np.zeros(prob.n_cqp), threads=nt)
        best_t = min(best_t, time.perf_counter() - t)
    o.close()
    cpu_threads = {"value": m / best_t, "unit": "elements/s", "cores": nt, "kind": "port", "cpu_model": model,
                   "sample": f"the same {m} elements, colour by colour on {nt} OpenMP threads, best of 3"}
    log(f"[bench] cpu baseline, {nt} threads: {cpu_threads['value']:.4g} elements/s")
    return cpu, cpu_threads


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--config", default="C3")
    ap.add_argument("--extra", default="C4,C5", help="further workloads measured after the main one ('' for none)")
    ap.add_argument("--cpu-seconds", type=float, default=12.0, help="target CPU-baseline sample time")
    ap.add_argument("--skip-cpu", action="store_true")
    ap.add_argument("--per-instance", action="store_true", help="C5 main config: also time the instances one by one")
    ap.add_argument("--no-per-instance", action="store_true",
                    help="C5 as a further workload: skip timing its instances one by one")
    ap.add_argument("--calib-copy", action="store_true", help="also run a 1 GiB device copy (PMC calibration)")
    ap.add_argument("--host-pattern", action="store_true",
                    help="run the set-up (detection, enrichment, pattern) on the host instead of the GPU")
    ap.add_argument("--tiles", default=None,
                    help="group_nodes,cut_variant,group_variant for every schedule instead of autotuning "
                         "(short profiler runs)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (got --gpus {args.gpus}); multi-GPU sharding is design-only",
              file=sys.stderr)
        sys.exit(2)

    prob, cfg, cfgs, setup = build_problem(args.config, not args.host_pattern)
    batch = cfgs is not None
    st = prob.stats()
    m = measure(prob, cfg, args, full=True)
    res = m["results"]
    per_instance = None
    if batch and args.per_instance:
        per_instance = per_instance_rates(cfgs, m["tuned"]["mixed"], m["uscale"], res["mixed"]["ms"])

    if args.calib_copy:
        from pigfem_amd import DeviceBuffer, synchronize
        a = DeviceBuffer(1 << 30)
        b = DeviceBuffer(1 << 30)
        for _ in range(3):
            b.copy_from(a)
        synchronize()
        a.free()
        b.free()
        log("[bench] calibration copy of 1 GiB done (3x)")

    cpu = cpu_threads = None
    if not args.skip_cpu:
        cpu, cpu_threads = cpu_baselines(prob, cfg, m["U"], args)

    out = {
        "metric": "K+R assembled elements/s, % HBM peak; FP32+batched vs FP64 at <=1e-6 rel err",
        "value": res["mixed"]["eps"],
        "unit": "elements/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": res["mixed"]["ms"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64",
        "data": "synthetic",
        "config": {"workload": workload(cfg, batch, m["uscale"]),
                   "elements": prob.n_elem, "cut_elements": prob.n_cut, "cohesive_qps": prob.n_cqp,
                   "dofs": prob.n_dof, "nnz": prob.nnz_total, "parallelism": "single GPU"},
        "roofline": m["roofline"],
        "cpu_baseline": cpu,
        "cpu_baseline_threads": cpu_threads,
        "fp64_elements_per_s": res["fp64"]["eps"],
        "fp64_ms_per_step": res["fp64"]["ms"],
        "speedup_vs_fp64": res["fp64"]["ms"] / res["mixed"]["ms"],
        "speedup_vs_binned64": res["binned64"]["ms"] / res["mixed"]["ms"],
        "binned64_elements_per_s": res["binned64"]["eps"],
        "binned64_ms_per_step": res["binned64"]["ms"],
        "mixed_cut64_elements_per_s": res["mixed_cut64"]["eps"],
        "mixed_cut64_ms_per_step": res["mixed_cut64"]["ms"],
        "mixed_cut64_rel_err_K": m["acc"]["mixed_cut64_rel_err_K"],
        "mixed_cut64_rel_err_P": m["acc"]["mixed_cut64_rel_err_P"],
        "rel_err_K": m["acc"]["rel_err_K"],
        "rel_err_P": m["acc"]["rel_err_P"],
        "isv_max_abs_diff": m["acc"]["isv_max_abs_diff"],
        "whole_assembly_GBps": m["whole_GBps"],
        "whole_assembly_frac": m["whole_GBps"] / HBM_PEAK_GBS,
        "algorithmic_bytes": m["B_alg"],
        "cut_kernel_ms": m["cut_kernel_ms"],
        "tuned_schedule_ms": m["tuned_schedule_ms"],
        "grid_nodes": m["grid_nodes"],
        "setup_s": setup["total"],
        "setup_stages_s": setup,
        "host_api_ms_per_call": m["host_ms"],
        "device_bytes": m["device_bytes"],
        "tiles": m["tuned"],
    }
    if per_instance:
        out["per_instance"] = per_instance
    del prob, m
    gc.collect()

    # further workloads, measured the same way (mixed, binned FP64, plain FP64)
    for name in [x.strip().upper() for x in args.extra.split(",") if x.strip()]:
        if name == args.config.upper():
            continue
        p2, c2, cf2, s2 = build_problem(name, not args.host_pattern)
        m2 = measure(p2, c2, args, full=False)
        r2 = m2["results"]
        out[name] = {
            "workload": workload(c2, cf2 is not None, m2["uscale"]),
            "elements": p2.n_elem, "cut_elements": p2.n_cut, "nnz": p2.nnz_total,
            "ms_per_step": r2["mixed"]["ms"], "elements_per_s": r2["mixed"]["eps"],
            "fp64_ms_per_step": r2["fp64"]["ms"], "binned64_ms_per_step": r2["binned64"]["ms"],
            "speedup_vs_fp64": r2["fp64"]["ms"] / r2["mixed"]["ms"],
            "speedup_vs_binned64": r2["binned64"]["ms"] / r2["mixed"]["ms"],
            "rel_err_K": m2["acc"]["rel_err_K"], "rel_err_P": m2["acc"]["rel_err_P"],
            "roofline": m2["roofline"], "whole_assembly_frac": m2["whole_GBps"] / HBM_PEAK_GBS,
            "cut_kernel_ms": m2["cut_kernel_ms"], "tuned_schedule_ms": m2["tuned_schedule_ms"],
            "grid_nodes": m2["grid_nodes"], "setup_s": s2["total"],
            "setup_stages_s": s2, "tiles": m2["tuned"],
        }
        if cf2 is not None and not args.no_per_instance:
            # the batch config quotes per-instance rates too: each RVE on its own
            out[name]["per_instance"] = per_instance_rates(cf2, m2["tuned"]["mixed"], m2["uscale"], r2["mixed"]["ms"])
        del p2, m2
        gc.collect()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
