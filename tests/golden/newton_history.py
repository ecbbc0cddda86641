"""Newton load-step histories of the cohesive RVEs, restating PIGFEM's load
stepping (src/SolverNonlinear.cpp:70-510, src/SolverNonlinearKSP.cpp:55-155).

Cases:
  c2    the 500 x 500 bilinear-cohesive RVE with 200 fibers, 10 load steps to
        2 % uniaxial strain;
  c3s5  the 2000 x 2000 Ortiz-Pandolfi RVE with 3000 fibers, the same load
        history, first five load steps only (each Newton iteration factorises
        a 9.9 M-dof system on the host);
  c5x4  the four smallest RVEs (64-66 elements a side) of the 256-RVE batch,
        merged into one block-diagonal problem as the batched assembly sees
        them, the same load history.

  python tests/golden/newton_history.py --case c2 --oracle
      CPU oracle assembly; writes tests/golden/newton_<case>_oracle.json (the
      committed fixture: iteration counts of every attempted load step, the
      explicit steps, the time reached and how the run ended).
  python tests/golden/newton_history.py --case c2 --modes fp64,mixed [--out DIR]
      the same history with the GPU assembly (one MI355X); prints each step,
      writes DIR/newton_<case>_<mode>.json and compares with the fixture.
"""
import argparse
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from pigfem_amd import RVEProblem, batch_config, config  # noqa: E402
from pigfem_amd.newton import NewtonDriver, NewtonError  # noqa: E402


def fixture(case):
    return os.path.join(HERE, f"newton_{case}_oracle.json")


def problem(case):
    if case == "c2":
        return RVEProblem(config("C2"))
    if case == "c3s5":
        return RVEProblem(config("C3"))
    if case == "c5x4":
        cfgs = batch_config(256)
        pick = sorted(range(len(cfgs)), key=lambda i: (cfgs[i].n, i))[:4]
        return RVEProblem.batch([cfgs[i] for i in pick])
    raise ValueError(f"unknown case {case}")


def run(prob, assemble, label, max_steps=None):
    t0 = time.time()
    log = lambda m: print(f"[{label}] {m} [{time.time() - t0:.0f}s]", flush=True)
    d = NewtonDriver(prob, assemble, log=log, max_steps=max_steps)
    status = "converged"
    try:
        d.solve()
    except NewtonError as e:
        status = f"stopped: {e}"
    t_end = max([r.t for r in d.history if r.converged] + [0.0])
    return {"label": label, "status": status, "t_reached": t_end, "solver": d.lin.kind,
            "counts": d.iteration_counts(), "seconds": time.time() - t0}


def compare(ref, out):
    a, b = ref["counts"], out["counts"]
    first = next((i for i, (x, y) in enumerate(zip(a, b)) if list(x) != list(y)), None)
    if first is None and len(a) != len(b):
        first = min(len(a), len(b))
    return {"identical": first is None, "first_difference": first,
            "ref_len": len(a), "len": len(b)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--case", default="c2", choices=("c2", "c3s5", "c5x4"))
    ap.add_argument("--oracle", action="store_true")
    ap.add_argument("--modes", default="")
    ap.add_argument("--out", default="gpurun_out")
    args = ap.parse_args()
    prob = problem(args.case)
    max_steps = 5 if args.case == "c3s5" else None
    print(f"{args.case}:", prob.stats(), flush=True)
    if args.oracle:
        from oracle.oracle import Oracle
        o = Oracle(prob)
        res = run(prob, lambda U, isv: o.assemble(U, isv), "oracle", max_steps)
        with open(fixture(args.case), "w") as f:
            json.dump(res, f)
        print(json.dumps({k: res[k] for k in ("status", "t_reached", "seconds")}))
    if args.modes:
        from pigfem_amd import AssemblerGPU
        from pigfem_amd.newton import gpu_assemble_fn
        ref = json.load(open(fixture(args.case)))
        os.makedirs(args.out, exist_ok=True)
        for mode in args.modes.split(","):
            res = run(prob, gpu_assemble_fn(AssemblerGPU(prob, mode)), mode, max_steps)
            res["vs_oracle"] = compare(ref, res)
            with open(os.path.join(args.out, f"newton_{args.case}_{mode}.json"), "w") as f:
                json.dump(res, f)
            print(json.dumps({"mode": mode, "status": res["status"], "t_reached": res["t_reached"],
                              "vs_oracle": res["vs_oracle"], "seconds": res["seconds"]}), flush=True)


if __name__ == "__main__":
    main()
