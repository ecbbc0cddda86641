"""Newton history of the 500 x 500 bilinear-cohesive RVE (C2: 200 fibers,
10 load steps to 2 % uniaxial strain), restating PIGFEM's load stepping
(src/SolverNonlinear.cpp:70-510, src/SolverNonlinearKSP.cpp:55-155).

  python tests/golden/newton_c2.py --oracle
      CPU oracle assembly; writes tests/golden/newton_c2_oracle.json (the
      committed fixture: iteration counts of every attempted load step, the
      explicit steps, the time reached and how the run ended).
  python tests/golden/newton_c2.py --modes fp64,mixed [--out DIR]
      the same history with the GPU assembly (one MI355X); prints each step,
      writes DIR/newton_c2_<mode>.json and compares with the fixture.
"""
import argparse
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from pigfem_amd import RVEProblem, config  # noqa: E402
from pigfem_amd.newton import NewtonDriver, NewtonError  # noqa: E402

FIXTURE = os.path.join(HERE, "newton_c2_oracle.json")


def run(prob, assemble, label):
    t0 = time.time()
    log = lambda m: print(f"[{label}] {m} [{time.time() - t0:.0f}s]", flush=True)
    d = NewtonDriver(prob, assemble, log=log)
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
    ap.add_argument("--oracle", action="store_true")
    ap.add_argument("--modes", default="")
    ap.add_argument("--out", default="gpurun_out")
    args = ap.parse_args()
    prob = RVEProblem(config("C2"))
    print("C2:", prob.stats(), flush=True)
    if args.oracle:
        from oracle.oracle import Oracle
        o = Oracle(prob)
        res = run(prob, lambda U, isv: o.assemble(U, isv), "oracle")
        with open(FIXTURE, "w") as f:
            json.dump(res, f)
        print(json.dumps({k: res[k] for k in ("status", "t_reached", "seconds")}))
    if args.modes:
        from pigfem_amd import AssemblerGPU
        from pigfem_amd.newton import gpu_assemble_fn
        ref = json.load(open(FIXTURE))
        os.makedirs(args.out, exist_ok=True)
        for mode in args.modes.split(","):
            res = run(prob, gpu_assemble_fn(AssemblerGPU(prob, mode)), mode)
            res["vs_oracle"] = compare(ref, res)
            with open(os.path.join(args.out, f"newton_c2_{mode}.json"), "w") as f:
                json.dump(res, f)
            print(json.dumps({"mode": mode, "status": res["status"], "t_reached": res["t_reached"],
                              "vs_oracle": res["vs_oracle"], "seconds": res["seconds"]}), flush=True)


if __name__ == "__main__":
    main()
