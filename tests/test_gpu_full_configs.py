"""Parity at the full sizes of the workloads: the 500 x 500 bilinear RVE
(500 k elements), the 2000 x 2000 Ortiz-Pandolfi RVE (8 M), the 120^3 x 6 tet4
cube (10.4 M) and the batch of 256 RVEs (31.4 M) -- one K + R assembly per
schedule against the oracle, which runs element range by element range into
the harness's own pattern (bounded memory) on the host threads.  Bars: K within
1e-12 (FP64 schedules) or 1e-6 (mixed) relative Frobenius, P within 1e-12, trial
ISVs identical; committed histories random, so that the Ortiz-Pandolfi RVE
takes both its loading and its unloading branch."""
import os
import time

import numpy as np
import pytest

from pigfem_amd import AssemblerGPU, PMatrix, PVector, RVEProblem, batch_config, config
from oracle.oracle import assemble_in_ranges
from tests.helpers import rel, u_scale

pytestmark = pytest.mark.gpu

MODES = ["fp64", "binned64", "mixed", "mixed_cut64"]
THREADS = max(1, min(16, int(os.environ.get("OMP_NUM_THREADS", "0")) or (os.cpu_count() or 1)))


@pytest.mark.parametrize("name", ["C2", "C3", "C4", "C5"])
def test_full_size_parity(name):
    t0 = time.time()
    p = RVEProblem.batch(batch_config(256)) if name == "C5" else RVEProblem(config(name))
    cfg = p.cfg
    U = p.random_u(7, u_scale(cfg))
    ic = np.random.default_rng(8).uniform(0.0, 1e-3, p.n_cqp) if p.n_cqp else None
    print(f"full size {name}: {p.stats()}, built in {time.time() - t0:.1f}s", flush=True)
    t = time.time()
    Kr, Pr, Ir = assemble_in_ranges(p, U, ic, threads=THREADS)
    print(f"full size {name}: oracle on {THREADS} threads in {time.time() - t:.1f}s", flush=True)
    asm = AssemblerGPU(p, "fp64")
    if p.n_cqp:
        asm.committed_isvs[:] = ic
    for mode in MODES:
        asm.mode = mode
        K, P = PMatrix(p), PVector(p)
        it = np.zeros(p.n_cqp)
        asm.assemble(K, P, 0.0, it, U, True, True)
        eK, eP = rel(K.values, Kr), rel(P.values, Pr)
        same = np.array_equal(it, Ir)
        print(f"full size {name} {mode}: rel K {eK:.3e}, rel P {eP:.3e}, ISVs identical {same}", flush=True)
        tol = 1e-12 if mode in ("fp64", "binned64") else 1e-6
        assert eK <= tol, (name, mode, eK)
        assert eP <= 1e-12, (name, mode, eP)
        assert same, (name, mode)
        del K, P
    asm.close()
