"""Batches of independent RVEs merged into one block-diagonal problem
(pgh_build_batch): each instance's slice of the merged K and P equals the
instance assembled on its own, bit for bit, on the oracle; the merged pattern
is block diagonal."""
import dataclasses

import numpy as np
import pytest

from pigfem_amd import RVEProblem, batch_config, small_config
from oracle.oracle import Oracle


def small_batch():
    cfgs = []
    for i, (n, ni, kind) in enumerate([(16, 2, "bl"), (24, 3, "bl"), (20, 2, "bl")]):
        cfgs.append(dataclasses.replace(small_config(kind, n=n, n_incl=ni, seed=30 + i), name=f"b{i}"))
    return cfgs


def _instance_u(B, singles, seed=3):
    """Random U on the merged dofs and the same values on each instance's dofs."""
    rng = np.random.default_rng(seed)
    U = rng.uniform(-1e-5, 1e-5, B.n_dof)
    Us = []
    for i, p in enumerate(singles):
        o = B.inst_off[i]
        u = np.empty(p.n_dof)
        u[: p.n_free] = U[o[3]: o[3] + p.n_free]
        u[p.n_free:] = U[B.n_free + o[4]: B.n_free + o[4] + p.n_pres]
        Us.append(u)
    return U, Us


def test_batch_offsets_and_block_diagonal_pattern():
    cfgs = small_batch()
    B = RVEProblem.batch(cfgs)
    singles = [RVEProblem(c) for c in cfgs]
    assert B.n_inst == 3
    assert B.n_elem == sum(p.n_elem for p in singles) and B.n_cut == sum(p.n_cut for p in singles)
    assert B.n_cqp == sum(p.n_cqp for p in singles) and B.nnz == [sum(p.nnz[b] for p in singles) for b in range(4)]
    for i, p in enumerate(singles):
        (kr, pr) = B.instance_ranges(i)
        o = B.inst_off[i]
        # rows of instance i only reference columns of instance i, shifted
        col0 = [o[3], o[4], o[3], o[4]]
        for b in range(4):
            lo, hi = kr[b]
            base = int(np.cumsum([0] + B.nnz)[b])
            assert hi - lo == p.nnz[b]
            assert np.array_equal(B.ci[b][lo - base: hi - base] - col0[b], p.ci[b])
        assert np.array_equal(np.diff(B.inst_off[:, 7]), [c.n_incl for c in cfgs])


def test_batch_slices_equal_single_instances_on_the_oracle():
    cfgs = small_batch()
    B = RVEProblem.batch(cfgs)
    singles = [RVEProblem(c) for c in cfgs]
    U, Us = _instance_u(B, singles)
    ic = np.random.default_rng(4).uniform(0, 5e-5, B.n_cqp)
    K, P, It = Oracle(B).assemble(U, ic)
    for i, p in enumerate(singles):
        o, q = B.inst_off[i], B.inst_off[i + 1]
        Ki, Pi, Ii = Oracle(p).assemble(Us[i], ic[o[5]: q[5]])
        kr, pr = B.instance_ranges(i)
        assert np.array_equal(np.concatenate([K[a:b] for a, b in kr]), Ki)
        assert np.array_equal(np.concatenate([P[a:b] for a, b in pr]), Pi)
        assert np.array_equal(It[o[5]: q[5]], Ii)


def test_batch_config_rules():
    cfgs = batch_config(256)
    ns = np.array([c.n for c in cfgs])
    assert len(cfgs) == 256 and ns.min() >= 64 and ns.max() <= 512
    for c in cfgs:
        h = 1.0 / c.n
        assert c.r_min == pytest.approx(4 * h) and c.r_max == pytest.approx(10 * h)
        assert 1 <= c.n_incl <= 150 and c.cohesive == 1
        mean_area = np.pi * (c.r_min ** 2 + c.r_min * c.r_max + c.r_max ** 2) / 3
        assert c.n_incl * mean_area <= 0.45 or c.n_incl == 1
