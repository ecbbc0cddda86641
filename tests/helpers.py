"""Shared helpers of the test suite (dense / sparse views of the 4-block K)."""
import numpy as np
import scipy.sparse as sp


def full_matrix(problem, values, rp=None, ci=None):
    """Assemble the global n_dof x n_dof matrix (global dof numbering) from the
    concatenated [ff|fp|pf|pp] values of the given (or the harness) pattern."""
    rp = rp if rp is not None else problem.rp
    ci = ci if ci is not None else problem.ci
    nf, npres = problem.n_free, problem.n_pres
    shapes = [(nf, nf), (nf, npres), (npres, nf), (npres, npres)]
    off = 0
    blocks = []
    for b in range(4):
        n = int(rp[b][-1])
        blocks.append(sp.csr_matrix((values[off:off + n], ci[b], rp[b]), shape=shapes[b]))
        off += n
    return sp.bmat([[blocks[0], blocks[1]], [blocks[2], blocks[3]]], format="csr")


def rel(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (nb if nb > 0 else 1.0)


def u_scale(cfg):
    """Random-displacement scale: the bilinear law's contact branch overflows
    (sinh of alpha*delta_n/delta_c1) for openings near 1e-3, as the reference does."""
    from pigfem_amd import _ffi
    return 1e-5 if cfg.cohesive in (_ffi.PGH_COH_BL, _ffi.PGH_COH_BL_NU) else 1e-3
