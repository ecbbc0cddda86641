"""Host mirror of PIGFEM's material-parameter sensitivity over the C ABI.

`SensitivityMaterialParameterGPU` stands for the SensitivityMaterialParameter
that a material's getSensitivityParameter returns
(src/material_lei.cpp:365-385, src/Material_OPCohesive*.cpp): it names a
material and a parameter, and `assemble_dPdd` computes dP/dd on the GPU
(pg_asm_sensitivity_dPdd), replacing SensitivityParameter::assemble_dPdd
(src/SensitivityParameter.cpp:108-151).  For an LEI parameter the derivative of
D is formed here, as LinearElasticIsotropicMaterial::computeDmatSensitivities
does it (src/material_lei.cpp:106-210), from the material's E and nu.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _ffi
from ._ffi import PgSensParam, check, lib

_LEI_NAMES = {
    "E": "E", "YOUNG'S MODULUS": "E", "YOUNGS MODULUS": "E",
    "NU": "NU", "POISSON'S RATIO": "NU", "POISSONS RATIO": "NU",
    "LAMBDA": "LAMBDA", "FIRST LAME CONSTANT": "LAMBDA",
    "MU": "MU", "G": "MU", "SECOND LAME CONSTANT": "MU", "SHEAR MODULUS": "MU",
    "K": "K", "BULK MODULUS": "K",
}
_OP_NAMES = {
    "SIGMA": "SIGMA_C", "SIGMA_C": "SIGMA_C", "SIGMA C": "SIGMA_C", "SIGMA_CRITICAL": "SIGMA_C",
    "SIGMA CRITICAL": "SIGMA_C",
    "DELTA": "DELTA_C", "DELTA_C": "DELTA_C", "DELTA C": "DELTA_C", "DELTA_CRITICAL": "DELTA_C",
    "DELTA CRITICAL": "DELTA_C",
}


def lei_dmat_sensitivity(E: float, nu: float, dim: int, param: str, plane_strain: bool = True) -> np.ndarray:
    """dD/dd of an isotropic material (nstr x nstr), as
    LinearElasticIsotropicMaterial::computeDmatSensitivities forms it
    (src/material_lei.cpp:106-210): for E, lambda, mu and K the matrix of D's
    pattern scaled by coefs[p] (D / d at fixed nu); for nu the derivative at
    fixed E."""
    p = {"E": 0, "LAMBDA": 1, "MU": 2, "K": 3, "NU": 4}[param]
    if dim == 2:
        if plane_strain:
            coefs = [1.0 / ((1.0 + nu) * (1.0 - 2.0 * nu)), 1.0 / nu, 2.0 / (1.0 - 2.0 * nu), 3.0 / (1.0 + nu)]
        else:
            coefs = [1.0 / (1.0 - nu * nu), (1.0 - 2.0 * nu) / (nu * (1.0 - nu)), 2.0 / (1.0 - nu),
                     3.0 * (1.0 - 2.0 * nu) / (1.0 - nu * nu)]
        S = np.zeros((3, 3))
        if p < 4:
            c = coefs[p]
            if plane_strain:
                S[0, 0] = c * (1.0 - nu); S[0, 1] = c * nu
                S[1, 0] = c * nu;         S[1, 1] = c * (1.0 - nu)
                S[2, 2] = c * (1.0 - 2.0 * nu) / 2.0
            else:
                S[0, 0] = c;       S[0, 1] = c * nu
                S[1, 0] = c * nu;  S[1, 1] = c
                S[2, 2] = c * (1.0 - nu) / 2.0
        elif plane_strain:
            coef = E / ((1.0 + nu) * (1.0 - 2.0 * nu))
            dcoef = E * (1.0 + 4.0 * nu) / pow(1.0 - nu - 2.0 * nu * nu, 2)
            S[0, 0] = dcoef * (1.0 - nu) - coef; S[0, 1] = dcoef * nu + coef
            S[1, 0] = dcoef * nu + coef;         S[1, 1] = dcoef * (1.0 - nu) - coef
            S[2, 2] = dcoef * (1.0 - 2.0 * nu) / 2.0 - coef
        else:
            coef = E / (1.0 - nu * nu)
            dcoef = 2.0 * E * nu / pow(1.0 - nu * nu, 2)
            S[0, 0] = dcoef;              S[0, 1] = dcoef * nu + coef
            S[1, 0] = dcoef * nu + coef;  S[1, 1] = dcoef
            S[2, 2] = dcoef * (1.0 - nu) / 2.0 - 0.5 * coef
        return S
    S = np.zeros((6, 6))
    if p < 4:
        c = [1.0 / ((1.0 + nu) * (1.0 - 2.0 * nu)), 1.0 / nu, 2.0 / (1.0 - 2.0 * nu), 3.0 / (1.0 + nu)][p]
        for i in range(3):
            for j in range(3):
                S[i, j] = c * (1.0 - nu) if i == j else c * nu
        for i in range(3, 6):
            S[i, i] = c * (1.0 - 2.0 * nu) / 2
    else:
        coef = E / ((1.0 + nu) * (1.0 - 2.0 * nu))
        dcoef = E * (1.0 + 4.0 * nu) / pow(1.0 - nu - 2.0 * nu * nu, 2)
        for i in range(3):
            for j in range(3):
                S[i, j] = dcoef * (1.0 - nu) - coef if i == j else dcoef * nu + coef
        for i in range(3, 6):
            S[i, i] = dcoef * (1.0 - 2.0 * nu) / 2.0 - coef
    return S


class SensitivityMaterialParameterGPU:
    """One material parameter d of the problem assembled by `assembler` (an
    AssemblerGPU): `mat` is the material's index in the problem description,
    `param_name` a name its getSensitivityParameter accepts.  LEI parameters
    need the material's E and nu (and plane_strain in 2D) to form dD/dd.
    own_child_b: give each child qp of a cut element its own B matrix instead
    of the reference's child-local index (PG_SENS_OWN_CHILD_B)."""

    def __init__(self, assembler, mat: int, param_name: str, E: float = None, nu: float = None,
                 plane_strain: bool = True, own_child_b: bool = False):
        p = assembler.problem
        self.assembler = assembler
        self.problem = p
        m = p.desc.mats[mat]
        self.prm = PgSensParam()
        self.prm.mat = int(mat)
        name = param_name.upper()
        if m.type == _ffi.PG_MAT_LEI:
            if name not in _LEI_NAMES:
                raise ValueError(f"{param_name} is an invalid parameter name for a linear elastic isotropic material!")
            if E is None or nu is None:
                raise ValueError("an LEI parameter needs the material's E and nu")
            self.param_name = _LEI_NAMES[name]
            S = lei_dmat_sensitivity(float(E), float(nu), p.dim, self.param_name, plane_strain)
            ns = S.shape[0]
            for i in range(ns):
                for j in range(ns):
                    self.prm.dD[i * ns + j] = S[i, j]
            self.prm.kind = _ffi.PG_SENS_LEI
        elif m.type in (_ffi.PG_MAT_OP, _ffi.PG_MAT_OP_NU):
            if name not in _OP_NAMES:
                raise ValueError(f"{param_name} is not a valid parameter name")
            self.param_name = _OP_NAMES[name]
            self.prm.kind = _ffi.PG_SENS_SIGMA_C if self.param_name == "SIGMA_C" else _ffi.PG_SENS_DELTA_C
        else:
            raise ValueError("Sensitivity is not implemented for the bilinear cohesive model")
        self.flags = _ffi.PG_SENS_OWN_CHILD_B if own_child_b else 0

    def assemble_dPdd(self, solution, isv=None, isv_sens=None) -> np.ndarray:
        """dP/dd = [dP_f | dP_p] at the host solution (global-dof order), with the
        committed cohesive internal variables and their sensitivities for this
        parameter (per cohesive qp; zeros if None)."""
        from .device import DeviceBuffer, synchronize
        p = self.problem
        U = np.ascontiguousarray(solution, dtype=np.float64)
        if U.shape != (p.n_dof,):
            raise ValueError(f"solution must have {p.n_dof} entries")
        dU = DeviceBuffer.from_numpy(U)
        dO = DeviceBuffer.zeros(p.n_dof)
        di = ds = None
        if p.n_cqp:
            di = DeviceBuffer.from_numpy(np.zeros(p.n_cqp) if isv is None else np.asarray(isv, np.float64))
            ds = DeviceBuffer.from_numpy(np.zeros(p.n_cqp) if isv_sens is None else np.asarray(isv_sens, np.float64))
        check(lib().pg_asm_sensitivity_dPdd(self.assembler._h, C.byref(self.prm), self.flags, dU.ptr,
                                            di.ptr if di else None, ds.ptr if ds else None, dO.ptr, None),
              "pg_asm_sensitivity_dPdd")
        synchronize()
        return dO.download(p.n_dof)
