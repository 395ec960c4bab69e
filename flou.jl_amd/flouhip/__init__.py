"""flouhip -- MI355X (gfx950) DGSEM hot path of Flou.jl behind a C ABI.

Host-side mirror of the Flou types a user composes to reach rhs! / timeintegrate; every
kernel runs in libflouhip.so (flou.jl_amd/libflouhip.so, see include/flouhip.h).  VTKHDF
files are written by libflouvtkhdf.so (include/flouvtkhdf.h), loaded on first use.  Importing
this package fails with ImportError when the library is missing: there is no CPU path.
"""
from . import _capi
from ._capi import DomainError
from .disc import (CarpenterKennedy2N54, EquationConfig, MonitorOutput, MultielementDisc, ORK256,
                   SSPRK33, ZhangShuLimiter, get_limiter, get_limiter_callback, get_max_dt,
                   get_monitor, get_monitor_callback, integrate, list_limiters, list_monitors,
                   make_callback_list, rhs_, rk_stage, timeintegrate)
from .equations import (ChandrasekharAverage, EulerEquation, EulerInflowBC, EulerOutflowBC,
                        EulerSlipBC, GenericBC, HybridDivOperator, KennedyGruber, LinearAdvection, LxF,
                        MatrixDissipation, ScalarDissipation, SplitDivOperator, StdAverage, StrongDivOperator,
                        normal_shockwave, nvariables, pressure, soundvelocity, variablenames,
                        vars_cons2prim, vars_prim2cons)
from .io import (FlouFile, add_celldata_, add_fielddata_, add_pointdata_, add_solution_, close_file_,
                 get_save_callback, open_for_write, pointdata2VTKHDF)
from .mesh import CartesianMesh, apply_periodicBCs
from .stdregions import DGSEMrec, LagrangeBasis, StdHex, StdQuad, StdSegment

LIB_PATH = _capi.LIB_PATH


def version():
    return _capi.lib.flou_hip_version().decode()


__all__ = [
    "CarpenterKennedy2N54", "CartesianMesh", "ChandrasekharAverage", "DGSEMrec", "DomainError",
    "EquationConfig", "EulerEquation", "EulerInflowBC", "EulerOutflowBC", "EulerSlipBC",
    "FlouFile", "GenericBC", "HybridDivOperator", "KennedyGruber", "LagrangeBasis", "LinearAdvection", "LxF", "MatrixDissipation",
    "MonitorOutput", "MultielementDisc", "ORK256", "SSPRK33", "ScalarDissipation", "SplitDivOperator", "StdAverage", "StdHex",
    "StdQuad", "StdSegment", "StrongDivOperator", "ZhangShuLimiter", "add_celldata_", "add_fielddata_",
    "add_pointdata_", "add_solution_", "apply_periodicBCs", "close_file_", "get_limiter", "get_limiter_callback", "get_max_dt", "get_monitor", "get_monitor_callback",
    "integrate", "list_limiters", "list_monitors", "make_callback_list", "normal_shockwave",
    "nvariables", "open_for_write", "pointdata2VTKHDF", "pressure", "rhs_", "rk_stage", "soundvelocity",
    "timeintegrate", "variablenames", "vars_cons2prim", "vars_prim2cons", "version",
]
