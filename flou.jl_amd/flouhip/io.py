"""VTKHDF output of device states: FlouBiz's file interface (src/FlouBiz/FlouBiz.jl:16-115),
its MultielementDisc methods (src/FlouSpatial/IO.jl:16-126) and the save callback of
src/FlouTime/FlouTime.jl:59-90.

The point data are projected onto the equispaced points of every element on the GPU
(flou_hip_equispaced in libflouhip.so) and cross PCIe once; the file itself is written by
libflouvtkhdf.so (include/flouvtkhdf.h) over the HDF5 C library.  Julia's `!` suffix is
`_` here (add_solution! -> add_solution_), as rhs! is rhs_.
"""
import ctypes
import logging
import os

import numpy as np

from . import _capi as C

try:
    import torch
except ImportError:  # pragma: no cover
    torch = None

VTKHDF_LIB_PATH = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "libflouvtkhdf.so")

OK, EINVAL, EHDF5 = 0, -1, -2
F64, F32, I64, U8 = 0, 1, 2, 3

_log = logging.getLogger("flouhip")
_vt = None
_i64p = ctypes.POINTER(ctypes.c_int64)


def _lib():
    """libflouvtkhdf.so, loaded on first use (only output needs HDF5)."""
    global _vt
    if _vt is None:
        if not os.path.exists(VTKHDF_LIB_PATH):
            raise RuntimeError(f"{VTKHDF_LIB_PATH} is not built (no HDF5 at build time): VTKHDF output is unavailable")
        L = ctypes.CDLL(VTKHDF_LIB_PATH)
        L.flou_vtkhdf_open.restype = ctypes.c_int
        L.flou_vtkhdf_open.argtypes = [ctypes.c_char_p, ctypes.c_int64, ctypes.c_void_p, ctypes.c_int,
                                       ctypes.c_int64, _i64p, ctypes.c_int64, ctypes.POINTER(ctypes.c_uint8),
                                       _i64p, _i64p, ctypes.POINTER(ctypes.c_void_p)]
        L.flou_vtkhdf_write.restype = ctypes.c_int
        L.flou_vtkhdf_write.argtypes = [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_int, ctypes.c_int, _i64p,
                                        ctypes.c_void_p]
        L.flou_vtkhdf_close.restype = ctypes.c_int
        L.flou_vtkhdf_close.argtypes = [ctypes.c_void_p]
        L.flou_vtkhdf_last_error.restype = ctypes.c_char_p
        L.flou_vtkhdf_last_error.argtypes = []
        L.flou_vtkhdf_hdf5_version.restype = ctypes.c_char_p
        L.flou_vtkhdf_hdf5_version.argtypes = []
        _vt = L
    return _vt


def hdf5_version():
    return _lib().flou_vtkhdf_hdf5_version().decode()


def _check(rc, what):
    if rc != OK:
        msg = _lib().flou_vtkhdf_last_error().decode(errors="replace")
        exc = ValueError if rc == EINVAL else OSError
        raise exc(f"{what}: {msg}")


def _elem_type(a):
    """Element type of a dataset: float64/float32 as given, uint8 and bool as uint8, any
    other integer as int64."""
    if a.dtype == np.float64:
        return a, F64
    if a.dtype == np.float32:
        return a, F32
    if a.dtype == np.uint8 or a.dtype == np.bool_:
        return a.astype(np.uint8, copy=False), U8
    if np.issubdtype(a.dtype, np.integer):
        return a.astype(np.int64, copy=False), I64
    raise TypeError(f"cannot write {a.dtype} data to a VTKHDF file")


class FlouFile:
    """FlouFile (FlouBiz.jl:25-28): the file name and the open writer handle."""

    def __init__(self, name, handler):
        self.name, self.handler = name, handler

    def _write(self, path, data):
        if self.handler is None:
            raise ValueError(f"{self.name} is closed")
        a = np.asarray(data)
        a, code = _elem_type(np.ascontiguousarray(a) if a.ndim else a.copy())
        dims = np.asarray(a.shape, dtype=np.int64)
        _check(_lib().flou_vtkhdf_write(self.handler, path.encode(), code, a.ndim,
                                        dims.ctypes.data_as(_i64p), a.ctypes.data_as(ctypes.c_void_p)),
               f"writing {path} to {self.name}")

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        if self.handler is not None:
            close_file_(self)


def mesh_arrays(mesh, std, dtype=np.float64):
    """The mesh datasets of open_for_write! (IO.jl:35-74) as host arrays: points
    (ne*m^nd, 3) of `dtype` (unused coordinates 0), connectivity, offsets (ne+1), types and
    regions (region(::CartesianMesh, i) = 1, CartesianMesh.jl:32)."""
    nd, ne = mesh.nd, mesh.nelements
    X = mesh.equispaced_coords(std)
    pts = np.zeros((X.shape[0], 3), dtype=dtype)
    pts[:, :nd] = X
    offsets = np.arange(ne + 1, dtype=np.int64) * std.nequispaced()
    conn = np.ascontiguousarray((std.vtk_connectivities()[None, :] + offsets[:-1, None]).reshape(-1))
    types = np.full(ne, std.vtk_type(), dtype=np.uint8)
    regions = np.ones(ne, dtype=np.int64)
    return pts, conn, offsets, types, regions


def open_mesh(filename, points, connectivity, offsets, types, regions):
    """Create (or rewrite) `filename` with the /VTKHDF group, its attributes and the given
    mesh datasets (flou_vtkhdf_open); returns the FlouFile."""
    pts = np.ascontiguousarray(points)
    if pts.dtype not in (np.float64, np.float32) or pts.ndim != 2 or pts.shape[1] != 3:
        raise ValueError("points must be an (npoints, 3) float64 or float32 array")
    conn = np.ascontiguousarray(connectivity, dtype=np.int64)
    offs = np.ascontiguousarray(offsets, dtype=np.int64)
    typ = np.ascontiguousarray(types, dtype=np.uint8)
    reg = np.ascontiguousarray(regions, dtype=np.int64)
    if offs.size != typ.size + 1 or reg.size != typ.size:
        raise ValueError("offsets must have one entry more than types and regions")
    h = ctypes.c_void_p()
    _check(_lib().flou_vtkhdf_open(os.fsencode(filename), pts.shape[0], pts.ctypes.data_as(ctypes.c_void_p),
                                   F64 if pts.dtype == np.float64 else F32, conn.size, conn.ctypes.data_as(_i64p),
                                   typ.size, typ.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)),
                                   offs.ctypes.data_as(_i64p), reg.ctypes.data_as(_i64p), ctypes.byref(h)),
           f"open_for_write({filename})")
    return FlouFile(filename, h)


def open_for_write(filename, disc):
    """open_for_write(filename, disc) (FlouBiz.jl:38-42, IO.jl:16-76): create (or rewrite) a
    VTKHDF file holding the mesh of `disc` as one VTK Lagrange cell per element: the
    equispaced points of every element (x, y, z) in the discretisation's precision, the
    cells' connectivity, offsets and types, and their regions."""
    return open_mesh(filename, *mesh_arrays(disc.mesh, disc.std, disc.dtype))


def add_fielddata_(file, data, name):
    """add_fielddata!(file, data, name) (FlouBiz.jl:49-52): /VTKHDF/FieldData/<name>."""
    file._write("FieldData/" + name, data)


def add_celldata_(file, data, name):
    """add_celldata!(file, data, name) (FlouBiz.jl:60-63): /VTKHDF/CellData/<name>, one value
    per element."""
    file._write("CellData/" + name, data)


def _columns(Q, disc):
    """Device (ncols, ndofs) view of a state: a device tensor (nvars, ndofs), its transpose,
    or a (ND, nvars, ndofs) block; or a host array (ndofs, nvars) / (ndofs, nvars, ND) /
    (ndofs,) uploaded first.  Columns keep Flou's column-major order (variable fastest)."""
    nd = disc.ndofs
    if isinstance(Q, torch.Tensor):
        if not Q.is_cuda:
            Q = Q.detach().cpu().numpy()
        else:
            if Q.dtype != disc.torch_dtype():
                raise TypeError(f"point data must be {disc.torch_dtype()}")
            if Q.numel() % nd:
                raise ValueError("point data must hold whole columns of ndofs values")
            if Q.dim() == 2 and Q.shape[0] == nd and Q.shape[1] != nd and Q.stride() == (1, nd):
                return Q.t()
            if Q.shape[-1] == nd and Q.is_contiguous():
                return Q.reshape(-1, nd)
            raise ValueError("device point data must be contiguous (..., ndofs) or a transposed (ndofs, ncols)")
    A = np.asarray(Q)
    if A.shape[0] != nd:
        raise ValueError(f"point data must have {nd} rows (one per solution node)")
    A = A.reshape(nd, -1, order="F")
    return torch.from_numpy(np.ascontiguousarray(A.T, dtype=disc.dtype)).to("cuda")


def pointdata2VTKHDF(Q, disc):
    """pointdata2VTKHDF(Q, disc) (IO.jl:78-121): the values of every column of Q at the
    equispaced points of every element, one host vector per column (element-major, points
    x-fastest).  The projection runs on the GPU; the result crosses to the host once."""
    disc._sync_stream()
    cols = _columns(Q, disc)
    ncols = cols.shape[0]
    std = disc.std
    A = np.ascontiguousarray(std.node2eq1, dtype=np.float64)
    m = A.shape[0]
    out = torch.empty((ncols, disc.nelements * m ** disc.nd), dtype=disc.torch_dtype(), device="cuda")
    C.check(C.lib.flou_hip_equispaced(disc.h, m, A.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                      ctypes.c_void_p(cols.data_ptr()), ncols, ctypes.c_void_p(out.data_ptr())),
            disc.h, "pointdata2VTKHDF")
    host = out.cpu().numpy()
    return [host[c] for c in range(ncols)]


def add_pointdata_(file, data, disc, name):
    """add_pointdata!(file, data, disc, name) (FlouBiz.jl:72-81): the first column of
    pointdata2VTKHDF(data, disc) at /VTKHDF/PointData/<name>."""
    file._write("PointData/" + name, pointdata2VTKHDF(data, disc)[0])


def add_solution_(file, sol, disc, equation):
    """add_solution!(file, sol, disc, equation) (FlouBiz.jl:91-103): every variable of the
    state at /VTKHDF/PointData/<variable name> (variablenames(equation))."""
    names = equation.variablenames()
    for iv, data in enumerate(pointdata2VTKHDF(sol, disc)):
        file._write("PointData/" + names[iv], data)


def close_file_(file):
    """close_file!(file) (FlouBiz.jl:110-113)."""
    h, file.handler = file.handler, None
    if h is not None:
        _check(_lib().flou_vtkhdf_close(h), f"closing {file.name}")


class SaveCallback:
    """The save callback of get_save_callback(basename; iter=true) (FlouTime.jl:59-90):
    after every selected step (and once at the start) it writes
    "<basename>_<iteration, 10 digits>.hdf" with the field Time = [t] and the solution.
    `files` lists what it wrote."""

    def __init__(self, basename, iter=True):
        self.basename, self.iter = basename, iter
        self.files = []

    def selected(self, it):
        return self.iter is True or it in self.iter

    def _save(self, Q, disc, equation, t, it):
        filename = "%s_%010d.hdf" % (self.basename, it)
        f = open_for_write(filename, disc)
        try:
            add_fielddata_(f, np.array([t], dtype=np.float64), "Time")
            add_solution_(f, Q, disc, equation)
        finally:
            close_file_(f)
        self.files.append(filename)
        _log.info("Saved solution at t=%.7g in `%s`", t, filename)

    def initialize(self, Q, disc, equation, t, it):
        self._save(Q, disc, equation, t, it)

    def __call__(self, Q, disc, equation, t, it):
        if self.selected(it):
            self._save(Q, disc, equation, t, it)


def get_save_callback(basename, iter=True):
    """get_save_callback(basename; iter=true) (FlouTime.jl:82-90); pass the result to
    timeintegrate(..., callback=...)."""
    return SaveCallback(basename, iter)
