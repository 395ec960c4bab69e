"""CPU tests of the drop-in boundary and of the host-side mirror (no GPU needed).

* libflouhip.so loads and exports every function include/flouhip.h declares, and
  libflouvtkhdf.so every function of include/flouvtkhdf.h;
* flou_hip_create rejects invalid descriptions before touching a device;
* the host mirror's operators, mesh connectivity, geometry and node coordinates equal the
  oracle's restatement of Flou's bit for bit (index work is exact).
"""
import ctypes
import os
import re

import numpy as np
import pytest

import flouhip as F
import oracle_lib as O
from flouhip import _capi as C

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "flouhip.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(flou_hip_\w+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(C.LIB_PATH)
    names = declared_functions()
    assert len(names) >= 20
    for name in names:
        assert hasattr(lib, name), name
    # and the ctypes table covers them all
    assert set(names) <= set(C.SYMBOLS), set(names) - set(C.SYMBOLS)


def test_vtkhdf_library_exports_every_declared_symbol():
    from flouhip import io as FIO
    if not os.path.exists(FIO.VTKHDF_LIB_PATH):
        pytest.skip("libflouvtkhdf.so not built (no HDF5)")
    text = re.sub(r"/\*.*?\*/", "", open(os.path.join(ROOT, "include", "flouvtkhdf.h")).read(), flags=re.S)
    names = sorted(set(re.findall(r"\b(flou_vtkhdf_\w+)\s*\(", text)))
    assert len(names) == 5
    lib = FIO._lib()
    for name in names:
        assert hasattr(lib, name), name
    assert FIO.hdf5_version().startswith("1.")


def test_version_string():
    assert "gfx950" in F.version()


def _valid_desc(keep):
    b = F.LagrangeBasis("GLL", 3)
    std = F.StdQuad(b, F.DGSEMrec(b), 4)
    m = F.CartesianMesh(2, (0, 0), (1, 1), (2, 2)).apply_periodicBCs(("1", "2"), ("3", "4"))
    d = C.Desc()
    d.nd, d.n, d.nvars, d.precision = 2, 3, 4, C.FP64
    d.nelements, d.nfaces, d.nboundaries = m.nelements, m.nfaces, 0
    d.equation, d.gamma = C.EULER, 1.4
    d.volume_op, d.tp_flux, d.surface_avg, d.surface_diss, d.intensity = C.SPLIT, 1, 0, 1, 1.0

    def p(x, t, ct):
        a = np.ascontiguousarray(x, dtype=t)
        keep.append(a)
        return a.ctypes.data_as(ctypes.POINTER(ct))

    for name, arr in [("D", std.D), ("Ds", std.Ds), ("Dsharp", std.Dsharp), ("l_minus", std.l[0]),
                      ("l_plus", std.l[1]), ("dg_minus", std.dg[0]), ("dg_plus", std.dg[1]),
                      ("omega", std.ω1)]:
        setattr(d, name, p(arr, np.float64, ctypes.c_double))
    d.elem_faceinds = p(m.elem_faces, np.int64, ctypes.c_int64)
    d.elem_facepos = p(m.elem_side, np.int8, ctypes.c_int8)
    d.face_eleminds = p(m.face_elems, np.int64, ctypes.c_int64)
    d.face_elempos = p(m.face_pos, np.int8, ctypes.c_int8)
    d.face_boundary = p(m.face_boundary, np.int32, ctypes.c_int32)
    ej, em, ev = m.element_geometry(std)
    fr, fj = m.face_geometry()
    d.elem_jac = p(ej, np.float64, ctypes.c_double)
    d.elem_metric = p(em, np.float64, ctypes.c_double)
    d.elem_volume = p(ev, np.float64, ctypes.c_double)
    d.face_frames = p(fr, np.float64, ctypes.c_double)
    d.face_jac = p(fj, np.float64, ctypes.c_double)
    return d


@pytest.mark.parametrize("field,value,code", [
    ("nd", 4, C.EINVAL),
    ("n", 9, C.EUNSUPPORTED),
    ("nvars", 3, C.EINVAL),
    ("precision", 7, C.EINVAL),
    ("volume_op", 5, C.EINVAL),
    ("tp_flux", 9, C.EINVAL),
    ("surface_diss", 4, C.EINVAL),
    ("geometry", 7, C.EINVAL),
    ("gamma", 0.5, C.EINVAL),
])
def test_create_rejects_invalid_descriptions(field, value, code):
    keep = []
    d = _valid_desc(keep)
    setattr(d, field, value)
    h = ctypes.c_void_p()
    rc = C.lib.flou_hip_create(ctypes.byref(d), ctypes.byref(h))
    assert rc == code
    assert not h.value
    assert C.lib.flou_hip_create_error()


@pytest.mark.parametrize("setup,code,text", [
    (lambda d, keep: setattr(d, "blend", -1.0), C.EINVAL, b"blending"),
    (lambda d, keep: setattr(d, "blend", float("inf")), C.EINVAL, b"blending"),
    (lambda d, keep: setattr(d, "geometry", C.GEOM_CURVED), C.EUNSUPPORTED, b"Cartesian"),
    (lambda d, keep: keep[-5].__setitem__(0, 2.0 * keep[-5][0]), C.EUNSUPPORTED, b"share one geometry"),
], ids=["negative-blend", "inf-blend", "curved", "non-uniform"])
def test_create_rejects_unsupported_hybrid(setup, code, text):
    keep = []
    d = _valid_desc(keep)
    d.volume_op, d.blend = C.HYBRID, 1.0
    setup(d, keep)
    h = ctypes.c_void_p()
    assert C.lib.flou_hip_create(ctypes.byref(d), ctypes.byref(h)) == code
    assert text in C.lib.flou_hip_create_error()


def test_create_rejects_gauss_split_advection():
    keep = []
    d = _valid_desc(keep)
    d.equation, d.nvars, d.tp_flux, d.surface_avg, d.surface_diss = C.ADVECTION, 1, 0, 0, 1
    keep[3][:] = [0.5, 0.5, 0.0]  # traces that are not node values: Gauss-type nodes
    h = ctypes.c_void_p()
    assert C.lib.flou_hip_create(ctypes.byref(d), ctypes.byref(h)) == C.EUNSUPPORTED
    assert b"Gauss" in C.lib.flou_hip_create_error()


def test_create_rejects_inconsistent_connectivity():
    keep = []
    d = _valid_desc(keep)
    fi = keep[8]  # elem_faceinds
    fi[0, 0], fi[0, 1] = fi[0, 1], fi[0, 0]
    h = ctypes.c_void_p()
    assert C.lib.flou_hip_create(ctypes.byref(d), ctypes.byref(h)) == C.EINVAL
    assert b"disagree" in C.lib.flou_hip_create_error()


@pytest.mark.parametrize("value,text", [(2, b"out of range"), (7, b"out of range")])
def test_create_rejects_bad_face_orientations(value, text):
    """Segment faces (2-D) take orientations 0 and 1 (StdSegment.jl:161-167)."""
    keep = []
    d = _valid_desc(keep)
    fo = np.zeros(d.nfaces, dtype=np.uint8)
    fo[0] = value
    keep.append(fo)
    d.face_orientation = fo.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8))
    h = ctypes.c_void_p()
    assert C.lib.flou_hip_create(ctypes.byref(d), ctypes.byref(h)) == C.EINVAL
    assert text in C.lib.flou_hip_create_error()


def test_null_arguments():
    assert C.lib.flou_hip_rhs(None, None, None, 0.0) == C.EINVAL
    assert C.lib.flou_hip_destroy(None) == C.EINVAL


# ------------------------------------------------------------------ host mirror vs oracle
@pytest.mark.parametrize("nodes", ["GLL", "GL"])
@pytest.mark.parametrize("n", [2, 3, 5, 8])
def test_operators_match_oracle(nodes, n):
    b = F.LagrangeBasis(nodes, n)
    std = F.StdSegment(b, F.DGSEMrec(b), 1)
    o = O.build_ops(n, nodes)
    for mine, ref in [(b.ξ, "xi"), (b.ω, "w"), (std.D, "D"), (std.Ds, "Ds"), (std.Dsharp, "Dsh"),
                      (std.l[0], "lm"), (std.l[1], "lp"), (std.dg[0], "gm"), (std.dg[1], "gp")]:
        assert np.array_equal(mine, o[ref]), ref


MESHES = [
    (1, [5], [0]), (1, [5], [1]), (2, [3, 4], [0, 0]), (2, [3, 4], [1, 1]), (2, [3, 4], [0, 1]),
    (2, [3, 4], [1, 0]), (3, [3, 4, 2], [0, 0, 0]), (3, [3, 4, 2], [1, 1, 1]),
    (3, [2, 3, 4], [1, 0, 1]), (3, [4, 4, 4], [0, 1, 0]),
]


@pytest.mark.parametrize("nd,nel,per", MESHES)
def test_mesh_connectivity_geometry_match_oracle(nd, nel, per):
    n = 3
    start, fin = [0.0] * nd, [1.0, 1.5, 2.0][:nd]
    b = F.LagrangeBasis("GLL", n)
    std = [None, F.StdSegment, F.StdQuad, F.StdHex][nd](b, F.DGSEMrec(b), nd + 2)
    m = F.CartesianMesh(nd, start, fin, nel)
    m.apply_periodicBCs(*[(str(2 * k + 1), str(2 * k + 2)) for k in range(nd) if per[k]])
    od = O.OracleDisc(nd, n, nel, start, fin, per, O.build_ops(n, "GLL"), O.EQ_EULER, [1.4], 1,
                      1, 0, 1, 1.0)
    c = od.connectivity()
    assert np.array_equal(c["e_face"], m.elem_faces)
    assert np.array_equal(c["e_side"], m.elem_side)
    assert np.array_equal(c["f_elem"], m.face_elems)
    assert np.array_equal(c["f_pos"], m.face_pos)
    assert np.array_equal(c["f_bd"], m.face_boundary)
    assert od.nbd == m.nboundaries()
    g = od.geometry()
    ej, em, ev = m.element_geometry(std)
    fr, fj = m.face_geometry()
    assert np.array_equal(g["e_jac"], ej)
    assert np.array_equal(g["e_metric"].transpose(0, 2, 1).reshape(len(ej), -1), em)
    assert np.array_equal(g["e_vol"], ev)
    assert np.array_equal(g["f_frame"], fr)
    assert np.array_equal(g["f_jac"], fj)
    X, FX = od.coords()
    assert np.array_equal(X, m.node_coords(std))


def test_reference_cartesian_numbering_3d():
    """Spot values of CartesianMesh.jl:206-233 (element 2 of a 2x2x2 box, 1-based there)."""
    m = F.CartesianMesh(3, (0, 0, 0), (1, 1, 1), (2, 2, 2))
    nx = ny = nz = 2
    npx, npy = 3, 3
    i, j, k = 2, 1, 1  # 1-based element (i, j, k) = (2, 1, 1) is element 2
    ref = [(k - 1) * npx * ny + (j - 1) * npx + i, (k - 1) * npx * ny + (j - 1) * npx + i + 1,
           npx * ny * nz + (k - 1) * nx * npy + (j - 1) * nx + i,
           npx * ny * nz + (k - 1) * nx * npy + (j - 1) * nx + i + nx,
           npx * ny * nz + nx * npy * nz + (k - 1) * nx * ny + (j - 1) * nx + i,
           npx * ny * nz + nx * npy * nz + (k - 1) * nx * ny + (j - 1) * nx + i + nx * ny]
    assert list(m.elem_faces[1] + 1) == ref
    assert list(m.elem_side[1] + 1) == [2, 1, 1, 1, 1, 1]


def test_periodic_rewiring_pairs_opposite_boundaries():
    m = F.CartesianMesh(2, (0, 0), (1, 1), (3, 2)).apply_periodicBCs(("1", "2"))
    # every x-min face now has the x-max element of its row as slave, position 2 (0-based 1)
    for f in np.nonzero(m.face_pos[:, 0] == 0)[0]:
        e1, e2 = m.face_elems[f]
        assert e1 % 3 == 0 and e2 == e1 + 2 and m.face_pos[f, 1] == 1
    assert m.nboundaries() == 2 and m.bdalive == [2, 3]
    with pytest.raises(ValueError):
        F.CartesianMesh(2, (0, 0), (1, 1), (3, 2)).apply_periodicBCs(("2", "3"))


def test_missing_library_fails_loudly(tmp_path, monkeypatch):
    import importlib
    monkeypatch.setattr(C, "LIB_PATH", str(tmp_path / "nope.so"))
    with pytest.raises(ImportError):
        C._load()


def test_source_term_is_refused_explicitly():
    """The reference's apply_sourceterm! (MultielementDiscontinuous.jl:139-146) runs a host
    function per node; the GPU path refuses one instead of dropping it."""
    b = F.LagrangeBasis("GLL", 3)
    std = F.StdQuad(b, F.DGSEMrec(b), 4)
    m = F.CartesianMesh(2, (0, 0), (1, 1), (2, 2)).apply_periodicBCs(("1", "2"), ("3", "4"))
    op = F.SplitDivOperator(F.ChandrasekharAverage(), F.LxF(F.StdAverage(), 1.0))

    def source(dQ, Q, x, t):
        dQ[0] += 1.0

    with pytest.raises(NotImplementedError, match="source"):
        F.MultielementDisc(m, std, F.EulerEquation(2, 1.4), op, {}, source)

