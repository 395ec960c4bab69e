"""VTKHDF output (src/FlouBiz/FlouBiz.jl:16-115, src/FlouSpatial/IO.jl:16-126) on the CPU:
the oracle's equispaced projection and VTK listing, the host-side mesh arrays and the file
writer (libflouvtkhdf.so).

No reference fixture holds a VTKHDF file, so the pieces are pinned by what fixes them:
* interp_matrix reproduces every polynomial of degree < n at the equispaced points (closed
  form), and its GLL rows at ±1 are unit vectors;
* the oracle's dense node2eq product equals an independent sum-factorised product and is
  exact on tensor-product polynomials;
* the cell listing follows VTK's Lagrange cell convention with the reference's edge list:
  corners at the reference cell's corners in VTK order, each edge's points on the segment
  between its corners in increasing distance from the first, face points inside their
  face with the first free axis fastest, interior points x-fastest;
* the files are read back with h5dump, the HDF5 project's own tool (shares no code with the
  writer): groups, attributes, element types, shapes and bit-exact values.
"""
import os

import numpy as np
import pytest

import h5read
import oracle_lib as O

import flouhip as F
from flouhip import io as FIO
from flouhip.stdregions import _equispaced_1d

STD = {1: F.StdSegment, 2: F.StdQuad, 3: F.StdHex}


def std_of(nd, n, nodes="GLL", m=None):
    b = F.LagrangeBasis(nodes, n)
    return STD[nd](b, F.DGSEMrec(b), nd + 2, nequispaced=m)


# ---------------------------------------------------------------- equispaced projection
@pytest.mark.parametrize("nodes", ["GLL", "GL"])
@pytest.mark.parametrize("n", [2, 3, 5, 7, 9])
def test_interp_matrix_is_exact_on_polynomials(nodes, n):
    xi = O.build_ops(n, nodes)["xi"]
    for m in (n, n + 3, 2):
        xe = _equispaced_1d(m)
        A = O.interp_matrix(n, xe, nodes)
        for p in range(n):
            assert np.max(np.abs(A @ xi ** p - xe ** p)) <= 2e-14
        if nodes == "GLL":
            assert np.array_equal(A[0], np.eye(n)[0]) and np.array_equal(A[-1], np.eye(n)[-1])
        # the package's own matrix is the same one
        assert np.max(np.abs(std_of(1, n, nodes, m).node2eq1 - A)) <= 1e-15


def test_equispaced_points():
    for m in range(2, 12):
        x = _equispaced_1d(m)
        assert x[0] == -1.0 and x[-1] == 1.0 and np.array_equal(x, -x[::-1])
        assert np.max(np.abs(x - np.linspace(-1, 1, m))) <= 2.3e-16
    assert np.array_equal(_equispaced_1d(5), [-1.0, -0.5, 0.0, 0.5, 1.0])
    assert np.array_equal(_equispaced_1d(1), [0.0])
    s = std_of(3, 4)
    assert s.nequispaced() == 64 and s.ξe.shape == (64, 3)
    assert np.array_equal(s.ξe[1], [s.ξe1[1], -1.0, -1.0]) and np.array_equal(s.ξe[4], [-1.0, s.ξe1[1], -1.0])


def _sumfact(nd, n, A, ne, Q):
    """Qe = (A ⊗ .. ⊗ A) Q_e by one contraction per direction (numpy einsum)."""
    m = A.shape[0]
    out = []
    for c in range(Q.shape[1]):
        q = Q[:, c].reshape((ne,) + (n,) * nd)  # q[e, (k, j,) i], i fastest
        if nd == 1:
            r = np.einsum("ai,ei->ea", A, q)
        elif nd == 2:
            r = np.einsum("ai,bj,eji->eba", A, A, q)
        else:
            r = np.einsum("ai,bj,ck,ekji->ecba", A, A, A, q)
        out.append(r.reshape(ne * m ** nd))
    return np.stack(out, 1)


@pytest.mark.parametrize("nd", [1, 2, 3])
@pytest.mark.parametrize("nodes", ["GLL", "GL"])
def test_oracle_projection(nd, nodes):
    rng = np.random.default_rng(nd)
    n, ne = 4, 3
    for m in (n, n + 2):
        A = O.interp_matrix(n, _equispaced_1d(m), nodes)
        Q = rng.standard_normal((ne * n ** nd, 3))
        got = O.equispaced(nd, n, A, ne, Q)
        assert np.max(np.abs(got - _sumfact(nd, n, A, ne, Q))) <= 1e-14
        # tensor polynomial of degree n-1 per direction is reproduced exactly
        xi, xe = O.build_ops(n, nodes)["xi"], _equispaced_1d(m)

        def field(x1, nd_, shape):
            g = np.meshgrid(*([x1] * nd_), indexing="ij")
            v = 1.0
            for k, gk in enumerate(g):
                v = v * (gk ** (n - 1) + (k + 1) * gk - 0.5)
            return np.tile(v.ravel(order="F"), shape)

        Qp = field(xi, nd, ne)[:, None]
        assert np.max(np.abs(O.equispaced(nd, n, A, ne, Qp)[:, 0] - field(xe, nd, ne))) <= 1e-13


# ---------------------------------------------------------------- VTK cell listing
def _vtk_expected(nd, n):
    """VTK Lagrange cell point order as geometry: corners, edges between corners, faces,
    interior; lattice coordinates 0..n-1 -> x-fastest linear index."""
    def idx(p):
        return int(sum(int(round(c)) * n ** k for k, c in enumerate(p)))

    t = np.arange(1, n - 1) / (n - 1)
    hi = n - 1
    if nd == 1:
        corners = [(0,), (hi,)]
        edges = [(0, 1)]
    elif nd == 2:
        corners = [(0, 0), (hi, 0), (hi, hi), (0, hi)]
        edges = [(0, 1), (1, 2), (3, 2), (0, 3)]
    else:
        corners = [(0, 0, 0), (hi, 0, 0), (hi, hi, 0), (0, hi, 0), (0, 0, hi), (hi, 0, hi), (hi, hi, hi), (0, hi, hi)]
        edges = [(0, 1), (1, 2), (3, 2), (0, 3), (4, 5), (5, 6), (7, 6), (4, 7), (0, 4), (1, 5), (2, 6), (3, 7)]
    out = [idx(c) for c in corners]
    if nd == 1:
        out += [idx((hi * s,)) for s in t]
        return out
    C = np.array(corners, dtype=float)
    for a, b in edges:
        out += [idx(C[a] + s * (C[b] - C[a])) for s in t]
    inner = range(1, n - 1)
    if nd == 2:
        out += [idx((i, j)) for j in inner for i in inner]
        return out
    for axis in range(3):  # faces x=0, x=hi, y=0, y=hi, z=0, z=hi
        free = [k for k in range(3) if k != axis]
        for val in (0, hi):
            for s in inner:
                for f in inner:  # first free axis fastest
                    p = [0, 0, 0]
                    p[axis], p[free[0]], p[free[1]] = val, f, s
                    out.append(idx(p))
    out += [idx((i, j, k)) for k in inner for j in inner for i in inner]
    return out


@pytest.mark.parametrize("nd", [1, 2, 3])
@pytest.mark.parametrize("n", [2, 3, 4, 6, 8])
def test_vtk_listing(nd, n):
    ref = O.vtk_connectivities(nd, n)
    assert sorted(ref) == list(range(n ** nd))
    assert ref.tolist() == _vtk_expected(nd, n)
    assert np.array_equal(std_of(nd, n).vtk_connectivities(), ref)
    assert std_of(nd, n).vtk_type() == {1: 68, 2: 70, 3: 72}[nd]


def test_mesh_arrays_cartesian():
    for nd, nel in ((1, [5]), (2, [3, 2]), (3, [2, 3, 2])):
        n = 4
        std = std_of(nd, n)
        start, finish = [0.0, -1.0, 2.0][:nd], [1.0, 1.0, 5.0][:nd]
        mesh = F.CartesianMesh(nd, start, finish, nel)
        pts, conn, offs, types, reg = FIO.mesh_arrays(mesh, std)
        ne = int(np.prod(nel))
        assert pts.shape == (ne * n ** nd, 3) and np.all(pts[:, nd:] == 0)
        assert np.array_equal(offs, np.arange(ne + 1) * n ** nd)
        assert np.array_equal(types, np.full(ne, 66 + 2 * nd)) and np.array_equal(reg, np.ones(ne))
        # element e, point p: start + (i_e + (ξe_i + 1)/2) Δ per direction
        xe = _equispaced_1d(n)
        for e in (0, ne - 1):
            ei = [e % nel[0], (e // nel[0]) % nel[1] if nd > 1 else 0, e // (nel[0] * nel[1]) if nd > 2 else 0]
            for p in range(n ** nd):
                ii = [p % n, (p // n) % n, p // (n * n)]
                for k in range(nd):
                    dx = (finish[k] - start[k]) / nel[k]
                    assert abs(pts[e * n ** nd + p, k] - (start[k] + (ei[k] + (xe[ii[k]] + 1) / 2) * dx)) <= 1e-14
        # a cell's connectivity names its own points only
        assert np.array_equal(conn[:n ** nd], O.vtk_connectivities(nd, n))
        assert np.array_equal(conn[-n ** nd:], O.vtk_connectivities(nd, n) + (ne - 1) * n ** nd)


# ---------------------------------------------------------------- the file
needs_h5 = pytest.mark.skipif(not (h5read.available() and os.path.exists(FIO.VTKHDF_LIB_PATH)),
                              reason="h5dump or libflouvtkhdf.so not present")

MESH_DATASETS = {"NumberOfPoints", "Points", "NumberOfConnectivityIds", "Connectivity", "NumberOfCells", "Types",
                 "Offsets", "CellData/Region"}


@needs_h5
@pytest.mark.parametrize("nd", [1, 2, 3])
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_file_layout_and_values(tmp_path, nd, dtype):
    n = 3
    std = std_of(nd, n)
    mesh = F.CartesianMesh(nd, [0.0] * nd, [1.0] * nd, [2] * nd)
    arrays = FIO.mesh_arrays(mesh, std, dtype)
    pts, conn, offs, types, reg = arrays
    path = str(tmp_path / "sol_0000000003.hdf")
    f = FIO.open_mesh(path, *arrays)
    rho = np.random.default_rng(1).standard_normal(pts.shape[0]).astype(dtype)
    FIO.add_fielddata_(f, np.array([0.25]), "Time")
    FIO.add_celldata_(f, np.arange(mesh.nelements), "Id")
    f._write("PointData/rho", rho)
    f._write("FieldData/Scalar", np.float64(3.5))
    FIO.close_file_(f)

    objs = h5read.contents(path)
    want = {"/", "/VTKHDF", "/VTKHDF/CellData", "/VTKHDF/FieldData", "/VTKHDF/PointData"}
    want |= {"/VTKHDF/" + d for d in MESH_DATASETS | {"FieldData/Time", "FieldData/Scalar", "CellData/Id",
                                                      "PointData/rho"}}
    want |= {"/VTKHDF/Version", "/VTKHDF/Type"}
    assert set(objs) == want
    assert objs["/VTKHDF/Version"] == objs["/VTKHDF/Type"] == "attribute"

    ty, dims, data = h5read.attribute(path, "/VTKHDF/Version")
    assert (ty, dims) == ("H5T_STD_I64LE", (2,)) and data.replace(" ", "") == "1,0"
    ty, dims, data = h5read.attribute(path, "/VTKHDF/Type")
    assert ty == ("string", 16, "H5T_STR_NULLTERM", "H5T_CSET_ASCII") and dims is None
    assert data == '"UnstructuredGrid"'

    real = "H5T_IEEE_F64LE" if dtype == np.float64 else "H5T_IEEE_F32LE"
    ne, npts = mesh.nelements, pts.shape[0]
    expect = {"NumberOfPoints": ("H5T_STD_I64LE", (1,), [npts]),
              "Points": (real, (npts, 3), pts),
              "NumberOfConnectivityIds": ("H5T_STD_I64LE", (1,), [conn.size]),
              "Connectivity": ("H5T_STD_I64LE", (conn.size,), conn),
              "NumberOfCells": ("H5T_STD_I64LE", (1,), [ne]),
              "Types": ("H5T_STD_U8LE", (ne,), types),
              "Offsets": ("H5T_STD_I64LE", (ne + 1,), offs),
              "CellData/Region": ("H5T_STD_I64LE", (ne,), reg),
              "CellData/Id": ("H5T_STD_I64LE", (ne,), np.arange(ne)),
              "FieldData/Time": ("H5T_IEEE_F64LE", (1,), [0.25]),
              "FieldData/Scalar": ("H5T_IEEE_F64LE", None, 3.5),
              "PointData/rho": (real, (npts,), rho)}
    for name, (ty, dims, vals) in expect.items():
        assert h5read.dataset_header(path, "/VTKHDF/" + name) == (ty, dims), name
        got = h5read.dataset(path, "/VTKHDF/" + name)
        assert np.array_equal(got, np.asarray(vals)), name


@needs_h5
def test_file_errors_and_rewrite(tmp_path):
    std = std_of(2, 3)
    mesh = F.CartesianMesh(2, [0.0, 0.0], [1.0, 1.0], [2, 2])
    arrays = FIO.mesh_arrays(mesh, std)
    path = str(tmp_path / "a.hdf")
    f = FIO.open_mesh(path, *arrays)
    FIO.add_fielddata_(f, [1.0], "Time")
    with pytest.raises(ValueError, match="already exists"):
        FIO.add_fielddata_(f, [2.0], "Time")
    with pytest.raises(TypeError):
        FIO.add_fielddata_(f, np.array([1 + 2j]), "Complex")
    FIO.close_file_(f)
    with pytest.raises(ValueError, match="closed"):
        FIO.add_fielddata_(f, [1.0], "Late")
    # "w" truncates: a rewrite holds only what the new handle wrote
    FIO.close_file_(FIO.open_mesh(path, *arrays))
    assert "/VTKHDF/FieldData/Time" not in h5read.contents(path)
    pts, conn, offs, types, reg = arrays
    bad = offs.copy()
    bad[-1] += 1
    with pytest.raises(ValueError, match="offsets"):
        FIO.open_mesh(str(tmp_path / "b.hdf"), pts, conn, bad, types, reg)
    badc = conn.copy()
    badc[5] = pts.shape[0]
    with pytest.raises(ValueError, match="connectivity"):
        FIO.open_mesh(str(tmp_path / "c.hdf"), pts, badc, offs, types, reg)
    with pytest.raises(OSError):
        FIO.open_mesh(str(tmp_path / "missing_dir" / "d.hdf"), *arrays)
