"""Cartesian meshes, periodic rewiring and Cartesian geometry (host side, setup only).

Mirrors CartesianMesh{ND,RT}(start, finish, nxyz) and apply_periodicBCs!
(src/FlouCommon/CartesianMesh.jl:38-150, 152-500; src/FlouCommon/Mesh.jl:236-322) and the
Cartesian PhysicalElement / PhysicalFace constructors (src/FlouSpatial/PhysicalRegions.jl:
370-408, 541-696), vectorised with numpy so that 64^3-element meshes build in seconds.
Numbering is the reference's, shifted to 0-based.
"""
import numpy as np

LD = np.longdouble


class CartesianMesh:
    """CartesianMesh(ND, start, finish, nxyz)."""

    def __init__(self, nd, start, finish, nxyz):
        if nd not in (1, 2, 3):
            raise ValueError("The mesh can only have 1, 2 or 3 dimensions.")
        start = np.atleast_1d(np.asarray(start, dtype=np.float64))
        finish = np.atleast_1d(np.asarray(finish, dtype=np.float64))
        nxyz = tuple(int(x) for x in np.atleast_1d(nxyz))
        if len(start) != nd:
            raise ValueError(f"The `start` point must have {nd} coordinates.")
        if len(finish) != nd:
            raise ValueError(f"The `finish` point must have {nd} coordinates.")
        if len(nxyz) != nd:
            raise ValueError("The number of elements in all directions must be given.")
        if not np.all(start < finish):
            raise ValueError("All components of `start` must be lower than those of `finish`.")
        self.nd = nd
        self.start, self.finish, self.nelements_dir = start, finish, nxyz
        self.Δx = tuple((finish[k] - start[k]) / nxyz[k] for k in range(nd))
        # vertex coordinates: range(start, finish, n+1), one rounding from extended precision
        self.xyz = [np.array([LD(start[k]) + (LD(finish[k]) - LD(start[k])) * i / nxyz[k]
                              for i in range(nxyz[k] + 1)], dtype=LD).astype(np.float64)
                    for k in range(nd)]
        self._build()
        self.bdnames = [str(i + 1) for i in range(2 * nd)]
        self.periodic = {}

    # ------------------------------------------------------------------ connectivity
    def _build(self):
        nd, n = self.nd, self.nelements_dir
        if nd == 1:
            nx, = n
            i = np.arange(nx)
            self.elem_faces = np.stack([i, i + 1], axis=1)
            self.elem_side = np.stack([np.where(i == 0, 0, 1), np.zeros(nx, int)], axis=1)
            nf = nx + 1
            f = np.arange(nf)
            fe = np.stack([f - 1, f], axis=1)
            fp = np.tile([1, 0], (nf, 1))
            fe[0] = [0, -1]; fp[0] = [0, -1]
            fe[-1] = [nx - 1, -1]; fp[-1] = [1, -1]
            self.face_elems, self.face_pos = fe, fp
            self.bdfaces = [np.array([0]), np.array([nf - 1])]
            self.face_vertices = f[:, None]
        elif nd == 2:
            nx, ny = n
            npx, npy = nx + 1, ny + 1
            jj, ii = np.meshgrid(np.arange(ny), np.arange(nx), indexing="ij")
            ii, jj = ii.ravel(), jj.ravel()
            fx = lambda i, j: j * npx + i
            fy = lambda i, j: npx * ny + j * nx + i
            self.elem_faces = np.stack([fx(ii, jj), fx(ii + 1, jj), fy(ii, jj), fy(ii, jj + 1)], axis=1)
            self.elem_side = np.stack([np.where(ii == 0, 0, 1), np.zeros_like(ii),
                                       np.where(jj == 0, 0, 1), np.zeros_like(ii)], axis=1)
            el = lambda i, j: j * nx + i
            nfx, nfy = npx * ny, nx * npy
            fe = np.zeros((nfx + nfy, 2), dtype=np.int64)
            fp = np.zeros((nfx + nfy, 2), dtype=np.int64)
            J, I = np.meshgrid(np.arange(ny), np.arange(npx), indexing="ij")
            I, J = I.ravel(), J.ravel()
            fe[:nfx, 0] = np.where(I == 0, el(0, J), el(np.maximum(I - 1, 0), J))
            fe[:nfx, 1] = np.where((I == 0) | (I == nx), -1, el(np.minimum(I, nx - 1), J))
            fp[:nfx, 0] = np.where(I == 0, 0, 1)
            fp[:nfx, 1] = np.where((I == 0) | (I == nx), -1, 0)
            J, I = np.meshgrid(np.arange(npy), np.arange(nx), indexing="ij")
            I, J = I.ravel(), J.ravel()
            fe[nfx:, 0] = np.where(J == 0, el(I, 0), el(I, np.maximum(J - 1, 0)))
            fe[nfx:, 1] = np.where((J == 0) | (J == ny), -1, el(I, np.minimum(J, ny - 1)))
            fp[nfx:, 0] = np.where(J == 0, 2, 3)
            fp[nfx:, 1] = np.where((J == 0) | (J == ny), -1, 2)
            self.face_elems, self.face_pos = fe, fp
            allf = np.arange(nfx + nfy)
            Ix = allf[:nfx] % npx
            Jy = (allf[nfx:] - nfx) // nx
            self.bdfaces = [allf[:nfx][Ix == 0], allf[:nfx][Ix == nx],
                            allf[nfx:][Jy == 0], allf[nfx:][Jy == ny]]
            # face vertex lists (for face-node coordinates)
            v = lambda i, j: j * npx + i
            fv = np.zeros((nfx + nfy, 2), dtype=np.int64)
            J, I = np.meshgrid(np.arange(ny), np.arange(npx), indexing="ij")
            fv[:nfx] = np.stack([v(I.ravel(), J.ravel()), v(I.ravel(), J.ravel() + 1)], axis=1)
            J, I = np.meshgrid(np.arange(npy), np.arange(nx), indexing="ij")
            fv[nfx:] = np.stack([v(I.ravel(), J.ravel()), v(I.ravel() + 1, J.ravel())], axis=1)
            self.face_vertices = fv
        else:
            nx, ny, nz = n
            npx, npy, npz = nx + 1, ny + 1, nz + 1
            kk, jj, ii = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
            ii, jj, kk = ii.ravel(), jj.ravel(), kk.ravel()
            nfx, nfy, nfz = npx * ny * nz, nx * npy * nz, nx * ny * npz
            fx = lambda i, j, k: k * npx * ny + j * npx + i
            fy = lambda i, j, k: nfx + k * nx * npy + j * nx + i
            fz = lambda i, j, k: nfx + nfy + k * nx * ny + j * nx + i
            self.elem_faces = np.stack([fx(ii, jj, kk), fx(ii + 1, jj, kk), fy(ii, jj, kk),
                                        fy(ii, jj + 1, kk), fz(ii, jj, kk), fz(ii, jj, kk + 1)], axis=1)
            z = np.zeros_like(ii)
            self.elem_side = np.stack([np.where(ii == 0, 0, 1), z, np.where(jj == 0, 0, 1), z,
                                       np.where(kk == 0, 0, 1), z], axis=1)
            el = lambda i, j, k: k * nx * ny + j * nx + i
            nf = nfx + nfy + nfz
            fe = np.zeros((nf, 2), dtype=np.int64)
            fp = np.zeros((nf, 2), dtype=np.int64)
            bd = [None] * 6
            # X faces
            K, J, I = [a.ravel() for a in np.meshgrid(np.arange(nz), np.arange(ny), np.arange(npx), indexing="ij")]
            s = slice(0, nfx)
            fe[s, 0] = np.where(I == 0, el(0, J, K), el(np.maximum(I - 1, 0), J, K))
            fe[s, 1] = np.where((I == 0) | (I == nx), -1, el(np.minimum(I, nx - 1), J, K))
            fp[s, 0] = np.where(I == 0, 0, 1)
            fp[s, 1] = np.where((I == 0) | (I == nx), -1, 0)
            idx = np.arange(nfx)
            bd[0], bd[1] = idx[I == 0], idx[I == nx]
            # Y faces
            K, J, I = [a.ravel() for a in np.meshgrid(np.arange(nz), np.arange(npy), np.arange(nx), indexing="ij")]
            s = slice(nfx, nfx + nfy)
            fe[s, 0] = np.where(J == 0, el(I, 0, K), el(I, np.maximum(J - 1, 0), K))
            fe[s, 1] = np.where((J == 0) | (J == ny), -1, el(I, np.minimum(J, ny - 1), K))
            fp[s, 0] = np.where(J == 0, 2, 3)
            fp[s, 1] = np.where((J == 0) | (J == ny), -1, 2)
            idx = np.arange(nfx, nfx + nfy)
            bd[2], bd[3] = idx[J == 0], idx[J == ny]
            # Z faces
            K, J, I = [a.ravel() for a in np.meshgrid(np.arange(npz), np.arange(ny), np.arange(nx), indexing="ij")]
            s = slice(nfx + nfy, nf)
            fe[s, 0] = np.where(K == 0, el(I, J, 0), el(I, J, np.maximum(K - 1, 0)))
            fe[s, 1] = np.where((K == 0) | (K == nz), -1, el(I, J, np.minimum(K, nz - 1)))
            fp[s, 0] = np.where(K == 0, 4, 5)
            fp[s, 1] = np.where((K == 0) | (K == nz), -1, 4)
            idx = np.arange(nfx + nfy, nf)
            bd[4], bd[5] = idx[K == 0], idx[K == nz]
            self.face_elems, self.face_pos, self.bdfaces = fe, fp, bd
            self.face_vertices = None  # 3-D face-node coordinates are computed on demand
        self.face_elems = self.face_elems.astype(np.int64)
        self.face_pos = self.face_pos.astype(np.int64)
        self.elem_faces = self.elem_faces.astype(np.int64)
        self.elem_side = self.elem_side.astype(np.int64)
        self.nelements = self.elem_faces.shape[0]
        self.nfaces = self.face_elems.shape[0]
        self.face_boundary = np.full(self.nfaces, -1, dtype=np.int64)
        for b, fl in enumerate(self.bdfaces):
            self.face_boundary[fl] = b
        self.bdalive = list(range(2 * self.nd))  # original ids of the remaining boundaries
        self._orig_face = np.arange(self.nfaces)

    def nboundaries(self):
        return len(self.bdalive)

    # ------------------------------------------------------------------ periodicity
    def apply_periodicBCs(self, *pairs):
        """apply_periodicBCs!(mesh, "1" => "2", ...): pairs of boundary names ("1", "2")."""
        todo = []
        for p in pairs:
            try:
                b1, b2 = int(p[0]), int(p[1])
            except (TypeError, ValueError):
                raise ValueError("Boundary IDs must be integers.")
            nb = 2 * self.nd
            if not (1 <= b1 <= nb and 1 <= b2 <= nb):
                raise ValueError(f"Boundary IDs must be between 1 and {nb}.")
            if not (b2 % 2 == 0 and b1 + 1 == b2):
                raise ValueError(f"Boundaries {p[0]} and {p[1]} cannot be made periodic.")
            if (b1 - 1) not in self.bdalive or (b2 - 1) not in self.bdalive:
                raise ValueError(f"Boundaries {p[0]} and {p[1]} cannot be periodic.")
            todo.append((b1 - 1, b2 - 1))
        fe, fp = self.face_elems, self.face_pos
        delete = []
        for b1, b2 in todo:
            f1 = self.bdfaces[self.bdalive.index(b1)]
            f2 = self.bdfaces[self.bdalive.index(b2)]
            e2 = fe[f2, 0]
            p2 = fp[f2, 0]
            fe[f1, 1] = e2
            fp[f1, 1] = p2
            self.elem_faces[e2, p2] = f1
            self.elem_side[e2, p2] = 1
            delete.append(f2)
            self.periodic[str(b1 + 1)] = str(b2 + 1)
        for b1, b2 in todo:
            i1 = self.bdalive.index(b1)
            del self.bdfaces[i1], self.bdalive[i1]
            i2 = self.bdalive.index(b2)
            del self.bdfaces[i2], self.bdalive[i2]
        if not delete:
            return self
        dele = np.sort(np.concatenate(delete))
        keep = np.ones(self.nfaces, dtype=bool)
        keep[dele] = False
        newidx = np.cumsum(keep) - 1
        self.face_elems = fe[keep]
        self.face_pos = fp[keep]
        self._orig_face = self._orig_face[keep]
        self.nfaces = int(keep.sum())
        self.bdfaces = [newidx[b] for b in self.bdfaces]
        self.face_boundary = np.full(self.nfaces, -1, dtype=np.int64)
        for b, fl in enumerate(self.bdfaces):
            self.face_boundary[fl] = b
        # Mesh.jl:308-320: element face indices from the (renumbered) face table
        f = np.arange(self.nfaces)
        self.elem_faces[self.face_elems[:, 0], self.face_pos[:, 0]] = f
        sl = self.face_elems[:, 1] >= 0
        self.elem_faces[self.face_elems[sl, 1], self.face_pos[sl, 1]] = f[sl]
        return self

    # ------------------------------------------------------------------ geometry
    def apply_mapping(self, fn):
        """Curve the mesh: every node x of the Cartesian mesh moves to fn(x) (fn maps an
        (m, nd) array of points to an (m, nd) array).  Geometry then becomes per node,
        computed from the mapped nodes of each element (curved_geometry).  A periodic
        mesh stays periodic when fn(x + L e_k) = fn(x) + L e_k."""
        self.mapping = fn
        return self

    def element_geometry(self, std):
        """Per-element J, metric (column-major nd x nd) and volume (PhysicalRegions.jl:370-408)."""
        nd, dx = self.nd, self.Δx
        pdx = dx[0]
        for k in range(1, nd):
            pdx = pdx * dx[k]
        jac = pdx / (2 ** nd)
        M = np.zeros((nd, nd))
        if nd == 1:
            M[0, 0] = 1.0
        elif nd == 2:
            M[0, 0], M[1, 1] = dx[1] / 2, dx[0] / 2
        else:
            M[0, 0], M[1, 1], M[2, 2] = dx[1] * dx[2] / 4, dx[0] * dx[2] / 4, dx[0] * dx[1] / 4
        vol = 0.0
        for w in std.ω:
            vol = vol + jac * w
        ne = self.nelements
        return (np.full(ne, jac), np.tile(M.T.ravel(), (ne, 1)), np.full(ne, vol))

    def face_geometry(self):
        """Per-face frames (n, t, b) and surface J from the master position."""
        nd, dx = self.nd, self.Δx
        frames = np.zeros((2 * nd, 3, nd))
        fjac = np.zeros(2 * nd)
        if nd == 1:
            frames[0, 0, 0], frames[1, 0, 0] = -1, 1
            fjac[:] = 1.0
        elif nd == 2:
            frames[0, 0] = [-1, 0]; frames[0, 1] = [0, -1]
            frames[1, 0] = [1, 0]; frames[1, 1] = [0, 1]
            frames[2, 0] = [0, -1]; frames[2, 1] = [1, 0]
            frames[3, 0] = [0, 1]; frames[3, 1] = [-1, 0]
            fjac[0:2], fjac[2:4] = dx[1] / 2, dx[0] / 2
        else:
            frames[0] = [[-1, 0, 0], [0, -1, 0], [0, 0, 1]]
            frames[1] = [[1, 0, 0], [0, 1, 0], [0, 0, 1]]
            frames[2] = [[0, -1, 0], [0, 0, -1], [1, 0, 0]]
            frames[3] = [[0, 1, 0], [0, 0, 1], [1, 0, 0]]
            frames[4] = [[0, 0, -1], [-1, 0, 0], [0, 1, 0]]
            frames[5] = [[0, 0, 1], [1, 0, 0], [0, 1, 0]]
            fjac[0:2] = dx[1] * dx[2] / 4
            fjac[2:4] = dx[0] * dx[2] / 4
            fjac[4:6] = dx[0] * dx[1] / 4
        mp = self.face_pos[:, 0]
        return frames[mp], fjac[mp]

    def node_coords(self, std, elements=None):
        """Coordinates of every solution node, (ndofs, nd) (Mesh.jl:380-446), mapped when
        the mesh carries a mapping; `elements` (an index range) restricts them to those
        elements' nodes."""
        X = self._cartesian_coords(std.basis.ξ, elements)
        fn = getattr(self, "mapping", None)
        return X if fn is None else np.asarray(fn(X), dtype=np.float64)

    def equispaced_coords(self, std, elements=None):
        """Coordinates of the equispaced output points of every element, (ne*m^nd, nd):
        phys_coords(ξ, mesh, ie) for ξ in std.ξe (src/FlouSpatial/IO.jl:41-47), mapped when
        the mesh carries a mapping."""
        X = self._cartesian_coords(std.ξe1, elements)
        fn = getattr(self, "mapping", None)
        return X if fn is None else np.asarray(fn(X), dtype=np.float64)

    def _cartesian_coords(self, xi1, elements=None):
        nd = self.nd
        n = len(xi1)
        r = (np.asarray(xi1, dtype=np.float64) + 1.0) / 2.0
        nel = self.nelements_dir
        eidx = np.arange(self.nelements) if elements is None else np.asarray(elements)
        ne = eidx.size
        ei = [eidx % nel[0]]
        if nd >= 2:
            ei.append((eidx // nel[0]) % nel[1])
        if nd == 3:
            ei.append(eidx // (nel[0] * nel[1]))
        # node tensor indices, x fastest
        nidx = np.arange(n ** nd)
        ni = [nidx % n, (nidx // n) % n, nidx // (n * n)][:nd]
        X = np.zeros((ne, n ** nd, nd))
        for c in range(nd):
            lo = self.xyz[c][ei[c]][:, None]
            hi = self.xyz[c][ei[c] + 1][:, None]
            R = [r[ni[k]][None, :] for k in range(nd)]
            if nd == 1:
                X[:, :, c] = lo * (1.0 - R[0]) + hi * R[0]
            elif nd == 2:
                V = [lo, hi, hi, lo] if c == 0 else [lo, lo, hi, hi]
                X[:, :, c] = ((V[0] * (1 - R[0]) * (1 - R[1]) + V[1] * R[0] * (1 - R[1]))
                              + V[2] * R[0] * R[1]) + V[3] * (1 - R[0]) * R[1]
            else:
                if c == 0:
                    V = [lo, hi, hi, lo, lo, hi, hi, lo]
                elif c == 1:
                    V = [lo, lo, hi, hi, lo, lo, hi, hi]
                else:
                    V = [lo, lo, lo, lo, hi, hi, hi, hi]
                x = V[0] * (1 - R[0]) * (1 - R[1]) * (1 - R[2])
                x = x + V[1] * R[0] * (1 - R[1]) * (1 - R[2])
                x = x + V[2] * R[0] * R[1] * (1 - R[2])
                x = x + V[3] * (1 - R[0]) * R[1] * (1 - R[2])
                x = x + V[4] * (1 - R[0]) * (1 - R[1]) * R[2]
                x = x + V[5] * R[0] * (1 - R[1]) * R[2]
                x = x + V[6] * R[0] * R[1] * R[2]
                x = x + V[7] * (1 - R[0]) * R[1] * R[2]
                X[:, :, c] = x
        return X.reshape(ne * n ** nd, nd)

    def face_node_coords(self, std):
        """Coordinates of the nodes of every face, (nfaces, n^(nd-1), nd)."""
        nd, n = self.nd, std.n
        npf = n ** (nd - 1)
        out = np.zeros((self.nfaces, npf, nd))
        # a face node coincides with the corresponding trace node of the master element
        X = self.node_coords(std).reshape(self.nelements, n ** nd, nd)
        e = self.face_elems[:, 0]
        pos = self.face_pos[:, 0]
        for f_ in range(npf):
            for p in range(2 * nd):
                sel = pos == p
                if not np.any(sel):
                    continue
                node = _line_node(nd, n, p // 2, f_, 0 if p % 2 == 0 else n - 1)
                out[sel, f_] = X[e[sel], node]
        return out


def _line_node(nd, n, d, f, p):
    if nd == 1:
        return p
    if nd == 2:
        return p + n * f if d == 0 else f + n * p
    fi, fj = f % n, f // n
    if d == 0:
        return p + n * (fi + n * fj)
    if d == 1:
        return fi + n * (p + n * fj)
    return fi + n * (fj + n * p)


def apply_periodicBCs(mesh, *pairs):
    return mesh.apply_periodicBCs(*pairs)


def _dref(u, D, nd, n, k):
    """Derivative along reference direction k of nodal fields u (ne, n^nd), x fastest."""
    ne = u.shape[0]
    if nd == 1:
        return u @ D.T
    if nd == 2:
        v = u.reshape(ne, n, n)  # [e, iy, ix]
        r = np.einsum("ij,eyj->eyi", D, v) if k == 0 else np.einsum("ij,ejx->eix", D, v)
        return r.reshape(ne, n * n)
    v = u.reshape(ne, n, n, n)  # [e, iz, iy, ix]
    if k == 0:
        r = np.einsum("ij,ezyj->ezyi", D, v)
    elif k == 1:
        r = np.einsum("ij,ezjx->ezix", D, v)
    else:
        r = np.einsum("ij,ejyx->eiyx", D, v)
    return r.reshape(ne, n ** 3)


def curved_geometry(mesh, std):
    """Per-node geometry of a mapped mesh from its mapped GLL nodes.

    Metric terms Ja^i (columns of `metric`, PhysicalRegions.jl:1001-1033 convention) in the
    conservative curl form  Ja^i_n = -x̂_i · ∇_ξ × I^N(X_l ∇_ξ X_m), (n, m, l) cyclic
    (Kopriva, J. Sci. Comput. 26 (2006)), which satisfies the discrete metric identities
    Σ_i D_i Ja^i = 0 and so preserves free streams; in 2-D Ja^1 = (y_η, -x_η),
    Ja^2 = (-y_ξ, x_ξ).  J = a_1·(a_2 × a_3) from the interpolant's covariant basis.  Faces:
    per face node of the master element, n = ±Ja^d/|Ja^d| (outward), surface J = |Ja^d|,
    t, b completing an orthonormal frame.

    Returns (node_jac (ndofs,), node_metric (ndofs, nd*nd) column-major per node,
    elem_volume (ne,), face_frames (nfaces*npf, 3, nd), face_jac (nfaces*npf,))."""
    nd, n = mesh.nd, std.n
    ne, npe = mesh.nelements, n ** nd
    D = np.asarray(std.D, dtype=np.float64)
    Ja = np.zeros((ne, npe, nd, nd))  # Ja[..., r, c] = (Ja^c)_r = metric[r, c]
    J = np.zeros((ne, npe))
    # element chunks keep the temporaries of the metric products in cache
    chunk = max(1, (1 << 16) // npe)
    for e0 in range(0, ne, chunk):
        es = slice(e0, min(ne, e0 + chunk))
        Ja[es], J[es] = _curved_chunk(mesh.node_coords(std, range(es.start, es.stop)), D, nd, n)
    if np.any(J <= 0):
        raise ValueError("mapped mesh has a non-positive Jacobian")
    w = np.asarray(std.ω, dtype=np.float64)
    evol = (J * w[None, :]).sum(axis=1)
    metric = np.transpose(Ja, (0, 1, 3, 2)).reshape(ne * npe, nd * nd)  # column-major blocks
    # faces, from the master element
    npf = n ** (nd - 1)
    nf = mesh.nfaces
    frames = np.zeros((nf, npf, 3, nd))
    fjac = np.zeros((nf, npf))
    e_m, pos = mesh.face_elems[:, 0], mesh.face_pos[:, 0]
    for p in range(2 * nd):
        sel = np.nonzero(pos == p)[0]
        if sel.size == 0:
            continue
        d, lr = p // 2, p % 2
        nodes = np.array([_line_node(nd, n, d, f_, n - 1 if lr else 0) for f_ in range(npf)])
        v = Ja[e_m[sel][:, None], nodes[None, :], :, d]  # Ja^d at the face nodes (nsel, npf, nd)
        nrm = np.sqrt(np.sum(v * v, axis=2))
        nv = (1.0 if lr else -1.0) * v / nrm[..., None]
        fjac[sel] = nrm
        frames[sel, :, 0] = nv
        if nd == 2:
            frames[sel, :, 1, 0], frames[sel, :, 1, 1] = -nv[..., 1], nv[..., 0]
        elif nd == 3:
            ref = np.zeros_like(nv)
            big = np.abs(nv[..., 0]) < 0.9
            ref[..., 0] = big
            ref[..., 1] = ~big
            t = ref - np.sum(ref * nv, axis=2)[..., None] * nv
            t /= np.sqrt(np.sum(t * t, axis=2))[..., None]
            frames[sel, :, 1] = t
            frames[sel, :, 2] = np.cross(nv, t)
    return (J.reshape(-1), metric, evol, frames.reshape(nf * npf, 3, nd), fjac.reshape(-1))


def _curved_chunk(Xc, D, nd, n):
    """Curl-form metric terms and J of a chunk of elements from their mapped nodes Xc
    (nelem*n^nd, nd); see curved_geometry."""
    npe = n ** nd
    ne = Xc.shape[0] // npe
    X = Xc.reshape(ne, npe, nd)
    # The curl form is invariant under a constant shift of X; shifting each element to its
    # first node keeps the products X_l ∂X_m small and their round-off at the metric's scale.
    x = [X[:, :, c] - X[:, :1, c] for c in range(nd)]
    dX = [[_dref(x[c], D, nd, n, k) for k in range(nd)] for c in range(nd)]  # dX[c][k] = ∂_k X_c
    Ja = np.zeros((ne, npe, nd, nd))
    if nd == 1:
        Ja[:, :, 0, 0] = 1.0
        J = dX[0][0]
    elif nd == 2:
        Ja[:, :, 0, 0], Ja[:, :, 1, 0] = dX[1][1], -dX[0][1]
        Ja[:, :, 0, 1], Ja[:, :, 1, 1] = -dX[1][0], dX[0][0]
        J = dX[0][0] * dX[1][1] - dX[0][1] * dX[1][0]
    else:
        for nn in range(3):
            m, l = (nn + 1) % 3, (nn + 2) % 3
            V = [x[l] * dX[m][k] for k in range(3)]
            Ja[:, :, nn, 0] = -(_dref(V[2], D, 3, n, 1) - _dref(V[1], D, 3, n, 2))
            Ja[:, :, nn, 1] = -(_dref(V[0], D, 3, n, 2) - _dref(V[2], D, 3, n, 0))
            Ja[:, :, nn, 2] = -(_dref(V[1], D, 3, n, 0) - _dref(V[0], D, 3, n, 1))
        # J = a_1·(a_2 × a_3) with a_k = (∂_k X_0, ∂_k X_1, ∂_k X_2)
        a = [[dX[c][k] for c in range(3)] for k in range(3)]
        J = (a[0][0] * (a[1][1] * a[2][2] - a[1][2] * a[2][1])
             + a[0][1] * (a[1][2] * a[2][0] - a[1][0] * a[2][2])
             + a[0][2] * (a[1][0] * a[2][1] - a[1][1] * a[2][0]))
    return Ja, J
