"""Meshes with non-zero face orientations, for the tests.

Flou's CartesianMesh numbers every element's nodes along the global axes, so all of its
faces have orientation 0; meshes read from files do not.  `ReorientedMesh` takes a
periodic CartesianMesh and renumbers chosen elements by a rotation of the reference element
(a signed permutation of the reference axes with determinant +1): the element covers the
same cell, its nodes, local faces and metric are permuted, and the faces it shares with
differently numbered neighbours get the orientation that makes master node i meet slave
node master2slave(i, o) (StdSegment.jl:161-167, StdQuad.jl:161-181).  The orientation of
each face is the code whose master2slave (restated here from the reference's table) pairs
face nodes with equal coordinates (modulo the periods); exactly one code must.

The object has the attributes MultielementDisc reads (connectivity, `face_orientation`,
`node_coords`, a non-None `mapping` so that the per-node geometry comes from
`flouhip.mesh.curved_geometry`); after `make_affine()` it reports per-element and per-face
geometry instead (`element_geometry`, `face_geometry`), taken from the same computation.
"""
import itertools

import numpy as np

from flouhip import mesh as M


def rotations(nd):
    """Proper signed permutations of nd reference axes: (perm, signs) with
    ξ_old[perm[k]] = signs[k] ξ_new[k]."""
    out = []
    for perm in itertools.permutations(range(nd)):
        for signs in itertools.product((1, -1), repeat=nd):
            R = np.zeros((nd, nd))
            for k in range(nd):
                R[perm[k], k] = signs[k]
            if round(np.linalg.det(R)) == 1:
                out.append((tuple(perm), tuple(signs)))
    return out


def node_permutation(nd, n, perm, signs):
    """old[new]: the old node index of each new node index (x fastest)."""
    idx = np.indices((n,) * nd).reshape(nd, -1)[::-1]  # idx[k] = index along axis k, x fastest
    old = np.zeros_like(idx)
    for k in range(nd):
        old[perm[k]] = idx[k] if signs[k] == 1 else n - 1 - idx[k]
    return sum(old[k] * n ** k for k in range(nd))


def face_permutation(nd, perm, signs):
    """old_face[new_face]: local face 2d + lr of the new numbering is old face
    2 perm[d] + lr' with lr' = lr when signs[d] = 1, else 1 - lr."""
    out = []
    for k in range(2 * nd):
        d, lr = divmod(k, 2)
        out.append(2 * perm[d] + (lr if signs[d] == 1 else 1 - lr))
    return out


def master2slave_py(nd, n, i, o):
    """The reference's master2slave (0-based), for the orientation search."""
    if o == 0 or nd == 1:
        return i
    if nd == 2:
        return n - 1 - i
    a, b, r = i % n, i // n, n - 1
    x, y = {1: (b, r - a), 2: (r - a, r - b), 3: (r - b, a), 4: (b, a), 5: (r - a, b),
            6: (r - b, r - a), 7: (a, r - b)}[o]
    return x + n * y


class ReorientedMesh:
    def __init__(self, base, std, elem_rot, lengths):
        """base: a periodic CartesianMesh; elem_rot: {element: (perm, signs)};
        lengths: the domain's period per direction."""
        nd, n = base.nd, std.n
        self.nd, self.base, self.std = nd, base, std
        self.nelements, self.nfaces = base.nelements, base.nfaces
        self.bdalive = list(base.bdalive)
        self.face_boundary = np.array(base.face_boundary)
        self.elem_faces = np.array(base.elem_faces).copy()
        self.elem_side = np.array(base.elem_side).copy()
        self.face_elems = np.array(base.face_elems).copy()
        self.face_pos = np.array(base.face_pos).copy()
        self.mapping = lambda X: X
        self.affine = False
        npe = n ** nd
        self.node_perm = np.tile(np.arange(npe), (self.nelements, 1))
        for e, (perm, signs) in elem_rot.items():
            self.node_perm[e] = node_permutation(nd, n, perm, signs)
            fp = face_permutation(nd, perm, signs)
            ef, es = self.elem_faces[e].copy(), self.elem_side[e].copy()
            for knew, kold in enumerate(fp):
                self.elem_faces[e, knew], self.elem_side[e, knew] = ef[kold], es[kold]
                self.face_pos[ef[kold], es[kold]] = knew
        self._X = base.node_coords(std).reshape(self.nelements, npe, nd)
        self._X = np.take_along_axis(self._X, self.node_perm[:, :, None], axis=1)
        self.face_orientation = self._orientations(n, np.asarray(lengths, dtype=np.float64))
        self._geom = M.curved_geometry(self, std)

    def nboundaries(self):
        return self.base.nboundaries()

    def node_coords(self, std, elements=None):
        X = self._X if elements is None else self._X[list(elements)]
        return X.reshape(-1, self.nd)

    def face_node_coords(self, std):
        return self.base.face_node_coords(std)

    def _face_nodes(self, e, k, n):
        d, lr = divmod(k, 2)
        npf = n ** (self.nd - 1)
        return np.array([M._line_node(self.nd, n, d, f, n - 1 if lr else 0) for f in range(npf)])

    def _orientations(self, n, L):
        nd = self.nd
        npf = n ** (nd - 1)
        nor = 8 if nd == 3 else (2 if nd == 2 else 1)
        out = np.zeros(self.nfaces, dtype=np.uint8)
        for f in range(self.nfaces):
            (em, es), (km, ks) = self.face_elems[f], self.face_pos[f]
            if es < 0:
                continue
            Xm = self._X[em][self._face_nodes(em, km, n)]
            Xs = self._X[es][self._face_nodes(es, ks, n)]
            found = []
            for o in range(nor):
                j = np.array([master2slave_py(nd, n, i, o) for i in range(npf)])
                d = Xm - Xs[j]
                d -= L * np.round(d / L)
                if np.max(np.abs(d)) < 1e-9:
                    found.append(o)
            if len(found) != 1:
                raise AssertionError(f"face {f}: orientations {found} match")
            out[f] = found[0]
        return out

    # per-node geometry for MultielementDisc (mapping is not None) and the oracle
    def geometry(self):
        return self._geom

    def make_affine(self):
        """Report per-element and per-face geometry (node 0 and face node 0 of the per-node
        geometry), and hold the per-node geometry as those values broadcast, so that an
        oracle given geometry() sees exactly what the affine kernels see."""
        J, met, ev, fr, fj = self._geom
        npe, npf = self.std.n ** self.nd, self.std.n ** (self.nd - 1)
        self._geom = (np.repeat(J[::npe], npe), np.repeat(met[::npe], npe, axis=0), ev,
                      np.repeat(fr[::npf], npf, axis=0), np.repeat(fj[::npf], npf))
        self.mapping, self.affine = None, True
        return self

    # affine form: per-element metric and J from node 0, per-face frame from face node 0
    def element_geometry(self, std):
        J, met, ev, fr, fj = self._geom
        npe = std.n ** self.nd
        return J[::npe].copy(), met[::npe].copy(), ev.copy()

    def face_geometry(self):
        J, met, ev, fr, fj = self._geom
        npf = self.std.n ** (self.nd - 1)
        return fr[::npf].copy(), fj[::npf].copy()

    def permute_state(self, Q_base):
        """(ndofs, nv) state of the base mesh -> the same field in this mesh's numbering."""
        nv = Q_base.shape[1]
        npe = self.node_perm.shape[1]
        Qe = Q_base.reshape(self.nelements, npe, nv)
        return np.take_along_axis(Qe, self.node_perm[:, :, None], axis=1).reshape(-1, nv)
