"""ctypes wrapper of the CPU oracle (oracle/liboracle.so).

Test infrastructure: only the tests, ``__graft_entry__.smoke()`` and bench.py's
``cpu_baseline`` leg use it, always as the reference the HIP results are compared with.
"""
import ctypes
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(ROOT, "oracle", "liboracle.so")

EQ_ADVECTION, EQ_EULER = 0, 1
VOL_STRONG, VOL_SPLIT, VOL_HYBRID = 0, 1, 2
TP_STDAVG, TP_CHANDRASEKHAR, TP_KENNEDY_GRUBER = 0, 1, 2
AVG_STD, AVG_CHANDRASEKHAR = 0, 1
DISS_NONE, DISS_LXF, DISS_MATRIX, DISS_SCALAR = 0, 1, 2, 3
BC_DIRICHLET, BC_INFLOW, BC_OUTFLOW, BC_SLIP = 0, 1, 2, 3
NODES_GLL, NODES_GL = 0, 1
RK_SSPRK33, RK_CK2N54, RK_ORK256 = 0, 1, 2
MON_KINETIC_ENERGY, MON_ENTROPY = 0, 1

_lib = None

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)
_lp = ctypes.POINTER(ctypes.c_long)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("oracle not built: run `make -C oracle`")
        L = ctypes.CDLL(LIB_PATH)
        L.oracle_create.restype = ctypes.c_void_p
        L.oracle_create.argtypes = [ctypes.c_int, ctypes.c_int, _ip, _dp, _dp, _ip, _dp,
                                    ctypes.c_int, _dp, ctypes.c_int, ctypes.c_int,
                                    ctypes.c_int, ctypes.c_int, ctypes.c_double]
        L.oracle_destroy.argtypes = [ctypes.c_void_p]
        L.oracle_set_option.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int]
        L.oracle_set_curved.argtypes = [ctypes.c_void_p, _dp, _dp, _dp, _dp, _dp]
        L.oracle_set_limiter.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_double]
        L.oracle_set_blend.argtypes = [ctypes.c_void_p, ctypes.c_double]
        L.oracle_limit.argtypes = [ctypes.c_void_p, _dp, ctypes.c_double]
        L.oracle_monitor.argtypes = [ctypes.c_void_p, _dp, ctypes.c_int, _dp]
        L.oracle_sizes.argtypes = [ctypes.c_void_p, _lp]
        L.oracle_connectivity.argtypes = [ctypes.c_void_p, _lp, _ip, _lp, _ip, _ip]
        L.oracle_geometry.argtypes = [ctypes.c_void_p, _dp, _dp, _dp, _dp, _dp]
        L.oracle_coords.argtypes = [ctypes.c_void_p, _dp, _dp]
        L.oracle_set_connectivity.argtypes = [ctypes.c_void_p, _lp, _ip, _lp, _ip, _ip]
        L.oracle_master2slave.argtypes = [ctypes.c_int] * 4
        L.oracle_slave2master.argtypes = [ctypes.c_int] * 4
        L.oracle_set_bc.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, _dp]
        L.oracle_rhs.argtypes = [ctypes.c_void_p, _dp, _dp]
        L.oracle_max_dt.restype = ctypes.c_double
        L.oracle_max_dt.argtypes = [ctypes.c_void_p, _dp, ctypes.c_double]
        L.oracle_integrate.restype = ctypes.c_double
        L.oracle_integrate.argtypes = [ctypes.c_void_p, _dp, ctypes.c_int]
        L.oracle_rk_run.argtypes = [ctypes.c_void_p, ctypes.c_int, _dp, ctypes.c_double,
                                    ctypes.c_int]
        L.oracle_build_ops.argtypes = [ctypes.c_int, ctypes.c_int] + [_dp] * 9
        L.oracle_tableau_2n.argtypes = [ctypes.c_int, _dp, _dp, _dp]
        L.oracle_rhs_jvp.argtypes = [ctypes.c_void_p, _dp, _dp, _dp, _dp]
        L.oracle_rk_run_jvp.argtypes = [ctypes.c_void_p, ctypes.c_int, _dp, _dp,
                                        ctypes.c_double, ctypes.c_int]
        L.oracle_logmean.restype = ctypes.c_double
        L.oracle_logmean.argtypes = [ctypes.c_double, ctypes.c_double]
        L.oracle_twopoint.argtypes = [ctypes.c_void_p, _dp, _dp, _dp, _dp, _dp]
        L.oracle_numflux.argtypes = [ctypes.c_void_p, _dp, _dp, _dp, _dp]
        L.oracle_set_threads.argtypes = [ctypes.c_int]
        L.oracle_bc_point.argtypes = [ctypes.c_void_p, ctypes.c_int, _dp, _dp, _dp]
        L.oracle_cons2entropy.argtypes = [ctypes.c_int, ctypes.c_double, _dp, _dp]
        L.oracle_entropy2cons.argtypes = [ctypes.c_int, ctypes.c_double, _dp, _dp]
        L.oracle_interp_matrix.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, _dp, _dp]
        L.oracle_equispaced.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, _dp, ctypes.c_int64,
                                        ctypes.c_int, _dp, _dp]
        L.oracle_vtk_connectivities.restype = ctypes.c_int64
        L.oracle_vtk_connectivities.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_int64)]
        L.oracle_last_error.restype = ctypes.c_char_p
        L.oracle_last_error.argtypes = [ctypes.c_void_p]
        _lib = L
    return _lib


def _p(a, t=_dp):
    return a.ctypes.data_as(t)


def build_ops(n, nodes="GLL"):
    """Operators (xi, w, D, Ds, Dsh, lm, lp, gm, gp) from the oracle's long-double build."""
    nt = NODES_GLL if nodes == "GLL" else NODES_GL
    out = [np.zeros(n), np.zeros(n), np.zeros((n, n)), np.zeros((n, n)), np.zeros((n, n)),
           np.zeros(n), np.zeros(n), np.zeros(n), np.zeros(n)]
    rc = lib().oracle_build_ops(nt, n, *[_p(a) for a in out])
    if rc != 0:
        raise ValueError("oracle_build_ops failed")
    return dict(zip(["xi", "w", "D", "Ds", "Dsh", "lm", "lp", "gm", "gp"], out))


def pack_ops(ops):
    return np.concatenate([np.ravel(np.asarray(ops[k], dtype=np.float64))
                           for k in ["xi", "w", "D", "Ds", "Dsh", "lm", "lp", "gm", "gp"]])


class OracleDisc:
    """A Cartesian multi-element DGSEM discretisation held by the oracle."""

    def __init__(self, nd, n, nel, start, finish, periodic, ops, eq, eqparams,
                 volop, tpflux, savg, sdiss, intensity=1.0):
        self.nd, self.n = nd, n
        packed = pack_ops(ops)
        nel_a = np.asarray(nel, dtype=np.int32)
        st = np.asarray(start, dtype=np.float64)
        fi = np.asarray(finish, dtype=np.float64)
        per = np.asarray(periodic, dtype=np.int32)
        eqp = np.asarray(eqparams, dtype=np.float64)
        self.h = lib().oracle_create(nd, n, _p(nel_a, _ip), _p(st), _p(fi), _p(per, _ip),
                                     _p(packed), eq, _p(eqp), volop, tpflux, savg, sdiss,
                                     float(intensity))
        if not self.h:
            raise ValueError("oracle_create failed")
        sz = np.zeros(6, dtype=np.int64)
        lib().oracle_sizes(self.h, _p(sz, _lp))
        self.ne, self.nf, self.npe, self.npf, self.nv, self.nbd = [int(x) for x in sz]
        self.ndofs = self.ne * self.npe

    def __del__(self):
        if getattr(self, "h", None):
            lib().oracle_destroy(self.h)
            self.h = None

    def set_curved(self, node_metric, node_jac, face_frames, face_jac, elem_volume):
        """Per-node geometry of a mapped mesh (arrays as from flouhip.mesh.curved_geometry)."""
        a = [np.ascontiguousarray(x, dtype=np.float64)
             for x in (node_metric, node_jac, face_frames, face_jac, elem_volume)]
        self._keep_curved = a
        lib().oracle_set_curved(self.h, *[_p(x) for x in a])

    def set_option(self, which, value):
        lib().oracle_set_option(self.h, which, value)

    def set_blend(self, blend):
        """Blending constant c of the HybridDivOperator (volop VOL_HYBRID)."""
        lib().oracle_set_blend(self.h, float(blend))

    def set_limiter(self, on, minval=0.0):
        """Zhang & Shu stage limiter applied by rk_run after every stage."""
        lib().oracle_set_limiter(self.h, int(bool(on)), float(minval))

    def limit(self, Q, minval):
        u = np.array(Q, dtype=np.float64, order="F", copy=True)
        self._check(lib().oracle_limit(self.h, _p(u), float(minval)))
        return u

    def monitor(self, Q, kind):
        Q = np.asfortranarray(Q, dtype=np.float64)
        out = np.zeros(1)
        self._check(lib().oracle_monitor(self.h, _p(Q), int(kind), _p(out)))
        return float(out[0])

    def connectivity(self):
        nd = self.nd
        e_face = np.zeros(self.ne * 2 * nd, dtype=np.int64)
        e_side = np.zeros(self.ne * 2 * nd, dtype=np.int32)
        f_elem = np.zeros(self.nf * 2, dtype=np.int64)
        f_pos = np.zeros(self.nf * 2, dtype=np.int32)
        f_bd = np.zeros(self.nf, dtype=np.int32)
        lib().oracle_connectivity(self.h, _p(e_face, _lp), _p(e_side, _ip), _p(f_elem, _lp),
                                  _p(f_pos, _ip), _p(f_bd, _ip))
        return dict(e_face=e_face.reshape(self.ne, 2 * nd), e_side=e_side.reshape(self.ne, 2 * nd),
                    f_elem=f_elem.reshape(self.nf, 2), f_pos=f_pos.reshape(self.nf, 2), f_bd=f_bd)

    def set_connectivity(self, e_face, e_side, f_elem, f_pos, f_orient):
        """Replace the connectivity (same element and face counts), with face orientations."""
        self._conn = [np.ascontiguousarray(e_face, dtype=np.int64).reshape(-1),
                      np.ascontiguousarray(e_side, dtype=np.int32).reshape(-1),
                      np.ascontiguousarray(f_elem, dtype=np.int64).reshape(-1),
                      np.ascontiguousarray(f_pos, dtype=np.int32).reshape(-1),
                      np.ascontiguousarray(f_orient, dtype=np.int32).reshape(-1)]
        c = self._conn
        assert c[0].size == self.ne * 2 * self.nd and c[2].size == 2 * self.nf and c[4].size == self.nf
        lib().oracle_set_connectivity(self.h, _p(c[0], _lp), _p(c[1], _ip), _p(c[2], _lp), _p(c[3], _ip),
                                      _p(c[4], _ip))

    def geometry(self):
        nd = self.nd
        e_jac = np.zeros(self.ne)
        e_metric = np.zeros(self.ne * nd * nd)
        e_vol = np.zeros(self.ne)
        f_frame = np.zeros(self.nf * 3 * nd)
        f_jac = np.zeros(self.nf)
        lib().oracle_geometry(self.h, _p(e_jac), _p(e_metric), _p(e_vol), _p(f_frame), _p(f_jac))
        return dict(e_jac=e_jac, e_metric=e_metric.reshape(self.ne, nd, nd).transpose(0, 2, 1),
                    e_vol=e_vol, f_frame=f_frame.reshape(self.nf, 3, nd), f_jac=f_jac)

    def coords(self):
        c = np.zeros(self.ndofs * self.nd)
        fc = np.zeros(self.nf * self.npf * self.nd)
        lib().oracle_coords(self.h, _p(c), _p(fc))
        return c.reshape(self.ndofs, self.nd), fc.reshape(self.nf, self.npf, self.nd)

    def set_bc(self, ib, bctype, data=None):
        if data is None:
            data = np.zeros(1)
        data = np.ascontiguousarray(data, dtype=np.float64)
        rc = lib().oracle_set_bc(self.h, ib, bctype, _p(data))
        if rc != 0:
            raise ValueError("bad boundary index")

    def _check(self, rc):
        if rc != 0:
            raise ArithmeticError(lib().oracle_last_error(self.h).decode())

    def rhs(self, Q):
        Q = np.asfortranarray(Q, dtype=np.float64)
        dQ = np.zeros_like(Q, order="F")
        self._check(lib().oracle_rhs(self.h, _p(Q), _p(dQ)))
        return dQ

    def rhs_jvp(self, Q, dQdir):
        Q = np.asfortranarray(Q, dtype=np.float64)
        dr = np.asfortranarray(dQdir, dtype=np.float64)
        out = np.zeros_like(Q, order="F")
        dout = np.zeros_like(Q, order="F")
        self._check(lib().oracle_rhs_jvp(self.h, _p(Q), _p(dr), _p(out), _p(dout)))
        return out, dout

    def twopoint(self, Q1, Q2, Ja1, Ja2):
        a = [np.ascontiguousarray(x, dtype=np.float64) for x in (Q1, Q2, Ja1, Ja2)]
        F = np.zeros(self.nv)
        self._check(lib().oracle_twopoint(self.h, *[_p(x) for x in a], _p(F)))
        return F

    def numflux(self, Ql, Qr, frame):
        a = [np.ascontiguousarray(x, dtype=np.float64) for x in (Ql, Qr, frame)]
        F = np.zeros(self.nv)
        self._check(lib().oracle_numflux(self.h, *[_p(x) for x in a], _p(F)))
        return F

    def bc_point(self, ib, Qin, frame):
        """Exterior state boundary ib assigns to the trace Qin at a face node with frame
        (n, t, b) (inflow, outflow and slip boundaries)."""
        a = [np.ascontiguousarray(x, dtype=np.float64) for x in (Qin, frame)]
        Q = np.zeros(self.nv)
        if lib().oracle_bc_point(self.h, int(ib), _p(a[0]), _p(a[1]), _p(Q)) != 0:
            raise ValueError("no point form for this boundary")
        return Q

    def max_dt(self, Q, cfl):
        Q = np.asfortranarray(Q, dtype=np.float64)
        return lib().oracle_max_dt(self.h, _p(Q), float(cfl))

    def integrate(self, Q, ivar=0):
        Q = np.asfortranarray(Q, dtype=np.float64)
        return lib().oracle_integrate(self.h, _p(Q), int(ivar))

    def rk_run(self, scheme, Q, dt, nsteps):
        u = np.array(Q, dtype=np.float64, order="F", copy=True)
        self._check(lib().oracle_rk_run(self.h, scheme, _p(u), float(dt), int(nsteps)))
        return u

    def rk_run_jvp(self, scheme, Q, dQ, dt, nsteps):
        u = np.array(Q, dtype=np.float64, order="F", copy=True)
        du = np.array(dQ, dtype=np.float64, order="F", copy=True)
        self._check(lib().oracle_rk_run_jvp(self.h, scheme, _p(u), _p(du), float(dt), int(nsteps)))
        return u, du


def master2slave(nd, n, i, o):
    return lib().oracle_master2slave(nd, n, i, o)


def slave2master(nd, n, i, o):
    return lib().oracle_slave2master(nd, n, i, o)


def set_threads(n):
    """OpenMP threads of the oracle's loops started from the calling thread."""
    lib().oracle_set_threads(int(n))


def logmean(a, b):
    return lib().oracle_logmean(float(a), float(b))


def cons2entropy(nd, gamma, Q):
    Q = np.ascontiguousarray(Q, dtype=np.float64)
    W = np.zeros(nd + 2)
    lib().oracle_cons2entropy(nd, float(gamma), _p(Q), _p(W))
    return W


def entropy2cons(nd, gamma, W):
    W = np.ascontiguousarray(W, dtype=np.float64)
    Q = np.zeros(nd + 2)
    if lib().oracle_entropy2cons(nd, float(gamma), _p(W), _p(Q)) != 0:
        raise ArithmeticError("entropy2cons: pow of a negative number")
    return Q


def interp_matrix(n, xe, nodes="GLL"):
    """A[i, j] = l_j(xe[i]) of the n-node Lagrange basis (interp_matrix, ApproximationBases.jl:85-87)."""
    xe = np.ascontiguousarray(xe, dtype=np.float64)
    A = np.zeros((xe.size, n))
    if lib().oracle_interp_matrix(NODES_GLL if nodes == "GLL" else NODES_GL, n, xe.size, _p(xe), _p(A)) != 0:
        raise ValueError("oracle_interp_matrix failed")
    return A


def equispaced(nd, n, A, ne, Q):
    """Element-wise Qe = node2eq * Q (StdRegions.jl:232-238) for Q of shape (ne*n^nd, ncols),
    column-major as the state matrix; returns (ne*m^nd, ncols)."""
    A = np.ascontiguousarray(A, dtype=np.float64)
    m = A.shape[0]
    Q = np.asfortranarray(Q, dtype=np.float64)
    ncols = 1 if Q.ndim == 1 else Q.shape[1]
    Qe = np.zeros((ne * m ** nd, ncols), order="F")
    if lib().oracle_equispaced(nd, n, m, _p(A), ne, ncols, _p(Q), _p(Qe)) != 0:
        raise ValueError("oracle_equispaced failed")
    return Qe


def vtk_connectivities(nd, n):
    """vtk_connectivities(std) (StdSegment.jl:171-175, StdQuad.jl:185-196, StdHex.jl:174-199)."""
    out = np.zeros(n ** nd, dtype=np.int64)
    cnt = lib().oracle_vtk_connectivities(nd, n, out.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)))
    assert cnt == n ** nd, cnt
    return out


def tableau_2n(scheme):
    A, B, c = np.zeros(5), np.zeros(5), np.zeros(5)
    lib().oracle_tableau_2n(scheme, _p(A), _p(B), _p(c))
    return A, B, c
