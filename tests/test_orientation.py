"""Face orientations (master2slave, StdSegment.jl:153-167, StdQuad.jl:138-181; used by
interface_fluxes!, Interfaces.jl:111-136) in the oracle.

* The oracle's tables: on quad faces, master2slave(·, o) is ρ^o for o = 0..3 and τ∘ρ^(o-4)
  for o = 4..7, with ρ(a, b) = (b, r - a) and τ(a, b) = (b, a) (r = n - 1) -- the dihedral
  group of the square, as the reference's table lists it -- and slave2master(·, o) is its
  inverse; on segment faces o = 1 reverses the nodes.  A transposed or mistyped entry
  breaks these identities.
* A mesh whose elements number their nodes by rotations of the reference element
  (tests/mesh_orient.py) describes the same discretisation as the Cartesian one, so the
  oracle's rhs on it must be the Cartesian rhs with the nodes renumbered (up to the
  round-off of the permuted metric and summation order).  Ignoring the orientations (a
  mistake planted in the oracle's input) breaks that by O(1).
"""
import numpy as np
import pytest

import flouhip as F
import oracle_lib as O
from cases import GAMMA, periodic_euler_state, warp
from flouhip import mesh as M
from mesh_orient import ReorientedMesh, master2slave_py, rotations


def _rho(n, f):
    a, b, r = f % n, f // n, n - 1
    return b + n * (r - a)


def _tau(n, f):
    return f // n + n * (f % n)


@pytest.mark.parametrize("n", [2, 3, 4, 6])
def test_quad_tables_are_the_dihedral_group(n):
    npf = n * n
    maps = []
    for o in range(8):
        m2s = [O.master2slave(3, n, i, o) for i in range(npf)]
        s2m = [O.slave2master(3, n, i, o) for i in range(npf)]
        ref = list(range(npf))
        for _ in range(o % 4):
            ref = [_rho(n, x) for x in ref]
        if o >= 4:
            ref = [_tau(n, x) for x in ref]
        assert m2s == ref, o
        assert sorted(m2s) == list(range(npf))
        assert [s2m[m2s[i]] for i in range(npf)] == list(range(npf)), o
        assert [m2s[s2m[i]] for i in range(npf)] == list(range(npf)), o
        assert m2s == [master2slave_py(3, n, i, o) for i in range(npf)]
        maps.append(tuple(m2s))
    if n > 1:
        assert len(set(maps)) == 8


@pytest.mark.parametrize("n", [2, 5])
def test_segment_tables(n):
    for o in (0, 1):
        m2s = [O.master2slave(2, n, i, o) for i in range(n)]
        assert m2s == (list(range(n)) if o == 0 else list(range(n - 1, -1, -1)))
        assert [O.slave2master(2, n, i, o) for i in range(n)] == m2s


def build(nd, n, nel, vol, tp, seed=3, curved=False, frac=0.5, affine=False):
    """(base mesh, reoriented mesh, oracles on both, state on both numberings); affine:
    the reoriented mesh reports per-element geometry and its oracle gets the same."""
    b = F.LagrangeBasis("GLL", n)
    std = [None, None, F.StdQuad, F.StdHex][nd](b, F.DGSEMrec(b), nd + 2)
    base = F.CartesianMesh(nd, (0.0,) * nd, (1.0,) * nd, (nel,) * nd)
    base.apply_periodicBCs(*[(str(2 * k + 1), str(2 * k + 2)) for k in range(nd)])
    base.apply_mapping(warp if curved else (lambda X: X))
    rng = np.random.default_rng(seed)
    rots = rotations(nd)
    elem_rot = {e: rots[rng.integers(1, len(rots))] for e in range(base.nelements) if rng.random() < frac}
    mesh = ReorientedMesh(base, std, elem_rot, [1.0] * nd)
    if affine:
        mesh.make_affine()
    volc, tpc = (O.VOL_STRONG, 0) if vol == "strong" else (O.VOL_SPLIT, tp)

    def oracle(geom):
        orc = O.OracleDisc(nd, n, [nel] * nd, [0.0] * nd, [1.0] * nd, [1] * nd, O.build_ops(n, "GLL"),
                           O.EQ_EULER, [GAMMA], volc, tpc, O.AVG_STD, O.DISS_LXF, 1.0)
        J, met, ev, fr, fj = geom
        orc.set_curved(met, J, fr.reshape(-1), fj, ev)
        return orc

    ob = oracle(M.curved_geometry(base, std))
    orr = oracle(mesh.geometry())
    orr.set_connectivity(mesh.elem_faces, mesh.elem_side, mesh.face_elems, mesh.face_pos, mesh.face_orientation)
    Qb = periodic_euler_state(base.node_coords(std), nd, 1.0, seed=seed, amp=0.2)
    Qr = mesh.permute_state(Qb)
    return base, mesh, ob, orr, Qb, Qr, std


def test_reoriented_meshes_use_every_orientation():
    _, mesh3, *_ = build(3, 3, 3, "split", O.TP_CHANDRASEKHAR, frac=0.7)
    assert set(np.unique(mesh3.face_orientation)) == set(range(8))
    _, mesh2, *_ = build(2, 3, 4, "split", O.TP_CHANDRASEKHAR)
    assert set(np.unique(mesh2.face_orientation)) == {0, 1}


CASES = [(2, 4, 4, "split", O.TP_CHANDRASEKHAR, False), (2, 5, 3, "strong", 0, True),
         (3, 3, 3, "split", O.TP_CHANDRASEKHAR, False), (3, 4, 3, "split", O.TP_KENNEDY_GRUBER, True),
         (3, 3, 2, "strong", 0, False)]


@pytest.mark.parametrize("case", CASES, ids=lambda c: "-".join(map(str, c)))
def test_oracle_rhs_is_the_cartesian_rhs_renumbered(case):
    nd, n, nel, vol, tp, curved = case
    base, mesh, ob, orr, Qb, Qr, std = build(nd, n, nel, vol, tp, curved=curved)
    ref = mesh.permute_state(ob.rhs(Qb))
    got = orr.rhs(Qr)
    scale = np.max(np.abs(ref), axis=0)
    err = np.max(np.abs(got - ref) / scale)
    print(f"{case}: reoriented vs renumbered Cartesian rhs, max rel err {err:.2e}; "
          f"orientations used {sorted(set(mesh.face_orientation.tolist()))}")
    assert err <= 1e-11
    # planted mistake: the same mesh with its orientations ignored
    orr.set_connectivity(mesh.elem_faces, mesh.elem_side, mesh.face_elems, mesh.face_pos,
                         np.zeros_like(mesh.face_orientation))
    bad = np.max(np.abs(orr.rhs(Qr) - ref) / scale)
    print(f"  orientations ignored: {bad:.2e}")
    assert bad > 1e-3


def test_oracle_steps_on_reoriented_mesh():
    base, mesh, ob, orr, Qb, Qr, std = build(3, 4, 2, "split", O.TP_CHANDRASEKHAR, curved=True)
    dt = 0.2 * ob.max_dt(Qb, 1.0)
    ref = mesh.permute_state(ob.rk_run(O.RK_SSPRK33, Qb, dt, 3))
    got = orr.rk_run(O.RK_SSPRK33, Qr, dt, 3)
    err = np.max(np.abs(got - ref)) / np.max(np.abs(ref))
    print(f"3 SSPRK33 steps: {err:.2e}")
    assert err <= 1e-12
