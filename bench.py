"""Headline benchmark: DOF-updates/s per RK stage of Flou's DGSEM rhs + RK stage on one
MI355X, FP32 (chosen form) vs FP64 (plain form).

Workload (default): 3-D compressible Euler, periodic 64^3 hexahedra on [0, 2π]^3, N = 5
(216 nodes per element, 56,623,104 nodes x 5 conserved variables), Taylor-Green initial
state (M0 = 0.1) plus seeded N(0, 1e-3^2) noise on every primitive variable,
SplitDivOperator(ChandrasekharAverage(), LxF(StdAverage(), 1.0)), CarpenterKennedy2N54 with
a fixed dt (CFL 0.3 of the initial state).  One "step" = one CK2N54 time step = 5 fused
rhs+stage kernels.  value = nodes x stages / time for the FP32 form; the FP64 plain form is
timed on the same steps and reported next to it, with the density L2 error of FP32 against
FP64 after the same number of steps.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-fp64] [--no-cpu]

Prints one JSON line.
"""
import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "flou.jl_amd"))

GAMMA = 1.4
M0 = 0.1
NEL = 64
N_NODES = 6  # N = 5


def build_disc(precision, nel=NEL, n=N_NODES):
    import numpy as np  # noqa: F401

    import flouhip as F
    b = F.LagrangeBasis("GLL", n)
    std = F.StdHex(b, F.DGSEMrec(b), 5)
    L = 2 * math.pi
    mesh = F.CartesianMesh(3, (0, 0, 0), (L, L, L), (nel, nel, nel))
    mesh.apply_periodicBCs(("1", "2"), ("3", "4"), ("5", "6"))
    eq = F.EulerEquation(3, GAMMA)
    op = F.SplitDivOperator(F.ChandrasekharAverage(), F.LxF(F.StdAverage(), 1.0))
    return F.MultielementDisc(mesh, std, eq, op, precision=precision), eq, std


def taylor_green_state(disc, std, seed=1234):
    """Device FP64 state (5, ndofs) of the Taylor-Green vortex plus seeded noise."""
    import torch
    nel, n = disc.mesh.nelements_dir[0], std.n
    L = 2 * math.pi
    dx = L / nel
    xi = torch.tensor(std.basis.ξ, dtype=torch.float64, device="cuda")
    e = torch.arange(disc.nelements, device="cuda")
    ei, ej, ek = e % nel, (e // nel) % nel, e // (nel * nel)
    a = torch.arange(n ** 3, device="cuda")
    ia, ib, ic = a % n, (a // n) % n, a // (n * n)
    x = ((ei[:, None] + (xi[ia][None, :] + 1) / 2) * dx).reshape(-1)
    y = ((ej[:, None] + (xi[ib][None, :] + 1) / 2) * dx).reshape(-1)
    z = ((ek[:, None] + (xi[ic][None, :] + 1) / 2) * dx).reshape(-1)
    rho = torch.ones_like(x)
    u = M0 * torch.sin(x) * torch.cos(y) * torch.cos(z)
    v = -M0 * torch.cos(x) * torch.sin(y) * torch.cos(z)
    w = torch.zeros_like(x)
    p = 1 / GAMMA + (rho * M0 ** 2 / 16) * (torch.cos(2 * x) + torch.cos(2 * y)) * (torch.cos(2 * z) + 2)
    g = torch.Generator(device="cuda")
    g.manual_seed(seed)
    prim = torch.stack([rho, u, v, w, p])
    prim += 1e-3 * torch.randn(prim.shape, generator=g, device="cuda", dtype=torch.float64)
    Q = torch.empty_like(prim)
    Q[0] = prim[0]
    Q[1:4] = prim[0] * prim[1:4]
    Q[4] = prim[4] / (GAMMA - 1) + prim[0] * (prim[1] ** 2 + prim[2] ** 2 + prim[3] ** 2) / 2
    return Q.contiguous()


def run_steps(disc, Q, dt, nsteps, bufs):
    """nsteps CK2N54 steps with ping-pong state buffers; returns the buffer holding the state."""
    import flouhip as F
    cur, oth, reg = Q, bufs[0], bufs[1]
    for _ in range(nsteps):
        for s in range(5):
            F.rk_stage(disc, F.CarpenterKennedy2N54.code, s, dt, cur, oth, reg=reg)
            cur, oth = oth, cur
    return cur


def time_form(precision, Q64, dt, steps, warmup, plain=False):
    """CK2N54 steps of the C3 workload in one precision; plain: the reference-order kernel
    even where the fast kernel applies (option 1 of flou_hip_set_option)."""
    import torch
    disc, eq, std = build_disc(precision)
    if plain:
        disc.set_option(1, 1)
    Q = Q64.to(disc.torch_dtype(), copy=True).contiguous()
    bufs = [disc.new_state(), disc.new_state()]
    cur = run_steps(disc, Q, dt, warmup, bufs)
    if cur is not Q:
        Q.copy_(cur)
    torch.cuda.synchronize()
    stream = torch.cuda.current_stream()
    t0, t1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    tic = time.perf_counter()
    t0.record(stream)
    cur = run_steps(disc, Q, dt, steps, bufs)
    t1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - tic
    disc.check()
    gpu_ms = t0.elapsed_time(t1)
    rho = cur[0].to(torch.float64).clone()
    ndofs = disc.ndofs
    fast = disc.uses_fast_kernel()
    del bufs, Q, cur, disc
    torch.cuda.empty_cache()
    return dict(wall_s=wall, gpu_ms=gpu_ms, ndofs=ndofs, rho=rho, fast=fast)


TRAFFIC_FILE = os.path.join(ROOT, "profiles", "k_fast_3_6_chandrasekhar_64cube.md")


def measured_traffic():
    """HBM bytes per launch of the fast FP32 kernel on this workload, from the committed
    rocprofv3 FETCH_SIZE / WRITE_SIZE summary (line 'traffic_bytes_per_launch: X')."""
    try:
        for line in open(TRAFFIC_FILE):
            if line.startswith("traffic_bytes_per_launch:"):
                return float(line.split(":", 1)[1])
    except OSError:
        pass
    return None


def adjoint_ratio(nel=128, n=6, nsteps=200, checkpoint_every=20, seed=(101, 202)):
    """Adjoint workload: 2-D Euler on a periodic 128^2 quad mesh over [0,2]^2, N=5, FP64,
    split Chandrasekhar + LxF(StdAverage), SSPRK33 at a fixed dt (CFL 0.15 of state A; 0.3
    is beyond the SSPRK33 stability limit of this N=5 split form).
    Loss ||ρ_A(T) - ρ_B(T)||²_{L2,Jω} at T = 200 steps; gradient w.r.t. the initial state
    by the checkpointed reverse sweep (checkpoints every 20 steps, the last segment's stage
    inputs kept from the forward run, recompute of the others), and with every stage input
    kept (no recompute).  Returns gradient time (checkpointed forward + reverse sweep) /
    forward time over the same steps, both ways."""
    import torch
    import flouhip as F
    from flouhip import adjoint as A
    b = F.LagrangeBasis("GLL", n)
    std = F.StdQuad(b, F.DGSEMrec(b), 4)
    mesh = F.CartesianMesh(2, (0, 0), (2.0, 2.0), (nel, nel))
    mesh.apply_periodicBCs(("1", "2"), ("3", "4"))
    eq = F.EulerEquation(2, GAMMA)
    op = F.SplitDivOperator(F.ChandrasekharAverage(), F.LxF(F.StdAverage(), 1.0))
    disc = F.MultielementDisc(mesh, std, eq, op, precision="fp64")
    xi = torch.tensor(std.basis.ξ, dtype=torch.float64, device="cuda")
    e = torch.arange(disc.nelements, device="cuda")
    a = torch.arange(n * n, device="cuda")
    dx = 2.0 / nel
    x = ((e % nel)[:, None] + (xi[a % n][None, :] + 1) / 2).reshape(-1) * dx
    y = ((e // nel)[:, None] + (xi[a // n][None, :] + 1) / 2).reshape(-1) * dx

    def state(sd):
        g = torch.Generator(device="cpu")
        g.manual_seed(sd)
        prim = []
        for v in range(4):
            f = torch.zeros_like(x)
            for _ in range(4):
                kx, ky = (torch.randint(1, 3, (2,), generator=g) * math.pi).tolist()
                ph = float(torch.rand(1, generator=g)) * 2 * math.pi
                f += torch.sin(kx * x + ky * y + ph) / 4
            prim.append(f)
        rho, u, v, p = 1 + 0.1 * prim[0], 0.3 * prim[1], 0.3 * prim[2], 1 + 0.1 * prim[3]
        return torch.stack([rho, rho * u, rho * v, p / (GAMMA - 1) + rho * (u * u + v * v) / 2]).contiguous()

    QA, QB = state(seed[0]), state(seed[1])
    dt = F.get_max_dt(QA, disc, eq, 0.15)
    w1 = torch.tensor(std.basis.ω, dtype=torch.float64, device="cuda")
    wnode = ((w1[a % n] * w1[a // n]) * (dx / 2) ** 2).repeat(disc.nelements)
    uB, _ = A.forward(disc, F.SSPRK33, QB, dt, nsteps)
    disc.check()
    rhoB = uB[0].clone()
    stream = torch.cuda.current_stream()
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(5)]
    Q = QA.clone()
    F.timeintegrate(Q.clone(), disc, eq, F.SSPRK33, 5 * dt, dt=dt)  # warm-up
    torch.cuda.synchronize()
    ev[0].record(stream)
    F.timeintegrate(Q, disc, eq, F.SSPRK33, nsteps * dt, dt=dt)
    ev[1].record(stream)
    torch.cuda.synchronize()
    nseg = -(-nsteps // checkpoint_every)

    def grad(keep):
        """The gradient, end to end: checkpointed forward, loss cotangent, reverse sweep;
        returns (forward ms, reverse ms, gradient)"""
        ev[2].record(stream)
        uT, cps = A.forward(disc, F.SSPRK33, QA, dt, nsteps, checkpoint_every, keep_segments=keep)
        ubar = torch.zeros_like(uT)
        ubar[0] = 2.0 * wnode * (uT[0] - rhoB)
        ev[3].record(stream)
        A.reverse(disc, F.SSPRK33, cps, dt, nsteps, ubar, checkpoint_every)
        ev[4].record(stream)
        torch.cuda.synchronize()
        disc.check()
        return ev[2].elapsed_time(ev[3]), ev[3].elapsed_time(ev[4]), ubar

    grad(1)  # warm-up: the allocator's cache then holds the segment buffers
    cfwd_ms, rev_ms, ubar = grad(1)
    grad(nseg)
    afwd_ms, arev_ms, ubar_all = grad(nseg)
    fwd_ms = ev[0].elapsed_time(ev[1])
    gnorm = float(torch.linalg.norm(ubar))
    if not torch.equal(ubar, ubar_all):
        raise RuntimeError("gradient from kept stage inputs differs from the recomputed one")
    return {"value": (cfwd_ms + rev_ms) / fwd_ms, "forward_ms": fwd_ms,
            "checkpointed_forward_ms": cfwd_ms, "reverse_ms": rev_ms,
            "reverse_only_ratio": rev_ms / fwd_ms, "gradient_norm": gnorm,
            "all_stages_kept": {"ratio": (afwd_ms + arev_ms) / fwd_ms, "forward_ms": afwd_ms,
                                "reverse_ms": arev_ms,
                                "device_bytes": nsteps * 3 * 4 * disc.ndofs * 8},
            "workload": "2-D Euler periodic 128^2 quads on [0,2]^2, N=5, FP64, split Chandrasekhar + "
                        "LxF(StdAverage), SSPRK33 x 200 fixed-dt steps (CFL 0.15), L2 density misfit, "
                        "checkpoints every 20 steps, the last segment's stage inputs kept from the "
                        "forward run, the other segments recomputed; value = one gradient "
                        "(checkpointed forward + reverse sweep with recompute and VJP) / one forward "
                        "run; all_stages_kept: every step's stage inputs kept in HBM (no recompute)"}


def _ssprk33_steps(disc, Q, dt, nsteps):
    """nsteps fused SSPRK33 steps in place on Q (two work states)."""
    import flouhip as F
    u1, u2 = disc.new_state(), disc.new_state()
    for _ in range(nsteps):
        F.rk_stage(disc, F.SSPRK33.code, 0, dt, Q, u1)
        F.rk_stage(disc, F.SSPRK33.code, 1, dt, u1, u2, u_base=Q)
        F.rk_stage(disc, F.SSPRK33.code, 2, dt, u2, Q, u_base=Q)
    return Q


SECONDARY = {
    "c2": "2-D Euler, periodic 256^2 quads on [0,2]^2, N=7 (4,194,304 nodes), split KennedyGruber volume "
          "flux + LxF(StdAverage) surface, SSPRK33 fused stages, fixed dt (CFL 0.1)",
    "c5": "3-D Euler, periodic curved 32^3 hex (x + 0.05 sin 2πx sin 2πy sin 2πz on the unit cube), N=6 "
          "(11,239,424 nodes, per-node metric and J), split ChandrasekharAverage volume flux + "
          "LxF(StdAverage) surface, SSPRK33 fused stages, fixed dt (CFL 0.1)",
}


def _c2_prim(X, g):
    """C2's seeded primitive state (ρ, u, v, p) at the node coordinates X (2, nodes):
    rho = 1 + 0.2 sin(π(x+y)) + 8 Fourier modes of 0.01, u, v = 0.5 + 0.05 smooth, p = 1."""
    import torch
    x, y = X[0], X[1]
    rho = 1 + 0.2 * torch.sin(math.pi * (x + y))
    for _ in range(8):
        kx, ky = torch.randint(1, 4, (2,), generator=g).tolist()
        rho = rho + 0.01 / 8 * torch.sin(math.pi * (kx * x + ky * y) + float(torch.rand(1, generator=g)) * 6.283)
    ph = (torch.rand(2, generator=g) * 6.283).tolist()
    u = 0.5 + 0.05 * torch.sin(math.pi * x + ph[0]) * torch.cos(math.pi * y)
    v = 0.5 + 0.05 * torch.cos(math.pi * x) * torch.sin(math.pi * y + ph[1])
    return torch.stack([rho, u, v, torch.ones_like(x)])


def _c2_case():
    """The C2 leg's discretisation parts and a function building (disc, state) per precision:
    periodic 256^2 quads on [0,2]^2, N=7, split Kennedy-Gruber + LxF(StdAverage), and its
    seeded FP64 state (built once)."""
    import torch
    import flouhip as F
    b = F.LagrangeBasis("GLL", 8)
    std = F.StdQuad(b, F.DGSEMrec(b), 4)
    eq = F.EulerEquation(2, GAMMA)
    op = F.SplitDivOperator(F.KennedyGruber(), F.LxF(F.StdAverage(), 1.0))
    state = {}

    def build(form):
        mesh = F.CartesianMesh(2, (0.0, 0.0), (2.0, 2.0), (256, 256))
        mesh.apply_periodicBCs(("1", "2"), ("3", "4"))
        disc = F.MultielementDisc(mesh, std, eq, op, precision=form)
        if "Q64" not in state:
            X = torch.tensor(mesh.node_coords(std), device="cuda").T
            g = torch.Generator(device="cpu")
            g.manual_seed(7)
            prim = _c2_prim(X, g)
            Q64 = torch.empty_like(prim)
            Q64[0] = prim[0]
            Q64[1:3] = prim[0] * prim[1:3]
            Q64[3] = prim[3] / (GAMMA - 1) + prim[0] * (prim[1:3] ** 2).sum(0) / 2
            state["Q64"] = Q64
        return disc, state["Q64"]
    return std, eq, build


def secondary_workload(name, steps=None, warmup=2):
    """C2: 2-D Euler on a periodic 256^2 quad mesh over [0,2]^2, N=7, split Kennedy-Gruber +
    LxF(StdAverage), SSPRK33 (CFL 0.1 of the initial state), seeded state rho = 1 +
    0.2 sin(π(x+y)) + 8 Fourier modes of 0.01, u, v = 0.5 + 0.05 smooth, p = 1.
    C5: 3-D Euler on the curved periodic 32^3 mesh x + 0.05 sin 2πx sin 2πy sin 2πz of the
    unit cube, N=6, split Chandrasekhar + LxF(StdAverage), SSPRK33 (CFL 0.1), seeded smooth
    state rho in [0.9, 1.1], |u| <= 0.3, p in [0.9, 1.1].
    FP32 and FP64 over the same steps from the same FP64 state: DOF-updates/s per stage,
    average stage-kernel time against the HBM roofline of the algorithmic bytes (SSPRK33
    stages move 2, 3 and 3 states of NV values; C5 adds the per-node metric and J, (d²+1)
    values per node, to every stage), and the density L2 error of FP32 against FP64."""
    import numpy as np
    import torch
    import flouhip as F
    if name == "c2":
        nd, n, nel, L, nv = 2, 8, 256, 2.0, 4
        nsteps = steps or 50
        op = F.SplitDivOperator(F.KennedyGruber(), F.LxF(F.StdAverage(), 1.0))
        mapping = None
    else:
        nd, n, nel, L, nv = 3, 7, 32, 1.0, 5
        nsteps = steps or 20
        op = F.SplitDivOperator(F.ChandrasekharAverage(), F.LxF(F.StdAverage(), 1.0))

        def mapping(X):
            s = np.ones((X.shape[0], 1))
            for k in range(3):
                s = s * np.sin(2 * np.pi * X[:, [k]])
            return X + 0.05 * s
    b = F.LagrangeBasis("GLL", n)
    std = [None, None, F.StdQuad, F.StdHex][nd](b, F.DGSEMrec(b), nv)
    eq = F.EulerEquation(nd, GAMMA)
    res, Q64, dt = {}, None, None
    for form in ("fp64", "fp32"):
        mesh = F.CartesianMesh(nd, (0.0,) * nd, (L,) * nd, (nel,) * nd)
        mesh.apply_periodicBCs(*[(str(2 * k + 1), str(2 * k + 2)) for k in range(nd)])
        if mapping is not None:
            mesh.apply_mapping(mapping)
        disc = F.MultielementDisc(mesh, std, eq, op, precision=form)
        if Q64 is None:
            X = torch.tensor(mesh.node_coords(std), device="cuda").T
            g = torch.Generator(device="cpu")
            g.manual_seed(7)
            if name == "c2":
                prim = _c2_prim(X, g)
            else:
                ph = (torch.rand(5, 3, generator=g) * 6.283).tolist()
                modes = [torch.sin(2 * math.pi * X[0] + ph[v][0]) * torch.sin(2 * math.pi * X[1] + ph[v][1])
                         * torch.sin(2 * math.pi * X[2] + ph[v][2]) for v in range(5)]
                prim = torch.stack([1 + 0.1 * modes[0], 0.17 * modes[1], 0.17 * modes[2], 0.17 * modes[3],
                                    1 + 0.1 * modes[4]])
            Q64 = torch.empty_like(prim)
            Q64[0] = prim[0]
            Q64[1:nd + 1] = prim[0] * prim[1:nd + 1]
            Q64[nd + 1] = prim[nd + 1] / (GAMMA - 1) + prim[0] * (prim[1:nd + 1] ** 2).sum(0) / 2
            del X, prim
            dt = F.get_max_dt(Q64, disc, eq, 0.1)
        Q = Q64.to(disc.torch_dtype(), copy=True).contiguous()  # the steps run in place
        _ssprk33_steps(disc, Q.clone(), dt, warmup)
        torch.cuda.synchronize()
        stream = torch.cuda.current_stream()
        t0, t1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        t0.record(stream)
        _ssprk33_steps(disc, Q, dt, nsteps)
        t1.record(stream)
        torch.cuda.synchronize()
        disc.check()
        ms = t0.elapsed_time(t1)
        res[form] = dict(ms=ms, rho=Q[0].to(torch.float64).clone(), ndofs=disc.ndofs,
                         fast=disc.uses_fast_kernel() if form == "fp32" else False)
        del disc, Q
        torch.cuda.empty_cache()
    nn = res["fp32"]["ndofs"]
    stages = 3 * nsteps
    out = {"workload": SECONDARY[name], "nodes": nn, "steps": nsteps, "scheme": "SSPRK33", "dt": dt}
    for form in ("fp32", "fp64"):
        s_bytes = 4 if form == "fp32" else 8
        geo = (nd * nd + 1) * s_bytes if name == "c5" else 0
        bytes_per_stage = ((2 + 3 + 3) / 3 * nv * s_bytes + geo) * nn
        launch_ms = res[form]["ms"] / stages
        achieved = bytes_per_stage / (launch_ms / 1e3) / 1e9
        out[form] = {"value": nn * stages / (res[form]["ms"] / 1e3), "unit": "DOF-updates/s",
                     "avg_stage_ms": launch_ms, "roofline_frac": achieved / 8000.0,
                     "achieved_GBps": achieved, "bytes_per_stage": bytes_per_stage}
    out["fp32"]["fast_kernel"] = res["fp32"]["fast"]
    out["fp32_speedup_vs_fp64"] = res["fp64"]["ms"] / res["fp32"]["ms"]
    d = res["fp32"]["rho"] - res["fp64"]["rho"]
    out["density_rel_l2_err_fp32_vs_fp64"] = float(torch.linalg.norm(d) / torch.linalg.norm(res["fp64"]["rho"]))
    return out


def _c5_case():
    """The C5 leg's discretisation parts and a function building (disc, state) per precision:
    the curved periodic 32^3 mesh x + 0.05 sin 2πx sin 2πy sin 2πz, N=6, split Chandrasekhar
    + LxF(StdAverage), and its seeded smooth FP64 state (built once)."""
    import numpy as np
    import torch
    import flouhip as F
    op = F.SplitDivOperator(F.ChandrasekharAverage(), F.LxF(F.StdAverage(), 1.0))
    b = F.LagrangeBasis("GLL", 7)
    std = F.StdHex(b, F.DGSEMrec(b), 5)
    eq = F.EulerEquation(3, GAMMA)

    def mapping(X):
        sn = np.ones((X.shape[0], 1))
        for k in range(3):
            sn = sn * np.sin(2 * np.pi * X[:, [k]])
        return X + 0.05 * sn
    state = {}

    def build(form):
        mesh = F.CartesianMesh(3, (0.0,) * 3, (1.0,) * 3, (32,) * 3)
        mesh.apply_periodicBCs(("1", "2"), ("3", "4"), ("5", "6"))
        mesh.apply_mapping(mapping)
        disc = F.MultielementDisc(mesh, std, eq, op, precision=form)
        if "Q64" not in state:
            X = torch.tensor(mesh.node_coords(std), device="cuda").T
            g = torch.Generator(device="cpu")
            g.manual_seed(7)
            ph = (torch.rand(5, 3, generator=g) * 6.283).tolist()
            modes = [torch.sin(2 * math.pi * X[0] + ph[v][0]) * torch.sin(2 * math.pi * X[1] + ph[v][1])
                     * torch.sin(2 * math.pi * X[2] + ph[v][2]) for v in range(5)]
            prim = torch.stack([1 + 0.1 * modes[0], 0.17 * modes[1], 0.17 * modes[2], 0.17 * modes[3],
                                1 + 0.1 * modes[4]])
            Q64 = torch.empty_like(prim)
            Q64[0] = prim[0]
            Q64[1:4] = prim[0] * prim[1:4]
            Q64[4] = prim[4] / (GAMMA - 1) + prim[0] * (prim[1:4] ** 2).sum(0) / 2
            state["Q64"] = Q64
        return disc, state["Q64"]
    return std, eq, build


def vjp_kernel(name, reps=20):
    """An FP32 rhs VJP alone (profiling run of the fast adjoint kernel): reps calls of
    flou_hip_rhs_vjp at the workload's state with a seeded cotangent, and reps fast rhs
    evaluations at the same state, timed with HIP events on the handle's stream.
    c5: the C5 mesh and state; c3: the C3 (64^3, N=5, Taylor-Green) mesh and state; c2: the
    C2 (2-D 256^2, N=7, Kennedy-Gruber) mesh and state.
    Algorithmic bytes of one VJP call per node: Q, Y and the accumulated Qbar read
    (3 x NV x 4 B), Qbar written (NV x 4 B), and on the curved mesh the metric and J (40 B)."""
    import torch
    import flouhip as F
    from flouhip import adjoint as A
    if name == "c5":
        std, eq, build = _c5_case()
        disc, Q64 = build("fp32")
    elif name == "c2":
        std, eq, build = _c2_case()
        disc, Q64 = build("fp32")
    else:
        disc, eq, std = build_disc("fp32")
        Q64 = taylor_green_state(disc, std)
    assert disc.uses_fast_vjp(), "this FP32 adjoint should run the fast kernel"
    Q = Q64.to(torch.float32).contiguous()
    g = torch.Generator(device="cpu")
    g.manual_seed(11)
    Y = torch.randn(Q.shape, generator=g).to(device="cuda", dtype=torch.float32)
    Qbar = torch.zeros_like(Q)
    dQ = torch.empty_like(Q)
    cfg = F.EquationConfig(disc, eq)
    for _ in range(3):
        A.rhs_vjp(Qbar, Y, Q, disc)
        F.rhs_(dQ, Q, cfg)
    torch.cuda.synchronize()
    stream = torch.cuda.current_stream()
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(3)]
    ev[0].record(stream)
    for _ in range(reps):
        A.rhs_vjp(Qbar, Y, Q, disc)
    ev[1].record(stream)
    for _ in range(reps):
        F.rhs_(dQ, Q, cfg)
    ev[2].record(stream)
    torch.cuda.synchronize()
    disc.check()
    vms, fms = ev[0].elapsed_time(ev[1]) / reps, ev[1].elapsed_time(ev[2]) / reps
    nv, nn = Q.shape
    vbytes = nn * (4 * nv * 4 + (40 if name == "c5" else 0))
    label = {"c5": "C5 curved 32^3 N=6", "c3": "C3 64^3 N=5 Taylor-Green", "c2": "C2 2-D 256^2 N=7 Kennedy-Gruber"}
    return {"vjp_ms": vms, "rhs_ms": fms, "vjp_over_rhs": vms / fms, "nodes": nn,
            "vjp_algorithmic_bytes": vbytes, "vjp_hbm_fraction": vbytes / (vms * 1e-3) / 8e12,
            "workload": label[name] + " FP32 (the bench's mesh and state), seeded cotangent; fast adjoint kernel vs fast rhs"}


def secondary_adjoint(name, nsteps=100, every=10):
    """Adjoint of a secondary workload (c5: the curved 32^3, N=6 mesh and state of the C5
    leg; c2: the 2-D 256^2, N=7 Kennedy-Gruber mesh and state of the C2 leg), loss
    Σ Jω (ρ(T) - 1)² after 100 SSPRK33 steps (CFL 0.1 of the FP64 state), gradient w.r.t.
    the initial state with FP64 checkpoints every 10 steps (the last segment's stage inputs
    kept; FP32 also with every stage input kept).  FP32 (fast forward and adjoint kernels)
    and FP64; per precision the forward and gradient times (HIP events, after one untimed
    gradient) and their ratio, and the FP32 gradient's relative L2 error against the FP64
    one."""
    import numpy as np
    import torch
    import flouhip as F
    from flouhip import adjoint as A
    std, eq, build = _c5_case() if name == "c5" else _c2_case()
    out, grads, dt = {}, {}, None
    stream = torch.cuda.current_stream()
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(4)]
    for form in ("fp64", "fp32"):
        disc, Q64 = build(form)
        if dt is None:
            dt = F.get_max_dt(Q64, disc, eq, 0.1)
        Q = Q64.to(disc.torch_dtype(), copy=True).contiguous()
        jac = torch.tensor(np.asarray(disc.elem_jac, dtype=np.float64).reshape(-1), device="cuda")
        if jac.numel() == disc.nelements:  # affine elements: one J per element
            jac = jac.repeat_interleave(len(std.ω))
        wq = (jac * torch.tensor(np.tile(std.ω, disc.nelements), device="cuda")).to(disc.torch_dtype())

        def seed(u):
            gr = torch.zeros_like(u)
            gr[0] = 2.0 * wq * (u[0] - 1.0)
            return gr

        def grad(keep=1):
            return A.gradient(disc, F.SSPRK33, Q, dt, nsteps, seed, checkpoint_every=every,
                              checkpoint_dtype=torch.float64, keep_segments=keep)
        grad()  # warm-up (allocator cache, first launches)
        torch.cuda.synchronize()
        ev[0].record(stream)
        A.forward(disc, F.SSPRK33, Q, dt, nsteps)
        ev[1].record(stream)
        gq = grad()
        ev[2].record(stream)
        torch.cuda.synchronize()
        disc.check()
        fwd, gms = ev[0].elapsed_time(ev[1]), ev[1].elapsed_time(ev[2])
        out[form] = {"forward_ms": fwd, "gradient_ms": gms, "gradient_over_forward": gms / fwd,
                     "fast_adjoint_kernel": disc.uses_fast_vjp()}
        if form == "fp32":
            # every stage input kept in HBM (3 x 100 states): no recomputation
            nseg = -(-nsteps // every)
            grad(nseg)
            torch.cuda.synchronize()
            ev[2].record(stream)
            gk = grad(nseg)
            ev[3].record(stream)
            torch.cuda.synchronize()
            disc.check()
            gka = ev[2].elapsed_time(ev[3])
            out[form]["all_stages_kept"] = {"gradient_ms": gka, "gradient_over_forward": gka / fwd,
                                            "max_abs_diff_vs_recomputed": float((gk - gq).abs().max())}
            del gk
            torch.cuda.empty_cache()
        grads[form] = gq.to(torch.float64)
        del disc, Q, wq, gq
        torch.cuda.empty_cache()
    out["fp32_gradient_rel_l2_err_vs_fp64"] = float(torch.linalg.norm(grads["fp32"] - grads["fp64"])
                                                  / torch.linalg.norm(grads["fp64"]))
    out["fp32_gradient_speedup_vs_fp64"] = out["fp64"]["gradient_ms"] / out["fp32"]["gradient_ms"]
    out["workload"] = (("C5 curved 32^3 N=6" if name == "c5" else "C2 2-D 256^2 N=7 Kennedy-Gruber")
                       + f" (the {name} leg's mesh and state), loss sum Jw (rho(T) - 1)^2 after "
                       "100 SSPRK33 steps, gradient with FP64 checkpoints every 10 steps (last segment kept; "
                       "FP32 also with every stage input kept); FP32: fast forward, recomputation and adjoint "
                       "kernels, FP64: plain kernels")
    return out


def c1_steps(nsteps=2000):
    """C1: 2-D linear advection (1, 0.5) on 16x16 periodic quads over [0,1]^2, N=4 GLL,
    strong form, upwind LxF, FP64, Gaussian pulse + seeded U(-1e-3, 1e-3) noise, SSPRK33 at
    a fixed dt (CFL 0.3 of the initial state): one flou_hip_timeintegrate call over nsteps
    steps (every stage launched from the host), HIP-event timing after a warm-up call.
    6400 nodes: the stage kernels are launch-bound."""
    import ctypes
    import numpy as np
    import torch
    import flouhip as F
    from flouhip import _capi as C
    b = F.LagrangeBasis("GLL", 5)
    std = F.StdQuad(b, F.DGSEMrec(b), 1)
    mesh = F.CartesianMesh(2, (0, 0), (1, 1), (16, 16)).apply_periodicBCs(("1", "2"), ("3", "4"))
    eq = F.LinearAdvection(1.0, 0.5)
    disc = F.MultielementDisc(mesh, std, eq, F.StrongDivOperator(F.LxF(F.StdAverage(), 1.0)))
    X = mesh.node_coords(std)
    rng = np.random.default_rng(1)
    Q0 = np.asfortranarray((np.exp(-((X[:, 0] - 0.5) ** 2 + (X[:, 1] - 0.5) ** 2) / 0.01)
                            + rng.uniform(-1e-3, 1e-3, X.shape[0]))[:, None])
    dt = F.get_max_dt(disc.to_device(Q0), disc, eq, 0.3)
    code = F.disc._solver_code(F.SSPRK33)
    stream = torch.cuda.current_stream()
    Q = disc.to_device(Q0)
    disc._sync_stream()
    t = ctypes.c_double(0.0)
    C.check(C.lib.flou_hip_timeintegrate(disc.h, code, disc._dptr(Q), dt, 50, 0.0, ctypes.byref(t)), disc.h)
    Q = disc.to_device(Q0)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    C.check(C.lib.flou_hip_timeintegrate(disc.h, code, disc._dptr(Q), dt, nsteps, 0.0, ctypes.byref(t)), disc.h)
    e1.record(stream)
    torch.cuda.synchronize()
    disc.check()
    ms = e0.elapsed_time(e1)
    return {"us_per_step": 1e3 * ms / nsteps, "us_per_stage": 1e3 * ms / (3 * nsteps),
            "DOF_updates_per_s": disc.ndofs * 3 * nsteps / (ms * 1e-3),
            "workload": (f"C1 2-D advection 16x16 N=4 FP64 strong/upwind, {disc.ndofs} nodes, {nsteps} SSPRK33 "
                         "steps at a fixed dt through flou_hip_timeintegrate")}


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(nel=32, steps=6):
    """Oracle (FP64 C++ restatement of Flou's rhs!, OpenMP) on a 32^3 sample of the same
    workload (about 10 s of CPU work), CK2N54 steps, DOF-updates/s per stage."""
    import numpy as np
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as O
    d = O.OracleDisc(3, N_NODES, [nel] * 3, [0.0] * 3, [2 * math.pi] * 3, [1, 1, 1],
                     O.build_ops(N_NODES, "GLL"), O.EQ_EULER, [GAMMA], O.VOL_SPLIT,
                     O.TP_CHANDRASEKHAR, O.AVG_STD, O.DISS_LXF, 1.0)
    X, _ = d.coords()
    rng = np.random.default_rng(1234)
    x, y, z = X[:, 0], X[:, 1], X[:, 2]
    prim = np.stack([np.ones_like(x), M0 * np.sin(x) * np.cos(y) * np.cos(z),
                     -M0 * np.cos(x) * np.sin(y) * np.cos(z), np.zeros_like(x),
                     1 / GAMMA + (M0 ** 2 / 16) * (np.cos(2 * x) + np.cos(2 * y)) * (np.cos(2 * z) + 2)], 1)
    prim += 1e-3 * rng.standard_normal(prim.shape)
    Q = np.zeros_like(prim)
    Q[:, 0] = prim[:, 0]
    Q[:, 1:4] = prim[:, :1] * prim[:, 1:4]
    Q[:, 4] = prim[:, 4] / (GAMMA - 1) + prim[:, 0] * (prim[:, 1:4] ** 2).sum(1) / 2
    Q = np.asfortranarray(Q)
    dt = d.max_dt(Q, 0.3)
    tic = time.perf_counter()
    d.rk_run(O.RK_CK2N54, Q, dt, steps)
    t = time.perf_counter() - tic
    cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    return {"value": d.ndofs * 5 * steps / t, "unit": "DOF-updates/s", "cores": cores,
            "kind": "port", "cpu": cpu_model(),
            "sample": f"oracle (FP64 C++ restatement of Flou's rhs!, OpenMP) on {nel}^3 elements, N=5 "
                      f"({d.ndofs} nodes), {steps} CK2N54 steps = {5 * steps} stages, {t:.1f} s"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--only", choices=["fp32", "fp64", "adjoint", "c1", "c2", "c5", "c2adj", "c5adj", "c5vjp",
                                       "c3vjp", "c2vjp"],
                    default=None,
                    help="time one form only, only the adjoint/forward measurement, or only the "
                         "2-D (c2) or curved 3-D (c5) workload (profiling runs)")
    ap.add_argument("--no-secondary", action="store_true", help="skip the C2 and C5 workloads")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the adjoint/forward measurement")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="also run 3 plain copies of 1 GiB (FP32), the reference kernel for "
                         "scaling FETCH_SIZE / WRITE_SIZE counters")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this library runs on one GPU; --gpus {args.gpus} is not supported "
              "(one GPU per run; the element-slab sharding across GPUs is not implemented)", file=sys.stderr)
        sys.exit(2)
    import torch
    import flouhip as F

    torch.cuda.set_device(0)
    if args.only == "adjoint":  # profiling run of the adjoint workload alone
        print(json.dumps({"adjoint": adjoint_ratio()}), flush=True)
        return
    if args.only == "c1":  # the C1 time-step loop alone
        print(json.dumps({"c1": c1_steps()}), flush=True)
        return
    if args.only in ("c2adj", "c5adj"):  # one secondary adjoint alone
        print(json.dumps({args.only[:2] + "_adjoint": secondary_adjoint(args.only[:2])}), flush=True)
        return
    if args.only in ("c5vjp", "c3vjp", "c2vjp"):  # profiling run of an FP32 adjoint kernel alone
        print(json.dumps({args.only[:2] + "_vjp": vjp_kernel(args.only[:2])}), flush=True)
        return
    if args.only in ("c2", "c5"):  # profiling run of one secondary workload alone
        print(json.dumps({args.only: secondary_workload(args.only, args.steps, args.warmup)}), flush=True)
        return
    if args.copy_calibration:
        src = torch.ones(1 << 28, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            torch.mul(src, 1.0, out=dst)  # an elementwise kernel: 1 GiB read + 1 GiB write
        torch.cuda.synchronize()
        del src, dst
    # initial state and dt (FP64)
    disc64, eq, std = build_disc("fp64")
    Q64 = taylor_green_state(disc64, std)
    dt = F.get_max_dt(Q64, disc64, eq, 0.3)
    nn = disc64.ndofs
    del disc64
    torch.cuda.empty_cache()

    res = {}
    forms = ["fp32", "fp64"]
    if args.no_fp64:
        forms = ["fp32"]
    if args.only:
        forms = [args.only]
    for form in forms:
        res[form] = time_form(form, Q64, dt, args.steps, args.warmup)
        print(f"# {form}: {res[form]['gpu_ms'] / args.steps:.3f} ms/step "
              f"({res[form]['gpu_ms'] / (5 * args.steps):.3f} ms/stage)", file=sys.stderr, flush=True)

    main_form = "fp32" if "fp32" in res else forms[0]
    r = res[main_form]
    stages = 5 * args.steps
    ms_step = r["gpu_ms"] / args.steps
    value = nn * stages / (r["gpu_ms"] / 1e3)
    s_bytes = 4 if main_form == "fp32" else 8
    # CK2N54 stage: read u_in, write register, write u_out; stages 2-5 also read the register
    # -> 3 + 4 + 4 + 4 + 4 arrays of nn x 5 values per step
    bytes_per_launch = (3 + 4 * 4) / 5 * 5 * s_bytes * nn
    launch_ms = r["gpu_ms"] / stages
    achieved = bytes_per_launch / (launch_ms / 1e3) / 1e9
    traffic_src = None
    if main_form == "fp32" and r["fast"]:
        kname = "k_fast<3, 6, CHANDRASEKHAR, MODE_2N> (FP32 fused rhs + CK2N54 stage)"
        traffic = measured_traffic()
        if traffic is not None:
            traffic_src = ("rocprofv3 FETCH_SIZE + WRITE_SIZE of this kernel on this workload, "
                           "from " + os.path.relpath(TRAFFIC_FILE, ROOT) + " (not measured in this run)")
    else:
        kname = f"k_plain<{'float' if main_form == 'fp32' else 'double'},3,6,EULER,SPLIT,CHANDRASEKHAR> (fused rhs + CK2N54 stage)"
        traffic = None
    out = {
        "metric": "DOF-updates/s per RK stage, 3D Euler N=5, FP32 vs FP64 GPU; adjoint/forward ratio",
        "value": value,
        "unit": "DOF-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if main_form == "fp32" else "f64",
        "data": "synthetic (seeded Taylor-Green + noise)",
        "config": {"workload": "3-D Euler, periodic 64^3 hex, N=5 (56,623,104 nodes), split "
                               "ChandrasekharAverage volume flux + LxF(StdAverage) surface, "
                               "CarpenterKennedy2N54 fused stages, fixed dt (CFL 0.3)",
                   "form": main_form, "nodes": nn, "stages_per_step": 5, "dt": dt},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": 8000.0, "unit": "GB/s",
                     "frac": achieved / 8000.0, "traffic": traffic, "traffic_source": traffic_src,
                     "kernel": kname, "bytes_per_launch": bytes_per_launch, "avg_launch_ms": launch_ms},
    }
    if "fp64" in res and "fp32" in res:
        import torch as _t
        r64 = res["fp64"]
        out["fp64"] = {"value": nn * stages / (r64["gpu_ms"] / 1e3), "ms_per_step": r64["gpu_ms"] / args.steps}
        out["fp32_speedup_vs_fp64"] = r64["gpu_ms"] / r["gpu_ms"]
        d = res["fp32"]["rho"] - r64["rho"]
        out["density_rel_l2_err_fp32_vs_fp64"] = float(_t.linalg.norm(d) / _t.linalg.norm(r64["rho"]))
        out["density_err_after_steps"] = args.steps + args.warmup
        # the speedup split into its two parts: precision (FP64 -> FP32, both in the
        # reference-order plain kernel) and algorithm (FP32 plain -> FP32 fast kernel)
        rp = time_form("fp32", Q64, dt, args.steps, args.warmup, plain=True)
        print(f"# fp32 plain: {rp['gpu_ms'] / (5 * args.steps):.3f} ms/stage", file=sys.stderr, flush=True)
        out["fp32_plain"] = {"value": nn * stages / (rp["gpu_ms"] / 1e3), "ms_per_step": rp["gpu_ms"] / args.steps,
                             "speedup_fp64_plain_to_fp32_plain": r64["gpu_ms"] / rp["gpu_ms"],
                             "speedup_fp32_plain_to_fp32_fast": rp["gpu_ms"] / r["gpu_ms"]}
    if not args.no_secondary and not args.only:
        out["c1"] = c1 = c1_steps()
        print(f"# c1: {c1['us_per_step']:.1f} us/step ({c1['us_per_stage']:.2f} us/stage)", file=sys.stderr,
              flush=True)
        for name in ("c2", "c5"):
            out[name] = secondary_workload(name)
            print(f"# {name}: fp32 {out[name]['fp32']['avg_stage_ms']:.3f} ms/stage, "
                  f"fp64 {out[name]['fp64']['avg_stage_ms']:.3f} ms/stage", file=sys.stderr, flush=True)
    out["adjoint_forward_ratio"] = None
    if not args.no_adjoint:
        adj = adjoint_ratio()
        out["adjoint_forward_ratio"] = adj["value"]
        out["adjoint"] = adj
        for name in ("c2", "c5") if not args.no_secondary else ():
            out[name + "_adjoint"] = sa = secondary_adjoint(name)
            print(f"# {name} adjoint: fp32 gradient/forward {sa['fp32']['gradient_over_forward']:.2f} "
                  f"(all stages kept {sa['fp32']['all_stages_kept']['gradient_over_forward']:.2f}), "
                  f"fp64 {sa['fp64']['gradient_over_forward']:.2f}, fp32 error "
                  f"{sa['fp32_gradient_rel_l2_err_vs_fp64']:.2e}", file=sys.stderr, flush=True)
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
