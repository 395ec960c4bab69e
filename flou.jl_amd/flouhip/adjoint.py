"""Adjoint of rhs! and of the RK time loop, with checkpointed recomputation.

Flou has no AD rules; its Julia binding puts a ChainRulesCore ``rrule`` around ``rhs!``
and ``timeintegrate``.  Here the same rules are ``torch.autograd.Function``s and
plain functions over the C-ABI (include/flouhip.h):

* ``rhs_vjp``        -- Qbar += (d rhs/dQ)^T dQbar            (flou_hip_rhs_vjp)
* ``rk_step_record`` -- one RK step keeping its stage inputs   (flou_hip_rk_step_record)
* ``rk_step_vjp``    -- cotangent of u_n from that of u_{n+1}  (flou_hip_rk_step_vjp)
* ``rk_step_vjp_noalloc`` -- the same without device allocation (flou_hip_rk_step_vjp_noalloc)
* ``gradient``       -- reverse sweep over nsteps fixed-dt steps: checkpoints every
  ``checkpoint_every`` steps during the forward run, which also keeps the stage inputs
  of the last ``keep_segments`` segments; then per segment (unless kept) one forward
  recomputation that keeps the stage inputs, followed by the reverse steps.

dt is a fixed schedule (no gradient through the CFL min).  All arithmetic runs in
libflouhip; torch supplies device buffers.
"""
import ctypes

import torch

from . import _capi as C
from .disc import _solver


def _solver_code(solver):
    solver = _solver(solver)
    if getattr(solver, "stage_limiter", None) is not None:
        raise NotImplementedError("adjoint: the stage limiter is not differentiated")
    return solver.code


def _nstages(code):
    return 3 if code == C.SSPRK33 else 5


def rhs_vjp(Qbar, dQbar, Q, disc):
    """Qbar += (d rhs!(·)/dQ at Q)^T dQbar (in place on Qbar)."""
    disc._sync_stream()
    C.check(C.lib.flou_hip_rhs_vjp(disc.h, disc._dptr(Q), disc._dptr(dQbar), disc._dptr(Qbar)),
            disc.h, "rhs_vjp")
    return Qbar


def _ptr_array(disc, bufs):
    arr = (ctypes.c_void_p * len(bufs))()
    for i, b in enumerate(bufs):
        arr[i] = disc._dptr(b).value
    return arr


def rk_step_record(disc, solver, dt, stages, u_next, t=0.0):
    """One RK step from stages[0] = u_n; stages[1:] receive the stage inputs, u_next u_{n+1}."""
    code = _solver_code(solver)
    if len(stages) != _nstages(code):
        raise ValueError(f"{_nstages(code)} stage buffers expected")
    disc._sync_stream()
    C.check(C.lib.flou_hip_rk_step_record(disc.h, code, float(dt), float(t), _ptr_array(disc, stages),
                                          disc._dptr(u_next)), disc.h, "rk_step_record")


def rk_step_vjp(disc, solver, dt, stages, ubar, t=0.0):
    """In place: ubar (cotangent of u_{n+1}) -> cotangent of u_n, at the recorded stages."""
    code = _solver_code(solver)
    if len(stages) != _nstages(code):
        raise ValueError(f"{_nstages(code)} stage buffers expected")
    disc._sync_stream()
    C.check(C.lib.flou_hip_rk_step_vjp(disc.h, code, float(dt), float(t), _ptr_array(disc, stages),
                                       disc._dptr(ubar)), disc.h, "rk_step_vjp")
    return ubar


def rk_step_vjp_noalloc(disc, solver, dt, stages, ubar, t=0.0):
    """rk_step_vjp with its three work states in the library's static per-GPU scratch
    buffer (flou_hip_rk_step_vjp_noalloc): no device allocation, so the call can be
    captured in a HIP graph.  Concurrent calls on one GPU race."""
    code = _solver_code(solver)
    if len(stages) != _nstages(code):
        raise ValueError(f"{_nstages(code)} stage buffers expected")
    disc._sync_stream()
    C.check(C.lib.flou_hip_rk_step_vjp_noalloc(disc.h, code, float(dt), float(t), _ptr_array(disc, stages),
                                               disc._dptr(ubar)), disc.h, "rk_step_vjp_noalloc")
    return ubar


class Checkpoints(list):
    """States at the segment starts (segment i starts at step i*every), plus in ``stages``
    the stage inputs of the trailing segments the forward run kept resident: stages[i] is
    one (steps, stages, nvars, ndofs) tensor, [k, s] = stage input s of step i*every + k
    (s = 0: the state u_n)."""

    def __init__(self, every):
        super().__init__()
        self.every = every
        self.stages = {}


def _step_ptrs(buf, k):
    """ctypes array of the stage buffers of step k of a kept segment (no tensor views: the
    loops below launch one step per call, and creating hundreds of view objects per
    sweep stalls the host long enough to starve the GPU)."""
    S = buf.shape[1]
    sb = buf.shape[2] * buf.shape[3] * buf.element_size()
    base = buf.data_ptr() + k * S * sb
    return (ctypes.c_void_p * S)(*[base + s * sb for s in range(S)])


# share of the device memory free at the start of a gradient that keep_segments="auto" may
# fill with kept stage inputs
AUTO_KEEP_FRACTION = 0.5


def auto_keep_segments(disc, solver, nsteps, checkpoint_every, fraction=AUTO_KEEP_FRACTION):
    """The number of trailing segments whose stage inputs fit in ``fraction`` of the free
    device memory (each segment: checkpoint_every x stages states), at least 1."""
    K = max(1, int(checkpoint_every or nsteps or 1))
    nseg = -(-nsteps // K)
    seg_bytes = K * _nstages(_solver_code(solver)) * disc.nvars * disc.ndofs * disc.torch_dtype().itemsize
    free, _ = torch.cuda.mem_get_info()
    return max(1, min(nseg, int(fraction * free // seg_bytes)))


def forward(disc, solver, Q0, dt, nsteps, checkpoint_every=None, checkpoint_dtype=None, keep_segments=0):
    """Fixed-dt forward run from Q0 (not modified).  Returns (u_T, checkpoints) where
    checkpoints[i] is the state at step i*checkpoint_every, stored as checkpoint_dtype
    (default: the working precision).  The stage inputs of the last ``keep_segments``
    segments stay in device memory (checkpoints.stages; one allocation of
    checkpoint_every x stages states per segment), so that the reverse sweep reverses those
    segments without recomputing them; "auto" keeps as many as fit in half of the free
    device memory (auto_keep_segments), which on a 288 GB device is usually all of them."""
    code = _solver_code(solver)
    K = max(1, int(checkpoint_every or nsteps or 1))
    cdt = checkpoint_dtype or disc.torch_dtype()
    S = _nstages(code)
    nseg = -(-nsteps // K)
    if keep_segments == "auto":
        keep_segments = auto_keep_segments(disc, solver, nsteps, K)
    first_kept = max(0, nseg - int(keep_segments))
    cur = [Q0.clone()] + [disc.new_state() for _ in range(S - 1)]
    nxt = disc.new_state()
    cps = Checkpoints(K)
    for sg in range(nseg):
        nloc = min(K, nsteps - sg * K)
        cps.append(cur[0].to(cdt, copy=True))
        if sg < first_kept:
            for _ in range(nloc):
                rk_step_record(disc, solver, dt, cur, nxt)
                cur[0], nxt = nxt, cur[0]
            continue
        buf = torch.empty((nloc, S, disc.nvars, disc.ndofs), dtype=disc.torch_dtype(), device=Q0.device)
        buf[0, 0].copy_(cur[0])
        disc._sync_stream()
        for i in range(nloc):
            dst = _step_ptrs(buf, i + 1)[0] if i + 1 < nloc else disc._dptr(nxt).value
            C.check(C.lib.flou_hip_rk_step_record(disc.h, code, float(dt), 0.0, _step_ptrs(buf, i),
                                                  ctypes.c_void_p(dst)), disc.h, "rk_step_record")
        cur[0], nxt = nxt, cur[0]
        cps.stages[sg] = buf
    return cur[0], cps


def reverse(disc, solver, checkpoints, dt, nsteps, ubar, checkpoint_every):
    """Reverse sweep: ubar (cotangent of u_T, modified in place) -> cotangent of u_0, from
    the checkpoints of ``forward``.  A segment whose stage inputs the forward run kept is
    reversed directly; any other gets one forward recomputation that keeps the stage
    inputs, then its reverse steps."""
    code = _solver_code(solver)
    S = _nstages(code)
    K = max(1, int(checkpoint_every))
    kept = getattr(checkpoints, "stages", {})
    seg_bufs = nxt = None
    ub = disc._dptr(ubar)
    for sgi in reversed(range(len(checkpoints))):
        if sgi in kept:
            buf = kept[sgi]
            if buf.shape[1] != S or buf.shape[2:] != (disc.nvars, disc.ndofs) or buf.dtype != disc.torch_dtype():
                raise ValueError("kept stage inputs do not match this scheme and discretisation")
            disc._sync_stream()
            for i in reversed(range(buf.shape[0])):
                C.check(C.lib.flou_hip_rk_step_vjp(disc.h, code, float(dt), 0.0, _step_ptrs(buf, i), ub),
                        disc.h, "rk_step_vjp")
            continue
        if seg_bufs is None:
            seg_bufs = [[disc.new_state() for _ in range(S)] for _ in range(min(K, nsteps))]
            nxt = disc.new_state()
        n0 = sgi * K
        nloc = min(K, nsteps - n0)
        seg_bufs[0][0].copy_(checkpoints[sgi])
        for i in range(nloc):
            dst = seg_bufs[i + 1][0] if i + 1 < nloc else nxt
            rk_step_record(disc, solver, dt, seg_bufs[i], dst)
        for i in reversed(range(nloc)):
            rk_step_vjp(disc, solver, dt, seg_bufs[i], ubar)
    return ubar


def gradient(disc, solver, Q0, dt, nsteps, seed, checkpoint_every=20, checkpoint_dtype=None,
             return_state=False, keep_segments=1):
    """Cotangent of the initial state for a loss of the final state.

    seed(u_T) returns dL/du_T (a device state tensor).  Memory: nsteps/checkpoint_every
    checkpoints, the stage inputs of the last keep_segments segments (checkpoint_every x
    stages states each, kept from the forward run, so those segments are not recomputed;
    "auto": as many as fit in half of the free device memory), and checkpoint_every x
    stages states for a segment being recomputed."""
    K = max(1, int(checkpoint_every))
    uT, cps = forward(disc, solver, Q0, dt, nsteps, K, checkpoint_dtype, keep_segments)
    ubar = seed(uT).to(disc.torch_dtype()).contiguous().clone()
    reverse(disc, solver, cps, dt, nsteps, ubar, K)
    return (ubar, uT) if return_state else ubar


class Rhs(torch.autograd.Function):
    """dQ = rhs!(Q) with the device adjoint as its backward (the rrule of rhs!)."""

    @staticmethod
    def forward(ctx, Q, disc):
        dQ = torch.empty_like(Q)
        disc._sync_stream()
        C.check(C.lib.flou_hip_rhs(disc.h, disc._dptr(dQ), disc._dptr(Q), 0.0), disc.h, "rhs!")
        ctx.save_for_backward(Q)
        ctx.disc = disc
        return dQ

    @staticmethod
    def backward(ctx, gdQ):
        (Q,) = ctx.saved_tensors
        Qbar = torch.zeros_like(Q)
        rhs_vjp(Qbar, gdQ.contiguous(), Q, ctx.disc)
        return Qbar, None


class TimeIntegrate(torch.autograd.Function):
    """u_T = timeintegrate(Q0) over nsteps fixed steps, backward by checkpointed reverse
    sweep (the rrule of timeintegrate)."""

    @staticmethod
    def forward(ctx, Q0, disc, solver, dt, nsteps, checkpoint_every=20, keep_segments=1):
        uT, cps = forward(disc, solver, Q0, dt, nsteps, checkpoint_every, keep_segments=keep_segments)
        ctx.disc, ctx.solver, ctx.dt, ctx.nsteps, ctx.K = disc, solver, dt, nsteps, checkpoint_every
        ctx.cps = cps
        return uT

    @staticmethod
    def backward(ctx, guT):
        disc = ctx.disc
        ubar = guT.to(disc.torch_dtype()).contiguous().clone()
        reverse(disc, ctx.solver, ctx.cps, ctx.dt, ctx.nsteps, ubar, ctx.K)
        return ubar, None, None, None, None, None, None
