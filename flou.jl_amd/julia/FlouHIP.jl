# FlouHIP.jl -- Julia binding of libflouhip.so (include/flouhip.h): Flou's DGSEM rhs!, its
# explicit RK stages, get_max_dt, integrate, the Zhang & Shu stage limiter, the monitors,
# the equispaced output projection, and the adjoints of rhs! and of the RK step, on one
# MI355X (gfx950) GPU.
#
# Nothing in Flou changes.  The reference's methods are typed on `::Matrix`
# (src/FlouSpatial/Equations/Hyperbolic.jl:31-36, MultielementDiscontinuous.jl:148-178), so
# the methods below, typed on the device state `HIPState`, are new methods that Julia's
# dispatch selects for device states only; Flou's CPU path is untouched.
#
#     using Flou, FlouHIP
#     disc = MultielementDisc(mesh, std, eq, op, bcs)          # Flou's own constructor
#     hd   = HIPDisc(disc, eq; precision=Float32)              # flattened once, uploaded
#     Q    = device(hd, Q0)                                    # (ndofs, nvars) on the GPU
#     sol, t = timeintegrate(Q, disc, eq, CarpenterKennedy2N54(williamson_condition=false),
#                            tf; dt=1e-3)                      # fused HIP stages
#     host(sol.u[end])
#
# The library is found at $FLOUHIP_LIB, else next to this directory
# (flou.jl_amd/libflouhip.so, built by flou.jl_amd/build_lib.py).
module FlouHIP

using LinearAlgebra
using StaticArrays
using ChainRulesCore
using OrdinaryDiffEq: OrdinaryDiffEq, SSPRK33, CarpenterKennedy2N54, ORK256
using Flou
using Flou: FlouCommon, FlouSpatial, FlouTime, FlouBiz

export HIPDisc, HIPState, KennedyGruber
export device, host, monitor, rhs, integrate_steps, gpu_limiter_callback, uses_fast_kernel, uses_fast_vjp

const libflouhip = get(ENV, "FLOUHIP_LIB", normpath(joinpath(@__DIR__, "..", "libflouhip.so")))
const libhip = "libamdhip64"

# ----------------------------------------------------------------------------------------
# Constants of include/flouhip.h
const OK, EINVAL, EUNSUPPORTED, EDEVICE, ENOMEM, EDOMAIN = 0, -1, -2, -3, -4, -5
const FP64, FP32 = Int32(0), Int32(1)
const ADVECTION, EULER = Int32(0), Int32(1)
const STRONG, SPLIT, HYBRID = Int32(0), Int32(1), Int32(2)
const TP_STDAVG, TP_CHANDRASEKHAR, TP_KENNEDY_GRUBER = Int32(0), Int32(1), Int32(2)
const AVG_STD, AVG_CHANDRASEKHAR = Int32(0), Int32(1)
const DISS_NONE, DISS_LXF, DISS_MATRIX, DISS_SCALAR = Int32(0), Int32(1), Int32(2), Int32(3)
const BC_DIRICHLET, BC_INFLOW, BC_OUTFLOW, BC_SLIP = Int32(0), Int32(1), Int32(2), Int32(3)
const SCHEME_SSPRK33, SCHEME_CK2N54, SCHEME_ORK256 = Int32(0), Int32(1), Int32(2)
const GEOM_AFFINE, GEOM_CURVED = Int32(0), Int32(1)
const LIMITER_NONE, LIMITER_ZHANG_SHU = Int32(0), Int32(1)
const MONITOR_KINETIC_ENERGY, MONITOR_ENTROPY = Int32(0), Int32(1)

# flou_hip_desc, field for field (same order, same C types: Julia lays out an isbits
# struct with C's alignment rules)
struct Desc
    nd::Int32
    n::Int32
    nvars::Int32
    precision::Int32
    nelements::Int64
    nfaces::Int64
    nboundaries::Int32
    equation::Int32
    gamma::Float64
    velocity::NTuple{3,Float64}
    volume_op::Int32
    tp_flux::Int32
    surface_avg::Int32
    surface_diss::Int32
    intensity::Float64
    D::Ptr{Float64}
    Ds::Ptr{Float64}
    Dsharp::Ptr{Float64}
    l_minus::Ptr{Float64}
    l_plus::Ptr{Float64}
    dg_minus::Ptr{Float64}
    dg_plus::Ptr{Float64}
    omega::Ptr{Float64}
    elem_faceinds::Ptr{Int64}
    elem_facepos::Ptr{Int8}
    face_eleminds::Ptr{Int64}
    face_elempos::Ptr{Int8}
    face_orientation::Ptr{UInt8}
    face_boundary::Ptr{Int32}
    geometry::Int32
    elem_jac::Ptr{Float64}
    elem_metric::Ptr{Float64}
    elem_volume::Ptr{Float64}
    face_frames::Ptr{Float64}
    face_jac::Ptr{Float64}
    blend::Float64
end

# ----------------------------------------------------------------------------------------
# Errors: the library's status codes mapped to the exceptions Flou raises
# (ArgumentError from the constructors, DomainError from sqrt/log of a negative number,
# FlouTime.jl:39-51 "Simulation crashed!").
last_error(h::Ptr{Cvoid}) =
    unsafe_string(ccall((:flou_hip_last_error, libflouhip), Cstring, (Ptr{Cvoid},), h))

function check(h::Ptr{Cvoid}, rc::Integer, what::AbstractString="libflouhip")
    rc == OK && return nothing
    msg = what * ": " * last_error(h)
    rc == EDOMAIN && throw(DomainError(msg))
    rc == EINVAL && throw(ArgumentError(msg))
    rc == EUNSUPPORTED && throw(ArgumentError("not available on the GPU path: " * msg))
    rc == ENOMEM && throw(OutOfMemoryError())
    error(msg)
end

hipcheck(rc::Integer, what) = rc == 0 || error("HIP error $(rc) in $(what)")

# ----------------------------------------------------------------------------------------
# The Kennedy & Gruber two-point flux (JCP 227, 2008), absent from Flou v0.1.0: the type
# that selects it in SplitDivOperator / HybridDivOperator, and its CPU method so that
# Flou's own path can run it too (same operation order as the library's FP64 kernels).
struct KennedyGruber <: FlouSpatial.AbstractNumericalFlux end

function FlouSpatial.twopointflux(Q1, Q2, Ja1, Ja2, eq::EulerEquation{ND}, ::KennedyGruber) where {ND}
    P1 = FlouCommon.vars_cons2prim(Q1, eq)
    P2 = FlouCommon.vars_cons2prim(Q2, eq)
    ρ = (P1[1] + P2[1]) / 2
    u = SVector(ntuple(k -> (P1[1 + k] + P2[1 + k]) / 2, ND))
    p = (P1[ND + 2] + P2[ND + 2]) / 2
    e = (Q1[ND + 2] / Q1[1] + Q2[ND + 2] / Q2[1]) / 2
    n = SVector(ntuple(k -> (Ja1[k] + Ja2[k]) / 2, ND))
    # f[v][k]: component v of the flux along physical direction k, then F_v = Σ_k f[v][k] n_k
    f = ntuple(k -> begin
        m = ρ * u[k]
        SVector(m, ntuple(c -> c == k ? m * u[c] + p : m * u[c], ND)..., m * e + p * u[k])
    end, ND)
    return SVector(ntuple(v -> begin
        s = f[1][v] * n[1]
        for k in 2:ND
            s = s + f[k][v] * n[k]
        end
        s
    end, ND + 2))
end

# ----------------------------------------------------------------------------------------
# Flattening of Flou's types into the descriptor

# 1-D standard segment of a tensor-product element (its ω is the 1-D quadrature)
segment(std::StdSegment) = std
segment(std::StdQuad) = std.face
segment(std::StdHex) = std.edge

rowmajor(A::AbstractMatrix) = Vector{Float64}(vec(permutedims(A)))

equation_codes(eq::EulerEquation{ND}) where {ND} = (EULER, Float64(eq.γ), (0.0, 0.0, 0.0))
function equation_codes(eq::LinearAdvection{ND}) where {ND}
    a = ntuple(k -> k <= ND ? Float64(eq.a[k]) : 0.0, 3)
    return (ADVECTION, 1.4, a)
end
equation_codes(eq) = throw(ArgumentError("the GPU path has the Euler and linear advection equations, not $(typeof(eq))"))

tp_code(::StdAverage) = TP_STDAVG
tp_code(::ChandrasekharAverage) = TP_CHANDRASEKHAR
tp_code(::KennedyGruber) = TP_KENNEDY_GRUBER
tp_code(f) = throw(ArgumentError("two-point flux $(typeof(f)) is not available on the GPU path"))

avg_code(::StdAverage) = AVG_STD
avg_code(::ChandrasekharAverage) = AVG_CHANDRASEKHAR
avg_code(f) = throw(ArgumentError("surface average $(typeof(f)) is not available on the GPU path"))

# (average, dissipation, intensity) of a numerical flux (include/flouhip.h)
surface_codes(f::StdAverage) = (AVG_STD, DISS_NONE, 1.0)
surface_codes(f::ChandrasekharAverage) = (AVG_CHANDRASEKHAR, DISS_NONE, 1.0)
surface_codes(f::LxF) = (avg_code(f.avg), DISS_LXF, Float64(f.intensity))
surface_codes(f::MatrixDissipation) = (avg_code(f.avg), DISS_MATRIX, Float64(f.intensity))
surface_codes(f::ScalarDissipation) = (avg_code(f.avg), DISS_SCALAR, Float64(f.intensity))
surface_codes(f) = throw(ArgumentError("numerical flux $(typeof(f)) is not available on the GPU path"))

# (volume operator, two-point flux, numerical flux, blend)
operator_codes(op::StrongDivOperator) = (STRONG, TP_STDAVG, op.numflux, 1.0)
operator_codes(op::SplitDivOperator) = (SPLIT, tp_code(op.tpflux), op.numflux, 1.0)
function operator_codes(op::HybridDivOperator)
    op.fvflux === op.numflux || throw(ArgumentError(
        "HybridDivOperator: the GPU path uses the surface flux as the finite-volume flux"))
    return (HYBRID, tp_code(op.tpflux), op.numflux, Float64(op.blend))
end
operator_codes(op) = throw(ArgumentError("operator $(typeof(op)) is not available on the GPU path"))

# Flou's default source is the no-op closure of MultielementDiscontinuous.jl:75-79, the only
# closure defined in FlouSpatial that lands in `source!`; a user source (any other
# function) runs arbitrary Julia per node and has no device form.
has_default_source(disc) = parentmodule(typeof(disc.source!)) === FlouSpatial

frame_values(fr, nd) = vcat(Vector{Float64}(fr.n), Vector{Float64}(fr.t), Vector{Float64}(fr.b))

"""
    HIPDisc(disc::MultielementDisc, equation; precision=Float64)

The GPU twin of a Flou `MultielementDisc`: Flou's own operators (FP64), connectivity,
geometry and boundary conditions flattened into a `flou_hip_desc` and uploaded once
(`flou_hip_create`).  `precision=Float32` selects the FP32 chosen form (FP32 state and
bulk arithmetic, FP64 dt and integrals).
"""
mutable struct HIPDisc
    ptr::Ptr{Cvoid}
    disc::Any
    equation::Any
    T::DataType
    nd::Int
    nvars::Int
    ndofs::Int
end

function HIPDisc(disc::MultielementDisc, equation; precision::Type=Float64)
    precision in (Float64, Float32) || throw(ArgumentError("precision must be Float64 or Float32"))
    has_default_source(disc) || throw(ArgumentError(
        "source terms run arbitrary Julia per node (apply_sourceterm!, " *
        "MultielementDiscontinuous.jl:139-146) and are not available on the GPU path"))
    (; mesh, std, geometry) = disc
    nd = spatialdim(disc)
    NF = 2nd
    nv = nvariables(equation)
    npe = length(std.ξ)
    npf = nd == 1 ? 1 : length(std.face.ξ)
    ne, nf = nelements(mesh), nfaces(mesh)
    volop, tp, numflux, blend = operator_codes(disc.operators[1])
    savg, sdiss, intensity = surface_codes(numflux)
    eqc, γ, a = equation_codes(equation)

    # operators of Flou's own standard element (StdSegment.jl:76-89)
    seg = segment(std)
    D, Ds, Dsh = rowmajor(std.D), rowmajor(std.Ds), rowmajor(std.D♯)
    lm, lp = Vector{Float64}(std.l[1]), Vector{Float64}(std.l[2])
    gm, gp = Vector{Float64}(std.∂g[1]), Vector{Float64}(std.∂g[2])
    ω = Vector{Float64}(seg.ω)

    # connectivity, 0-based (Mesh.jl:19-113)
    efi = Vector{Int64}(undef, ne * NF)
    efp = Vector{Int8}(undef, ne * NF)
    for ie in 1:ne
        el = mesh.elements[ie]
        for k in 1:NF
            efi[(ie - 1) * NF + k] = el.faceinds[k] - 1
            efp[(ie - 1) * NF + k] = el.facepos[k] - 1
        end
    end
    fei = Vector{Int64}(undef, 2nf)
    fep = Vector{Int8}(undef, 2nf)
    fori = Vector{UInt8}(undef, nf)
    for f in 1:nf
        fc = mesh.faces[f]
        for s in 1:2
            e = s <= length(fc.eleminds) ? fc.eleminds[s] : 0
            p = s <= length(fc.elempos) ? fc.elempos[s] : 0
            fei[2(f - 1) + s] = e - 1
            fep[2(f - 1) + s] = p - 1
        end
        fori[f] = UInt8(fc.orientation)
    end
    fbd = fill(Int32(-1), nf)
    for ib in eachboundary(mesh), f in eachbdface(mesh, ib)
        fbd[f] = Int32(ib - 1)
    end

    # geometry (PhysicalRegions.jl:306-364, 485-535): per element and per face when every
    # node of an element (face) carries the same metric and J (frame and surface J), as on
    # Flou's CartesianMesh; per node otherwise (trilinear maps)
    ev, fv = geometry.elements, geometry.faces
    elem_uniform = all(1:ne) do ie
        o = (ie - 1) * npe
        all(i -> ev.metric[o + i] == ev.metric[o + 1] && ev.jac[o + i] == ev.jac[o + 1], 1:npe)
    end
    face_uniform = all(1:nf) do f
        o = (f - 1) * npf
        all(i -> fv.frames[o + i] == fv.frames[o + 1] && fv.jac[o + i] == fv.jac[o + 1], 1:npf)
    end
    affine = elem_uniform && face_uniform
    if affine
        ejac = Float64[ev.jac[(ie - 1) * npe + 1] for ie in 1:ne]
        emet = Float64[ev.metric[(ie - 1) * npe + 1][c] for ie in 1:ne for c in 1:nd^2]
        ffr = reduce(vcat, (frame_values(fv.frames[(f - 1) * npf + 1], nd) for f in 1:nf))
        fjac = Float64[fv.jac[(f - 1) * npf + 1] for f in 1:nf]
    else
        ejac = Vector{Float64}(ev.jac)
        emet = Float64[m[c] for m in ev.metric for c in 1:nd^2]  # SMatrix: column-major
        ffr = reduce(vcat, (frame_values(fr, nd) for fr in fv.frames))
        fjac = Vector{Float64}(fv.jac)
    end
    evol = Vector{Float64}(ev.volume)

    keep = Any[D, Ds, Dsh, lm, lp, gm, gp, ω, efi, efp, fei, fep, fori, fbd, ejac, emet, evol, ffr, fjac]
    out = Ref{Ptr{Cvoid}}(C_NULL)
    GC.@preserve keep begin
        d = Desc(Int32(nd), Int32(length(seg.ω)), Int32(nv), precision === Float64 ? FP64 : FP32,
                 Int64(ne), Int64(nf), Int32(nboundaries(mesh)),
                 eqc, γ, a,
                 volop, tp, savg, sdiss, intensity,
                 pointer(D), pointer(Ds), pointer(Dsh), pointer(lm), pointer(lp), pointer(gm), pointer(gp),
                 pointer(ω),
                 pointer(efi), pointer(efp), pointer(fei), pointer(fep), pointer(fori), pointer(fbd),
                 affine ? GEOM_AFFINE : GEOM_CURVED,
                 pointer(ejac), pointer(emet), pointer(evol), pointer(ffr), pointer(fjac),
                 blend)
        rc = ccall((:flou_hip_create, libflouhip), Cint, (Ref{Desc}, Ref{Ptr{Cvoid}}), d, out)
    end
    if rc != OK
        msg = unsafe_string(ccall((:flou_hip_create_error, libflouhip), Cstring, ()))
        rc == EUNSUPPORTED && throw(ArgumentError("not available on the GPU path: " * msg))
        throw(ArgumentError(msg))
    end
    hd = HIPDisc(out[], disc, equation, precision, nd, nv, ne * npe)
    finalizer(h -> h.ptr == C_NULL || ccall((:flou_hip_destroy, libflouhip), Cint, (Ptr{Cvoid},), h.ptr), hd)
    set_bcs!(hd, disc, equation, npf)
    return hd
end

# disc.bcs (MultielementDiscontinuous.jl:40-51) -> flou_hip_set_bc
function set_bcs!(hd::HIPDisc, disc, equation, npf)
    (; mesh, geometry) = disc
    nv, nf = hd.nvars, nfaces(mesh)
    for ib in eachboundary(mesh)
        bc = disc.bcs[ib]
        if bc isa EulerInflowBC
            data = Vector{Float64}(bc.Qext)
            rc = ccall((:flou_hip_set_bc, libflouhip), Cint, (Ptr{Cvoid}, Int32, Int32, Ptr{Float64}),
                       hd.ptr, ib - 1, BC_INFLOW, data)
        elseif bc isa EulerOutflowBC || bc isa EulerSlipBC
            rc = ccall((:flou_hip_set_bc, libflouhip), Cint, (Ptr{Cvoid}, Int32, Int32, Ptr{Float64}),
                       hd.ptr, ib - 1, bc isa EulerSlipBC ? BC_SLIP : BC_OUTFLOW, C_NULL)
        elseif bc isa GenericBC
            # the GPU holds one exterior state per face node: Qext(Qin, x, frame, t, eq) is
            # evaluated once at t = 0 with a NaN interior state, so a condition that reads
            # Qin (or would change in time) cannot pass for a fixed state unnoticed
            data = zeros(Float64, nf * npf * nv)
            qnan = SVector(ntuple(_ -> NaN, nv))
            for f in eachbdface(mesh, ib), i in 1:npf
                k = (f - 1) * npf + i
                q = bc(qnan, geometry.faces.coords[k], geometry.faces.frames[k], 0.0, equation)
                any(isnan, q) && throw(ArgumentError(
                    "GenericBC on boundary $(ib) depends on the interior state: not available on the GPU path"))
                data[(k - 1) * nv + 1:k * nv] .= q
            end
            rc = ccall((:flou_hip_set_bc, libflouhip), Cint, (Ptr{Cvoid}, Int32, Int32, Ptr{Float64}),
                       hd.ptr, ib - 1, BC_DIRICHLET, data)
        else
            throw(ArgumentError("boundary condition $(typeof(bc)) is not available on the GPU path"))
        end
        check(hd.ptr, rc, "flou_hip_set_bc")
    end
    return nothing
end

"""
    set_option!(hd, which, value)

Options of `flou_hip_set_option`: 0 = reproduce the reference's two 3-D surface-flux
slips (default 1), 1 = force the FP32 plain kernels (rhs and adjoint).
"""
set_option!(hd::HIPDisc, which::Integer, value::Integer) =
    check(hd.ptr, ccall((:flou_hip_set_option, libflouhip), Cint, (Ptr{Cvoid}, Int32, Int32),
                        hd.ptr, which, value), "flou_hip_set_option")

"Does this FP32 handle run the fast rhs kernel?  (query 2 of `flou_hip_set_option`)"
uses_fast_kernel(hd::HIPDisc) =
    ccall((:flou_hip_set_option, libflouhip), Cint, (Ptr{Cvoid}, Int32, Int32), hd.ptr, 2, 0) == 1

"Does this FP32 handle's adjoint run the fast kernel?  (query 3 of `flou_hip_set_option`)"
uses_fast_vjp(hd::HIPDisc) =
    ccall((:flou_hip_set_option, libflouhip), Cint, (Ptr{Cvoid}, Int32, Int32), hd.ptr, 3, 0) == 1

# ----------------------------------------------------------------------------------------
# Device state: Flou's (ndofs, nvars) column-major state matrix in HIP device memory

"""
    HIPState{T} <: AbstractMatrix{T}

A state matrix in device memory, same layout as Flou's `Q.data` (rows grouped per
element, nodes x-fastest, variables in columns).  Elements are not indexable from the
host; `host(Q)` copies it back.
"""
mutable struct HIPState{T} <: AbstractMatrix{T}
    ptr::Ptr{Cvoid}
    nn::Int
    nv::Int
    owner::HIPDisc
end

Base.size(Q::HIPState) = (Q.nn, Q.nv)
Base.getindex(::HIPState, ::Int...) = error("HIPState lives in device memory: copy it with host(Q)")
Base.IndexStyle(::Type{<:HIPState}) = IndexCartesian()
Base.show(io::IO, Q::HIPState{T}) where {T} = print(io, "HIPState{", T, "}(", Q.nn, "×", Q.nv, " on the GPU)")
Base.show(io::IO, ::MIME"text/plain", Q::HIPState) = show(io, Q)
nbytes(Q::HIPState{T}) where {T} = Q.nn * Q.nv * sizeof(T)

function alloc(hd::HIPDisc)
    p = Ref{Ptr{Cvoid}}(C_NULL)
    bytes = hd.ndofs * hd.nvars * sizeof(hd.T)
    check(hd.ptr, ccall((:flou_hip_alloc, libflouhip), Cint, (Ptr{Cvoid}, Ref{Ptr{Cvoid}}, Csize_t),
                        hd.ptr, p, bytes), "flou_hip_alloc")
    Q = HIPState{hd.T}(p[], hd.ndofs, hd.nvars, hd)
    finalizer(Q) do x
        x.ptr == C_NULL || ccall((:flou_hip_free, libflouhip), Cint, (Ptr{Cvoid}, Ptr{Cvoid}), x.owner.ptr, x.ptr)
        x.ptr = C_NULL
    end
    return Q
end

Base.similar(Q::HIPState) = alloc(Q.owner)

"""
    device(hd, Q::AbstractMatrix) -> HIPState

Upload a host state (converted to the handle's precision).
"""
function device(hd::HIPDisc, Q::AbstractMatrix)
    size(Q) == (hd.ndofs, hd.nvars) || throw(DimensionMismatch("state must be ($(hd.ndofs), $(hd.nvars))"))
    A = Matrix{hd.T}(Q)
    D = alloc(hd)
    check(hd.ptr, ccall((:flou_hip_upload, libflouhip), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Csize_t),
                        hd.ptr, D.ptr, A, sizeof(A)), "flou_hip_upload")
    return D
end

"""
    host(Q::HIPState) -> Matrix

Download a device state.
"""
function host(Q::HIPState{T}) where {T}
    A = Matrix{T}(undef, Q.nn, Q.nv)
    check(Q.owner.ptr, ccall((:flou_hip_download, libflouhip), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Csize_t),
                             Q.owner.ptr, A, Q.ptr, sizeof(A)), "flou_hip_download")
    return A
end

function Base.copyto!(dst::HIPState, src::HIPState)
    size(dst) == size(src) || throw(DimensionMismatch("states differ in size"))
    sync(dst.owner)
    hipcheck(ccall((:hipMemcpy, libhip), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Csize_t, Cint),
                   dst.ptr, src.ptr, nbytes(src), 3), "hipMemcpy")  # hipMemcpyDeviceToDevice
    return dst
end

Base.copy(Q::HIPState) = copyto!(similar(Q), Q)

function zero!(Q::HIPState)
    sync(Q.owner)
    hipcheck(ccall((:hipMemset, libhip), Cint, (Ptr{Cvoid}, Cint, Csize_t), Q.ptr, 0, nbytes(Q)), "hipMemset")
    return Q
end

sync(hd::HIPDisc) = check(hd.ptr, ccall((:flou_hip_synchronize, libflouhip), Cint, (Ptr{Cvoid},), hd.ptr),
                          "flou_hip_synchronize")

function owner_of(disc, states::HIPState...)
    hd = states[1].owner
    hd.disc === disc || throw(ArgumentError("the device state belongs to another discretisation"))
    all(s -> s.owner === hd, states) || throw(ArgumentError("device states of different handles"))
    return hd
end

# ----------------------------------------------------------------------------------------
# rhs!, get_max_dt, integrate (FlouCommon / FlouSpatial methods for device states)

"""
    rhs!(dQ::HIPState, Q::HIPState, p::EquationConfig, t)

`rhs!` (src/FlouSpatial/Equations/Hyperbolic.jl:31-69) on the GPU: `flou_hip_rhs`.
"""
function FlouCommon.rhs!(dQ::HIPState, Q::HIPState,
                         p::FlouCommon.EquationConfig{<:MultielementDisc,<:FlouCommon.HyperbolicEquation},
                         time::Real)
    hd = owner_of(p.disc, dQ, Q)
    check(hd.ptr, ccall((:flou_hip_rhs, libflouhip), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Float64),
                        hd.ptr, dQ.ptr, Q.ptr, time), "rhs!")
    return nothing
end

function FlouCommon.get_max_dt(Q::HIPState, disc::MultielementDisc, ::FlouCommon.AbstractEquation, cfl)
    hd = owner_of(disc, Q)
    out = Ref(0.0)
    check(hd.ptr, ccall((:flou_hip_max_dt, libflouhip), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Float64, Ref{Float64}),
                        hd.ptr, Q.ptr, cfl, out), "get_max_dt")
    return out[]
end

function FlouSpatial.integrate(Q::HIPState, disc::MultielementDisc, ivar::Integer=1)
    hd = owner_of(disc, Q)
    out = Ref(0.0)
    check(hd.ptr, ccall((:flou_hip_integrate, libflouhip), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Int32, Ref{Float64}),
                        hd.ptr, Q.ptr, ivar - 1, out), "integrate")
    return out[]
end

"""
    monitor(Q::HIPState, name::Symbol)

The value `get_monitor(disc, eq, name)` gives for `Q` (Equations/Euler.jl:545-590):
`:energy` / `:kinetic_energy` or `:entropy`, accumulated in FP64 on the device.
"""
function monitor(Q::HIPState, name::Symbol)
    kind = name in (:energy, :kinetic_energy) ? MONITOR_KINETIC_ENERGY :
           name === :entropy ? MONITOR_ENTROPY : throw(ArgumentError("Unknown monitor '$(name)'."))
    out = Ref(0.0)
    check(Q.owner.ptr, ccall((:flou_hip_monitor, libflouhip), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Int32, Ref{Float64}),
                             Q.owner.ptr, Q.ptr, kind, out), "monitor")
    return out[]
end

# ----------------------------------------------------------------------------------------
# Stage limiter and time integration

"""
    gpu_limiter_callback(disc, equation, :zhang_shu, minval)

The stage limiter for the GPU path, passed to the solver as Flou's
`get_limiter_callback` result is (`ORK256(stage_limiter! = ...)`, FlouTime.jl:136-148):
`timeintegrate` then applies the Zhang & Shu limiter after every stage on the device.
"""
struct ZhangShuStage
    minval::Float64
end
(::ZhangShuStage)(u, integrator, p, t) = throw(ArgumentError(
    "the GPU Zhang & Shu stage limiter runs inside timeintegrate(::HIPState, ...)"))

function gpu_limiter_callback(::MultielementDisc, ::EulerEquation, name::Symbol, p::Real)
    name === :zhang_shu || throw(ArgumentError("Unknown limiter '$(name)'."))
    isfinite(p) || throw(ArgumentError("The minimum value must be specified when using the limiter of Zhang & Shu."))
    return ZhangShuStage(Float64(p))
end

# the scheme code of an OrdinaryDiffEq solver instance (the structs are parametric, so
# test with isa, not by type identity)
function scheme_code(solver)
    solver isa SSPRK33 && return SCHEME_SSPRK33
    solver isa CarpenterKennedy2N54 && return SCHEME_CK2N54
    solver isa ORK256 && return SCHEME_ORK256
    throw(ArgumentError("solver $(typeof(solver)) is not available on the GPU path " *
                        "(SSPRK33, CarpenterKennedy2N54, ORK256)"))
end
nstages(code) = code == SCHEME_SSPRK33 ? 3 : 5

function stage_limiter(solver)
    hasfield(typeof(solver), :stage_limiter!) || return nothing
    lim = getfield(solver, :stage_limiter!)
    lim isa ZhangShuStage && return lim
    lim === OrdinaryDiffEq.trivial_limiter! && return nothing
    throw(ArgumentError("a CPU stage limiter cannot act on a device state: use gpu_limiter_callback"))
end

"""
    timeintegrate(Q0::HIPState, disc, equation, solver, tfinal; dt, cfl=nothing, Δtmax=dt,
                  monitors=(), alias_u0=false)

`timeintegrate` (FlouTime.jl:34-56) for a device state, with fixed-step explicit RK
stages fused into the HIP rhs kernels.  `cfl` recomputes dt before every step as
`min(get_max_dt(Q, disc, eq, cfl), Δtmax)` (the get_cfl_callback pattern,
FlouTime.jl:92-104); `monitors` names monitors recorded after every step (the
get_monitor_callback pattern, FlouTime.jl:113-134).  Returns `(sol, exetime)` with
`sol.u[end]` the final state, `sol.t[end]` the final time and `sol.monitors[name]` the
recorded `(time, iter, value)`.  A non-physical state raises nothing: like Flou it logs
"Simulation crashed!" and returns what was reached.
"""
function FlouTime.timeintegrate(Q0::HIPState, disc::MultielementDisc, equation, solver, tfinal;
                                dt, cfl=nothing, Δtmax=dt, monitors=(), alias_u0=false, kwargs...)
    isempty(kwargs) || throw(ArgumentError("keywords not available on the GPU path: $(keys(kwargs))"))
    hd = owner_of(disc, Q0)
    code = scheme_code(solver)
    lim = stage_limiter(solver)
    Q = alias_u0 ? Q0 : copy(Q0)
    t = Ref(0.0)
    mons = Dict{Symbol,FlouTime.MonitorOutput}(m => FlouTime.MonitorOutput(Float64[], Int[], Float64[])
                                                for m in monitors)
    limcall(kind, v) = check(hd.ptr, ccall((:flou_hip_set_limiter, libflouhip), Cint,
                                           (Ptr{Cvoid}, Int32, Float64), hd.ptr, kind, v), "set_limiter")
    lim === nothing || limcall(LIMITER_ZHANG_SHU, lim.minval)
    exetime = @elapsed begin try
        if cfl === nothing && isempty(monitors)
            nsteps = round(Int64, tfinal / dt)
            check(hd.ptr, ccall((:flou_hip_timeintegrate, libflouhip), Cint,
                                (Ptr{Cvoid}, Int32, Ptr{Cvoid}, Float64, Int64, Float64, Ref{Float64}),
                                hd.ptr, code, Q.ptr, dt, nsteps, 0.0, t), "timeintegrate")
        else
            iter = 0
            while t[] < tfinal * (1 - 1e-12)
                h = cfl === nothing ? dt : min(Δtmax, FlouCommon.get_max_dt(Q, disc, equation, cfl))
                h = min(h, tfinal - t[])
                check(hd.ptr, ccall((:flou_hip_timeintegrate, libflouhip), Cint,
                                    (Ptr{Cvoid}, Int32, Ptr{Cvoid}, Float64, Int64, Float64, Ref{Float64}),
                                    hd.ptr, code, Q.ptr, h, 1, 0.0, t), "timeintegrate")
                iter += 1
                for (name, out) in mons
                    push!(out.time, t[]); push!(out.iter, iter); push!(out.value, monitor(Q, name))
                end
            end
        end
    catch e
        e isa DomainError || rethrow()
        @error "Simulation crashed!"
    finally
        lim === nothing || limcall(LIMITER_NONE, 0.0)
    end end
    return ((u=[Q], t=[t[]], monitors=mons), exetime)
end

# ----------------------------------------------------------------------------------------
# Output: only the projection to the equispaced points moves to the device; FlouBiz writes
# the VTKHDF file from the returned vectors exactly as before.
function FlouBiz.pointdata2VTKHDF(Q::HIPState{T}, disc::MultielementDisc) where {T}
    hd = owner_of(disc, Q)
    (; std) = disc
    seg = segment(std)
    ne, nd = nelements(disc), hd.nd
    m = length(seg.ξe)
    n = length(seg.ω)
    A = rowmajor(seg.node2eq)  # m x n, node2eq = A ⊗ ... ⊗ A (StdQuad.jl:53, StdHex.jl:61)
    Qe = Vector{T}(undef, ne * m^nd * hd.nvars)
    p = Ref{Ptr{Cvoid}}(C_NULL)
    check(hd.ptr, ccall((:flou_hip_alloc, libflouhip), Cint, (Ptr{Cvoid}, Ref{Ptr{Cvoid}}, Csize_t),
                        hd.ptr, p, sizeof(Qe)), "flou_hip_alloc")
    try
        check(hd.ptr, ccall((:flou_hip_equispaced, libflouhip), Cint,
                            (Ptr{Cvoid}, Int32, Ptr{Float64}, Ptr{Cvoid}, Int32, Ptr{Cvoid}),
                            hd.ptr, m, A, Q.ptr, hd.nvars, p[]), "pointdata2VTKHDF")
        check(hd.ptr, ccall((:flou_hip_download, libflouhip), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Csize_t),
                            hd.ptr, Qe, p[], sizeof(Qe)), "flou_hip_download")
    finally
        ccall((:flou_hip_free, libflouhip), Cint, (Ptr{Cvoid}, Ptr{Cvoid}), hd.ptr, p[])
    end
    npts = ne * m^nd
    return [Qe[(c - 1) * npts + 1:c * npts] for c in 1:hd.nvars]
end

# ----------------------------------------------------------------------------------------
# Differentiation: ChainRulesCore rules over the device adjoint

"""
    rhs(Q::HIPState, p::EquationConfig, t) -> HIPState

Out-of-place `rhs!`; its `rrule` pulls a cotangent back with `flou_hip_rhs_vjp`.
"""
function rhs(Q::HIPState, p::FlouCommon.EquationConfig, t::Real)
    dQ = similar(Q)
    FlouCommon.rhs!(dQ, Q, p, t)
    return dQ
end

function ChainRulesCore.rrule(::typeof(rhs), Q::HIPState, p::FlouCommon.EquationConfig, t::Real)
    dQ = rhs(Q, p, t)
    function rhs_pullback(dQbar)
        W = unthunk(dQbar)
        W isa HIPState || throw(ArgumentError("the cotangent of a HIPState must be a HIPState"))
        Qbar = zero!(similar(Q))
        check(Q.owner.ptr, ccall((:flou_hip_rhs_vjp, libflouhip), Cint,
                                 (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}),
                                 Q.owner.ptr, Q.ptr, W.ptr, Qbar.ptr), "rhs_vjp")
        return (NoTangent(), Qbar, NoTangent(), NoTangent())
    end
    return dQ, rhs_pullback
end

function step_record!(hd, code, dt, t, stages::Vector{<:HIPState}, unext::HIPState)
    ptrs = Ptr{Cvoid}[s.ptr for s in stages]
    check(hd.ptr, ccall((:flou_hip_rk_step_record, libflouhip), Cint,
                        (Ptr{Cvoid}, Int32, Float64, Float64, Ptr{Ptr{Cvoid}}, Ptr{Cvoid}),
                        hd.ptr, code, dt, t, ptrs, unext.ptr), "rk_step_record")
end

function step_vjp!(hd, code, dt, t, stages::Vector{<:HIPState}, ubar::HIPState)
    ptrs = Ptr{Cvoid}[s.ptr for s in stages]
    check(hd.ptr, ccall((:flou_hip_rk_step_vjp, libflouhip), Cint,
                        (Ptr{Cvoid}, Int32, Float64, Float64, Ptr{Ptr{Cvoid}}, Ptr{Cvoid}),
                        hd.ptr, code, dt, t, ptrs, ubar.ptr), "rk_step_vjp")
end

"""
    integrate_steps(Q0::HIPState, disc, equation, solver, dt, nsteps; checkpoint_every=20,
                    keep_segments=1)

`nsteps` fixed-dt RK steps from `Q0` (not modified); the differentiable form of
`timeintegrate`.  Its `rrule` keeps a checkpoint every `checkpoint_every` steps in the
forward run and, for the last `keep_segments` segments, every step's stage inputs; the
pullback walks a kept segment backwards with `flou_hip_rk_step_vjp` directly and recomputes
any other segment once with `flou_hip_rk_step_record` (keeping the stage inputs) first, so
memory is (nsteps / K + (keep_segments + 1)·K·stages) states.
"""
function integrate_steps(Q0::HIPState, disc::MultielementDisc, equation, solver, dt, nsteps;
                         checkpoint_every=20, keep_segments=1)
    Q, _, _ = forward_checkpointed(Q0, disc, solver, dt, nsteps, nsteps + 1, 0)
    return Q
end

function forward_checkpointed(Q0::HIPState, disc, solver, dt, nsteps, K, keep_segments)
    hd = owner_of(disc, Q0)
    code = scheme_code(solver)
    stage_limiter(solver) === nothing || throw(ArgumentError("adjoint: the stage limiter is not differentiated"))
    S = nstages(code)
    nseg = cld(nsteps, K)
    first_kept = max(0, nseg - keep_segments)      # 0-based index of the first kept segment
    stages = HIPState[copy(Q0)]
    append!(stages, [similar(Q0) for _ in 2:S])
    nxt = similar(Q0)
    cps = HIPState[]
    kept = Dict{Int,Vector{Vector{HIPState}}}()    # segment (1-based) => stage inputs per step
    for n in 0:nsteps - 1
        sg = n ÷ K
        n % K == 0 && push!(cps, copy(stages[1]))
        if sg >= first_kept
            steps = get!(kept, sg + 1, Vector{Vector{HIPState}}())
            st = HIPState[copy(stages[1])]
            append!(st, [similar(Q0) for _ in 2:S])
            step_record!(hd, code, dt, n * dt, st, nxt)
            push!(steps, st)
        else
            step_record!(hd, code, dt, n * dt, stages, nxt)
        end
        stages[1], nxt = nxt, stages[1]
    end
    return stages[1], cps, kept
end

function ChainRulesCore.rrule(::typeof(integrate_steps), Q0::HIPState, disc::MultielementDisc, equation,
                              solver, dt, nsteps; checkpoint_every=20, keep_segments=1)
    K = max(1, checkpoint_every)
    QT, cps, kept = forward_checkpointed(Q0, disc, solver, dt, nsteps, K, keep_segments)
    hd = owner_of(disc, Q0)
    code = scheme_code(solver)
    S = nstages(code)
    function integrate_steps_pullback(QTbar)
        ubar = copy(unthunk(QTbar))
        seg = nothing
        nxt = similar(Q0)
        for sgi in length(cps):-1:1
            n0 = (sgi - 1) * K
            nloc = min(K, nsteps - n0)
            if haskey(kept, sgi)  # stage inputs kept by the forward run
                for j in nloc:-1:1
                    step_vjp!(hd, code, dt, (n0 + j - 1) * dt, kept[sgi][j], ubar)
                end
                continue
            end
            seg === nothing && (seg = [[similar(Q0) for _ in 1:S] for _ in 1:min(K, nsteps)])
            copyto!(seg[1][1], cps[sgi])
            for j in 1:nloc  # recompute the segment, keeping every stage input
                step_record!(hd, code, dt, (n0 + j - 1) * dt, seg[j], j < nloc ? seg[j + 1][1] : nxt)
            end
            for j in nloc:-1:1
                step_vjp!(hd, code, dt, (n0 + j - 1) * dt, seg[j], ubar)
            end
        end
        return (NoTangent(), ubar, NoTangent(), NoTangent(), NoTangent(), NoTangent(), NoTangent())
    end
    return QT, integrate_steps_pullback
end

end # module FlouHIP
