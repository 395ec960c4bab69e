/*
 * flouhip.h -- C ABI of libflouhip.so, the MI355X (gfx950) implementation of Flou.jl's
 * DGSEM right-hand side, its explicit Runge-Kutta stage update, the CFL time-step
 * reduction, the conservation integral and the adjoint (vector-Jacobian products) of the
 * rhs and of the RK stage.
 *
 * Conventions
 *  - Plain C types only; device buffers are passed as raw pointers to HIP device memory.
 *  - State layout is Flou's ODE state matrix: Q is (ndofs, nvars), column-major
 *    (variable-major SoA), rows grouped per element (row = e*n^d + i, element nodes
 *    x-fastest).  Element numbering and face numbering are those of the host mesh.
 *    References: src/FlouSpatial/DofHandler.jl:53-68, src/FlouSpatial/Containers.jl:19-151.
 *  - The scalar type of every state buffer is the handle's precision: double for
 *    FLOU_HIP_FP64, float for FLOU_HIP_FP32.  Reductions always return FP64.
 *  - Every call that launches work is asynchronous on the handle's stream; outputs are
 *    valid once the stream is synchronised.  A handle is not thread safe.
 *  - Return codes: 0 on success, negative FLOU_HIP_E* on failure; flou_hip_last_error()
 *    gives the message.
 */
#ifndef FLOUHIP_H
#define FLOUHIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct flou_hip_handle flou_hip_handle;

/* Status codes */
enum {
    FLOU_HIP_OK = 0,
    FLOU_HIP_EINVAL = -1,       /* invalid argument (Julia: ArgumentError) */
    FLOU_HIP_EUNSUPPORTED = -2, /* valid in Flou but not implemented on the GPU path */
    FLOU_HIP_EDEVICE = -3,      /* HIP runtime error */
    FLOU_HIP_ENOMEM = -4,       /* device allocation failed */
    FLOU_HIP_EDOMAIN = -5       /* non-physical state met (Julia: DomainError) */
};

/* Precision of the state and of the bulk arithmetic */
enum { FLOU_HIP_FP64 = 0, FLOU_HIP_FP32 = 1 };

/* Equations: LinearAdvection{ND} and EulerEquation{ND} */
enum { FLOU_HIP_ADVECTION = 0, FLOU_HIP_EULER = 1 };

/* Volume operators: StrongDivOperator, SplitDivOperator, HybridDivOperator (the
 * telescopic subcell form blending two-point and finite-volume fluxes,
 * src/FlouSpatial/Equations/OpDivergence.jl:441-801; on GLL nodes in the volume term, on
 * Gauss nodes entirely in the surface term with entropy-projected end states; Cartesian
 * elements, Euler; the FV flux is the surface flux, as in the two- and three-argument
 * constructors) */
enum { FLOU_HIP_STRONG = 0, FLOU_HIP_SPLIT = 1, FLOU_HIP_HYBRID = 2 };

/* Two-point volume fluxes: StdAverage, ChandrasekharAverage, Kennedy-Gruber */
enum { FLOU_HIP_TP_STDAVG = 0, FLOU_HIP_TP_CHANDRASEKHAR = 1, FLOU_HIP_TP_KENNEDY_GRUBER = 2 };

/* Surface numerical flux = average + dissipation:
 *   StdAverage()                      -> (AVG_STD, DISS_NONE)
 *   LxF(StdAverage(), c)              -> (AVG_STD, DISS_LXF), intensity c
 *   MatrixDissipation(Chandrasekhar(), c) -> (AVG_CHANDRASEKHAR, DISS_MATRIX), intensity c
 *   ScalarDissipation(avg, c)             -> (avg, DISS_SCALAR), intensity c (Euler.jl:228-301) */
enum { FLOU_HIP_AVG_STD = 0, FLOU_HIP_AVG_CHANDRASEKHAR = 1 };
enum { FLOU_HIP_DISS_NONE = 0, FLOU_HIP_DISS_LXF = 1, FLOU_HIP_DISS_MATRIX = 2, FLOU_HIP_DISS_SCALAR = 3 };

/* Boundary conditions (applyBC!, src/FlouSpatial/Interfaces.jl:25-49;
 * src/FlouSpatial/Equations/Euler.jl:69-94) */
enum {
    FLOU_HIP_BC_DIRICHLET = 0, /* GenericBC with a time-independent Qext(x): per face-node state */
    FLOU_HIP_BC_INFLOW = 1,    /* EulerInflowBC(Qext) */
    FLOU_HIP_BC_OUTFLOW = 2,   /* EulerOutflowBC() */
    FLOU_HIP_BC_SLIP = 3       /* EulerSlipBC() */
};

/* Explicit RK schemes (OrdinaryDiffEq 6.49.1 perform_step! forms) */
enum { FLOU_HIP_SSPRK33 = 0, FLOU_HIP_CK2N54 = 1, FLOU_HIP_ORK256 = 2 };

/* Geometry storage */
enum {
    FLOU_HIP_GEOM_AFFINE = 0, /* metric, J constant per element; frames, jac constant per face */
    FLOU_HIP_GEOM_CURVED = 1  /* metric, J per node; frames, jac per face node */
};

/*
 * Flattened MultielementDisc + equation + operator
 * (src/FlouSpatial/MultielementDiscontinuous.jl:16-92).  All indices are 0-based.
 * The library copies everything it needs; the caller may free the arrays after create.
 */
typedef struct flou_hip_desc {
    int32_t nd;         /* spatial dimension, 1..3 */
    int32_t n;          /* nodes per direction (polynomial order + 1), 2..8 */
    int32_t nvars;      /* 1 (advection) or nd + 2 (Euler) */
    int32_t precision;  /* FLOU_HIP_FP64 | FLOU_HIP_FP32 */
    int64_t nelements;
    int64_t nfaces;
    int32_t nboundaries;

    int32_t equation;   /* FLOU_HIP_ADVECTION | FLOU_HIP_EULER */
    double gamma;       /* EulerEquation.γ */
    double velocity[3]; /* LinearAdvection.a */

    int32_t volume_op;  /* FLOU_HIP_STRONG | FLOU_HIP_SPLIT | FLOU_HIP_HYBRID */
    int32_t tp_flux;    /* two-point flux of SplitDivOperator */
    int32_t surface_avg;
    int32_t surface_diss;
    double intensity;

    /* 1-D operators of the standard element (StdSegment.jl:76-89), FP64.
     * Matrices row-major n x n: M[i*n + j] = M_ij. */
    const double* D;
    const double* Ds;
    const double* Dsharp;
    const double* l_minus;   /* std.l[1]: interpolation at ξ = -1 */
    const double* l_plus;    /* std.l[2] */
    const double* dg_minus;  /* std.∂g[1] */
    const double* dg_plus;   /* std.∂g[2] */
    const double* omega;     /* std.ω (1-D) */

    /* Connectivity (src/FlouCommon/Mesh.jl:19-113, after apply_periodicBCs!) */
    const int64_t* elem_faceinds;  /* nelements * 2nd: mesh.elements[ie].faceinds - 1 */
    const int8_t* elem_facepos;    /* nelements * 2nd: side of the element on that face (0 = master) */
    const int64_t* face_eleminds;  /* nfaces * 2: eleminds - 1 (-1: boundary exterior) */
    const int8_t* face_elempos;    /* nfaces * 2: elempos - 1 (-1: boundary exterior) */
    const uint8_t* face_orientation; /* nfaces, NULL = all 0: mesh.faces[i].orientation, the
                                        master2slave map of the face's nodes (StdSegment.jl:
                                        161-167: 0..1; StdQuad.jl:161-181: 0..7); 0 on boundary
                                        faces.  Non-zero orientations run the plain kernels. */
    const int32_t* face_boundary;  /* nfaces: boundary index of boundary faces, else -1 */

    /* Geometry (src/FlouSpatial/PhysicalRegions.jl:306-364, 485-535) */
    int32_t geometry;          /* FLOU_HIP_GEOM_AFFINE | FLOU_HIP_GEOM_CURVED */
    const double* elem_jac;    /* AFFINE: nelements; CURVED: ndofs */
    const double* elem_metric; /* AFFINE: nelements*nd*nd; CURVED: ndofs*nd*nd.  Each block
                                  column-major: m[r + nd*c] = metric[r, c] (Ja^c in column c) */
    const double* elem_volume; /* nelements */
    const double* face_frames; /* AFFINE: nfaces*3*nd; CURVED: nfacedofs*3*nd; (n, t, b) */
    const double* face_jac;    /* AFFINE: nfaces; CURVED: nfacedofs */

    double blend;              /* HybridDivOperator.blend (c >= 0); ignored otherwise */
} flou_hip_desc;

/* Library version string. */
const char* flou_hip_version(void);

/* Replaces MultielementDisc construction for the GPU (MultielementDiscontinuous.jl:29-92)
 * and HyperbolicCache / construct_cache (Equations/Hyperbolic.jl:16-29): validates the
 * description, uploads operators, connectivity and geometry, and builds the per-element
 * face tables.  Binds the handle to the current device and the default stream. */
int flou_hip_create(const flou_hip_desc* desc, flou_hip_handle** out);
int flou_hip_destroy(flou_hip_handle* h);
const char* flou_hip_last_error(const flou_hip_handle* h);
/* Message of the last failed flou_hip_create on this thread (the constructor ArgumentErrors
 * of MultielementDiscontinuous.jl:42-44 and OpDivergence.jl:211-214 map here). */
const char* flou_hip_create_error(void);

/* Options.
 *  0: reproduce two 3-D surface-flux typos of src/FlouSpatial/Equations/Euler.jl: the
 *     ChandrasekharAverage numerical flux of line 216 (used alone and as the average of
 *     MatrixDissipation), which squares wl where wr is meant, and the ScalarDissipation
 *     left momenta of line 257, taken from Qr.  Default 1: the reference's results;
 *     0 gives the corrected (entropy-consistent, symmetric) fluxes.  1-D and 2-D runs
 *     are unaffected.
 *  1: FP32 handles: 1 = always run the plain (reference-order) kernel, 0 = use the fast
 *     FP32 kernel where it applies (default).
 *  2: query, returns 1 if this FP32 handle has the fast kernel, else 0 (value ignored).
 *  3: query, returns 1 if this FP32 handle's adjoint (flou_hip_rhs_vjp, flou_hip_rk_step_vjp)
 *     has the fast kernel (wherever the fast rhs runs; option 1 switches it off with the
 *     fast rhs), else 0. */
int flou_hip_set_option(flou_hip_handle* h, int32_t which, int32_t value);

/* Stream the handle launches on (hipStream_t; NULL = default stream). */
int flou_hip_set_stream(flou_hip_handle* h, void* stream);

/* Boundary condition of boundary `ib` (disc.bcs, MultielementDiscontinuous.jl:40-51).
 * DIRICHLET: `data` = nfaces * n^(nd-1) * nvars host doubles (only faces of boundary ib are
 * read); INFLOW: `data` = nvars host doubles; OUTFLOW, SLIP: data ignored. */
int flou_hip_set_bc(flou_hip_handle* h, int32_t ib, int32_t type, const double* data);

/* Number of DOFs (nodes) and bytes of one state buffer. */
int64_t flou_hip_ndofs(const flou_hip_handle* h);
size_t flou_hip_state_bytes(const flou_hip_handle* h);

/* Device state helpers for hosts without their own device allocator (the Julia binding). */
int flou_hip_alloc(flou_hip_handle* h, void** dptr, size_t bytes);
int flou_hip_free(flou_hip_handle* h, void* dptr);
int flou_hip_upload(flou_hip_handle* h, void* dst_device, const void* src_host, size_t bytes);
int flou_hip_download(flou_hip_handle* h, void* dst_host, const void* src_device, size_t bytes);
int flou_hip_synchronize(flou_hip_handle* h);

/* rhs!(dQ, Q, EquationConfig(disc, eq), t) -- src/FlouSpatial/Equations/Hyperbolic.jl:31-69.
 * dQ is fully overwritten; Q is read only.  dQ and Q must not overlap. */
int flou_hip_rhs(flou_hip_handle* h, void* dQ, const void* Q, double t);

/* One fused RK stage = rhs of the stage input + the stage combination
 * (OrdinaryDiffEq perform_step!, called from FlouTime.timeintegrate, FlouTime.jl:34-56).
 *   SSPRK33, stage 0: u_out = u_in + dt*L(u_in)                          (u_base unused)
 *            stage 1: u_out = (3*u_base + u_in + dt*L(u_in)) / 4
 *            stage 2: u_out = (u_base + 2*u_in + 2*dt*L(u_in)) / 3       (u_out may be u_base)
 *   CK2N54 / ORK256, stage s: reg = A_s*reg + dt*L(u_in) (reg = dt*L(u_in) at s = 0);
 *            u_out = u_in + B_s*reg                                       (u_base unused)
 * u_out must not overlap u_in (neighbouring elements read u_in); u_base must not overlap
 * u_out except as the exact alias allowed at SSPRK33 stage 2 (EINVAL otherwise). */
int flou_hip_rk_stage(flou_hip_handle* h, int32_t scheme, int32_t stage, double dt, double t,
                      const void* u_in, const void* u_base, void* reg, void* u_out);

/* Full fixed-dt time integration of nsteps steps in place on Q, with handle-owned work
 * buffers (timeintegrate(Q0, disc, eq, solver, tf; dt, adaptive=false), FlouTime.jl:34-56).
 * If cfl > 0, dt is recomputed before every step as min(get_max_dt(Q, cfl), dt) (the
 * get_cfl_callback pattern, FlouTime.jl:92-104) and `dt` acts as Δtmax.
 * Returns FLOU_HIP_EDOMAIN if a non-physical state was met ("Simulation crashed!"). */
int flou_hip_timeintegrate(flou_hip_handle* h, int32_t scheme, void* Q, double dt,
                           int64_t nsteps, double cfl, double* t_inout);

/* get_max_dt(Q, disc, eq, cfl) -- MultielementDiscontinuous.jl:162-178 (FP64 result). */
int flou_hip_max_dt(flou_hip_handle* h, const void* Q, double cfl, double* out);

/* integrate(Q, disc, ivar) -- MultielementDiscontinuous.jl:148-160 (FP64 result); ivar
 * 0-based. */
int flou_hip_integrate(flou_hip_handle* h, const void* Q, int32_t ivar, double* out);

/* Stage limiter and monitors of the Euler equations (src/FlouSpatial/Equations/Euler.jl). */
enum { FLOU_HIP_LIMITER_NONE = 0, FLOU_HIP_LIMITER_ZHANG_SHU = 1 };
enum { FLOU_HIP_MONITOR_KINETIC_ENERGY = 0, FLOU_HIP_MONITOR_ENTROPY = 1 };

/* get_limiter(disc, eq, :zhang_shu, minval) handed to the RK method as its stage limiter
 * (get_limiter_callback, FlouTime.jl:136-148; zhang_shu_limiter, Euler.jl:612-654): once
 * set, flou_hip_rk_stage (and so flou_hip_timeintegrate) applies the limiter to u_out
 * after every stage.  FLOU_HIP_LIMITER_NONE switches it off.  minval must be finite (the
 * reference raises ArgumentError without one); EUNSUPPORTED for linear advection.  The
 * adjoint entry points refuse a handle with a limiter set. */
int flou_hip_set_limiter(flou_hip_handle* h, int32_t kind, double minval);

/* The Zhang & Shu limiter applied once, in place on Q (Euler.jl:619-653). */
int flou_hip_limit(flou_hip_handle* h, void* Q, double minval);

/* get_monitor(disc, eq, name) evaluated on Q (Euler.jl:545-590): Σ_e ∫ s(Q) with s the
 * kinetic energy |ρu|²/2ρ (reference name :energy) or the mathematical entropy
 * -ρ(log p - γ log ρ)/(γ-1); FP64 accumulation (FP64 result).  EDOMAIN (entropy) when
 * p or ρ is negative somewhere (log of a negative number). */
int flou_hip_monitor(flou_hip_handle* h, const void* Q, int32_t kind, double* out);

/* The point values of a VTKHDF file: FlouBiz.pointdata2VTKHDF(Q, disc)
 * (src/FlouSpatial/IO.jl:78-121) projects every element's nodal values onto m equispaced
 * points per direction, Qe_e = node2eq * Q_e with node2eq = A ⊗ ... ⊗ A
 * (StdQuad.jl:53, StdHex.jl:61; StdRegions.jl:232-238), for each column of Q.
 *   node2eq  host, m x n row-major FP64: A[i][j] = l_j(ξe_i), the basis interpolated at
 *            the equispaced points ξe (interp_matrix(ξe, basis), StdSegment.jl:56-58)
 *   Q        device, (ndofs, ncols) column-major in the handle's precision: a state
 *            (ncols = nvars) or a block of nvars*ND columns (IO.jl:102-121)
 *   Qe       device, (ne*m^d, ncols) column-major, same precision: column c is the point
 *            data of variable c in the order the reference appends it (element-major,
 *            equispaced points x-fastest)
 * Asynchronous on the handle's stream (A is copied to the device first; a call with a
 * different A synchronises the stream).  EUNSUPPORTED when an element's points do not fit
 * the kernel's 64 KB of LDS (max(n, m)^d > 4096 / bytes per value). */
int flou_hip_equispaced(flou_hip_handle* h, int32_t m, const double* node2eq, const void* Q,
                        int32_t ncols, void* Qe);

/* Synchronises, returns FLOU_HIP_EDOMAIN if a kernel met a non-physical state since the
 * last call (sqrt or log of a negative number in the reference), and clears the flag. */
int flou_hip_check(flou_hip_handle* h);

/* Adjoint of rhs! (Equations/Hyperbolic.jl:31-69): Qbar += (d rhs / dQ)^T dQbar, evaluated
 * at Q.  Flou has no AD rule for rhs!; this backs the ChainRulesCore rrule of the binding.
 * Supported: the strong and split forms with every surface flux (StdAverage or
 * Chandrasekhar average, no dissipation, LxF, Scalar or Matrix dissipation); the
 * HybridDivOperator, the split form on Gauss nodes and a handle with a stage limiter set
 * return FLOU_HIP_EUNSUPPORTED.  Qbar must not overlap Q or dQbar. */
int flou_hip_rhs_vjp(flou_hip_handle* h, const void* Q, const void* dQbar, void* Qbar);

/* One RK step of FlouTime.timeintegrate (FlouTime.jl:34-56; the OrdinaryDiffEq
 * perform_step! forms of flou_hip_rk_stage) that keeps the stage inputs for the adjoint:
 * stages[0] = u_n (read), stages[1..S-1] receive the inputs of stages 1..S-1, u_next the
 * result (S = 3 for SSPRK33, 5 for CK2N54 and ORK256).  All buffers distinct. */
int flou_hip_rk_step_record(flou_hip_handle* h, int32_t scheme, double dt, double t,
                            void* const* stages, void* u_next);

/* Adjoint of one RK step at fixed dt (no reference counterpart; the reverse sweep of the
 * checkpointed rrule of timeintegrate): on entry ubar = cotangent of u_{n+1}, on return
 * the cotangent of u_n.  stages[] as written by flou_hip_rk_step_record.  Uses three
 * handle-owned work buffers; ubar must not overlap any stage buffer. */
int flou_hip_rk_step_vjp(flou_hip_handle* h, int32_t scheme, double dt, double t,
                         const void* const* stages, void* ubar);

/* Bytes of the static device buffer behind flou_hip_rk_step_vjp_noalloc, per GPU. */
#define FLOU_HIP_NOALLOC_SCRATCH_BYTES (64u << 20)

/* flou_hip_rk_step_vjp without device allocation.  No counterpart in the reference: it is
 * for callers that cannot allocate (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  The three work states live in one static device buffer
 * of FLOU_HIP_NOALLOC_SCRATCH_BYTES that each GPU holds for the life of the process, instead
 * of handle-owned buffers allocated on first use; the call itself allocates nothing and does
 * not synchronise.  Same arguments, results and checks as flou_hip_rk_step_vjp, plus
 * FLOU_HIP_EINVAL when three states (each rounded up to 256 bytes) exceed the buffer: up
 * to 22.3 MB per state, e.g. 699,050 nodes of FP64 2-D Euler (a 128 x 128 mesh at N = 5
 * has 589,824) or 1.1 M nodes of FP32 3-D Euler.  Every call on
 * one GPU shares that buffer, whatever the handle: calls that overlap in time on one GPU
 * (other streams or host threads) race and give wrong results. */
int flou_hip_rk_step_vjp_noalloc(flou_hip_handle* h, int32_t scheme, double dt, double t,
                                 const void* const* stages, void* ubar);

#ifdef __cplusplus
}
#endif

#endif /* FLOUHIP_H */
