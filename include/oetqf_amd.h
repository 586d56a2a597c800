/*
 * oetqf_amd.h -- C ABI of the MI355X (gfx950) earthquake-cycle hot path.
 *
 * Drop-in for the CPU hot path of Oetqf.jl v0.3.6: the stress Green's-function
 * builders (src/BEM/GF.jl), the ODE right-hand side that `assemble` hands to
 * OrdinaryDiffEq (src/BEM/equation.jl) and the Tsit5 time loop the solver runs
 * over it.  Plain C: pointers, sizes and status codes; no exceptions cross it.
 *
 * Conventions (those of the reference, Julia column-major):
 *  - fault fields: nx*nxi doubles, strike index fastest (vec(v), equation.jl:35-42);
 *  - mantle fields: ne*6 doubles, component-major [xx xy xz yy yz zz]
 *    (equation.jl:207-222, GF.jl:163-169);
 *  - tuple vectors (ax, axi: Vector{NTuple{2,Float64}}) are 2n interleaved doubles;
 *  - ftype: OQ_STRIKE_SLIP / OQ_DIP_SLIP (GF.jl:3-5);
 *  - every function returns OQ_OK (0) or an error code; oq_last_error() gives
 *    a message (thread-local).  Argument checks mirror the reference's
 *    @assert's (GF.jl:36,131,326; property.jl:20-24; mesh.jl:19-20,70-71).
 *  - host-pointer entry points are synchronous; *_device entry points enqueue
 *    on the handle's HIP stream and take device pointers.
 */
#ifndef OETQF_AMD_H
#define OETQF_AMD_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define OQ_OK 0
#define OQ_ERR_ARG 1     /* invalid argument (reference: AssertionError / ArgumentError) */
#define OQ_ERR_HIP 2     /* HIP runtime failure */
#define OQ_ERR_STATE 3   /* handle in an invalid state */
#define OQ_ERR_NOMEM 4   /* device allocation failed */

#define OQ_STRIKE_SLIP 0 /* StrikeSlip (GF.jl:4) */
#define OQ_DIP_SLIP 1    /* DipSlip (GF.jl:5) */

/* Storage type of the dense operator K streamed by the RHS. */
#define OQ_K_FP64 0
#define OQ_K_FP32 1
#define OQ_K_BF16 2
/* Fault-only problems (ne = 0): K_ff stays its Toeplitz kernel st
 * (nx*nxi*nxi FP64, GF.jl:41-58) instead of the dense (nx*nxi)^2 operator,
 * and is applied along strike directly (the reference applies it by FFT
 * convolution, equation.jl:44-61). */
#define OQ_K_TOEPLITZ 3
/* Ensembles only: each row of K as two FP16 halves (hi + lo, 22 significant
 * bits) times a power-of-two row scale, multiplied on the FP16 matrix cores
 * with the members' source vectors split the same way (three products per
 * pair, FP32 chunk sums, FP64 totals).  Same bytes as FP32, FP32-class
 * error, eight times the FP32 matrix rate. */
#define OQ_K_FP16X2 4
/* K_ff applied as the reference applies it, by DFT convolution along
 * strike (equation.jl:44-61 fault only, :200 coupled): FFT of the
 * zero-padded source, per-frequency contraction with the DFT of the
 * symmetric circulant of st (GF.jl:60-67, real), inverse FFT.  The padded
 * length is the power of two L >= 2nx-1 (the reference pads to 2nx-1; both
 * give the same first nx rows of the circular convolution).  st stays
 * resident as well (operator expansion).  Coupled (ne > 0): K_fm
 * (nf x 6ne), K_mf and K_mm (6ne x N) stay dense FP64; the fault rows add
 * the DFT traction rates to K_fm s_m. */
#define OQ_K_FFT 5
/* oq_rhs_create_from_mesh only, coupled model: OQ_K_FFT, and K_mm applied
 * as 2D convolutions over a mantle whose cells form a uniform horizontal
 * grid (cell ix + nx (iy + ny iz), equal spacing and extents along x and y,
 * the same layers under every column; OQ_ERR_ARG otherwise, and for more
 * than about 60 x 60 cells per layer).  The half-space stress between two
 * cells then depends on their horizontal offset and their two layers only,
 * so K_mm is held as 36 nz^2 tiles of (2ny-1) x (2nx-1) FP64 entries
 * (17.7 MB for 16 x 16 x 8 cells instead of 1.21 GB), assembled by the same
 * pair kernel for the shifted receivers.  When the mantle's x spacing is m
 * times the fault's strike spacing (m a whole number, at most 32 cells
 * along x), K_fm and K_mf depend on a patch-cell pair's x offset only
 * through a - m jx and are held as tables of nx_f + m (nx - 1) offsets per
 * fault row and mantle line (49 MB each at the 128x32 fault with 16x16x8
 * cells instead of 403 MB); otherwise they stay dense FP64.  All FP64.
 * No counterpart in the reference (which multiplies the dense blocks,
 * equation.jl:201-203); entries equal the dense blocks' to rounding. */
#define OQ_K_FFT_GRID 6

/* RectOkadaMesh (src/BEM/mesh.jl:5-21).  Only x, ax, y, z, axi, nx, nxi, dx,
 * dep and dip enter the Green's functions (GF.jl:38,49-53,150-160). */
typedef struct oq_fault_mesh {
    const double *x;   /* nx strike centroids */
    const double *ax;  /* 2*nx: (x - dx/2, x + dx/2) */
    const double *y;   /* nxi */
    const double *z;   /* nxi */
    const double *axi; /* 2*nxi: down-dip extents */
    int64_t nx, nxi;
    double dx, dep, dip; /* dip in degrees */
} oq_fault_mesh;

/* BEMHex8Mesh (src/BEM/mesh.jl:58-72). */
typedef struct oq_hex8_mesh {
    const double *cx, *cy, *cz; /* barycentres */
    const double *qx, *qy, *qz; /* reference points (cx, cy - dy/2, cz + dz/2) */
    const double *dx, *dy, *dz; /* cell extents (T, L, W) */
    int64_t ne;
} oq_hex8_mesh;

/* Quadrature over a hex8 cell in [-1,1]^3, weights normalised to sum 1
 * (get_quadrature, GF.jl:318-328).  coords: 3*nq. */
typedef struct oq_quadrature {
    const double *coords;
    const double *weights;
    int64_t nq;
} oq_quadrature;

/* RateStateQuasiDynamicProperty (src/BEM/property.jl:10-25), aging law
 * (DieterichStateLaw, property.jl:4-5). a, b, L, sigma: nx*nxi each. */
typedef struct oq_rs_props {
    const double *a, *b, *L, *sigma;
    double eta, vpl, f0, v0;
} oq_rs_props;

/* PowerLawViscosityProperty (property.jl:34-41), or with nlaws > 1 the
 * CompositePowerLawViscosityProperty (property.jl:43-48) whose laws are
 * stacked: gamma[law*ne + i], n[law*ne + i] (n = power - 1). */
typedef struct oq_visc_props {
    const double *gamma;
    const double *n;
    int64_t nlaws;
    double deps0[6];
} oq_visc_props;

/* DilatancyProperty (property.jl:27-32): per patch, nx*nxi each. */
typedef struct oq_dila_props {
    const double *tp;   /* characteristic diffusion time */
    const double *eps;  /* dilatancy coefficient */
    const double *beta; /* gouge bulk compressibility */
    const double *p0;   /* ambient pore pressure */
} oq_dila_props;

/* Solver options (the reference passes these to OrdinaryDiffEq.solve:
 * examples/otf-with-mantle.jl:160-161). dt0 <= 0 selects the automatic
 * initial step; dtmax <= 0 means unbounded. */
typedef struct oq_solver_opts {
    double reltol, abstol, dt0, dtmax, dtmin;
    int64_t maxiters;
} oq_solver_opts;

/* Stepper return codes (OrdinaryDiffEq ReturnCode analogues). */
#define OQ_RET_RUNNING 0
#define OQ_RET_SUCCESS 1
#define OQ_RET_MAXITERS 2
#define OQ_RET_DTMIN 3
#define OQ_RET_NONFINITE 4

typedef struct oq_rhs oq_rhs;
typedef struct oq_stepper oq_stepper;

/* ---- library ---- */
const char *oq_last_error(void);
int oq_version(void);
int oq_device_count(int *count);

/* ---- Green's functions (host output, caller-owned, Julia layout) ---- */

/* Replaces stress_greens_function(mesh::RectOkadaMesh, λ, μ; ftype,
 * fourier=false, nrept, buffer_ratio) (src/BEM/GF.jl:31-58).
 * st: nx*nxi*nxi, st[i + nx*j + nx*nxi*l]. */
int oq_gf_okada_ff(const oq_fault_mesh *mesh, double lambda, double mu,
                   int ftype, int64_t nrept, double buffer_ratio, double *st);

/* The fourier=true branch (src/BEM/GF.jl:60-68): the length-(2nx-1) real DFT
 * of [st[:,j,l]; reverse(st[2:end,j,l])]; st_dft: 2*nx*nxi*nxi doubles,
 * interleaved (re, im), Complex{Float64} layout.  Both the kernel and its
 * DFT are computed on the device (one workgroup per (j, l) column; the
 * sequence is symmetric, so the imaginary parts are exactly 0).  nx <= 5120. */
int oq_gf_okada_ff_dft(const oq_fault_mesh *mesh, double lambda, double mu,
                       int ftype, int64_t nrept, double buffer_ratio,
                       double *st_dft);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops): the
 * fourier=true K_ff of oq_gf_okada_ff_dft with nothing allocated.  The mesh
 * arrays (mesh->x, ax, y, z, axi) and st_dft (2*nx*nxi*nxi doubles, same
 * layout) are DEVICE pointers; the work is enqueued on `stream` (a
 * hipStream_t; NULL = the default stream) and the call returns without
 * waiting.  The real kernel st between the two launches goes through a
 * static device scratch array of 1048576 doubles (8 MiB, held by each GPU
 * the library runs on until the process ends): nx*nxi*nxi above that is
 * rejected with OQ_ERR_ARG before anything is enqueued.  Calls on one GPU
 * share that array, so calls that overlap on one GPU (different streams or
 * threads) race.  The first call on each device looks the array's address
 * up (hipGetSymbolAddress, no allocation); make it outside graph capture. */
int oq_gf_okada_ff_dft_noalloc(const oq_fault_mesh *mesh_dev, double lambda, double mu,
                               int ftype, int64_t nrept, double buffer_ratio,
                               double *st_dft_dev, void *stream);

/* Replaces stress_greens_function(mf::RectOkadaMesh, ma::BEMHex8Mesh, λ, μ;
 * ftype, qtype, nrept, buffer_ratio) (src/BEM/GF.jl:123-174), gf₁₂.
 * st: (6ne) x (nx*nxi), column-major. */
int oq_gf_okada_mf(const oq_fault_mesh *mf, const oq_hex8_mesh *ma,
                   double lambda, double mu, int ftype,
                   const oq_quadrature *quad, int64_t nrept,
                   double buffer_ratio, double *st);

/* Replaces stress_greens_function(ma::BEMHex8Mesh, mf::RectOkadaMesh, λ, μ;
 * ftype) (src/BEM/GF.jl:194-227), gf₂₁.  st: (nx*nxi) x (6ne), column-major. */
int oq_gf_hex8_fm(const oq_hex8_mesh *ma, const oq_fault_mesh *mf,
                  double lambda, double mu, int ftype, double *st);

/* Replaces stress_greens_function(ma::BEMHex8Mesh, λ, μ; qtype,
 * checkeigvals=false) (src/BEM/GF.jl:250-290), gf₂₂.
 * st: (6ne) x (6ne), column-major.  The eigenvalue diagnostic of
 * GF.jl:291-294 stays with the caller. */
int oq_gf_hex8_mm(const oq_hex8_mesh *ma, double lambda, double mu,
                  const oq_quadrature *quad, double *st);

/* ---- right-hand side ---- */

/* Replaces assemble(gf, p, u0, tspan) (equation.jl:81-89; fault only: ne = 0,
 * visc, gf12, gf21, gf22 NULL) and assemble(gf₁₁, gf₁₂, gf₂₁, gf₂₂, pf, pa,
 * u0, tspan) (equation.jl:141-154): binds the Green's functions and
 * properties to a device-resident handle.  gf11 is the real Toeplitz kernel
 * (nx,nxi,nxi) when gf11_is_dft = 0, or its complex DFT (2*nx*nxi*nxi
 * doubles, the default output of the reference builder) when 1.  The
 * library expands K into one dense receiver-major operator stored as
 * kdtype and owns it until oq_rhs_destroy. */
int oq_rhs_create(int device, int64_t nx, int64_t nxi, int64_t ne,
                  const oq_rs_props *fault, const oq_visc_props *visc,
                  const double *gf11, int gf11_is_dft, const double *gf12,
                  const double *gf21, const double *gf22, int kdtype,
                  oq_rhs **out);

/* As oq_rhs_create, but the four blocks are assembled directly on the device
 * (GF.jl:31-58,123-174,194-227,250-290) -- no host copy of K. */
int oq_rhs_create_from_mesh(int device, const oq_fault_mesh *mf,
                            const oq_hex8_mesh *ma, double lambda, double mu,
                            int ftype, int64_t nrept, double buffer_ratio,
                            const oq_quadrature *quad,
                            const oq_rs_props *fault,
                            const oq_visc_props *visc, int kdtype,
                            oq_rhs **out);

/* Replaces assemble(gf, p, dila, u0, tspan) (equation.jl:108-117) on a
 * fault-only handle: the RHS becomes the dilatancy `ode` (equation.jl:168-183,
 * update_fault_with_dilatancy! :248-276) and the packed state [v theta
 * delta p] (4 nx*nxi).  Call before creating a stepper on the handle. */
int oq_rhs_set_dilatancy(oq_rhs *h, const oq_dila_props *dila);
/* The dilatancy `ode(du,u,p,t)` on host arrays (u = (v, theta, delta, p)). */
int oq_rhs_eval_dila(oq_rhs *h, double t, const double *v, const double *theta,
                     const double *delta, const double *p, double *dv, double *dtheta,
                     double *ddelta, double *dp);

/* The RHS `ode(du,u,p,t)` (equation.jl:156-166 fault only, :185-205 coupled)
 * on host arrays: the ArrayPartition components are separate pointers.
 * Fault only: eps, sigma, deps, dsigma are ignored (may be NULL). */
int oq_rhs_eval(oq_rhs *h, double t, const double *v, const double *theta,
                const double *eps, const double *sigma, const double *delta,
                double *dv, double *dtheta, double *deps, double *dsigma,
                double *ddelta);

/* Same on device arrays in packed ArrayPartition order
 * ([v θ δ] or [v θ ϵ σ δ]), enqueued on the handle's stream. */
int oq_rhs_eval_device(oq_rhs *h, double t, const double *u, double *du);

/* Sizes: number of state entries, matrix order N = nf + 6ne, bytes of the
 * resident operator (dense K; the Toeplitz kernel; or, coupled DFT, the
 * kernel plus the dense K_fm, K_mf and K_mm blocks). */
int oq_rhs_info(oq_rhs *h, int64_t *nstate, int64_t *n, int64_t *kbytes);

/* Copies the dense operator back (N*N doubles, row r = receiver), for tests. */
int oq_rhs_get_k(oq_rhs *h, double *k_rowmajor);

/* Times `iters` device RHS evaluations on the handle's stream with HIP
 * events; reports the mean per evaluation and the mean of the dominant
 * (matrix-streaming) kernel. */
int oq_rhs_time(oq_rhs *h, int iters, double *ms_eval, double *ms_matvec);

int oq_rhs_destroy(oq_rhs *h);

/* ---- on-device Tsit5 stepper (replaces solve(prob, Tsit5(); reltol,
 * abstol, dt, dtmax, maxiters), io.jl:125-130, test/tests.jl:11) ---- */
int oq_stepper_create(oq_rhs *h, const double *u0, double t0, double t1,
                      const oq_solver_opts *opts, oq_stepper **out);
/* Advances up to max_steps attempted steps (or to t1).  *retcode gets one of
 * OQ_RET_*. */
int oq_stepper_run(oq_stepper *s, int64_t max_steps, int *retcode);
/* Current state: t, dt, u (packed, host), step counters. */
int oq_stepper_state(oq_stepper *s, double *t, double *dt, double *u,
                     int64_t *naccept, int64_t *nreject);
/* Device recorder behind wsolve (src/io.jl:51-82): after every accepted
 * step the time and max|v| over the fault go to a step ring of
 * step_capacity entries; when the accepted-step count is a multiple of
 * `stride` the whole packed state goes to a snapshot ring of snap_capacity
 * states (the caller saves the initial state itself, as wsolve's callback
 * does at t0).  Rings must be drained (oq_stepper_drain) before they fill:
 * running at most min(step_capacity, snap_capacity) attempted steps between
 * drains guarantees it. */
int oq_stepper_record(oq_stepper *s, int64_t stride, int64_t snap_capacity,
                      int64_t step_capacity);
/* Copies out and clears the rings: *nsteps step records (t, max|v|) and
 * *nsnaps snapshots (t, packed u).  Buffers must hold the capacities.  On a
 * ring overrun the newest records are handed out and OQ_ERR_STATE names
 * how many were lost (see oq_stepper_drain_du). */
int oq_stepper_drain(oq_stepper *s, int64_t *nsteps, double *step_t, double *step_vmax,
                     int64_t *nsnaps, double *snap_t, double *snap_u);
/* As oq_stepper_record; with du != 0 every snapshot also keeps the
 * derivative f(u) at the saved state (Tsit5: the FSAL stage k7 of the
 * accepted step; VCABM5: the corrector's f(y_{n+1})), which is what the
 * reference's saving handlers read as integrator(integrator.t, Val{1})
 * (src/io.jl:86, examples/otf-with-mantle.jl:157-159). */
int oq_stepper_record_du(oq_stepper *s, int64_t stride, int64_t snap_capacity,
                         int64_t step_capacity, int du);
/* As oq_stepper_drain, plus the derivative snapshots (snap_du: snap_capacity
 * packed states; may be NULL).  On a ring overrun (more records since the
 * last drain than a ring holds) both drains hand out the newest records in
 * order, restart the recorder and return OQ_ERR_STATE, with the number of
 * lost records in oq_last_error(). */
int oq_stepper_drain_du(oq_stepper *s, int64_t *nsteps, double *step_t, double *step_vmax,
                        int64_t *nsnaps, double *snap_t, double *snap_u, double *snap_du);
int oq_stepper_destroy(oq_stepper *s);
/* As oq_stepper_create, with VCABM5 instead of Tsit5 (the solver of the
 * reference's example, examples/otf-with-mantle.jl:160,167-168): variable-
 * coefficient Adams-Bashforth-Moulton of order 5 in divided-difference form
 * (Hairer, Norsett & Wanner I, Sec. III.5), PECE, two RHS per step; the
 * order ramps from 1 to 5 over the first steps.  run/state/record/drain as
 * for Tsit5. */
int oq_stepper_create_vcabm5(oq_rhs *h, const double *u0, double t0, double t1,
                             const oq_solver_opts *opts, oq_stepper **out);
/* Persistent mode (no counterpart in the reference; the graph stepper's
 * semantics are kept): for fault-only problems with an FP64 operator whose
 * rows fit in LDS of at most 128 workgroups (N <= 1024), oq_stepper_run executes all of
 * its steps in one cooperative launch -- rows of K resident in LDS, grid
 * barriers between the RHS evaluations (Tsit5: 6 per step, the next step's
 * first stage riding on the error-norm barrier; VCABM5: 2 per step) --
 * instead of replaying the per-step graph.  Tsit5 and VCABM5 steppers both
 * qualify.  on = 0 returns to the graph.  OQ_ERR_ARG if the problem does not
 * qualify. */
int oq_stepper_set_persistent(oq_stepper *s, int on);

/* ---- ensembles of frictional-parameter realizations sharing one
 * geometry.  No entry point of the reference corresponds: a parameter study
 * there is one `assemble` + `solve` per realization (equation.jl:141-154,
 * examples/otf-with-mantle.jl:154-161).  Here the members' RHS is one MFMA
 * GEMM over a shared FP32/BF16 operator and a lockstep Tsit5 advances all
 * members with per-member dt and acceptance.  Member states are packed
 * ArrayPartitions laid one after another (member-major, n entries each). */
typedef struct oq_ensemble oq_ensemble;

/* members: nmembers oq_rs_props (per-member a, b, L, sigma arrays and
 * scalars); visc is shared.  kdtype: OQ_K_FP32, OQ_K_BF16 or OQ_K_FP16X2. */
int oq_ensemble_create(int device, const oq_fault_mesh *mf, const oq_hex8_mesh *ma,
                       double lambda, double mu, int ftype, int64_t nrept,
                       double buffer_ratio, const oq_quadrature *quad,
                       int64_t nmembers, const oq_rs_props *members,
                       const oq_visc_props *visc, int kdtype, oq_ensemble **out);
/* du = f(u) for all members (host arrays, nmembers * nstate each). */
int oq_ensemble_eval(oq_ensemble *e, const double *u, double *du);
/* Sets the members' initial states and options (opts->dt0 > 0 required)
 * and captures one lockstep step into a HIP graph. */
int oq_ensemble_start(oq_ensemble *e, const double *u0, double t0, double t1,
                      const oq_solver_opts *opts);
/* The same with VCABM5 (the solver of examples/otf-with-mantle.jl:160, as
 * oq_stepper_create_vcabm5: order 1 rising to 5, two RHS per attempted
 * step) for every member, each with its own order, step history and
 * coefficients. */
int oq_ensemble_start_vcabm5(oq_ensemble *e, const double *u0, double t0, double t1,
                             const oq_solver_opts *opts);
/* Advances up to max_steps lockstep attempts; *nrunning = members not done. */
int oq_ensemble_run(oq_ensemble *e, int64_t max_steps, int64_t *nrunning);
/* Per-member t[M], u[M*nstate], step counters and OQ_RET_* codes (0 = running). */
int oq_ensemble_state(oq_ensemble *e, double *t, double *u, int64_t *naccept,
                      int64_t *nreject, int *retcode);

/* Checkpoint / resume of a lockstep run (no counterpart in the reference,
 * whose solves are one per member; a restart there is a new ODEProblem from
 * the saved u).  Per member: t, the next step dt, the PI controller's qold,
 * the accepted / rejected / attempted step counts, done and the OQ_RET_*
 * code.  Any pointer of oq_*_get_ctrl may be NULL (not copied). */
typedef struct oq_lockstep_ctrl {
    double *t, *dt, *qold;
    int64_t *naccept, *nreject, *iter;
    int *done, *retcode;
} oq_lockstep_ctrl;
/* Copies the controller state of the current (or last) run to host arrays
 * of M entries. */
int oq_ensemble_get_ctrl(oq_ensemble *e, oq_lockstep_ctrl *ctrl);
/* Starts stepping from saved states u[M*nstate] and controller arrays (all
 * required) towards t1 with the tolerances and limits of opts (opts->dt0 is
 * not used).  With the same operator the continuation is bit for bit the
 * uninterrupted run: k1 = f(u) is recomputed with the same kernels.  The
 * recorder restarts empty.  Tsit5 runs (a VCABM5 run's divided-difference
 * history is not part of the checkpoint). */
int oq_ensemble_resume(oq_ensemble *e, const double *u, const oq_lockstep_ctrl *ctrl,
                       double t1, const oq_solver_opts *opts);
/* Per-member record of accepted steps, the ensemble form of the single-state
 * recorder (the saving callback of wsolve, io.jl:51-58, reduced to (t,
 * max |v|)): after each accepted step of member m, its new t and the
 * largest slip rate of its new state go into a ring of step_capacity
 * entries.  0 turns it off.  Takes effect at the next oq_ensemble_start;
 * while on, oq_ensemble_run accepts at most step_capacity steps per call. */
int oq_ensemble_record(oq_ensemble *e, int64_t step_capacity);
/* Hands out the records made since the last drain: counts[m] entries of
 * member m at step_t[m*cap + j], step_vmax[m*cap + j] (arrays of
 * nmembers*step_capacity).  OQ_ERR_STATE if a ring was overrun. */
int oq_ensemble_drain(oq_ensemble *e, int64_t *counts, double *step_t, double *step_vmax);
/* Mean time of one ensemble RHS (all members) over `iters` evaluations. */
int oq_ensemble_time(oq_ensemble *e, int iters, double *ms_eval);
int oq_ensemble_destroy(oq_ensemble *e);

/* ---- grouped batches: geometries of uneven size, up to 16 frictional
 * realizations each.  No entry point of the reference corresponds (one
 * `assemble` + `solve` per geometry and realization, equation.jl:141-154).
 * Each geometry's operator is FP32; one RHS is one grouped MFMA launch per
 * tile shape in use over all geometries, and a lockstep Tsit5 advances all
 * ngeo*nmembers members with per-member dt and acceptance.  Member states
 * are packed ArrayPartitions, geometry after geometry, member after member
 * (geometry g: nmembers states of nstate[g] entries). */
typedef struct oq_group oq_group;

/* mf[ngeo]; ma[ngeo] (NULL entries: fault only; ma itself may be NULL);
 * visc[ngeo] (used where ma[g] is set); members[ngeo*nmembers], member m of
 * geometry g at g*nmembers + m, its arrays of that geometry's nx*nxi
 * patches.  nmembers in [1, 16]; kdtype OQ_K_FP32. */
int oq_group_create(int device, int64_t ngeo, const oq_fault_mesh *mf,
                    const oq_hex8_mesh *const *ma, const oq_visc_props *const *visc,
                    double lambda, double mu, int ftype, int64_t nrept, double buffer_ratio,
                    const oq_quadrature *quad, int64_t nmembers, const oq_rs_props *members,
                    int kdtype, oq_group **out);
/* Total state entries, per-geometry state size nstate[ngeo] and the bytes
 * of all operators (any pointer may be NULL). */
int oq_group_info(oq_group *e, int64_t *ntot, int64_t *nstate, int64_t *kbytes);
/* Times every tile shape (15: register-streaming kernels with 16/32/64 rows
 * per wavefront, each with and without a software-pipelined K stream,
 * LDS-staged kernels with 16/32 rows per wavefront and one or two slices in
 * flight, the two-slice kernel on column-split units of ~2 MB balanced
 * over the resident workgroups, and four configurations of a kernel on the
 * same kind of units that copies K and S from HBM straight into an LDS ring
 * (3 slices of 64 rows x 128 columns, 1 workgroup per CU; 3 x 32 x 128,
 * 2 per CU; 3 x 64 x 64, 2 per CU; 2 x 64 x 64, 4 per CU)) on the
 * geometries of each size bucket
 * (bucket b: N < 2048 << b;
 * 8 buckets) and keeps the fastest; then times the whole grouped GEMM (one
 * launch per shape in use) with that per-bucket choice and with each single
 * shape for all buckets (one launch), and keeps the faster configuration.
 * shape_of_bucket[8] and ms_of_bucket[8] receive the choice (-1 for an empty
 * bucket), ms_table[8*15] (may be NULL) every per-bucket timing.  Call before
 * oq_group_start (a running stepper must be started again). */
int oq_group_autotune(oq_group *e, int iters, int *shape_of_bucket, double *ms_of_bucket,
                      double *ms_table);
/* Pins the tile shape of each size bucket (shape_of_bucket[8]: 0..14, -1 =
 * keep), e.g. to reuse an earlier autotuning.  Like oq_group_autotune, call
 * before oq_group_start. */
int oq_group_set_shapes(oq_group *e, const int *shape_of_bucket);
/* du = f(u) for all members (host arrays of ntot entries). */
int oq_group_eval(oq_group *e, const double *u, double *du);
/* Mean time of one grouped RHS (all geometries and members). */
int oq_group_time(oq_group *e, int iters, double *ms_eval);
/* Sets the initial states (opts->dt0 > 0 required) and captures one lockstep
 * step into a HIP graph. */
int oq_group_start(oq_group *e, const double *u0, double t0, double t1,
                   const oq_solver_opts *opts);
/* The same with VCABM5 for every member (as oq_ensemble_start_vcabm5). */
int oq_group_start_vcabm5(oq_group *e, const double *u0, double t0, double t1,
                          const oq_solver_opts *opts);
/* Advances up to max_steps lockstep attempts; *nrunning = members not done. */
int oq_group_run(oq_group *e, int64_t max_steps, int64_t *nrunning);
/* Per-member t, u (ntot), step counters and OQ_RET_* codes (ngeo*nmembers). */
int oq_group_state(oq_group *e, double *t, double *u, int64_t *naccept, int64_t *nreject,
                   int *retcode);
/* Checkpoint / resume as oq_ensemble_get_ctrl / oq_ensemble_resume (Tsit5
 * runs), over the ngeo*nmembers members (u in the layout of
 * oq_group_start).  A bit-for-bit continuation also needs the same tile
 * shapes (oq_group_set_shapes). */
int oq_group_get_ctrl(oq_group *e, oq_lockstep_ctrl *ctrl);
int oq_group_resume(oq_group *e, const double *u, const oq_lockstep_ctrl *ctrl, double t1,
                    const oq_solver_opts *opts);
int oq_group_destroy(oq_group *e);

#ifdef __cplusplus
}
#endif
#endif
