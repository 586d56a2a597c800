/*
 * oracle/oracle.h -- TEST INFRASTRUCTURE.  The CPU FP64 oracle for the
 * Oetqf.jl earthquake-cycle hot path: mesh generation, Okada and hex8
 * Green's functions, the ODE right-hand side and a Tsit5 integrator.
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg load
 * this library, and only as the checker.  The product (the HIP library under
 * oetqf.jl_amd/) never links or calls it.
 *
 * Array layouts follow the reference (Julia, column-major):
 *   fault fields   nx * nxi, strike index fastest        (equation.jl:35-42)
 *   mantle fields  ne * 6, component-major [xx xy xz yy yz zz] (equation.jl:207-222)
 *   Toeplitz K_ff  st[i + nx*j + nx*nxi*l]                (GF.jl:41,56)
 *   dense blocks   Julia column-major                     (GF.jl:140,202,258)
 */
#ifndef OQ_ORACLE_H
#define OQ_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Degree trig with exact values at multiples of 90 degrees, as Julia's
 * sind/cosd (used at mesh.jl:48 and GF.jl:79,86,91,95). */
double oq_sind(double deg);
double oq_cosd(double deg);

/* Okada (1992) DC3D.  Returns 0, 1 (receiver on a fault edge: u = 0) or 2
 * (z > 0). */
int oq_dc3d(double alpha, double x, double y, double z, double depth,
            double dip, double al1, double al2, double aw1, double aw2,
            double disl1, double disl2, double disl3, double u[12]);

/* Stress at (x,y,z) due to a uniform eigenstrain (tensor components
 * e = [xx xy xz yy yz zz]) in the cuboid [qx-T/2,qx+T/2] x [qy,qy+L] x
 * [qz-W,qz] of a half-space z <= 0 (reference call GF.jl:215-221,277-283,
 * "T-L-W <=> x-y-z").  s = [xx xy xz yy yz zz].  nquad is the Gauss order per
 * axis used for the smooth image part. */
void oq_stress_vol_hex8(double x, double y, double z, double qx, double qy,
                        double qz, double T, double L, double W,
                        const double e[6], double mu, double nu, int nquad,
                        double s[6]);

/* ---------------- mesh (mesh.jl) ---------------- */
/* gen_mesh(Val(:RectOkada), ...) (mesh.jl:39-56).  Fills up to cap entries;
 * returns nx or nxi. */
int64_t oq_range_len(double start, double stop, double step);
void oq_mesh_strike(double xlen, double dx, int64_t nx, double *x, double *ax);
void oq_mesh_downdip(double xilen, double dxi, double dip, int64_t nxi,
                     double *xi, double *axi, double *y, double *z);
/* Structured hex8 box equivalent of gen_gmsh_mesh + gen_mesh(:BEMHex8Mesh)
 * (mesh.jl:95-186), cells ordered x fastest, then y, then layer. */
void oq_mesh_hex8(double llx, double lly, double llz, double dx, double dy,
                  double dz, int64_t nx, int64_t ny, int64_t nz,
                  const double *rfzh, double *cx, double *cy, double *cz,
                  double *qx, double *qy, double *qz, double *ddx, double *ddy,
                  double *ddz);

/* ---------------- Green's functions (GF.jl) ---------------- */
typedef struct {
    const double *x;   /* nx strike centres */
    const double *ax;  /* 2*nx interleaved (lo, hi) */
    const double *y;   /* nxi */
    const double *z;   /* nxi */
    const double *axi; /* 2*nxi interleaved (lo, hi) */
    int64_t nx, nxi;
    double dx, dep, dip;
} oq_fault_mesh_t;

typedef struct {
    const double *cx, *cy, *cz, *qx, *qy, *qz, *dx, *dy, *dz;
    int64_t ne;
} oq_hex8_mesh_t;

/* ftype: 0 = StrikeSlip, 1 = DipSlip (GF.jl:3-5). */
void oq_gf_ff(const oq_fault_mesh_t *m, double lambda, double mu, int ftype,
              int64_t nrept, double buffer_ratio, double *st);
void oq_gf_mf(const oq_fault_mesh_t *mf, const oq_hex8_mesh_t *ma,
              double lambda, double mu, int ftype, const double *qc,
              const double *qw, int64_t nq, int64_t nrept,
              double buffer_ratio, double *st);
void oq_gf_fm(const oq_hex8_mesh_t *ma, const oq_fault_mesh_t *mf,
              double lambda, double mu, int ftype, int nquad, double *st);
void oq_gf_mm(const oq_hex8_mesh_t *ma, double lambda, double mu,
              const double *qc, const double *qw, int64_t nq, int nquad,
              double *st);
/* Rows of K_mm for receiver cells rcv over source cells src:
 * (6 nr) x (6 ns) column-major; oq_gf_mm is rcv = src. */
void oq_gf_mm_rows(const oq_hex8_mesh_t *src, const oq_hex8_mesh_t *rcv,
                   double lambda, double mu, const double *qc, const double *qw,
                   int64_t nq, int nquad, double *st);

/* ---------------- RHS (equation.jl) ---------------- */
typedef struct {
    const double *a, *b, *L, *sigma; /* N_f each */
    double eta, vpl, f0, v0;
} oq_rs_t;

typedef struct {
    const double *gamma; /* ne */
    const double *n;     /* ne, power - 1 */
    double deps0[6];
} oq_visc_t;

/* DilatancyProperty (property.jl:27-32): per patch. */
typedef struct {
    const double *tp, *eps, *beta, *p0;
} oq_dila_t;

/* update_fault! (equation.jl:233-246) given dtau. */
void oq_update_fault(const oq_rs_t *p, int64_t nf, const double *v,
                     const double *theta, const double *dtau, double *dv,
                     double *dtheta, double *ddelta);
/* update_fault_with_dilatancy! (equation.jl:248-276) given dtau. */
void oq_update_fault_dila(const oq_rs_t *p, const oq_dila_t *d, int64_t nf,
                          const double *v, const double *theta, const double *pp,
                          const double *dtau, double *dv, double *dtheta,
                          double *ddelta, double *dpp);
/* Fault-only RHS with dilatancy (equation.jl:168-183). */
void oq_rhs_fault_dila(const oq_rs_t *p, const oq_dila_t *d, const double *st,
                       int64_t nx, int64_t nxi, const double *v, const double *theta,
                       const double *pp, double *dv, double *dtheta, double *ddelta,
                       double *dpp);
/* update_strain_rate! (equation.jl:207-222). */
void oq_update_strain_rate(const oq_visc_t *p, int64_t ne, const double *sigma,
                           double *deps);
/* Toeplitz-by-DFT dtau = K_ff s (equation.jl:44-61), with st real Toeplitz
 * kernel; the DFT of length 2nx-1 is formed here. */
void oq_dtau_fft(const double *st, int64_t nx, int64_t nxi, const double *s,
                 double *dtau);
/* Same product by direct Toeplitz sums (the test's dense contraction,
 * test/BEM/tests.jl:46-58). */
void oq_dtau_toeplitz(const double *st, int64_t nx, int64_t nxi,
                      const double *s, double *dtau);
/* Fault-only RHS (equation.jl:156-166). */
void oq_rhs_fault(const oq_rs_t *p, const double *st, int64_t nx, int64_t nxi,
                  const double *v, const double *theta, double *dv,
                  double *dtheta, double *ddelta);
/* Coupled RHS (equation.jl:185-205); gf12 6ne x nf, gf21 nf x 6ne,
 * gf22 6ne x 6ne, all column-major. */
void oq_rhs_coupled(const oq_rs_t *pf, const oq_visc_t *pa, const double *st,
                    const double *gf12, const double *gf21, const double *gf22,
                    int64_t nx, int64_t nxi, int64_t ne, const double *v,
                    const double *theta, const double *sigma, double *dv,
                    double *dtheta, double *deps, double *dsigma,
                    double *ddelta);
/* Same RHS with a dense receiver-major N x N operator (row r = receiver).
 * The precision variants of the GPU path are checked against this. */
void oq_rhs_dense(const oq_rs_t *pf, const oq_visc_t *pa, const double *K,
                  int64_t nf, int64_t ne, const double *v, const double *theta,
                  const double *sigma, double *dv, double *dtheta,
                  double *deps, double *dsigma, double *ddelta);

/* ---------------- Tsit5 (OrdinaryDiffEq semantics) ---------------- */
typedef struct {
    double reltol, abstol, dtmax, dtmin;
    double dt0; /* <= 0: automatic initial step */
    int64_t maxiters;
} oq_ode_opts_t;

typedef struct {
    int64_t naccept, nreject, nf;
    double t, dt;
    int retcode; /* 0 success, 1 maxiters, 2 dt < dtmin, 3 non-finite */
} oq_ode_stats_t;

typedef void (*oq_rhs_fn)(void *ctx, double t, const double *u, double *du);
/* Integrates u (length n) from t0 to t1 in place. */
void oq_tsit5(oq_rhs_fn f, void *ctx, int64_t n, double *u, double t0,
              double t1, const oq_ode_opts_t *o, oq_ode_stats_t *st);
void oq_tsit5_record(oq_rhs_fn f, void *ctx, int64_t n, double *u, double t0,
                     double t1, const oq_ode_opts_t *o, oq_ode_stats_t *st,
                     int64_t nrec_n, double *step_t, double *step_v, int64_t cap,
                     int64_t *nrec);
const double *oq_tsit5_tableau(void); /* c[7], a[7][7], b[7], btilde[7] */

/* VCABM5 (variable-coefficient Adams-Bashforth-Moulton, order 5, PECE),
 * vcabm.c; same interface and recorder as oq_tsit5_record. */
void oq_vcabm5(oq_rhs_fn f, void *ctx, int64_t n, double *u, double t0,
               double t1, const oq_ode_opts_t *o, oq_ode_stats_t *st);
void oq_vcabm5_record(oq_rhs_fn f, void *ctx, int64_t n, double *u, double t0,
                      double t1, const oq_ode_opts_t *o, oq_ode_stats_t *st,
                      int64_t nrec_n, double *step_t, double *step_v, int64_t cap,
                      int64_t *nrec);
/* g[0..k] and beta[0..k-1] of a step h with past steps hist[0..k]. */
void oq_vcabm_coeffs(int k, double h, const double *hist, double *g, double *beta);

/* A whole earthquake cycle in C: Tsit5 over oq_rhs_dense (packed state
 * [v θ δ] or [v θ ϵ σ δ]), recording (t, max v) per accepted step. */
void oq_cycle_dense(const oq_rs_t *pf, const oq_visc_t *pa, const double *K,
                    int64_t nf, int64_t ne, double *u, double t0, double t1,
                    const oq_ode_opts_t *o, oq_ode_stats_t *st, double *step_t,
                    double *step_v, int64_t cap, int64_t *nrec);
/* The same with alg = 0 (Tsit5) or 1 (VCABM5). */
void oq_cycle_dense_alg(const oq_rs_t *pf, const oq_visc_t *pa, const double *K,
                        int64_t nf, int64_t ne, double *u, double t0, double t1,
                        const oq_ode_opts_t *o, oq_ode_stats_t *st, double *step_t,
                        double *step_v, int64_t cap, int64_t *nrec, int alg);

#ifdef __cplusplus
}
#endif
#endif
