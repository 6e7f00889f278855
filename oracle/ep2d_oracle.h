/*
 * ep2d_oracle.h -- CPU FP64 restatement of the EffPlast2D pseudo-transient
 * elastoplastic RVE solver (reference: beacon1143/EffPlast2D).
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the parity checker for the HIP
 * product path and the timed CPU baseline.  Only tests/, the smoke() entry
 * and the bench's cpu_baseline leg may load it.  The product library never
 * links, loads or calls it.
 *
 * Every function cites the reference file:line it restates.  Arithmetic is
 * written in the same operation order as the reference source and the
 * library is compiled with -ffp-contract=off, so results are the IEEE-754
 * evaluation of the reference expressions as written (no FMA contraction).
 *
 * Array layout (reference convention, EffPlast2D.cu:3-177): A[j*m + i],
 * i along x (fastest), m = row length.
 *   cell arrays   (K, G, P0, P, tauXX, tauYY, tauXYav, J2): m = nX, nY rows
 *   vertex arrays (tauXY, J2XY):                             m = nX-1, nY-1 rows
 *   Ux, Vx: m = nX+1, nY rows;  Uy, Vy: m = nX, nY+1 rows
 */
#ifndef EP2D_ORACLE_H
#define EP2D_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

#define ORA_NPARS 11
/* pa[] layout, get_sigma_2D.m:32 */
enum {
  ORA_PA_DX = 0, ORA_PA_DY, ORA_PA_DT, ORA_PA_K0, ORA_PA_G0, ORA_PA_RHO,
  ORA_PA_DAMPX, ORA_PA_DAMPY, ORA_PA_COH, ORA_PA_RAD, ORA_PA_NPORES
};

typedef struct {
  double *K, *G, *P0;                         /* nX*nY                     */
  double *P, *tauXX, *tauYY, *tauXYav, *J2;   /* nX*nY                     */
  double *tauXY, *J2XY;                       /* (nX-1)*(nY-1)             */
  double *Ux, *Vx;                            /* (nX+1)*nY                 */
  double *Uy, *Vy;                            /* nX*(nY+1)                 */
  unsigned char *cellPore;                    /* nX*nY: ComputeStress cell pore test   */
  unsigned char *vertPore;                    /* (nX-1)*(nY-1): vertex pore test       */
} ora_fields;

typedef struct {
  long nX, nY;
  double pa[ORA_NPARS];
  int nl;            /* NL: number of load tasks (EffPlast2D.h:25-27)        */
  long nIter;        /* NITER (EffPlast2D.h:29-31)                          */
  double eIter;      /* EITER (EffPlast2D.h:33-35)                          */
  long outputStep;   /* output_step = 10'000 (EffPlast2D.cu:1608)           */
  int nThreads;      /* OpenMP threads for the stencil passes (<=0: default) */
  /* Extension (not in the reference): when useShearLoadType != 0 the third
   * task applies shearLoadType instead of loadType - spherical part. */
  int useShearLoadType;
  double shearLoadType[3];
} ora_config;

/* One record per (load step, time step), EffPlast2D.cu:314-322, 453-620. */
typedef struct {
  double epsilon[3], epsilonPer[3];
  double sigma[4], sigmaPer[4];
  double deltaP, deltaPper, tauInfty;
  double dPhi, dPhiPer;
  double holeAreaPi, internalHoleAreaPi;
  double Phi0, Phi, PhiPer;
  double error;      /* last residual evaluated in this time step            */
  double iters;      /* iterations executed in this time step                */
  double sigmaXYclean; /* extension: mean of tauXY over all (nX-1)*(nY-1) vertices */
} ora_step_record;

typedef struct { double Kphi, Kd, G; } ora_moduli;

typedef struct {
  ora_moduli an, an_per, num, num_per;   /* EffPlast2D.h:135-138 */
  /* values the reference only logs */
  double KphiExactElast, KdExactElast, GexactElast;
  double KphiExactPlast, KdExactPlast, GexactPlast;
  double KphiExactPlastPer, KdExactPlastPer, GexactPlastPer;
  /* extension: shear modulus from the xy components (used when the
   * reference xx-yy formula is 0/0, e.g. a pure xy shear task) */
  double Gxy, GxyPer;
} ora_result;

/* ---- parameters (get_sigma_2D.m:3-32) ---------------------------------- */
/* nX = nY = 32*nGrid in the reference; any nX >= 4 is accepted here. */
void ora_derive_params(double Lx, double Ly, long nX, long nY, double nPores,
                       double porosity, double Y, double K0, double G0,
                       double rho0, double CFL, double pa[ORA_NPARS]);

/* ---- setup (EffPlast2D.cu:708-783) -------------------------------------- */
void ora_set_materials(const ora_config* c, ora_fields* f);
void ora_set_pore_masks(const ora_config* c, ora_fields* f);
void ora_set_init_pressure(const ora_config* c, ora_fields* f);

/* ---- the four kernels (EffPlast2D.cu:3-177) ----------------------------- */
void ora_compute_stress(const ora_config* c, ora_fields* f);
void ora_compute_j2(const ora_config* c, ora_fields* f);
void ora_compute_plasticity(const ora_config* c, ora_fields* f);
void ora_compute_disp(const ora_config* c, ora_fields* f);
/* one full iteration, EffPlast2D.cu:390-399 (plasticity passes when nl > 1) */
void ora_iterate(const ora_config* c, ora_fields* f, long n);

/* ---- host reductions ---------------------------------------------------- */
double ora_find_max_abs(const double* arr, long size);          /* :805-813 */
double ora_find_max_abs_signed(const double* v, long size);     /* :814-822 */
double ora_residual(const ora_config* c, const ora_fields* f,
                    const double curEffStrain[3]);              /* :411-412 */
void ora_apply_affine(const ora_config* c, ora_fields* f,
                      double dUxdx, double dUydy, double dUxdy, double dUydx); /* :370-381 */
void ora_averages(const ora_config* c, const ora_fields* f, ora_step_record* r); /* :453-503 */
double ora_delta_p_honest(const ora_config* c, const ora_fields* f);   /* :825-853 */
double ora_delta_p_periodic(const ora_config* c, const ora_fields* f); /* :854-882 */
double ora_tau_infty_honest(const ora_config* c, const ora_fields* f); /* :925-942 */
int ora_pore_in_u_node(const ora_config* c, int isUy, long i, long j); /* :517-547 */
int ora_hole_area(const ora_config* c, const ora_fields* f, ora_step_record* r); /* :549-620 */

/* ---- analytic static fields (SaveAnStatic2D, EffPlast2D.cu:988-1292) ------
 * Point solutions of the single-pore elastoplastic problem (complex
 * potentials for the elastic zone around an elliptic plastic zone, the
 * logarithmic plastic-zone stress, the hydrostatic radial displacement) and
 * the field comparison against the numerical fields.  Y = pa[8]/sqrt(2)
 * (EffPlast2D.cu:1572); |.| on doubles is taken as the magnitude. */
double ora_an_ur_hydro(double r, double deltaP, double Y, double rad, double G0); /* :1030-1032 */
void ora_an_u_elast(double x, double y, double tauInfty, double xi, double kappa, double c0,
                    double Y, double rad, double G0, double K0, double* re, double* im); /* :1016-1029 */
void ora_an_j_elast(double x, double y, double xi, double kappa, double c0, double Y,
                    double rad, double* J1, double* J2);                          /* :1039-1051 */
void ora_an_j_plast(double r, double xi, double Y, double rad, double* J1, double* J2); /* :1052-1059 */
typedef struct {
  double *UnuAbs, *J1nu, *J2nu;                          /* always            */
  double *UanAbs, *errorUabs, *J1an, *J2an, *errorJ1, *errorJ2; /* nPores == 1 */
  double errorUabsMax, errorUabsAvg, errorJ1Max, errorJ1Avg, errorJ2Max, errorJ2Avg;
} ora_an_out;
/* The fields of the current state (U zeroed in the pores on a copy, as the
 * reference's host copies are at that point); arrays are caller-owned, nX*nY
 * for U, (nX-1)*(nY-1) for J; any may be NULL.  Returns 1 when the analytic
 * fields were computed (nPores == 1). */
int ora_an_static(const ora_config* c, const ora_fields* f, double deltaP, double tauInfty,
                  const double loadType[3], ora_an_out* out);

/* ---- moduli ------------------------------------------------------------- */
/* Per-step analytic moduli (EffPlast2D.cu:643-696) from one record of load
 * step `step` into res (elastic when c->nl <= 1, plastic at step 0). */
void ora_analytic_moduli(const ora_config* c, int step, const ora_step_record* r,
                         ora_result* res);
/* The numerical moduli getters (EffPlast2D.cu:1412-1539) from the records of
 * load steps 0, 1 and 2 (r1, r2 unused when NL is 1, r2 unused when NL is 2). */
void ora_moduli_getters(int NL, const ora_step_record* r0, const ora_step_record* r1,
                        const ora_step_record* r2, ora_result* res);

/* ---- drivers ------------------------------------------------------------ */
/* ComputeEffParams (EffPlast2D.cu:311-698).  recs has nTimeSteps entries. */
int ora_compute_eff_params(const ora_config* c, ora_fields* f, int step,
                           double loadStepValue, const double loadType[3],
                           long nTimeSteps, double curEffStrain[3],
                           ora_step_record* recs, ora_result* res);
/* ComputeEffModuli (EffPlast2D.cu:179-229) followed by the moduli getters
 * (:1412-1539).  recs is [nl][max(nTimeSteps,1)] row-major. */
int ora_compute_eff_moduli(const ora_config* c, ora_fields* f,
                           double initLoadValue, const double loadType[3],
                           long nTimeSteps, ora_step_record* recs,
                           ora_result* res);
/* ComputeEffDamping (EffPlast2D.cu:230-310): nl > 3 load steps. */
int ora_compute_eff_damping(const ora_config* c, ora_fields* f,
                            double initLoadValue, const double loadType[3],
                            double loadValue, long nTimeSteps,
                            ora_step_record* recs, ora_result* res);

#ifdef __cplusplus
}
#endif
#endif
