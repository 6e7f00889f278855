/*
 * ep2d.h -- C ABI of the MI355X (gfx950) EffPlast2D solver: libep2d.so.
 *
 * The library replaces the reference's CUDA class EffPlast2D
 * (EffPlast2D.h:53-198, EffPlast2D.cu:179-1658) and its four kernels
 * (EffPlast2D.cu:3-177).  The reference configures itself from compile-time
 * macros (NGRID, NL, NITER, EITER, DEVICE_IDX; EffPlast2D.h:17-39) plus the
 * binary parameter file data/pa.dat (get_sigma_2D.m:32,41-43); here every
 * one of those is a runtime field of ep2d_options / ep2d_physics.
 *
 * Conventions
 *  - Every entry point returns EP2D_OK (0) or a negative ep2d_status; the
 *    message of the last failure on the calling thread is ep2d_last_error().
 *    Nothing calls exit() (the reference's gpuAssert does, EffPlast2D.h:42-49).
 *  - Arrays use the reference layout A[j*m + i] (x fastest), FP64:
 *      cell fields   P, tauXX, tauYY, tauXYav, J2, K, G:  m = nX,   n = nY
 *      vertex fields tauXY, J2XY:                         m = nX-1, n = nY-1
 *      Ux, Vx: m = nX+1, n = nY;   Uy, Vy: m = nX, n = nY+1
 *    which is also the byte layout of the reference's *.dat files
 *    (SaveMatrix, EffPlast2D.cu:795-799; read_data_2D.m:1-12).
 *  - A handle owns its device buffers and a HIP stream of default priority.
 *    One handle must not be used from two threads at once; distinct handles
 *    are independent.  Handles are not copyable (the reference class is
 *    copyable and would double-free).
 */
#ifndef EP2D_H
#define EP2D_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define EP2D_NPARS 11   /* NPARS, EffPlast2D.h:21-23 */

/* pa[] indices: [dX, dY, dt, K0, G0, rho0, dampX, dampY, coh, rad, nPores]
 * (get_sigma_2D.m:32; read at EffPlast2D.cu:1563-1573). */
enum ep2d_pa_index {
  EP2D_PA_DX = 0, EP2D_PA_DY, EP2D_PA_DT, EP2D_PA_K0, EP2D_PA_G0, EP2D_PA_RHO,
  EP2D_PA_DAMPX, EP2D_PA_DAMPY, EP2D_PA_COH, EP2D_PA_RAD, EP2D_PA_NPORES
};

typedef enum ep2d_status {
  EP2D_OK = 0,
  EP2D_ERR_ARG = -1,        /* invalid argument (std::invalid_argument in the reference) */
  EP2D_ERR_HIP = -2,        /* HIP runtime error (gpuAssert in the reference)            */
  EP2D_ERR_IO = -3,         /* file error (std::runtime_error in ReadParams)             */
  EP2D_ERR_GEOMETRY = -4,   /* pore geometry puts a sample outside the grid               */
  EP2D_ERR_NONFINITE = -5,  /* NaN/Inf in the velocity field at a residual check          */
  EP2D_ERR_MEMORY = -6,     /* device memory budget exceeded                              */
  EP2D_ERR_STATE = -7       /* call not valid in the handle's current state               */
} ep2d_status;

/* Iteration schemes.  All three compute the reference's algorithm. */
typedef enum ep2d_mode {
  EP2D_MODE_PLAIN = 0,  /* four kernels per iteration, FP64, stress fields in HBM:
                           the reference's kernel structure (EffPlast2D.cu:390-399) */
  EP2D_MODE_FP64 = 1,   /* one fused FP64 kernel per iteration (U ping-pong)          */
  EP2D_MODE_SPLIT = 2   /* fused kernel: FP32 stress/velocity, FP64 displacement;
                           residual, averages and moduli in FP64                     */
} ep2d_mode;

/* Field identifiers for get/set/save. */
typedef enum ep2d_field {
  EP2D_FIELD_P = 0, EP2D_FIELD_TAUXX, EP2D_FIELD_TAUYY, EP2D_FIELD_TAUXY,
  EP2D_FIELD_TAUXYAV, EP2D_FIELD_J2, EP2D_FIELD_J2XY,
  EP2D_FIELD_UX, EP2D_FIELD_UY, EP2D_FIELD_VX, EP2D_FIELD_VY,
  EP2D_FIELD_K, EP2D_FIELD_G,
  EP2D_FIELD_COUNT
} ep2d_field;

/* Physical inputs from which pa[] is derived (get_sigma_2D.m:1-32 and the
 * driver scalars of eff_props_2D.m:7-26). */
typedef struct ep2d_physics {
  double Lx, Ly;        /* RVE size                                  */
  int nGrid;            /* nX = nY = 32*nGrid unless nX is given     */
  int nX;               /* 0: 32*nGrid                               */
  int nPores;           /* nPores x nPores pores                     */
  double porosity;      /* -> rad = sqrt(porosity*Lx*Ly/(pi*nPores^2)) */
  double Y;             /* yield stress; coh = Y*sqrt(2)             */
  double K0, G0, rho0;  /* moduli and density                        */
  double CFL;           /* dt = CFL*min(dX,dY)/sqrt((K0+4G0/3)/rho0) */
} ep2d_physics;

/* Solver options: the reference's compile-time macros made runtime, plus
 * the scheme selection. */
typedef struct ep2d_options {
  int nTasks;           /* NL: 1 elastic, 2 bulk moduli, 3 + shear, >3 damping */
  long nIter;           /* NITER: iteration cap per time step        */
  double eIter;         /* EITER: residual target                    */
  long checkInterval;   /* output_step (reference: 10000, :1608); 0 -> 10000 */
  int mode;             /* ep2d_mode                                  */
  int device;           /* DEVICE_IDX                                 */
  int nTimeSteps;       /* time steps of the preload (EffPlast2D ctor argument) */
  /* Extension: when useShearLoadType != 0 the third task applies
   * shearLoadType instead of loadType minus its spherical part
   * (EffPlast2D.cu:188-189, 203), e.g. {0,0,1} for a pure xy shear. */
  int useShearLoadType;
  double shearLoadType[3];
  int verbose;          /* 1: print the reference's console lines to stdout */
  const char* logPath;  /* non-NULL: append the reference's log lines here
                           (EffPlast2D.log, :1607) */
  int tileRows;         /* fused kernels: tile height 6, 8, 14 or 30 (0: tuned default) */
  int marchRows;        /* > 0: row-marching kernel, segments of this many rows  */
  int itersPerLaunch;   /* with marchRows: 2 iterations per launch on wave pairs,
                           3 or 4 on chains of 3 or 4 stages, one wave per
                           stage unless stageWaves = 2 (nX, nY <= 16382:
                           fields under 2 GiB); 1: one; 0 (default) with
                           tileRows = marchRows = 0: the shape is chosen by grid
                           size and scheme (same results) */
  int adaptiveCheck;    /* extension, 0: check every checkInterval iterations;
                           1: first check after checkInterval, then where the
                           residual's decay rate over the last two checks
                           predicts eIter (between checkInterval/64 and
                           checkInterval later)                           */
  int laneColumns;      /* split scheme with itersPerLaunch 3 or 4: 2 runs the
                           wide chains (two columns per lane, 128-column
                           strips), 1 the chains of one column per lane;
                           0: chosen with the shape (same results)        */
  int keepFp32Stress;   /* split scheme, 0 (default): a time step whose residual,
                           within 1000 eIter of the target, has not halved
                           over eight checks (the FP32 stress no longer
                           resolves it) finishes with the FP64 arithmetic on
                           the same FP64 U and FP32 V; 1: never           */
  int stageWaves;       /* with itersPerLaunch 3 or 4 and one column per lane:
                           2 runs each iteration of a chain on two waves side
                           by side (124-column strips, each stage's rows
                           handed on across both), 1 on one wave; 0: chosen
                           with the shape (same results)                  */
} ep2d_options;

typedef struct ep2d_moduli { double Kphi, Kd, G; } ep2d_moduli;  /* EffPlast2D.h:120-134 */

/* Per load step and time step, the quantities ComputeEffParams stores
 * (EffPlast2D.h:107-117, EffPlast2D.cu:344-620). */
typedef struct ep2d_record {
  double epsilon[3], epsilonPer[3];
  double sigma[4], sigmaPer[4];
  double deltaP, deltaPper, tauInfty;
  double dPhi, dPhiPer;
  double holeAreaPi, internalHoleAreaPi;
  double Phi0, Phi, PhiPer;
  double error;          /* residual at the last check of this time step */
  double iters;          /* iterations executed                          */
  double sigmaXYclean;   /* extension: mean tauXY over all vertices      */
} ep2d_record;

typedef struct ep2d_result {
  ep2d_moduli an, an_per, num, num_per;      /* EffPlast2D.h:135-138 */
  double KphiExactElast, KdExactElast, GexactElast;         /* logged only */
  double KphiExactPlast, KdExactPlast, GexactPlast;
  double KphiExactPlastPer, KdExactPlastPer, GexactPlastPer;
  double Gxy, GxyPer;   /* extension: G = d(sigma_xy) / (2 d(eps_xy)) between tasks 2 and 3 */
  int nRecords;         /* records available from ep2d_get_records   */
  long totalIters;      /* iterations over all load and time steps   */
  double seconds;       /* wall time of the call                     */
  int converged;        /* 1 when every time step reached eIter      */
} ep2d_result;

typedef struct ep2d_handle ep2d_handle;

/* ---- library ------------------------------------------------------------ */
const char* ep2d_last_error(void);

/* sizeof of one of this header's structs by its name ("ep2d_options",
 * "ep2d_result", ...; 0 for an unknown name), so a binding in another
 * language can check its mirror of the layout.  No reference counterpart. */
size_t ep2d_struct_size(const char* name);
const char* ep2d_version(void);
/* Number of visible HIP devices (0 when none; never an error). */
int ep2d_device_count(void);

/* ---- parameters (host only; no device needed) ----------------------------
 * get_sigma_2D.m:3-32: derive pa[] and nX from the physical inputs. */
int ep2d_derive_params(const ep2d_physics* phys, double pa[EP2D_NPARS], int* nX);
/* ReadParams (EffPlast2D.cu:700-707) / the fwrite of get_sigma_2D.m:41-43:
 * 11 little-endian float64 values, no header. */
int ep2d_read_pa(const char* path, double pa[EP2D_NPARS]);
int ep2d_write_pa(const char* path, const double pa[EP2D_NPARS]);
/* Default options: the reference macro defaults NL=3, NITER=100000,
 * EITER=1e-10, output_step=10000, DEVICE_IDX=0, nTimeSteps=1. */
void ep2d_default_options(ep2d_options* opt);

/* ---- solver handle -------------------------------------------------------
 * EffPlast2D::EffPlast2D (EffPlast2D.cu:1541-1612): allocate and set up the
 * fields for an nX x nY grid described by pa[] (8 <= nX = nY <= 23169: every
 * field under 4 GiB). */
int ep2d_create(const double pa[EP2D_NPARS], int nX, int nY,
                const ep2d_options* opt, ep2d_handle** out);
/* The same, with every piece of the handle's device work ordered on the
 * caller's stream (a hipStream_t of device opt->device, passed as void*),
 * which the handle uses but does not own or destroy: the reference's
 * kernel-launch boundary (EffPlast2D.cu:390-399) with an explicit stream
 * instead of the legacy default stream and its device-wide synchronisation.
 * Calls that return host results synchronise that stream. */
int ep2d_create_on_stream(const double pa[EP2D_NPARS], int nX, int nY,
                          const ep2d_options* opt, void* stream, ep2d_handle** out);
/* Convenience: derive pa[] from phys, then ep2d_create. */
int ep2d_create_from_physics(const ep2d_physics* phys, const ep2d_options* opt,
                             ep2d_handle** out);
/* EffPlast2D::~EffPlast2D (EffPlast2D.cu:1613-1658). */
void ep2d_destroy(ep2d_handle* h);

/* EffPlast2D::ComputeEffModuli (EffPlast2D.cu:179-229): preload, then the
 * small hydrostatic (and deviatoric) increments, then the moduli getters
 * (:1412-1539).  loadValue is accepted and ignored, as in the reference
 * (deprecated parameter).  Results go to *out (the reference returns 0.0
 * and prints them). */
int ep2d_compute_eff_moduli(ep2d_handle* h, double initLoadValue,
                            const double loadType[3], double loadValue,
                            ep2d_result* out);
/* EffPlast2D::ComputeEffDamping (EffPlast2D.cu:230-310), nTasks > 3. */
int ep2d_compute_eff_damping(ep2d_handle* h, double initLoadValue,
                             const double loadType[3], double loadValue,
                             ep2d_result* out);
/* Records of the last compute call, [step][timeStep] row-major with
 * nTimeSteps entries for step 0 and one for each later step. */
int ep2d_get_records(ep2d_handle* h, ep2d_record* out, int capacity, int* count);

/* ---- fields ---------------------------------------------------------------*/
int ep2d_field_shape(ep2d_handle* h, int field, int* m, int* n);
int ep2d_get_field(ep2d_handle* h, int field, double* host, size_t count);
int ep2d_set_field(ep2d_handle* h, int field, const double* host, size_t count);
/* SaveMatrix (EffPlast2D.cu:795-799): raw FP64 dump of one field. */
int ep2d_save_field(ep2d_handle* h, int field, const char* path);
/* Device buffer of a field of the current state in the layout above (FP64;
 * Vx, Vy FP32 in EP2D_MODE_SPLIT), valid until the handle's next call that
 * iterates or is destroyed: the reference's raw device fields (the kernel
 * arguments of EffPlast2D.cu:3-177).  Stress fields are finalised first and
 * the handle's stream is synchronised.  K and G are not stored (EP2D_ERR_ARG). */
int ep2d_field_device(ep2d_handle* h, int field, const void** ptr);
/* Pore masks: bit0 SetMaterials pore (K, G reduced, :708-728), bit1
 * ComputeStress cell zeroing (:77-87), bit2 vertex zeroing (:88-95). */
int ep2d_get_mask(ep2d_handle* h, unsigned char* host, size_t count);

/* ---- step-level entry points (the pieces ComputeEffParams is made of) ----*/
/* n iterations of the handle's scheme, no residual checks (:389-399). */
int ep2d_iterate(ep2d_handle* h, long n);
/* U += affine field of the macro strain increment (:370-381), on device. */
int ep2d_apply_affine(ep2d_handle* h, double dUxdx, double dUydy,
                      double dUxdy, double dUydx);
/* Residual (:407-412): (max|Vx|/(dX(nX-1)) + max|Vy|/(dY(nY-1)))*dt /
 * max(|e0|, max(e1, e2)); nonfinite receives the count of NaN/Inf values. */
int ep2d_residual(ep2d_handle* h, const double curEffStrain[3], double* error,
                  long* nonfinite);
/* No counterpart in the reference: for callers that cannot allocate device
 * memory (code under HIP graph capture, serving loops), the residual of
 * ep2d_residual on velocity fields the caller owns.  Vx is (nX+1)*nY and Vy
 * nX*(nY+1) values (float when velocityF32, else double) in the reference
 * layout; dX, dY, dT as in pa.  Enqueued on `stream` (a hipStream_t, NULL:
 * the null stream): writes the residual to the device double *error and,
 * when nonfinite is not NULL, the count of NaN/Inf values to the device
 * word *nonfinite; nothing is allocated and nothing is synchronised.  The
 * reduction's scratch is a fixed device array of 4 words per GPU (32 bytes,
 * held until the process ends) that every call on that GPU shares: calls
 * that overlap on one GPU race.  nX, nY in [3, EP2D_NOALLOC_MAX_SIDE]. */
#define EP2D_NOALLOC_MAX_SIDE 32768
int ep2d_residual_noalloc(const void* Vx, const void* Vy, int nX, int nY, int velocityF32,
                          double dX, double dY, double dT, const double curEffStrain[3],
                          double* error, unsigned long long* nonfinite, void* stream);
/* Volume averages (:453-503) on the device in FP64, plus sigmaXYclean. */
int ep2d_averages(ep2d_handle* h, ep2d_record* rec);
/* getDeltaP_honest, getDeltaP_periodic, getTauInfty_honest (:825-882, :925-942). */
int ep2d_boundary(ep2d_handle* h, double* deltaP, double* deltaPper, double* tauInfty);
/* Pore-edge hole area, Phi0/Phi/PhiPer, dPhi/dPhiPer (:514-620). */
int ep2d_hole_area(ep2d_handle* h, ep2d_record* rec);

/* ---- timing ----------------------------------------------------------------
 * Sample the duration of the dominant iteration kernel with HIP events on
 * the handle's stream: every sampleEvery-th launch is bracketed (0: off). */
int ep2d_set_timing(ep2d_handle* h, long sampleEvery);
int ep2d_get_timing(ep2d_handle* h, double* avgKernelMs, long* samples,
                    const char** kernelName);
/* Measurement aid, no counterpart in the reference: runs a plain
 * double-precision device-to-device copy kernel over `bytes` bytes `reps`
 * times on `device` and reports the mean kernel time.  Used to calibrate HBM
 * traffic counters and as the achievable-bandwidth reference. */
int ep2d_diag_copy(int device, size_t bytes, int reps, double* avgMs);
/* Diagnostic, no counterpart in the reference (host only, no device call):
 * the launch order the engine gives a work list of the chain kernels of
 * stageWaves waves per stage.  segs: n entries of four ints (RVE index,
 * first column, first row, end row), dims: nRves pairs (nX, nY), resident:
 * workgroups the device holds at once.  Column-edge segments move to the
 * front of each XCD's share and, when n > resident, the end of each share is
 * cut into half- and quarter-height pieces.  *nOut receives the length of
 * the ordered list, written to out when out is not null (outCapacity
 * entries of four ints).  For tests of the order's coverage. */
int ep2d_diag_chain_order(const int* segs, int n, int stageWaves, const int* dims, int nRves,
                          int resident, int* out, int outCapacity, int* nOut);
/* Analytic static solution of the single-pore problem beside the numerical
 * fields of the current state (SaveAnStatic2D, EffPlast2D.cu:1060-1292; the
 * reference calls it after ComputeEffModuli with nTasks > 1 and after
 * ComputeEffDamping, with deltaP and tauInfty of load step 0, last time
 * step).  Caller-owned arrays, any may be NULL: nX*nY for UnuAbs, UanAbs,
 * errorUabs; (nX-1)*(nY-1) for the J arrays.  The analytic fields and the
 * error statistics need nPores == 1 (analytic = 1). */
typedef struct ep2d_an_static {
  double *UnuAbs, *J1nu, *J2nu;
  double *UanAbs, *errorUabs, *J1an, *J2an, *errorJ1, *errorJ2;
  double errorUabsMax, errorUabsAvg, errorJ1Max, errorJ1Avg, errorJ2Max, errorJ2Avg;
  int analytic;
} ep2d_an_static;
int ep2d_an_static_fields(ep2d_handle* h, double deltaP, double tauInfty,
                          const double loadType[3], ep2d_an_static* out);
/* The same, written as the reference's files <dir>/<name>_<nX>_.dat (UnuAbs,
 * J1nu, J2nu; with one pore also UanAbs, errorUabs, J1an, J2an, errorJ1,
 * errorJ2) and the error lines printed/logged as the reference does;
 * stats may be NULL. */
int ep2d_save_an_static(ep2d_handle* h, double deltaP, double tauInfty,
                        const double loadType[3], const char* dir, ep2d_an_static* stats);

/* Synchronise the handle's stream. */
int ep2d_sync(ep2d_handle* h);
/* Device bytes held by the handle. */
size_t ep2d_device_bytes(ep2d_handle* h);

/* ---- many RVEs on one GPU: batched launches with admission and backfill ----
 * Extension, no single counterpart in the reference (which runs one RVE per
 * process, eff_props_2D.m:44-46).  Each request is one ComputeEffModuli
 * (nTasks <= 3) or ComputeEffDamping (nTasks > 3) of its own RVE, with the
 * arithmetic of the single-RVE path: a result does not depend on what else is
 * in the batch.  Every launch advances all admitted RVEs by one iteration;
 * every checkInterval iterations each RVE's residual is checked, converged
 * load steps are post-processed, finished RVEs retire and queued requests are
 * admitted while the device memory budget allows. */
typedef struct ep2d_batch_options {
  int mode;              /* EP2D_MODE_FP64 or EP2D_MODE_SPLIT                 */
  long checkInterval;    /* default residual check interval of requests         */
  int tileRows;          /* fixed tile kernel of this height (6, 8, 14, 30)   */
  int marchRows;         /* fixed row-marching kernel, segment rows (> 0)      */
  int persistRows;       /* fixed persistent tile kernel of this height        */
  int march2Rows;        /* fixed two-iteration marching kernel, segment rows  */
  int pairRows;          /* fixed two-iteration marching on wave pairs, rows   */
  int tripleRows;        /* fixed three-iteration marching on chains of three
                            waves, segment rows                                */
  int quadRows;          /* fixed four-iteration marching on chains of four
                            waves, segment rows                                */
                         /* (the multi-iteration shapes, here and below:
                            requests up to nX = 16382)                         */
                         /* all 0: kernel and shape tuned on the running batch  */
                         /* (multi-iteration shapes while every RVE has nX <= 16382) */
  int device;
  size_t memoryBudget;   /* bytes of device memory for RVE state (0: 90% of free) */
  int maxActive;         /* cap on admitted RVEs (0: none)                     */
  int useGraphs;         /* capture iteration chunks in HIP graphs             */
  long timingEvery;      /* time every n-th launch with dispatch events (0: off) */
  int adaptiveCheck;     /* residual checks scheduled by the decay rate
                            (ep2d_options.adaptiveCheck) for every request   */
  int wide3Rows;         /* fixed three-iteration wide chains (two columns per
                            lane; split scheme only), segment rows          */
  int wide4Rows;         /* fixed four-iteration wide chains, segment rows   */
  int keepFp32Stress;    /* ep2d_options.keepFp32Stress for every request: an
                            RVE that finishes a time step in the FP64
                            arithmetic leaves the batched launches for the
                            rest of that step                                */
  int triple2Rows;       /* fixed three-iteration chains with two waves per
                            iteration (ep2d_options.stageWaves), segment rows */
  int quad2Rows;         /* fixed four-iteration chains with two waves per
                            iteration, segment rows                          */
} ep2d_batch_options;

typedef struct ep2d_request {
  double pa[EP2D_NPARS]; /* parameter vector of this RVE                       */
  int nX;                /* grid (nY = nX)                                     */
  int nTasks;            /* NL                                                 */
  long nIter;
  double eIter;
  int nTimeSteps;
  double initLoadValue;
  double loadType[3];
  double loadValue;      /* damping amplitude (nTasks > 3)                     */
  int useShearLoadType;
  double shearLoadType[3];
  long checkInterval;    /* residual check interval (0: the engine's)          */
  long long tag;         /* caller's identifier, returned with the result      */
} ep2d_request;

typedef struct ep2d_batch_result {
  long long id, tag;
  int status;            /* EP2D_OK or the error of this request               */
  ep2d_result result;
  double tSubmit, tAdmit, tDone;  /* seconds on the batch clock               */
  char message[160];     /* error text when status != EP2D_OK                  */
} ep2d_batch_result;

#define EP2D_MAX_CAND 32

typedef struct ep2d_batch_stats {
  long long launches, chunks, tilesLaunched, submitted, finished, failed;
  long long cellIterations;       /* sum over launches of RVE cells advanced  */
  double bytesAlgorithmic;        /* U and V read and written, all launches   */
  double timedMs, timedBytes;     /* sampled launches: duration and bytes     */
  long long timedLaunches;        /* iterations in the sampled launches       */
  double graphMs;                 /* timed graph replays: total duration      */
  long long graphLaunches;        /* launches inside those replays            */
  int active, queued;
  size_t deviceBytes;             /* device memory held by admitted RVEs      */
  int kernel;                     /* kernel in use: 1 tiles, 2 row marching,
                                     3 persistent tiles, 4 row marching with
                                     two iterations per launch, 5 the same on
                                     pairs of waves (one iteration each), 6
                                     three iterations per launch on chains of
                                     three waves, 7 four on chains of four, 8
                                     and 9 the wide chains of three and four,
                                     11 and 12 the chains of three and four
                                     stages on two waves per stage            */
  int kernelRows;                 /* its tile height or segment rows          */
  int tunings;                    /* kernel tunings run                       */
  int nCand;                      /* kernel shapes the tuner compares        */
  int candKind[EP2D_MAX_CAND];    /* candidate k: kernel (as `kernel`)       */
  int candRows[EP2D_MAX_CAND];    /* candidate k: tile height / segment rows */
  double candUs[EP2D_MAX_CAND];   /* last tuning, mean us per iteration of
                                     candidate k (0 if not timed)             */
  long long fp64Finishes;         /* time steps finished in the FP64 arithmetic
                                     (ep2d_options.keepFp32Stress)            */
  long long fp64CellIterations;   /* cells x iterations advanced in the FP64
                                     arithmetic (part of cellIterations)      */
} ep2d_batch_stats;

typedef struct ep2d_batch ep2d_batch;

void ep2d_batch_default_options(ep2d_batch_options* opt);
/* Device bytes the engine counts against memoryBudget for one request in the
 * given options' scheme (host only, no device needed): what the request's
 * RVE state allocates when admitted.  No reference counterpart. */
int ep2d_batch_request_bytes(const ep2d_batch_options* opt, const ep2d_request* rq,
                             size_t* bytes);
int ep2d_batch_create(const ep2d_batch_options* opt, ep2d_batch** out);
void ep2d_batch_destroy(ep2d_batch* b);
/* Queue a request; *id receives its identifier.  Validation errors return
 * immediately; device errors are reported in the request's result. */
int ep2d_batch_submit(ep2d_batch* b, const ep2d_request* rq, long long* id);
/* One chunk: admit, run every admitted RVE to its next residual check (the
 * nearest one over the batch), check, post-process, retire, admit.  Unless a
 * kernel is fixed, the first chunk after the batch's cell count changes by
 * more than a quarter times each kernel shape on real iterations (the shape
 * does not change results) and keeps the fastest.  Returns the number
 * admitted and queued. */
int ep2d_batch_step(ep2d_batch* b, int* nActive, int* nQueued);
/* Step until nothing is queued or admitted. */
int ep2d_batch_run(ep2d_batch* b);
/* Move up to capacity finished results into out. */
int ep2d_batch_poll(ep2d_batch* b, ep2d_batch_result* out, int capacity, int* count);
double ep2d_batch_clock(ep2d_batch* b);
int ep2d_batch_stats_get(ep2d_batch* b, ep2d_batch_stats* out);

#ifdef __cplusplus
}
#endif
#endif
