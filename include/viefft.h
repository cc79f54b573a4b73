/*
 * viefft.h -- C ABI of the MI355X (gfx950) VIE-FFT operator and Krylov solvers.
 *
 * Drop-in boundary for the reference's hot path.  The reference is one C++
 * translation unit whose entry points are void(void) functions over global
 * arrays; each entry point below replaces one of them (cited file:line,
 * relative to the reference root) with explicit arguments:
 *
 *   Green_Function_Compute   vie_fft/GreenFunction.h:8-141   -> vie_green_compute
 *   Incident_Wave            vie_fft/GreenFunction.h:144-248 -> vie_incident
 *   GetMaterial/GetFlag out  vie_fft/Model.h:305-398         -> vie_set_medium[_eps]
 *   X_FFT/Y_FFT/Z_FFT +      vie_fft/CG_Iteration.h:131-173, -> vie_apply
 *     Current_Reset            vie_fft/fft.h:13-35
 *   X_FFT/Y_FFT/Z_FFT alone  vie_fft/CG_Iteration.h:131-173   -> vie_apply_raw
 *   *_FFT_Conjugate          vie_fft/CG_Iteration.h:86-128   -> vie_apply(adjoint=1)
 *   FFT_Program              vie_fft/CG_Iteration.h:1365-1388-> vie_solve
 *   FarField1                vie_fft/NearFar.h:16-109        -> vie_forward_scatter
 *   NearField (Ene_Metal)    vie_fft/NearFar.h:180-250       -> vie_absorbed
 *   FarField2 (RCS cuts)     vie_fft/NearFar.h:302-456       -> vie_rcs, vie_far_field
 *
 * Data layout is the reference's: complex numbers are interleaved
 * {double re, double im} (MKL_Complex16, vie_fft/common.h:68-94), 3-D arrays
 * are C-order [i][j][k] with k fastest, and a field is the three components
 * x|y|z stored back to back (3*V complex, V = Nx*Ny*Nz).  Host buffers are
 * owned by the caller and never freed by the library; device buffers are
 * owned by the context.  Every call returns a vie_status; no exception
 * crosses the ABI.  One context per host thread; calls on a context are
 * synchronous unless named *_async.
 */
#ifndef VIEFFT_H
#define VIEFFT_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct vie_ctx vie_ctx;

typedef enum {
    VIE_OK = 0,
    VIE_ERR_ARG = 1,          /* bad argument or call order            */
    VIE_ERR_DEVICE = 2,       /* HIP runtime error / no device         */
    VIE_ERR_OOM = 3,          /* device allocation failed              */
    VIE_NOT_CONVERGED = 4,    /* maxit reached before the stop rule    */
    VIE_BREAKDOWN = 5,        /* zero or non-finite Krylov denominator */
    VIE_ERR_NAN = 6,          /* non-finite input                      */
    VIE_ERR_UNSUPPORTED = 7,  /* grid size without a compiled kernel   */
    VIE_DIVERGED = 8          /* divergence stop (vie_set_divergence_stop) */
} vie_status;

typedef enum { VIE_FP64 = 0, VIE_FP32 = 1 } vie_precision;

/* Iteration_Type of vie_fft/CG_Iteration.h:47 */
typedef enum {
    VIE_CG = 1,        /* CG on the normal equations (CG_FFT)     */
    VIE_BICG = 2,      /* complex-symmetric BiCG (BICG_FFT)       */
    VIE_BICGSTAB = 3,  /* BiCGSTAB (BICGSTAB_FFT), the default    */
    VIE_CORS = 4,      /* CORS (CORS_FFT)                         */
    VIE_BICORSTAB = 5  /* BiCOR-STAB (BICORSTAB_FFT)              */
} vie_method;

typedef struct {
    int n[3];          /* Nx, Ny, Nz: cells per axis (Model.h:163-165)          */
    double d[3];       /* delta_x, delta_y, delta_z (Model.h:139-160)           */
    double k0;         /* 2*pi/lambda in the same length unit (Model.h:106)     */
    int precision;     /* vie_precision of the operator and the plain solves    */
} vie_desc;

typedef struct {
    int iterations;        /* loop passes run (the reference prints iterations-1) */
    int converged;         /* method's own stop rule met                          */
    double err;            /* last value of the method's own stop quantity        */
    double true_relres;    /* ||b - A x|| / ||b||, FP64 operator, after the solve */
    int outer;             /* refinement steps (mixed solve), else 0              */
    double seconds;        /* device time of the solve                            */
} vie_info;

/* ---- context ---------------------------------------------------------- */
/* Axis lengths 1..256 are accepted.  A length without compiled pass kernels
   runs embedded at the origin of the next compiled length with background
   cells (eps 1.1, no active face) around it; every host array keeps the
   caller's n, and results are those of the caller's grid (the padded period
   only grows).  Such grids need no active face on a component's own-axis last
   plane (the 3-point flag rule never makes one).  Longer axes:
   VIE_ERR_UNSUPPORTED. */
int vie_create(vie_ctx** ctx, const vie_desc* desc);
int vie_destroy(vie_ctx* ctx);
const char* vie_last_error(const vie_ctx* ctx);
/* 1 when n has compiled pass kernels (products of 2, 3 and 5 from a fixed list
   up to 256; the padded transform length is 2n), 0 when n runs embedded. */
int vie_supported_n(int n);

/* ---- medium ------------------------------------------------------------- */
/* material: V bytes (Material, vie_fft/Model.h:17), eps_table: n_mat complex
   (Real_EPR/Imag_EPR, vie_fft/Model.h:82-83, engineering sign), flags: 3*V
   bytes Flag_X|Flag_Y|Flag_Z (vie_fft/Model.h:21-23).  The library may order
   its internal grid differently (the axis whose active faces span the fewest
   planes first, see vie_plan); every array at this interface keeps the
   caller's layout.  A medium with another internal order drops the resident
   right-hand side and solution. */
int vie_set_medium(vie_ctx* ctx, const unsigned char* material, const double* eps_table,
                   int n_mat, const unsigned char* flags);
/* per-voxel permittivity, V complex */
int vie_set_medium_eps(vie_ctx* ctx, const double* eps_voxel, const unsigned char* flags);
/* A call with the material ids (or, for vie_set_medium_eps, the flags) of the
   resident medium -- a wavelength sweep with dispersive tables -- only
   rebuilds the permittivity and the diagonal term on the device. */

/* Analytic shapes voxelised on the device with the reference's rules
   (GetMaterial / GetFlag, vie_fft/Model.h:305-398): material ids at the
   cell centres (i+1/2, j+1/2, k+1/2) in cell units, the first shape holding
   a point gives its id (0 = background), a sphere holds a point when
   |p - c|^2 <= r^2 (Model.h:229), a z band when z0 <= z < z1, and points
   outside the box [0, n] are background; a face is active when its centre
   and the centres of the two cells it joins are all non-background
   (Model.h:345-394).  material (V bytes, may be NULL) and flags (3 V bytes)
   are written in the caller's layout. */
enum { VIE_SPHERE = 0, VIE_ZBAND = 1 };
typedef struct {
    int kind;        /* VIE_SPHERE or VIE_ZBAND            */
    int mat;         /* material id 0..255                  */
    double c[3];     /* sphere centre (cells)               */
    double r;        /* sphere radius (cells)               */
    double z0, z1;   /* band [z0, z1) (cells)               */
} vie_shape;
int vie_voxelize(vie_ctx* ctx, int nshape, const vie_shape* shapes, unsigned char* material,
                 unsigned char* flags);
/* vie_voxelize then vie_set_medium with the eps table */
int vie_set_medium_shapes(vie_ctx* ctx, int nshape, const vie_shape* shapes, const double* eps_table,
                          int n_mat);

/* ---- set-up ------------------------------------------------------------- */
/* Green_Function_Compute (vie_fft/GreenFunction.h:8-141): computed on the
   device; only the octant that the even extension makes unique is kept. */
int vie_green_compute(vie_ctx* ctx);

/* Change the free-space wavenumber of an existing context (one grid swept over
   wavelengths, as the reference's frequency loop does by re-running the
   program per wavelength, vie_fft/main.cpp:27-36): the spectral tables and the
   diagonal term follow; G~ must be recomputed (vie_green_compute) and the
   right-hand side set again.  The medium stays; call vie_set_medium again for
   dispersive materials. */
int vie_set_k0(vie_ctx* ctx, double k0);
/* Incident_Wave (vie_fft/GreenFunction.h:144-248); angles in degrees;
   E: host, 3V complex, written unmasked like Ex/Ey/Ez_inc. */
int vie_incident(vie_ctx* ctx, double theta, double phi, double psi, double* E);

/* ---- operator ----------------------------------------------------------- */
/* y = A x (adjoint=0) or A^H x (adjoint=1) including the output mask;
   host buffers, 3V complex each, FP64 interface whatever the precision. */
int vie_apply(vie_ctx* ctx, const double* x, double* y, int adjoint);
/* y = A x or A^H x without the output mask: the products of X_FFT/Y_FFT/
   Z_FFT and of the _Conjugate set before Current_Reset
   (vie_fft/CG_Iteration.h:86-173; the reference's solvers mask after each,
   e.g. vie_fft/CG_Iteration.h:643-649) */
int vie_apply_raw(vie_ctx* ctx, const double* x, double* y, int adjoint);
/* same on device buffers of the context's precision (complex<double> for
   VIE_FP64, complex<float> for VIE_FP32), enqueued on `stream` (a hipStream_t,
   NULL = the context's stream) without a host synchronisation; both buffers
   16-byte aligned. */
int vie_apply_async(vie_ctx* ctx, const void* x_dev, void* y_dev, int adjoint, void* stream);

/* ---- solve (FFT_Program, vie_fft/CG_Iteration.h:1365-1388) --------------- */
/* b is masked by the flags first (Current_Reset of E_inc, e.g.
   vie_fft/CG_Iteration.h:559-562); x starts from 0 like Current_* does
   (vie_fft/Model.h:189-191).  maxit <= 0 means V (the reference's limit). */
int vie_solve(vie_ctx* ctx, int method, double tol, int maxit, const double* b, double* x,
              vie_info* info);
/* resident variant: b is uploaded once by vie_set_rhs and the solution stays
   on the device (vie_get_solution copies it out); nothing crosses PCIe inside. */
int vie_set_rhs(vie_ctx* ctx, const double* b);
/* Several right-hand sides on one operator (e.g. the polarisations of one
   wavelength): B and X hold nrhs fields of 3 V complex back to back.
   strategy VIE_BATCH_PLAIN solves each with `method` in the context's
   precision, VIE_BATCH_MIXED with mixed-precision BiCGSTAB (reliable updates;
   needs an FP64 context).  The right-hand sides are solved together as
   batches of as many as the device holds (vie_batch_solve; this reserves the
   batch storage); infos (may be NULL) receives nrhs records.  Returns
   VIE_OK, or the first non-OK solver status (VIE_NOT_CONVERGED /
   VIE_BREAKDOWN) after finishing every right-hand side, or a hard error at
   once. */
typedef enum { VIE_BATCH_PLAIN = 0, VIE_BATCH_MIXED = 1 } vie_batch_strategy;
int vie_solve_batched(vie_ctx* ctx, int nrhs, int strategy, int method, double tol, int maxit, const double* B,
                      double* X, vie_info* infos);

/* ---- batches: many instances of one geometry in one set of launches ----
   The reference's only parallelism is independent frequencies, one process
   each (vie_fft/main.cpp:27-36, re-running vie_fft/CG_FFT_VIE.cpp:45-61 per
   wavelength).  Here the instances of one medium's geometry -- wavelengths,
   polarisations, incidence angles -- are solved together: every pass and
   vector launch of a batched solve covers all instances still iterating,
   each with its own solver scalars and done flag, so an instance that
   converges (or reaches its limit) stops costing work without stopping the
   others.  Instance 0 is the resident right-hand side / solution of the
   single-instance calls.
   vie_batch_reserve: device storage for nb instances (VIE_ERR_OOM if it does
   not fit; vie_batch_capacity estimates the largest nb); drops resident
   right-hand sides, solutions and per-instance operators, and returns to one
   operator (the context's k0 and permittivity) for every instance. */
int vie_batch_reserve(vie_ctx* ctx, int nb);
int vie_batch_capacity(vie_ctx* ctx, int* nb_max);
/* Per-instance operators for instances [i0, i0 + count): wavenumbers k0[]
   and, for a medium set with material ids, permittivity tables eps_tables
   (count x n_mat complex, engineering sign like Real_EPR/Imag_EPR; NULL: the
   context's permittivity).  Green_Function_Compute (vie_fft/GreenFunction.h:
   8-141) for all count wavenumbers runs in one set of launches; the
   diagonal term (vie_fft/fft.h:405-588) follows each table.  Needs
   vie_set_medium first (a new medium drops them). */
int vie_batch_set_operators(vie_ctx* ctx, int i0, int count, const double* k0, const double* eps_tables,
                            int n_mat);
/* Instance i's right-hand side: the masked plane wave at its wavenumber
   (Incident_Wave, vie_fft/GreenFunction.h:144-248), or a host field (3 V
   complex, masked like vie_set_rhs). */
int vie_batch_set_rhs_incident(vie_ctx* ctx, int i, double theta, double phi, double psi);
int vie_batch_set_rhs(vie_ctx* ctx, int i, const double* b);
/* Solve instances 0..nb-1 together: VIE_BATCH_PLAIN runs `method` in the
   context's precision, VIE_BATCH_MIXED the mixed-precision BiCGSTAB of
   vie_solve_mixed_resident (reliable updates, FP64 context).  tols / maxits
   (may be NULL) give each instance its own tolerance and limit, else tol /
   maxit (<= 0: V).  infos (may be NULL) receives nb records.  Returns VIE_OK,
   or the first instance's non-OK solver status after every instance ended. */
int vie_batch_solve(vie_ctx* ctx, int nb, int strategy, int method, double tol, int maxit, const double* tols,
                    const int* maxits, vie_info* infos);
int vie_batch_get_solution(vie_ctx* ctx, int i, double* x);
/* vie_forward_scatter / vie_absorbed of instance i's solution, at its own
   wavenumber and permittivity */
int vie_batch_forward_scatter(vie_ctx* ctx, int i, double theta, double phi, double psi, double* value);
int vie_batch_absorbed(vie_ctx* ctx, int i, int mat, double* value);
/* Both of the above for instances 0..nb-1 with one read-back (no counterpart
   in the reference, whose processes each print their own): theta, phi, psi
   (nb each, degrees) per instance; forward[i] = Forward_RCS; absorbed (may be
   NULL, else nb) = Ene_Metal of material mat. */
int vie_batch_outputs(vie_ctx* ctx, int nb, const double* theta, const double* phi, const double* psi, int mat,
                      double* forward, double* absorbed);

/* Resident right-hand side computed on the device: Incident_Wave
   (vie_fft/GreenFunction.h:144-248) for angles theta, phi, psi (degrees)
   masked by the flags (vie_fft/CG_Iteration.h:559-562); no host transfer. */
int vie_set_rhs_incident(vie_ctx* ctx, double theta, double phi, double psi);
int vie_solve_resident(vie_ctx* ctx, int method, double tol, int maxit, vie_info* info);
/* Divergence stop (off by default, as in the reference, whose solvers run
   to Nx*Ny*Nz iterations whatever the residual does, e.g.
   vie_fft/CG_Iteration.h:591,734): with factor >= 1 every later solve of the
   context ends with VIE_DIVERGED once the method's stop quantity exceeds
   factor times the smallest value it reached; 0 turns it off. */
int vie_set_divergence_stop(vie_ctx* ctx, double factor);
/* Jacobi preconditioner of the BiCGSTAB solves (no counterpart in the
   reference, whose BICGSTAB_FFT, vie_fft/CG_Iteration.h:531-776, is
   unpreconditioned; VIE_PRECOND_NONE, the default, keeps its iterates).  With
   VIE_PRECOND_JACOBI every later BiCGSTAB solve of the context -- vie_solve,
   vie_solve_x0, vie_solve_resident and the mixed solve -- iterates on the
   right-preconditioned system A (m (.) y) = b with m = 1/d_f, the reciprocal
   of the diagonal face term (LXX_DO/LYY_DO/LZZ_DO, vie_fft/fft.h:405-588),
   and returns x = m (.) y: the stop rule still reads ||b - A x|| / ||b||, so
   the solution meets the same tolerance in fewer iterations on media whose
   permittivities differ strongly.  Other methods, batches, and media with an
   active face whose d_f vanishes return VIE_ERR_ARG. */
enum { VIE_PRECOND_NONE = 0, VIE_PRECOND_JACOBI = 1 };
int vie_set_preconditioner(vie_ctx* ctx, int kind);

/* ---- device memory ------------------------------------------------------ */
/* The library keeps the device buffers of 64 MB and more that a context frees
   (up to a third of the device's memory) and reuses them for later requests
   of the same size, so contexts and batches created one after another (sweeps)
   allocate nothing; kept buffers still count as used for hipMemGetInfo.  This
   call (no counterpart in the reference) hands the current device's kept
   buffers back to the HIP runtime, for callers that need the memory for other
   allocations; *released (may be NULL) receives the byte count. */
int vie_release_buffers(size_t* released);
/* Warm start (the reference always starts from x = 0, vie_fft/Model.h:189-191):
   the next vie_solve_resident / vie_solve_mixed_resident starts from x0
   (host, 3 V complex, masked by the flags like the solution), or from the
   resident solution when x0 is NULL (a neighbouring wavelength's, in a
   sweep).  The stop rule still measures ||b - A x|| / ||b||; a guess that
   already meets it returns at once (0 iterations).  Not for VIE_BICG, whose
   stop quantity is not a residual norm. */
int vie_set_guess(vie_ctx* ctx, const double* x0);
/* vie_solve with x as the initial guess on entry and the solution on exit */
int vie_solve_x0(vie_ctx* ctx, int method, double tol, int maxit, const double* b, double* x, vie_info* info);
int vie_get_solution(vie_ctx* ctx, double* x);
/* Mixed precision: the Krylov recurrences, FFTs and spectral product run in
   FP32 complex while the solution accumulates in FP64 and the residual that
   decides convergence is the true FP64 residual b - A64 x (stop rule
   ||b - A64 x|| / ||b|| < tol).  Needs a context created with VIE_FP64.
     VIE_MIXED_RELIABLE (BiCGSTAB only): one FP32 BiCGSTAB whose residual is
       replaced by the FP64 residual whenever it has fallen by inner_param
       (<= 0: 0.1) since the last replacement, without restarting;
     VIE_MIXED_REFINE: classic iterative refinement, each step an FP32 solve
       (BiCGSTAB, CORS or BiCORSTAB) of A e = r to inner_param (<= 0: 1e-4)
       followed by x += e in FP64.
   maxit bounds the FP32 iterations; info->outer counts residual replacements
   or refinement steps. */
typedef enum { VIE_MIXED_REFINE = 0, VIE_MIXED_RELIABLE = 1 } vie_mixed_strategy;
int vie_solve_mixed_resident(vie_ctx* ctx, int strategy, int inner_method, double tol, double inner_param,
                             int maxit, vie_info* info);

/* ---- post-processing ----------------------------------------------------- */
/* Forward_RCS of FarField1 (vie_fft/NearFar.h:16-109) of the resident
   solution; divide by the geometric cross-section for Q_ext. */
int vie_forward_scatter(vie_ctx* ctx, double theta, double phi, double psi, double* value);
/* Ene_Metal of NearField (vie_fft/NearFar.h:180-250) for material id `mat`
   (needs vie_set_medium); times Z0 / area = Q_abs. */
int vie_absorbed(vie_ctx* ctx, int mat, double* value);

/* Far-field amplitudes of the resident solution for ndir directions (degrees):
   S[3q + c] = dV exp(j k0 delta_c v_c / 2) sum_cells J_c exp(j k0 r.v), r the
   cell centre, v the unit direction -- the sums of FarField1 and FarField2
   (vie_fft/NearFar.h:59-96, 321-380); S holds ndir x 3 complex (interleaved
   doubles).  One pass over J serves up to 1024 directions. */
int vie_far_field(vie_ctx* ctx, int ndir, const double* theta_deg, const double* phi_deg, double* S);

/* vie_far_field without any device allocation (no counterpart in the
   reference): for callers that cannot allocate device memory -- code under
   HIP graph capture, tight serving loops, harnesses that forbid hipMalloc.
   Its per-block partials, sums and directions live in a fixed 6.3 MB static
   device array (one per GPU, held until the process ends), so a call takes
   at most 512 directions (more: VIE_ERR_ARG).  Calls on one GPU share that
   array: calls that overlap on one GPU race.  Same S as vie_far_field (its
   sums in another block order, so equal to roundoff). */
int vie_far_field_noalloc(vie_ctx* ctx, int ndir, const double* theta_deg, const double* phi_deg, double* S);

/* FarField2 (vie_fft/NearFar.h:302-456): RCS in dB relative to lambda^2 of the
   E plane (phi_s = phi, theta component) and the H plane (phi_s = phi + 90,
   phi component) for scattering angles 0..180 degrees (181 values each). */
int vie_rcs(vie_ctx* ctx, double phi_deg, double* rcs_theta, double* rcs_phi);

/* Absorbed power of the resident solution over every active face (flag set,
   own-axis index < N_f - 1): P = sum w_f |J_f|^2, w_f = dV Z0/(2 k0)
   Im(1/(eps_a - 1) + 1/(eps_b - 1)) for the two cells a, b of the face.  This
   is the per-face term of NearField's Ene_Metal (vie_fft/NearFar.h:186-250)
   without its material and i,j,k < N-1 selection.  Needs eps and flags that
   are 0 on own-axis last planes (the 3-point flag rule gives that). */
int vie_absorption(vie_ctx* ctx, double* value);

/* Adjoint gradient of vie_absorption with respect to the permittivity of every
   cell (no reference counterpart; the design-gradient workload): one more
   solve with the same operator, A nu = delta_f w conj(J), by `method` to
   `tol` (diag(1/delta_f) A is complex symmetric), then one gather over the six
   faces of each cell.  grad_re / grad_im (host, V doubles each; grad_im may be
   NULL) receive dP/dRe(eps) and dP/dIm(eps); *value (may be NULL) receives P;
   info describes the adjoint solve.  The resident solution must solve A J = b
   for the current medium (vie_solve_resident or vie_solve before). */
int vie_absorption_grad(vie_ctx* ctx, int method, double tol, int maxit, double* value, double* grad_re,
                        double* grad_im, vie_info* info);
/* vie_absorption_grad with the adjoint system solved by the mixed-precision
   BiCGSTAB of vie_solve_mixed_resident (FP32 operator and recurrences, FP64
   solution, residual replacements and stop rule ||g - A nu|| / ||g|| < tol;
   FP64 contexts).  info->outer = residual replacements, info->true_relres and
   info->err = the adjoint solve's true FP64 residual.  No reference
   counterpart (the design-gradient workload). */
int vie_absorption_grad_mixed(vie_ctx* ctx, double tol, int maxit, double* value, double* grad_re,
                              double* grad_im, vie_info* info);

/* ---- instrumentation ----------------------------------------------------- */
/* The last single solve's stop quantity after each iteration (the values the
   reference prints as "CG Error", e.g. vie_fft/CG_Iteration.h:737-747): the
   first min(max, count) into err, *count = iterations recorded (at most
   65536).  For a mixed solve these are the FP32 recursive residuals. */
int vie_history(vie_ctx* ctx, int max, double* err, int* count);
/* Per-launch HIP-event timing of one kernel class while solving:
   0 off, 1 z-forward, 2 y-forward, 3 x spectral pass, 4 y-inverse,
   5 z-inverse of the FP64 operator, 6 FP64 vector update (linear
   combination + dots), 7 solver scalar step; 11-16 the same kernels in FP32
   (17 = 7).  Every 8th launch of the class is bracketed by events (timing
   each one would slow the solve being measured).  vie_kernel_time returns the
   timed launches and their mean milliseconds (launches enqueued after
   convergence, which exit at once, are left out). */
int vie_time_kernel(vie_ctx* ctx, int kernel_class);
int vie_kernel_time(vie_ctx* ctx, int* launches, double* mean_ms);
/* bytes the algorithmic model charges one operator application inside a
   solve (the passes over the x window, the x pass on its pencil length) */
double vie_model_bytes_matvec(const vie_ctx* ctx);
/* Internal plan of the resident medium (no counterpart in the reference; for
   profiling and tests): out[0..2] internal axis order (internal axis a is the
   caller's axis out[a]); out[3], out[4] the x window [x0, x1) of the products
   inside solves, on the internal grid; out[5] the pencil length of their x
   pass (the short period, or the internal Nx); out[6..8] the internal grid;
   out[9], out[10] the y window [y0, y1) of the products inside solves;
   out[11] their y pencil length (the short y period, or the internal Ny);
   out[12], out[13] their z window [z0, z1) (z0 even); out[14] their z
   pencil length (the short z period, or the internal Nz).  out holds 15
   ints. */
int vie_plan(const vie_ctx* ctx, int* out);
/* The autotuner's plan of the FP64 (out[0..4], us[0]) and FP32 (out[5..9],
   us[1]) operators (no counterpart in the reference, whose MKL plans are
   built per call, vie_fft/fft.h:65-73): tuned (1; 0 fixed defaults, -1 no
   such operator), pass variants (bit 0 register z forward, bit 1 LDS-staged
   y/z passes, bit 2 LDS-staged x pass; -1 defaults), the x, y and z pencil
   lengths of the windowed products; us (may be NULL, 4 values) the measured
   microseconds per product of the FP64 and FP32 plans, then of the fixed
   defaults (the tuner's starting point).  vie_set_medium picks the plan by timing the
   candidates on the device once per geometry and process (VIE_TUNE=0: the
   fixed defaults; VIE_TUNE_CACHE=file keeps plans across processes). */
int vie_tuning(const vie_ctx* ctx, int* out, double* us);

#ifdef __cplusplus
}
#endif
#endif
