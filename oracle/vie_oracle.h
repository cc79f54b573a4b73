/*
 * vie_oracle.h -- CPU FP64 restatement of the VIE-FFT hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP product
 * path: only tests, the smoke check and the bench's cpu_baseline leg may load
 * it.  Nothing in the product library links or calls it.
 *
 * Conventions (all follow the reference, paths relative to its root):
 *   - complex data are interleaved {re, im} doubles (MKL_Complex16 layout,
 *     vie_fft/common.h:68-94), 3-D arrays are C-order [i][j][k], k fastest;
 *   - a "field" is three components x|y|z stored back to back (3*V complex,
 *     V = Nx*Ny*Nz), the same order as Current_X/Y/Z (vie_fft/Model.h:13-15);
 *   - engineering time convention exp(+j w t): loss is negative Im(eps),
 *     the Green's function is exp(-j k0 R) (vie_fft/GreenFunction.h:65-67);
 *   - Z0 = 120*pi, c = 3e8 (vie_fft/common.h:17-19).
 */
#ifndef VIE_ORACLE_H
#define VIE_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    int n[3];      /* Nx, Ny, Nz (cells)                                   */
    double d[3];   /* delta_x, delta_y, delta_z (length unit = 100 nm)     */
    double k0;     /* free-space wavenumber 2*pi/lambda (same length unit) */
} orc_grid;

/* ---- FFT (unnormalised; sign -1: exp(-j..) forward, +1: backward) ------- */
void orc_fft1d(double* a, int n, int sign);
void orc_fft3d(double* a, int p0, int p1, int p2, int sign);

/* ---- set-up ----------------------------------------------------------- */
/* g on the (Nx+1)(Ny+1)(Nz+1) offsets, C-order (vie_fft/GreenFunction.h:53-88) */
void orc_green_offsets(const orc_grid* g, double* goff);
/* G~ = forward DFT of the even extension of g to P = 2N per axis
   (vie_fft/GreenFunction.h:8-141); G has 8V complex entries */
void orc_green(const orc_grid* g, double* G);
/* plane-wave right-hand side, 3V complex (vie_fft/GreenFunction.h:144-248);
   angles in degrees */
void orc_incident(const orc_grid* g, double theta, double phi, double psi, double* E);
/* per-face diagonal d_f (vie_fft/fft.h:405-624); eps is per-voxel, V complex;
   adjoint != 0 gives the conjugate (LXX_DO_Conjugate etc.) */
void orc_diag(const orc_grid* g, const double* eps, double* d, int adjoint);

/* ---- operator ----------------------------------------------------------- */
typedef struct orc_op orc_op;
enum { ORC_SCHED_REFERENCE = 0, ORC_SCHED_FUSED = 1 };
/* G: 8V complex (orc_green); eps: V complex per voxel; flags: 3V bytes
   (Flag_X|Flag_Y|Flag_Z).  schedule REFERENCE repeats the reference's 9
   forward + 9 inverse FFTs per product with per-element trigonometry
   (vie_fft/CG_Iteration.h:131-173, vie_fft/fft.h:41-397); FUSED computes the
   same operator with 3 forward + 3 inverse FFTs. */
orc_op* orc_op_create(const orc_grid* g, const double* G, const double* eps,
                      const unsigned char* flags, int schedule);
void orc_op_apply(orc_op* op, const double* x, double* y, int adjoint);
void orc_op_destroy(orc_op* op);
/* wall seconds of one output component f of the product (the work of one
   X_FFT/Y_FFT/Z_FFT call in the reference schedule); y receives that
   component.  Used to time a bounded sample of the CPU path. */
double orc_op_time_component(orc_op* op, const double* x, double* y, int f);
/* matvec callback usable with orc_solve */
void orc_op_matvec(void* op, const double* x, double* y, int adjoint);

/* ---- Krylov solvers (vie_fft/CG_Iteration.h:176-1388) ------------------- */
typedef void (*orc_matvec_fn)(void* ctx, const double* x, double* y, int adjoint);
typedef struct {
    int iterations;  /* loop passes executed (reference prints m = iterations-1) */
    int converged;   /* stop rule met                                          */
    double err;      /* last error value of the method's own stop rule         */
} orc_info;
/* method: 1 CG(NR), 2 BiCG, 3 BiCGSTAB, 4 CORS, 5 BiCORSTAB
   (Iteration_Type, vie_fft/CG_Iteration.h:47).  b is used as given (callers
   mask it like Current_Reset(E_inc) does); x is overwritten, starting from 0.
   hist (may be NULL) receives the error after every iteration (maxit entries). */
int orc_solve(int method, orc_matvec_fn A, void* ctx, long n, const double* b,
              double* x, double tol, int maxit, double* hist, orc_info* info);

/* ---- post-processing (vie_fft/NearFar.h) -------------------------------- */
/* Forward_RCS of FarField1 (vie_fft/NearFar.h:16-109); Q_ext = value / area */
double orc_forward_scatter(const orc_grid* g, const double* J, double theta,
                           double phi, double psi);
/* Ene_Metal-style absorbed power over faces whose two cells both carry
   material id `mat` (vie_fft/NearFar.h:180-250); times Z0 / area = Q_abs */
double orc_absorbed(const orc_grid* g, const double* J, const double* eps,
                    const unsigned char* material, int mat);
/* FarField2 cuts (vie_fft/NearFar.h:302-456): 181 values each, dB */
void orc_rcs(const orc_grid* g, const double* J, double theta, double phi,
             double* rcs_theta, double* rcs_phi);

/* ---- absorbed power over all active faces and its eps gradient ----------
   P = sum over faces (f, i) with flags_f[i] = 1 and i_f < N_f - 1 of
   w_f[i] |J_f[i]|^2, w_f[i] = (dV Z0 / (2 k0)) Im(1/(eps[i]-1) + 1/(eps[i+e_f]-1)):
   the per-face term of NearField's Ene_Metal (vie_fft/NearFar.h:186-250)
   without its material and i,j,k < N-1 selection. */
double orc_absorption(const orc_grid* g, const double* J, const double* eps,
                      const unsigned char* flags);
/* dP/dRe(eps_v) and dP/dIm(eps_v) for every voxel, J the solution of A J = b
   for this operator: one more solve, of A nu = delta_f w conj(J) with the same
   operator (diag(1/delta_f) A is complex symmetric), by `method` to `tol`.
   Returns P; info (may be NULL) describes the extra solve. */
double orc_absorption_grad(const orc_grid* g, orc_op* op, const double* J, const double* eps,
                           const unsigned char* flags, int method, double tol, int maxit,
                           double* grad_re, double* grad_im, orc_info* info);

#ifdef __cplusplus
}
#endif
#endif
