// vie_scalar.h -- device-resident scalar state of the Krylov loops of
// vie_fft/CG_Iteration.h (paths relative to the reference root) and the
// scalar programs that update it from reduced dot products, run by the
// one-workgroup scalar kernel.  (Running the step in the tail of the
// dot-producing kernel instead needs a device-scope release per workgroup,
// which on gfx950 writes back the L2: it made the z-inverse pass 9x slower.)
#pragma once

#include "vie_passes.h"

namespace vie {

constexpr int MAX_HIST = 1 << 16;

// Solver scalars.  cf[] holds the coefficients that Lin prologues (3 entries)
// and LinComb kernels (8 entries: 4 for out, 4 for out2) read at fixed
// per-solver offsets.
struct KState {
    double2 rho, rho1, alpha, omega, beta;
    double2 cf[48];
    double2 sums[4];
    double bnrm2, err, tol;
    double pause;         // recursive residual below which the iterations stop for a residual replacement (0: never)
    int iter, maxit, done, converged, status;
    int fail_op;          // scalar program that stopped the solve, if any
    // mixed-precision batches: the instance is finished, its residual
    // replacements and its last true FP64 residual
    int fin, nrep;
    double terr;
    // divergence stop (0: off): the solve ends once its stop quantity exceeds
    // dvg times the smallest value it had (emin)
    double dvg, emin;
};

enum { KS_OK = 0, KS_BREAKDOWN = 5, KS_NAN = 6, KS_DIVERGED = 8 };

// Coefficient offsets per solver (see the host drivers).
enum {
    CF_BS_P = 0, CF_BS_Q = 3, CF_BS_U = 8,                                        // BiCGSTAB
    CF_CG_U = 8, CF_CG_P = 16,                                                    // CG(NR)
    CF_BI_S = 0, CF_BI_U = 8, CF_BI_P = 16,                                       // BiCG
    CF_CO_L1 = 0, CF_CO_Q = 8, CF_CO_L3 = 12, CF_CO_X = 20, CF_CO_R = 28,         // CORS
    CF_BC_P = 0, CF_BC_Q = 8, CF_BC_L2 = 12, CF_BC_X = 20, CF_BC_R = 28           // BiCORSTAB
};

// ------------------------------------------------------------ scalar helpers
__device__ __forceinline__ double2 zmul(double2 a, double2 b)
{
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ double2 zdiv(double2 a, double2 b)
{
    // Smith's algorithm for a / b
    if (fabs(b.x) >= fabs(b.y)) {
        double r = b.y / b.x, den = b.x + b.y * r;
        return make_double2((a.x + a.y * r) / den, (a.y - a.x * r) / den);
    }
    double r = b.x / b.y, den = b.x * r + b.y;
    return make_double2((a.x * r + a.y) / den, (a.y * r - a.x) / den);
}
__device__ __forceinline__ double2 zneg(double2 a) { return make_double2(-a.x, -a.y); }
__device__ __forceinline__ double2 zsc(double2 a, double s) { return make_double2(a.x * s, a.y * s); }
__device__ __forceinline__ bool zbad(double2 a)
{
    return !(isfinite(a.x) && isfinite(a.y)) || (a.x == 0.0 && a.y == 0.0);
}
__device__ __forceinline__ double2 zc(double re, double im) { return make_double2(re, im); }

// Sum nd partial arrays (each nblk long) in a fixed order into out[0..nd).
// The scalar step runs as one workgroup of NTS threads.
constexpr int NTS = 256;

// out[d] = sum over blocks of part[d * nblk + block], d < ND (<= 4), in a fixed
// order (strided per thread, xor-butterfly per wave, waves in index order).
// Each thread issues the loads of RU strided entries of every dot before it
// adds any of them, so up to RU NTS partials per dot cost one memory latency
// (the partial sums come from a kernel that has just ended, so they are read
// from HBM or a remote cache, not from this XCD's L2); the additions keep the
// strided order, so the sums do not depend on RU.  Offsets are 32-bit from a
// per-dot uniform base (nblk < 2^27 in every launch).
template <int ND>
__device__ __forceinline__ void reduce_nd(const double2* part, int nblk, double2* out, double2 (*red)[NTS / 64])
{
    constexpr int RU = 8;
    double2 a[ND];
#pragma unroll
    for (int d = 0; d < ND; ++d) a[d] = make_double2(0, 0);
    for (int i0 = threadIdx.x; i0 < nblk; i0 += RU * NTS) {
        double2 p[ND][RU];
#pragma unroll
        for (int d = 0; d < ND; ++d)
#pragma unroll
            for (int u = 0; u < RU; ++u) {
                const int i = i0 + u * NTS;
                p[d][u] = i < nblk ? part[(long)d * nblk + i] : make_double2(0, 0);
            }
#pragma unroll
        for (int d = 0; d < ND; ++d)
#pragma unroll
            for (int u = 0; u < RU; ++u)
                if (i0 + u * NTS < nblk) {
                    a[d].x += p[d][u].x;
                    a[d].y += p[d][u].y;
                }
    }
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
    for (int d = 0; d < ND; ++d) {
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) {
            a[d].x += __shfl_xor(a[d].x, off);
            a[d].y += __shfl_xor(a[d].y, off);
        }
        if (lane == 0) red[d][wid] = a[d];
    }
    __syncthreads();
    if (threadIdx.x < ND) {
        double2 t = make_double2(0, 0);
        for (int w = 0; w < NTS / 64; ++w) {
            t.x += red[threadIdx.x][w].x;
            t.y += red[threadIdx.x][w].y;
        }
        out[threadIdx.x] = t;
    }
    __syncthreads();
}
__device__ __forceinline__ void reduce_partials(const double2* part, long nblk, int nd, double2* out,
                                                double2 (*red)[NTS / 64])
{
    switch (nd) {
    case 1: reduce_nd<1>(part, (int)nblk, out, red); break;
    case 2: reduce_nd<2>(part, (int)nblk, out, red); break;
    case 3: reduce_nd<3>(part, (int)nblk, out, red); break;
    default: reduce_nd<4>(part, (int)nblk, out, red); break;
    }
}

// Scalar programs, one per solver step; v[] are the reduced dot products.
enum ScalarOp {
    SC_BNRM = 1,          // v0 = b^H b -> bnrm2
    SC_BICGSTAB_INIT,     // v0 = b^H b -> bnrm2, rho, p = r
    SC_BICGSTAB_ALPHA,    // v0 = b^H Ap
    SC_BICGSTAB_OMEGA,    // v0 = Aq^H q, v1 = Aq^H Aq
    SC_BICGSTAB_END,      // v0 = r^H r, v1 = b^H r
    SC_RELRES,            // v0 = z^H z -> err = sqrt(re / bnrm2), no iteration count
    SC_CG_INIT,           // v0 = p^H p (p = A^H b), v1 = b^H b
    SC_CG_ALPHA,          // v0 = Ap^H Ap
    SC_CG_V4,             // v0 = r^H r
    SC_CG_BETA,           // v0 = q^H q
    SC_BICG_INIT,         // v0 = r^T r
    SC_BICG_ALPHA,        // v0 = Ap^T p
    SC_BICG_END,          // v0 = r^T r
    SC_CORS_RHO,          // v0 = b^H Ap
    SC_CORS_ALPHA,        // v0 = b^H Aq
    SC_RES_END,           // v0 = r^H r: err, iteration end, rho1 = rho
    SC_BICOR_RHO,         // v0 = b^H Ap
    SC_BICOR_ALPHA,       // v0 = b^H Aq
    SC_BICOR_OMEGA,       // v0 = qc^H pc, v1 = qc^H qc
    SC_BICGSTAB_REPLACE   // v0 = b^H r, v1 = r^H r after a residual replacement
};

__device__ __forceinline__ void ks_fail(KState* s, int status)
{
    s->status = status;
    s->done = 1;
}
__device__ __forceinline__ void ks_fail_at(KState* s, int status, int op)
{
    s->fail_op = op;
    ks_fail(s, status);
}

// End-of-iteration bookkeeping common to all methods.
__device__ __forceinline__ void ks_finish(KState* s, double err, double* hist)
{
    s->err = err;
    if (hist && s->iter < MAX_HIST) hist[s->iter] = err;
    s->iter += 1;
    if (!isfinite(err)) { ks_fail(s, KS_NAN); return; }
    if (err < s->tol) { s->converged = 1; s->done = 1; return; }
    if (s->dvg > 0.0) {
        if (s->iter == 1 || err < s->emin) {
            s->emin = err;
        } else if (err > s->dvg * s->emin) {
            ks_fail(s, KS_DIVERGED);
            return;
        }
    }
    if (err < s->pause || s->iter >= s->maxit) s->done = 1;
}

__device__ __forceinline__ void set3(double2* c, double2 a, double2 b, double2 d)
{
    c[0] = a; c[1] = b; c[2] = d;
}
__device__ __forceinline__ void set8(double2* c, double2 a0, double2 a1, double2 a2, double2 a3,
                                     double2 b0, double2 b1, double2 b2, double2 b3)
{
    c[0] = a0; c[1] = a1; c[2] = a2; c[3] = a3;
    c[4] = b0; c[5] = b1; c[6] = b2; c[7] = b3;
}

__device__ inline void scalar_program(int op, KState* s, const double2* v, int nd, double* hist)
{
    for (int d = 0; d < nd; ++d) s->sums[d] = v[d];
    const double2 one = zc(1, 0), O = zc(0, 0);
    double2* cf = s->cf;
    switch (op) {
    case SC_BNRM:
        s->bnrm2 = v[0].x;
        s->iter = 0; s->converged = 0; s->status = 0;
        s->omega = one; s->alpha = O; s->rho1 = O; s->beta = O;
        if (!(v[0].x > 0.0) || !isfinite(v[0].x)) ks_fail(s, v[0].x == 0.0 ? KS_BREAKDOWN : KS_NAN);
        break;
    // ---- BiCGSTAB (vie_fft/CG_Iteration.h:564-748)
    case SC_BICGSTAB_INIT:
        s->bnrm2 = v[0].x;
        s->rho = v[0];  // b^H r with r = b
        s->omega = one; s->alpha = O;
        s->iter = 0; s->converged = 0; s->status = 0;
        set3(cf + CF_BS_P, one, O, O);  // p = r
        if (!(v[0].x > 0.0) || !isfinite(v[0].x)) ks_fail(s, v[0].x == 0.0 ? KS_BREAKDOWN : KS_NAN);
        break;
    case SC_BICGSTAB_ALPHA:
        if (zbad(v[0])) { ks_fail_at(s, KS_BREAKDOWN, op); break; }
        s->alpha = zdiv(s->rho, v[0]);
        set3(cf + CF_BS_Q, one, zneg(s->alpha), O);  // q = r - alpha Ap
        break;
    case SC_BICGSTAB_OMEGA:
        if (zbad(v[1])) { ks_fail_at(s, KS_BREAKDOWN, op); break; }
        s->omega = zdiv(v[0], v[1]);
        // over {x, p, q, Aq}: x = x + alpha p + omega q ; r = q - omega Aq
        set8(cf + CF_BS_U, one, s->alpha, s->omega, O, O, O, one, zneg(s->omega));
        break;
    case SC_BICGSTAB_END: {
        // rho_1 = rho and the next p-update coefficients are formed even on
        // the last iteration, so a residual replacement can resume from here
        s->rho1 = s->rho;
        s->rho = v[1];
        const bool bad = zbad(s->rho1) || zbad(s->omega);
        if (!bad) {
            s->beta = zmul(zdiv(s->rho, s->rho1), zdiv(s->alpha, s->omega));
            // over {r, p, Ap}: p = r + beta p - beta omega Ap
            set3(cf + CF_BS_P, one, s->beta, zneg(zmul(s->beta, s->omega)));
        }
        ks_finish(s, sqrt(v[0].x / s->bnrm2), hist);
        if (!s->done && bad) ks_fail_at(s, KS_BREAKDOWN, op);
        break;
    }
    case SC_BICGSTAB_REPLACE: {
        // the recursive residual was replaced by the true one: v0 = b^H r,
        // v1 = r^H r (reliable update, mixed precision)
        s->rho = v[0];
        if (zbad(s->rho1) || zbad(s->omega)) { ks_fail_at(s, KS_BREAKDOWN, op); break; }
        s->beta = zmul(zdiv(s->rho, s->rho1), zdiv(s->alpha, s->omega));
        set3(cf + CF_BS_P, one, s->beta, zneg(zmul(s->beta, s->omega)));
        s->err = sqrt(v[1].x / s->bnrm2);
        if (s->err < s->tol) { s->converged = 1; s->done = 1; }
        break;
    }
    case SC_RELRES:
        s->err = sqrt(v[0].x / s->bnrm2);
        break;
    // ---- CG on the normal equations (vie_fft/CG_Iteration.h:204-332)
    case SC_CG_INIT:
        s->sums[2] = v[0];     // v3 = ||A^H b||^2
        s->bnrm2 = v[1].x;     // vx = ||b||^2
        s->iter = 0; s->converged = 0; s->status = 0;
        if (!(v[1].x > 0.0) || !isfinite(v[1].x)) ks_fail(s, v[1].x == 0.0 ? KS_BREAKDOWN : KS_NAN);
        break;
    case SC_CG_ALPHA: {
        if (zbad(v[0])) { ks_fail_at(s, KS_BREAKDOWN, op); break; }
        const double a = s->sums[2].x / v[0].x;  // alpha = v1 / v2, v1 = v3
        s->alpha = zc(a, 0);
        s->sums[3] = s->sums[2];
        // over {x, p, r, Ap}: x = x + a p ; r = r + a Ap
        set8(cf + CF_CG_U, one, zc(a, 0), O, O, O, O, one, zc(a, 0));
        break;
    }
    case SC_CG_V4:
        s->sums[1] = v[0];     // v4 = ||r||^2
        break;
    case SC_CG_BETA: {
        const double v4 = s->sums[1].x;
        const double beta = v[0].x / s->sums[3].x;  // v3 / v1
        s->sums[2] = v[0];
        set8(cf + CF_CG_P, zc(beta, 0), zc(-1, 0), O, O, O, O, O, O);  // p = beta p - q
        ks_finish(s, v4 / s->bnrm2, hist);  // squared ratio (vie_fft/CG_Iteration.h:323)
        break;
    }
    // ---- complex-symmetric BiCG (vie_fft/CG_Iteration.h:397-506)
    case SC_BICG_INIT:
        s->sums[2] = v[0];  // vx = r^T r
        s->iter = 0; s->converged = 0; s->status = 0;
        if (zbad(v[0])) { ks_fail_at(s, KS_BREAKDOWN, op); break; }
        set8(cf + CF_BI_S, zdiv(one, v[0]), O, O, O, O, O, O, O);  // p = r / v1
        break;
    case SC_BICG_ALPHA:
        if (zbad(v[0])) { ks_fail_at(s, KS_BREAKDOWN, op); break; }
        s->alpha = zdiv(one, v[0]);
        // over {x, p, r, Ap}: x = x + a p ; r = r - a Ap
        set8(cf + CF_BI_U, one, s->alpha, O, O, O, O, one, zneg(s->alpha));
        break;
    case SC_BICG_END: {
        const double2 vx = s->sums[2];
        const double err = sqrt((v[0].x * v[0].x + v[0].y * v[0].y) / (vx.x * vx.x + vx.y * vx.y));
        ks_finish(s, err, hist);
        if (s->done) break;
        if (zbad(v[0])) { ks_fail_at(s, KS_BREAKDOWN, op); break; }
        s->beta = zdiv(one, v[0]);
        set8(cf + CF_BI_P, one, s->beta, O, O, O, O, O, O);  // p = p + beta r
        break;
    }
    // ---- CORS (vie_fft/CG_Iteration.h:848-1036)
    case SC_CORS_RHO: {
        s->rho = v[0];
        if (s->iter > 0) {
            if (zbad(s->rho1)) { ks_fail_at(s, KS_BREAKDOWN, op); break; }
            s->beta = zdiv(s->rho, s->rho1);
        } else {
            s->beta = O;
        }
        const double2 b = s->beta;
        // over {r, pc, Ap, rc}: p = r + b pc ; qc = Ap + b rc
        set8(cf + CF_CO_L1, one, b, O, O, O, O, one, b);
        // over {qc, rc, q}: q = qc + b rc + b^2 q
        set3(cf + CF_CO_Q, one, b, zmul(b, b));
        break;
    }
    case SC_CORS_ALPHA: {
        if (zbad(v[0])) { ks_fail_at(s, KS_BREAKDOWN, op); break; }
        s->alpha = zdiv(s->rho, v[0]);
        const double2 a = s->alpha, a2 = zmul(a, a);
        // over {p, q, qc, Aq}: pc = p - a q ; rc = qc - a Aq
        set8(cf + CF_CO_L3, one, zneg(a), O, O, O, O, one, zneg(a));
        // over {x, p, q}: x = x + 2a p - a^2 q
        set8(cf + CF_CO_X, one, zsc(a, 2.0), zneg(a2), O, O, O, O, O);
        // over {r, qc, Aq}: r = r - 2a qc + a^2 Aq
        set8(cf + CF_CO_R, one, zsc(a, -2.0), a2, O, O, O, O, O);
        break;
    }
    case SC_RES_END:
        ks_finish(s, sqrt(v[0].x / s->bnrm2), hist);
        s->rho1 = s->rho;
        break;
    // ---- BiCOR-STAB (vie_fft/CG_Iteration.h:1140-1328)
    case SC_BICOR_RHO: {
        s->rho = v[0];
        if (s->iter > 0) {
            if (zbad(s->rho1) || zbad(s->omega)) { ks_fail_at(s, KS_BREAKDOWN, op); break; }
            s->beta = zmul(zdiv(s->rho, s->rho1), zdiv(s->alpha, s->omega));
        } else {
            s->beta = O;
        }
        const double2 b = s->beta, bw = zneg(zmul(s->beta, s->omega));
        // over {r, p, q}: p = r + b p - b w q
        set8(cf + CF_BC_P, one, b, bw, O, O, O, O, O);
        // over {Ap, q, Aq}: q = Ap + b q - b w Aq
        set3(cf + CF_BC_Q, one, b, bw);
        break;
    }
    case SC_BICOR_ALPHA:
        if (zbad(v[0])) { ks_fail_at(s, KS_BREAKDOWN, op); break; }
        s->alpha = zdiv(s->rho, v[0]);
        // over {r, q, Ap, Aq}: pc = r - a q ; qc = Ap - a Aq
        set8(cf + CF_BC_L2, one, zneg(s->alpha), O, O, O, O, one, zneg(s->alpha));
        break;
    case SC_BICOR_OMEGA:
        if (zbad(v[1])) { ks_fail_at(s, KS_BREAKDOWN, op); break; }
        s->omega = zdiv(v[0], v[1]);
        // over {x, p, pc}: x = x + a p + w pc
        set8(cf + CF_BC_X, one, s->alpha, s->omega, O, O, O, O, O);
        // over {pc, qc}: r = pc - w qc
        set8(cf + CF_BC_R, one, zneg(s->omega), O, O, O, O, O, O);
        break;
    default:
        break;
    }
}

}  // namespace vie
