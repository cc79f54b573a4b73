/*
 * vie_oracle.c -- CPU FP64 restatement of the VIE-FFT hot path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker, never the thing measured or
 * shipped).  Every function names the reference lines it restates; paths are
 * relative to the reference root.  Written from the mathematics of those
 * lines, not copied from them.
 */
#define _POSIX_C_SOURCE 199309L
#include "vie_oracle.h"

#include <complex.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

typedef double complex cplx;

/* vie_fft/common.h:17-18 */
#define ORC_PI 3.1415926535897932384626433832795
#define ORC_Z0 (120.0 * ORC_PI)
/* vie_fft/common.h:16 (EPS used by Incident_Wave's k-component test) */
#define ORC_EPS 1.0e-7

static long vol(const orc_grid* g) { return (long)g->n[0] * g->n[1] * g->n[2]; }

/* ======================================================================== */
/* FFT: recursive mixed-radix decimation in time, any length.  Pinned in the */
/* tests against numpy.fft and a brute-force DFT.                            */
/* ======================================================================== */

typedef struct {
    int n;
    int nf;
    int f[40];
    cplx* tw; /* tw[t] = exp(sign * 2 pi j t / n) */
} fft_plan;

static void plan_init(fft_plan* p, int n, int sign)
{
    p->n = n;
    p->nf = 0;
    int m = n;
    while (m % 4 == 0) { p->f[p->nf++] = 4; m /= 4; }
    while (m % 2 == 0) { p->f[p->nf++] = 2; m /= 2; }
    for (int q = 3; m > 1; q += 2)
        while (m % q == 0) { p->f[p->nf++] = q; m /= q; }
    p->tw = (cplx*)malloc(sizeof(cplx) * (size_t)(n > 0 ? n : 1));
    for (int t = 0; t < n; ++t) {
        double a = 2.0 * ORC_PI * (double)t / (double)n;
        p->tw[t] = cos(a) + sign * sin(a) * I;
    }
}

static void plan_free(fft_plan* p) { free(p->tw); }

/* out[0..n) = DFT of in[0], in[is], ..., level indexes the factor list */
static void fft_rec(const cplx* in, long is, cplx* out, int n, int level,
                    const fft_plan* P)
{
    if (n == 1) { out[0] = in[0]; return; }
    const int p = P->f[level];
    const int m = n / p;
    for (int q = 0; q < p; ++q)
        fft_rec(in + q * is, is * p, out + (long)q * m, m, level + 1, P);
    const int N = P->n;
    const int tws = N / n;        /* w_n^e = tw[e * N/n]   */
    const int wps = N / p;        /* w_p^e = tw[e * N/p]   */
    cplx tloc[16];
    cplx* t = p <= 16 ? tloc : (cplx*)malloc(sizeof(cplx) * (size_t)p);
    for (int k = 0; k < m; ++k) {
        for (int q = 0; q < p; ++q) {
            cplx y = out[k + (long)q * m];
            t[q] = (q == 0) ? y : y * P->tw[((long)q * k * tws) % N];
        }
        if (p == 2) {
            out[k] = t[0] + t[1];
            out[k + m] = t[0] - t[1];
        } else if (p == 4) {
            /* w_4 = tw[N/4] = +-j depending on the sign */
            cplx j4 = P->tw[N / 4];
            cplx a0 = t[0] + t[2], a1 = t[0] - t[2];
            cplx b0 = t[1] + t[3], b1 = (t[1] - t[3]) * j4;
            out[k] = a0 + b0;
            out[k + m] = a1 + b1;
            out[k + 2 * (long)m] = a0 - b0;
            out[k + 3 * (long)m] = a1 - b1;
        } else {
            for (int s = 0; s < p; ++s) {
                cplx acc = 0;
                for (int q = 0; q < p; ++q)
                    acc += t[q] * P->tw[((long)q * s % p) * wps];
                out[k + (long)s * m] = acc;
            }
        }
    }
    if (t != tloc) free(t);
}

static void fft_run(const fft_plan* P, cplx* a, cplx* tmp)
{
    memcpy(tmp, a, sizeof(cplx) * (size_t)P->n);
    fft_rec(tmp, 1, a, P->n, 0, P);
}

void orc_fft1d(double* a, int n, int sign)
{
    if (n <= 1) return;
    fft_plan P;
    plan_init(&P, n, sign);
    cplx* tmp = (cplx*)malloc(sizeof(cplx) * (size_t)n);
    fft_run(&P, (cplx*)a, tmp);
    free(tmp);
    plan_free(&P);
}

/* 3-D transform: one pass per axis; strided axes gather blocks of BLK
   neighbouring pencils so rows are read contiguously. */
#define BLK 8
void orc_fft3d(double* ad, int p0, int p1, int p2, int sign)
{
    cplx* a = (cplx*)ad;
    int pmax = p0 > p1 ? p0 : p1;
    if (p2 > pmax) pmax = p2;
    cplx* line = (cplx*)malloc(sizeof(cplx) * (size_t)pmax * BLK);
    cplx* tmp = (cplx*)malloc(sizeof(cplx) * (size_t)pmax);
    /* axis 2: contiguous */
    if (p2 > 1) {
        fft_plan P;
        plan_init(&P, p2, sign);
        for (long r = 0; r < (long)p0 * p1; ++r) fft_run(&P, a + r * p2, tmp);
        plan_free(&P);
    }
    /* axis 1: stride p2, blocks over k */
    if (p1 > 1) {
        fft_plan P;
        plan_init(&P, p1, sign);
        for (int i = 0; i < p0; ++i)
            for (int k0 = 0; k0 < p2; k0 += BLK) {
                int nb = p2 - k0 < BLK ? p2 - k0 : BLK;
                cplx* base = a + (long)i * p1 * p2 + k0;
                for (int j = 0; j < p1; ++j)
                    for (int b = 0; b < nb; ++b) line[(long)b * p1 + j] = base[(long)j * p2 + b];
                for (int b = 0; b < nb; ++b) fft_run(&P, line + (long)b * p1, tmp);
                for (int j = 0; j < p1; ++j)
                    for (int b = 0; b < nb; ++b) base[(long)j * p2 + b] = line[(long)b * p1 + j];
            }
        plan_free(&P);
    }
    /* axis 0: stride p1*p2, blocks over the (j,k) plane */
    if (p0 > 1) {
        fft_plan P;
        plan_init(&P, p0, sign);
        long plane = (long)p1 * p2;
        for (long c0 = 0; c0 < plane; c0 += BLK) {
            int nb = plane - c0 < BLK ? (int)(plane - c0) : BLK;
            for (int i = 0; i < p0; ++i)
                for (int b = 0; b < nb; ++b) line[(long)b * p0 + i] = a[(long)i * plane + c0 + b];
            for (int b = 0; b < nb; ++b) fft_run(&P, line + (long)b * p0, tmp);
            for (int i = 0; i < p0; ++i)
                for (int b = 0; b < nb; ++b) a[(long)i * plane + c0 + b] = line[(long)b * p0 + i];
        }
        plan_free(&P);
    }
    free(tmp);
    free(line);
}

/* ======================================================================== */
/* Set-up                                                                    */
/* ======================================================================== */

/* Equivalent-sphere integrated scalar Green's function
   (vie_fft/GreenFunction.h:23-25 parameters, :59-85 closed forms).
   With D = 2(3 dV / 4 pi)^(1/3) and p = k0 D / 2:
     R > 0 : g = exp(-j k0 R) (sin p / p - cos p) / ((pi/3) k0^2 D^2 R) * dV
     R = 0 : g = ((1 + j p) exp(-j p) - 1) / ((pi/6) k0^2 D^3) * dV          */
void orc_green_offsets(const orc_grid* g, double* goffd)
{
    cplx* goff = (cplx*)goffd;
    const double dV = g->d[0] * g->d[1] * g->d[2];
    const double D = 2.0 * cbrt(dV / (4.0 / 3.0 * ORC_PI));
    const double p = 0.5 * g->k0 * D;
    const double k0 = g->k0;
    const double shape = sin(p) / p - cos(p);
    const int n1 = g->n[1] + 1, n2 = g->n[2] + 1;
    for (int i = 0; i <= g->n[0]; ++i)
        for (int j = 0; j <= g->n[1]; ++j)
            for (int k = 0; k <= g->n[2]; ++k) {
                cplx v;
                if (i + j + k != 0) {
                    double x = i * g->d[0], y = j * g->d[1], z = k * g->d[2];
                    double R = sqrt(x * x + y * y + z * z);
                    cplx ph = cos(k0 * R) - I * sin(k0 * R);
                    v = ph * shape / (ORC_PI / 3.0 * k0 * k0 * D * D * R) * dV;
                } else {
                    cplx e = cos(p) - I * sin(p);
                    v = ((1.0 + I * p) * e - 1.0) / (ORC_PI / 6.0 * k0 * k0 * D * D * D) * dV;
                }
                goff[((long)i * n1 + j) * n2 + k] = v;
            }
}

/* Even extension g^[n] = g[min(n, P - n)] per axis onto P = 2N, then the
   unnormalised forward DFT (vie_fft/GreenFunction.h:91-139). */
void orc_green(const orc_grid* g, double* Gd)
{
    cplx* G = (cplx*)Gd;
    const int N0 = g->n[0], N1 = g->n[1], N2 = g->n[2];
    const int P0 = 2 * N0, P1 = 2 * N1, P2 = 2 * N2;
    cplx* goff = (cplx*)malloc(sizeof(cplx) * (size_t)(N0 + 1) * (N1 + 1) * (N2 + 1));
    orc_green_offsets(g, (double*)goff);
    for (int i = 0; i < P0; ++i) {
        int ii = i >= N0 ? P0 - i : i;
        for (int j = 0; j < P1; ++j) {
            int jj = j >= N1 ? P1 - j : j;
            for (int k = 0; k < P2; ++k) {
                int kk = k >= N2 ? P2 - k : k;
                G[((long)i * P1 + j) * P2 + k] = goff[((long)ii * (N1 + 1) + jj) * (N2 + 1) + kk];
            }
        }
    }
    free(goff);
    orc_fft3d(Gd, P0, P1, P2, -1);
}

/* Pulse-integrated plane wave on each component's staggered points
   (vie_fft/GreenFunction.h:144-248).  The integral of exp(-j k_a s) over a
   cell of length delta_a centred on the face is
   (exp(-j k_a d/2) - exp(+j k_a d/2)) / (-j k_a), or delta_a when |k_a| < EPS. */
static cplx pulse(double ka, double da)
{
    if (fabs(ka) < ORC_EPS) return da;
    cplx e1 = cos(ka * da / 2.0) - I * sin(ka * da / 2.0);
    cplx e2 = cos(ka * da / 2.0) + I * sin(ka * da / 2.0);
    return (e1 - e2) / (-I * ka);
}

void orc_incident(const orc_grid* g, double theta, double phi, double psi, double* Ed)
{
    cplx* E = (cplx*)Ed;
    const long V = vol(g);
    const double th = ORC_PI * theta / 180.0, ph = ORC_PI * phi / 180.0, ps = ORC_PI * psi / 180.0;
    const double kv[3] = {g->k0 * sin(th) * cos(ph), g->k0 * sin(th) * sin(ph), g->k0 * cos(th)};
    const double pol[3] = {cos(th) * cos(ph) * cos(ps) - sin(ph) * sin(ps),
                           cos(th) * sin(ph) * cos(ps) + cos(ph) * sin(ps),
                           -sin(th) * cos(ps)};
    for (int c = 0; c < 3; ++c) {
        cplx pw = pulse(kv[c], g->d[c]) * pol[c];
        for (int i = 0; i < g->n[0]; ++i)
            for (int j = 0; j < g->n[1]; ++j)
                for (int k = 0; k < g->n[2]; ++k) {
                    double pos[3] = {(i + 0.5) * g->d[0], (j + 0.5) * g->d[1], (k + 0.5) * g->d[2]};
                    pos[c] += 0.5 * g->d[c]; /* the face of component c */
                    double a = kv[0] * pos[0] + kv[1] * pos[1] + kv[2] * pos[2];
                    cplx e = cos(a) - I * sin(a);
                    E[c * V + ((long)i * g->n[1] + j) * g->n[2] + k] = e * pw;
                }
    }
}

/* d_f[i] = -j (delta_f/2)(Z0/k0)(1/(eps[i+e_f]-1) + 1/(eps[i]-1)) for
   i_f < N_f - 1, else 0 (vie_fft/fft.h:405-438, 480-513, 555-588); the
   adjoint is the complex conjugate (vie_fft/fft.h:441-474, 516-549, 591-624). */
void orc_diag(const orc_grid* g, const double* epsd, double* dd, int adjoint)
{
    const cplx* eps = (const cplx*)epsd;
    cplx* d = (cplx*)dd;
    const long V = vol(g);
    const int N0 = g->n[0], N1 = g->n[1], N2 = g->n[2];
    const long st[3] = {(long)N1 * N2, N2, 1};
    for (int f = 0; f < 3; ++f) {
        cplx c = -I * (g->d[f] / 2.0) * ORC_Z0 / g->k0;
        for (int i = 0; i < N0; ++i)
            for (int j = 0; j < N1; ++j)
                for (int k = 0; k < N2; ++k) {
                    long v = ((long)i * N1 + j) * N2 + k;
                    int idx[3] = {i, j, k};
                    cplx val = 0;
                    if (idx[f] < g->n[f] - 1) {
                        cplx a = eps[v + st[f]] - 1.0, b = eps[v] - 1.0;
                        val = c * (1.0 / a + 1.0 / b);
                    }
                    d[f * V + v] = adjoint ? conj(val) : val;
                }
    }
}

/* ======================================================================== */
/* Operator                                                                  */
/* ======================================================================== */

struct orc_op {
    orc_grid g;
    int P[3];
    long V, NP;
    int sched;
    cplx* G;       /* 8V */
    cplx* d;       /* 3V forward diagonal */
    unsigned char* flags; /* 3V */
    cplx* work;    /* sched REFERENCE: 1 x 8V; FUSED: 3 x 8V */
};

orc_op* orc_op_create(const orc_grid* g, const double* G, const double* eps,
                      const unsigned char* flags, int schedule)
{
    orc_op* op = (orc_op*)calloc(1, sizeof(orc_op));
    op->g = *g;
    for (int a = 0; a < 3; ++a) op->P[a] = 2 * g->n[a];
    op->V = vol(g);
    op->NP = 8 * op->V;
    op->sched = schedule;
    op->G = (cplx*)malloc(sizeof(cplx) * (size_t)op->NP);
    memcpy(op->G, G, sizeof(cplx) * (size_t)op->NP);
    op->d = (cplx*)malloc(sizeof(cplx) * 3 * (size_t)op->V);
    orc_diag(g, eps, (double*)op->d, 0);
    op->flags = (unsigned char*)malloc(3 * (size_t)op->V);
    memcpy(op->flags, flags, 3 * (size_t)op->V);
    size_t nw = (schedule == ORC_SCHED_FUSED ? 3 : 1) * (size_t)op->NP;
    op->work = (cplx*)malloc(sizeof(cplx) * nw);
    return op;
}

void orc_op_destroy(orc_op* op)
{
    if (!op) return;
    free(op->G);
    free(op->d);
    free(op->flags);
    free(op->work);
    free(op);
}

/* Current_Expansion (vie_fft/CG_Iteration.h:50-83): zero the P^3 buffer and
   copy component s, dropping its own-axis last plane. */
static void expand(const orc_op* op, const cplx* xs, int s, cplx* w)
{
    const int* N = op->g.n;
    const int* P = op->P;
    memset(w, 0, sizeof(cplx) * (size_t)op->NP);
    int lim[3] = {N[0], N[1], N[2]};
    lim[s] -= 1;
    for (int i = 0; i < lim[0]; ++i)
        for (int j = 0; j < lim[1]; ++j)
            for (int k = 0; k < lim[2]; ++k)
                w[((long)i * P[1] + j) * P[2] + k] = xs[((long)i * N[1] + j) * N[2] + k];
}

/* truncation + accumulation of the inverse transform onto [0,N)
   (vie_fft/fft.h:113-123 and the three copies) */
static void accumulate(const orc_op* op, const cplx* w, cplx* yf)
{
    const int* N = op->g.n;
    const int* P = op->P;
    for (int i = 0; i < N[0]; ++i)
        for (int j = 0; j < N[1]; ++j)
            for (int k = 0; k < N[2]; ++k)
                yf[((long)i * N[1] + j) * N[2] + k] += w[((long)i * P[1] + j) * P[2] + k];
}

/* Reference schedule: one output component f is the sum of three padded
   convolutions, each with its own forward and inverse FFT and a per-element
   trigonometric symbol (X_FFT/Y_FFT/Z_FFT, vie_fft/CG_Iteration.h:131-173;
   L_DC_Dq vie_fft/fft.h:41-127; L_Dq :224-309; conjugates :131-219, :312-397). */
static void apply_reference_f(orc_op* op, const cplx* x, cplx* y, int adj, int fsel);

static void apply_reference(orc_op* op, const cplx* x, cplx* y, int adj)
{
    for (int f = 0; f < 3; ++f) apply_reference_f(op, x, y, adj, f);
}

static void apply_reference_component(orc_op* op, const cplx* x, cplx* y, int f)
{
    apply_reference_f(op, x, y, 0, f);
}

static void apply_reference_f(orc_op* op, const cplx* x, cplx* y, int adj, int fsel)
{
    const long V = op->V;
    const int* P = op->P;
    const double k0 = op->g.k0;
    const double NP = (double)op->NP;
    const double sg = adj ? -1.0 : 1.0; /* conjugation sign */
    /* source order of X_FFT, Y_FFT, Z_FFT */
    static const int order[3][3] = {{0, 1, 2}, {1, 0, 2}, {2, 0, 1}};
    cplx* w = op->work;
    for (int f = fsel; f == fsel; ++f) {
        cplx* yf = y + f * V;
        memset(yf, 0, sizeof(cplx) * (size_t)V);
        for (int o = 0; o < 3; ++o) {
            int s = order[f][o];
            expand(op, x + s * V, s, w);
            if (s == f) {
                /* LXX_DO / LYY_DO / LZZ_DO: diagonal term on [0, N_f-1) */
                for (long v = 0; v < V; ++v) {
                    cplx dv = op->d[f * V + v];
                    if (adj) dv = conj(dv);
                    int i = (int)(v / ((long)op->g.n[1] * op->g.n[2]));
                    int j = (int)((v / op->g.n[2]) % op->g.n[1]);
                    int k = (int)(v % op->g.n[2]);
                    yf[v] += dv * w[((long)i * P[1] + j) * P[2] + k];
                }
            }
            orc_fft3d((double*)w, P[0], P[1], P[2], -1);
            if (s == f) {
                const double df = op->g.d[f];
                const cplx V5 = sg * I * k0 * ORC_Z0 * df / NP;
                const cplx V6 = sg * I * ORC_Z0 / k0 / df / NP;
                for (int i = 0; i < P[0]; ++i)
                    for (int j = 0; j < P[1]; ++j)
                        for (int k = 0; k < P[2]; ++k) {
                            int idx[3] = {i, j, k};
                            double th = 2.0 * ORC_PI * idx[s] / (double)P[s];
                            cplx V3 = (1.0 - cos(th)) + sg * sin(th) * I;
                            cplx V4 = (cos(th) - 1.0) + sg * sin(th) * I;
                            long e = ((long)i * P[1] + j) * P[2] + k;
                            cplx Gv = adj ? conj(op->G[e]) : op->G[e];
                            w[e] = w[e] * Gv * (V5 + V6 * V3 * V4);
                        }
            } else {
                /* forward: V5 uses the source's delta; adjoint: the field's */
                const double dd = adj ? op->g.d[f] : op->g.d[s];
                const cplx V5 = sg * I * ORC_Z0 / k0 / dd / NP;
                for (int i = 0; i < P[0]; ++i)
                    for (int j = 0; j < P[1]; ++j)
                        for (int k = 0; k < P[2]; ++k) {
                            int idx[3] = {i, j, k};
                            long e = ((long)i * P[1] + j) * P[2] + k;
                            cplx V3, V4;
                            if (!adj) {
                                double ts = 2.0 * ORC_PI * idx[s] / (double)P[s];
                                double tf = 2.0 * ORC_PI * idx[f] / (double)P[f];
                                V3 = (1.0 - cos(ts)) + sin(ts) * I;  /* 1 - e^{-j ts} */
                                V4 = (cos(tf) - 1.0) + sin(tf) * I;  /* e^{j tf} - 1  */
                            } else {
                                double tf = 2.0 * ORC_PI * idx[f] / (double)P[f];
                                double ts = 2.0 * ORC_PI * idx[s] / (double)P[s];
                                V3 = (1.0 - cos(tf)) - sin(tf) * I;  /* 1 - e^{+j tf} */
                                V4 = (cos(ts) - 1.0) - sin(ts) * I;  /* e^{-j ts} - 1 */
                            }
                            cplx Gv = adj ? conj(op->G[e]) : op->G[e];
                            w[e] = V5 * w[e] * Gv * V3 * V4;
                        }
            }
            orc_fft3d((double*)w, P[0], P[1], P[2], +1);
            accumulate(op, w, yf);
        }
        /* Current_Reset (vie_fft/fft.h:13-35) */
        for (long v = 0; v < V; ++v)
            if (!op->flags[f * V + v]) yf[v] = 0;
    }
}

/* Same operator with 3 + 3 transforms: per spectral point
   Y_f = G~ (D_f X_f + u_f sum_s v_s X_s), D_f = j k0 Z0 delta_f / NP,
   u_f = e^{j th_f} - 1, v_s = (j Z0 / (k0 delta_s NP)) (1 - e^{-j th_s});
   the adjoint conjugates every factor and swaps the roles of u and v.  This
   rank-one-plus-diagonal factorisation is the product
   (1 - e^{-j th_s})(e^{j th_f} - 1) of vie_fft/fft.h:95-100 and :276-281. */
static void apply_fused(orc_op* op, const cplx* x, cplx* y, int adj)
{
    const long V = op->V, NPl = op->NP;
    const int* P = op->P;
    const double k0 = op->g.k0;
    const double NP = (double)NPl;
    cplx* X[3] = {op->work, op->work + NPl, op->work + 2 * NPl};
    for (int s = 0; s < 3; ++s) {
        expand(op, x + s * V, s, X[s]);
        orc_fft3d((double*)X[s], P[0], P[1], P[2], -1);
    }
    cplx* u[3];
    cplx* vv[3];
    cplx Dd[3];
    for (int a = 0; a < 3; ++a) {
        u[a] = (cplx*)malloc(sizeof(cplx) * (size_t)P[a]);
        vv[a] = (cplx*)malloc(sizeof(cplx) * (size_t)P[a]);
        for (int t = 0; t < P[a]; ++t) {
            double th = 2.0 * ORC_PI * t / (double)P[a];
            cplx ep = cos(th) + I * sin(th);
            u[a][t] = ep - 1.0;
            vv[a][t] = I * ORC_Z0 / (k0 * op->g.d[a] * NP) * (1.0 - conj(ep));
            if (adj) { /* conj and swap: u' = conj(v), v' = conj(u) */
                cplx tu = conj(vv[a][t]);
                vv[a][t] = conj(u[a][t]);
                u[a][t] = tu;
            }
        }
        Dd[a] = I * k0 * ORC_Z0 * op->g.d[a] / NP;
        if (adj) Dd[a] = conj(Dd[a]);
    }
    for (int i = 0; i < P[0]; ++i)
        for (int j = 0; j < P[1]; ++j)
            for (int k = 0; k < P[2]; ++k) {
                long e = ((long)i * P[1] + j) * P[2] + k;
                int t[3] = {i, j, k};
                cplx Gv = adj ? conj(op->G[e]) : op->G[e];
                cplx sig = vv[0][t[0]] * X[0][e] + vv[1][t[1]] * X[1][e] + vv[2][t[2]] * X[2][e];
                cplx Y[3];
                for (int f = 0; f < 3; ++f) Y[f] = Gv * (Dd[f] * X[f][e] + u[f][t[f]] * sig);
                for (int f = 0; f < 3; ++f) X[f][e] = Y[f];
            }
    for (int a = 0; a < 3; ++a) { free(u[a]); free(vv[a]); }
    for (int f = 0; f < 3; ++f) {
        orc_fft3d((double*)X[f], P[0], P[1], P[2], +1);
        cplx* yf = y + f * V;
        const int* N = op->g.n;
        int lim = N[f] - 1;
        for (int i = 0; i < N[0]; ++i)
            for (int j = 0; j < N[1]; ++j)
                for (int k = 0; k < N[2]; ++k) {
                    long v = ((long)i * N[1] + j) * N[2] + k;
                    int idx[3] = {i, j, k};
                    cplx xv = idx[f] < lim ? x[f * V + v] : 0;
                    cplx dv = adj ? conj(op->d[f * V + v]) : op->d[f * V + v];
                    cplx val = X[f][((long)i * P[1] + j) * P[2] + k] + dv * xv;
                    yf[v] = op->flags[f * V + v] ? val : 0;
                }
    }
}

static void apply_reference_component(orc_op* op, const cplx* x, cplx* y, int f);

double orc_op_time_component(orc_op* op, const double* x, double* y, int f)
{
    struct timespec t0, t1;
    clock_gettime(CLOCK_MONOTONIC, &t0);
    apply_reference_component(op, (const cplx*)x, (cplx*)y, f);
    clock_gettime(CLOCK_MONOTONIC, &t1);
    return (double)(t1.tv_sec - t0.tv_sec) + 1e-9 * (double)(t1.tv_nsec - t0.tv_nsec);
}

void orc_op_apply(orc_op* op, const double* x, double* y, int adjoint)
{
    if (op->sched == ORC_SCHED_REFERENCE)
        apply_reference(op, (const cplx*)x, (cplx*)y, adjoint);
    else
        apply_fused(op, (const cplx*)x, (cplx*)y, adjoint);
}

void orc_op_matvec(void* op, const double* x, double* y, int adjoint)
{
    orc_op_apply((orc_op*)op, x, y, adjoint);
}

/* ======================================================================== */
/* Krylov solvers (vie_fft/CG_Iteration.h)                                   */
/* ======================================================================== */

static cplx dotc(long n, const cplx* a, const cplx* b) /* sum conj(a) b */
{
    cplx s = 0;
    for (long i = 0; i < n; ++i) s += conj(a[i]) * b[i];
    return s;
}
static cplx dotu(long n, const cplx* a, const cplx* b) /* sum a b */
{
    cplx s = 0;
    for (long i = 0; i < n; ++i) s += a[i] * b[i];
    return s;
}
static void axpy(long n, cplx a, const cplx* x, cplx* y)
{
    for (long i = 0; i < n; ++i) y[i] += a * x[i];
}
static void scal(long n, cplx a, cplx* x)
{
    for (long i = 0; i < n; ++i) x[i] *= a;
}
static void copy(long n, const cplx* x, cplx* y) { memcpy(y, x, sizeof(cplx) * (size_t)n); }

static cplx* vec(long n) { return (cplx*)calloc((size_t)n, sizeof(cplx)); }

#define MV(in, out, adj) A(ctx, (const double*)(in), (double*)(out), (adj))

/* CG on the normal equations (vie_fft/CG_Iteration.h:176-357): r = -b,
   p = A^H b, stop when ||r||^2/||b||^2 (squared, :323) < tol. */
static void solve_cg(orc_matvec_fn A, void* ctx, long n, const cplx* b, cplx* x,
                     double tol, int maxit, double* hist, orc_info* info)
{
    cplx *r = vec(n), *q = vec(n), *p = vec(n), *Ap = vec(n);
    copy(n, b, r);
    scal(n, -1.0, r);
    MV(b, p, 1);
    double v3 = creal(dotc(n, p, p));
    double vx = creal(dotc(n, b, b));
    for (int m = 0; m < maxit; ++m) {
        double v1 = v3;
        MV(p, Ap, 0);
        double v2 = creal(dotc(n, Ap, Ap));
        double alpha = v1 / v2;
        axpy(n, alpha, p, x);
        axpy(n, alpha, Ap, r);
        double v4 = creal(dotc(n, r, r));
        MV(r, q, 1);
        v3 = creal(dotc(n, q, q));
        double beta = v3 / v1;
        scal(n, beta, p);
        axpy(n, -1.0, q, p);
        double err = v4 / vx;
        if (hist) hist[m] = err;
        info->iterations = m + 1;
        info->err = err;
        if (err < tol) { info->converged = 1; break; }
    }
    free(r); free(q); free(p); free(Ap);
}

/* Complex-symmetric BiCG with unconjugated products
   (vie_fft/CG_Iteration.h:360-528); stop on |r^T r| / |b^T b| (:495-496). */
static void solve_bicg(orc_matvec_fn A, void* ctx, long n, const cplx* b, cplx* x,
                       double tol, int maxit, double* hist, orc_info* info)
{
    cplx *r = vec(n), *p = vec(n), *Ap = vec(n);
    copy(n, b, r);
    cplx v1 = dotu(n, r, r);
    copy(n, r, p);
    scal(n, 1.0 / v1, p);
    cplx vx = v1;
    for (int m = 0; m < maxit; ++m) {
        MV(p, Ap, 0);
        v1 = dotu(n, Ap, p);
        cplx alpha = 1.0 / v1;
        axpy(n, alpha, p, x);
        axpy(n, -alpha, Ap, r);
        cplx v2 = dotu(n, r, r);
        cplx beta = 1.0 / v2;
        axpy(n, beta, r, p);
        double err = sqrt((creal(v2) * creal(v2) + cimag(v2) * cimag(v2)) /
                          (creal(vx) * creal(vx) + cimag(vx) * cimag(vx)));
        if (hist) hist[m] = err;
        info->iterations = m + 1;
        info->err = err;
        if (err < tol) { info->converged = 1; break; }
    }
    free(r); free(p); free(Ap);
}

/* BiCGSTAB, shadow residual b (vie_fft/CG_Iteration.h:531-776). */
static void solve_bicgstab(orc_matvec_fn A, void* ctx, long n, const cplx* b, cplx* x,
                           double tol, int maxit, double* hist, orc_info* info)
{
    cplx *r = vec(n), *q = vec(n), *p = vec(n), *Ap = vec(n), *Aq = vec(n);
    copy(n, b, r);
    double bnrm2 = creal(dotc(n, b, b));
    cplx omega = 1.0, alpha = 0, rho1 = 0;
    for (int m = 0; m < maxit; ++m) {
        cplx rho = dotc(n, b, r);
        if (m > 0) {
            cplx beta = (rho / rho1) * (alpha / omega);
            axpy(n, -omega, Ap, p);
            scal(n, beta, p);
            axpy(n, 1.0, r, p);
        } else {
            copy(n, r, p);
        }
        MV(p, Ap, 0);
        alpha = rho / dotc(n, b, Ap);
        copy(n, r, q);
        axpy(n, -alpha, Ap, q);
        MV(q, Aq, 0);
        omega = dotc(n, Aq, q) / dotc(n, Aq, Aq);
        axpy(n, alpha, p, x);
        axpy(n, omega, q, x);
        copy(n, q, r);
        axpy(n, -omega, Aq, r);
        double err = sqrt(creal(dotc(n, r, r)) / bnrm2);
        if (hist) hist[m] = err;
        info->iterations = m + 1;
        info->err = err;
        if (err < tol) { info->converged = 1; break; }
        rho1 = rho;
    }
    free(r); free(q); free(p); free(Ap); free(Aq);
}

/* CORS, shadow b (vie_fft/CG_Iteration.h:780-1073). */
static void solve_cors(orc_matvec_fn A, void* ctx, long n, const cplx* b, cplx* x,
                       double tol, int maxit, double* hist, orc_info* info)
{
    cplx *r = vec(n), *q = vec(n), *p = vec(n), *Ap = vec(n), *Aq = vec(n);
    cplx *rc = vec(n), *qc = vec(n), *pc = vec(n);
    copy(n, b, r);
    double bnrm2 = creal(dotc(n, b, b));
    cplx rho1 = 0;
    for (int m = 0; m < maxit; ++m) {
        MV(r, Ap, 0);
        cplx rho = dotc(n, b, Ap);
        if (m > 0) {
            cplx beta = rho / rho1;
            copy(n, r, p);
            axpy(n, beta, pc, p);          /* p  = r + beta pc           */
            copy(n, Ap, qc);
            axpy(n, beta, rc, qc);         /* qc = Ap + beta rc          */
            scal(n, beta, q);
            axpy(n, 1.0, rc, q);
            scal(n, beta, q);
            axpy(n, 1.0, qc, q);           /* q = qc + beta (rc + beta q) */
        } else {
            copy(n, r, p);
            copy(n, Ap, qc);
            copy(n, Ap, q);
        }
        MV(q, Aq, 0);
        cplx alpha = rho / dotc(n, b, Aq);
        copy(n, p, pc);
        axpy(n, -alpha, q, pc);            /* pc = p - alpha q           */
        copy(n, qc, rc);
        axpy(n, -alpha, Aq, rc);           /* rc = qc - alpha Aq         */
        axpy(n, 2.0 * alpha, p, x);
        axpy(n, -alpha * alpha, q, x);     /* x += alpha (2p - alpha q)  */
        axpy(n, -2.0 * alpha, qc, r);
        axpy(n, alpha * alpha, Aq, r);     /* r -= alpha (2qc - alpha Aq) */
        double err = sqrt(creal(dotc(n, r, r)) / bnrm2);
        if (hist) hist[m] = err;
        info->iterations = m + 1;
        info->err = err;
        if (err < tol) { info->converged = 1; break; }
        rho1 = rho;
    }
    free(r); free(q); free(p); free(Ap); free(Aq); free(rc); free(qc); free(pc);
}

/* BiCOR-STAB, shadow b (vie_fft/CG_Iteration.h:1076-1362). */
static void solve_bicorstab(orc_matvec_fn A, void* ctx, long n, const cplx* b, cplx* x,
                            double tol, int maxit, double* hist, orc_info* info)
{
    cplx *r = vec(n), *q = vec(n), *p = vec(n), *Ap = vec(n), *Aq = vec(n);
    cplx *qc = vec(n), *pc = vec(n);
    copy(n, b, r);
    double bnrm2 = creal(dotc(n, b, b));
    cplx rho1 = 0, alpha = 0, omega = 1.0;
    for (int m = 0; m < maxit; ++m) {
        MV(r, Ap, 0);
        cplx rho = dotc(n, b, Ap);
        if (m > 0) {
            cplx beta = (rho / rho1) * (alpha / omega);
            axpy(n, -omega, q, p);
            scal(n, beta, p);
            axpy(n, 1.0, r, p);            /* p = r + beta (p - omega q)  */
            axpy(n, -omega, Aq, q);
            scal(n, beta, q);
            axpy(n, 1.0, Ap, q);           /* q = Ap + beta (q - omega Aq) */
        } else {
            copy(n, r, p);
            copy(n, Ap, q);
        }
        MV(q, Aq, 0);
        alpha = rho / dotc(n, b, Aq);
        copy(n, r, pc);
        axpy(n, -alpha, q, pc);            /* pc = r - alpha q            */
        copy(n, Ap, qc);
        axpy(n, -alpha, Aq, qc);           /* qc = Ap - alpha Aq          */
        omega = dotc(n, qc, pc) / dotc(n, qc, qc);
        axpy(n, alpha, p, x);
        axpy(n, omega, pc, x);
        copy(n, pc, r);
        axpy(n, -omega, qc, r);
        double err = sqrt(creal(dotc(n, r, r)) / bnrm2);
        if (hist) hist[m] = err;
        info->iterations = m + 1;
        info->err = err;
        if (err < tol) { info->converged = 1; break; }
        rho1 = rho;
    }
    free(r); free(q); free(p); free(Ap); free(Aq); free(qc); free(pc);
}

int orc_solve(int method, orc_matvec_fn A, void* ctx, long n, const double* bd,
              double* xd, double tol, int maxit, double* hist, orc_info* info)
{
    const cplx* b = (const cplx*)bd;
    cplx* x = (cplx*)xd;
    orc_info tmp;
    if (!info) info = &tmp;
    memset(info, 0, sizeof(*info));
    memset(x, 0, sizeof(cplx) * (size_t)n);
    switch (method) {
    case 1: solve_cg(A, ctx, n, b, x, tol, maxit, hist, info); break;
    case 2: solve_bicg(A, ctx, n, b, x, tol, maxit, hist, info); break;
    case 3: solve_bicgstab(A, ctx, n, b, x, tol, maxit, hist, info); break;
    case 4: solve_cors(A, ctx, n, b, x, tol, maxit, hist, info); break;
    case 5: solve_bicorstab(A, ctx, n, b, x, tol, maxit, hist, info); break;
    default: return -1;
    }
    return info->converged ? 0 : 1;
}

/* ======================================================================== */
/* Post-processing (vie_fft/NearFar.h)                                       */
/* ======================================================================== */

/* Far-field sums S_c = dV e^{j k0 delta_c v_c / 2} sum J_c e^{j k0 r.v} over
   cell centres (vie_fft/NearFar.h:59-96). */
static void far_sums(const orc_grid* g, const cplx* J, double th_deg, double ph_deg, cplx S[3])
{
    const long V = vol(g);
    const double th = ORC_PI * th_deg / 180.0, ph = ORC_PI * ph_deg / 180.0;
    const double vd[3] = {sin(th) * cos(ph), sin(th) * sin(ph), cos(th)};
    const double dV = g->d[0] * g->d[1] * g->d[2];
    S[0] = S[1] = S[2] = 0;
    for (int i = 0; i < g->n[0]; ++i)
        for (int j = 0; j < g->n[1]; ++j)
            for (int k = 0; k < g->n[2]; ++k) {
                double a = g->k0 * ((i + 0.5) * g->d[0] * vd[0] + (j + 0.5) * g->d[1] * vd[1] +
                                    (k + 0.5) * g->d[2] * vd[2]);
                cplx e = cos(a) + I * sin(a);
                long v = ((long)i * g->n[1] + j) * g->n[2] + k;
                for (int c = 0; c < 3; ++c) S[c] += e * J[c * V + v];
            }
    for (int c = 0; c < 3; ++c) {
        double a = g->k0 * 0.5 * g->d[c] * vd[c];
        S[c] = (cos(a) + I * sin(a)) * S[c] * dV;
    }
}

double orc_forward_scatter(const orc_grid* g, const double* Jd, double theta, double phi, double psi)
{
    cplx S[3];
    far_sums(g, (const cplx*)Jd, theta, phi, S);
    const double th = ORC_PI * theta / 180.0, ph = ORC_PI * phi / 180.0, ps = ORC_PI * psi / 180.0;
    double tm = ORC_Z0 * (creal(S[0]) * cos(th) * cos(ph) + creal(S[1]) * cos(th) * sin(ph) -
                          creal(S[2]) * sin(th));
    double te = ORC_Z0 * (creal(S[0]) * sin(ph) - creal(S[1]) * cos(ph));
    return cos(ps) * tm - sin(ps) * te;
}

/* vie_fft/NearFar.h:180-250: faces with i,j,k < N-1 on every axis whose two
   cells both carry material `mat`; |E_f| = |J_f| |(-j Z0/(2k0)) S|,
   S = 1/(eps_a-1) + 1/(eps_b-1), eps_eff = 2/S + 1, power
   sum (k0/Z0) Im(-eps_eff) |E_f|^2 dV. */
double orc_absorbed(const orc_grid* g, const double* Jd, const double* epsd,
                    const unsigned char* mat, int mid)
{
    const cplx* J = (const cplx*)Jd;
    const cplx* eps = (const cplx*)epsd;
    const long V = vol(g);
    const int N0 = g->n[0], N1 = g->n[1], N2 = g->n[2];
    const long st[3] = {(long)N1 * N2, N2, 1};
    const double dV = g->d[0] * g->d[1] * g->d[2];
    const cplx V5 = -0.5 * I * ORC_Z0 / g->k0;
    double P = 0;
    for (int i = 0; i < N0 - 1; ++i)
        for (int j = 0; j < N1 - 1; ++j)
            for (int k = 0; k < N2 - 1; ++k) {
                long v = ((long)i * N1 + j) * N2 + k;
                for (int f = 0; f < 3; ++f) {
                    long w = v + st[f];
                    if (!(mat[w] == mid && mat[v] == mid)) continue;
                    cplx S = 1.0 / (eps[w] - 1.0) + 1.0 / (eps[v] - 1.0);
                    double aE = cabs(J[f * V + v]) * cabs(S * V5);
                    cplx ee = 2.0 / S + 1.0;
                    P += g->k0 / ORC_Z0 * cimag(-ee) * aE * aE * dV;
                }
            }
    return P;
}

void orc_rcs(const orc_grid* g, const double* Jd, double theta, double phi,
             double* rcs_theta, double* rcs_phi)
{
    (void)theta;
    const cplx* J = (const cplx*)Jd;
    const double lam = 2.0 * ORC_PI / g->k0;
    for (int a = 0; a <= 180; ++a) {
        /* E plane, phi_s = Phi (vie_fft/NearFar.h:321-390) */
        cplx S[3];
        far_sums(g, J, (double)a, phi, S);
        double t = ORC_PI * a / 180.0, p = ORC_PI * phi / 180.0;
        double m1 = creal(S[0]) * cos(t) * cos(p) + creal(S[1]) * cos(t) * sin(p) - creal(S[2]) * sin(t);
        double m2 = cimag(S[0]) * cos(t) * cos(p) + cimag(S[1]) * cos(t) * sin(p) - cimag(S[2]) * sin(t);
        double r = g->k0 * g->k0 * ORC_Z0 * ORC_Z0 * (m1 * m1 + m2 * m2) / (4.0 * ORC_PI);
        rcs_theta[a] = 10.0 * log10(r / lam / lam);
        /* H plane, phi_s = Phi + 90, x and y currents only (:395-456) */
        double ph2 = phi + 90.0;
        far_sums(g, J, (double)a, ph2, S);
        p = ORC_PI * ph2 / 180.0;
        m1 = creal(S[0]) * sin(p) - creal(S[1]) * cos(p);
        m2 = cimag(S[0]) * sin(p) - cimag(S[1]) * cos(p);
        r = g->k0 * g->k0 * ORC_Z0 * ORC_Z0 * (m1 * m1 + m2 * m2) / (4.0 * ORC_PI);
        rcs_phi[a] = 10.0 * log10(r / lam / lam);
    }
}

/* ---- absorbed power over active faces and its adjoint gradient ----------
 * Per face (f, i) with neighbour i + e_f, S = 1/(eps_i - 1) + 1/(eps_{i+e_f} - 1),
 * NearField's absorbed power (vie_fft/NearFar.h:192-205) is
 *   (k0/Z0) Im(-eps_eff) |J|^2 |(-j Z0/(2 k0)) S|^2 dV,  eps_eff = 2/S + 1,
 * and Im(-eps_eff) |S|^2 = -Im(2 conj(S)) = 2 Im(S), so the weight reduces to
 * w = dV Z0/(2 k0) Im(S).
 *
 * Gradient: the product is A = conv + diag(d), d_f[i] = c_f S_f[i] with
 * c_f = -j (delta_f/2)(Z0/k0) (vie_fft/fft.h:419-434).  With A J = b and
 * g = w conj(J),  dP = sum dw |J|^2 + 2 Re(g^T dJ),  dJ = -A^{-1} (dd J), so
 *   dP/da = sum (dw/da) |J|^2 - 2 Re(mu^T (dd/da J)),  mu = A^{-T} g.
 * diag(1/delta_f) A is complex symmetric (masked rows and columns included),
 * hence A^{-T} g = diag(1/delta) A^{-1} diag(delta) g: one solve of A nu =
 * delta g with the same operator, mu = nu / delta.  For a = Re(eps_v):
 * dS/da = -1/(eps_v - 1)^2 on every face touching v; for a = Im(eps_v) the
 * same times j (d is holomorphic in eps; w is not). */
static int face_active(const orc_grid* g, const unsigned char* flags, int f, int i, int j, int k)
{
    const int c[3] = {i, j, k};
    if (c[f] >= g->n[f] - 1) return 0;
    return flags[(long)f * vol(g) + ((long)i * g->n[1] + j) * g->n[2] + k] != 0;
}

double orc_absorption(const orc_grid* g, const double* Jd, const double* epsd, const unsigned char* flags)
{
    const cplx* J = (const cplx*)Jd;
    const cplx* eps = (const cplx*)epsd;
    const long V = vol(g);
    const long st[3] = {(long)g->n[1] * g->n[2], g->n[2], 1};
    const double K = g->d[0] * g->d[1] * g->d[2] * ORC_Z0 / (2.0 * g->k0);
    double P = 0;
    for (int i = 0; i < g->n[0]; ++i)
        for (int j = 0; j < g->n[1]; ++j)
            for (int k = 0; k < g->n[2]; ++k)
                for (int f = 0; f < 3; ++f) {
                    if (!face_active(g, flags, f, i, j, k)) continue;
                    const long v = ((long)i * g->n[1] + j) * g->n[2] + k;
                    const cplx S = 1.0 / (eps[v] - 1.0) + 1.0 / (eps[v + st[f]] - 1.0);
                    const double a = cabs(J[f * V + v]);
                    P += K * cimag(S) * a * a;
                }
    return P;
}

double orc_absorption_grad(const orc_grid* g, orc_op* op, const double* Jd, const double* epsd,
                           const unsigned char* flags, int method, double tol, int maxit,
                           double* grad_re, double* grad_im, orc_info* info)
{
    const cplx* J = (const cplx*)Jd;
    const cplx* eps = (const cplx*)epsd;
    const long V = vol(g);
    const int N1 = g->n[1], N2 = g->n[2];
    const long st[3] = {(long)N1 * N2, N2, 1};
    const double K = g->d[0] * g->d[1] * g->d[2] * ORC_Z0 / (2.0 * g->k0);
    cplx* rhs = calloc(3 * V, sizeof(cplx));
    cplx* nu = calloc(3 * V, sizeof(cplx));
    double P = 0;
    for (int i = 0; i < g->n[0]; ++i)
        for (int j = 0; j < N1; ++j)
            for (int k = 0; k < N2; ++k)
                for (int f = 0; f < 3; ++f) {
                    if (!face_active(g, flags, f, i, j, k)) continue;
                    const long v = ((long)i * N1 + j) * N2 + k;
                    const cplx S = 1.0 / (eps[v] - 1.0) + 1.0 / (eps[v + st[f]] - 1.0);
                    const double w = K * cimag(S);
                    P += w * creal(J[f * V + v] * conj(J[f * V + v]));
                    rhs[f * V + v] = g->d[f] * w * conj(J[f * V + v]);
                }
    orc_info inf;
    orc_solve(method, orc_op_matvec, op, 3 * V, (const double*)rhs, (double*)nu, tol, maxit, NULL, &inf);
    if (info) *info = inf;
    for (long v = 0; v < V; ++v) grad_re[v] = grad_im[v] = 0;
    for (int i = 0; i < g->n[0]; ++i)
        for (int j = 0; j < N1; ++j)
            for (int k = 0; k < N2; ++k)
                for (int f = 0; f < 3; ++f) {
                    if (!face_active(g, flags, f, i, j, k)) continue;
                    const long v = ((long)i * N1 + j) * N2 + k;
                    const long F = f * V + v;
                    const cplx cf = -I * (g->d[f] / 2.0) * ORC_Z0 / g->k0;
                    const cplx mu = nu[F] / g->d[f];
                    const double aJ2 = creal(J[F] * conj(J[F]));
                    const long cells[2] = {v, v + st[f]};
                    for (int s = 0; s < 2; ++s) {
                        const long u = cells[s];
                        const cplx dS = -1.0 / ((eps[u] - 1.0) * (eps[u] - 1.0)); /* dS/de_u */
                        /* a = Re(eps_u): dS/da = dS; a = Im(eps_u): dS/da = j dS */
                        grad_re[u] += K * cimag(dS) * aJ2 - 2.0 * creal(mu * cf * dS * J[F]);
                        grad_im[u] += K * cimag(I * dS) * aJ2 - 2.0 * creal(mu * cf * I * dS * J[F]);
                    }
                }
    free(rhs);
    free(nu);
    return P;
}
