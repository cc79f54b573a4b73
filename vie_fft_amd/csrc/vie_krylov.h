// vie_krylov.h -- vector kernels and device-resident scalar updates of the
// Krylov loops of vie_fft/CG_Iteration.h (paths relative to the reference).
//
// The reference calls CBLAS level-1 routines three times per logical vector
// operation (one per component) and keeps the scalars on the host.  Here a
// vector operation covers all 3V entries in one kernel, dot products are
// fused into the kernel that produces their operand (per-workgroup partial
// sums, summed in a fixed order by a one-workgroup scalar kernel, so results
// are deterministic), and scalars live in device memory (KState), so an
// iteration needs no host synchronisation.  Every kernel exits at once when
// KState::done is set, which lets the host enqueue iterations ahead.
#pragma once

#include "vie_scalar.h"

namespace vie {

// ------------------------------------------------------------ lincomb
// out = sum_i c_i x_i and optionally out2 = sum_i c2_i x_i over n entries
// (c from device coef[0..3], c2 from coef[4..7], or from hc/hc2), with up to
// three fused dot products (operands 0 = out, 1 = e1, 2 = e2, 3 = out2).
// mask: nullptr, or face flags when every input and dot operand is zero on
// inactive faces: those entries are then not read (the outputs there are
// written as zero).
template <class R> struct LinComb {
    long n;
    const cx<R>* x[4];
    const double2* coef;
    double hc[4][2];
    double hc2[4][2];
    cx<R>* out;
    cx<R>* out2;
    Dots<R> dots;
    const uint8_t* mask;
    int wskip;  // with mask: out and out2 already hold zero on inactive faces
    // with mask and wskip, a window: only the entries [wlo, whi) of each of
    // the three components (length wV, all even) hold active faces; the
    // kernel then runs over them alone (grid.y = 3).  wV = 0: every entry.
    long wV, wlo, whi;
    // with a window, a box instead (bzp > 0): only the z-row pairs [bz0, bz0 +
    // 2 bzp) of the rows (x, y), x in [bx0, bx0 + bwx), y in [by0, by0 + bwy),
    // of each component (grid ny x nz = bny x bnz, rows per component bnx bny)
    // hold active faces; 2^blg lanes serve a row (2^blg >= bzp, <= NT)
    int bzp, bx0, bwx, by0, bwy, bz0, bny, bnz, blg;
    // batches (grid.z = launch slots): the vectors hold the instances n
    // entries apart; mask and window are shared
    Bat bt{};
};

// Active range of each z-row of the face flags (rows = 3 Nx Ny, component
// major): zr = lo | hi << 16 with every flag outside [lo, hi) zero; 0 for a
// row without active faces.
__global__ __launch_bounds__(NT) void k_zrange(long rows, int nz, const uint8_t* __restrict__ flags,
                                               uint32_t* __restrict__ zr)
{
    for (long r = (long)blockIdx.x * NT + threadIdx.x; r < rows; r += (long)gridDim.x * NT) {
        const uint8_t* f = flags + r * nz;
        int lo = nz, hi = 0;
        for (int k = 0; k < nz; ++k)
            if (f[k]) {
                lo = lo < k ? lo : k;
                hi = k + 1;
            }
        zr[r] = lo < hi ? (uint32_t)lo | ((uint32_t)hi << 16) : 0u;
    }
}

template <class R>
__global__ __launch_bounds__(NT) void k_lincomb(LinComb<R> L, const int* done)
{
    using T = cx<R>;
    __shared__ double2 red[NT / 64];
    const int inst = bat_inst(L.bt, blockIdx.z);
    done = bat_done(L.bt, done, inst);
    if (done && *done) return;
    {   // the instance's vectors, coefficients and partial sums
        const long o = (long)inst * L.n;
#pragma unroll
        for (int q = 0; q < 4; ++q)
            if (L.x[q]) L.x[q] += o;
        if (L.out) L.out += o;
        if (L.out2) L.out2 += o;
        if (L.dots.e1) L.dots.e1 += o;
        if (L.dots.e2) L.dots.e2 += o;
        if (L.coef) L.coef += (long)inst * L.bt.cstr;
        L.dots.partial += (long)blockIdx.z * 3 * gridDim.x * gridDim.y;
    }
    T c[4], c2[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        if (L.coef) {
            c[q] = mkc((R)L.coef[q].x, (R)L.coef[q].y);
            c2[q] = mkc((R)L.coef[4 + q].x, (R)L.coef[4 + q].y);
        } else {
            c[q] = mkc((R)L.hc[q][0], (R)L.hc[q][1]);
            c2[q] = mkc((R)L.hc2[q][0], (R)L.hc2[q][1]);
        }
    }
    double2 acc[3] = {make_double2(0, 0), make_double2(0, 0), make_double2(0, 0)};
    const long stride = (long)gridDim.x * NT;
    // pairs of entries (16-byte accesses for FP32), then a possible odd tail;
    // with a window, the pairs of component blockIdx.y within it
    const long npair = L.wV ? (L.whi - L.wlo) / 2 : L.n / 2;
    const long base = L.wV ? (long)blockIdx.y * L.wV + L.wlo : 0;
    bool use1 = false, use2 = false;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        if (d >= L.dots.n) break;
        use1 = use1 || L.dots.a[d] == 1 || L.dots.b[d] == 1;
        use2 = use2 || L.dots.a[d] == 2 || L.dots.b[d] == 2;
    }
    // operand selection by value (a conditional over lvalues would select an
    // address and push the candidates to scratch memory)
    auto pick = [](int sel, T v, T e1, T e2, T v2) {
        T r = v2;
        if (sel == 0) r = v;
        if (sel == 1) r = e1;
        if (sel == 2) r = e2;
        return r;
    };
    auto pair = [&](long i) {
        // every load of the pair first (more bytes in flight per thread)
        T a0[4], a1[4], e10 = czero<T>(), e11 = czero<T>(), e20 = czero<T>(), e21 = czero<T>();
        const bool act = !L.mask || *reinterpret_cast<const uint16_t*>(L.mask + i) != 0;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            a0[q] = a1[q] = czero<T>();
            if (L.x[q] && act) ld2(L.x[q] + i, a0[q], a1[q]);
        }
        if (use1 && act) ld2(L.dots.e1 + i, e10, e11);
        if (use2 && act) ld2(L.dots.e2 + i, e20, e21);
        T v = czero<T>(), v2 = czero<T>(), w = czero<T>(), w2 = czero<T>();
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            v = cadd(v, cmul(c[q], a0[q]));
            w = cadd(w, cmul(c[q], a1[q]));
            v2 = cadd(v2, cmul(c2[q], a0[q]));
            w2 = cadd(w2, cmul(c2[q], a1[q]));
        }
        if (L.dots.n > 0) {
            dot_acc(acc[0], pick(L.dots.a[0], v, e10, e20, v2), pick(L.dots.b[0], v, e10, e20, v2), L.dots.nc[0]);
            dot_acc(acc[0], pick(L.dots.a[0], w, e11, e21, w2), pick(L.dots.b[0], w, e11, e21, w2), L.dots.nc[0]);
        }
        if (L.dots.n > 1) {
            dot_acc(acc[1], pick(L.dots.a[1], v, e10, e20, v2), pick(L.dots.b[1], v, e10, e20, v2), L.dots.nc[1]);
            dot_acc(acc[1], pick(L.dots.a[1], w, e11, e21, w2), pick(L.dots.b[1], w, e11, e21, w2), L.dots.nc[1]);
        }
        if (L.dots.n > 2) {
            dot_acc(acc[2], pick(L.dots.a[2], v, e10, e20, v2), pick(L.dots.b[2], v, e10, e20, v2), L.dots.nc[2]);
            dot_acc(acc[2], pick(L.dots.a[2], w, e11, e21, w2), pick(L.dots.b[2], w, e11, e21, w2), L.dots.nc[2]);
        }
        if (act || !L.wskip) {
            if (L.out) st2(L.out + i, v, w);
            if (L.out2) st2(L.out2 + i, v2, w2);
        }
    };
    if (L.wV && L.bzp) {  // the window box: rows of the x and y windows, the z window's pairs
        const int lanes = 1 << L.blg, rpb = NT >> L.blg;
        const int lane = threadIdx.x & (lanes - 1);
        const long nrows = (long)L.bwx * L.bwy;
        const long cbase = (long)blockIdx.y * L.wV + L.bz0;
        for (long r = (long)blockIdx.x * rpb + (threadIdx.x >> L.blg); r < nrows; r += (long)gridDim.x * rpb) {
            const long row = (long)(L.bx0 + (int)(r / L.bwy)) * L.bny + L.by0 + (int)(r % L.bwy);
            for (int zp = lane; zp < L.bzp; zp += lanes) pair(cbase + row * L.bnz + 2 * zp);
        }
    } else {
        for (long i2 = (long)blockIdx.x * NT + threadIdx.x; i2 < npair; i2 += stride) pair(base + 2 * i2);
    }
    if (!L.wV && blockIdx.x == 0 && threadIdx.x == 0 && (L.n & 1)) {
        const long i = L.n - 1;
        T v = czero<T>(), v2 = czero<T>();
        for (int q = 0; q < 4; ++q)
            if (L.x[q]) {
                const T xv = L.x[q][i];
                v = cadd(v, cmul(c[q], xv));
                if (L.out2) v2 = cadd(v2, cmul(c2[q], xv));
            }
        const T e1 = use1 ? L.dots.e1[i] : czero<T>(), e2 = use2 ? L.dots.e2[i] : czero<T>();
        if (L.dots.n > 0) dot_acc(acc[0], pick(L.dots.a[0], v, e1, e2, v2), pick(L.dots.b[0], v, e1, e2, v2), L.dots.nc[0]);
        if (L.dots.n > 1) dot_acc(acc[1], pick(L.dots.a[1], v, e1, e2, v2), pick(L.dots.b[1], v, e1, e2, v2), L.dots.nc[1]);
        if (L.dots.n > 2) dot_acc(acc[2], pick(L.dots.a[2], v, e1, e2, v2), pick(L.dots.b[2], v, e1, e2, v2), L.dots.nc[2]);
        if (L.out) L.out[i] = v;
        if (L.out2) L.out2[i] = v2;
    }
    #pragma unroll
    for (int d = 0; d < 3; ++d) {
        if (d >= L.dots.n) break;
        double2 s = block_sum(acc[d], red);
        if (threadIdx.x == 0)
            L.dots.partial[d * (long)gridDim.x * gridDim.y + (long)blockIdx.y * gridDim.x + blockIdx.x] = s;
    }
}

// One solver scalar step.  The state is staged through LDS: every thread
// loads one word of it alongside the partial sums (one memory latency for
// both), thread 0 runs the step on the LDS copy, and every thread writes its
// word back -- instead of a chain of dependent global accesses by one thread.
// Batches: workgroup b runs the step of launch slot b's instance (partial
// sums of slot b: 3 nblk entries after slot b - 1's), unless that instance's
// flag in mask (nullptr: none; spaced like the done flags) is set.
__global__ __launch_bounds__(NTS) void k_scalar(int op, KState* gs, const double2* part, long nblk,
                                                 int nd, double* hist, Bat bt, const int* mask)
{
    constexpr int W = sizeof(KState) / sizeof(unsigned long long);
    static_assert(sizeof(KState) % sizeof(unsigned long long) == 0 && W <= NTS, "state words");
    const int inst = bat_inst(bt, blockIdx.x);
    mask = bat_done(bt, mask, inst);
    if (mask && *mask) return;
    gs += inst;
    part += (long)blockIdx.x * 3 * nblk;
    __shared__ KState sh;
    __shared__ double2 v[4];
    __shared__ double2 red[4][NTS / 64];
    unsigned long long* shw = reinterpret_cast<unsigned long long*>(&sh);
    unsigned long long* gw = reinterpret_cast<unsigned long long*>(gs);
    const unsigned long long mine = threadIdx.x < W ? gw[threadIdx.x] : 0ull;
    reduce_partials(part, nblk, nd, v, red);
    if (threadIdx.x < W) shw[threadIdx.x] = mine;
    __syncthreads();
    if (sh.done) return;
    if (threadIdx.x == 0) scalar_program(op, &sh, v, nd, hist);
    __syncthreads();
    if (threadIdx.x < W) gw[threadIdx.x] = shw[threadIdx.x];
}


// Graph-safe fill (used instead of hipMemsetAsync).
template <class R>
__global__ void k_fill_zero(long n, cx<R>* x)
{
    for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (long)gridDim.x * blockDim.x)
        x[t] = czero<cx<R>>();
}

// x *= s over n complex FP64 values.
__global__ void k_scale_c64(long n, double s, double2* x)
{
    for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (long)gridDim.x * blockDim.x) {
        x[t].x *= s;
        x[t].y *= s;
    }
}

// Zero a 3-component field (V entries per component) outside the entries
// [lo, hi) of each component: the x planes outside a product's x window.
template <class R>
__global__ void k_zero_outside(long V, long lo, long hi, cx<R>* x, const int* done)
{
    if (done && *done) return;
    const long out = V - (hi - lo);
    for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < 3 * out; t += (long)gridDim.x * blockDim.x) {
        const long c = t / out, r = t % out;
        x[c * V + (r < lo ? r : r + (hi - lo))] = czero<cx<R>>();
    }
}

// x64 += s * e (e in the inner precision): the correction update of the
// FP64 iterative refinement.
template <class R>
__global__ void k_axpy_to64(long n, double s, const cx<R>* e, double2* x)
{
    for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (long)gridDim.x * blockDim.x) {
        x[t].x += s * (double)e[t].x;
        x[t].y += s * (double)e[t].y;
    }
}

// x64 += m (.) e (m and e in the inner precision): the correction update of
// the mixed solve when its inner iterations are right-preconditioned (the
// inner iterate e lives in the preconditioned variables, x = m (.) e).
template <class R>
__global__ void k_axpy_scaled_to64(long n, const cx<R>* m, const cx<R>* e, double2* x)
{
    for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (long)gridDim.x * blockDim.x) {
        const double mx = (double)m[t].x, my = (double)m[t].y, ex = (double)e[t].x, ey = (double)e[t].y;
        x[t].x += mx * ex - my * ey;
        x[t].y += mx * ey + my * ex;
    }
}

// x = m (.) x: a right-preconditioned solve's iterate back to the unknowns.
template <class R>
__global__ void k_cmul_diag(long n, const cx<R>* m, cx<R>* x)
{
    for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (long)gridDim.x * blockDim.x)
        x[t] = cmul(m[t], x[t]);
}

// Jacobi preconditioner of the diagonal term: m_f = 1/d_f on active faces, 0
// elsewhere (computed in FP64 from the operator's own d_f).  *bad counts the
// active faces whose d_f is zero or whose reciprocal is not finite.
template <class R>
__global__ void k_jacobi(long n, const cx<R>* d, const uint8_t* flags, cx<R>* m, int* bad)
{
    int nb = 0;
    for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (long)gridDim.x * blockDim.x) {
        double2 r = make_double2(0, 0);
        if (flags[t]) {
            r = zdiv(make_double2(1, 0), make_double2((double)d[t].x, (double)d[t].y));
            if (!(isfinite(r.x) && isfinite(r.y)) || (d[t].x == 0 && d[t].y == 0)) {
                ++nb;
                r = make_double2(0, 0);
            }
        }
        m[t] = mkc((R)r.x, (R)r.y);
    }
    if (nb) atomicAdd(bad, nb);
}

// out = (R)(s * in): scaled conversion of an FP64 residual to the inner precision.
template <class R>
__global__ void k_scale_from64(long n, double s, const double2* in, cx<R>* out)
{
    for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (long)gridDim.x * blockDim.x)
        out[t] = mkc((R)(s * in[t].x), (R)(s * in[t].y));
}

// Let a solve that stopped on its recursive residual continue (after a
// residual replacement); a breakdown or the iteration limit stays final.
__global__ void k_state_resume(KState* s, double pause)
{
    if (threadIdx.x == 0 && blockIdx.x == 0 && s->status == KS_OK && s->iter < s->maxit) {
        s->done = 0;
        s->converged = 0;
        s->pause = pause;
    }
}

// KState initialisation from the host-side tolerance, limit and pause level.
__global__ void k_state_init(KState* s, double tol, int maxit, double pause, double dvg)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        KState z = {};
        z.tol = tol;
        z.maxit = maxit;
        z.pause = pause;
        z.dvg = dvg;
        *s = z;
    }
}
// The same for nb instances: tolerance and limit tolv[i] and maxitv[i], or
// tol and maxit for every instance when those are nullptr.
__global__ void k_state_init_n(KState* s, int nb, const double* tolv, const int* maxitv, double tol, int maxit,
                               double pause, double dvg)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < nb) {
        KState z = {};
        z.tol = tolv ? tolv[i] : tol;
        z.maxit = maxitv ? maxitv[i] : maxit;
        z.pause = pause;
        z.dvg = dvg;
        s[i] = z;
    }
}

// Launch slots: the instances whose records are not done first (in index
// order), then the rest, so that a launch over the first m slots, m at least
// the number not done, serves every instance still iterating.  One workgroup.
__global__ __launch_bounds__(NTS) void k_ilist(const KState* s, int nb, int* ilist)
{
    __shared__ int pre[NTS + 1];
    // thread t owns the instances [lo, hi)
    const int per = (nb + NTS - 1) / NTS;
    const int lo = min(nb, (int)threadIdx.x * per), hi = min(nb, lo + per);
    int mine = 0;
    for (int i = lo; i < hi; ++i) mine += s[i].done ? 0 : 1;
    pre[threadIdx.x + 1] = mine;
    __syncthreads();
    if (threadIdx.x == 0) {  // prefix sums of the active counts
        pre[0] = 0;
        for (int t = 0; t < NTS; ++t) pre[t + 1] += pre[t];
    }
    __syncthreads();
    const int nact = pre[NTS];
    int a = pre[threadIdx.x];          // active instances before lo
    int d = nact + (lo - a);           // slot of the first done instance at or after lo
    for (int i = lo; i < hi; ++i) {
        if (s[i].done) ilist[d++] = i;
        else ilist[a++] = i;
    }
}
// Mixed-precision batches (reliable updates, one FP32 BiCGSTAB per
// instance): the instances whose FP32 iterations have stopped and that are
// not finished get a residual replacement.  ax[i].done = 0 selects them for
// the FP64 launches that follow (the others' workgroups exit at once).
__global__ void k_mix_select(const KState* st, KState* ax, int nb)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < nb) ax[i].done = (st[i].done && !st[i].fin) ? 0 : 1;
}
// After the replacement residual of the selected instances is known
// (ax[i].err): an instance whose true residual meets its tolerance, or whose
// iterations ended for good (breakdown, limit), is finished; every other
// selected one resumes its FP32 iterations with pause level delta * err, and
// keeps ax[i].done = 0 for the launches that load its new residual.
__global__ void k_mix_decide(KState* st, KState* ax, int nb, double delta)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nb || ax[i].done) return;
    const double t = ax[i].err;
    st[i].nrep += 1;
    st[i].terr = t;
    const bool stop = st[i].status != KS_OK || st[i].iter >= st[i].maxit;
    if (t < st[i].tol || stop || !isfinite(t)) {
        st[i].fin = 1;
        ax[i].done = 1;
    } else {
        st[i].done = 0;
        st[i].converged = 0;
        st[i].pause = delta * t;
    }
}

// Batched element-wise updates of the residual replacements (grid.z = launch
// slots, done as in the pass kernels; n entries per instance): over the
// entries [wlo, wlo + wn) of each of the three components (grid.y = 3, V
// entries per component) whose face is active -- every other entry of the
// operands is zero and stays so.  x64 += e, e = 0 (the reliable update), and
// out = (R) in.
template <class R>
__global__ void k_fold64_b(long n, long V, long wlo, long wn, const uint8_t* __restrict__ mask, Bat bt,
                           const int* done, cx<R>* e, double2* x)
{
    const int inst = bat_inst(bt, blockIdx.z);
    done = bat_done(bt, done, inst);
    if (done && *done) return;
    const long base = (long)inst * n + (long)blockIdx.y * V + wlo;
    const uint8_t* m = mask + (long)blockIdx.y * V + wlo;
    for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < wn; t += (long)gridDim.x * blockDim.x) {
        if (!m[t]) continue;
        const cx<R> v = e[base + t];
        x[base + t].x += (double)v.x;
        x[base + t].y += (double)v.y;
        e[base + t] = czero<cx<R>>();
    }
}
template <class R>
__global__ void k_convert_b(long n, long V, long wlo, long wn, const uint8_t* __restrict__ mask, Bat bt,
                            const int* done, const double2* in, cx<R>* out)
{
    const int inst = bat_inst(bt, blockIdx.z);
    done = bat_done(bt, done, inst);
    if (done && *done) return;
    const long base = (long)inst * n + (long)blockIdx.y * V + wlo;
    const uint8_t* m = mask + (long)blockIdx.y * V + wlo;
    for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < wn; t += (long)gridDim.x * blockDim.x)
        if (m[t]) out[base + t] = mkc((R)in[base + t].x, (R)in[base + t].y);
}

// ------------------------------------------------------------ Green's function
// g on the (Nx+1)(Ny+1)(Nz+1) offsets (vie_fft/GreenFunction.h:53-88).
// Several wavenumbers in one launch: with k0v, workgroup row y computes
// k0 = k0v[y] into out + y * total.
__global__ void k_green_offsets(int n0, int n1, int n2, double d0, double d1, double d2,
                                double k0, const double* k0v, double2* out)
{
    const long total = (long)(n0 + 1) * (n1 + 1) * (n2 + 1);
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= total) return;
    if (k0v) {
        k0 = k0v[blockIdx.y];
        out += blockIdx.y * total;
    }
    const int k = t % (n2 + 1);
    const int j = (t / (n2 + 1)) % (n1 + 1);
    const int i = t / ((long)(n1 + 1) * (n2 + 1));
    const double pi = 3.1415926535897932384626433832795;
    const double dV = d0 * d1 * d2;
    const double D = 2.0 * cbrt(dV / (4.0 / 3.0 * pi));
    const double p = 0.5 * k0 * D;
    double2 v;
    if (i + j + k != 0) {
        const double x = i * d0, y = j * d1, z = k * d2;
        const double R = sqrt(x * x + y * y + z * z);
        const double sh = sin(p) / p - cos(p);
        const double a = sh / (pi / 3.0 * k0 * k0 * D * D * R) * dV;
        double sr, cr;
        sincos(k0 * R, &sr, &cr);
        v = make_double2(cr * a, -sr * a);
    } else {
        double sp, cp;
        sincos(p, &sp, &cp);
        // (1 + j p)(cos p - j sin p) - 1
        const double re = cp + p * sp - 1.0, im = p * cp - sp;
        const double den = pi / 6.0 * k0 * k0 * D * D * D;
        v = make_double2(re / den * dV, im / den * dV);
    }
    out[t] = v;
}

// Cosine transform along one axis of an (L0, L1, L2) array, L = N + 1 on that
// axis: out[p] = sum_n w_n cos(pi k(p) n / N) in[n], w_0 = w_N = 1, else 2,
// where k(p) is the frequency stored at parity position p (evens then odds).
// Applied once per axis this is the forward DFT of the even extension of g
// to 2N points, restricted to the octant 0 <= k <= N that evenness makes
// unique (vie_fft/GreenFunction.h:91-139).  ctab[j] = cos(pi j / N), j < 2N.
// Workgroup row y transforms array y of a batch (in, out + y * total).
__global__ void k_green_cos(int L0, int L1, int L2, int axis, const double* ctab,
                            const double2* in, double2* out)
{
    const long total = (long)L0 * L1 * L2;
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= total) return;
    in += blockIdx.y * total;
    out += blockIdx.y * total;
    const int c2 = t % L2, c1 = (t / L2) % L1, c0 = t / ((long)L1 * L2);
    const int L = axis == 0 ? L0 : axis == 1 ? L1 : L2;
    const int N = L - 1;
    const int pos = axis == 0 ? c0 : axis == 1 ? c1 : c2;
    const int neven = N / 2 + 1;
    const int kf = pos < neven ? 2 * pos : 2 * (pos - neven) + 1;
    const long st = axis == 0 ? (long)L1 * L2 : axis == 1 ? L2 : 1;
    const long base = t - (long)pos * st;
    double2 acc = make_double2(0, 0);
    for (int n = 0; n <= N; ++n) {
        const double w = (n == 0 || n == N) ? 1.0 : 2.0;
        const double c = w * ctab[((long)kf * n) % (2 * N)];
        const double2 a = in[base + (long)n * st];
        acc.x += c * a.x;
        acc.y += c * a.y;
    }
    out[t] = acc;
}

// Element-wise kernels are grid-stride loops: launches cap the grid size.
template <class R>
__global__ void k_convert(long n, const double2* in, cx<R>* out)
{
    for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (long)gridDim.x * blockDim.x)
        out[t] = mkc((R)in[t].x, (R)in[t].y);
}

template <class R>
__global__ void k_to_f64(long n, const cx<R>* in, double2* out)
{
    for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (long)gridDim.x * blockDim.x)
        out[t] = make_double2((double)in[t].x, (double)in[t].y);
}

// Mask a field by the flags: y = flag ? x : 0 (Current_Reset, vie_fft/fft.h:13-35)
template <class R>
__global__ void k_mask(long n, const cx<R>* x, const uint8_t* flags, cx<R>* y)
{
    for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (long)gridDim.x * blockDim.x)
        y[t] = flags[t] ? x[t] : czero<cx<R>>();
}

// Far-field sums of FarField1 (vie_fft/NearFar.h:59-96): per-block partials of
// sum_cells J_c exp(j k0 r.v) for the three components.
__global__ __launch_bounds__(NT) void k_farsum(int n0, int n1, int n2, double d0, double d1,
                                               double d2, double k0, double vx, double vy,
                                               double vz, const double2* J, double2* partial)
{
    __shared__ double2 red[NT / 64];
    const long V = (long)n0 * n1 * n2;
    double2 acc[3] = {make_double2(0, 0), make_double2(0, 0), make_double2(0, 0)};
    for (long v = (long)blockIdx.x * NT + threadIdx.x; v < V; v += (long)gridDim.x * NT) {
        const int k = v % n2, j = (v / n2) % n1, i = v / ((long)n1 * n2);
        const double a = k0 * ((i + 0.5) * d0 * vx + (j + 0.5) * d1 * vy + (k + 0.5) * d2 * vz);
        double sa, ca;
        sincos(a, &sa, &ca);
        for (int c = 0; c < 3; ++c) {
            const double2 jv = J[c * V + v];
            acc[c].x += ca * jv.x - sa * jv.y;
            acc[c].y += ca * jv.y + sa * jv.x;
        }
    }
    for (int c = 0; c < 3; ++c) {
        double2 s = block_sum(acc[c], red);
        if (threadIdx.x == 0) partial[c * (long)gridDim.x + blockIdx.x] = s;
    }
}

// Far-field sums for many directions in one pass over J (FarField2,
// vie_fft/NearFar.h:302-456): per direction v, sum_cells J_c exp(j k0 r.v)
// with r the cell centre.  Each block walks z-pencils (i, j); each thread owns
// up to FF_MAXD directions and evaluates a pencil by Horner's rule in
// w = exp(j k0 dz vz), then applies the pencil's phase.  Partials are laid out
// [block][component][direction].
constexpr int FF_MAXD = 4;
__global__ __launch_bounds__(NT) void k_farfield(int n0, int n1, int n2, double d0, double d1, double d2, double k0,
                                                 int ndir, const double* __restrict__ vdir,
                                                 const double2* __restrict__ J, double2* __restrict__ partial)
{
    extern __shared__ double2 pen[];  // 3 * n2
    const long V = (long)n0 * n1 * n2;
    double2 acc[FF_MAXD][3];
    double2 w[FF_MAXD];
    double vx[FF_MAXD], vy[FF_MAXD], vz[FF_MAXD];
#pragma unroll
    for (int q = 0; q < FF_MAXD; ++q) {
        const int dir = threadIdx.x + q * NT;
        vx[q] = vy[q] = vz[q] = 0;
        if (dir < ndir) {
            vx[q] = vdir[3 * dir];
            vy[q] = vdir[3 * dir + 1];
            vz[q] = vdir[3 * dir + 2];
        }
        double sn, cs;
        sincos(k0 * d2 * vz[q], &sn, &cs);
        w[q] = make_double2(cs, sn);
        for (int c = 0; c < 3; ++c) acc[q][c] = make_double2(0, 0);
    }
    const long npen = (long)n0 * n1;
    for (long p = blockIdx.x; p < npen; p += gridDim.x) {
        const int i = (int)(p / n1), j = (int)(p % n1);
        __syncthreads();
        for (int e = threadIdx.x; e < 3 * n2; e += NT) {
            const int c = e / n2, k = e % n2;
            pen[e] = J[c * V + p * n2 + k];
        }
        __syncthreads();
#pragma unroll
        for (int q = 0; q < FF_MAXD; ++q) {
            if (threadIdx.x + q * NT >= ndir) break;
            double2 s[3] = {make_double2(0, 0), make_double2(0, 0), make_double2(0, 0)};
            for (int k = n2 - 1; k >= 0; --k)
                for (int c = 0; c < 3; ++c) {
                    const double2 t = zmul(s[c], w[q]);
                    const double2 jv = pen[c * n2 + k];
                    s[c] = make_double2(t.x + jv.x, t.y + jv.y);
                }
            double sn, cs;
            sincos(k0 * ((i + 0.5) * d0 * vx[q] + (j + 0.5) * d1 * vy[q] + 0.5 * d2 * vz[q]), &sn, &cs);
            const double2 ph = make_double2(cs, sn);
            for (int c = 0; c < 3; ++c) {
                const double2 t = zmul(s[c], ph);
                acc[q][c].x += t.x;
                acc[q][c].y += t.y;
            }
        }
    }
#pragma unroll
    for (int q = 0; q < FF_MAXD; ++q) {
        const int dir = threadIdx.x + q * NT;
        if (dir >= ndir) break;
        for (int c = 0; c < 3; ++c) partial[((long)blockIdx.x * 3 + c) * ndir + dir] = acc[q][c];
    }
}

// Column sums of a [rows][cols] double2 array in row order (deterministic).
__global__ __launch_bounds__(NT) void k_colsum(long rows, int cols, const double2* __restrict__ in,
                                               double2* __restrict__ out)
{
    const int col = blockIdx.x * NT + threadIdx.x;
    if (col >= cols) return;
    double2 s = make_double2(0, 0);
    for (long r = 0; r < rows; ++r) {
        const double2 v = in[r * cols + col];
        s.x += v.x;
        s.y += v.y;
    }
    out[col] = s;
}

// Absorbed power of NearField (vie_fft/NearFar.h:186-250) for faces with
// i,j,k < N-1 whose two cells both carry material `mat`.
// (b0, b1, b2): the caller's grid, whose N - 1 bounds the loops (the compiled
// grid n may be larger).
__global__ __launch_bounds__(NT) void k_absorb(int n0, int n1, int n2, int b0, int b1, int b2, double dV, double k0,
                                               double z0, const double2* J, const double2* eps,
                                               const uint8_t* mat, int mid, double2* partial)
{
    __shared__ double2 red[NT / 64];
    const long V = (long)n0 * n1 * n2;
    const long st[3] = {(long)n1 * n2, n2, 1};
    double acc = 0;
    for (long v = (long)blockIdx.x * NT + threadIdx.x; v < V; v += (long)gridDim.x * NT) {
        const int k = v % n2, j = (v / n2) % n1, i = v / ((long)n1 * n2);
        if (i >= b0 - 1 || j >= b1 - 1 || k >= b2 - 1) continue;
        for (int f = 0; f < 3; ++f) {
            const long w = v + st[f];
            if (!(mat[w] == mid && mat[v] == mid)) continue;
            const double2 a = zdiv(make_double2(1, 0), make_double2(eps[w].x - 1.0, eps[w].y));
            const double2 b = zdiv(make_double2(1, 0), make_double2(eps[v].x - 1.0, eps[v].y));
            const double2 S = make_double2(a.x + b.x, a.y + b.y);
            const double absS = sqrt(S.x * S.x + S.y * S.y);
            const double2 jv = J[f * V + v];
            const double aE = sqrt(jv.x * jv.x + jv.y * jv.y) * absS * (0.5 * z0 / k0);
            const double2 ee = zdiv(make_double2(2, 0), S);
            // Im(-(ee + 1)) = -Im(ee)
            acc += k0 / z0 * (-ee.y) * aE * aE * dV;
        }
    }
    double2 s = block_sum(make_double2(acc, 0), red);
    if (threadIdx.x == 0) partial[blockIdx.x] = s;
}


// ------------------------------------------------- caller grid <-> internal grid
// The internal grid holds the caller's grid at its origin with the axes in the
// order perm (internal axis a is the caller's axis perm[a], and component a,
// the faces normal to that axis, is the caller's component perm[a]); a length
// without compiled pass kernels is embedded in the next compiled one, whose
// extra cells are inactive background.
struct Remap {
    int n[3];      // internal grid
    int u[3];      // caller grid, caller axis order
    int perm[3];   // internal axis -> caller axis
    int iperm[3];  // caller axis -> internal axis
    long V, Vu;
};

// k_embed writes ncomp internal fields from the caller's (zero on the
// embedding cells); k_crop writes the caller's cells back.  A field of three
// components follows perm, a single one is component 0 on both sides.
template <class T>
__global__ __launch_bounds__(NT) void k_embed(Remap m, int ncomp, const T* __restrict__ in, T* __restrict__ out)
{
    const long cs[3] = {(long)m.u[1] * m.u[2], m.u[2], 1};
    for (long e = (long)blockIdx.x * NT + threadIdx.x; e < ncomp * m.V; e += (long)gridDim.x * NT) {
        const long c = e / m.V, v = e % m.V;
        const int i[3] = {(int)(v / ((long)m.n[1] * m.n[2])), (int)((v / m.n[2]) % m.n[1]), (int)(v % m.n[2])};
        bool in_grid = true;
        long off = (ncomp == 3 ? m.perm[c] : c) * m.Vu;
        for (int a = 0; a < 3; ++a) {
            in_grid = in_grid && i[a] < m.u[m.perm[a]];
            off += i[a] * cs[m.perm[a]];
        }
        // a field's component c leaves its own-axis last plane of the caller's
        // grid out of the operator (Current_Expansion, vie_fft/CG_Iteration.h:
        // 59-63); embedded, that plane is interior, so it is dropped here
        if (ncomp == 3 && i[c] == m.u[m.perm[c]] - 1) in_grid = false;
        T r{};
        if (in_grid) r = in[off];
        out[e] = r;
    }
}
template <class T>
__global__ __launch_bounds__(NT) void k_crop(Remap m, int ncomp, const T* __restrict__ in, T* __restrict__ out)
{
    const long is[3] = {(long)m.n[1] * m.n[2], m.n[2], 1};
    for (long e = (long)blockIdx.x * NT + threadIdx.x; e < ncomp * m.Vu; e += (long)gridDim.x * NT) {
        const long c = e / m.Vu, v = e % m.Vu;
        const int r[3] = {(int)(v / ((long)m.u[1] * m.u[2])), (int)((v / m.u[2]) % m.u[1]), (int)(v % m.u[2])};
        long off = (ncomp == 3 ? m.iperm[c] : c) * m.V;
        for (int b = 0; b < 3; ++b) off += r[b] * is[m.iperm[b]];
        out[e] = in[off];
    }
}

// ------------------------------------------------------------ set-up kernels
// Analytic shapes (GetMaterial / GetFlag, vie_fft/Model.h:305-398): a sphere
// holds a point when |p - c|^2 <= r^2 (Model.h:229), a band of z planes when
// z0 <= z < z1; points are in cell units, the first shape that holds a point
// gives its material, and points outside the box [0, n] are background (0).
struct VoxShape {
    int kind;  // 0 sphere, 1 z band
    int mat;
    double cx, cy, cz, r, z0, z1;
};
__device__ inline int vox_material(const VoxShape* sh, int ns, int n0, int n1, int n2, double x, double y, double z)
{
#pragma clang fp contract(off)
    if (!(x >= 0.0 && x <= n0 && y >= 0.0 && y <= n1 && z >= 0.0 && z <= n2)) return 0;
    for (int s = 0; s < ns; ++s) {
        const VoxShape& q = sh[s];
        if (q.kind == 0) {
            // rounded like ((x - cx)^2 + (y - cy)^2) + (z - cz)^2 <= r r, every
            // product and sum rounded on its own (no fused multiply-add)
            const double dx = x - q.cx, dy = y - q.cy, dz = z - q.cz;
            const double d2 = (dx * dx + dy * dy) + dz * dz;
            if (d2 <= q.r * q.r) return q.mat;
        } else if (z >= q.z0 && z < q.z1) {
            return q.mat;
        }
    }
    return 0;
}
// Material at the cell centres ((i, j, k) + 1/2) and the 3-point face flags:
// face f of cell v is active when the face centre and the centres of the two
// cells it joins are all non-background (GetFlag, vie_fft/Model.h:345-394).
__global__ __launch_bounds__(NT) void k_voxelize(int n0, int n1, int n2, int ns, const VoxShape* __restrict__ sh,
                                                 uint8_t* __restrict__ mat, uint8_t* __restrict__ flags)
{
    const long V = (long)n0 * n1 * n2;
    for (long v = (long)blockIdx.x * NT + threadIdx.x; v < V; v += (long)gridDim.x * NT) {
        const int i = (int)(v / ((long)n1 * n2)), j = (int)((v / n2) % n1), k = (int)(v % n2);
        const double c[3] = {i + 0.5, j + 0.5, k + 0.5};
        const int m = vox_material(sh, ns, n0, n1, n2, c[0], c[1], c[2]);
        if (mat) mat[v] = (uint8_t)m;
        for (int f = 0; f < 3; ++f) {
            double p[3] = {c[0], c[1], c[2]};
            p[f] = c[f] + 0.5;
            const bool face = vox_material(sh, ns, n0, n1, n2, p[0], p[1], p[2]) != 0;
            p[f] = c[f] + 1.0;
            const bool up = vox_material(sh, ns, n0, n1, n2, p[0], p[1], p[2]) != 0;
            flags[f * V + v] = (uint8_t)(m != 0 && face && up);
        }
    }
}


// Per-voxel permittivity from material ids and a table (GetMaterial's
// Material with Real_EPR/Imag_EPR, vie_fft/Model.h:17,82-83).
__global__ __launch_bounds__(NT) void k_eps_table(long V, const uint8_t* __restrict__ mat,
                                                  const double2* __restrict__ tab, double2* __restrict__ eps)
{
    for (long v = (long)blockIdx.x * NT + threadIdx.x; v < V; v += (long)gridDim.x * NT) eps[v] = tab[mat[v]];
}

// Diagonal face term d_f[v] = c_f (1/(eps[v + e_f] - 1) + 1/(eps[v] - 1)) for
// v_f < N_f - 1, else 0 (LXX_DO/LYY_DO/LZZ_DO, vie_fft/fft.h:405-588);
// c_f = j cim[f].
template <class R>
__global__ __launch_bounds__(NT) void k_diag(int n0, int n1, int n2, double c0, double c1, double c2,
                                             const double2* __restrict__ eps, cx<R>* __restrict__ d)
{
    const long V = (long)n0 * n1 * n2;
    const long st[3] = {(long)n1 * n2, n2, 1};
    const int nn[3] = {n0, n1, n2};
    const double cim[3] = {c0, c1, c2};
    for (long v = (long)blockIdx.x * NT + threadIdx.x; v < V; v += (long)gridDim.x * NT) {
        const int c[3] = {(int)(v / ((long)n1 * n2)), (int)((v / n2) % n1), (int)(v % n2)};
        const double2 a = zdiv(make_double2(1, 0), make_double2(eps[v].x - 1.0, eps[v].y));
        for (int f = 0; f < 3; ++f) {
            double2 r = make_double2(0, 0);
            if (c[f] < nn[f] - 1) {
                const double2 e = eps[v + st[f]];
                const double2 b = zdiv(make_double2(1, 0), make_double2(e.x - 1.0, e.y));
                const double2 sm = make_double2(a.x + b.x, a.y + b.y);
                r = make_double2(-cim[f] * sm.y, cim[f] * sm.x);  // j cim * sm
            }
            d[f * V + v] = mkc((R)r.x, (R)r.y);
        }
    }
}

// k_diag for the instances of a batch in one launch (instance blockIdx.y +
// k gridDim.y), both precisions: instance i's permittivity is read through
// the material ids from its table tabs[256 i ..], its c_f is j cims[3 i + f],
// and its d_f goes to d64 / d32 + i istr.  Same arithmetic as k_diag.
__global__ __launch_bounds__(NT) void k_diag_batch(int cnt, int n0, int n1, int n2, const double* __restrict__ cims,
                                                   const uint8_t* __restrict__ mat, const double2* __restrict__ tabs,
                                                   double2* __restrict__ d64, float2* __restrict__ d32, long istr)
{
    const long V = (long)n0 * n1 * n2;
    const long st[3] = {(long)n1 * n2, n2, 1};
    const int nn[3] = {n0, n1, n2};
    for (int i = blockIdx.y; i < cnt; i += gridDim.y) {
        const double2* tab = tabs + 256L * i;
        const double cim[3] = {cims[3 * i], cims[3 * i + 1], cims[3 * i + 2]};
        double2* o64 = d64 + (long)i * istr;
        float2* o32 = d32 + (long)i * istr;
        for (long v = (long)blockIdx.x * NT + threadIdx.x; v < V; v += (long)gridDim.x * NT) {
            const int c[3] = {(int)(v / ((long)n1 * n2)), (int)((v / n2) % n1), (int)(v % n2)};
            const double2 ev = tab[mat[v]];
            const double2 a = zdiv(make_double2(1, 0), make_double2(ev.x - 1.0, ev.y));
            for (int f = 0; f < 3; ++f) {
                double2 r = make_double2(0, 0);
                if (c[f] < nn[f] - 1) {
                    const double2 e = tab[mat[v + st[f]]];
                    const double2 b = zdiv(make_double2(1, 0), make_double2(e.x - 1.0, e.y));
                    const double2 sm = make_double2(a.x + b.x, a.y + b.y);
                    r = make_double2(-cim[f] * sm.y, cim[f] * sm.x);  // j cim * sm
                }
                o64[f * V + v] = r;
                o32[f * V + v] = make_float2((float)r.x, (float)r.y);
            }
        }
    }
}

// Incident_Wave (vie_fft/GreenFunction.h:144-248) masked by the flags
// (vie_fft/CG_Iteration.h:559-562): b_c = Flag_c exp(-j k.r_c) seg_c, r_c the
// face centre of component c, seg_c its pulse integral times the polarisation.
__global__ __launch_bounds__(NT) void k_incident(int n0, int n1, int n2, double d0, double d1, double d2,
                                                 double k0x, double k0y, double k0z, double2 s0, double2 s1,
                                                 double2 s2, const uint8_t* __restrict__ flags,
                                                 double2* __restrict__ b)
{
    const long V = (long)n0 * n1 * n2;
    for (long v = (long)blockIdx.x * NT + threadIdx.x; v < V; v += (long)gridDim.x * NT) {
        const int i = (int)(v / ((long)n1 * n2)), j = (int)((v / n2) % n1), k = (int)(v % n2);
        const double px = (i + 0.5) * d0, py = (j + 0.5) * d1, pz = (k + 0.5) * d2;
        const double base = k0x * px + k0y * py + k0z * pz;
        const double2 seg[3] = {s0, s1, s2};
        const double shift[3] = {0.5 * d0 * k0x, 0.5 * d1 * k0y, 0.5 * d2 * k0z};
        for (int c = 0; c < 3; ++c) {
            const long F = c * V + v;
            double2 r = make_double2(0, 0);
            if (flags[F]) {
                double sn, cs;
                sincos(-(base + shift[c]), &sn, &cs);
                r = zmul(make_double2(cs, sn), seg[c]);
            }
            b[F] = r;
        }
    }
}

// ---------------------------------------------------- absorbed power, gradient
// Face (f, v) is active when flags_f[v] != 0 and v_f < N_f - 1; its absorbed
// power is w |J|^2 with w = K Im(S), K = dV Z0 / (2 k0),
// S = 1/(eps_v - 1) + 1/(eps_{v+e_f} - 1): NearField's per-face Ene_Metal term
// (vie_fft/NearFar.h:192-205) with Im(-eps_eff)|S|^2 = 2 Im(S) folded in.
__device__ __forceinline__ double2 zinv1(double2 e)  // 1 / (e - 1)
{
    return zdiv(make_double2(1, 0), make_double2(e.x - 1.0, e.y));
}

// P partial sums per block; rhs (optional) = delta_f w conj(J) on active
// faces, 0 elsewhere: the right-hand side of the adjoint solve.
__global__ __launch_bounds__(NT) void k_abs_rhs(int n0, int n1, int n2, double K, double d0, double d1, double d2,
                                                const double2* __restrict__ J, const double2* __restrict__ eps,
                                                const uint8_t* __restrict__ flags, double2* __restrict__ rhs,
                                                double2* partial)
{
    __shared__ double2 red[NT / 64];
    const long V = (long)n0 * n1 * n2;
    const long st[3] = {(long)n1 * n2, n2, 1};
    const int nn[3] = {n0, n1, n2};
    const double dd[3] = {d0, d1, d2};
    double acc = 0;
    for (long v = (long)blockIdx.x * NT + threadIdx.x; v < V; v += (long)gridDim.x * NT) {
        const int c[3] = {(int)(v / ((long)n1 * n2)), (int)((v / n2) % n1), (int)(v % n2)};
        const double2 iv = zinv1(eps[v]);
        for (int f = 0; f < 3; ++f) {
            const long F = f * V + v;
            double2 r = make_double2(0, 0);
            if (c[f] < nn[f] - 1 && flags[F]) {
                const double2 iw = zinv1(eps[v + st[f]]);
                const double w = K * (iv.y + iw.y);
                const double2 j = J[F];
                acc += w * (j.x * j.x + j.y * j.y);
                r = make_double2(dd[f] * w * j.x, -dd[f] * w * j.y);
            }
            if (rhs) rhs[F] = r;
        }
    }
    double2 s = block_sum(make_double2(acc, 0), red);
    if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

// dP/dRe(eps_u), dP/dIm(eps_u) for every cell u, gathering the (up to six)
// active faces that touch u.  nu solves A nu = rhs (k_abs_rhs); the adjoint
// field is mu_f = nu_f / delta_f.  With dS = -1/(eps_u - 1)^2 and
// c_f = -j (delta_f/2)(Z0/k0) (the diagonal's factor, vie_fft/fft.h:419-434):
//   d/dRe: K Im(dS) |J|^2 - 2 Re(mu c_f dS J);  d/dIm: same with dS -> j dS.
__global__ __launch_bounds__(NT) void k_abs_grad(int n0, int n1, int n2, double K, double z0k0, double d0,
                                                 double d1, double d2, const double2* __restrict__ J,
                                                 const double2* __restrict__ nu, const double2* __restrict__ eps,
                                                 const uint8_t* __restrict__ flags, double* __restrict__ gre,
                                                 double* __restrict__ gim)
{
    const long V = (long)n0 * n1 * n2;
    const long st[3] = {(long)n1 * n2, n2, 1};
    const int nn[3] = {n0, n1, n2};
    const double dd[3] = {d0, d1, d2};
    for (long u = (long)blockIdx.x * NT + threadIdx.x; u < V; u += (long)gridDim.x * NT) {
        const int c[3] = {(int)(u / ((long)n1 * n2)), (int)((u / n2) % n1), (int)(u % n2)};
        const double2 iu = zinv1(eps[u]);
        const double2 dS = zneg(zmul(iu, iu));
        double ar = 0, ai = 0;
        for (int f = 0; f < 3; ++f) {
            const double2 cf = make_double2(0, -0.5 * dd[f] * z0k0);
            for (int side = 0; side < 2; ++side) {
                // side 0: face (f, u) (u is its lower cell); side 1: face (f, u - e_f)
                const int cf_idx = c[f] - side;
                if (cf_idx < 0 || cf_idx >= nn[f] - 1) continue;
                const long F = f * V + u - side * st[f];
                if (!flags[F]) continue;
                const double2 j = J[F];
                const double2 m = nu[F];
                const double2 mu = make_double2(m.x / dd[f], m.y / dd[f]);
                const double aJ2 = j.x * j.x + j.y * j.y;
                const double2 t = zmul(zmul(mu, cf), j);  // mu c_f J
                const double2 tr = zmul(t, dS);           // times dS
                ar += K * dS.y * aJ2 - 2.0 * tr.x;
                // j dS: Im(j dS) = Re(dS); Re(t j dS) = -Im(t dS)
                ai += K * dS.x * aJ2 + 2.0 * tr.y;
            }
        }
        gre[u] = ar;
        if (gim) gim[u] = ai;
    }
}

}  // namespace vie
