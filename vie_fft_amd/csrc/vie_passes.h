// vie_passes.h -- the five pass kernels of one VIE operator product.
//
// Reference correspondence (paths relative to the reference root):
//   Current_Expansion       vie_fft/CG_Iteration.h:50-83   zfwd load (own-axis
//                                                          last plane dropped)
//   L_DC_Dq / L_Dq symbols  vie_fft/fft.h:76-100, :256-281 xmix product
//   *_Conjugate symbols     vie_fft/fft.h:166-191, :344-369 xmix, adjoint tables
//   truncation + sum        vie_fft/fft.h:113-123           yinv / zinv
//   LXX_DO / LYY_DO / LZZ_DO vie_fft/fft.h:405-588          zinv diagonal term
//   Current_Reset           vie_fft/fft.h:13-35             zinv output mask
#pragma once

#include "vie_fft_device.h"

namespace vie {

// [x0, x1) is the x window the passes cover (0 <= x0 < x1 <= nx, set by the
// owner of the Geo): the x planes outside it carry no active face, so a
// product whose input vanishes there needs neither its z/y transforms on
// those planes nor its x-pencil entries there.  [y0, y1) (0 <= y0 < y1 <= ny)
// likewise bounds the z rows (x, y) the z passes transform and the y entries
// the y passes load and store.  The x pass's pencil entry n is plane n + xs
// (xs = x0 when the pencils have the short period).  W2 holds nyp y pencils
// of length 2 nyp per (x, kz): nyp = ny, or the short y period, whose entry
// n is y plane n + ys (ys = y0).  W1 and W2 hold 2 nzp z frequencies: nzp =
// nz, or the short z period, whose entry n is z plane n + zs (nz and zs
// even, so pairs of entries stay 16-byte aligned and never straddle nz).
//
// Batches: a launch may serve nb instances of one geometry (right-hand sides,
// wavelengths, polarisations).  Launch slot b (workgroup row b; rows 3b..3b+2
// for the per-component passes) serves instance ilist[b] (b without a list).
// Field vectors hold the instances back to back (3V entries each), the pass
// buffers W1/W2 hold the launch slots back to back, and every instance has its
// own solver state record: its done flag is dstr ints and its coefficient
// block cstr complex values after the previous instance's.
struct Bat {
    int nb = 1;
    const int* ilist = nullptr;
    int dstr = 0, cstr = 0;
};
__device__ __forceinline__ int bat_inst(const Bat& bt, int b) { return bt.ilist ? bt.ilist[b] : b; }
__device__ __forceinline__ const int* bat_done(const Bat& bt, const int* done, int inst)
{
    return done ? done + (long)inst * bt.dstr : nullptr;
}

struct Geo {
    int nx, ny, nz;
    long V;
    int x0 = 0, x1 = 0;
    int y0 = 0, y1 = 0;
    int xs = 0;
    int nyp = 0, ys = 0;
    int nzp = 0, zs = 0;
    Bat bt{};
    // pass variants (a tuned plan): bit 0 register z forward, bit 1 LDS-staged
    // y and z passes, bit 2 LDS-staged x pass; -1: the fixed defaults
    int var = -1;
};

// z row of window row r, r < (x1 - x0)(y1 - y0): x0 + r / wy, y0 + r % wy
__device__ __forceinline__ long win_row(const Geo& g, int r)
{
    const int wy = g.y1 - g.y0;
    return (long)(g.x0 + r / wy) * g.ny + g.y0 + r % wy;
}
__host__ __device__ __forceinline__ int win_rows(const Geo& g) { return (g.x1 - g.x0) * (g.y1 - g.y0); }

// Tables of the spectral product, per axis a in split order t = e*N_a + m
// (frequency k = 2m + e, theta = 2 pi k / (2 N_a)):
//   U[a][t] = exp(j theta) - 1,  Vv[a][t] = (j Z0/(k0 delta_a NP)) (1 - exp(-j theta))
// (adjoint: U' = conj(Vv), Vv' = conj(U)), D[f] = j k0 Z0 delta_f / NP (adjoint
// conj), fold[a][t] = position of min(k, P - k) in the parity order of the
// Green's-function octant G (evens 0,2,..,N then odds 1,3,..).
// Batches with one operator per instance (several wavelengths): instance i
// reads G + i gstr, U[a] + i ustr, Vv[a] + i vstr and, with Dd set, D_f =
// Dd[3 i + f] (strides 0: one operator for every instance).
template <class R> struct MixTabs {
    const cx<R>* U[3];
    const cx<R>* Vv[3];
    const int* fold[3];
    const cx<R>* G;
    R D[3][2];
    int gconj;
    long gstr = 0;
    int ustr = 0, vstr = 0;
    const cx<R>* Dd = nullptr;
};

// The instance's tables of a batched product.
template <class R> struct InstTabs {
    const cx<R>* U[3];
    const cx<R>* Vv[3];
    const cx<R>* G;
    cx<R> D[3];
};
template <class R> __device__ __forceinline__ InstTabs<R> inst_tabs(const MixTabs<R>& mt, int inst)
{
    InstTabs<R> t;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        t.U[a] = mt.U[a] + (long)inst * mt.ustr;
        t.Vv[a] = mt.Vv[a] + (long)inst * mt.vstr;
        t.D[a] = mt.Dd ? mt.Dd[3L * inst + a] : mkc(mt.D[a][0], mt.D[a][1]);
    }
    t.G = mt.G + inst * mt.gstr;
    return t;
}

template <class T> __device__ __forceinline__ T czero() { return mkc(decltype(T().x)(0), decltype(T().x)(0)); }

// Diagonal face term d_f = c_f (chi[i + e_f] + chi[i]) for i_f < N_f - 1
// (vie_fft/fft.h:405-588), chi = 1/(eps - 1), c_f = -j (delta_f/2)(Z0/k0),
// precomputed on the host into a 3V array (component-major like the field).
template <class R> struct Diag {
    const cx<R>* d;       // per-face coefficients (3V)
    int conj;             // adjoint: use conj(d_f)
    int skip;             // y already holds zero on inactive faces: leave them unwritten
    long dstr = 0;        // batches: instance i reads d + i dstr (0: shared)
    int nomask = 0;       // every face counts as active: no output mask (Current_Reset left out)
    int unit = 0;         // the diagonal term is x itself (d (.) m = 1: the Jacobi-preconditioned
                          // product, whose input the z forward pass scaled by m = 1/d); d is not read
};

// Block-wide sum of a double2; sh holds NT/64 entries.  Result valid in thread 0.
__device__ __forceinline__ double2 block_sum(double2 v, double2* sh)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        v.x += __shfl_xor(v.x, off);
        v.y += __shfl_xor(v.y, off);
    }
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) sh[wid] = v;
    __syncthreads();
    double2 s = make_double2(0.0, 0.0);
    if (threadIdx.x == 0)
        for (int w = 0; w < NT / 64; ++w) {
            s.x += sh[w].x;
            s.y += sh[w].y;
        }
    return s;
}

// acc += conj(a) b, or a b when nc is set
template <class T> __device__ __forceinline__ void dot_acc(double2& acc, T a, T b, int nc)
{
    const double ay = nc ? -(double)a.y : (double)a.y;
    acc.x += (double)a.x * (double)b.x + ay * (double)b.y;
    acc.y += (double)a.x * (double)b.y - ay * (double)b.x;
}

// Dot-product operands: selector 0 is the value being produced, 1 is e1,
// 2 is e2.  Operands are picked by value: a conditional over lvalues selects
// an address and pushes the candidates to scratch memory.
template <class R> __device__ __forceinline__ bool dots_use(const Dots<R>& d, int sel)
{
    bool u = false;
#pragma unroll
    for (int k = 0; k < 3; ++k)
        if (k < d.n && (d.a[k] == sel || d.b[k] == sel)) u = true;
    return u;
}
template <class T> __device__ __forceinline__ T dsel(int sel, T v, T e1, T e2)
{
    T r = v;
    if (sel == 1) r = e1;
    if (sel == 2) r = e2;
    return r;
}
template <class R>
__device__ __forceinline__ void dots_add(double2 (&acc)[3], const Dots<R>& d, cx<R> v, cx<R> e1, cx<R> e2)
{
    if (d.n > 0) dot_acc(acc[0], dsel(d.a[0], v, e1, e2), dsel(d.b[0], v, e1, e2), d.nc[0]);
    if (d.n > 1) dot_acc(acc[1], dsel(d.a[1], v, e1, e2), dsel(d.b[1], v, e1, e2), d.nc[1]);
    if (d.n > 2) dot_acc(acc[2], dsel(d.a[2], v, e1, e2), dsel(d.b[2], v, e1, e2), d.nc[2]);
}

template <class R>
__device__ __forceinline__ void load_coef(const double2* coef, cx<R>* c)
{
    if (coef) {
        for (int q = 0; q < 3; ++q) c[q] = mkc((R)coef[q].x, (R)coef[q].y);
    } else {
        c[0] = mkc((R)1, (R)0);
        c[1] = c[2] = mkc((R)0, (R)0);
    }
}

// Two consecutive complex values per lane: one 16-byte access for FP32,
// two for FP64 (already 16 bytes each).
__device__ __forceinline__ void ld2(const float2* p, float2& a, float2& b)
{
    const float4 v = *reinterpret_cast<const float4*>(p);
    a = make_float2(v.x, v.y);
    b = make_float2(v.z, v.w);
}
__device__ __forceinline__ void ld2(const double2* p, double2& a, double2& b)
{
    a = p[0];
    b = p[1];
}
__device__ __forceinline__ void st2(float2* p, float2 a, float2 b)
{
    *reinterpret_cast<float4*>(p) = make_float4(a.x, a.y, b.x, b.y);
}
__device__ __forceinline__ void st2(double2* p, double2 a, double2 b)
{
    p[0] = a;
    p[1] = b;
}

// ------------------------------------------------------------------ zfwd
// grid (ceil(Nx*Ny / B), 3 components); N == Nz.
template <class R, int N, int B, bool ZS = false>
__global__ __launch_bounds__(NT) void k_zfwd(Geo g, Lin<R> lin, cx<R>* __restrict__ W1,
                                             const cx<R>* __restrict__ twg,
                                             const cx<R>* __restrict__ twistg, const int* done)
{
    using T = cx<R>;
    constexpr int COLS = 2 * B, LD = COLS + 1;
    __shared__ T buf[N * LD];
    __shared__ T tw[N];
    __shared__ T tws[N];
    // component s of instance inst (field entries of component sb)
    const int s = blockIdx.y % 3, inst = bat_inst(g.bt, blockIdx.y / 3);
    const long sb = 3L * inst + s;
    done = bat_done(g.bt, done, inst);
    if (done && *done) return;
    // z row length and the z plane of entry 0: without ZS the pencil is the
    // whole row (N = nz, entry n is plane n), fixed at compile time
    const int nzr = ZS ? g.nz : N, zs = ZS ? g.zs : 0;
    const long rows = (long)g.nx * g.ny;
    const int nw = win_rows(g);  // the window's rows (column col: window row r0 + col)
    const int r0 = (int)xcd_tile(blockIdx.x, gridDim.x) * B;
    for (int t = threadIdx.x; t < N; t += NT) {
        tw[t] = twg[t];
        tws[t] = twistg[t];
    }
    __shared__ unsigned zrs[B];  // active range of each row (the whole row without lin.zr)
    __shared__ long rw[B];       // z row of each column
    for (int t = threadIdx.x; t < B; t += NT) {
        const bool ok = r0 + t < nw;
        rw[t] = ok ? win_row(g, r0 + t) : 0;
        zrs[t] = lin.zr && ok ? lin.zr[s * rows + rw[t]] : ((unsigned)nzr << 16);
    }
    T c[3];
    load_coef<R>(lin.coef ? lin.coef + (long)inst * g.bt.cstr : nullptr, c);
    __syncthreads();
    if constexpr (N % 2 == 0) {
        // all loads first, then the combination, stores and LDS writes
        constexpr int IT = (B * N / 2 + NT - 1) / NT;
        T x0[IT][2], x1[IT][2], x2[IT][2], sc[IT][2];
#pragma unroll
        for (int it = 0; it < IT; ++it) {
            const int e = threadIdx.x + it * NT;
            const int col = e / (N / 2), n = 2 * (e % (N / 2));
            const bool rok = e < B * N / 2 && r0 + col < nw;
            const long row = rok ? rw[col] : 0;
            x0[it][0] = x0[it][1] = x1[it][0] = x1[it][1] = x2[it][0] = x2[it][1] = czero<T>();
            sc[it][0] = sc[it][1] = czero<T>();
            // with row ranges, pairs that miss the row's active range are zero
            const unsigned rg = e < B * N / 2 ? zrs[col] : 0u;
            const int z = n + zs;  // z plane of entry n
            if (rok && z + 1 >= (int)(rg & 0xffffu) && z < (int)(rg >> 16)) {
                const long idx = sb * g.V + row * nzr + z;
                ld2(lin.x[0] + idx, x0[it][0], x0[it][1]);
                if (lin.x[1]) ld2(lin.x[1] + idx, x1[it][0], x1[it][1]);
                if (lin.x[2]) ld2(lin.x[2] + idx, x2[it][0], x2[it][1]);
                if (lin.scale) ld2(lin.scale + idx, sc[it][0], sc[it][1]);
            }
        }
#pragma unroll
        for (int it = 0; it < IT; ++it) {
            const int e = threadIdx.x + it * NT;
            if (e >= B * N / 2) break;
            const int col = e / (N / 2), n = 2 * (e % (N / 2));
            const long row = rw[col];
            T v0 = czero<T>(), v1 = czero<T>();
            const int z = n + zs;
            if (r0 + col < nw) {
                const long idx = sb * g.V + row * nzr + z;
                v0 = cmul(c[0], x0[it][0]);
                v1 = cmul(c[0], x0[it][1]);
                if (lin.x[1]) {
                    v0 = cadd(v0, cmul(c[1], x1[it][0]));
                    v1 = cadd(v1, cmul(c[1], x1[it][1]));
                }
                if (lin.x[2]) {
                    v0 = cadd(v0, cmul(c[2], x2[it][0]));
                    v1 = cadd(v1, cmul(c[2], x2[it][1]));
                }
                // entries past Nz (short periods only) are not on the grid
                if (lin.out && (!ZS || z < nzr) &&
                    (!lin.wskip || (z + 1 >= (int)(zrs[col] & 0xffffu) && z < (int)(zrs[col] >> 16))))
                    st2(lin.out + idx, v0, v1);
                if (lin.scale) {
                    v0 = cmul(v0, sc[it][0]);
                    v1 = cmul(v1, sc[it][1]);
                }
                const bool dropall = (s == 0 && row >= rows - g.ny) || (s == 1 && (row % g.ny) == g.ny - 1);
                if (dropall) v0 = v1 = czero<T>();
                if (s == 2 && z + 1 == nzr - 1) v1 = czero<T>();
            }
            buf[n * LD + col] = v0;
            buf[n * LD + B + col] = cmul(v0, tws[n]);
            buf[(n + 1) * LD + col] = v1;
            buf[(n + 1) * LD + B + col] = cmul(v1, tws[n + 1]);
        }
    } else {
        for (int e = threadIdx.x; e < B * N; e += NT) {
            const int col = e / N, n = e % N;
            const long row = rw[col];
            T v = czero<T>();
            const int z = n + zs;
            if (r0 + col < nw) {
                const long idx = sb * g.V + row * nzr + z;
                const unsigned rg = zrs[col];
                if (z >= (int)(rg & 0xffffu) && z < (int)(rg >> 16)) {
                    v = cmul(c[0], lin.x[0][idx]);
                    if (lin.x[1]) v = cadd(v, cmul(c[1], lin.x[1][idx]));
                    if (lin.x[2]) v = cadd(v, cmul(c[2], lin.x[2][idx]));
                }
                if (lin.out && (!ZS || z < nzr) &&
                    (!lin.wskip || (z >= (int)(zrs[col] & 0xffffu) && z < (int)(zrs[col] >> 16))))
                    lin.out[idx] = v;
                if (lin.scale && z >= (int)(zrs[col] & 0xffffu) && z < (int)(zrs[col] >> 16))
                    v = cmul(v, lin.scale[idx]);
                const bool drop = (s == 2 && z == nzr - 1) || (s == 0 && row >= rows - g.ny) ||
                                  (s == 1 && (row % g.ny) == g.ny - 1);
                if (drop) v = czero<T>();
            }
            buf[n * LD + col] = v;
            buf[n * LD + B + col] = cmul(v, tws[n]);
        }
    }
    __syncthreads();
    fft_lds<T, N, COLS, LD, -1>(buf, tw);
    // t and t+1 share the half e when N is even (pairs never straddle N)
    for (int e = threadIdx.x; e < B * N; e += NT) {
        const int col = e / N, t = 2 * (e % N);
        const long row = rw[col];
        if (r0 + col < nw) {
            const int ep = t >= N, m = t - ep * N;
            T* dst = W1 + ((long)blockIdx.y * rows + row) * (2 * N) + t;
            if constexpr (N % 2 == 0)
                st2(dst, buf[m * LD + ep * B + col], buf[(m + 1) * LD + ep * B + col]);
            else {
                // odd N: t+1 may fall in the other half
                dst[0] = buf[m * LD + ep * B + col];
                const int t1 = t + 1, ep1 = t1 >= N, m1 = t1 - ep1 * N;
                dst[1] = buf[m1 * LD + ep1 * B + col];
            }
        }
    }
}

// ------------------------------------------------------------------ yfwd
// grid (Nx * ceil(2Nz / B), 3); N == Ny.
template <class R, int N, int B>
__global__ __launch_bounds__(NT) void k_yfwd(Geo g, const cx<R>* __restrict__ W1,
                                             cx<R>* __restrict__ W2,
                                             const cx<R>* __restrict__ twg,
                                             const cx<R>* __restrict__ twistg, const int* done)
{
    using T = cx<R>;
    constexpr int COLS = 2 * B, LD = COLS;
    static_assert(B % 2 == 0, "column pairs");
    __shared__ T buf[N * LD];
    __shared__ T tw[N];
    __shared__ T tws[N];
    done = bat_done(g.bt, done, bat_inst(g.bt, blockIdx.y / 3));
    if (done && *done) return;
    const int s = blockIdx.y;  // launch slot (component of one instance) of W1 and W2
    const int P2 = 2 * g.nzp;
    const int nzt = (P2 + B - 1) / B;
    const unsigned tile = xcd_tile(blockIdx.x, gridDim.x);
    const int i = g.x0 + (int)(tile / nzt), z0 = (tile % nzt) * B;
    for (int t = threadIdx.x; t < N; t += NT) {
        tw[t] = twg[t];
        tws[t] = twistg[t];
    }
    __syncthreads();
    const long in0 = ((long)s * g.nx + i) * g.ny * P2 + z0;
    // column pairs (col, col+1): P2 and z0 are even, so a pair is in or out together
    for (int e = threadIdx.x; e < N * B / 2; e += NT) {
        const int j = e / (B / 2), col = 2 * (e % (B / 2)), pl = j + g.ys;  // entry j: y plane pl
        T v0 = czero<T>(), v1 = czero<T>();
        if (z0 + col < P2 && pl >= g.y0 && pl < g.y1) ld2(W1 + in0 + (long)pl * P2 + col, v0, v1);
        buf[j * LD + col] = v0;
        buf[j * LD + col + 1] = v1;
        buf[j * LD + B + col] = cmul(v0, tws[j]);
        buf[j * LD + B + col + 1] = cmul(v1, tws[j]);
    }
    __syncthreads();
    fft_lds<T, N, COLS, LD, -1>(buf, tw);
    const long out0 = ((long)s * g.nx + i) * (2 * N) * P2 + z0;
    for (int e = threadIdx.x; e < N * B; e += NT) {
        const int t = e / (B / 2), col = 2 * (e % (B / 2));
        if (z0 + col < P2) {
            const int ep = t >= N, m = t - ep * N;
            st2(W2 + out0 + (long)t * P2 + col, buf[m * LD + ep * B + col], buf[m * LD + ep * B + col + 1]);
        }
    }
}

// ------------------------------------------------------------------ xmix
// grid (2Ny * ceil(2Nz / B)); N == Nx.  In place on W2.
template <class R, int N, int B>
__global__ __launch_bounds__(NT) void k_xmix(Geo g, cx<R>* __restrict__ W2,
                                             const cx<R>* __restrict__ twg,
                                             const cx<R>* __restrict__ twistg, MixTabs<R> mt,
                                             const int* done)
{
    using T = cx<R>;
    constexpr int COLS = 6 * B, LD = COLS;
    __shared__ T buf[N * LD];
    __shared__ T tw[N];
    __shared__ T tws[N];
    __shared__ T ux[2 * N];
    __shared__ T vx[2 * N];
    __shared__ int fx[2 * N];
    __shared__ T uz[B];
    __shared__ T vz[B];
    __shared__ int fz[B];
    const int inst = bat_inst(g.bt, blockIdx.y);  // launch slot blockIdx.y of W2
    done = bat_done(g.bt, done, inst);
    if (done && *done) return;
    const InstTabs<R> it = inst_tabs(mt, inst);
    const int P1 = 2 * g.nyp, P2 = 2 * g.nzp;
    const int nzt = (P2 + B - 1) / B;
    const unsigned tile = xcd_tile(blockIdx.x, gridDim.x);
    const int ty = tile / nzt, z0 = (tile % nzt) * B;
    W2 += (long)blockIdx.y * 3 * g.nx * P1 * P2;
    for (int t = threadIdx.x; t < N; t += NT) {
        tw[t] = twg[t];
        tws[t] = twistg[t];
    }
    for (int t = threadIdx.x; t < 2 * N; t += NT) {
        ux[t] = it.U[0][t];
        vx[t] = it.Vv[0][t];
        fx[t] = mt.fold[0][t];
    }
    for (int t = threadIdx.x; t < B; t += NT) {
        const bool ok = z0 + t < P2;
        uz[t] = ok ? it.U[2][z0 + t] : czero<T>();
        vz[t] = ok ? it.Vv[2][z0 + t] : czero<T>();
        fz[t] = ok ? mt.fold[2][z0 + t] : 0;
    }
    const T uy = it.U[1][ty], vy = it.Vv[1][ty];
    const long fyoff = (long)mt.fold[1][ty] * (g.nzp + 1);
    const long gs = (long)(g.nyp + 1) * (g.nzp + 1);
    __syncthreads();
    // x-pencils of the three components: rows i < Nx, columns z0..z0+B
    for (int e = threadIdx.x; e < 3 * N * B; e += NT) {
        const int s = e / (N * B), r = e % (N * B);
        const int i = r / B, col = r % B, pl = i + g.xs;
        const bool in = z0 + col < P2 && pl >= g.x0 && pl < g.x1;
        T v = in ? W2[(((long)s * g.nx + pl) * P1 + ty) * P2 + z0 + col] : czero<T>();
        buf[i * LD + (2 * s) * B + col] = v;
        buf[i * LD + (2 * s + 1) * B + col] = cmul(v, tws[i]);
    }
    __syncthreads();
    fft_lds<T, N, COLS, LD, -1>(buf, tw);
    const T D0 = it.D[0], D1 = it.D[1], D2 = it.D[2];
    // Y_f = G (D_f X_f + U_f sum_s Vv_s X_s)
    for (int e = threadIdx.x; e < 2 * N * B; e += NT) {
        const int ep = e / (N * B), r = e % (N * B);
        const int m = r / B, col = r % B;
        const int tx = ep * N + m;
        T* p0 = &buf[m * LD + ep * B + col];
        const T X0 = p0[0], X1 = p0[2 * B], X2 = p0[4 * B];
        T Gv = it.G[(long)fx[tx] * gs + fyoff + fz[col]];
        if (mt.gconj) Gv = cconj(Gv);
        const T sig = cadd(cadd(cmul(vx[tx], X0), cmul(vy, X1)), cmul(vz[col], X2));
        p0[0] = cmul(Gv, cadd(cmul(D0, X0), cmul(ux[tx], sig)));
        p0[2 * B] = cmul(Gv, cadd(cmul(D1, X1), cmul(uy, sig)));
        p0[4 * B] = cmul(Gv, cadd(cmul(D2, X2), cmul(uz[col], sig)));
    }
    __syncthreads();
    fft_lds<T, N, COLS, LD, +1>(buf, tw);
    for (int e = threadIdx.x; e < 3 * N * B; e += NT) {
        const int s = e / (N * B), r = e % (N * B);
        const int i = r / B, col = r % B, pl = i + g.xs;
        if (z0 + col < P2 && pl >= g.x0 && pl < g.x1) {
            const T a = buf[i * LD + (2 * s) * B + col];
            const T b = buf[i * LD + (2 * s + 1) * B + col];
            W2[(((long)s * g.nx + pl) * P1 + ty) * P2 + z0 + col] = cadd(a, cmulc(b, tws[i]));
        }
    }
}

// ------------------------------------------------------------------ yinv
// grid (Nx * ceil(2Nz / B), 3); N == Ny.
template <class R, int N, int B>
__global__ __launch_bounds__(NT) void k_yinv(Geo g, const cx<R>* __restrict__ W2,
                                             cx<R>* __restrict__ W1,
                                             const cx<R>* __restrict__ twg,
                                             const cx<R>* __restrict__ twistg, const int* done)
{
    using T = cx<R>;
    constexpr int COLS = 2 * B, LD = COLS;
    static_assert(B % 2 == 0, "column pairs");
    __shared__ T buf[N * LD];
    __shared__ T tw[N];
    __shared__ T tws[N];
    done = bat_done(g.bt, done, bat_inst(g.bt, blockIdx.y / 3));
    if (done && *done) return;
    const int s = blockIdx.y;  // launch slot (component of one instance) of W1 and W2
    const int P2 = 2 * g.nzp;
    const int nzt = (P2 + B - 1) / B;
    const unsigned tile = xcd_tile(blockIdx.x, gridDim.x);
    const int i = g.x0 + (int)(tile / nzt), z0 = (tile % nzt) * B;
    for (int t = threadIdx.x; t < N; t += NT) {
        tw[t] = twg[t];
        tws[t] = twistg[t];
    }
    const long in0 = ((long)s * g.nx + i) * (2 * N) * P2 + z0;
    for (int e = threadIdx.x; e < N * B; e += NT) {
        const int t = e / (B / 2), col = 2 * (e % (B / 2));
        const int ep = t >= N, m = t - ep * N;
        T v0 = czero<T>(), v1 = czero<T>();
        if (z0 + col < P2) ld2(W2 + in0 + (long)t * P2 + col, v0, v1);
        buf[m * LD + ep * B + col] = v0;
        buf[m * LD + ep * B + col + 1] = v1;
    }
    __syncthreads();
    fft_lds<T, N, COLS, LD, +1>(buf, tw);
    const long out0 = ((long)s * g.nx + i) * g.ny * P2 + z0;
    for (int e = threadIdx.x; e < N * B / 2; e += NT) {
        const int j = e / (B / 2), col = 2 * (e % (B / 2)), pl = j + g.ys;  // entry j: y plane pl
        if (z0 + col < P2 && pl >= g.y0 && pl < g.y1)
            st2(W1 + out0 + (long)pl * P2 + col, cadd(buf[j * LD + col], cmulc(buf[j * LD + B + col], tws[j])),
                cadd(buf[j * LD + col + 1], cmulc(buf[j * LD + B + col + 1], tws[j])));
    }
}

// ------------------------------------------------------------------ zinv
// grid (ceil(Nx*Ny / B), 3); N == Nz.  y = flag (trunc(IDFT) + d x), dots.
template <class R, int N, int B, bool ZS = false>
__global__ __launch_bounds__(NT) void k_zinv(Geo g, const cx<R>* __restrict__ W1,
                                             const cx<R>* __restrict__ src, cx<R>* __restrict__ y,
                                             Diag<R> dg, const uint8_t* __restrict__ flags, Dots<R> dots,
                                             const cx<R>* __restrict__ twg,
                                             const cx<R>* __restrict__ twistg, const int* done)
{
    using T = cx<R>;
    constexpr int COLS = 2 * B, LD = COLS + 1;
    __shared__ T buf[N * LD];
    __shared__ T tw[N];
    __shared__ T tws[N];
    __shared__ double2 red[NT / 64];
    // component f of instance inst (field entries of component fb)
    const int f = blockIdx.y % 3, inst = bat_inst(g.bt, blockIdx.y / 3);
    const long fb = 3L * inst + f;
    done = bat_done(g.bt, done, inst);
    if (done && *done) return;
    const cx<R>* dgd = dg.d + inst * dg.dstr;
    // z row length and the z plane of entry 0: without ZS the pencil is the
    // whole row (N = nz, entry n is plane n), fixed at compile time
    const int nzr = ZS ? g.nz : N, zs = ZS ? g.zs : 0;
    const long rows = (long)g.nx * g.ny;
    const int nw = win_rows(g);  // the window's rows (column col: window row r0 + col)
    const int r0 = (int)xcd_tile(blockIdx.x, gridDim.x) * B;
    __shared__ long rw[B];       // z row of each column
    for (int t = threadIdx.x; t < B; t += NT) rw[t] = r0 + t < nw ? win_row(g, r0 + t) : 0;
    for (int t = threadIdx.x; t < N; t += NT) {
        tw[t] = twg[t];
        tws[t] = twistg[t];
    }
    __syncthreads();
    for (int e = threadIdx.x; e < B * N; e += NT) {
        const int col = e / N, t = 2 * (e % N);
        const long row = rw[col];
        T v0 = czero<T>(), v1 = czero<T>();
        if (r0 + col < nw) ld2(W1 + ((long)blockIdx.y * rows + row) * (2 * N) + t, v0, v1);
        const int ep = t >= N, m = t - ep * N;
        buf[m * LD + ep * B + col] = v0;
        const int t1 = t + 1, ep1 = t1 >= N, m1 = t1 - ep1 * N;
        buf[m1 * LD + ep1 * B + col] = v1;
    }
    __syncthreads();
    fft_lds<T, N, COLS, LD, +1>(buf, tw);
    double2 acc[3] = {make_double2(0, 0), make_double2(0, 0), make_double2(0, 0)};
    const bool need1 = dots_use(dots, 1), need2 = dots_use(dots, 2);
    // ia: entry of the instance's field vectors; idx - ia is the same entry
    // of the shared flags (and of a shared diagonal)
    const long ioff = (fb - f) * g.V;
    auto term = [&](long ia, T out) {  // y = flag (out + d x), then its dot products
        const long idx = ia - ioff;
        const bool on = dg.nomask || flags[idx] != 0;
        T yv = czero<T>();
        if (on) {  // d and x are read only where the face is active
            if (dg.unit) {
                yv = cadd(out, src[ia]);
            } else {
                const T dv = dg.conj ? cconj(dgd[idx]) : dgd[idx];
                yv = cadd(out, cmul(dv, src[ia]));
            }
        }
        if (on || !dg.skip) y[ia] = yv;
        const bool rd = on || !dots.masked;
        dots_add(acc, dots, yv, need1 && rd ? dots.e1[ia] : czero<T>(), need2 && rd ? dots.e2[ia] : czero<T>());
    };
    if constexpr (N % 2 == 0) {
        for (int e = threadIdx.x; e < B * N / 2; e += NT) {
            const int col = e / (N / 2), k = 2 * (e % (N / 2));
            const long row = rw[col];
            if (r0 + col < nw && k + zs < nzr) {  // entry k: z plane k + zs (pairs in or out together)
                const long idx = fb * g.V + row * nzr + zs + k;
                term(idx, cadd(buf[k * LD + col], cmulc(buf[k * LD + B + col], tws[k])));
                term(idx + 1, cadd(buf[(k + 1) * LD + col], cmulc(buf[(k + 1) * LD + B + col], tws[k + 1])));
            }
        }
    } else {
        for (int e = threadIdx.x; e < B * N; e += NT) {
            const int col = e / N, k = e % N;
            const long row = rw[col];
            if (r0 + col < nw && k + zs < nzr)
                term(fb * g.V + row * nzr + zs + k, cadd(buf[k * LD + col], cmulc(buf[k * LD + B + col], tws[k])));
        }
    }
    if (dots.n > 0) {
        // per launch slot: 3 arrays of the slot's 3 gridDim.x partial sums
        const long nblk = 3L * gridDim.x;
        const long blk = (long)f * gridDim.x + blockIdx.x;
        double2* part = dots.partial + (long)(blockIdx.y / 3) * 3 * nblk;
        #pragma unroll
        for (int d = 0; d < 3; ++d) {
            if (d >= dots.n) break;
            double2 s = block_sum(acc[d], red);
            if (threadIdx.x == 0) part[d * nblk + blk] = s;
        }
    }
}

// LDS-budgeted column count per pass: the largest power of two <= bmax whose
// buffer fits the budget.
template <class R> constexpr int pick_b(int n, int cols_per_b, int bmax, int budget)
{
    int b = bmax;
    while (b > 1 && (long)n * cols_per_b * b * (long)sizeof(cx<R>) > budget) b /= 2;
    return b;
}
// z passes: at most 8 rows per workgroup in either precision (FP32 128^3 z
// forward: 44 us per launch against 48 us with 16 rows, from more workgroups
// per CU)
template <class R> constexpr int b_z(int n) { return pick_b<R>(n, 2, 8, 40 * 1024); }
template <class R> constexpr int b_y(int n) { return pick_b<R>(n, 2, sizeof(R) == 8 ? 8 : 16, 40 * 1024); }
template <class R> constexpr int b_x(int n) { return pick_b<R>(n, 6, sizeof(R) == 8 ? 8 : 16, 56 * 1024); }

}  // namespace vie
