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

struct Geo {
    int nx, ny, nz;
    long V;
};

// Tables of the spectral product, per axis a in split order t = e*N_a + m
// (frequency k = 2m + e, theta = 2 pi k / (2 N_a)):
//   U[a][t] = exp(j theta) - 1,  Vv[a][t] = (j Z0/(k0 delta_a NP)) (1 - exp(-j theta))
// (adjoint: U' = conj(Vv), Vv' = conj(U)), D[f] = j k0 Z0 delta_f / NP (adjoint
// conj), fold[a][t] = position of min(k, P - k) in the parity order of the
// Green's-function octant G (evens 0,2,..,N then odds 1,3,..).
template <class R> struct MixTabs {
    const cx<R>* U[3];
    const cx<R>* Vv[3];
    const int* fold[3];
    const cx<R>* G;
    R D[3][2];
    int gconj;
};

template <class T> __device__ __forceinline__ T czero() { return mkc(decltype(T().x)(0), decltype(T().x)(0)); }

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

// ------------------------------------------------------------------ zfwd
// grid (ceil(Nx*Ny / B), 3 components); N == Nz.
template <class R, int N, int B>
__global__ __launch_bounds__(NT) void k_zfwd(Geo g, Lin<R> lin, cx<R>* __restrict__ W1,
                                             const cx<R>* __restrict__ twg,
                                             const cx<R>* __restrict__ twistg, const int* done)
{
    using T = cx<R>;
    constexpr int COLS = 2 * B, LD = COLS + 1;
    __shared__ T buf[N * LD];
    __shared__ T tw[N];
    __shared__ T tws[N];
    if (done && *done) return;
    const int s = blockIdx.y;
    const long rows = (long)g.nx * g.ny;
    const long row0 = (long)xcd_tile(blockIdx.x, gridDim.x) * B;
    for (int t = threadIdx.x; t < N; t += NT) {
        tw[t] = twg[t];
        tws[t] = twistg[t];
    }
    T c[3];
    load_coef<R>(lin.coef, c);
    __syncthreads();
    for (int e = threadIdx.x; e < B * N; e += NT) {
        const int col = e / N, n = e % N;
        const long row = row0 + col;
        T v = czero<T>();
        if (row < rows) {
            const long idx = s * g.V + row * N + n;
            v = cmul(c[0], lin.x[0][idx]);
            if (lin.x[1]) v = cadd(v, cmul(c[1], lin.x[1][idx]));
            if (lin.x[2]) v = cadd(v, cmul(c[2], lin.x[2][idx]));
            if (lin.out) lin.out[idx] = v;
            const bool drop = (s == 2 && n == N - 1) || (s == 0 && row >= rows - g.ny) ||
                              (s == 1 && (row % g.ny) == g.ny - 1);
            if (drop) v = czero<T>();
        }
        buf[n * LD + col] = v;
        buf[n * LD + B + col] = cmul(v, tws[n]);
    }
    __syncthreads();
    fft_lds<T, N, COLS, LD, -1>(buf, tw);
    for (int e = threadIdx.x; e < B * 2 * N; e += NT) {
        const int col = e / (2 * N), t = e % (2 * N);
        const long row = row0 + col;
        if (row < rows) {
            const int ep = t >= N, m = t - ep * N;
            W1[(s * rows + row) * (2 * N) + t] = buf[m * LD + ep * B + col];
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
    __shared__ T buf[N * LD];
    __shared__ T tw[N];
    __shared__ T tws[N];
    if (done && *done) return;
    const int s = blockIdx.y;
    const int P2 = 2 * g.nz;
    const int nzt = (P2 + B - 1) / B;
    const unsigned tile = xcd_tile(blockIdx.x, gridDim.x);
    const int i = tile / nzt, z0 = (tile % nzt) * B;
    for (int t = threadIdx.x; t < N; t += NT) {
        tw[t] = twg[t];
        tws[t] = twistg[t];
    }
    __syncthreads();
    const long in0 = ((long)s * g.nx + i) * N * P2 + z0;
    for (int e = threadIdx.x; e < N * B; e += NT) {
        const int j = e / B, col = e % B;
        T v = (z0 + col < P2) ? W1[in0 + (long)j * P2 + col] : czero<T>();
        buf[j * LD + col] = v;
        buf[j * LD + B + col] = cmul(v, tws[j]);
    }
    __syncthreads();
    fft_lds<T, N, COLS, LD, -1>(buf, tw);
    const long out0 = ((long)s * g.nx + i) * (2 * N) * P2 + z0;
    for (int e = threadIdx.x; e < 2 * N * B; e += NT) {
        const int t = e / B, col = e % B;
        if (z0 + col < P2) {
            const int ep = t >= N, m = t - ep * N;
            W2[out0 + (long)t * P2 + col] = buf[m * LD + ep * B + col];
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
    if (done && *done) return;
    const int P1 = 2 * g.ny, P2 = 2 * g.nz;
    const int nzt = (P2 + B - 1) / B;
    const unsigned tile = xcd_tile(blockIdx.x, gridDim.x);
    const int ty = tile / nzt, z0 = (tile % nzt) * B;
    for (int t = threadIdx.x; t < N; t += NT) {
        tw[t] = twg[t];
        tws[t] = twistg[t];
    }
    for (int t = threadIdx.x; t < 2 * N; t += NT) {
        ux[t] = mt.U[0][t];
        vx[t] = mt.Vv[0][t];
        fx[t] = mt.fold[0][t];
    }
    for (int t = threadIdx.x; t < B; t += NT) {
        const bool ok = z0 + t < P2;
        uz[t] = ok ? mt.U[2][z0 + t] : czero<T>();
        vz[t] = ok ? mt.Vv[2][z0 + t] : czero<T>();
        fz[t] = ok ? mt.fold[2][z0 + t] : 0;
    }
    const T uy = mt.U[1][ty], vy = mt.Vv[1][ty];
    const long fyoff = (long)mt.fold[1][ty] * (g.nz + 1);
    const long gs = (long)(g.ny + 1) * (g.nz + 1);
    __syncthreads();
    // x-pencils of the three components: rows i < Nx, columns z0..z0+B
    for (int e = threadIdx.x; e < 3 * N * B; e += NT) {
        const int s = e / (N * B), r = e % (N * B);
        const int i = r / B, col = r % B;
        T v = (z0 + col < P2) ? W2[(((long)s * g.nx + i) * P1 + ty) * P2 + z0 + col] : czero<T>();
        buf[i * LD + (2 * s) * B + col] = v;
        buf[i * LD + (2 * s + 1) * B + col] = cmul(v, tws[i]);
    }
    __syncthreads();
    fft_lds<T, N, COLS, LD, -1>(buf, tw);
    const T D0 = mkc(mt.D[0][0], mt.D[0][1]);
    const T D1 = mkc(mt.D[1][0], mt.D[1][1]);
    const T D2 = mkc(mt.D[2][0], mt.D[2][1]);
    // Y_f = G (D_f X_f + U_f sum_s Vv_s X_s)
    for (int e = threadIdx.x; e < 2 * N * B; e += NT) {
        const int ep = e / (N * B), r = e % (N * B);
        const int m = r / B, col = r % B;
        const int tx = ep * N + m;
        T* p0 = &buf[m * LD + ep * B + col];
        const T X0 = p0[0], X1 = p0[2 * B], X2 = p0[4 * B];
        T Gv = mt.G[(long)fx[tx] * gs + fyoff + fz[col]];
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
        const int i = r / B, col = r % B;
        if (z0 + col < P2) {
            const T a = buf[i * LD + (2 * s) * B + col];
            const T b = buf[i * LD + (2 * s + 1) * B + col];
            W2[(((long)s * g.nx + i) * P1 + ty) * P2 + z0 + col] = cadd(a, cmulc(b, tws[i]));
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
    __shared__ T buf[N * LD];
    __shared__ T tw[N];
    __shared__ T tws[N];
    if (done && *done) return;
    const int s = blockIdx.y;
    const int P2 = 2 * g.nz;
    const int nzt = (P2 + B - 1) / B;
    const unsigned tile = xcd_tile(blockIdx.x, gridDim.x);
    const int i = tile / nzt, z0 = (tile % nzt) * B;
    for (int t = threadIdx.x; t < N; t += NT) {
        tw[t] = twg[t];
        tws[t] = twistg[t];
    }
    const long in0 = ((long)s * g.nx + i) * (2 * N) * P2 + z0;
    for (int e = threadIdx.x; e < 2 * N * B; e += NT) {
        const int t = e / B, col = e % B;
        const int ep = t >= N, m = t - ep * N;
        buf[m * LD + ep * B + col] = (z0 + col < P2) ? W2[in0 + (long)t * P2 + col] : czero<T>();
    }
    __syncthreads();
    fft_lds<T, N, COLS, LD, +1>(buf, tw);
    const long out0 = ((long)s * g.nx + i) * N * P2 + z0;
    for (int e = threadIdx.x; e < N * B; e += NT) {
        const int j = e / B, col = e % B;
        if (z0 + col < P2)
            W1[out0 + (long)j * P2 + col] = cadd(buf[j * LD + col], cmulc(buf[j * LD + B + col], tws[j]));
    }
}

// ------------------------------------------------------------------ zinv
// grid (ceil(Nx*Ny / B), 3); N == Nz.  y = flag (trunc(IDFT) + d x), dots.
template <class R, int N, int B>
__global__ __launch_bounds__(NT) void k_zinv(Geo g, const cx<R>* __restrict__ W1,
                                             const cx<R>* __restrict__ src, cx<R>* __restrict__ y,
                                             const cx<R>* __restrict__ dg, int dconj,
                                             const uint8_t* __restrict__ flags, Dots<R> dots,
                                             const cx<R>* __restrict__ twg,
                                             const cx<R>* __restrict__ twistg, const int* done)
{
    using T = cx<R>;
    constexpr int COLS = 2 * B, LD = COLS + 1;
    __shared__ T buf[N * LD];
    __shared__ T tw[N];
    __shared__ T tws[N];
    __shared__ double2 red[NT / 64];
    if (done && *done) return;
    const int f = blockIdx.y;
    const long rows = (long)g.nx * g.ny;
    const long row0 = (long)xcd_tile(blockIdx.x, gridDim.x) * B;
    for (int t = threadIdx.x; t < N; t += NT) {
        tw[t] = twg[t];
        tws[t] = twistg[t];
    }
    for (int e = threadIdx.x; e < B * 2 * N; e += NT) {
        const int col = e / (2 * N), t = e % (2 * N);
        const long row = row0 + col;
        const int ep = t >= N, m = t - ep * N;
        buf[m * LD + ep * B + col] = row < rows ? W1[(f * rows + row) * (2 * N) + t] : czero<T>();
    }
    __syncthreads();
    fft_lds<T, N, COLS, LD, +1>(buf, tw);
    double2 acc[3] = {make_double2(0, 0), make_double2(0, 0), make_double2(0, 0)};
    for (int e = threadIdx.x; e < B * N; e += NT) {
        const int col = e / N, k = e % N;
        const long row = row0 + col;
        if (row < rows) {
            const long idx = f * g.V + row * N + k;
            const T out = cadd(buf[k * LD + col], cmulc(buf[k * LD + B + col], tws[k]));
            T yv = czero<T>();
            if (flags[idx]) {
                const T dv = dconj ? cconj(dg[idx]) : dg[idx];
                yv = cadd(out, cmul(dv, src[idx]));
            }
            y[idx] = yv;
            for (int d = 0; d < dots.n; ++d) {
                const T a = dots.a[d] == 0 ? yv : dots.a[d] == 1 ? dots.e1[idx] : dots.e2[idx];
                const T b = dots.b[d] == 0 ? yv : dots.b[d] == 1 ? dots.e1[idx] : dots.e2[idx];
                dot_acc(acc[d], a, b, dots.nc[d]);
            }
        }
    }
    if (dots.n > 0) {
        const long nblk = (long)gridDim.x * gridDim.y;
        const long blk = (long)blockIdx.y * gridDim.x + blockIdx.x;
        for (int d = 0; d < dots.n; ++d) {
            double2 s = block_sum(acc[d], red);
            if (threadIdx.x == 0) dots.partial[d * nblk + blk] = s;
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
template <class R> constexpr int b_z(int n) { return pick_b<R>(n, 2, sizeof(R) == 8 ? 8 : 16, 40 * 1024); }
template <class R> constexpr int b_y(int n) { return pick_b<R>(n, 2, sizeof(R) == 8 ? 8 : 16, 40 * 1024); }
template <class R> constexpr int b_x(int n) { return pick_b<R>(n, 6, sizeof(R) == 8 ? 8 : 16, 56 * 1024); }

}  // namespace vie
