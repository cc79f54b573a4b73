// vie_xpass.h -- register-resident x pass: forward transform along x, the
// 3x3 spectral product with the Green's function, inverse transform and
// truncation, for B k-columns of the three components at once.
//
// Each column is one full padded transform of length P = 2N (input zero on
// [N, P), only rows n < N kept on output) spread over T threads that hold
// E = P/T values each in registers.  A Stockham stage with radix r runs E/r
// butterflies per thread (butterfly i = t + T b reads positions i + q P/r and
// writes (i - i mod S) r + i mod S + q S, S the product of earlier radices);
// stages exchange through LDS, one component at a time.  After the last
// forward stage thread t holds frequencies k = i + q P/r, applies the product
// there, and runs the inverse as the reverse dataflow of the forward (conjugate
// DFT, then conjugate twiddle, read and written at swapped positions), so every
// thread ends with the rows n = i + q P/r1 it loaded.  This replaces the
// L_DC_Dq / L_Dq spectral loops and the x-direction part of their transforms
// (vie_fft/fft.h:41-397, paths relative to the reference root).
#pragma once

#include "vie_passes.h"

namespace vie {

template <int... Rs> struct XPlan {
    static constexpr int n = sizeof...(Rs);
    static constexpr int r[n] = {Rs...};
    static constexpr int stride(int s)
    {
        int v = 1;
        for (int u = 0; u < s; ++u) v *= r[u];
        return v;
    }
    static constexpr int product = (Rs * ...);
};

// CS: column-slow thread map (col = tid / T, the z passes, whose lanes walk
// along a contiguous row); otherwise col = tid % B (lanes walk across
// columns).  Exchange slots are padded by one every M positions, M = E for
// power-of-two E (else 4): with that padding every exchange of the plans in
// use is free of LDS bank conflicts in the column-fast map and within 1.75x
// of conflict-free in the column-slow map (64 lanes, 32 or 64 banks, 8- and
// 16-byte elements, checked by enumeration).
template <class R, int N, int T, class PL, bool CS = false> struct XCfg {
    static constexpr int P = 2 * N;
    static constexpr int E = P / T;
    static constexpr int B = NT / T;     // columns per workgroup
    static constexpr int M = (E & (E - 1)) == 0 ? E : 4;
    static constexpr int LM = M == 2 ? 1 : M == 4 ? 2 : M == 8 ? 3 : M == 16 ? 4 : M == 32 ? 5 : 6;
    static constexpr int PP = P + P / M;  // padded positions per column
    static constexpr int LDS = PP * B;    // exchange buffer elements
    // padded slot of position p, and the LDS index of slot l in column col
    __device__ static constexpr unsigned pad(unsigned p) { return p + (p >> LM); }
    __device__ static constexpr unsigned idx(unsigned l, unsigned col) { return CS ? col * PP + l : l * B + col; }
    __device__ static constexpr unsigned at(unsigned p, unsigned col) { return idx(pad(p), col); }
    // pad(p0 + d) == pad(p0) + pad(d) whenever p0 is a multiple of M: then the
    // slots of one butterfly are a common base plus compile-time offsets
    static constexpr int step(int d) { return CS ? d + d / M : (d + d / M) * B; }
    static_assert((1 << LM) == M, "pad period must be a power of two");
    static_assert(P % T == 0, "T must divide P");
    static_assert(PL::product == P, "the radices must multiply to the padded length");
    static_assert(B * T == NT, "T must divide the workgroup");
};

// forward stage s on one component's registers
template <class T_, class C, class PL, int s>
__device__ __forceinline__ void xf_stage(T_* v, const T_* tw, int t)
{
    constexpr int r = PL::r[s], S = PL::stride(s), P = C::P, E = C::E, T = P / E;
#pragma unroll
    for (int b = 0; b < E / r; ++b) {
        const int i = t + T * b;
        const int k = i % S;
        T_* a = v + b * r;
        if (S > 1) {
#pragma unroll
            for (int q = 1; q < r; ++q) a[q] = cmul(a[q], tw[q * k * (P / (S * r))]);
        }
        dft_impl<r, -1, T_>::run(a);
    }
}

// inverse stage s (reverse dataflow of xf_stage)
template <class T_, class C, class PL, int s>
__device__ __forceinline__ void xi_stage(T_* v, const T_* tw, int t)
{
    constexpr int r = PL::r[s], S = PL::stride(s), P = C::P, E = C::E, T = P / E;
#pragma unroll
    for (int b = 0; b < E / r; ++b) {
        const int i = t + T * b;
        const int k = i % S;
        T_* a = v + b * r;
        dft_impl<r, +1, T_>::run(a);
        if (S > 1) {
#pragma unroll
            for (int q = 1; q < r; ++q) a[q] = cmulc(a[q], tw[q * k * (P / (S * r))]);
        }
    }
}

// exchange after forward stage s: write its outputs (xf_put), then -- after
// a barrier -- read stage s+1 inputs (xf_get)
template <class T_, class C, class PL, int s>
__device__ __forceinline__ void xf_put(const T_* v, T_* lds, int t, int col)
{
    constexpr int r = PL::r[s], S = PL::stride(s), P = C::P, E = C::E, T = P / E;
    constexpr bool wsep = (S * r) % C::M == 0 && (S % C::M == 0 || C::M % S == 0);
#pragma unroll
    for (int b = 0; b < E / r; ++b) {
        const int i = t + T * b, k = i % S, j = (i - k) * r + k;
        if constexpr (wsep) {
            T_* w = lds + C::idx(C::pad(j), col);
#pragma unroll
            for (int q = 0; q < r; ++q) w[C::step(q * S)] = v[b * r + q];
        } else {
#pragma unroll
            for (int q = 0; q < r; ++q) lds[C::at(j + q * S, col)] = v[b * r + q];
        }
    }
}
template <class T_, class C, class PL, int s>
__device__ __forceinline__ void xf_get(T_* v, const T_* lds, int t, int col)
{
    constexpr int r2 = PL::r[s + 1], P = C::P, E = C::E, T = P / E;
    constexpr bool rsep = (P / r2) % C::M == 0;
#pragma unroll
    for (int b = 0; b < E / r2; ++b) {
        const int i = t + T * b;
        if constexpr (rsep) {
            const T_* w = lds + C::idx(C::pad(i), col);
#pragma unroll
            for (int q = 0; q < r2; ++q) v[b * r2 + q] = w[C::step(q * (P / r2))];
        } else {
#pragma unroll
            for (int q = 0; q < r2; ++q) v[b * r2 + q] = lds[C::at(i + q * (P / r2), col)];
        }
    }
}
// DB: the caller alternates between two exchange buffers, so the barrier
// after the reads is not needed (a buffer is written again only two
// exchanges later, after every thread has passed the barrier in between,
// which waits for its reads)
template <class T_, class C, class PL, int s, bool DB = false>
__device__ __forceinline__ void xf_xchg(T_* v, T_* lds, int t, int col)
{
    xf_put<T_, C, PL, s>(v, lds, t, col);
    __syncthreads();
    xf_get<T_, C, PL, s>(v, lds, t, col);
    if constexpr (!DB) __syncthreads();
}

// exchange after inverse stage s (s > 0): write at stage-s input positions
// (xi_put), read at stage s-1 output positions (xi_get)
template <class T_, class C, class PL, int s>
__device__ __forceinline__ void xi_put(const T_* v, T_* lds, int t, int col)
{
    constexpr int r = PL::r[s], P = C::P, E = C::E, T = P / E;
    constexpr bool wsep = (P / r) % C::M == 0;
#pragma unroll
    for (int b = 0; b < E / r; ++b) {
        const int i = t + T * b;
        if constexpr (wsep) {
            T_* w = lds + C::idx(C::pad(i), col);
#pragma unroll
            for (int q = 0; q < r; ++q) w[C::step(q * (P / r))] = v[b * r + q];
        } else {
#pragma unroll
            for (int q = 0; q < r; ++q) lds[C::at(i + q * (P / r), col)] = v[b * r + q];
        }
    }
}
template <class T_, class C, class PL, int s>
__device__ __forceinline__ void xi_get(T_* v, const T_* lds, int t, int col)
{
    constexpr int rp = PL::r[s - 1], Sp = PL::stride(s - 1), P = C::P, E = C::E, T = P / E;
    constexpr bool rsep = (Sp * rp) % C::M == 0 && (Sp % C::M == 0 || C::M % Sp == 0);
#pragma unroll
    for (int b = 0; b < E / rp; ++b) {
        const int i = t + T * b, k = i % Sp, j = (i - k) * rp + k;
        if constexpr (rsep) {
            const T_* w = lds + C::idx(C::pad(j), col);
#pragma unroll
            for (int q = 0; q < rp; ++q) v[b * rp + q] = w[C::step(q * Sp)];
        } else {
#pragma unroll
            for (int q = 0; q < rp; ++q) v[b * rp + q] = lds[C::at(j + q * Sp, col)];
        }
    }
}
template <class T_, class C, class PL, int s, bool DB = false>
__device__ __forceinline__ void xi_xchg(T_* v, T_* lds, int t, int col)
{
    xi_put<T_, C, PL, s>(v, lds, t, col);
    __syncthreads();
    xi_get<T_, C, PL, s>(v, lds, t, col);
    if constexpr (!DB) __syncthreads();
}

// A copy of v the compiler cannot see through.  With OPQ each component gets
// its own copy of the lane coordinates, so its twiddle loads and exchange
// addresses are recomputed instead of shared (and held live) across the three
// components: fewer registers for a few more LDS reads and integer ops.
__device__ __forceinline__ int opaque(int v)
{
    asm volatile("" : "+v"(v));
    // lane coordinates lie in [0, NT): with that known again, i % S of the
    // stages is a mask rather than a signed remainder with sign fix-ups
    __builtin_assume(v >= 0 && v < NT);
    return v;
}

// DB: exchange k (in order: forward stages, then inverse stages, three
// components each) uses buffer k mod 2 of lds (2 C::LDS entries)
template <class T_, class C, class PL, bool OPQ, bool DB = false, int s = 0>
__device__ __forceinline__ void x_forward(T_ (*v)[C::E], T_* lds, const T_* tw, int t, int col)
{
    if constexpr (s < PL::n) {
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const int tc = OPQ ? opaque(t) : t, cc = OPQ ? opaque(col) : col;
            xf_stage<T_, C, PL, s>(v[c], tw, tc);
            if constexpr (s + 1 < PL::n)
                xf_xchg<T_, C, PL, s, DB>(v[c], DB ? lds + ((3 * s + c) & 1) * C::LDS : lds, tc, cc);
        }
        x_forward<T_, C, PL, OPQ, DB, s + 1>(v, lds, tw, t, col);
    }
}

template <class T_, class C, class PL, bool OPQ, bool DB = false, int s = PL::n - 1>
__device__ __forceinline__ void x_inverse(T_ (*v)[C::E], T_* lds, const T_* tw, int t, int col)
{
    if constexpr (s >= 0) {
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const int tc = OPQ ? opaque(t) : t, cc = OPQ ? opaque(col) : col;
            xi_stage<T_, C, PL, s>(v[c], tw, tc);
            constexpr int k0 = 3 * (PL::n - 1) + 3 * (PL::n - 1 - s);  // exchanges before this stage's
            if constexpr (s > 0) xi_xchg<T_, C, PL, s, DB>(v[c], DB ? lds + ((k0 + c) & 1) * C::LDS : lds, tc, cc);
        }
        x_inverse<T_, C, PL, OPQ, DB, s - 1>(v, lds, tw, t, col);
    }
}

// Parity-order position of min(k, 2n - k) for the split-order index t of an
// axis of length n (frequency k = 2 (t mod n) + t / n): evens 0, 2, .., then odds.
__device__ __forceinline__ int fold_split(int t, int n)
{
    const int ep = t >= n ? 1 : 0;
    const int k = 2 * (t - ep * n) + ep;
    const int kf = k < 2 * n - k ? k : 2 * n - k;
    return (kf & 1) ? (n / 2 + 1) + (kf >> 1) : (kf >> 1);
}

// Waves per SIMD (= workgroups per CU: a workgroup has one wave per SIMD)
// that the x pass's register allocation allows, by configuration (the
// compiler's VGPR counts: 4 for the FP32 passes up to 64 and the FP64 ones up
// to 32, 2 for the 16-thread FP64 64 pass at 238 VGPRs; others are not
// double-buffered).
template <class R, int N, int T, int WPE> struct XWaves {
    static constexpr int v = (sizeof(R) == 4 && N <= 64) || (sizeof(R) == 8 && N <= 32) ? 4
                             : (sizeof(R) == 8 && N == 64 && T == 16 && WPE == 0) ? 2 : 8;
};

// grid (2Ny * ceil(2Nz / B)); N == Nx; in place on W2.
// twP[e] = exp(-2 pi j e / P), e < P.  WPE: waves per SIMD the register
// allocation aims for (0: compiler default).  GM: how the G octant is read --
// 0 per frequency from memory in the product, 1 a slice staged in LDS, 2 per
// frequency into registers before the forward transform.  OPQ: per-component
// lane coordinates (x_forward).  SUP: only pencil entries n < SUP can hold
// data (the caller's window lies within them), so the loads and stores of the
// rest, and the butterflies that only feed them, drop out at compile time.
template <class R, int N, int T, class PL, int WPE, int GM = (sizeof(R) == 8 ? 1 : 0), bool OPQ = false,
          int SUP = N>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(WPE > 0 ? WPE : 1)))
void k_xreg(Geo g, cx<R>* __restrict__ W2, const cx<R>* __restrict__ twP, MixTabs<R> mt, const int* done)
{
    using T_ = cx<R>;
    using C = XCfg<R, N, T, PL>;
    constexpr int P = C::P, E = C::E, B = C::B;
    constexpr int r1 = PL::r[0], rL = PL::r[PL::n - 1];
    constexpr bool GSL = GM == 1;
    // two exchange buffers (one barrier per exchange) where the workgroups
    // per CU that the registers allow (XWaves) still fit in LDS with them
    constexpr bool DB = XWaves<R, N, T, WPE>::v * (2 * C::LDS + 3 * P) * sizeof(T_) <= 160 * 1024;
    __shared__ T_ lds[(DB ? 2 : 1) * C::LDS];
    __shared__ T_ tw[P];
    __shared__ T_ uv[2 * P];                      // U and Vv of the x axis by frequency
    __shared__ T_ gsl[GSL ? (N + 1) * B : 1];    // G octant slice [fx][col]
    const int inst = bat_inst(g.bt, blockIdx.y);  // launch slot blockIdx.y of W2
    done = bat_done(g.bt, done, inst);
    if (done && *done) return;
    const InstTabs<R> it = inst_tabs(mt, inst);
    const int P1 = 2 * g.nyp, P2 = 2 * g.nzp;
    const int nzt = (P2 + B - 1) / B;
    const unsigned tile = xcd_tile(blockIdx.x, gridDim.x);
    const int ty = tile / nzt, z0 = (tile % nzt) * B;
    const int col = threadIdx.x % B, t = threadIdx.x / B;
    const bool colok = z0 + col < P2;
    // element offsets within a slot fit 32 bits (the host checks 3 x 4V <
    // 2^31 before launching), so the per-access address arithmetic stays 32-bit
    const unsigned rowst = (unsigned)P1 * P2;  // stride between x rows
    const unsigned cst = (unsigned)g.nx * rowst;  // stride between components
    // row n = i + m of a pencil sits at t rowst + (n - t) rowst: the lane's part
    // once (a 24-bit multiply: t < 64 and rowst <= 512^2, exact below 2^32),
    // the rest workgroup-uniform (scalar multiplies); a 32-bit vector multiply
    // issues at a quarter of the rate
    const unsigned trow = __umul24((unsigned)t, rowst);
    // entry 0 (plane xs) of the workgroup's first column: a workgroup-uniform
    // base, so the pencil accesses take a 32-bit per-lane byte offset from it
    // (a slot's 3 x Nx x 2Ny x 2Nz entries stay below 2^32 bytes up to 256^3)
    char* base = reinterpret_cast<char*>(W2 + (long)blockIdx.y * 3 * cst + ((long)g.xs * P1 + ty) * P2 + z0);
    const unsigned lane0 = trow + (unsigned)col;
    auto at = [base](unsigned e) -> T_& { return *reinterpret_cast<T_*>(base + e * (unsigned)sizeof(T_)); };

    // issue the pencil loads first, then the tables, so their latencies overlap
    T_ v[3][E];
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
        for (int b = 0; b < E / r1; ++b) {
            const int i = t + (P / E) * b;
#pragma unroll
            for (int q = 0; q < r1; ++q) {
                const int n = i + q * (P / r1);
                const unsigned off = (unsigned)((P / E) * b + q * (P / r1)) * rowst;
                v[c][b * r1 + q] = (n < SUP && colok && n + g.xs >= g.x0 && n + g.xs < g.x1)
                                       ? at(c * cst + lane0 + off)
                                       : czero<T_>();
            }
        }
    if constexpr (GSL) {
        const long gs = (long)(g.nyp + 1) * (g.nzp + 1);
        const long gy = (long)fold_split(ty, g.nyp) * (g.nzp + 1);
        for (int e = threadIdx.x; e < (N + 1) * B; e += NT) {
            const int fxe = e / B, ce = e % B;
            const int tze = z0 + ce < P2 ? z0 + ce : z0;
            T_ gv = it.G[(long)fxe * gs + gy + fold_split(tze, g.nzp)];
            gsl[e] = mt.gconj ? cconj(gv) : gv;
        }
    }
    for (int e = threadIdx.x; e < P; e += NT) {
        tw[e] = twP[e];
        const int ts = (e & 1) * N + (e >> 1);  // split-order index of frequency e
        uv[e] = it.U[0][ts];
        uv[P + e] = it.Vv[0][ts];
    }
    const T_ uy = it.U[1][ty], vy = it.Vv[1][ty];
    const int tz = colok ? z0 + col : z0;
    const T_ uz = it.U[2][tz], vz = it.Vv[2][tz];
    // G octant column of this thread: G[fx][fold(ky)][fold(kz)]
    // (a uniform row base and a 32-bit per-lane byte offset, as for the pencils:
    // the octant stays below 2^32 bytes up to 256^3)
    const char* grow = reinterpret_cast<const char*>(it.G + (long)fold_split(ty, g.nyp) * (g.nzp + 1));
    const unsigned gl = (unsigned)fold_split(tz, g.nzp);
    auto gat = [grow, gl](unsigned fxgs) -> const T_& {
        return *reinterpret_cast<const T_*>(grow + (gl + fxgs) * (unsigned)sizeof(T_));
    };
    const unsigned gs = (unsigned)(g.nyp + 1) * (g.nzp + 1);  // < 2^24; fx gs < 2^32 (24-bit multiplies)
    auto fxof = [](int kx) {  // parity-order index of min(kx, P - kx)
        const int kf = kx < P - kx ? kx : P - kx;
        return (kf & 1) ? (N / 2 + 1) + (kf >> 1) : (kf >> 1);
    };
    T_ gpre[GM == 2 ? E : 1];
    if constexpr (GM == 2) {
#pragma unroll
        for (int b = 0; b < E / rL; ++b)
#pragma unroll
            for (int q = 0; q < rL; ++q)
                gpre[b * rL + q] = gat(__umul24((unsigned)fxof(t + (P / E) * b + q * (P / rL)), gs));
    }
    __syncthreads();
    x_forward<T_, C, PL, OPQ, DB>(v, lds, tw, t, col);

    // D_f = j k0 Z0 delta_f / NP is imaginary: D_f X = (-Im D_f X.y, Im D_f X.x)
    const R D0 = it.D[0].y, D1 = it.D[1].y, D2 = it.D[2].y;
    // last forward stage: register e = b rL + q holds frequency i + q P/rL, i = t + T b
#pragma unroll
    for (int b = 0; b < E / rL; ++b) {
#pragma unroll
        for (int q = 0; q < rL; ++q) {
            const int kx = t + (P / E) * b + q * (P / rL);
            const int fx = fxof(kx);
            T_ Gv;
            if constexpr (GM == 1) {
                Gv = gsl[fx * B + col];
            } else {
                Gv = GM == 2 ? gpre[b * rL + q] : gat(__umul24((unsigned)fx, gs));
                if (mt.gconj) Gv = cconj(Gv);
            }
            const int e = b * rL + q;
            const T_ X0 = v[0][e], X1 = v[1][e], X2 = v[2][e];
            const T_ ux = uv[kx], vx = uv[P + kx];
            const T_ sig = cadd(cadd(cmul(vx, X0), cmul(vy, X1)), cmul(vz, X2));
            v[0][e] = cmul(Gv, cadd(mkc(-D0 * X0.y, D0 * X0.x), cmul(ux, sig)));
            v[1][e] = cmul(Gv, cadd(mkc(-D1 * X1.y, D1 * X1.x), cmul(uy, sig)));
            v[2][e] = cmul(Gv, cadd(mkc(-D2 * X2.y, D2 * X2.x), cmul(uz, sig)));
        }
    }
    x_inverse<T_, C, PL, OPQ, DB>(v, lds, tw, t, col);
    if (colok) {
#pragma unroll
        for (int c = 0; c < 3; ++c)
#pragma unroll
            for (int b = 0; b < E / r1; ++b) {
                const int i = t + (P / E) * b;
#pragma unroll
                for (int q = 0; q < r1; ++q) {
                    const int n = i + q * (P / r1);
                    const unsigned off = (unsigned)((P / E) * b + q * (P / r1)) * rowst;
                    if (n < SUP && n + g.xs >= g.x0 && n + g.xs < g.x1) at(c * cst + lane0 + off) = v[c][b * r1 + q];
                }
            }
    }
}

}  // namespace vie

namespace vie {

// ---------------------------------------------------------------------------
// Register-resident y and z passes (one component per thread).  Same engine as
// the x pass: a full padded transform of length P = 2N over T threads holding
// E = P/T values each.  Spectral data in HBM keeps the split order of the
// LDS-staged passes (t = (k mod 2) N + k / 2), so any pass variant can follow
// any other; the conversion is folded into the addressing.
// ---------------------------------------------------------------------------

template <class T_, class C, class PL, int s = 0>
__device__ __forceinline__ void r_forward(T_* v, T_* lds, const T_* tw, int t, int col)
{
    if constexpr (s < PL::n) {
        xf_stage<T_, C, PL, s>(v, tw, t);
        if constexpr (s + 1 < PL::n) xf_xchg<T_, C, PL, s>(v, lds, t, col);
        r_forward<T_, C, PL, s + 1>(v, lds, tw, t, col);
    }
}

template <class T_, class C, class PL, int s = PL::n - 1>
__device__ __forceinline__ void r_inverse(T_* v, T_* lds, const T_* tw, int t, int col)
{
    if constexpr (s >= 0) {
        xi_stage<T_, C, PL, s>(v, tw, t);
        if constexpr (s > 0) xi_xchg<T_, C, PL, s>(v, lds, t, col);
        r_inverse<T_, C, PL, s - 1>(v, lds, tw, t, col);
    }
}

template <int N> __device__ __forceinline__ int split_of(int k) { return (k & 1) * N + (k >> 1); }

// z forward: rows of Nz cells (optionally the Lin combination) -> W1 rows of
// 2Nz in split order.  grid (ceil(Nx*Ny / B), 3); N == Nz.
template <class R, int N, int T, class PL, bool ZS = false>
__global__ __launch_bounds__(NT) void k_zfwd_r(Geo g, Lin<R> lin, cx<R>* __restrict__ W1,
                                               const cx<R>* __restrict__ twP, const int* done)
{
    using T_ = cx<R>;
    using C = XCfg<R, N, T, PL, true>;
    constexpr int P = C::P, E = C::E, B = C::B;
    constexpr int r1 = PL::r[0], rL = PL::r[PL::n - 1];
    __shared__ T_ lds[C::LDS];
    __shared__ T_ tw[P];
    // component s of instance inst (field entries of component sb)
    const int s = blockIdx.y % 3, inst = bat_inst(g.bt, blockIdx.y / 3);
    const long sb = 3L * inst + s;
    done = bat_done(g.bt, done, inst);
    if (done && *done) return;
    // z row length and the z plane of entry 0: without ZS the pencil is the
    // whole row (N = nz, entry n is plane n), fixed at compile time
    const int nzr = ZS ? g.nz : N, zs = ZS ? g.zs : 0;
    const long rows = (long)g.nx * g.ny;
    const int col = threadIdx.x / T, t = threadIdx.x % T;  // consecutive lanes: consecutive cells
    // window row r of the x and y windows
    const int r = (int)xcd_tile(blockIdx.x, gridDim.x) * B + col;
    const bool rowok = r < win_rows(g);
    const long row = rowok ? win_row(g, r) : 0;
    T_ c[3];
    load_coef<R>(lin.coef ? lin.coef + (long)inst * g.bt.cstr : nullptr, c);
    const bool dropall = (s == 0 && row >= rows - g.ny) || (s == 1 && (row % g.ny) == g.ny - 1);
    // entry n is z plane n + zs of the row
    const long base = sb * g.V + row * nzr + zs;
    // z range outside which the inputs vanish (the whole row without lin.zr)
    int lo = 0, hi = nzr;
    if (lin.zr && rowok) {
        const unsigned rg = lin.zr[s * rows + row];
        lo = (int)(rg & 0xffffu);
        hi = (int)(rg >> 16);
    }
    // all loads before any store (lin.out may alias an input)
    T_ v[E], w1[E], w2[E], sc[E];
#pragma unroll
    for (int b = 0; b < E / r1; ++b) {
        const int i = t + T * b;
#pragma unroll
        for (int q = 0; q < r1; ++q) {
            const int n = i + q * (P / r1), e = b * r1 + q;
            v[e] = w1[e] = w2[e] = sc[e] = czero<T_>();
            if (n < N && rowok && n + zs >= lo && n + zs < hi) {
                v[e] = lin.x[0][base + n];
                if (lin.x[1]) w1[e] = lin.x[1][base + n];
                if (lin.x[2]) w2[e] = lin.x[2][base + n];
                if (lin.scale) sc[e] = lin.scale[base + n];
            }
        }
    }
#pragma unroll
    for (int b = 0; b < E / r1; ++b) {
        const int i = t + T * b;
#pragma unroll
        for (int q = 0; q < r1; ++q) {
            const int n = i + q * (P / r1), e = b * r1 + q;
            T_ z = cadd(cadd(cmul(c[0], v[e]), cmul(c[1], w1[e])), cmul(c[2], w2[e]));
            if (n < N && rowok && n + zs < nzr) {
                if (lin.out && (!lin.wskip || (n + zs >= lo && n + zs < hi))) lin.out[base + n] = z;
                if (lin.scale) z = cmul(z, sc[e]);  // zero outside the row's range
                if (dropall || (s == 2 && n + zs == nzr - 1)) z = czero<T_>();
            } else {
                z = czero<T_>();
            }
            v[e] = z;
        }
    }
    for (int e = threadIdx.x; e < P; e += NT) tw[e] = twP[e];
    __syncthreads();
    r_forward<T_, C, PL>(v, lds, tw, t, col);
    if (rowok) {
        T_* out = W1 + ((long)blockIdx.y * rows + row) * (2 * N);
#pragma unroll
        for (int b = 0; b < E / rL; ++b)
#pragma unroll
            for (int q = 0; q < rL; ++q) out[split_of<N>(t + T * b + q * (P / rL))] = v[b * rL + q];
    }
}

// z inverse: W1 rows (split order) -> y = mask (trunc(IDFT) + d x), dots.
// grid (ceil(Nx*Ny / B), 3); N == Nz.
template <class R, int N, int T, class PL, bool ZS = false>
__global__ __launch_bounds__(NT) void k_zinv_r(Geo g, const cx<R>* __restrict__ W1, const cx<R>* __restrict__ src,
                                               cx<R>* __restrict__ y, Diag<R> dg, const uint8_t* __restrict__ flags,
                                               Dots<R> dots, const cx<R>* __restrict__ twP, const int* done)
{
    using T_ = cx<R>;
    using C = XCfg<R, N, T, PL, true>;
    constexpr int P = C::P, E = C::E, B = C::B;
    constexpr int r1 = PL::r[0], rL = PL::r[PL::n - 1];
    __shared__ T_ lds[C::LDS];
    __shared__ T_ tw[P];
    __shared__ double2 red[NT / 64];
    // component f of instance inst (field entries of component fb)
    const int f = blockIdx.y % 3, inst = bat_inst(g.bt, blockIdx.y / 3);
    const long fb = 3L * inst + f;
    done = bat_done(g.bt, done, inst);
    if (done && *done) return;
    // z row length and the z plane of entry 0: without ZS the pencil is the
    // whole row (N = nz, entry n is plane n), fixed at compile time
    const int nzr = ZS ? g.nz : N, zs = ZS ? g.zs : 0;
    const long rows = (long)g.nx * g.ny;
    const int col = threadIdx.x / T, t = threadIdx.x % T;
    // window row r of the x and y windows
    const int r = (int)xcd_tile(blockIdx.x, gridDim.x) * B + col;
    const bool rowok = r < win_rows(g);
    const long row = rowok ? win_row(g, r) : 0;
    T_ v[E];
    {
        const T_* in = W1 + ((long)blockIdx.y * rows + (rowok ? row : 0)) * (2 * N);
#pragma unroll
        for (int b = 0; b < E / rL; ++b)
#pragma unroll
            for (int q = 0; q < rL; ++q)
                v[b * rL + q] = rowok ? in[split_of<N>(t + T * b + q * (P / rL))] : czero<T_>();
    }
    // epilogue operands of the kept rows n < N: loaded before the transform so
    // their latency overlaps it
    T_ sv[E], dv[E], e1v[E], e2v[E];
    uint8_t fl[E];
    const bool need1 = dots_use(dots, 1), need2 = dots_use(dots, 2);
    const bool e1src = dots.e1 == src;
    // entry n is z plane n + zs of the row (entries past Nz are not on the grid);
    // base indexes the shared flags (and a shared diagonal), bb the instance's vectors
    const long base = f * g.V + (rowok ? row : 0) * nzr + zs;
    const long bb = base + (fb - f) * g.V;
    const cx<R>* dgd = dg.d + inst * dg.dstr;
    if (rowok) {
        // flags first: x and d are read only on active faces (the output is
        // zero elsewhere), masked dot operands likewise
#pragma unroll
        for (int b = 0; b < E / r1; ++b)
#pragma unroll
            for (int q = 0; q < r1; ++q) {
                const int n = t + T * b + q * (P / r1);
                if (n < N && n + zs < nzr) fl[b * r1 + q] = dg.nomask ? 1 : flags[base + n];
            }
#pragma unroll
        for (int b = 0; b < E / r1; ++b) {
            const int i = t + T * b;
#pragma unroll
            for (int q = 0; q < r1; ++q) {
                const int n = i + q * (P / r1), e = b * r1 + q;
                if (n < N && n + zs < nzr) {
                    const long ia = bb + n;
                    sv[e] = dv[e] = e1v[e] = e2v[e] = czero<T_>();
                    if (fl[e] || (!dots.masked && need1 && e1src)) sv[e] = src[ia];
                    if (fl[e] && !dg.unit) dv[e] = dgd[base + n];
                    if (fl[e] || !dots.masked) {
                        if (need1 && !e1src) e1v[e] = dots.e1[ia];
                        if (need2) e2v[e] = dots.e2[ia];
                    }
                }
            }
        }
    }
    for (int e = threadIdx.x; e < P; e += NT) tw[e] = twP[e];
    __syncthreads();
    r_inverse<T_, C, PL>(v, lds, tw, t, col);
    double2 acc[3] = {make_double2(0, 0), make_double2(0, 0), make_double2(0, 0)};
    if (rowok) {
#pragma unroll
        for (int b = 0; b < E / r1; ++b) {
            const int i = t + T * b;
#pragma unroll
            for (int q = 0; q < r1; ++q) {
                const int n = i + q * (P / r1), e = b * r1 + q;
                if (n < N && n + zs < nzr) {
                    const long ia = bb + n;
                    T_ yv = czero<T_>();
                    if (fl[e]) yv = cadd(v[e], dg.unit ? sv[e] : cmul(dg.conj ? cconj(dv[e]) : dv[e], sv[e]));
                    if (fl[e] || !dg.skip) y[ia] = yv;
                    dots_add(acc, dots, yv, e1src ? sv[e] : need1 ? e1v[e] : czero<T_>(),
                             need2 ? e2v[e] : czero<T_>());
                }
            }
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
            double2 sm = block_sum(acc[d], red);
            if (threadIdx.x == 0) part[d * nblk + blk] = sm;
        }
    }
}

// y forward: W1 [Nx][Ny][2Nz] -> W2 [Nx][2Ny][2Nz], B k-columns per workgroup.
// grid (Nx * ceil(2Nz / B), 3); N == Ny.
template <class R, int N, int T, class PL>
__global__ __launch_bounds__(NT) void k_yfwd_r(Geo g, const cx<R>* __restrict__ W1, cx<R>* __restrict__ W2,
                                               const cx<R>* __restrict__ twP, const int* done)
{
    using T_ = cx<R>;
    using C = XCfg<R, N, T, PL>;
    constexpr int P = C::P, E = C::E, B = C::B;
    constexpr int r1 = PL::r[0], rL = PL::r[PL::n - 1];
    __shared__ T_ lds[C::LDS];
    __shared__ T_ tw[P];
    done = bat_done(g.bt, done, bat_inst(g.bt, blockIdx.y / 3));
    if (done && *done) return;
    const int s = blockIdx.y;  // launch slot (component of one instance) of W1 and W2
    const int P2 = 2 * g.nzp;
    const int nzt = (P2 + B - 1) / B;
    const unsigned tile = xcd_tile(blockIdx.x, gridDim.x);
    const int i = g.x0 + (int)(tile / nzt), z0 = (tile % nzt) * B;
    const int col = threadIdx.x % B, t = threadIdx.x / B;
    const bool colok = z0 + col < P2;
    // pencil entry n is y plane n + ys of W1 (Ny planes)
    const T_* in = W1 + (((long)s * g.nx + i) * g.ny + g.ys) * P2 + z0 + col;
    T_ v[E];
#pragma unroll
    for (int b = 0; b < E / r1; ++b) {
        const int ii = t + T * b;
#pragma unroll
        for (int q = 0; q < r1; ++q) {
            const int n = ii + q * (P / r1);
            v[b * r1 + q] = (n < N && colok && n + g.ys >= g.y0 && n + g.ys < g.y1) ? in[(long)n * P2]
                                                                                    : czero<T_>();
        }
    }
    for (int e = threadIdx.x; e < P; e += NT) tw[e] = twP[e];
    __syncthreads();
    r_forward<T_, C, PL>(v, lds, tw, t, col);
    if (colok) {
        T_* out = W2 + ((long)s * g.nx + i) * (2 * N) * P2 + z0 + col;
#pragma unroll
        for (int b = 0; b < E / rL; ++b)
#pragma unroll
            for (int q = 0; q < rL; ++q)
                out[(long)split_of<N>(t + T * b + q * (P / rL)) * P2] = v[b * rL + q];
    }
}

// y inverse: W2 -> W1 rows j < Ny.  grid (Nx * ceil(2Nz / B), 3); N == Ny.
template <class R, int N, int T, class PL>
__global__ __launch_bounds__(NT) void k_yinv_r(Geo g, const cx<R>* __restrict__ W2, cx<R>* __restrict__ W1,
                                               const cx<R>* __restrict__ twP, const int* done)
{
    using T_ = cx<R>;
    using C = XCfg<R, N, T, PL>;
    constexpr int P = C::P, E = C::E, B = C::B;
    constexpr int r1 = PL::r[0], rL = PL::r[PL::n - 1];
    __shared__ T_ lds[C::LDS];
    __shared__ T_ tw[P];
    done = bat_done(g.bt, done, bat_inst(g.bt, blockIdx.y / 3));
    if (done && *done) return;
    const int s = blockIdx.y;  // launch slot (component of one instance) of W1 and W2
    const int P2 = 2 * g.nzp;
    const int nzt = (P2 + B - 1) / B;
    const unsigned tile = xcd_tile(blockIdx.x, gridDim.x);
    const int i = g.x0 + (int)(tile / nzt), z0 = (tile % nzt) * B;
    const int col = threadIdx.x % B, t = threadIdx.x / B;
    const bool colok = z0 + col < P2;
    const T_* in = W2 + ((long)s * g.nx + i) * (2 * N) * P2 + z0 + col;
    T_ v[E];
#pragma unroll
    for (int b = 0; b < E / rL; ++b)
#pragma unroll
        for (int q = 0; q < rL; ++q)
            v[b * rL + q] = colok ? in[(long)split_of<N>(t + T * b + q * (P / rL)) * P2] : czero<T_>();
    for (int e = threadIdx.x; e < P; e += NT) tw[e] = twP[e];
    __syncthreads();
    r_inverse<T_, C, PL>(v, lds, tw, t, col);
    if (colok) {
        // pencil entry n is y plane n + ys of W1 (Ny planes)
        T_* out = W1 + (((long)s * g.nx + i) * g.ny + g.ys) * P2 + z0 + col;
#pragma unroll
        for (int b = 0; b < E / r1; ++b) {
            const int ii = t + T * b;
#pragma unroll
            for (int q = 0; q < r1; ++q) {
                const int n = ii + q * (P / r1);
                if (n < N && n + g.ys >= g.y0 && n + g.ys < g.y1) out[(long)n * P2] = v[b * r1 + q];
            }
        }
    }
}

}  // namespace vie
