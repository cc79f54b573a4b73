// vie_fft_device.h -- gfx950 device code for the VIE-FFT operator.
//
// The padded convolution of the reference (2N points per axis, zero input on
// [N, 2N), output kept on [0, N); vie_fft/CG_Iteration.h:50-83 and
// vie_fft/fft.h:113-123) is computed with length-N transforms only: along an
// axis of padded length P = 2N,
//     X[2m + e] = sum_n (x[n] w_P^{e n}) w_N^{m n},        e in {0, 1},
//     y[n]      = IDFT_N(Y[2m])[n] + w_P^{-n} IDFT_N(Y[2m + 1])[n],  n < N,
// so each pass reads N (or 2N) points per pencil and never touches padding.
// A spectral axis is stored as t = e*N + m ("split order").
//
// Five passes make one product y = A x (all three components at once):
//   zfwd : x (V per component)      -> W1 [Nx][Ny][2Nz]      pencils along k
//   yfwd : W1                       -> W2 [Nx][2Ny][2Nz]     pencils along j
//   xmix : W2 -> W2 in place: forward along i, the 3x3 spectral product with
//          the Green's function, inverse along i, truncation to i < Nx
//   yinv : W2                       -> W1
//   zinv : W1 -> y (V), truncation, diagonal term d_f x_f, output mask and
//          optional fused dot products for the Krylov loop.
// Strided passes take B consecutive k-columns per workgroup so that every
// global row access is B contiguous complex numbers.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace vie {

constexpr int NT = 256;  // threads per workgroup for every pass kernel

// ---------------------------------------------------------------- complex
template <typename R> struct cx_of;
template <> struct cx_of<double> { using type = double2; };
template <> struct cx_of<float> { using type = float2; };
template <typename R> using cx = typename cx_of<R>::type;

__device__ __forceinline__ double2 mkc(double a, double b) { return make_double2(a, b); }
__device__ __forceinline__ float2 mkc(float a, float b) { return make_float2(a, b); }

template <class T> __device__ __forceinline__ T cadd(T a, T b) { return mkc(a.x + b.x, a.y + b.y); }
template <class T> __device__ __forceinline__ T csub(T a, T b) { return mkc(a.x - b.x, a.y - b.y); }
template <class T> __device__ __forceinline__ T cmul(T a, T b)
{
    return mkc(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
// a * conj(b)
template <class T> __device__ __forceinline__ T cmulc(T a, T b)
{
    return mkc(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y);
}
template <class T> __device__ __forceinline__ T cconj(T a) { return mkc(a.x, -a.y); }
template <class T, class S> __device__ __forceinline__ T cscale(T a, S s) { return mkc(a.x * s, a.y * s); }
// multiply by s*j (s = +1 or -1)
template <int S, class T> __device__ __forceinline__ T mulj(T a)
{
    return S > 0 ? mkc(-a.y, a.x) : mkc(a.y, -a.x);
}

// ---------------------------------------------------------------- radix plan
constexpr int next_radix(int n)
{
    return n % 8 == 0 ? 8 : n % 4 == 0 ? 4 : n % 2 == 0 ? 2 : n % 3 == 0 ? 3 : n % 5 == 0 ? 5 : n;
}
constexpr bool plan_ok(int n)
{
    while (n > 1) {
        int r = next_radix(n);
        if (r != 2 && r != 3 && r != 4 && r != 5 && r != 8) return false;
        n /= r;
    }
    return true;
}

// In-register DFT of length R; D = -1 forward (exp(-j)), +1 inverse.
template <int R, int D, class T> struct dft_impl;

template <int D, class T> struct dft_impl<2, D, T> {
    static __device__ __forceinline__ void run(T* v)
    {
        T a = v[0], b = v[1];
        v[0] = cadd(a, b);
        v[1] = csub(a, b);
    }
};
template <int D, class T> struct dft_impl<4, D, T> {
    static __device__ __forceinline__ void run(T* v)
    {
        T a0 = cadd(v[0], v[2]), a1 = csub(v[0], v[2]);
        T b0 = cadd(v[1], v[3]), b1 = mulj<D>(csub(v[1], v[3]));
        v[0] = cadd(a0, b0);
        v[1] = cadd(a1, b1);
        v[2] = csub(a0, b0);
        v[3] = csub(a1, b1);
    }
};
template <int D, class T> struct dft_impl<8, D, T> {
    static __device__ __forceinline__ void run(T* v)
    {
        using R = decltype(v[0].x);
        const R h = (R)0.70710678118654752440084436210485;
        T e[4] = {v[0], v[2], v[4], v[6]};
        T o[4] = {v[1], v[3], v[5], v[7]};
        dft_impl<4, D, T>::run(e);
        dft_impl<4, D, T>::run(o);
        // w8^1 = (1 + D j)/sqrt2, w8^2 = D j, w8^3 = (-1 + D j)/sqrt2
        T o1 = cscale(cadd(o[1], mulj<D>(o[1])), h);
        T o2 = mulj<D>(o[2]);
        T o3 = cscale(csub(mulj<D>(o[3]), o[3]), h);
        v[0] = cadd(e[0], o[0]);
        v[4] = csub(e[0], o[0]);
        v[1] = cadd(e[1], o1);
        v[5] = csub(e[1], o1);
        v[2] = cadd(e[2], o2);
        v[6] = csub(e[2], o2);
        v[3] = cadd(e[3], o3);
        v[7] = csub(e[3], o3);
    }
};
template <int D, class T> struct dft_impl<3, D, T> {
    static __device__ __forceinline__ void run(T* v)
    {
        using R = decltype(v[0].x);
        const R s = (R)(D * 0.86602540378443864676372317075294);
        T t1 = cadd(v[1], v[2]);
        T t2 = csub(v[0], cscale(t1, (R)0.5));
        T t3 = cscale(csub(v[1], v[2]), s);
        v[0] = cadd(v[0], t1);
        v[1] = cadd(t2, mulj<1>(t3));
        v[2] = csub(t2, mulj<1>(t3));
    }
};
template <int D, class T> struct dft_impl<5, D, T> {
    static __device__ __forceinline__ void run(T* v)
    {
        using R = decltype(v[0].x);
        const R c1 = (R)0.30901699437494742410229341718282;   // cos(2pi/5)
        const R c2 = (R)-0.80901699437494742410229341718282;  // cos(4pi/5)
        const R s1 = (R)0.95105651629515357211643933337938;   // sin(2pi/5)
        const R s2 = (R)0.58778525229247312916870595463907;   // sin(4pi/5)
        T p1 = cadd(v[1], v[4]), m1 = csub(v[1], v[4]);
        T p2 = cadd(v[2], v[3]), m2 = csub(v[2], v[3]);
        T a1 = cadd(v[0], cadd(cscale(p1, c1), cscale(p2, c2)));
        T a2 = cadd(v[0], cadd(cscale(p1, c2), cscale(p2, c1)));
        T b1 = cadd(cscale(m1, s1), cscale(m2, s2));
        T b2 = csub(cscale(m1, s2), cscale(m2, s1));
        v[0] = cadd(v[0], cadd(p1, p2));
        // forward (D=-1): X1 = a1 - j b1; inverse flips the sign of j
        v[1] = cadd(a1, mulj<D>(b1));
        v[4] = csub(a1, mulj<D>(b1));
        v[2] = cadd(a2, mulj<D>(b2));
        v[3] = csub(a2, mulj<D>(b2));
    }
};

// radix 6 = 2 x 3, prime-factor (Good-Thomas) order, so no twiddles between
// the two: n = (3 n1 + 2 n2) mod 6, k = (3 k1 + 4 k2) mod 6 and
// w6^{n k} = w2^{n1 k1} w3^{n2 k2}.  Length-2 transforms over n1 for each n2,
// then length-3 transforms over n2 for each k1.
template <int D, class T> struct dft_impl<6, D, T> {
    static __device__ __forceinline__ void run(T* v)
    {
        T e[3] = {cadd(v[0], v[3]), cadd(v[2], v[5]), cadd(v[4], v[1])};
        T o[3] = {csub(v[0], v[3]), csub(v[2], v[5]), csub(v[4], v[1])};
        dft_impl<3, D, T>::run(e);
        dft_impl<3, D, T>::run(o);
        v[0] = e[0];
        v[4] = e[1];
        v[2] = e[2];
        v[3] = o[0];
        v[1] = o[1];
        v[5] = o[2];
    }
};

// radix 16 = 4 x 4: n = 4 n1 + n2, k = k1 + 4 k2,
// X[k] = sum_n2 w4^{n2 k2} (w16^{n2 k1} sum_n1 w4^{n1 k1} x[4 n1 + n2])
template <int D, class T> struct dft_impl<16, D, T> {
    static __device__ __forceinline__ void run(T* v)
    {
        using R = decltype(v[0].x);
        const R c1 = (R)0.92387953251128675612818318939679;  // cos(pi/8)
        const R s1 = (R)0.38268343236508977172845998403040;  // sin(pi/8)
        const R h = (R)0.70710678118654752440084436210485;
#pragma unroll
        for (int n2 = 0; n2 < 4; ++n2) {
            T a[4] = {v[n2], v[n2 + 4], v[n2 + 8], v[n2 + 12]};
            dft_impl<4, D, T>::run(a);
#pragma unroll
            for (int k1 = 0; k1 < 4; ++k1) v[n2 + 4 * k1] = a[k1];
        }
        // twiddles w16^{n2 k1}: w = c + D j s (D = -1: exp(-j ...), D = +1: its conjugate),
        // x w = (x.x c - D x.y s, x.y c + D x.x s)
#define VIE_TW(idx, c, s) v[idx] = mkc(v[idx].x * (c) - (R)D * v[idx].y * (s), v[idx].y * (c) + (R)D * v[idx].x * (s))
        VIE_TW(1 + 4 * 1, c1, s1);   // w^1
        VIE_TW(1 + 4 * 2, h, h);     // w^2
        VIE_TW(1 + 4 * 3, s1, c1);   // w^3
        VIE_TW(2 + 4 * 1, h, h);     // w^2
        v[2 + 4 * 2] = mulj<D>(v[2 + 4 * 2]);  // w^4 = D j
        VIE_TW(2 + 4 * 3, -h, h);    // w^6
        VIE_TW(3 + 4 * 1, s1, c1);   // w^3
        VIE_TW(3 + 4 * 2, -h, h);    // w^6
        VIE_TW(3 + 4 * 3, -c1, -s1); // w^9
#undef VIE_TW
        T o[16];
#pragma unroll
        for (int k1 = 0; k1 < 4; ++k1) {
            T a[4] = {v[4 * k1], v[4 * k1 + 1], v[4 * k1 + 2], v[4 * k1 + 3]};
            dft_impl<4, D, T>::run(a);
#pragma unroll
            for (int k2 = 0; k2 < 4; ++k2) o[k1 + 4 * k2] = a[k2];
        }
#pragma unroll
        for (int k = 0; k < 16; ++k) v[k] = o[k];
    }
};

// One Stockham radix-R stage over COLS independent columns held in LDS as
// buf[element * LD + column].  S is the product of the radices already done.
// Butterfly i reads elements i + q*N/R, applies w_{S R}^{q (i mod S)} and
// writes (i - i mod S)*R + i mod S + q*S.  tw[t] = exp(-2 pi j t / N).
template <class T, int N, int R, int S, int COLS, int LD, int D>
__device__ __forceinline__ void fft_stage(T* buf, const T* tw)
{
    constexpr int NBF = COLS * (N / R);
    constexpr int PER = (NBF + NT - 1) / NT;
    T v[PER][R];
#pragma unroll
    for (int b = 0; b < PER; ++b) {
        const int g = threadIdx.x + b * NT;
        if ((NBF % NT == 0) || g < NBF) {
            const int c = g % COLS, i = g / COLS, k = i % S;
#pragma unroll
            for (int q = 0; q < R; ++q) v[b][q] = buf[(i + q * (N / R)) * LD + c];
            if (S > 1) {
#pragma unroll
                for (int q = 1; q < R; ++q) {
                    const T w = tw[q * k * (N / (S * R))];
                    v[b][q] = D < 0 ? cmul(v[b][q], w) : cmulc(v[b][q], w);
                }
            }
            dft_impl<R, D, T>::run(v[b]);
        }
    }
    __syncthreads();
#pragma unroll
    for (int b = 0; b < PER; ++b) {
        const int g = threadIdx.x + b * NT;
        if ((NBF % NT == 0) || g < NBF) {
            const int c = g % COLS, i = g / COLS, k = i % S;
            const int j = (i - k) * R + k;
#pragma unroll
            for (int q = 0; q < R; ++q) buf[(j + q * S) * LD + c] = v[b][q];
        }
    }
    __syncthreads();
}

// Full length-N transform of COLS columns in LDS (natural order in and out).
// The caller synchronises before the call; the call ends synchronised.
template <class T, int N, int COLS, int LD, int D, int S = 1, int REM = N>
__device__ __forceinline__ void fft_lds(T* buf, const T* tw)
{
    if constexpr (REM > 1) {
        constexpr int R = next_radix(REM);
        static_assert(R == 2 || R == 3 || R == 4 || R == 5 || R == 8, "unsupported radix");
        fft_stage<T, N, R, S, COLS, LD, D>(buf, tw);
        fft_lds<T, N, COLS, LD, D, S * R, REM / R>(buf, tw);
    }
}

// Workgroup-order remap: gfx950 dispatches consecutive workgroups round-robin
// over its 8 XCDs, each with a private L2.  Remap so that consecutive logical
// tiles (which share cache lines of neighbouring rows) land on one XCD.
__device__ __forceinline__ unsigned xcd_tile(unsigned bid, unsigned nblk)
{
    const unsigned per = nblk / 8u;
    if (per == 0 || bid >= per * 8u) return bid;
    return (bid % 8u) * per + bid / 8u;
}

// ---------------------------------------------------------------- args
// Linear combination prologue: z = c0 x0 + c1 x1 + c2 x2 (coefficients in
// device memory, so a solver can update them without a host round trip).
// zr: nullptr, or per z-row active ranges (lo | hi << 16, k_zrange) when
// every input is zero on inactive faces (Krylov vectors: the masked
// right-hand side and masked products keep them so); entries outside a row's
// range are then not read.
// wskip (with zr): out, and the product this prologue feeds, already hold
// zero outside the active faces, so those entries are not written.
// scale (optional, 3V): the product transforms scale (.) z instead of z (out
// still receives z): the right-preconditioned product A (m (.) z).
template <class R> struct Lin {
    const cx<R>* x[3];
    const double2* coef;  // 3 complex; nullptr means (1, 0, 0)
    cx<R>* out;           // optional: store z (full vector, before padding)
    const uint32_t* zr;
    int wskip;
    const cx<R>* scale;
};

// Dot-product epilogue: partial[k * nblk + blk] = sum conj(a_k) b_k (or
// a_k b_k when nc[k] is set: the unconjugated product of zdotu) with operands
// chosen from {0: the output y, 1: e1, 2: e2, 3: a second output}.
// masked: e1 and e2 are zero on inactive faces, so they are read only where
// the output's face is active.
template <class R> struct Dots {
    int n;          // 0..3
    int a[3], b[3];
    int nc[3];
    const cx<R>* e1;
    const cx<R>* e2;
    double2* partial;
    int masked;
};

struct Flag {
    const int* done;  // nullptr or a device flag: kernels exit when set
};

}  // namespace vie
