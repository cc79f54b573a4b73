// vie_launch_impl.h -- launcher bodies; included once per precision with
// VIE_REAL defined as double or float.
#include "vie_launch.h"
#include "vie_xpass.h"

#include <algorithm>
#include <cstdlib>

// Axis lengths (products of 2, 3, 5) with compiled pass kernels.
#define VIE_SIZES(X) \
    X(2) X(3) X(4) X(5) X(6) X(8) X(10) X(12) X(16) X(20) X(24) X(32) X(40) X(48) X(64) \
    X(80) X(96) X(128) X(160) X(192) X(256)

namespace vie {

using R_ = VIE_REAL;

// Register-resident y/z pass configurations (one component per thread).
template <class R, int N> struct YZReg { static constexpr bool ok = false; };
template <> struct YZReg<float, 32> { static constexpr bool ok = true; static constexpr int T = 4; using PL = XPlan<16, 4>; };
template <> struct YZReg<float, 64> { static constexpr bool ok = true; static constexpr int T = 8; using PL = XPlan<16, 8>; };
template <> struct YZReg<float, 96> { static constexpr bool ok = true; static constexpr int T = 16; using PL = XPlan<4, 4, 4, 3>; };
template <> struct YZReg<float, 128> { static constexpr bool ok = true; static constexpr int T = 16; using PL = XPlan<16, 16>; };
template <> struct YZReg<float, 256> { static constexpr bool ok = true; static constexpr int T = 32; using PL = XPlan<16, 16, 2>; };
template <> struct YZReg<double, 32> { static constexpr bool ok = true; static constexpr int T = 8; using PL = XPlan<8, 8>; };
template <> struct YZReg<double, 64> { static constexpr bool ok = true; static constexpr int T = 16; using PL = XPlan<8, 4, 4>; };
template <> struct YZReg<double, 96> { static constexpr bool ok = true; static constexpr int T = 16; using PL = XPlan<4, 4, 4, 3>; };
template <> struct YZReg<double, 128> { static constexpr bool ok = true; static constexpr int T = 32; using PL = XPlan<8, 8, 4>; };
template <> struct YZReg<double, 256> { static constexpr bool ok = true; static constexpr int T = 64; using PL = XPlan<8, 8, 8>; };

// z-inverse register configuration (the YZReg one unless specialised).  FP32
// 128: T = 32 (half the LDS per workgroup of T = 16, so twice the workgroups
// per CU) takes 49.6 us per launch at 128^3 against 55.2 us.
template <class R, int N> struct ZInvReg : YZReg<R, N> {};
template <> struct ZInvReg<float, 128> { static constexpr bool ok = true; static constexpr int T = 32; using PL = XPlan<8, 8, 4>; };

// VIE_PASSES_LDS=1 selects the LDS-staged y/z passes for every size (A/B timing).
static bool yz_lds_forced()
{
    static const bool v = [] {
        const char* e = std::getenv("VIE_PASSES_LDS");
        return e && e[0] == '1';
    }();
    return v;
}

// Whether the z pencils are a short period (entry n: plane n + zs of rows of
// nz > nzp cells); the full-row instantiations fix nz = N and zs = 0, which
// keeps the 128-point z passes at their earlier cost (one more compare and a
// run-time row stride cost the FP32 z inverse pass 7%).
static inline bool short_z(const Geo& g) { return g.nzp != g.nz || g.zs != 0; }

// z forward pass: the LDS version for FP32, the register version for FP64.
// Over the 48-plane window of the 128^3 layered workload the LDS version
// takes 26.2 us FP32 / 46.8 us FP64 per launch, the register one 29.3 / 42.7
// us (on every plane, before windows: 52 / 96 against 63 / 100 us).
// VIE_ZFWD_REG=1 (=0) selects the register (LDS) version for both.
template <class R> static bool zfwd_reg()
{
    static const int v = [] {
        const char* e = std::getenv("VIE_ZFWD_REG");
        return e ? (e[0] == '1' ? 1 : 0) : -1;
    }();
    return v < 0 ? sizeof(R) == 8 : v == 1;
}

// The variant a launch uses: the Geo's tuned plan, else the fixed defaults
// (and their VIE_* overrides).
template <class R> static bool use_zfwd_reg(const Geo& g) { return g.var < 0 ? zfwd_reg<R>() : (g.var & 1) != 0; }
static bool use_yz_lds(const Geo& g) { return g.var < 0 ? yz_lds_forced() : (g.var & 2) != 0; }
static bool xpass_lds_forced();
static bool use_x_lds(const Geo& g) { return g.var < 0 ? xpass_lds_forced() : (g.var & 4) != 0; }

template <class R, int n, bool ZS>
static hipError_t launch_zfwd(const Geo& g, const Lin<R>& lin, cx<R>* W1, const cx<R>* tw, const cx<R>* twist,
                              const cx<R>* twP, const int* done, hipStream_t st)
{
    const long rows = win_rows(g);  // the rows of the x and y windows
    if constexpr (YZReg<R, n>::ok) {
        if (use_zfwd_reg<R>(g) && !use_yz_lds(g)) {
            using X = YZReg<R, n>;
            constexpr int B = NT / X::T;
            dim3 grid((unsigned)((rows + B - 1) / B), 3 * g.bt.nb);
            VIE_LAUNCH((k_zfwd_r<R, n, X::T, typename X::PL, ZS>), grid, dim3(NT), st, g, lin, W1, twP, done);
            return hipGetLastError();
        }
    }
    constexpr int B = b_z<R>(n);
    dim3 grid((unsigned)((rows + B - 1) / B), 3 * g.bt.nb);
    VIE_LAUNCH((k_zfwd<R, n, B, ZS>), grid, dim3(NT), st, g, lin, W1, tw, twist, done);
    return hipGetLastError();
}

template <class R, int n>
static hipError_t launch_yfwd(const Geo& g, const cx<R>* W1, cx<R>* W2, const cx<R>* tw, const cx<R>* twist,
                              const cx<R>* twP, const int* done, hipStream_t st)
{
    if constexpr (YZReg<R, n>::ok) {
        if (!use_yz_lds(g)) {
            using X = YZReg<R, n>;
            constexpr int B = NT / X::T;
            const long nzt = (2L * g.nzp + B - 1) / B;
            dim3 grid((unsigned)((g.x1 - g.x0) * nzt), 3 * g.bt.nb);
            VIE_LAUNCH((k_yfwd_r<R, n, X::T, typename X::PL>), grid, dim3(NT), st, g, W1, W2, twP, done);
            return hipGetLastError();
        }
    }
    constexpr int B = b_y<R>(n);
    const long nzt = (2L * g.nzp + B - 1) / B;
    dim3 grid((unsigned)((g.x1 - g.x0) * nzt), 3 * g.bt.nb);
    VIE_LAUNCH((k_yfwd<R, n, B>), grid, dim3(NT), st, g, W1, W2, tw, twist, done);
    return hipGetLastError();
}

template <class R, int n>
static hipError_t launch_yinv(const Geo& g, const cx<R>* W2, cx<R>* W1, const cx<R>* tw, const cx<R>* twist,
                              const cx<R>* twP, const int* done, hipStream_t st)
{
    if constexpr (YZReg<R, n>::ok) {
        if (!use_yz_lds(g)) {
            using X = YZReg<R, n>;
            constexpr int B = NT / X::T;
            const long nzt = (2L * g.nzp + B - 1) / B;
            dim3 grid((unsigned)((g.x1 - g.x0) * nzt), 3 * g.bt.nb);
            VIE_LAUNCH((k_yinv_r<R, n, X::T, typename X::PL>), grid, dim3(NT), st, g, W2, W1, twP, done);
            return hipGetLastError();
        }
    }
    constexpr int B = b_y<R>(n);
    const long nzt = (2L * g.nzp + B - 1) / B;
    dim3 grid((unsigned)((g.x1 - g.x0) * nzt), 3 * g.bt.nb);
    VIE_LAUNCH((k_yinv<R, n, B>), grid, dim3(NT), st, g, W2, W1, tw, twist, done);
    return hipGetLastError();
}

template <class R, int n, bool ZS>
static hipError_t launch_zinv(const Geo& g, const cx<R>* W1, const cx<R>* src, cx<R>* y, const Diag<R>& dg,
                              const uint8_t* flags, const Dots<R>& dots, const cx<R>* tw, const cx<R>* twist,
                              const cx<R>* twP, const int* done, hipStream_t st, long* nblk)
{
    const long rows = win_rows(g);  // the rows of the x and y windows
    if constexpr (ZInvReg<R, n>::ok) {
        if (!use_yz_lds(g) && dg.d) {
            using X = ZInvReg<R, n>;
            constexpr int B = NT / X::T;
            dim3 grid((unsigned)((rows + B - 1) / B), 3 * g.bt.nb);
            if (nblk) *nblk = 3L * grid.x;  // per launch slot
            VIE_LAUNCH((k_zinv_r<R, n, X::T, typename X::PL, ZS>), grid, dim3(NT), st, g, W1, src, y, dg,
                               flags, dots, twP, done);
            return hipGetLastError();
        }
    }
    constexpr int B = b_z<R>(n);
    dim3 grid((unsigned)((rows + B - 1) / B), 3 * g.bt.nb);
    if (nblk) *nblk = 3L * grid.x;  // per launch slot
    VIE_LAUNCH((k_zinv<R, n, B, ZS>), grid, dim3(NT), st, g, W1, src, y, dg, flags, dots, tw, twist, done);
    return hipGetLastError();
}

template <class R, int n> static long zinv_blocks_n(const Geo& g)
{
    const long rows = (long)g.nx * g.ny;
    long r = 3 * ((rows + b_z<R>(n) - 1) / b_z<R>(n));
    if constexpr (ZInvReg<R, n>::ok) {
        constexpr int B = NT / ZInvReg<R, n>::T;
        r = std::max(r, 3 * ((rows + B - 1) / B));
    }
    return r;
}

template <>
hipError_t run_zfwd<R_>(const Geo& g, const Lin<R_>& lin, cx<R_>* W1, const cx<R_>* tw,
                        const cx<R_>* twist, const cx<R_>* twP, const int* done, hipStream_t st)
{
    switch (g.nzp) {
#define CASE(n) \
    case n:                                                                                     \
        return short_z(g) ? launch_zfwd<R_, n, true>(g, lin, W1, tw, twist, twP, done, st)        \
                          : launch_zfwd<R_, n, false>(g, lin, W1, tw, twist, twP, done, st);
        VIE_SIZES(CASE)
#undef CASE
    default:
        return hipErrorInvalidValue;
    }
}

template <>
hipError_t run_yfwd<R_>(const Geo& g, const cx<R_>* W1, cx<R_>* W2, const cx<R_>* tw,
                        const cx<R_>* twist, const cx<R_>* twP, const int* done, hipStream_t st)
{
    switch (g.nyp) {
#define CASE(n) \
    case n: return launch_yfwd<R_, n>(g, W1, W2, tw, twist, twP, done, st);
        VIE_SIZES(CASE)
#undef CASE
    default:
        return hipErrorInvalidValue;
    }
}

// Register-resident x-pass configurations: T threads per column, radix plan,
// W waves-per-SIMD register target (0: compiler default), GM how G~ is read,
// OP per-component lane coordinates (k_xreg).  At 128 (the 128^3 workload,
// measured per launch on MI355X): FP32 T16 <16,16> with G~ prefetched into
// registers 133 us against 149 us for T32 <8,8,4> reading G~ in the product
// and 187 us for T64 <4,4,4,4>; with OP its registers fall from 208 to 164
// (three waves per SIMD instead of two) and it takes 118 us.  FP64 T32
// <8,8,4> with the prefetch 229 us against 276 us with an LDS slice; OP
// leaves it at two waves (169 registers) and costs 262 us, with G~ read in
// the product (148 registers) 264 us.  OP is on where it adds a wave per SIMD
// (FP32 96: 211 -> 159 registers, 128); FP32 256 also reads G~ in the
// product (240 -> 159 registers: 1578 against 1653 us at 256^3, where FP64
// that way takes 3884 against 2983 us); T64 <8,8,8> without OP (124
// registers, four waves, 15% fewer instructions per pencil) takes 2045
// against 1513 us.  FP32 64 (the short x period of the
// 128^3 layered workload, 48-plane window): G~ read in the product without
// OP is 128 registers, four waves, and 407 VALU instructions per pencil
// against 503 with OP: 56.9-58.0 us per launch against 60.8 us (G~
// prefetched with OP), 61.9 us (prefetched, no OP) and 60.2 us (T8 <16,8>).
// Of the T16 radix orders (window of 48) <8,4,4> issues the fewest VALU and
// LDS instructions per wave: 1532 and 142, against 1533-1726 and 157-223 for
// <8,8,2>, <8,2,8>, <4,8,4>, <4,4,8>, <2,8,8>, <4,4,4,2> and <2,4,4,4>.
// A five-wave register target (96 registers, 14 dwords spilled to scratch)
// takes 94.5 against 55.7 us; FP64 at three waves spills likewise.  FP64 64
// with OP and G~ read in the product reaches three waves without spilling
// (147 registers) but issues 23% more instructions: 117.7 against 105.0 us.
template <class R, int N> struct XReg { static constexpr bool ok = false; };
#define VIE_XREG(R, N, T_, W_, GM_, OP_, ...)                                                                \
    template <> struct XReg<R, N> {                                                                       \
        static constexpr bool ok = true;                                                                  \
        static constexpr int T = T_, W = W_, GM = GM_;                                                    \
        static constexpr bool OP = OP_;                                                                   \
        using PL = XPlan<__VA_ARGS__>;                                                                    \
    };
VIE_XREG(float, 32, 8, 0, 2, false, 8, 8)
VIE_XREG(float, 48, 8, 3, 0, false, 4, 4, 6)
VIE_XREG(float, 64, 16, 0, 0, false, 8, 4, 4)
VIE_XREG(float, 96, 16, 0, 2, true, 4, 4, 4, 3)
VIE_XREG(float, 128, 16, 0, 2, true, 16, 16)
VIE_XREG(float, 256, 32, 0, 0, true, 16, 16, 2)
VIE_XREG(double, 32, 8, 0, 2, false, 8, 8)
VIE_XREG(double, 64, 16, 0, 2, false, 8, 4, 4)
VIE_XREG(double, 96, 16, 0, 2, false, 4, 4, 4, 3)
VIE_XREG(double, 128, 32, 0, 2, false, 8, 8, 4)
VIE_XREG(double, 256, 64, 0, 2, false, 8, 8, 8)
#undef VIE_XREG

// VIE_XPASS_LDS=1 selects the LDS-staged x pass for every size (A/B timing).
static bool xpass_lds_forced()
{
    static const bool v = [] {
        const char* e = std::getenv("VIE_XPASS_LDS");
        return e && e[0] == '1';
    }();
    return v;
}

template <class R, int n>
static hipError_t launch_x(const Geo& g, cx<R>* W2, const cx<R>* tw, const cx<R>* twist, const cx<R>* twP,
                           const MixTabs<R>& mt, const int* done, hipStream_t st)
{
    if constexpr (XReg<R, n>::ok) {
        if (!use_x_lds(g)) {
            using X = XReg<R, n>;
            constexpr int B = NT / X::T;
            const long nzt = (2L * g.nzp + B - 1) / B;
            dim3 grid((unsigned)(2L * g.nyp * nzt), g.bt.nb);
            // FP32: pencils whose data lies in the first half or three quarters
            // (a window narrower than the pencil) use a kernel that knows it
            // (64-length pencils, 48-plane window: 1532 VALU instructions per
            // wave against 1628, 56.4 against 57.4 us per launch at 128^3)
            const int need = g.x1 - g.xs;
            if constexpr (sizeof(R) == 4 && n % 4 == 0) {
                if (need <= n / 2) {
                    VIE_LAUNCH((k_xreg<R, n, X::T, typename X::PL, X::W, X::GM, X::OP, n / 2>), grid, dim3(NT), st, g, W2, twP, mt, done);
                    return hipGetLastError();
                }
                if (need <= 3 * n / 4) {
                    VIE_LAUNCH((k_xreg<R, n, X::T, typename X::PL, X::W, X::GM, X::OP, 3 * n / 4>), grid, dim3(NT), st, g, W2, twP, mt, done);
                    return hipGetLastError();
                }
            }
            VIE_LAUNCH((k_xreg<R, n, X::T, typename X::PL, X::W, X::GM, X::OP>), grid, dim3(NT), st, g, W2, twP, mt,
                               done);
            return hipGetLastError();
        }
    }
    constexpr int B = b_x<R>(n);
    const long nzt = (2L * g.nzp + B - 1) / B;
    dim3 grid((unsigned)(2L * g.nyp * nzt), g.bt.nb);
    VIE_LAUNCH((k_xmix<R, n, B>), grid, dim3(NT), st, g, W2, tw, twist, mt, done);
    return hipGetLastError();
}

template <> bool y_register<R_>(int n)
{
    if (yz_lds_forced()) return false;
    switch (n) {
#define CASE(m) \
    case m: return YZReg<R_, m>::ok;
        VIE_SIZES(CASE)
#undef CASE
    default:
        return false;
    }
}

template <> bool x_register<R_>(int n)
{
    if (xpass_lds_forced()) return false;
    switch (n) {
#define CASE(m) \
    case m: return XReg<R_, m>::ok;
        VIE_SIZES(CASE)
#undef CASE
    default:
        return false;
    }
}

template <>
hipError_t run_xmix<R_>(const Geo& g, int nxp, cx<R_>* W2, const cx<R_>* tw, const cx<R_>* twist, const cx<R_>* twP,
                        const MixTabs<R_>& mt, const int* done, hipStream_t st)
{
    switch (nxp) {
#define CASE(n) \
    case n: return launch_x<R_, n>(g, W2, tw, twist, twP, mt, done, st);
        VIE_SIZES(CASE)
#undef CASE
    default:
        return hipErrorInvalidValue;
    }
}

template <>
hipError_t run_yinv<R_>(const Geo& g, const cx<R_>* W2, cx<R_>* W1, const cx<R_>* tw,
                        const cx<R_>* twist, const cx<R_>* twP, const int* done, hipStream_t st)
{
    switch (g.nyp) {
#define CASE(n) \
    case n: return launch_yinv<R_, n>(g, W2, W1, tw, twist, twP, done, st);
        VIE_SIZES(CASE)
#undef CASE
    default:
        return hipErrorInvalidValue;
    }
}

template <> long zinv_blocks<R_>(const Geo& g)
{
    switch (g.nz) {
#define CASE(n) \
    case n: return zinv_blocks_n<R_, n>(g);
        VIE_SIZES(CASE)
#undef CASE
    default:
        return 0;
    }
}

template <>
hipError_t run_zinv<R_>(const Geo& g, const cx<R_>* W1, const cx<R_>* src, cx<R_>* y,
                        const Diag<R_>& dg, const uint8_t* flags, const Dots<R_>& dots,
                        const cx<R_>* tw, const cx<R_>* twist, const cx<R_>* twP, const int* done,
                        hipStream_t st, long* nblk)
{
    switch (g.nzp) {
#define CASE(n) \
    case n:                                                                                           \
        return short_z(g) ? launch_zinv<R_, n, true>(g, W1, src, y, dg, flags, dots, tw, twist, twP, done, st, nblk) \
                          : launch_zinv<R_, n, false>(g, W1, src, y, dg, flags, dots, tw, twist, twP, done, st, nblk);
        VIE_SIZES(CASE)
#undef CASE
    default:
        return hipErrorInvalidValue;
    }
}

}  // namespace vie
