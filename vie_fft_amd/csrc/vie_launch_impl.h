// vie_launch_impl.h -- launcher bodies; included once per precision with
// VIE_REAL defined as double or float.
#include "vie_launch.h"

// Axis lengths (products of 2, 3, 5) with compiled pass kernels.
#define VIE_SIZES(X) \
    X(2) X(3) X(4) X(5) X(6) X(8) X(10) X(12) X(16) X(20) X(24) X(32) X(40) X(48) X(64) \
    X(80) X(96) X(128) X(160) X(192) X(256)

namespace vie {

using R_ = VIE_REAL;

template <>
hipError_t run_zfwd<R_>(const Geo& g, const Lin<R_>& lin, cx<R_>* W1, const cx<R_>* tw,
                        const cx<R_>* twist, const int* done, hipStream_t st)
{
    const long rows = (long)g.nx * g.ny;
    switch (g.nz) {
#define CASE(n)                                                                               \
    case n: {                                                                                 \
        constexpr int B = b_z<R_>(n);                                                         \
        dim3 grid((unsigned)((rows + B - 1) / B), 3);                                         \
        hipLaunchKernelGGL((k_zfwd<R_, n, B>), grid, dim3(NT), 0, st, g, lin, W1, tw, twist, done); \
        break;                                                                                \
    }
        VIE_SIZES(CASE)
#undef CASE
    default:
        return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

template <>
hipError_t run_yfwd<R_>(const Geo& g, const cx<R_>* W1, cx<R_>* W2, const cx<R_>* tw,
                        const cx<R_>* twist, const int* done, hipStream_t st)
{
    switch (g.ny) {
#define CASE(n)                                                                               \
    case n: {                                                                                 \
        constexpr int B = b_y<R_>(n);                                                         \
        const long nzt = (2L * g.nz + B - 1) / B;                                             \
        dim3 grid((unsigned)(g.nx * nzt), 3);                                                 \
        hipLaunchKernelGGL((k_yfwd<R_, n, B>), grid, dim3(NT), 0, st, g, W1, W2, tw, twist, done); \
        break;                                                                                \
    }
        VIE_SIZES(CASE)
#undef CASE
    default:
        return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

template <>
hipError_t run_xmix<R_>(const Geo& g, cx<R_>* W2, const cx<R_>* tw, const cx<R_>* twist,
                        const MixTabs<R_>& mt, const int* done, hipStream_t st)
{
    switch (g.nx) {
#define CASE(n)                                                                               \
    case n: {                                                                                 \
        constexpr int B = b_x<R_>(n);                                                         \
        const long nzt = (2L * g.nz + B - 1) / B;                                             \
        dim3 grid((unsigned)(2L * g.ny * nzt));                                               \
        hipLaunchKernelGGL((k_xmix<R_, n, B>), grid, dim3(NT), 0, st, g, W2, tw, twist, mt, done); \
        break;                                                                                \
    }
        VIE_SIZES(CASE)
#undef CASE
    default:
        return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

template <>
hipError_t run_yinv<R_>(const Geo& g, const cx<R_>* W2, cx<R_>* W1, const cx<R_>* tw,
                        const cx<R_>* twist, const int* done, hipStream_t st)
{
    switch (g.ny) {
#define CASE(n)                                                                               \
    case n: {                                                                                 \
        constexpr int B = b_y<R_>(n);                                                         \
        const long nzt = (2L * g.nz + B - 1) / B;                                             \
        dim3 grid((unsigned)(g.nx * nzt), 3);                                                 \
        hipLaunchKernelGGL((k_yinv<R_, n, B>), grid, dim3(NT), 0, st, g, W2, W1, tw, twist, done); \
        break;                                                                                \
    }
        VIE_SIZES(CASE)
#undef CASE
    default:
        return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

template <> long zinv_blocks<R_>(const Geo& g)
{
    const long rows = (long)g.nx * g.ny;
    switch (g.nz) {
#define CASE(n)                                                                               \
    case n: {                                                                                 \
        constexpr int B = b_z<R_>(n);                                                         \
        return 3 * ((rows + B - 1) / B);                                                      \
    }
        VIE_SIZES(CASE)
#undef CASE
    default:
        return 0;
    }
}

template <>
hipError_t run_zinv<R_>(const Geo& g, const cx<R_>* W1, const cx<R_>* src, cx<R_>* y,
                        const cx<R_>* dg, int dconj, const uint8_t* flags, const Dots<R_>& dots,
                        const cx<R_>* tw, const cx<R_>* twist, const int* done, hipStream_t st,
                        long* nblk)
{
    const long rows = (long)g.nx * g.ny;
    switch (g.nz) {
#define CASE(n)                                                                               \
    case n: {                                                                                 \
        constexpr int B = b_z<R_>(n);                                                         \
        dim3 grid((unsigned)((rows + B - 1) / B), 3);                                         \
        if (nblk) *nblk = (long)grid.x * grid.y;                                              \
        hipLaunchKernelGGL((k_zinv<R_, n, B>), grid, dim3(NT), 0, st, g, W1, src, y, dg, dconj, \
                           flags, dots, tw, twist, done);                                     \
        break;                                                                                \
    }
        VIE_SIZES(CASE)
#undef CASE
    default:
        return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

}  // namespace vie
