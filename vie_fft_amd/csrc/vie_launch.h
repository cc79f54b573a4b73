// vie_launch.h -- host launchers of the pass kernels and the vector kernel,
// instantiated for double and float in vie_kernels_f64.hip / _f32.hip.
#pragma once

#include <hip/hip_ext.h>

#include "vie_passes.h"

namespace vie {

// Kernel timing (vie_time_kernel): a sampled launch of the timed class takes
// its start and stop events from here, armed just before it, and records
// them with the dispatch itself (hipExtLaunchKernel), so the measured time is
// the kernel's own, without the stream time an event record between two
// launches adds.  Launches that find nothing armed go out as usual.
struct LaunchEvents {
    hipEvent_t start = nullptr, stop = nullptr;
};
extern thread_local LaunchEvents g_launch_ev;

template <class F, class... Args>
void launch_timed(F kernel, dim3 grid, dim3 block, hipStream_t st, Args... args)
{
    const LaunchEvents e = g_launch_ev;
    g_launch_ev = LaunchEvents{};
    hipExtLaunchKernelGGL(kernel, grid, block, 0, st, e.start, e.stop, 0, args...);
}
#define VIE_LAUNCH(K, G, B, S, ...)                               \
    do {                                                          \
        if (::vie::g_launch_ev.start)                             \
            ::vie::launch_timed(K, G, B, S, __VA_ARGS__);         \
        else                                                      \
            hipLaunchKernelGGL(K, G, B, 0, S, __VA_ARGS__);       \
    } while (0)

template <class R> struct LinComb;

// axis lengths with compiled pass kernels
bool size_supported(int n);

// tw/twist: length-N tables of the LDS-staged passes; twP: exp(-2 pi j e / 2N),
// e < 2N, of the register-resident passes (all along the pass's own axis).
template <class R>
hipError_t run_zfwd(const Geo& g, const Lin<R>& lin, cx<R>* W1, const cx<R>* tw, const cx<R>* twist,
                    const cx<R>* twP, const int* done, hipStream_t st);
template <class R>
hipError_t run_yfwd(const Geo& g, const cx<R>* W1, cx<R>* W2, const cx<R>* tw, const cx<R>* twist,
                    const cx<R>* twP, const int* done, hipStream_t st);
// Pencils of length nxp (Nx, or the short period with g.xs = g.x0); twP:
// exp(-2 pi j e / (2 nxp)), e < 2 nxp (register x pass)
template <class R>
hipError_t run_xmix(const Geo& g, int nxp, cx<R>* W2, const cx<R>* tw, const cx<R>* twist, const cx<R>* twP,
                    const MixTabs<R>& mt, const int* done, hipStream_t st);
// whether length n has a register-resident x (y) pass (else the LDS-staged one)
template <class R> bool x_register(int n);
template <class R> bool y_register(int n);
template <class R>
hipError_t run_yinv(const Geo& g, const cx<R>* W2, cx<R>* W1, const cx<R>* tw, const cx<R>* twist,
                    const cx<R>* twP, const int* done, hipStream_t st);
// *nblk receives the number of partial sums written per dot product
template <class R>
hipError_t run_zinv(const Geo& g, const cx<R>* W1, const cx<R>* src, cx<R>* y, const Diag<R>& dg,
                    const uint8_t* flags, const Dots<R>& dots, const cx<R>* tw,
                    const cx<R>* twist, const cx<R>* twP, const int* done, hipStream_t st, long* nblk);
// partial-sum count a zinv launch produces for these sizes
template <class R> long zinv_blocks(const Geo& g);

}  // namespace vie
