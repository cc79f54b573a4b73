// vie_api.hip -- C ABI of the gfx950 VIE-FFT path (declared in include/viefft.h).
//
// Host side of the operator and of the five Krylov loops.  Reference
// correspondence (paths relative to the reference root):
//   Green_Function_Compute  vie_fft/GreenFunction.h:8-141   green_compute()
//   Incident_Wave           vie_fft/GreenFunction.h:144-248 vie_incident()
//   X/Y/Z_FFT (+Conjugate)  vie_fft/CG_Iteration.h:86-173   Op<R>::apply()
//   CG_FFT                  vie_fft/CG_Iteration.h:176-357  solve_cg()
//   BICG_FFT                vie_fft/CG_Iteration.h:360-528  solve_bicg()
//   BICGSTAB_FFT            vie_fft/CG_Iteration.h:531-776  solve_bicgstab()
//   CORS_FFT                vie_fft/CG_Iteration.h:780-1073 solve_cors()
//   BICORSTAB_FFT           vie_fft/CG_Iteration.h:1076-1362 solve_bicorstab()
//   FFT_Program             vie_fft/CG_Iteration.h:1365-1388 vie_solve*()
//   FarField1, NearField    vie_fft/NearFar.h:16-109, :180-250
#include "../../include/viefft.h"
#include "vie_krylov.h"
#include "vie_launch.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <complex>
#include <cstring>
#include <cstdio>
#include <iterator>
#include <map>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

using namespace vie;

// the event pair armed for the next timed launch (see Timer)
thread_local vie::LaunchEvents vie::g_launch_ev;

namespace {

const double kPi = 3.1415926535897932384626433832795;
const double kZ0 = 120.0 * kPi;  // vie_fft/common.h:18
const double kEps = 1.0e-7;      // vie_fft/common.h:16
const double kEpsBackground = 1.1;  // dummy background permittivity, vie_fft/Model.h:82

struct Err : std::runtime_error {
    int code;
    Err(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

#define HIPCHK(x)                                                                             \
    do {                                                                                      \
        hipError_t e_ = (x);                                                                  \
        if (e_ != hipSuccess)                                                                 \
            throw Err(e_ == hipErrorOutOfMemory ? VIE_ERR_OOM : VIE_ERR_DEVICE,               \
                      std::string(#x) + ": " + hipGetErrorString(e_));                        \
    } while (0)

// ---- large device buffers are kept for reuse
// A freed buffer of at least kPoolMin bytes is kept (per device, up to a third
// of the device's memory) and handed out again for a request of the same
// rounded size, instead of going back to the HIP runtime: contexts created one
// after another with the same sizes (sweeps, batches of tens of GB) then
// allocate nothing.  The runtime's own release of such frees is deferred and
// has blocked a later hipMalloc for seconds after a few tens of GB of churn.
// A kept buffer holds the contents it had (as memory the runtime recycles
// does).  Freeing synchronises the device first, as hipFree does, so no
// kernel still using the buffer can overlap its next owner's work.
constexpr size_t kPoolMin = size_t(64) << 20;
constexpr size_t kPoolAlign = size_t(2) << 20;
struct PoolBuf {
    int dev;
    size_t bytes;
    void* p;
};
static std::mutex g_pool_mu;
static std::vector<PoolBuf> g_pool;          // kept buffers
static std::map<void*, PoolBuf> g_pool_live;  // handed-out buffers of pooled size
static std::map<int, size_t> g_pool_kept;     // kept bytes per device

// a third of device dev's memory; dev must be the current device
static size_t pool_cap(int dev)
{
    static std::map<int, size_t> cap;
    auto it = cap.find(dev);
    if (it != cap.end()) return it->second;
    size_t fr = 0, tot = 0;
    (void)hipMemGetInfo(&fr, &tot);
    return cap[dev] = tot / 3;
}

// bytes of the current device's kept buffers (memory that a new buffer can
// take although hipMemGetInfo counts it as used)
size_t pool_kept_bytes()
{
    int dev = 0;
    (void)hipGetDevice(&dev);
    std::lock_guard<std::mutex> lk(g_pool_mu);
    return g_pool_kept[dev];
}

static void* raw_alloc(size_t bytes)
{
    void* p = nullptr;
    if (bytes < kPoolMin) {
        if (hipMalloc(&p, bytes) != hipSuccess) p = nullptr;
        return p;
    }
    bytes = (bytes + kPoolAlign - 1) / kPoolAlign * kPoolAlign;
    int dev = 0;
    (void)hipGetDevice(&dev);
    std::lock_guard<std::mutex> lk(g_pool_mu);
    for (size_t i = 0; i < g_pool.size(); ++i)
        if (g_pool[i].dev == dev && g_pool[i].bytes == bytes) {
            const PoolBuf b = g_pool[i];
            g_pool.erase(g_pool.begin() + (long)i);
            g_pool_kept[dev] -= bytes;
            g_pool_live[b.p] = b;
            return b.p;
        }
    if (hipMalloc(&p, bytes) != hipSuccess) {
        // out of memory: give this device's kept buffers back and try again
        (void)hipGetLastError();
        for (size_t i = 0; i < g_pool.size();)
            if (g_pool[i].dev == dev) {
                (void)hipFree(g_pool[i].p);
                g_pool.erase(g_pool.begin() + (long)i);
            } else {
                ++i;
            }
        g_pool_kept[dev] = 0;
        if (hipMalloc(&p, bytes) != hipSuccess) return nullptr;
    }
    g_pool_live[p] = PoolBuf{dev, bytes, p};
    return p;
}

static void raw_free(void* p)
{
    if (!p) return;
    {
        std::lock_guard<std::mutex> lk(g_pool_mu);
        auto it = g_pool_live.find(p);
        if (it != g_pool_live.end()) {
            const PoolBuf b = it->second;
            g_pool_live.erase(it);
            int cur = 0;
            (void)hipGetDevice(&cur);
            if (cur != b.dev) (void)hipSetDevice(b.dev);
            const bool keep = g_pool_kept[b.dev] + b.bytes <= pool_cap(b.dev);
            if (keep) (void)hipDeviceSynchronize();
            if (cur != b.dev) (void)hipSetDevice(cur);
            if (keep) {
                g_pool.push_back(b);
                g_pool_kept[b.dev] += b.bytes;
                return;
            }
        }
    }
    (void)hipFree(p);
}

// hipFree every kept buffer of the current device; returns the bytes released
static size_t pool_release()
{
    int dev = 0;
    (void)hipGetDevice(&dev);
    std::lock_guard<std::mutex> lk(g_pool_mu);
    size_t rel = 0;
    for (size_t i = 0; i < g_pool.size();)
        if (g_pool[i].dev == dev) {
            (void)hipFree(g_pool[i].p);
            rel += g_pool[i].bytes;
            g_pool.erase(g_pool.begin() + (long)i);
        } else {
            ++i;
        }
    g_pool_kept[dev] = 0;
    return rel;
}

template <class T> T* dalloc(size_t n)
{
    if (n == 0) n = 1;
    void* p = raw_alloc(n * sizeof(T));
    if (!p) throw Err(VIE_ERR_OOM, "hipMalloc of " + std::to_string(n * sizeof(T)) + " bytes failed");
    return (T*)p;
}
template <class T> void dfree(T*& p)
{
    raw_free(p);
    p = nullptr;
}

// Every copy goes through one stream and completes before returning: the
// library stream is non-blocking, so null-stream copies would be unordered
// with its kernels (device-to-device hipMemcpy may also return early).
void copy_sync(void* dst, const void* src, size_t bytes, hipMemcpyKind kind, hipStream_t st)
{
    HIPCHK(hipMemcpyAsync(dst, src, bytes, kind, st));
    HIPCHK(hipStreamSynchronize(st));
}

unsigned grid1(long n, int nt = NT, long cap = 4096)
{
    long g = (n + nt - 1) / nt;
    if (g > cap) g = cap;
    if (g < 1) g = 1;
    return (unsigned)g;
}

template <class R> cx<R> hc(std::complex<double> z)
{
    cx<R> v;
    v.x = (R)z.real();
    v.y = (R)z.imag();
    return v;
}

// Per-launch HIP-event timing of one kernel class.  mark() is called before
// and after every launch of the class; for one launch in kEvery the first
// mark arms a start/stop event pair that the launch records with its own
// dispatch (VIE_LAUNCH, hipExtLaunchKernel), so the time is the kernel's
// alone.  A launch that leaves the pair armed (it did not go through
// VIE_LAUNCH) is not sampled.
struct Timer {
    static constexpr long kEvery = 8;
    int cls = 0;
    std::vector<hipEvent_t> ev;
    size_t used = 0;
    long seen = 0;     // launches of the class since reset
    bool open = false; // between the two marks of a launch
    void reset()
    {
        used = 0;
        seen = 0;
        open = false;
        vie::g_launch_ev = LaunchEvents{};
    }
    void mark(hipStream_t)
    {
        if (!cls) return;
        open = !open;
        if (!open) {  // after the launch
            if (vie::g_launch_ev.start) {
                vie::g_launch_ev = LaunchEvents{};
                used -= 2;
            }
            return;
        }
        if (seen++ % kEvery) return;
        while (used + 2 > ev.size()) {
            hipEvent_t e;
            HIPCHK(hipEventCreate(&e));
            ev.push_back(e);
        }
        vie::g_launch_ev = LaunchEvents{ev[used], ev[used + 1]};
        used += 2;
    }
    ~Timer()
    {
        for (auto e : ev) (void)hipEventDestroy(e);
    }
};

// ------------------------------------------------------------------ operator
template <class R> struct Op {
    using T = cx<R>;
    Geo g{};
    int n[3];
    double d[3], k0;
    // per-axis tables, slots 0-2 the grid's axes; slots 3, 4 and 5 the x, y
    // and z axes of the short periods (lengths nxs, nys, nzs)
    T* tw[6] = {};
    T* twist[6] = {};
    T* twP[6] = {};    // exp(-2 pi j e / (2 N_a)), e < 2 N_a, per axis
    T* U[6] = {};
    T* Vv[6] = {};
    T* Ua[6] = {};
    T* Va[6] = {};
    int* fold[6] = {};
    T* G = nullptr;
    T* dg = nullptr;           // per-face diagonal d_f (3V)
    // Jacobi preconditioner m = 1/d_f on active faces (3V), built on first use
    // after each change of dg (dgen counts the changes, pmgen is the dgen pm
    // was built for)
    T* pm = nullptr;
    long dgen = 0, pmgen = -1;
    uint8_t* flags = nullptr;
    uint32_t* zr = nullptr;    // per z-row active range lo | hi << 16 (3 Nx Ny)
    int wx0 = 0, wx1 = 0;      // x window: the x planes with an active face
    int wy0 = 0, wy1 = 0;      // y window: the y planes with an active face
    int wz0 = 0, wz1 = 0;      // z window: the z planes with an active face (wz0 even)
    // Short-period x pass: a product whose input and output live on the x
    // window [wx0, wx1) is a linear convolution along x of width
    // w = wx1 - wx0, which a padded transform of any length 2 nxs with
    // nxs >= w reproduces (the window shifted to plane 0).  nxs = 0: the
    // x pass runs at the grid's length.  Gs is the G~ octant of the
    // (nxs, Ny, Nz) grid scaled by Nx / nxs (the tables keep the grid's
    // normalisation 1/(8 Nx Ny Nz)).  Products whose outputs the solver keeps
    // zero (wskip) also use the y window [wy0, wy1) and the z window [wz0,
    // wz1): with nys > 0 (nzs > 0) their y (z) transforms have the short
    // period nys >= wy1 - wy0 (nzs >= wz1 - wz0), and their x pass reads the
    // octant Gw of the (nxs or Nx, nys or Ny, nzs or Nz) grid, scaled likewise.
    int nxs = 0, nys = 0, nzs = 0;
    T* Gs = nullptr;
    T* Gw = nullptr;
    T* W1 = nullptr;
    T* W2 = nullptr;
    bool own_w = true;
    MixTabs<R> mf{}, ma{}, mfs{}, mas{}, mfy{}, may{};
    // Batches: the launch slots of batched products (the context's current
    // slots) and, when the instances have their own operators (several
    // wavenumbers), their tables for the windowed products' grid (pencil
    // lengths bl[3] along x, y, z): G~ octants (gob values each), forward Vv
    // and adjoint U tables (2 lb values per axis each), D_f forward and
    // adjoint (3 each) and d_f (3V each), for nbt instances.
    // tuned plan (autotuner): pass variants of every product (Geo::var) and
    // the short periods of the windowed products; var -1 and tuned false:
    // the fixed defaults
    int var = -1;
    bool tuned = false;
    int tns = 0, tnys = 0, tnzs = 0;
    double tus = 0.0;      // microseconds per windowed product of the plan
    double tus_def = 0.0;  // and of the fixed defaults (measured when tuning)
    const Bat* bt = nullptr;
    bool bmulti = false;
    int nbt = 0, lb = 0, bl[3] = {0, 0, 0};
    long gob = 0;
    T* Gb = nullptr;
    T* Vb[3] = {};
    T* Uab[3] = {};
    T* Db = nullptr;
    T* Dab = nullptr;
    T* dgb = nullptr;
    MixTabs<R> mfb{}, mab{};

    // Length-N tables of slot a: LDS-pass twiddles, the register passes'
    // twP, U, the adjoint's V and the G~ fold (the k0-dependent Vv and Ua
    // come from axis_k0).
    void build_axis(int a, int N, hipStream_t st)
    {
        std::vector<T> h_tw(N), h_twist(N), h_u(2 * N), h_va(2 * N), h_p(2 * N);
        std::vector<int> h_fold(2 * N);
        for (int t = 0; t < N; ++t) {
            const long double a1 = -2.0L * (long double)kPi * t / N;
            const long double a2 = -2.0L * (long double)kPi * t / (2 * N);
            h_tw[t] = hc<R>(std::complex<double>((double)cosl(a1), (double)sinl(a1)));
            h_twist[t] = hc<R>(std::complex<double>((double)cosl(a2), (double)sinl(a2)));
        }
        for (int t = 0; t < 2 * N; ++t) {
            const int k = 2 * (t % N) + t / N;
            const long double th = (long double)kPi * k / N;
            const double s = (double)sinl(th), h = (double)sinl(th / 2);
            // e^{j th} - 1 = -2 sin^2(th/2) + j sin th
            const std::complex<double> u(-2.0 * h * h, s);
            h_u[t] = hc<R>(u);
            h_va[t] = hc<R>(std::conj(u));
            const int kf = std::min(k, 2 * N - k);
            h_fold[t] = (kf % 2 == 0) ? kf / 2 : (N / 2 + 1) + (kf - 1) / 2;
            const long double ang = -2.0L * (long double)kPi * t / (2 * N);
            h_p[t] = hc<R>(std::complex<double>((double)cosl(ang), (double)sinl(ang)));
        }
        tw[a] = dalloc<T>(N);
        twist[a] = dalloc<T>(N);
        twP[a] = dalloc<T>(2 * N);
        U[a] = dalloc<T>(2 * N);
        Vv[a] = dalloc<T>(2 * N);
        Ua[a] = dalloc<T>(2 * N);
        Va[a] = dalloc<T>(2 * N);
        fold[a] = dalloc<int>(2 * N);
        copy_sync(tw[a], h_tw.data(), N * sizeof(T), hipMemcpyHostToDevice, st);
        copy_sync(twist[a], h_twist.data(), N * sizeof(T), hipMemcpyHostToDevice, st);
        copy_sync(twP[a], h_p.data(), 2 * N * sizeof(T), hipMemcpyHostToDevice, st);
        copy_sync(U[a], h_u.data(), 2 * N * sizeof(T), hipMemcpyHostToDevice, st);
        copy_sync(Va[a], h_va.data(), 2 * N * sizeof(T), hipMemcpyHostToDevice, st);
        copy_sync(fold[a], h_fold.data(), 2 * N * sizeof(int), hipMemcpyHostToDevice, st);
    }
    void free_axis(int a)
    {
        dfree(tw[a]);
        dfree(twist[a]);
        dfree(twP[a]);
        dfree(U[a]);
        dfree(Vv[a]);
        dfree(Ua[a]);
        dfree(Va[a]);
        dfree(fold[a]);
    }

    // w1, w2: pass buffers shared with another operator, else cap slots' own
    Op(const int* nn, const double* dd, double kk, T* w1, T* w2, hipStream_t st, int cap = 1)
    {
        for (int a = 0; a < 3; ++a) {
            n[a] = nn[a];
            d[a] = dd[a];
        }
        k0 = kk;
        g.nx = n[0];
        g.ny = n[1];
        g.nz = n[2];
        g.V = (long)n[0] * n[1] * n[2];
        g.x0 = wx0 = 0;
        g.x1 = wx1 = n[0];
        g.y0 = wy0 = 0;
        g.y1 = wy1 = n[1];
        g.nyp = n[1];
        wz0 = 0;
        wz1 = n[2];
        g.nzp = n[2];
        for (int a = 0; a < 3; ++a) build_axis(a, n[a], st);
        const long gsz = (long)(n[0] + 1) * (n[1] + 1) * (n[2] + 1);
        G = dalloc<T>(gsz);
        flags = dalloc<uint8_t>(3 * g.V);
        zr = dalloc<uint32_t>(3L * g.nx * g.ny);
        if (w1) {
            W1 = w1;
            W2 = w2;
            own_w = false;
        } else {
            W1 = dalloc<T>(3 * 2 * g.V * cap);
            W2 = dalloc<T>(3 * 4 * g.V * cap);
        }
        for (int a = 0; a < 3; ++a) {
            mf.U[a] = U[a];
            mf.Vv[a] = Vv[a];
            mf.fold[a] = fold[a];
            ma.U[a] = Ua[a];
            ma.Vv[a] = Va[a];
            ma.fold[a] = fold[a];
        }
        mf.G = ma.G = G;
        mf.gconj = 0;
        ma.gconj = 1;
        set_k0(kk, st);
    }

    // Short periods of lengths ns (x), nsy (y) and nsz (z), 0: none:
    // slot-3/4/5 tables, the octants Gs and Gw (filled by the caller from
    // green_octant) and the MixTabs that use them.
    void set_short(int ns, int nsy, int nsz, hipStream_t st)
    {
        if (ns == nxs && nsy == nys && nsz == nzs) return;
        if (ns != nxs) {
            free_axis(3);
            nxs = ns;
            if (ns) {
                build_axis(3, ns, st);
                axis_k0(3, ns, d[0], st);
            }
        }
        if (nsy != nys) {
            free_axis(4);
            nys = nsy;
            if (nsy) {
                build_axis(4, nsy, st);
                axis_k0(4, nsy, d[1], st);
            }
        }
        if (nsz != nzs) {
            free_axis(5);
            nzs = nsz;
            if (nsz) {
                build_axis(5, nsz, st);
                axis_k0(5, nsz, d[2], st);
            }
        }
        dfree(Gs);
        dfree(Gw);
        if (nxs) Gs = dalloc<T>((long)(nxs + 1) * (n[1] + 1) * (n[2] + 1));
        if (nys || nzs)
            Gw = dalloc<T>((long)((nxs ? nxs : n[0]) + 1) * ((nys ? nys : n[1]) + 1) * ((nzs ? nzs : n[2]) + 1));
        short_tabs();
    }
    void short_tabs()
    {
        mfs = mf;
        mas = ma;
        if (nxs) {
            mfs.U[0] = U[3];
            mfs.Vv[0] = Vv[3];
            mfs.fold[0] = fold[3];
            mas.U[0] = Ua[3];
            mas.Vv[0] = Va[3];
            mas.fold[0] = fold[3];
            mfs.G = mas.G = Gs;
        }
        mfy = mfs;
        may = mas;
        for (int a = 1; a < 3; ++a) {
            if (!(a == 1 ? nys : nzs)) continue;
            mfy.U[a] = U[3 + a];
            mfy.Vv[a] = Vv[3 + a];
            mfy.fold[a] = fold[3 + a];
            may.U[a] = Ua[3 + a];
            may.Vv[a] = Va[3 + a];
            may.fold[a] = fold[3 + a];
        }
        if (nys || nzs) mfy.G = may.G = Gw;
    }

    // v_s = (j Z0/(k0 delta_s NP)) (1 - e^{-j theta_s}) of slot a (its
    // conjugate is the adjoint's u), NP = 8 Nx Ny Nz of the grid.
    void axis_k0(int a, int N, double da, hipStream_t st)
    {
        std::vector<T> h_v(2 * N), h_ua(2 * N);
        k0_tables(N, da, k0, h_v.data(), h_ua.data());
        copy_sync(Vv[a], h_v.data(), 2 * N * sizeof(T), hipMemcpyHostToDevice, st);
        copy_sync(Ua[a], h_ua.data(), 2 * N * sizeof(T), hipMemcpyHostToDevice, st);
    }
    // Vv (and its conjugate, the adjoint's U) of an axis of length N and cell
    // size da for wavenumber kk
    void k0_tables(int N, double da, double kk, T* v, T* ua) const
    {
        std::vector<std::complex<double>> w(2 * N);
        k0_phase(N, w.data());
        k0_scale(N, da, kk, w.data(), v, ua);
    }
    // its wavenumber-free factor w_t = 1 - e^{-j th}, th = pi k / N for split
    // index t (frequency k = 2 (t mod N) + t / N), from long-double sines --
    // computed once per axis for a whole batch of wavenumbers
    static void k0_phase(int N, std::complex<double>* w)
    {
        for (int t = 0; t < 2 * N; ++t) {
            const int k = 2 * (t % N) + t / N;
            const long double th = (long double)kPi * k / N;
            const double s = (double)sinl(th), h = (double)sinl(th / 2);
            w[t] = std::complex<double>(2.0 * h * h, s);
        }
    }
    void k0_scale(int N, double da, double kk, const std::complex<double>* w, T* v, T* ua) const
    {
        const double NP = 8.0 * (double)g.V;
        const std::complex<double> J(0, 1);
        for (int t = 0; t < 2 * N; ++t) {
            const std::complex<double> vv = J * kZ0 / (kk * da * NP) * w[t];
            v[t] = hc<R>(vv);
            ua[t] = hc<R>(std::conj(vv));
        }
    }
    // D_f = j kk Z0 delta_f / NP (forward; the adjoint's is its conjugate)
    T d_of(int a, double kk) const
    {
        const std::complex<double> D = std::complex<double>(0, 1) * kk * kZ0 * d[a] / (8.0 * (double)g.V);
        return hc<R>(D);
    }

    // Per-instance operator tables for cap instances on the current windows.
    void batch_free()
    {
        dfree(Gb);
        for (int a = 0; a < 3; ++a) {
            dfree(Vb[a]);
            dfree(Uab[a]);
        }
        dfree(Db);
        dfree(Dab);
        dfree(dgb);
        nbt = 0;
    }
    void batch_alloc(int cap)
    {
        const int l[3] = {nxs ? nxs : n[0], nys ? nys : n[1], nzs ? nzs : n[2]};
        if (!(cap <= nbt && l[0] == bl[0] && l[1] == bl[1] && l[2] == bl[2])) {
            batch_free();
            for (int a = 0; a < 3; ++a) bl[a] = l[a];
            lb = std::max(l[0], std::max(l[1], l[2]));
            gob = (long)(l[0] + 1) * (l[1] + 1) * (l[2] + 1);
            Gb = dalloc<T>(cap * gob);
            for (int a = 0; a < 3; ++a) {
                Vb[a] = dalloc<T>(2L * lb * cap);
                Uab[a] = dalloc<T>(2L * lb * cap);
            }
            Db = dalloc<T>(3L * cap);
            Dab = dalloc<T>(3L * cap);
            dgb = dalloc<T>(3 * g.V * cap);
            nbt = cap;
        }
        // The windowed products' slots of the shared (k0-free) tables, taken
        // again on every call although the per-instance buffers above are
        // kept: set_short rebuilds the short-period tables (every re-tune
        // does, also when it ends on the lengths it started from), so
        // addresses taken at an earlier call may be freed ones.
        const int sl[3] = {nxs ? 3 : 0, nys ? 4 : 1, nzs ? 5 : 2};
        for (int a = 0; a < 3; ++a) {
            mfb.U[a] = U[sl[a]];
            mfb.Vv[a] = Vb[a];
            mfb.fold[a] = fold[sl[a]];
            mab.U[a] = Uab[a];
            mab.Vv[a] = Va[sl[a]];
            mab.fold[a] = fold[sl[a]];
        }
        mfb.G = mab.G = Gb;
        mfb.gconj = 0;
        mab.gconj = 1;
        mfb.gstr = mab.gstr = gob;
        mfb.vstr = mab.ustr = 2 * lb;
        mfb.ustr = mab.vstr = 0;
        mfb.Dd = Db;
        mab.Dd = Dab;
    }
    // instance i's k0 tables (host part: Vv, adjoint U, D)
    // the k0 tables of instances [i0, i0 + cnt): built on the host for all of
    // them, one upload per table
    void batch_k0(int i0, int cnt, const double* ks, hipStream_t st)
    {
        std::vector<T> v(2L * lb * cnt), ua(2L * lb * cnt), df(3L * cnt), da(3L * cnt);
        std::vector<std::complex<double>> w(2L * lb);
        for (int a = 0; a < 3; ++a) {
            k0_phase(bl[a], w.data());
            for (int i = 0; i < cnt; ++i) k0_scale(bl[a], d[a], ks[i], w.data(), &v[2L * lb * i], &ua[2L * lb * i]);
            HIPCHK(hipMemcpyAsync(Vb[a] + 2L * lb * i0, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, st));
            HIPCHK(hipMemcpyAsync(Uab[a] + 2L * lb * i0, ua.data(), ua.size() * sizeof(T), hipMemcpyHostToDevice, st));
            HIPCHK(hipStreamSynchronize(st));  // v and ua are refilled for the next axis
        }
        for (int i = 0; i < cnt; ++i)
            for (int a = 0; a < 3; ++a) {
                df[3 * i + a] = d_of(a, ks[i]);
                da[3 * i + a] = df[3 * i + a];
                da[3 * i + a].y = -da[3 * i + a].y;
            }
        HIPCHK(hipMemcpyAsync(Db + 3L * i0, df.data(), df.size() * sizeof(T), hipMemcpyHostToDevice, st));
        HIPCHK(hipMemcpyAsync(Dab + 3L * i0, da.data(), da.size() * sizeof(T), hipMemcpyHostToDevice, st));
        HIPCHK(hipStreamSynchronize(st));
    }

    // The wavenumber enters the tables through v_s and D_f = j k0 Z0
    // delta_f / NP (vie_fft/fft.h:76-100, 256-281).
    void set_k0(double kk, hipStream_t st)
    {
        k0 = kk;
        for (int a = 0; a < 3; ++a) {
            axis_k0(a, n[a], d[a], st);
            const T D = d_of(a, k0);
            mf.D[a][0] = D.x;
            mf.D[a][1] = D.y;
            ma.D[a][0] = D.x;
            ma.D[a][1] = -D.y;
        }
        if (nxs) axis_k0(3, nxs, d[0], st);
        if (nys) axis_k0(4, nys, d[1], st);
        if (nzs) axis_k0(5, nzs, d[2], st);
        short_tabs();
    }
    ~Op()
    {
        batch_free();
        for (int a = 0; a < 6; ++a) free_axis(a);
        dfree(G);
        dfree(Gs);
        dfree(Gw);
        dfree(dg);
        dfree(pm);
        dfree(flags);
        dfree(zr);
        if (own_w) {
            dfree(W1);
            dfree(W2);
        }
    }

    // y = A z (or A^H z) where z is the Lin prologue; src is the vector the
    // diagonal term multiplies (z itself).  Returns the partial-sum count.
    // With row ranges (lin.zr: the inputs vanish on inactive faces) the passes
    // cover only the x window; outside it the output and lin.out are zero,
    // written here unless the caller keeps them zero (wskip), in which case
    // the z rows and y entries outside the y window are skipped too.
    long apply(const Lin<R>& lin, const T* src, T* y, Dots<R> dots, bool adj, const int* done,
               hipStream_t st, Timer* tm, bool mask = true)
    {
        long nblk = 0;
        Geo gw = g;
        gw.var = var;
        if (lin.zr) {
            gw.x0 = wx0;
            gw.x1 = wx1;
            if (lin.wskip) {
                gw.y0 = wy0;
                gw.y1 = wy1;
            }
        }
        // batched products are the solvers' (windowed, outputs kept zero)
        const bool batched = bt && (bt->nb > 1 || bt->ilist);
        if (batched && !(lin.zr && lin.wskip)) throw Err(VIE_ERR_ARG, "internal: unwindowed product in a batch");
        if (lin.scale && (adj || batched || !mask))
            throw Err(VIE_ERR_ARG, "internal: preconditioned product must be a forward, masked single product");
        if (batched) gw.bt = *bt;
        const bool multi = batched && bmulti;
        if (multi && (nbt == 0 || bl[0] != (nxs ? nxs : n[0]) || bl[1] != (nys ? nys : n[1]) ||
                      bl[2] != (nzs ? nzs : n[2])))
            throw Err(VIE_ERR_ARG, "the batch's operators were set for other windows: set them again");
        if ((gw.x0 > 0 || gw.x1 < g.nx) && !lin.wskip) {
            const long pl = (long)g.ny * g.nz;
            const unsigned gz = grid1(3 * (g.V - (gw.x1 - gw.x0) * pl));
            if (lin.out)
                hipLaunchKernelGGL(k_zero_outside<R>, dim3(gz), dim3(NT), 0, st, g.V, gw.x0 * pl, gw.x1 * pl,
                                   lin.out, done);
            hipLaunchKernelGGL(k_zero_outside<R>, dim3(gz), dim3(NT), 0, st, g.V, gw.x0 * pl, gw.x1 * pl, y, done);
            HIPCHK(hipGetLastError());
        }
        const int off = sizeof(R) == 8 ? 0 : 10;  // timing classes 1-5 FP64, 11-15 FP32
        auto t0 = [&](int c) { if (tm && tm->cls == c + off) tm->mark(st); };
        t0(1);
        // the y (z) window's pencils on the short y (z) period
        const bool ysh = lin.zr && lin.wskip && nys, zsh = lin.zr && lin.wskip && nzs;
        const int ay = ysh ? 4 : 1, az = zsh ? 5 : 2;
        if (ysh) {
            gw.nyp = nys;
            gw.ys = wy0;
        }
        if (zsh) {
            gw.nzp = nzs;
            gw.zs = wz0;
        }
        HIPCHK(run_zfwd<R>(gw, lin, W1, tw[az], twist[az], twP[az], done, st));
        t0(1);
        t0(2);
        HIPCHK(run_yfwd<R>(gw, W1, W2, tw[ay], twist[ay], twP[ay], done, st));
        t0(2);
        Geo gx = gw;
        int ax = 0, nxp = g.nx;
        if (lin.zr && nxs) {  // the window's pencils on the short period
            gx.xs = wx0;
            ax = 3;
            nxp = nxs;
        }
        const MixTabs<R>& mt = multi ? (adj ? mab : mfb)
                               : (ysh || zsh) ? (adj ? may : mfy) : ax ? (adj ? mas : mfs) : (adj ? ma : mf);
        t0(3);
        HIPCHK(run_xmix<R>(gx, nxp, W2, tw[ax], twist[ax], twP[ax], mt, done, st));
        t0(3);
        t0(4);
        HIPCHK(run_yinv<R>(gw, W2, W1, tw[ay], twist[ay], twP[ay], done, st));
        t0(4);
        Diag<R> dd{};
        dd.d = multi ? dgb : dg;
        dd.dstr = multi ? 3 * g.V : 0;
        dd.conj = adj ? 1 : 0;
        dd.skip = lin.wskip;
        dd.nomask = mask ? 0 : 1;
        dd.unit = lin.scale ? 1 : 0;  // lin.scale is this operator's pm: d (.) m = 1
        t0(5);
        HIPCHK(run_zinv<R>(gw, W1, src, y, dd, flags, dots, tw[az], twist[az], twP[az], done, st, &nblk));
        t0(5);
        return nblk;
    }
};

template <class R> Lin<R> lin1(const cx<R>* x)
{
    Lin<R> l{};
    l.x[0] = x;
    return l;
}

// Lin over Krylov vectors (zero on inactive faces: the masked right-hand side
// and the masked products keep them so), which the prologue then reads only
// within each z-row's active range.
template <class R> Lin<R> lin_k(const uint32_t* zr, const cx<R>* x)
{
    Lin<R> l = lin1<R>(x);
    l.zr = zr;
    return l;
}

// The same for a solver that zeroed its vectors at the start and writes only
// active faces from then on: the product leaves inactive output entries
// unwritten and uses the y window and short periods too.
template <class R> Lin<R> lin_kw(const uint32_t* zr, const cx<R>* x)
{
    Lin<R> l = lin_k<R>(zr, x);
    l.wskip = 1;
    return l;
}

template <class R> Dots<R> nodots()
{
    Dots<R> d{};
    return d;
}

}  // namespace

// ---------------------------------------------------------------------- ctx
struct vie_ctx {
    int nu[3] = {0, 0, 0};  // grid as the caller sees it (desc.n: the compiled grid it is embedded in)
    long Vu = 0;
    bool emb = false;         // the caller's grid is embedded in a larger compiled one
    // Internal axis order: internal axis a is the caller's axis perm[a]
    // (desc.n and desc.d are in internal order); set_medium orders the axes
    // by the fraction of their planes holding an active face, narrowest
    // first, so the passes' x window is the narrowest and the y window next.
    int perm[3] = {0, 1, 2};
    int nr[3] = {0, 0, 0};     // compiled grid in the caller's axis order
    double dr[3] = {0, 0, 0};  // cell sizes in the caller's axis order
    bool remap = false;        // caller layout != internal layout (embedding or axis order)
    bool windows = true;       // x/y windows and the axis order that favours them (VIE_WINDOW=0: off)
    bool no_graph = false;     // a round could not be captured as a graph: direct launches
    vie_desc desc{};
    hipStream_t stream = nullptr;
    std::string err;
    long V = 0;
    Op<double>* op64 = nullptr;
    Op<float>* op32 = nullptr;
    double2* Goff = nullptr;       // FP64 octant of G~ (parity order)
    double2* Gshort = nullptr;     // FP64 octant of the short-period x pass (sized as Goff)
    double2* gwork = nullptr;      // second octant buffer of green_compute
    double* gdct = nullptr;        // cosine tables of green_compute
    bool green_ready = false;
    bool medium_ready = false;
    std::vector<double> eps_h;     // per-voxel eps, 2V doubles
    std::vector<uint8_t> flags_h;  // 3V
    // the last medium's caller-layout material ids (empty for per-voxel eps)
    // and flags: a medium with the same ones only needs its eps and d_f rebuilt
    std::vector<uint8_t> mat_in, flags_in;
    int mat_max = 0;
    double2* eps_dev = nullptr;
    uint8_t* mat_dev = nullptr;
    // solver storage
    KState* state = nullptr;
    KState* aux = nullptr;
    double* hist = nullptr;
    double2* partial = nullptr;
    long partial_cap = 0;
    double2* b64 = nullptr;        // resident, masked right-hand side
    double2* x64 = nullptr;        // resident solution
    bool rhs_ready = false;
    bool sol_ready = false;
    std::vector<void*> vec64, vec32;
    double2* t64 = nullptr;        // FP64 scratch (residual / product)
    double2* r64 = nullptr;
    Timer timer;
    int timed_launches = 0;
    double timed_ms = 0.0;
    int ahead = 4;                 // iterations enqueued per host check
    KState* snap = nullptr;        // two pinned state snapshots (one batch in flight while the other is read)
    hipEvent_t snap_ev[2] = {};
    // batches of instances (one geometry, several right-hand sides or
    // wavelengths): the vectors, states and pass buffers hold nbcap
    // instances; bn instances take part in the solve being run and the
    // batched launches serve the slots of `cur`
    int nbcap = 1;
    int bn = 1;
    bool in_batch = false;         // a batched solve is being run (its records are per instance)
    Bat cur{};
    int* ilist = nullptr;          // launch slot -> instance (device, nbcap)
    bool bmulti = false;           // instances have their own operators
    std::vector<double> bk0;       // their wavenumbers
    double* btol = nullptr;        // per-instance tolerance and limit (device)
    int* bmaxit = nullptr;
    std::vector<KState> bstate;    // per-instance states after a batched solve
    std::vector<char> brhs;        // instance has a right-hand side
    std::vector<char> bop;         // instance has its own operator (2: with its own eps table)
    std::vector<double> btab;      // their eps tables (bnmat complex each)
    int bnmat = 0;
    double2* beps = nullptr;       // per-instance permittivity (V complex, post-processing scratch)
    double2* btabd = nullptr;      // the eps tables on the device, 256 entries per instance
    long btabd_cap = 0;            // instances btabd holds
    bool guess = false;            // the next solve starts from the guess in the VW vector
    bool x_held = false;           // x64 holds a solution in the current layout (a guess for the next solve)
    double dvg = 0.0;              // divergence stop factor of the solvers (0: off)
    int precond = 0;               // vie_preconditioner of the BiCGSTAB solves (0: none, the reference's)
};

// Krylov state records of a batch are this many ints / complex values apart.
static_assert(sizeof(KState) % sizeof(double2) == 0, "KState stride");
constexpr int kStateInts = (int)(sizeof(KState) / sizeof(int));
constexpr int kStateCx = (int)(sizeof(KState) / sizeof(double2));

namespace {

int fail(vie_ctx* c, const Err& e)
{
    if (c) c->err = e.what();
    return e.code;
}

template <class R> Op<R>* op_of(vie_ctx* c);
template <> Op<double>* op_of<double>(vie_ctx* c) { return c->op64; }
template <> Op<float>* op_of<float>(vie_ctx* c) { return c->op32; }

template <class R> std::vector<void*>& vecs_of(vie_ctx* c);
template <> std::vector<void*>& vecs_of<double>(vie_ctx* c) { return c->vec64; }
template <> std::vector<void*>& vecs_of<float>(vie_ctx* c) { return c->vec32; }

// Vectors: index constants into the per-precision work-vector pool.
// VT/VT2: staging of bare products; VG/VNU: adjoint right-hand side and
// solution; VW/VWB: a warm start's guess and its correction's right-hand side.
enum { VB = 0, VX, VR, VP, VQ, VAP, VAQ, VRC, VQC, VPC, VT, VT2, VG, VNU, VW, VWB, NVEC };

// k-th solver work vector (3V entries) of precision R, allocated on demand
template <class R> cx<R>* vec(vie_ctx* c, int k)
{
    auto& v = vecs_of<R>(c);
    while ((int)v.size() <= k) v.push_back(nullptr);
    if (!v[k]) v[k] = dalloc<cx<R>>(3 * c->V * c->nbcap);
    return (cx<R>*)v[k];
}

void ensure_partial(vie_ctx* c, long need)
{
    if (need <= c->partial_cap) return;
    dfree(c->partial);
    c->partial = dalloc<double2>(need);
    c->partial_cap = need;
}

// n entries per instance, for every instance of the batch being solved
template <class R> void fill0(vie_ctx* c, cx<R>* x, long n)
{
    n *= c->bn;
    hipLaunchKernelGGL(k_fill_zero<R>, dim3(grid1(n, NT, 16384)), dim3(NT), 0, c->stream, n, x);
    HIPCHK(hipGetLastError());
}

// Every linear combination the solvers form is over vectors that vanish on
// inactive faces, so the kernel reads inputs and dot operands only there.
template <class R> long lincomb(vie_ctx* c, LinComb<R> L, const int* done)
{
    Op<R>* o = op_of<R>(c);
    if (o) L.mask = o->flags;
    // outputs kept zero on inactive faces (wskip): run over the x window only
    const long pl = o ? (long)o->g.ny * o->g.nz : 0;
    if (o && L.wskip && L.n == 3 * c->V && (o->wx0 > 0 || o->wx1 < o->g.nx) && pl % 2 == 0) {
        L.wV = c->V;
        L.wlo = o->wx0 * pl;
        L.whi = o->wx1 * pl;
        // narrower y or z windows too (compact media): the box of the three
        // windows only; its z range starts even and holds whole pairs
        const int z1 = std::min(o->g.nz, (o->wz1 + 1) & ~1);
        if ((o->wy0 > 0 || o->wy1 < o->g.ny || o->wz0 > 0 || z1 < o->g.nz) && o->g.nz % 2 == 0) {
            L.bzp = (z1 - o->wz0) / 2;
            L.bx0 = o->wx0;
            L.bwx = o->wx1 - o->wx0;
            L.by0 = o->wy0;
            L.bwy = o->wy1 - o->wy0;
            L.bz0 = o->wz0;
            L.bny = o->g.ny;
            L.bnz = o->g.nz;
            L.blg = 0;
            while ((1 << L.blg) < L.bzp && (1 << L.blg) < NT) ++L.blg;
        }
    }
    // windowed: at most 1024 workgroups per component.  The FP32 BiCGSTAB
    // update over the 48-plane window at 128^3 takes 28.0 us per launch; 768
    // workgroups (two whole grid-stride rounds) 27.7, 1536 29.4, 2048 32.4 us.
    const long work = L.bzp ? (long)L.bwx * L.bwy * (1L << L.blg) : L.wV ? (L.whi - L.wlo) / 2 : (L.n + 1) / 2;
    const unsigned gsz = grid1(work, NT, L.wV ? 1024 : 2048);
    const unsigned gy = L.wV ? 3 : 1;
    L.dots.partial = c->partial;
    L.bt = c->cur;
    const int cls = sizeof(R) == 8 ? 6 : 16;
    if (c->timer.cls == cls) c->timer.mark(c->stream);
    VIE_LAUNCH(k_lincomb<R>, dim3(gsz, gy, c->cur.nb), dim3(NT), c->stream, L, done);
    HIPCHK(hipGetLastError());
    if (c->timer.cls == cls) c->timer.mark(c->stream);
    return (long)gsz * gy;
}

// One workgroup per launch slot of the current batch (one outside batches);
// the residual history is kept for single solves only.  mask (optional, one
// flag per instance state record): instances whose flag is set are skipped.
void scalar(vie_ctx* c, KState* s, int op, long nblk, int nd, const int* mask = nullptr)
{
    if (nd < 1 || nd > 4 || nblk <= 0 || nblk >= (1L << 27))
        throw Err(VIE_ERR_ARG, "internal: scalar step with " + std::to_string(nd) + " dots of " +
                                   std::to_string(nblk) + " partial sums");
    if (c->timer.cls == 7) c->timer.mark(c->stream);
    const bool single = c->cur.nb == 1 && !c->cur.ilist;
    VIE_LAUNCH(k_scalar, dim3(c->cur.nb), dim3(NTS), c->stream, op, s, c->partial, nblk, nd,
               single ? c->hist : nullptr, c->cur, mask);
    HIPCHK(hipGetLastError());
    if (c->timer.cls == 7) c->timer.mark(c->stream);
}

// pause > 0: the iterations also stop (done, not converged) once the
// recursive residual falls below it (reliable updates).  In a batch every
// instance's record is initialised; per_inst takes each instance's own
// tolerance and limit.
void state_init(vie_ctx* c, KState* s, double tol, int maxit, double pause = 0.0, bool per_inst = true)
{
    // the divergence stop applies to the solvers' own records (per_inst), not
    // to the residual bookkeeping records
    const double dvg = per_inst ? c->dvg : 0.0;
    if (c->in_batch) {
        hipLaunchKernelGGL(k_state_init_n, dim3((c->bn + 63) / 64), dim3(64), 0, c->stream, s, c->bn,
                           per_inst ? c->btol : nullptr, per_inst ? c->bmaxit : nullptr, tol, maxit, pause, dvg);
    } else {
        hipLaunchKernelGGL(k_state_init, dim3(1), dim3(1), 0, c->stream, s, tol, maxit, pause, dvg);
    }
    HIPCHK(hipGetLastError());
}

KState read_state(vie_ctx* c, KState* s)
{
    KState h;
    HIPCHK(hipMemcpyAsync(&h, s, sizeof(KState), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return h;
}

// the bn state records of a batch
void read_states(vie_ctx* c, KState* s, std::vector<KState>& out)
{
    out.resize(c->bn);
    HIPCHK(hipMemcpyAsync(out.data(), s, c->bn * sizeof(KState), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
}

// Enqueue a copy of the state into pinned snapshot `slot`, read later with
// wait_snap: the host checks one batch of iterations while the next runs.
// Snapshots hold the bn records of a batch (slot k at snap + k nbcap).
void snap_state(vie_ctx* c, const KState* s, int slot)
{
    if (!c->snap) {
        HIPCHK(hipHostMalloc((void**)&c->snap, 2 * c->nbcap * sizeof(KState), hipHostMallocDefault));
        for (hipEvent_t& e : c->snap_ev) HIPCHK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    }
    HIPCHK(hipMemcpyAsync(&c->snap[slot * c->nbcap], s, c->bn * sizeof(KState), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipEventRecord(c->snap_ev[slot], c->stream));
}
KState wait_snap(vie_ctx* c, int slot)
{
    HIPCHK(hipEventSynchronize(c->snap_ev[slot]));
    return c->snap[slot * c->nbcap];
}
const KState* snap_all(vie_ctx* c, int slot) { return &c->snap[slot * c->nbcap]; }

template <class R> Dots<R> dots_with(vie_ctx* c, int n, std::initializer_list<int> ab, const cx<R>* e1,
                                     const cx<R>* e2, std::initializer_list<int> nc = {})
{
    Dots<R> d{};
    d.n = n;
    int i = 0;
    for (int v : ab) {
        if (i % 2 == 0) d.a[i / 2] = v; else d.b[i / 2] = v;
        ++i;
    }
    i = 0;
    for (int v : nc) d.nc[i++] = v;
    d.e1 = e1;
    d.e2 = e2;
    d.partial = c->partial;
    d.masked = 1;  // solver dot products: every operand vanishes on inactive faces
    return d;
}

template <class R> Op<R>* make_op(vie_ctx* c);

void short_sync(vie_ctx* c);

// G~ octant of an (n[0], n[1], n[2]) grid with the context's cell sizes and
// k0, in parity order, into out ((n0+1)(n1+1)(n2+1) FP64 values); gwork is
// the other buffer of the ping-pong.
void green_octant(vie_ctx* c, const int* n, double2* out)
{
    const double* d = c->desc.d;
    const long L0 = n[0] + 1, L1 = n[1] + 1, L2 = n[2] + 1, tot = L0 * L1 * L2;
    const int* nf = c->desc.n;
    const long full = (long)(nf[0] + 1) * (nf[1] + 1) * (nf[2] + 1);
    if (!c->gwork) c->gwork = dalloc<double2>(full);
    const int nmax = std::max(nf[0], std::max(nf[1], nf[2]));
    if (!c->gdct) c->gdct = dalloc<double>(3 * 2 * nmax);
    std::vector<double> ctab(3 * 2 * nmax);
    for (int ax = 0; ax < 3; ++ax)
        for (int j = 0; j < 2 * n[ax]; ++j) ctab[ax * 2 * nmax + j] = (double)cosl((long double)kPi * j / n[ax]);
    copy_sync(c->gdct, ctab.data(), ctab.size() * sizeof(double), hipMemcpyHostToDevice, c->stream);
    double2* src = c->gwork;
    double2* dst = out;
    hipLaunchKernelGGL(k_green_offsets, dim3(grid1(tot, 256, 1L << 30)), dim3(256), 0, c->stream, n[0], n[1], n[2],
                       d[0], d[1], d[2], c->desc.k0, nullptr, src);
    HIPCHK(hipGetLastError());
    // x: gwork -> out, y: out -> gwork, z: gwork -> out
    for (int ax = 0; ax < 3; ++ax) {
        hipLaunchKernelGGL(k_green_cos, dim3(grid1(tot, 256, 1L << 30)), dim3(256), 0, c->stream, (int)L0, (int)L1,
                           (int)L2, ax, c->gdct + ax * 2 * nmax, src, dst);
        HIPCHK(hipGetLastError());
        std::swap(src, dst);
    }
}

void green_compute(vie_ctx* c)
{
    const int* n = c->desc.n;
    const long tot = (long)(n[0] + 1) * (n[1] + 1) * (n[2] + 1);
    if (!c->Goff) c->Goff = dalloc<double2>(tot);
    green_octant(c, n, c->Goff);
    c->green_ready = true;
    if (c->op64) HIPCHK(hipMemcpyAsync(c->op64->G, c->Goff, tot * sizeof(double2), hipMemcpyDeviceToDevice, c->stream));
    if (c->op32) {
        hipLaunchKernelGGL(k_convert<float>, dim3(grid1(tot, 256, 1L << 30)), dim3(256), 0, c->stream, tot,
                           c->Goff, c->op32->G);
        HIPCHK(hipGetLastError());
    }
    HIPCHK(hipStreamSynchronize(c->stream));
    short_sync(c);
}

template <class R> void upload_medium(vie_ctx* c, Op<R>* op)
{
    const int* n = c->desc.n;
    const double* d = c->desc.d;
    if (!op->dg) op->dg = dalloc<cx<R>>(3 * c->V);
    double cim[3];
    for (int f = 0; f < 3; ++f) cim[f] = -(d[f] / 2.0) * kZ0 / c->desc.k0;
    hipLaunchKernelGGL(k_diag<R>, dim3(grid1(c->V)), dim3(NT), 0, c->stream, n[0], n[1], n[2], cim[0], cim[1],
                       cim[2], c->eps_dev, op->dg);
    HIPCHK(hipGetLastError());
    ++op->dgen;
    copy_sync(op->flags, c->flags_h.data(), c->flags_h.size(), hipMemcpyHostToDevice, c->stream);
    // x, y and z windows: the planes holding an active face of any component
    // (all of them when windows are off or no face is active); the z window
    // starts on an even plane
    int lo = n[0], hi = 0, ylo = n[1], yhi = 0, zlo = n[2], zhi = 0;
    auto nz0 = [](uint8_t v) { return v != 0; };
    for (int f = 0; f < 3; ++f)
        for (int i = 0; i < n[0]; ++i)
            for (int j = 0; j < n[1]; ++j) {
                const uint8_t* p = c->flags_h.data() + f * c->V + ((long)i * n[1] + j) * n[2];
                const uint8_t* first = std::find_if(p, p + n[2], nz0);
                if (first == p + n[2]) continue;
                const uint8_t* last = std::find_if(std::make_reverse_iterator(p + n[2]),
                                                   std::make_reverse_iterator(p), nz0).base();
                lo = std::min(lo, i);
                hi = std::max(hi, i + 1);
                ylo = std::min(ylo, j);
                yhi = std::max(yhi, j + 1);
                zlo = std::min(zlo, (int)(first - p));
                zhi = std::max(zhi, (int)(last - p));
            }
    if (hi <= lo || !c->windows) {
        lo = ylo = zlo = 0;
        hi = n[0];
        yhi = n[1];
        zhi = n[2];
    }
    op->wx0 = lo;
    op->wx1 = hi;
    op->wy0 = ylo;
    op->wy1 = yhi;
    op->wz0 = zlo & ~1;
    op->wz1 = zhi;
    const long rows = 3L * n[0] * n[1];
    hipLaunchKernelGGL(k_zrange, dim3(grid1(rows)), dim3(NT), 0, c->stream, rows, n[2], op->flags, op->zr);
    HIPCHK(hipGetLastError());
}

// Length of the short-period x pass for a window of w planes: the
// shortest length >= w below Nx with a register-resident x pass in both
// precisions, else the shortest compiled one below Nx, else none (also with
// windows off).  At 128^3 with w = 48 (MI355X, per launch, FP32 / FP64):
// the 64-length register pass (60 / 105 us) beat the 48-length LDS-staged
// (83 / 153 us) and register (4-stage plan, 67 / 155 us) ones and the full
// 128 (84 / 196 us).  VIE_XSHORT=0 turns it off, VIE_XSHORT=m asks for
// length m.
int short_len(const vie_ctx* c, int w, int axis)
{
    const int nn = c->desc.n[axis];
    if (!c->windows || w >= nn) return 0;
    if (axis == 2 && nn % 2) return 0;  // z pairs must not straddle rows
    if (const char* e = std::getenv(axis == 0 ? "VIE_XSHORT" : axis == 1 ? "VIE_YSHORT" : "VIE_ZSHORT")) {
        const int m = std::atoi(e);
        return (m >= w && m < nn && size_supported(m)) ? m : 0;
    }
    // the z passes use the y passes' register configurations
    auto reg = [axis](int m) {
        return axis == 0 ? x_register<double>(m) && x_register<float>(m) : y_register<double>(m) && y_register<float>(m);
    };
    for (int m = w; m < nn; ++m)
        if (size_supported(m) && reg(m)) return m;
    for (int m = w; m < nn; ++m)
        if (size_supported(m)) return m;
    return 0;
}

// G~ octant of the (px, py, pz) grid, scaled by (Nx Ny Nz) / (px py pz),
// into the FP64 staging buffer Gshort and the operators' G64 / G32 copies.
void short_octant(vie_ctx* c, int px, int py, int pz, double2* g64, float2* g32)
{
    const int* n = c->desc.n;
    const int n3[3] = {px, py, pz};
    const long tot = (long)(px + 1) * (py + 1) * (pz + 1);
    if (!c->Gshort) c->Gshort = dalloc<double2>((long)(n[0] + 1) * (n[1] + 1) * (n[2] + 1));
    green_octant(c, n3, c->Gshort);
    hipLaunchKernelGGL(k_scale_c64, dim3(grid1(tot, 256, 1L << 30)), dim3(256), 0, c->stream, tot,
                       (double)n[0] * n[1] * n[2] / ((double)px * py * pz), c->Gshort);
    HIPCHK(hipGetLastError());
    if (g64) HIPCHK(hipMemcpyAsync(g64, c->Gshort, tot * sizeof(double2), hipMemcpyDeviceToDevice, c->stream));
    if (g32) {
        hipLaunchKernelGGL(k_convert<float>, dim3(grid1(tot, 256, 1L << 30)), dim3(256), 0, c->stream, tot,
                           c->Gshort, g32);
        HIPCHK(hipGetLastError());
    }
    HIPCHK(hipStreamSynchronize(c->stream));
}

// Short periods of an operator for the current windows (its tuned plan's,
// else the shortest lengths with register passes), with their G~ octants once
// G~ is computed.
template <class R> void short_set(vie_ctx* c, Op<R>* op, int ns, int nsy, int nsz)
{
    const int* n = c->desc.n;
    if (nsz) {  // partial sums of the z inverse at the short z length
        Geo gz{n[0], n[1], nsz, c->V};
        ensure_partial(c, 3L * c->nbcap * std::max<long>(std::max(zinv_blocks<double>(gz), zinv_blocks<float>(gz)), 4096));
    }
    op->set_short(ns, nsy, nsz, c->stream);
    if (!c->green_ready) return;
    double2* g64 = nullptr;
    float2* g32 = nullptr;
    if constexpr (sizeof(R) == 8) g64 = op->Gs; else g32 = op->Gs;
    if (ns) short_octant(c, ns, n[1], n[2], g64, g32);
    if (nsy || nsz) {
        g64 = nullptr;
        g32 = nullptr;
        if constexpr (sizeof(R) == 8) g64 = op->Gw; else g32 = op->Gw;
        short_octant(c, ns ? ns : n[0], nsy ? nsy : n[1], nsz ? nsz : n[2], g64, g32);
    }
}

template <class R> void short_plan(vie_ctx* c, Op<R>* op, int* ns, int* nsy, int* nsz)
{
    if (op->tuned) {
        *ns = op->tns;
        *nsy = op->tnys;
        *nsz = op->tnzs;
        return;
    }
    const bool m = c->medium_ready;
    *ns = m ? short_len(c, op->wx1 - op->wx0, 0) : 0;
    *nsy = m ? short_len(c, op->wy1 - op->wy0, 1) : 0;
    *nsz = m ? short_len(c, op->wz1 - op->wz0, 2) : 0;
}

void short_sync(vie_ctx* c)
{
    int a[3] = {0, 0, 0}, b[3] = {0, 0, 0};
    if (c->op64) short_plan<double>(c, c->op64, &a[0], &a[1], &a[2]);
    if (c->op32) short_plan<float>(c, c->op32, &b[0], &b[1], &b[2]);
    // one FP64 octant computation serves both operators when their lengths agree
    if (c->op64 && c->op32 && a[0] == b[0] && a[1] == b[1] && a[2] == b[2] && c->green_ready) {
        const int* n = c->desc.n;
        short_set<double>(c, c->op64, a[0], a[1], a[2]);
        c->op32->set_short(b[0], b[1], b[2], c->stream);
        if (b[0]) hipLaunchKernelGGL(k_convert<float>, dim3(grid1((long)(b[0] + 1) * (n[1] + 1) * (n[2] + 1), 256, 1L << 30)),
                                     dim3(256), 0, c->stream, (long)(b[0] + 1) * (n[1] + 1) * (n[2] + 1), c->op64->Gs,
                                     c->op32->Gs);
        if (b[1] || b[2]) {
            const long t = (long)((b[0] ? b[0] : n[0]) + 1) * ((b[1] ? b[1] : n[1]) + 1) * ((b[2] ? b[2] : n[2]) + 1);
            hipLaunchKernelGGL(k_convert<float>, dim3(grid1(t, 256, 1L << 30)), dim3(256), 0, c->stream, t,
                               c->op64->Gw, c->op32->Gw);
        }
        HIPCHK(hipGetLastError());
        HIPCHK(hipStreamSynchronize(c->stream));
        return;
    }
    if (c->op64) short_set<double>(c, c->op64, a[0], a[1], a[2]);
    if (c->op32) short_set<float>(c, c->op32, b[0], b[1], b[2]);
}

// ================================================================ autotuner
// Each operator's plan -- the variant of every pass (register-resident or
// LDS-staged x pass, y/z passes and z forward, Geo::var) and the short
// periods of its windowed products along x, y and z -- is chosen by timing
// the solvers' windowed product on the device for the candidates: the eight
// variant combinations, then, axis by axis, the full length and the compiled
// lengths from the window width up, then the variants again.  Plans depend
// on the grid, the precision and the three windows only (not on eps or k0),
// so they are kept per process for every geometry met (and in the file
// VIE_TUNE_CACHE names, if set).  VIE_TUNE=0 keeps the fixed defaults.
struct TunePlan {
    int var, ns, nys, nzs;
    double us, us_def;
};
static std::mutex g_tune_mu;
static std::map<std::string, TunePlan> g_tune;
static bool g_tune_loaded = false;

static bool tuning_on()
{
    const char* e = std::getenv("VIE_TUNE");
    return !(e && e[0] == '0');
}

template <class R> static std::string tune_key(const vie_ctx* c, const Op<R>* op)
{
    const int* n = c->desc.n;
    char b[160];
    std::snprintf(b, sizeof b, "%s %d %d %d %d %d %d %d %d %d", sizeof(R) == 8 ? "f64" : "f32", n[0], n[1], n[2],
                  op->wx0, op->wx1, op->wy0, op->wy1, op->wz0, op->wz1);
    return b;
}

static void tune_cache_load()
{
    if (g_tune_loaded) return;
    g_tune_loaded = true;
    const char* path = std::getenv("VIE_TUNE_CACHE");
    if (!path) return;
    FILE* f = std::fopen(path, "r");
    if (!f) return;
    char line[256];
    while (std::fgets(line, sizeof line, f)) {
        char pr[8];
        int v[10];
        TunePlan p;
        if (std::sscanf(line, "%7s %d %d %d %d %d %d %d %d %d | %d %d %d %d %lf %lf", pr, &v[0], &v[1], &v[2], &v[3],
                        &v[4], &v[5], &v[6], &v[7], &v[8], &p.var, &p.ns, &p.nys, &p.nzs, &p.us, &p.us_def) == 16) {
            char key[160];
            std::snprintf(key, sizeof key, "%s %d %d %d %d %d %d %d %d %d", pr, v[0], v[1], v[2], v[3], v[4], v[5],
                          v[6], v[7], v[8]);
            g_tune[key] = p;
        }
    }
    std::fclose(f);
}

static void tune_cache_store(const std::string& key, const TunePlan& p)
{
    const char* path = std::getenv("VIE_TUNE_CACHE");
    if (!path) return;
    FILE* f = std::fopen(path, "a");
    if (!f) return;
    std::fprintf(f, "%s | %d %d %d %d %.3f %.3f\n", key.c_str(), p.var, p.ns, p.nys, p.nzs, p.us, p.us_def);
    std::fclose(f);
}

// microseconds per windowed product (the solvers' kind) with the operator's
// current plan: the best of two rounds of `reps` products
template <class R> static double time_products(vie_ctx* c, Op<R>* op, int reps)
{
    const long n = 3 * c->V;
    cx<R>* x = vec<R>(c, VT);
    cx<R>* y = vec<R>(c, VT2);
    fill0<R>(c, x, n);
    fill0<R>(c, y, n);
    hipEvent_t e0, e1;
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    double best = 1e30;
    try {
        op->apply(lin_kw<R>(op->zr, x), x, y, nodots<R>(), false, nullptr, c->stream, nullptr);
        for (int round = 0; round < 2; ++round) {
            HIPCHK(hipEventRecord(e0, c->stream));
            for (int k = 0; k < reps; ++k)
                op->apply(lin_kw<R>(op->zr, x), x, y, nodots<R>(), false, nullptr, c->stream, nullptr);
            HIPCHK(hipEventRecord(e1, c->stream));
            HIPCHK(hipEventSynchronize(e1));
            float ms = 0;
            HIPCHK(hipEventElapsedTime(&ms, e0, e1));
            best = std::min(best, 1e3 * ms / reps);
        }
    } catch (...) {
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        throw;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return best;
}

// compiled lengths a window of w planes along an axis of N may use: the full
// length (0) and the supported lengths from w up to (but not including) N,
// the smallest three; z short periods need an even N
static std::vector<int> short_candidates(int w, int N, int axis)
{
    std::vector<int> v{0};
    if (w >= N || (axis == 2 && N % 2)) return v;
    for (int m = w; m < N && v.size() < 4; ++m)
        if (size_supported(m)) v.push_back(m);
    return v;
}

// A plan read from the cache file must be one the tuner could have picked
// for this operator: a known variant, and per axis the full length (0) or a
// compiled length that covers the window (a shorter period would change the
// operator) and is shorter than the axis.
template <class R> static bool plan_valid(const vie_ctx* c, const Op<R>* op, const TunePlan& p)
{
    if (p.var < -1 || p.var > 7) return false;
    const int w[3] = {op->wx1 - op->wx0, op->wy1 - op->wy0, op->wz1 - op->wz0};
    const int len[3] = {p.ns, p.nys, p.nzs};
    for (int a = 0; a < 3; ++a) {
        const int m = len[a], N = c->desc.n[a];
        if (m != 0 && !(m > 0 && size_supported(m) && m >= w[a] && m < N && (a != 2 || N % 2 == 0))) return false;
    }
    return true;
}

template <class R> void tune_op(vie_ctx* c, Op<R>* op)
{
    op->tuned = false;
    op->var = -1;
    if (!tuning_on() || !c->medium_ready) return;
    const std::string key = tune_key(c, op);
    {
        std::lock_guard<std::mutex> lk(g_tune_mu);
        tune_cache_load();
        auto it = g_tune.find(key);
        if (it != g_tune.end() && !plan_valid(c, op, it->second)) {
            g_tune.erase(it);  // a stale or damaged cache line: tune again
            it = g_tune.end();
        }
        if (it != g_tune.end()) {
            op->var = it->second.var;
            op->tns = it->second.ns;
            op->tnys = it->second.nys;
            op->tnzs = it->second.nzs;
            op->tus = it->second.us;
            op->tus_def = it->second.us_def;
            op->tuned = true;
            return;
        }
    }
    const bool green = c->green_ready;
    c->green_ready = false;  // timing does not need G~ values: skip the octants while trying
    try {
        // start from the defaults
        int ns = 0, nys = 0, nzs = 0;
        short_plan<R>(c, op, &ns, &nys, &nzs);
        op->set_short(ns, nys, nzs, c->stream);
        // reps: about 2 ms of products per round
        op->var = -1;
        const double t0 = time_products<R>(c, op, 2);
        const int reps = std::max(2, std::min(20, (int)(2000.0 / std::max(t0, 1.0))));
        op->tus_def = time_products<R>(c, op, reps);  // the fixed defaults, for comparison
        // two passes over the variants, each variant's best time of both: a
        // slowdown of the device lasting a few milliseconds (clocks settling
        // after heavy work) then cannot hit every timing of one variant
        double tv[8];
        auto time_vars = [&]() {
            for (int v = 0; v < 8; ++v) tv[v] = 1e30;
            for (int pass = 0; pass < 2; ++pass)
                for (int v = 0; v < 8; ++v) {
                    op->var = v;
                    tv[v] = std::min(tv[v], time_products<R>(c, op, reps));
                }
        };
        auto best_var = [&](double* tbest) {
            time_vars();
            int bv = 0;
            for (int v = 1; v < 8; ++v)
                if (tv[v] < tv[bv]) bv = v;
            op->var = bv;
            *tbest = tv[bv];
        };
        double tb = 0;
        best_var(&tb);
        // short periods, one axis at a time.  Variants and periods interact (a
        // variant that ties for best at one period can be far slower at
        // another), so each candidate period is timed with the three variants
        // that were fastest at the default periods
        int rank[8];
        for (int v = 0; v < 8; ++v) rank[v] = v;
        std::sort(rank, rank + 8, [&](int a, int b) { return tv[a] < tv[b]; });
        int bvar = op->var;
        const int w[3] = {op->wx1 - op->wx0, op->wy1 - op->wy0, op->wz1 - op->wz0};
        int len[3] = {ns, nys, nzs};
        for (int a = 0; a < 3; ++a) {
            int bl = len[a];
            for (int m : short_candidates(w[a], c->desc.n[a], a)) {
                if (m == len[a]) continue;
                int l2[3] = {len[0], len[1], len[2]};
                l2[a] = m;
                if (l2[2]) {
                    Geo gz{c->desc.n[0], c->desc.n[1], l2[2], c->V};
                    ensure_partial(c, 3L * c->nbcap *
                                          std::max<long>(std::max(zinv_blocks<double>(gz), zinv_blocks<float>(gz)), 4096));
                }
                op->set_short(l2[0], l2[1], l2[2], c->stream);
                for (int j = 0; j < 3; ++j) {
                    op->var = rank[j];
                    const double t = time_products<R>(c, op, reps);
                    if (t < tb) {
                        tb = t;
                        bl = m;
                        bvar = rank[j];
                    }
                }
            }
            len[a] = bl;
            op->set_short(len[0], len[1], len[2], c->stream);
        }
        op->var = bvar;
        best_var(&tb);
        // the pick against the fixed defaults, timed alternately three times
        // (best of each): the defaults come back when the pick is clearly
        // slower (a choice made under a transient slowdown); within 2% the pick
        // stays, so timing noise does not switch plans (and rounding) between runs
        const int pv = op->var;
        double tp = 1e30, td = 1e30;
        for (int round = 0; round < 3; ++round) {
            op->var = pv;
            op->set_short(len[0], len[1], len[2], c->stream);
            tp = std::min(tp, time_products<R>(c, op, reps));
            op->var = -1;
            op->set_short(ns, nys, nzs, c->stream);
            td = std::min(td, time_products<R>(c, op, reps));
        }
        const bool keep = tp <= 1.02 * td;
        if (keep) {
            op->var = pv;
            op->set_short(len[0], len[1], len[2], c->stream);
        } else {
            len[0] = ns;
            len[1] = nys;
            len[2] = nzs;
        }
        op->tns = len[0];
        op->tnys = len[1];
        op->tnzs = len[2];
        op->tus = keep ? tp : td;
        op->tus_def = td;
        op->tuned = true;
    } catch (...) {
        c->green_ready = green;
        throw;
    }
    c->green_ready = green;
    TunePlan p{op->var, op->tns, op->tnys, op->tnzs, op->tus, op->tus_def};
    std::lock_guard<std::mutex> lk(g_tune_mu);
    g_tune[key] = p;
    tune_cache_store(key, p);
}

template <> Op<double>* make_op<double>(vie_ctx* c)
{
    if (!c->op64) {
        double2* w1 = c->op32 ? (double2*)c->op32->W1 : nullptr;
        double2* w2 = c->op32 ? (double2*)c->op32->W2 : nullptr;
        // an FP32 operator's work buffers are too small to share; allocate fresh
        (void)w1;
        (void)w2;
        c->op64 = new Op<double>(c->desc.n, c->desc.d, c->desc.k0, nullptr, nullptr, c->stream, c->nbcap);
        c->op64->bt = &c->cur;
        if (c->green_ready)
            copy_sync(c->op64->G, c->Goff,
                      (long)(c->desc.n[0] + 1) * (c->desc.n[1] + 1) * (c->desc.n[2] + 1) * sizeof(double2),
                      hipMemcpyDeviceToDevice, c->stream);
        if (c->medium_ready) {
            upload_medium<double>(c, c->op64);
            tune_op<double>(c, c->op64);
            short_sync(c);
        }
    }
    return c->op64;
}

template <> Op<float>* make_op<float>(vie_ctx* c)
{
    if (!c->op32) {
        // the FP32 pass buffers fit in the FP64 ones: share them
        float2* w1 = c->op64 ? (float2*)c->op64->W1 : nullptr;
        float2* w2 = c->op64 ? (float2*)c->op64->W2 : nullptr;
        c->op32 = new Op<float>(c->desc.n, c->desc.d, c->desc.k0, w1, w2, c->stream, c->nbcap);
        c->op32->bt = &c->cur;
        const long tot = (long)(c->desc.n[0] + 1) * (c->desc.n[1] + 1) * (c->desc.n[2] + 1);
        if (c->green_ready) {
            hipLaunchKernelGGL(k_convert<float>, dim3(grid1(tot, 256, 1L << 30)), dim3(256), 0, c->stream, tot,
                               c->Goff, c->op32->G);
            HIPCHK(hipGetLastError());
            HIPCHK(hipStreamSynchronize(c->stream));
        }
        if (c->medium_ready) {
            upload_medium<float>(c, c->op32);
            tune_op<float>(c, c->op32);
            short_sync(c);
        }
    }
    return c->op32;
}

void require(vie_ctx* c, bool ok, const char* what)
{
    if (!ok) throw Err(VIE_ERR_ARG, what);
    (void)c;
}

// ---------------------------------------------------------------- solvers
// Drives iterations until the device flags done; `iter` enqueues one.  Two
// batches of c->ahead iterations are in flight: the snapshot taken after one
// is read while the next runs, so the device never waits for the host (the
// iterations enqueued after convergence exit at once).
template <class F> KState drive_batch(vie_ctx* c, KState* st, F iter);

// A round of c->ahead iterations, enqueued as one HIP graph: captured at the
// first round and replayed after (the iterations' kernel arguments -- buffers,
// the state's coefficient and done addresses, launch slots -- stay the same;
// every scalar the kernels use lives in device memory).  A new capture is
// made when the launch slots change (batches).  Off while kernels are being
// timed (host events) and with VIE_GRAPH=0.
struct IterGraph {
    hipGraphExec_t exec = nullptr;
    Bat slots{};
    ~IterGraph()
    {
        if (exec) (void)hipGraphExecDestroy(exec);
    }
};

static bool graphs_on()
{
    static const bool v = [] {
        const char* e = std::getenv("VIE_GRAPH");
        return !(e && e[0] == '0');
    }();
    return v;
}

// If a capture cannot be made (the runtime refuses a call the round makes
// while capturing, or the graph does not instantiate), the context runs its
// rounds as direct launches from then on; nothing captured has executed, so
// the round is simply enqueued again.
template <class F> void round_of(vie_ctx* c, IterGraph& g, F& iter)
{
    auto direct = [&] {
        for (int k = 0; k < c->ahead; ++k) iter();
    };
    if (!graphs_on() || c->timer.cls || c->no_graph) return direct();
    if (g.exec && (g.slots.nb != c->cur.nb || g.slots.ilist != c->cur.ilist)) {
        (void)hipGraphExecDestroy(g.exec);
        g.exec = nullptr;
    }
    if (!g.exec) {
        bool ok = hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal) == hipSuccess;
        if (ok) {
            hipGraph_t graph = nullptr;
            try {
                direct();
            } catch (...) {
                ok = false;
            }
            ok = hipStreamEndCapture(c->stream, &graph) == hipSuccess && ok && graph;
            ok = ok && hipGraphInstantiate(&g.exec, graph, nullptr, nullptr, 0) == hipSuccess;
            if (graph) (void)hipGraphDestroy(graph);
        }
        if (!ok) {
            if (g.exec) (void)hipGraphExecDestroy(g.exec);
            g.exec = nullptr;
            (void)hipGetLastError();  // the refused call's error, not the solve's
            c->no_graph = true;
            return direct();
        }
        g.slots = c->cur;
    }
    HIPCHK(hipGraphLaunch(g.exec, c->stream));
}

template <class F> KState drive(vie_ctx* c, KState* st, F iter)
{
    if (c->in_batch) return drive_batch(c, st, iter);
    IterGraph g;
    int slot = 0;
    round_of(c, g, iter);
    snap_state(c, st, slot);
    for (;;) {
        round_of(c, g, iter);
        snap_state(c, st, slot ^ 1);
        const KState h = wait_snap(c, slot);
        slot ^= 1;
        if (h.done) return read_state(c, st);
    }
}

// Launch slots of a batch: all bn instances until the snapshots show that a
// quarter of them have stopped, then the first m slots of a list that puts
// the instances still iterating first (k_ilist, rebuilt on the device at the
// start of every round, so it also places instances that resumed), m the
// number of instances that may still iterate at the last snapshot.
void batch_slots(vie_ctx* c, KState* st, int may_run, int* nact)
{
    if (c->cur.ilist || may_run <= *nact * 3 / 4) {
        hipLaunchKernelGGL(k_ilist, dim3(1), dim3(NTS), 0, c->stream, st, c->bn, c->ilist);
        HIPCHK(hipGetLastError());
        c->cur.ilist = c->ilist;
        c->cur.nb = std::max(1, may_run);
        *nact = c->cur.nb;
    }
}
// every instance of the batch in slot order (no list)
void batch_all(vie_ctx* c)
{
    c->cur.nb = c->bn;
    c->cur.ilist = nullptr;
}

// drive() for a batch: the same two rounds in flight; the solve ends when
// every instance's record is done.  c->bstate receives the bn final records.
template <class F> KState drive_batch(vie_ctx* c, KState* st, F iter)
{
    const int nb = c->bn;
    int nact = nb, slot = 0;
    IterGraph g;
    batch_all(c);
    round_of(c, g, iter);
    snap_state(c, st, slot);
    int live = nb;
    for (;;) {
        batch_slots(c, st, live, &nact);
        round_of(c, g, iter);
        snap_state(c, st, slot ^ 1);
        wait_snap(c, slot);
        const KState* h = snap_all(c, slot);
        slot ^= 1;
        live = 0;
        for (int i = 0; i < nb; ++i) live += h[i].done ? 0 : 1;
        if (live == 0) break;
    }
    batch_all(c);
    read_states(c, st, c->bstate);
    return c->bstate[0];
}

// The Jacobi preconditioner of op (m = 1/d_f on active faces), rebuilt when
// d_f has changed since it was built.  Right preconditioning: the solvers
// iterate on A (m (.) y) = b, whose residual is b - A x for x = m (.) y, so the
// stop rule reads the true residual as without it; with m = 1/d the
// preconditioned product has a unit diagonal term (Diag::unit), so its z
// inverse pass reads x instead of d and x, and its z forward pass reads m.
// An active face whose d_f vanishes has no Jacobi scale: VIE_ERR_ARG.
template <class R> const cx<R>* jacobi_of(vie_ctx* c, Op<R>* op)
{
    if (op->pmgen == op->dgen) return op->pm;
    const long n = 3 * c->V;
    if (!op->pm) op->pm = dalloc<cx<R>>(n);
    int* bad = dalloc<int>(1);
    int hbad = 0;
    try {
        HIPCHK(hipMemsetAsync(bad, 0, sizeof(int), c->stream));
        hipLaunchKernelGGL(k_jacobi<R>, dim3(grid1(n)), dim3(NT), 0, c->stream, n, op->dg, op->flags, op->pm, bad);
        HIPCHK(hipGetLastError());
        copy_sync(&hbad, bad, sizeof(int), hipMemcpyDeviceToHost, c->stream);
    } catch (...) {
        dfree(bad);
        throw;
    }
    dfree(bad);
    if (hbad) throw Err(VIE_ERR_ARG, "Jacobi preconditioner: the diagonal term d_f vanishes on " +
                                         std::to_string(hbad) + " active faces");
    op->pmgen = op->dgen;
    return op->pm;
}

// BiCGSTAB (vie_fft/CG_Iteration.h:531-776) on b -> x, both of precision R;
// with c->precond the products are A (m (.) y) and x = m (.) y at the end.
template <class R> KState solve_bicgstab(vie_ctx* c, Op<R>* op, const cx<R>* b, cx<R>* x, double tol, int maxit)
{
    KState* st = c->state;
    const long n = 3 * c->V;
    cx<R>*r = vec<R>(c, VR), *p = vec<R>(c, VP), *q = vec<R>(c, VQ), *Ap = vec<R>(c, VAP), *Aq = vec<R>(c, VAQ);
    const int* done = &st->done;
    if (c->precond && c->in_batch) throw Err(VIE_ERR_ARG, "the Jacobi preconditioner applies to single solves");
    const cx<R>* pm = c->precond ? jacobi_of<R>(c, op) : nullptr;
    state_init(c, st, tol, maxit);
    // p, q and their products hold zero on inactive faces from here on: the
    // iterations leave those entries unwritten (wskip)
    fill0<R>(c, p, n);
    fill0<R>(c, Ap, n);
    fill0<R>(c, q, n);
    fill0<R>(c, Aq, n);
    {   // x = 0 ; r = b ; b^H b
        LinComb<R> L{};
        L.n = n;
        L.x[0] = b;
        L.hc2[0][0] = 1.0;
        L.out = x;
        L.out2 = r;
        L.dots = dots_with<R>(c, 1, {1, 1}, b, nullptr);
        long nb = lincomb<R>(c, L, done);
        scalar(c, st, SC_BICGSTAB_INIT, nb, 1);
    }
    auto iter = [&]() {
        Lin<R> lp{};
        lp.zr = op->zr;
        lp.wskip = 1;
        lp.x[0] = r; lp.x[1] = p; lp.x[2] = Ap;
        lp.coef = &st->cf[CF_BS_P];
        lp.out = p;
        lp.scale = pm;
        long nb = op->apply(lp, p, Ap, dots_with<R>(c, 1, {1, 0}, b, nullptr), false, done, c->stream, &c->timer);
        scalar(c, st, SC_BICGSTAB_ALPHA, nb, 1);
        Lin<R> lq{};
        lq.zr = op->zr;
        lq.wskip = 1;
        lq.x[0] = r; lq.x[1] = Ap;
        lq.coef = &st->cf[CF_BS_Q];
        lq.out = q;
        lq.scale = pm;
        nb = op->apply(lq, q, Aq, dots_with<R>(c, 2, {0, 1, 0, 0}, q, nullptr), false, done, c->stream, &c->timer);
        scalar(c, st, SC_BICGSTAB_OMEGA, nb, 2);
        LinComb<R> L{};
        L.n = n;
        L.x[0] = x; L.x[1] = p; L.x[2] = q; L.x[3] = Aq;
        L.coef = &st->cf[CF_BS_U];
        L.out = x;
        L.out2 = r;
        L.wskip = 1;
        L.dots = dots_with<R>(c, 2, {3, 3, 1, 3}, b, nullptr);
        nb = lincomb<R>(c, L, done);
        scalar(c, st, SC_BICGSTAB_END, nb, 2);
    };
    const KState h = drive(c, st, iter);
    if (pm) {  // x = m (.) y
        hipLaunchKernelGGL(k_cmul_diag<R>, dim3(grid1(n)), dim3(NT), 0, c->stream, n, pm, x);
        HIPCHK(hipGetLastError());
    }
    return h;
}

// CG on the normal equations (vie_fft/CG_Iteration.h:176-357).
template <class R> KState solve_cg(vie_ctx* c, Op<R>* op, const cx<R>* b, cx<R>* x, double tol, int maxit)
{
    KState* st = c->state;
    const long n = 3 * c->V;
    cx<R>*r = vec<R>(c, VR), *p = vec<R>(c, VP), *q = vec<R>(c, VQ), *Ap = vec<R>(c, VAP);
    const int* done = &st->done;
    state_init(c, st, tol, maxit);
    // every vector starts at zero and is written only on active faces from here
    // on (wskip): products and updates skip the rest and run over the windows
    fill0<R>(c, x, n);
    fill0<R>(c, r, n);
    fill0<R>(c, p, n);
    fill0<R>(c, q, n);
    fill0<R>(c, Ap, n);
    {   // x = 0 ; r = -b
        LinComb<R> L{};
        L.n = n;
        L.wskip = 1;
        L.x[0] = b;
        L.hc2[0][0] = -1.0;
        L.out = x;
        L.out2 = r;
        lincomb<R>(c, L, done);
        // p = A^H b ; ||p||^2, ||b||^2
        long nb = op->apply(lin_kw<R>(op->zr, b), b, p, dots_with<R>(c, 2, {0, 0, 1, 1}, b, nullptr), true, done, c->stream,
                            &c->timer);
        scalar(c, st, SC_CG_INIT, nb, 2);
    }
    auto iter = [&]() {
        long nb = op->apply(lin_kw<R>(op->zr, p), p, Ap, dots_with<R>(c, 1, {0, 0}, nullptr, nullptr), false, done, c->stream,
                            &c->timer);
        scalar(c, st, SC_CG_ALPHA, nb, 1);
        LinComb<R> L{};
        L.n = n;
        L.wskip = 1;
        L.x[0] = x; L.x[1] = p; L.x[2] = r; L.x[3] = Ap;
        L.coef = &st->cf[CF_CG_U];
        L.out = x;
        L.out2 = r;
        L.dots = dots_with<R>(c, 1, {3, 3}, nullptr, nullptr);
        nb = lincomb<R>(c, L, done);
        scalar(c, st, SC_CG_V4, nb, 1);
        nb = op->apply(lin_kw<R>(op->zr, r), r, q, dots_with<R>(c, 1, {0, 0}, nullptr, nullptr), true, done, c->stream,
                       &c->timer);
        scalar(c, st, SC_CG_BETA, nb, 1);
        LinComb<R> P{};
        P.n = n;
        P.wskip = 1;
        P.x[0] = p; P.x[1] = q;
        P.coef = &st->cf[CF_CG_P];
        P.out = p;
        lincomb<R>(c, P, done);
    };
    return drive(c, st, iter);
}

// Complex-symmetric BiCG (vie_fft/CG_Iteration.h:360-528).
template <class R> KState solve_bicg(vie_ctx* c, Op<R>* op, const cx<R>* b, cx<R>* x, double tol, int maxit)
{
    KState* st = c->state;
    const long n = 3 * c->V;
    cx<R>*r = vec<R>(c, VR), *p = vec<R>(c, VP), *Ap = vec<R>(c, VAP);
    const int* done = &st->done;
    state_init(c, st, tol, maxit);
    // every vector starts at zero and is written only on active faces from here
    // on (wskip): products and updates skip the rest and run over the windows
    fill0<R>(c, x, n);
    fill0<R>(c, r, n);
    fill0<R>(c, p, n);
    fill0<R>(c, Ap, n);
    {
        LinComb<R> L{};
        L.n = n;
        L.wskip = 1;
        L.x[0] = b;
        L.hc2[0][0] = 1.0;
        L.out = x;
        L.out2 = r;
        L.dots = dots_with<R>(c, 1, {3, 3}, nullptr, nullptr, {1});
        long nb = lincomb<R>(c, L, done);
        scalar(c, st, SC_BICG_INIT, nb, 1);
        LinComb<R> S{};  // p = r / (r^T r)
        S.n = n;
        S.wskip = 1;
        S.x[0] = r;
        S.coef = &st->cf[CF_BI_S];
        S.out = p;
        lincomb<R>(c, S, done);
    }
    auto iter = [&]() {
        // Ap ; Ap^T p (unconjugated)
        long nb = op->apply(lin_kw<R>(op->zr, p), p, Ap, dots_with<R>(c, 1, {0, 1}, p, nullptr, {1}), false, done, c->stream,
                            &c->timer);
        scalar(c, st, SC_BICG_ALPHA, nb, 1);
        LinComb<R> L{};
        L.n = n;
        L.wskip = 1;
        L.x[0] = x; L.x[1] = p; L.x[2] = r; L.x[3] = Ap;
        L.coef = &st->cf[CF_BI_U];
        L.out = x;
        L.out2 = r;
        L.dots = dots_with<R>(c, 1, {3, 3}, nullptr, nullptr, {1});
        nb = lincomb<R>(c, L, done);
        scalar(c, st, SC_BICG_END, nb, 1);
        LinComb<R> P{};
        P.n = n;
        P.wskip = 1;
        P.x[0] = p; P.x[1] = r;
        P.coef = &st->cf[CF_BI_P];
        P.out = p;
        lincomb<R>(c, P, done);
    };
    return drive(c, st, iter);
}

// CORS (vie_fft/CG_Iteration.h:780-1073).
template <class R> KState solve_cors(vie_ctx* c, Op<R>* op, const cx<R>* b, cx<R>* x, double tol, int maxit)
{
    KState* st = c->state;
    const long n = 3 * c->V;
    cx<R>*r = vec<R>(c, VR), *p = vec<R>(c, VP), *q = vec<R>(c, VQ), *Ap = vec<R>(c, VAP), *Aq = vec<R>(c, VAQ);
    cx<R>*rc = vec<R>(c, VRC), *qc = vec<R>(c, VQC), *pc = vec<R>(c, VPC);
    const int* done = &st->done;
    state_init(c, st, tol, maxit);
    // every vector starts at zero and is written only on active faces from here
    // on (wskip): products and updates skip the rest and run over the windows
    fill0<R>(c, x, n);
    fill0<R>(c, r, n);
    fill0<R>(c, p, n);
    fill0<R>(c, q, n);
    fill0<R>(c, Ap, n);
    fill0<R>(c, Aq, n);
    fill0<R>(c, rc, n);
    fill0<R>(c, qc, n);
    fill0<R>(c, pc, n);
    {
        LinComb<R> L{};
        L.n = n;
        L.wskip = 1;
        L.x[0] = b;
        L.hc2[0][0] = 1.0;
        L.out = x;
        L.out2 = r;
        L.dots = dots_with<R>(c, 1, {1, 1}, b, nullptr);
        long nb = lincomb<R>(c, L, done);
        scalar(c, st, SC_BNRM, nb, 1);
    }
    auto iter = [&]() {
        long nb = op->apply(lin_kw<R>(op->zr, r), r, Ap, dots_with<R>(c, 1, {1, 0}, b, nullptr), false, done, c->stream,
                            &c->timer);
        scalar(c, st, SC_CORS_RHO, nb, 1);
        LinComb<R> L1{};  // p = r + b pc ; qc = Ap + b rc
        L1.n = n;
        L1.wskip = 1;
        L1.x[0] = r; L1.x[1] = pc; L1.x[2] = Ap; L1.x[3] = rc;
        L1.coef = &st->cf[CF_CO_L1];
        L1.out = p;
        L1.out2 = qc;
        lincomb<R>(c, L1, done);
        Lin<R> lq{};  // q = qc + b rc + b^2 q, then Aq
        lq.zr = op->zr;
        lq.wskip = 1;
        lq.x[0] = qc; lq.x[1] = rc; lq.x[2] = q;
        lq.coef = &st->cf[CF_CO_Q];
        lq.out = q;
        nb = op->apply(lq, q, Aq, dots_with<R>(c, 1, {1, 0}, b, nullptr), false, done, c->stream, &c->timer);
        scalar(c, st, SC_CORS_ALPHA, nb, 1);
        LinComb<R> L3{};  // pc = p - a q ; rc = qc - a Aq
        L3.n = n;
        L3.wskip = 1;
        L3.x[0] = p; L3.x[1] = q; L3.x[2] = qc; L3.x[3] = Aq;
        L3.coef = &st->cf[CF_CO_L3];
        L3.out = pc;
        L3.out2 = rc;
        lincomb<R>(c, L3, done);
        LinComb<R> LX{};  // x += 2a p - a^2 q
        LX.n = n;
        LX.wskip = 1;
        LX.x[0] = x; LX.x[1] = p; LX.x[2] = q;
        LX.coef = &st->cf[CF_CO_X];
        LX.out = x;
        lincomb<R>(c, LX, done);
        LinComb<R> LR{};  // r += -2a qc + a^2 Aq ; ||r||^2
        LR.n = n;
        LR.wskip = 1;
        LR.x[0] = r; LR.x[1] = qc; LR.x[2] = Aq;
        LR.coef = &st->cf[CF_CO_R];
        LR.out = r;
        LR.dots = dots_with<R>(c, 1, {0, 0}, nullptr, nullptr);
        nb = lincomb<R>(c, LR, done);
        scalar(c, st, SC_RES_END, nb, 1);
    };
    return drive(c, st, iter);
}

// BiCOR-STAB (vie_fft/CG_Iteration.h:1076-1362).
template <class R> KState solve_bicorstab(vie_ctx* c, Op<R>* op, const cx<R>* b, cx<R>* x, double tol, int maxit)
{
    KState* st = c->state;
    const long n = 3 * c->V;
    cx<R>*r = vec<R>(c, VR), *p = vec<R>(c, VP), *q = vec<R>(c, VQ), *Ap = vec<R>(c, VAP), *Aq = vec<R>(c, VAQ);
    cx<R>*qc = vec<R>(c, VQC), *pc = vec<R>(c, VPC);
    const int* done = &st->done;
    state_init(c, st, tol, maxit);
    // every vector starts at zero and is written only on active faces from here
    // on (wskip): products and updates skip the rest and run over the windows
    fill0<R>(c, x, n);
    fill0<R>(c, r, n);
    fill0<R>(c, p, n);
    fill0<R>(c, q, n);
    fill0<R>(c, Ap, n);
    fill0<R>(c, Aq, n);
    fill0<R>(c, qc, n);
    fill0<R>(c, pc, n);
    {
        LinComb<R> L{};
        L.n = n;
        L.wskip = 1;
        L.x[0] = b;
        L.hc2[0][0] = 1.0;
        L.out = x;
        L.out2 = r;
        L.dots = dots_with<R>(c, 1, {1, 1}, b, nullptr);
        long nb = lincomb<R>(c, L, done);
        scalar(c, st, SC_BNRM, nb, 1);
    }
    auto iter = [&]() {
        long nb = op->apply(lin_kw<R>(op->zr, r), r, Ap, dots_with<R>(c, 1, {1, 0}, b, nullptr), false, done, c->stream,
                            &c->timer);
        scalar(c, st, SC_BICOR_RHO, nb, 1);
        LinComb<R> L1{};  // p = r + b p - b w q
        L1.n = n;
        L1.wskip = 1;
        L1.x[0] = r; L1.x[1] = p; L1.x[2] = q;
        L1.coef = &st->cf[CF_BC_P];
        L1.out = p;
        lincomb<R>(c, L1, done);
        Lin<R> lq{};  // q = Ap + b q - b w Aq, then Aq
        lq.zr = op->zr;
        lq.wskip = 1;
        lq.x[0] = Ap; lq.x[1] = q; lq.x[2] = Aq;
        lq.coef = &st->cf[CF_BC_Q];
        lq.out = q;
        nb = op->apply(lq, q, Aq, dots_with<R>(c, 1, {1, 0}, b, nullptr), false, done, c->stream, &c->timer);
        scalar(c, st, SC_BICOR_ALPHA, nb, 1);
        LinComb<R> L2{};  // pc = r - a q ; qc = Ap - a Aq ; qc^H pc, qc^H qc
        L2.n = n;
        L2.wskip = 1;
        L2.x[0] = r; L2.x[1] = q; L2.x[2] = Ap; L2.x[3] = Aq;
        L2.coef = &st->cf[CF_BC_L2];
        L2.out = pc;
        L2.out2 = qc;
        L2.dots = dots_with<R>(c, 2, {3, 0, 3, 3}, nullptr, nullptr);
        nb = lincomb<R>(c, L2, done);
        scalar(c, st, SC_BICOR_OMEGA, nb, 2);
        LinComb<R> LX{};  // x += a p + w pc
        LX.n = n;
        LX.wskip = 1;
        LX.x[0] = x; LX.x[1] = p; LX.x[2] = pc;
        LX.coef = &st->cf[CF_BC_X];
        LX.out = x;
        lincomb<R>(c, LX, done);
        LinComb<R> LR{};  // r = pc - w qc ; ||r||^2
        LR.n = n;
        LR.wskip = 1;
        LR.x[0] = pc; LR.x[1] = qc;
        LR.coef = &st->cf[CF_BC_R];
        LR.out = r;
        LR.dots = dots_with<R>(c, 1, {0, 0}, nullptr, nullptr);
        nb = lincomb<R>(c, LR, done);
        scalar(c, st, SC_RES_END, nb, 1);
    };
    return drive(c, st, iter);
}

// Mixed-precision BiCGSTAB with reliable updates: the BiCGSTAB recurrences of
// solve_bicgstab run in FP32 (operator, FFTs, vectors), the solution is
// accumulated in FP64, and whenever the recursive residual has fallen by
// `delta` since the last replacement (and when it meets the tolerance) the
// FP32 residual is replaced by the true FP64 residual b - A64 x64 without
// restarting the Krylov process.  The stop rule is the reference's
// ||r|| / ||b|| < tol (vie_fft/CG_Iteration.h:727-740), checked on the true
// FP64 residual.  Returns the FP32 state; *true_err and *replacements report
// the final FP64 residual and the number of replacements.
KState solve_bicgstab_reliable(vie_ctx* c, double tol, int maxit, double delta, double* true_err,
                               int* replacements)
{
    using R = float;
    Op<float>* op = make_op<float>(c);
    Op<double>* op64 = make_op<double>(c);
    KState* st = c->state;
    KState* ax = c->aux;
    const long n = 3 * c->V;
    float2 *b = vec<R>(c, VB), *x = vec<R>(c, VX), *r = vec<R>(c, VR), *p = vec<R>(c, VP), *q = vec<R>(c, VQ),
           *Ap = vec<R>(c, VAP), *Aq = vec<R>(c, VAQ);
    const int* done = &st->done;
    // right-preconditioned inner iterations: the FP32 iterate lives in the
    // preconditioned variables and enters x64 as m (.) x32
    const float2* pm = c->precond ? jacobi_of<float>(c, op) : nullptr;
    hipLaunchKernelGGL(k_convert<float>, dim3(grid1(n)), dim3(NT), 0, c->stream, n, c->b64, b);
    HIPCHK(hipGetLastError());
    hipLaunchKernelGGL(k_fill_zero<double>, dim3(grid1(n)), dim3(NT), 0, c->stream, n, c->x64);
    HIPCHK(hipGetLastError());
    // the FP64 products of x64 write only active faces (wskip): t64 starts at zero
    hipLaunchKernelGGL(k_fill_zero<double>, dim3(grid1(n)), dim3(NT), 0, c->stream, n, c->t64);
    HIPCHK(hipGetLastError());
    // FP64 ||b||^2 for the true residual
    state_init(c, ax, 0.0, 1);
    {
        LinComb<double> L{};
        L.n = n;
        L.x[0] = c->b64;
        L.dots = dots_with<double>(c, 1, {1, 1}, c->b64, nullptr);
        long nb = lincomb<double>(c, L, nullptr);
        scalar(c, ax, SC_BNRM, nb, 1);
    }
    {   // a zero or non-finite right-hand side ends the solve here (x = 0)
        const KState a0 = read_state(c, ax);
        if (a0.status != KS_OK) {
            KState h{};
            h.status = a0.status;
            h.done = 1;
            *true_err = std::nan("");
            *replacements = 0;
            return h;
        }
    }
    state_init(c, st, tol, maxit, delta);  // err_ref = 1 until the first replacement
    // p, q and their products hold zero on inactive faces from here on: the
    // iterations leave those entries unwritten (wskip)
    fill0<R>(c, p, n);
    fill0<R>(c, Ap, n);
    fill0<R>(c, q, n);
    fill0<R>(c, Aq, n);
    {
        LinComb<R> L{};
        L.n = n;
        L.x[0] = b;
        L.hc2[0][0] = 1.0;
        L.out = x;
        L.out2 = r;
        L.dots = dots_with<R>(c, 1, {1, 1}, b, nullptr);
        long nb = lincomb<R>(c, L, done);
        scalar(c, st, SC_BICGSTAB_INIT, nb, 1);
    }
    auto iter = [&]() {
        Lin<R> lp{};
        lp.zr = op->zr;
        lp.wskip = 1;
        lp.x[0] = r; lp.x[1] = p; lp.x[2] = Ap;
        lp.coef = &st->cf[CF_BS_P];
        lp.out = p;
        lp.scale = pm;
        long nb = op->apply(lp, p, Ap, dots_with<R>(c, 1, {1, 0}, b, nullptr), false, done, c->stream, &c->timer);
        scalar(c, st, SC_BICGSTAB_ALPHA, nb, 1);
        Lin<R> lq{};
        lq.zr = op->zr;
        lq.wskip = 1;
        lq.x[0] = r; lq.x[1] = Ap;
        lq.coef = &st->cf[CF_BS_Q];
        lq.out = q;
        lq.scale = pm;
        nb = op->apply(lq, q, Aq, dots_with<R>(c, 2, {0, 1, 0, 0}, q, nullptr), false, done, c->stream, &c->timer);
        scalar(c, st, SC_BICGSTAB_OMEGA, nb, 2);
        LinComb<R> L{};
        L.n = n;
        L.x[0] = x; L.x[1] = p; L.x[2] = q; L.x[3] = Aq;
        L.coef = &st->cf[CF_BS_U];
        L.out = x;
        L.out2 = r;
        L.wskip = 1;
        L.dots = dots_with<R>(c, 2, {3, 3, 1, 3}, b, nullptr);
        nb = lincomb<R>(c, L, done);
        scalar(c, st, SC_BICGSTAB_END, nb, 2);
    };
    double err_ref = 1.0, terr = 1.0;
    int nrep = 0;
    KState h{};
    // As in drive(): one batch of iterations runs while the snapshot of the
    // previous one is checked; a replacement waits for the batch in flight.
    int slot = 0;
    bool primed = false;
    IterGraph g;
    for (;;) {
        if (!primed) {
            round_of(c, g, iter);
            snap_state(c, st, slot);
            primed = true;
        }
        round_of(c, g, iter);
        snap_state(c, st, slot ^ 1);
        h = wait_snap(c, slot);
        slot ^= 1;
        if (!(h.done || h.err < delta * err_ref || h.status != KS_OK || h.iter >= h.maxit)) continue;
        h = read_state(c, st);
        primed = false;
        const bool stop = h.status != KS_OK || h.iter >= h.maxit;
        // reliable update: x64 += x32 (m (.) x32 when preconditioned) ; x32 = 0 ;
        // r = b - A64 x64 (in FP64, then FP32)
        if (pm)
            hipLaunchKernelGGL(k_axpy_scaled_to64<float>, dim3(grid1(n)), dim3(NT), 0, c->stream, n, pm, x, c->x64);
        else
            hipLaunchKernelGGL(k_axpy_to64<float>, dim3(grid1(n)), dim3(NT), 0, c->stream, n, 1.0, x, c->x64);
        HIPCHK(hipGetLastError());
        fill0<R>(c, x, n);
        op64->apply(lin_kw<double>(op64->zr, c->x64), c->x64, c->t64, nodots<double>(), false, nullptr, c->stream,
                    &c->timer);
        LinComb<double> M{};
        M.n = n;
        M.x[0] = c->b64;
        M.x[1] = c->t64;
        M.hc[0][0] = 1.0;
        M.hc[1][0] = -1.0;
        M.out = c->r64;
        M.dots = dots_with<double>(c, 1, {0, 0}, nullptr, nullptr);
        long nb = lincomb<double>(c, M, nullptr);
        scalar(c, ax, SC_RELRES, nb, 1);
        ++nrep;
        terr = read_state(c, ax).err;
        if (terr < tol || stop || !std::isfinite(terr)) break;
        hipLaunchKernelGGL(k_convert<float>, dim3(grid1(n)), dim3(NT), 0, c->stream, n, c->r64, r);
        HIPCHK(hipGetLastError());
        hipLaunchKernelGGL(k_state_resume, dim3(1), dim3(1), 0, c->stream, st, delta * terr);
        HIPCHK(hipGetLastError());
        LinComb<R> Rr{};  // b^H r, r^H r of the replaced residual
        Rr.n = n;
        Rr.x[0] = r;
        Rr.hc[0][0] = 1.0;
        Rr.dots = dots_with<R>(c, 2, {1, 0, 0, 0}, b, nullptr);
        nb = lincomb<R>(c, Rr, nullptr);
        scalar(c, st, SC_BICGSTAB_REPLACE, nb, 2);
        err_ref = terr;
    }
    h = read_state(c, st);
    if (terr < tol) {
        h.converged = 1;
        h.status = KS_OK;
    } else if (h.status == KS_OK) {
        h.converged = 0;
    }
    h.err = terr;
    *true_err = terr;
    *replacements = nrep;
    return h;
}

// solve_bicgstab_reliable for the bn instances of a batch at once.  The
// decisions the single solve takes on the host -- when an instance's residual
// is replaced, and whether it is then finished or resumes -- are taken per
// instance on the device (k_mix_select, k_mix_decide), so one host loop
// drives every instance: each round enqueues `ahead` FP32 iterations over the
// instances that may still iterate and, when the last snapshot shows a
// stopped instance that is not finished, one replacement step over the
// stopped ones.  Each instance's arithmetic is that of its own single solve.
// Per-instance results are left in c->bstate (terr, nrep, fin).
void solve_bicgstab_reliable_batch(vie_ctx* c, double delta)
{
    using R = float;
    Op<float>* op = make_op<float>(c);
    Op<double>* op64 = make_op<double>(c);
    const int nb = c->bn;
    KState* st = c->state;
    KState* ax = c->aux;
    const long n = 3 * c->V, nt = n * nb;
    float2 *b = vec<R>(c, VB), *x = vec<R>(c, VX), *r = vec<R>(c, VR), *p = vec<R>(c, VP), *q = vec<R>(c, VQ),
           *Ap = vec<R>(c, VAP), *Aq = vec<R>(c, VAQ);
    const int* done = &st->done;
    const int* sel = &ax->done;  // replacement steps: instances with ax.done == 0
    if (c->precond) throw Err(VIE_ERR_ARG, "the Jacobi preconditioner applies to single solves");
    batch_all(c);
    hipLaunchKernelGGL(k_convert<float>, dim3(grid1(nt, NT, 16384)), dim3(NT), 0, c->stream, nt, c->b64, b);
    HIPCHK(hipGetLastError());
    fill0<double>(c, c->x64, n);
    fill0<double>(c, c->t64, n);
    state_init(c, ax, 0.0, 1, 0.0, false);
    {   // FP64 ||b||^2 per instance
        LinComb<double> L{};
        L.n = n;
        L.x[0] = c->b64;
        L.dots = dots_with<double>(c, 1, {1, 1}, c->b64, nullptr);
        long nbk = lincomb<double>(c, L, nullptr);
        scalar(c, ax, SC_BNRM, nbk, 1);
    }
    state_init(c, st, 0.0, 0, delta);  // each instance's tolerance and limit
    fill0<R>(c, p, n);
    fill0<R>(c, Ap, n);
    fill0<R>(c, q, n);
    fill0<R>(c, Aq, n);
    {
        LinComb<R> L{};
        L.n = n;
        L.x[0] = b;
        L.hc2[0][0] = 1.0;
        L.out = x;
        L.out2 = r;
        L.dots = dots_with<R>(c, 1, {1, 1}, b, nullptr);
        long nbk = lincomb<R>(c, L, done);
        scalar(c, st, SC_BICGSTAB_INIT, nbk, 1);
    }
    auto iter = [&]() {
        Lin<R> lp{};
        lp.zr = op->zr;
        lp.wskip = 1;
        lp.x[0] = r; lp.x[1] = p; lp.x[2] = Ap;
        lp.coef = &st->cf[CF_BS_P];
        lp.out = p;
        long nbk = op->apply(lp, p, Ap, dots_with<R>(c, 1, {1, 0}, b, nullptr), false, done, c->stream, &c->timer);
        scalar(c, st, SC_BICGSTAB_ALPHA, nbk, 1);
        Lin<R> lq{};
        lq.zr = op->zr;
        lq.wskip = 1;
        lq.x[0] = r; lq.x[1] = Ap;
        lq.coef = &st->cf[CF_BS_Q];
        lq.out = q;
        nbk = op->apply(lq, q, Aq, dots_with<R>(c, 2, {0, 1, 0, 0}, q, nullptr), false, done, c->stream, &c->timer);
        scalar(c, st, SC_BICGSTAB_OMEGA, nbk, 2);
        LinComb<R> L{};
        L.n = n;
        L.x[0] = x; L.x[1] = p; L.x[2] = q; L.x[3] = Aq;
        L.coef = &st->cf[CF_BS_U];
        L.out = x;
        L.out2 = r;
        L.wskip = 1;
        L.dots = dots_with<R>(c, 2, {3, 3, 1, 3}, b, nullptr);
        nbk = lincomb<R>(c, L, done);
        scalar(c, st, SC_BICGSTAB_END, nbk, 2);
    };
    // residual replacement of the stopped, unfinished instances:
    // x64 += x32 ; x32 = 0 ; r64 = b - A64 x64 ; then resume or finish
    auto replace = [&]() {
        const Bat keep = c->cur;
        batch_all(c);
        const unsigned gi = (unsigned)((nb + 63) / 64);
        hipLaunchKernelGGL(k_mix_select, dim3(gi), dim3(64), 0, c->stream, st, ax, nb);
        const Bat all = c->cur;
        // the x window's entries of each component (every active face lies in it)
        const long pl = (long)op64->g.ny * op64->g.nz, wlo = op64->wx0 * pl, wn = (op64->wx1 - op64->wx0) * pl;
        const dim3 gw(grid1(wn, NT, 512), 3, nb);
        hipLaunchKernelGGL(k_fold64_b<float>, gw, dim3(NT), 0, c->stream, n, c->V, wlo, wn, op64->flags, all, sel, x,
                           c->x64);
        HIPCHK(hipGetLastError());
        op64->apply(lin_kw<double>(op64->zr, c->x64), c->x64, c->t64, nodots<double>(), false, sel, c->stream,
                    &c->timer);
        LinComb<double> M{};
        M.n = n;
        M.x[0] = c->b64;
        M.x[1] = c->t64;
        M.hc[0][0] = 1.0;
        M.hc[1][0] = -1.0;
        M.out = c->r64;
        M.wskip = 1;  // only the active faces of r64 are read (below)
        M.dots = dots_with<double>(c, 1, {0, 0}, nullptr, nullptr);
        long nbk = lincomb<double>(c, M, sel);
        scalar(c, ax, SC_RELRES, nbk, 1);
        hipLaunchKernelGGL(k_mix_decide, dim3(gi), dim3(64), 0, c->stream, st, ax, nb, delta);
        HIPCHK(hipGetLastError());
        // the resumed instances (ax.done still 0): load the replaced residual
        hipLaunchKernelGGL(k_convert_b<float>, gw, dim3(NT), 0, c->stream, n, c->V, wlo, wn, op64->flags, all, sel,
                           c->r64, r);
        HIPCHK(hipGetLastError());
        LinComb<R> Rr{};  // b^H r, r^H r of the replaced residual
        Rr.n = n;
        Rr.x[0] = r;
        Rr.hc[0][0] = 1.0;
        Rr.dots = dots_with<R>(c, 2, {1, 0, 0, 0}, b, nullptr);
        nbk = lincomb<R>(c, Rr, sel);
        scalar(c, st, SC_BICGSTAB_REPLACE, nbk, 2, sel);
        c->cur = keep;
    };
    int nact = nb, slot = 0, may_run = nb;
    bool need = false;
    IterGraph g;
    round_of(c, g, iter);
    snap_state(c, st, slot);
    for (;;) {
        batch_slots(c, st, may_run, &nact);
        round_of(c, g, iter);
        if (need) replace();
        snap_state(c, st, slot ^ 1);
        wait_snap(c, slot);
        const KState* h = snap_all(c, slot);
        slot ^= 1;
        int fin = 0, stopped = 0;
        for (int i = 0; i < nb; ++i) {
            fin += h[i].fin ? 1 : 0;
            stopped += (h[i].done && !h[i].fin) ? 1 : 0;
        }
        if (fin == nb) break;
        // an instance stopped at the older snapshot may already have been
        // replaced by the step just enqueued; another step then selects nothing
        need = stopped > 0;
        may_run = nb - fin;
    }
    batch_all(c);
    read_states(c, st, c->bstate);
}

template <class R> KState solve_method(vie_ctx* c, int method, const cx<R>* b, cx<R>* x, double tol, int maxit)
{
    Op<R>* op = make_op<R>(c);
    if (c->precond && method != VIE_BICGSTAB)
        throw Err(VIE_ERR_ARG, "the Jacobi preconditioner applies to BiCGSTAB (method 3)");
    switch (method) {
    case VIE_CG: return solve_cg<R>(c, op, b, x, tol, maxit);
    case VIE_BICG: return solve_bicg<R>(c, op, b, x, tol, maxit);
    case VIE_BICGSTAB: return solve_bicgstab<R>(c, op, b, x, tol, maxit);
    case VIE_CORS: return solve_cors<R>(c, op, b, x, tol, maxit);
    case VIE_BICORSTAB: return solve_bicorstab<R>(c, op, b, x, tol, maxit);
    default: throw Err(VIE_ERR_ARG, "unknown method");
    }
}

// ||b - A64 x|| / ||b|| with the FP64 operator, on resident b64 / x64.
double true_relres(vie_ctx* c)
{
    Op<double>* op = make_op<double>(c);
    const long n = 3 * c->V;
    KState* s = c->aux;
    state_init(c, s, 0.0, 1);
    LinComb<double> L{};
    L.n = n;
    L.x[0] = c->b64;
    L.dots = dots_with<double>(c, 1, {1, 1}, c->b64, nullptr);
    long nb = lincomb<double>(c, L, nullptr);
    scalar(c, s, SC_BNRM, nb, 1);
    op->apply(lin_k<double>(op->zr, c->x64), c->x64, c->t64, nodots<double>(), false, nullptr, c->stream, nullptr);
    LinComb<double> M{};
    M.n = n;
    M.x[0] = c->b64;
    M.x[1] = c->t64;
    M.hc[0][0] = 1.0;
    M.hc[1][0] = -1.0;
    M.dots = dots_with<double>(c, 1, {0, 0}, nullptr, nullptr);
    nb = lincomb<double>(c, M, nullptr);
    scalar(c, s, SC_RELRES, nb, 1);
    return read_state(c, s).err;
}

int status_of(vie_ctx* c, const KState& h)
{
    if (h.status == KS_BREAKDOWN) {
        c->err = "Krylov breakdown (zero or non-finite denominator) in scalar step " + std::to_string(h.fail_op) +
                 " at iteration " + std::to_string(h.iter);
        return VIE_BREAKDOWN;
    }
    if (h.status == KS_NAN) {
        c->err = "non-finite residual at iteration " + std::to_string(h.iter);
        return VIE_ERR_NAN;
    }
    if (h.status == KS_DIVERGED) {
        c->err = "diverging: stop quantity " + std::to_string(h.err) + " at iteration " + std::to_string(h.iter) +
                 " exceeds " + std::to_string(h.dvg) + " x its minimum " + std::to_string(h.emin);
        return VIE_DIVERGED;
    }
    if (!h.converged) c->err = "maxit reached after " + std::to_string(h.iter) + " iterations";
    return h.converged ? VIE_OK : VIE_NOT_CONVERGED;
}

void check_ready(vie_ctx* c)
{
    require(c, c->green_ready, "vie_green_compute has not been called");
    require(c, c->medium_ready, "vie_set_medium has not been called");
}

void timing_begin(vie_ctx* c) { c->timer.reset(); }
void timing_end(vie_ctx* c)
{
    if (!c->timer.cls) return;
    HIPCHK(hipStreamSynchronize(c->stream));
    // launches enqueued after convergence exit at once; drop them (far below
    // the median) from the mean
    std::vector<float> t;
    for (size_t i = 0; i + 1 < c->timer.used; i += 2) {
        float ms = 0;
        HIPCHK(hipEventElapsedTime(&ms, c->timer.ev[i], c->timer.ev[i + 1]));
        t.push_back(ms);
    }
    if (t.empty()) return;
    std::vector<float> srt = t;
    std::nth_element(srt.begin(), srt.begin() + srt.size() / 2, srt.end());
    const float med = srt[srt.size() / 2];
    for (float v : t)
        if (v >= 0.25f * med) {
            c->timed_ms += v;
            c->timed_launches += 1;
        }
}

}  // namespace

namespace {
template <class R> void apply_dev(vie_ctx* c, const cx<R>* x, cx<R>* y, int adjoint, hipStream_t st,
                                  bool mask = true)
{
    Op<R>* op = make_op<R>(c);
    const bool timed = c->timer.cls != 0;  // vie_time_kernel also covers bare products
    if (timed) timing_begin(c);
    op->apply(lin1<R>(x), x, y, nodots<R>(), adjoint != 0, nullptr, st, timed ? &c->timer : nullptr, mask);
    if (timed) {
        HIPCHK(hipStreamSynchronize(st));
        timing_end(c);
    }
}

}  // namespace

Remap remap_of(const vie_ctx* c)
{
    Remap m{};
    for (int a = 0; a < 3; ++a) {
        m.n[a] = c->desc.n[a];
        m.u[a] = c->nu[a];
        m.perm[a] = c->perm[a];
        m.iperm[c->perm[a]] = a;
    }
    m.V = c->V;
    m.Vu = c->Vu;
    return m;
}

// Caller layout <-> internal layout on the host (k_embed / k_crop): the
// caller's grid sits at the origin of the internal one with its axes in the
// order perm, every other cell takes `fill`.
template <class T> void embed_h(const vie_ctx* c, const T* in, T* out, int ncomp, T fill)
{
    const Remap m = remap_of(c);
    const long cs[3] = {(long)m.u[1] * m.u[2], m.u[2], 1};
    const long s0 = cs[m.perm[0]], s1 = cs[m.perm[1]], s2 = cs[m.perm[2]];
    const int e0 = m.u[m.perm[0]], e1 = m.u[m.perm[1]], e2 = m.u[m.perm[2]];
    for (int cp = 0; cp < ncomp; ++cp) {
        const T* src = in + (ncomp == 3 ? m.perm[cp] : cp) * m.Vu;
        T* o = out + cp * m.V;
        for (int i = 0; i < m.n[0]; ++i)
            for (int j = 0; j < m.n[1]; ++j) {
                T* row = o + ((long)i * m.n[1] + j) * m.n[2];
                if (i < e0 && j < e1) {
                    const T* sr = src + i * s0 + j * s1;
                    for (int k = 0; k < e2; ++k) row[k] = sr[k * s2];
                    std::fill(row + e2, row + m.n[2], fill);
                } else {
                    std::fill(row, row + m.n[2], fill);
                }
            }
    }
}
template <class T> void crop_h(const vie_ctx* c, const T* in, T* out, int ncomp)
{
    const Remap m = remap_of(c);
    const long is[3] = {(long)m.n[1] * m.n[2], m.n[2], 1};
    const long s0 = is[m.iperm[0]], s1 = is[m.iperm[1]], s2 = is[m.iperm[2]];
    for (int cp = 0; cp < ncomp; ++cp) {
        const T* src = in + (ncomp == 3 ? m.iperm[cp] : cp) * m.V;
        T* o = out + cp * m.Vu;
        for (int i = 0; i < m.u[0]; ++i)
            for (int j = 0; j < m.u[1]; ++j) {
                const T* sr = src + i * s0 + j * s1;
                T* row = o + ((long)i * m.u[1] + j) * m.u[2];
                for (int k = 0; k < m.u[2]; ++k) row[k] = sr[k * s2];
            }
    }
}

// Internal axis order for a medium with these (caller-layout) flags: the
// axes by increasing fraction of their planes that hold an active face (the
// x window of the passes the narrowest, the y window next); the caller's
// order on ties, when no face is active or with windows off.
void choose_frame(const vie_ctx* c, const unsigned char* flags, int* perm)
{
    perm[0] = 0;
    perm[1] = 1;
    perm[2] = 2;
    if (!c->windows) return;
    const int* u = c->nu;
    int lo[3] = {u[0], u[1], u[2]}, hi[3] = {-1, -1, -1};
    for (int f = 0; f < 3; ++f)
        for (int i = 0; i < u[0]; ++i)
            for (int j = 0; j < u[1]; ++j) {
                const unsigned char* row = flags + f * c->Vu + ((long)i * u[1] + j) * u[2];
                int k0 = -1, k1 = -1;
                for (int k = 0; k < u[2]; ++k)
                    if (row[k]) {
                        if (k0 < 0) k0 = k;
                        k1 = k;
                    }
                if (k0 < 0) continue;
                lo[0] = std::min(lo[0], i);
                hi[0] = std::max(hi[0], i);
                lo[1] = std::min(lo[1], j);
                hi[1] = std::max(hi[1], j);
                lo[2] = std::min(lo[2], k0);
                hi[2] = std::max(hi[2], k1);
            }
    if (hi[0] < 0) return;
    double fr[3];
    for (int b = 0; b < 3; ++b) fr[b] = (double)(hi[b] - lo[b] + 1) / c->nr[b];
    std::stable_sort(perm, perm + 3, [&](int a, int b) { return fr[a] < fr[b]; });
}

using cd = std::complex<double>;

// On the caller's grid the own-axis last plane of each component is outside
// the operator (vie_fft/CG_Iteration.h:59-63); embedded, that plane becomes
// interior, so it must carry no active face (the 3-point flag rule never
// activates it).
void check_last_planes(const vie_ctx* c, const unsigned char* flags)
{
    const int* u = c->nu;
    for (int f = 0; f < 3; ++f)
        for (int i = 0; i < u[0]; ++i)
            for (int j = 0; j < u[1]; ++j)
                for (int k = 0; k < u[2]; ++k) {
                    const int idx[3] = {i, j, k};
                    if (idx[f] == u[f] - 1 && flags[f * c->Vu + ((long)i * u[1] + j) * u[2] + k])
                        throw Err(VIE_ERR_ARG, "active face on an own-axis last plane of an embedded grid");
                }
}

// ---- the far field without allocation (vie_far_field_noalloc) --------------
// Its scratch lives in a file-scope device array: the per-block partials of
// FF_NA_BLOCKS workgroups, the column sums and the direction vectors for up to
// FF_NA_MAXDIR directions (6.3 MB).  Host code never uses the array's name as
// an address: each device's copy is looked up once (hipGetSymbolAddress, at
// vie_create with that device current) and that pointer is what kernels and
// copies receive.  Every call on one device shares its copy and leaves it
// holding the last call's partials; nothing in it needs to start at zero
// (k_farfield writes every partial it later sums).  Calls that overlap on one
// device race.
constexpr int FF_NA_MAXDIR = 512;
constexpr int FF_NA_BLOCKS = 256;
constexpr long FF_NA_PART = (long)FF_NA_BLOCKS * 3 * FF_NA_MAXDIR;  // double2 entries
constexpr long FF_NA_WORDS = FF_NA_PART + 3 * FF_NA_MAXDIR + (3 * FF_NA_MAXDIR + 1) / 2;
static __device__ __align__(256) double2 scratch[FF_NA_WORDS];

constexpr int kMaxDevices = 64;
static void* g_scratch[kMaxDevices] = {};
static std::mutex g_scratch_mu;

// The current device's copy of `scratch` (looked up once per device).
static void* scratch_ptr()
{
    int dev = 0;
    HIPCHK(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) throw Err(VIE_ERR_DEVICE, "device index out of range");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            throw Err(VIE_ERR_DEVICE, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_scratch[dev] = p;
    }
    return g_scratch[dev];
}

// ============================================================== warm start
// A solve from a guess w (vie_set_guess) runs as the solve of the correction
// A e = r0 = b - A w from e = 0, x = w + e; the correction's tolerance is tol
// ||b|| / ||r0|| (squared ratios for CG's squared stop rule), so the stop rule
// still reads ||b - A x|| / ||b||.  warm_begin computes r0 into VWB, swaps it
// in as the resident right-hand side and returns ||r0|| / ||b|| (x64 = w and
// no solve is needed when that already meets tol); warm_end swaps b back and
// adds w to the correction.
double warm_begin(vie_ctx* c)
{
    Op<double>* op = make_op<double>(c);
    const long n = 3 * c->V;
    double2* w = vec<double>(c, VW);
    double2* rb = vec<double>(c, VWB);
    KState* s = c->aux;
    state_init(c, s, 0.0, 1, 0.0, false);
    LinComb<double> L{};
    L.n = n;
    L.x[0] = c->b64;
    L.dots = dots_with<double>(c, 1, {1, 1}, c->b64, nullptr);
    long nb = lincomb<double>(c, L, nullptr);
    scalar(c, s, SC_BNRM, nb, 1);
    op->apply(lin_k<double>(op->zr, w), w, c->t64, nodots<double>(), false, nullptr, c->stream, nullptr);
    LinComb<double> M{};
    M.n = n;
    M.x[0] = c->b64;
    M.x[1] = c->t64;
    M.hc[0][0] = 1.0;
    M.hc[1][0] = -1.0;
    M.out = rb;
    M.dots = dots_with<double>(c, 1, {0, 0}, nullptr, nullptr);
    nb = lincomb<double>(c, M, nullptr);
    scalar(c, s, SC_RELRES, nb, 1);
    const KState h = read_state(c, s);
    if (h.status != KS_OK) throw Err(h.status == KS_NAN ? VIE_ERR_NAN : VIE_BREAKDOWN, "zero or non-finite rhs");
    std::swap(c->b64, *reinterpret_cast<double2**>(&vecs_of<double>(c)[VWB]));
    return h.err;
}
void warm_end(vie_ctx* c)
{
    std::swap(c->b64, *reinterpret_cast<double2**>(&vecs_of<double>(c)[VWB]));
    const long n = 3 * c->V;
    LinComb<double> L{};  // x = e + w
    L.n = n;
    L.x[0] = c->x64;
    L.x[1] = vec<double>(c, VW);
    L.hc[0][0] = 1.0;
    L.hc[1][0] = 1.0;
    L.out = c->x64;
    lincomb<double>(c, L, nullptr);
    c->guess = false;
}
// The guess itself (it already meets the tolerance): x64 = w.
void warm_take_guess(vie_ctx* c)
{
    std::swap(c->b64, *reinterpret_cast<double2**>(&vecs_of<double>(c)[VWB]));
    HIPCHK(hipMemcpyAsync(c->x64, vec<double>(c, VW), 3 * c->V * sizeof(double2), hipMemcpyDeviceToDevice,
                          c->stream));
    c->guess = false;
}

// ================================================================ batches
// Incident_Wave (vie_fft/GreenFunction.h:144-248) for wavenumber k0 and angles
// in degrees, masked by the flags (vie_fft/CG_Iteration.h:559-562), into the
// internal-layout field dst on the device.
void rhs_incident(vie_ctx* c, double k0, double theta, double phi, double psi, double2* dst)
{
    const int* n = c->desc.n;
    const double* d = c->desc.d;
    const double th = kPi * theta / 180.0, ph = kPi * phi / 180.0, ps = kPi * psi / 180.0;
    const double kv[3] = {k0 * std::sin(th) * std::cos(ph), k0 * std::sin(th) * std::sin(ph), k0 * std::cos(th)};
    const double pol[3] = {std::cos(th) * std::cos(ph) * std::cos(ps) - std::sin(ph) * std::sin(ps),
                           std::cos(th) * std::sin(ph) * std::cos(ps) + std::cos(ph) * std::sin(ps),
                           -std::sin(th) * std::cos(ps)};
    // wave vector and per-component factors along the internal axes
    double ki[3];
    double2 seg[3];
    for (int a = 0; a < 3; ++a) {
        const int b = c->perm[a];
        ki[a] = kv[b];
        const double sg = std::fabs(kv[b]) < kEps ? d[a] : 2.0 * std::sin(kv[b] * d[a] / 2.0) / kv[b];
        seg[a] = make_double2(sg * pol[b], 0.0);
    }
    Op<double>* op = make_op<double>(c);
    hipLaunchKernelGGL(k_incident, dim3(grid1(c->V)), dim3(NT), 0, c->stream, n[0], n[1], n[2], d[0], d[1], d[2],
                       ki[0], ki[1], ki[2], seg[0], seg[1], seg[2], op->flags, dst);
    HIPCHK(hipGetLastError());
}

// A caller-layout host field b, masked by the flags, into dst (internal
// layout); t64 and r64 are staging.
void rhs_host(vie_ctx* c, const double* b, double2* dst)
{
    const long n = 3 * c->V;
    for (long i = 0; i < 6 * c->Vu; ++i)
        if (!std::isfinite(b[i])) throw Err(VIE_ERR_NAN, "non-finite right-hand side");
    if (c->remap) {  // caller layout staged in r64
        copy_sync(c->r64, b, 3 * c->Vu * sizeof(double2), hipMemcpyHostToDevice, c->stream);
        hipLaunchKernelGGL(k_embed<double2>, dim3(grid1(n)), dim3(NT), 0, c->stream, remap_of(c), 3, c->r64, c->t64);
        HIPCHK(hipGetLastError());
    } else {
        copy_sync(c->t64, b, n * sizeof(double2), hipMemcpyHostToDevice, c->stream);
    }
    Op<double>* op = make_op<double>(c);
    // Current_Reset of E_inc (vie_fft/CG_Iteration.h:559-562)
    hipLaunchKernelGGL(k_mask<double>, dim3(grid1(n, NT, 1L << 30)), dim3(NT), 0, c->stream, n, c->t64, op->flags,
                       dst);
    HIPCHK(hipGetLastError());
}

// FarField1's Forward_RCS (vie_fft/NearFar.h:16-109) of the field J (internal
// layout, device) for wavenumber k0, in two halves so that many fields can
// share one read-back: out_gsz() partial-sum blocks; far_launch writes 3 of
// them per field into part, far_finish turns them (on the host) into the value.
unsigned out_gsz(const vie_ctx* c) { return grid1(c->V, NT, 1024); }

void far_dir(const vie_ctx* c, double theta, double phi, double vd[3])
{
    const double th = kPi * theta / 180.0, ph = kPi * phi / 180.0;
    const double vc[3] = {std::sin(th) * std::cos(ph), std::sin(th) * std::sin(ph), std::cos(th)};
    for (int a = 0; a < 3; ++a) vd[a] = vc[c->perm[a]];  // internal axes
}

void far_launch(vie_ctx* c, const double2* J, double k0, double theta, double phi, double2* part)
{
    const int* n = c->desc.n;
    const double* d = c->desc.d;
    double vd[3];
    far_dir(c, theta, phi, vd);
    hipLaunchKernelGGL(k_farsum, dim3(out_gsz(c)), dim3(NT), 0, c->stream, n[0], n[1], n[2], d[0], d[1], d[2], k0,
                       vd[0], vd[1], vd[2], J, part);
    HIPCHK(hipGetLastError());
}

double far_finish(const vie_ctx* c, const double2* h, double k0, double theta, double phi, double psi)
{
    const double* d = c->desc.d;
    const unsigned gsz = out_gsz(c);
    const double th = kPi * theta / 180.0, ph = kPi * phi / 180.0, ps = kPi * psi / 180.0;
    double vd[3];
    far_dir(c, theta, phi, vd);
    std::complex<double> S[3];
    const double dV = d[0] * d[1] * d[2];
    for (int cpt = 0; cpt < 3; ++cpt) {
        std::complex<double> acc = 0;
        for (unsigned b = 0; b < gsz; ++b) acc += std::complex<double>(h[cpt * gsz + b].x, h[cpt * gsz + b].y);
        S[c->perm[cpt]] = std::polar(1.0, k0 * 0.5 * d[cpt] * vd[cpt]) * acc * dV;
    }
    const double tm = kZ0 * (S[0].real() * std::cos(th) * std::cos(ph) + S[1].real() * std::cos(th) * std::sin(ph) -
                             S[2].real() * std::sin(th));
    const double te = kZ0 * (S[0].real() * std::sin(ph) - S[1].real() * std::cos(ph));
    return std::cos(ps) * tm - std::sin(ps) * te;
}

double forward_scatter_of(vie_ctx* c, const double2* J, double k0, double theta, double phi, double psi)
{
    far_launch(c, J, k0, theta, phi, c->partial);
    std::vector<double2> h(3 * out_gsz(c));
    HIPCHK(hipMemcpyAsync(h.data(), c->partial, h.size() * sizeof(double2), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return far_finish(c, h.data(), k0, theta, phi, psi);
}

// NearField's Ene_Metal (vie_fft/NearFar.h:180-250) of the field J for
// material `mat`, permittivity eps (device, V) and wavenumber k0: out_gsz()
// partial sums into part (abs_launch), their sum (abs_finish).
void abs_launch(vie_ctx* c, const double2* J, const double2* eps, double k0, int mat, double2* part)
{
    const int* n = c->desc.n;
    const double* d = c->desc.d;
    const int* p = c->perm;  // the caller's extents along the internal axes
    hipLaunchKernelGGL(k_absorb, dim3(out_gsz(c)), dim3(NT), 0, c->stream, n[0], n[1], n[2], c->nu[p[0]],
                       c->nu[p[1]], c->nu[p[2]], d[0] * d[1] * d[2], k0, kZ0, J, eps, c->mat_dev, mat, part);
    HIPCHK(hipGetLastError());
}

double abs_finish(const vie_ctx* c, const double2* h)
{
    double acc = 0;
    for (unsigned b = 0; b < out_gsz(c); ++b) acc += h[b].x;
    return acc;
}

double absorbed_of(vie_ctx* c, const double2* J, const double2* eps, double k0, int mat)
{
    abs_launch(c, J, eps, k0, mat, c->partial);
    std::vector<double2> h(out_gsz(c));
    HIPCHK(hipMemcpyAsync(h.data(), c->partial, h.size() * sizeof(double2), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return abs_finish(c, h.data());
}

// Bytes of device storage one more instance of a batch needs (solver
// vectors of the mixed solve or of the longest plain method, pass buffers,
// per-instance operator tables), for the capacity check.
double batch_bytes_per_instance(const vie_ctx* c)
{
    const double V = (double)c->V;
    const double s = c->desc.precision == VIE_FP64 ? 16.0 : 8.0;
    const double fp64_fields = 4 * 3 * V * 16;           // b64, x64, t64, r64
    const double krylov = std::max(7 * 3 * V * 8, 10 * 3 * V * s);
    const double passes = 3 * 6 * V * (c->desc.precision == VIE_FP64 ? 16.0 : 8.0) + 3 * 6 * V * 8;
    const double tables = 3 * V * (16 + 8);
    return fp64_fields + krylov + passes + tables + 2 * sizeof(KState);
}

// Device storage for nb instances: solver records, resident right-hand sides
// and solutions, Krylov vectors, pass buffers and partial sums.  Growing it
// drops the resident right-hand sides, solutions and per-instance operators.
void batch_reserve(vie_ctx* c, int nb)
{
    if (nb <= c->nbcap) return;
    size_t fr = 0, tot = 0;
    HIPCHK(hipMemGetInfo(&fr, &tot));
    fr += pool_kept_bytes();  // kept buffers are available to this batch
    const double need = (nb - c->nbcap) * batch_bytes_per_instance(c);
    if (need > 0.9 * (double)fr)
        throw Err(VIE_ERR_OOM, "a batch of " + std::to_string(nb) + " instances needs about " +
                                   std::to_string((long)(need / 1e6)) + " MB more device memory, " +
                                   std::to_string((long)(fr / 1000000)) + " MB free");
    HIPCHK(hipStreamSynchronize(c->stream));
    const long n = 3 * c->V;
    dfree(c->state);
    dfree(c->aux);
    dfree(c->b64);
    dfree(c->x64);
    dfree(c->t64);
    dfree(c->r64);
    dfree(c->ilist);
    dfree(c->btol);
    dfree(c->bmaxit);
    for (void*& v : c->vec64) if (v) { raw_free(v); v = nullptr; }
    for (void*& v : c->vec32) if (v) { raw_free(v); v = nullptr; }
    if (c->snap) {
        (void)hipHostFree(c->snap);
        c->snap = nullptr;
    }
    c->nbcap = nb;
    c->rhs_ready = c->sol_ready = c->x_held = c->guess = false;
    c->state = dalloc<KState>(nb);
    c->aux = dalloc<KState>(nb);
    c->b64 = dalloc<double2>(n * nb);
    c->x64 = dalloc<double2>(n * nb);
    c->t64 = dalloc<double2>(n * nb);
    c->r64 = dalloc<double2>(n * nb);
    c->ilist = dalloc<int>(nb);
    c->btol = dalloc<double>(nb);
    c->bmaxit = dalloc<int>(nb);
    // pass buffers of every slot (an FP32 operator shares the FP64 one's)
    Op<double>* o64 = c->op64;
    Op<float>* o32 = c->op32;
    if (o64) {
        dfree(o64->W1);
        dfree(o64->W2);
        o64->W1 = dalloc<double2>(3 * 2 * c->V * nb);
        o64->W2 = dalloc<double2>(3 * 4 * c->V * nb);
        o64->own_w = true;
    }
    if (o32) {
        if (o32->own_w) {
            dfree(o32->W1);
            dfree(o32->W2);
            o32->W1 = dalloc<float2>(3 * 2 * c->V * nb);
            o32->W2 = dalloc<float2>(3 * 4 * c->V * nb);
        } else {
            o32->W1 = (float2*)o64->W1;
            o32->W2 = (float2*)o64->W2;
        }
        o32->batch_free();
    }
    if (o64) o64->batch_free();
    // partial sums: 3 per launch slot of the largest pass or vector launch
    const int* ni = c->desc.n;
    ensure_partial(c, 3L * nb * std::max<long>(std::max(zinv_blocks<double>(Geo{ni[0], ni[1], ni[2], c->V}),
                                                          zinv_blocks<float>(Geo{ni[0], ni[1], ni[2], c->V})),
                                                 4096));
    c->brhs.assign(nb, 0);
    c->bop.assign(nb, 0);
    c->bk0.assign(nb, c->desc.k0);
    c->bmulti = false;
}

// Operators of instances [i0, i0 + cnt): wavenumbers k0s and, for a medium
// with material ids, permittivity tables (cnt x n_mat complex; nullptr: the
// context's permittivity for every instance).  The G~ octants of all of them
// come from one set of launches (one workgroup row per instance); each
// instance's tables are computed as the single-instance path computes them.
template <class R> void batch_tables(vie_ctx* c, Op<R>* op, int i0, int cnt, const double* k0s, const double2* gall)
{
    op->batch_alloc(c->nbcap);
    const long tot = op->gob;
    if constexpr (sizeof(R) == 8) {
        HIPCHK(hipMemcpyAsync(op->Gb + i0 * tot, gall, cnt * tot * sizeof(double2), hipMemcpyDeviceToDevice,
                              c->stream));
    } else {
        hipLaunchKernelGGL(k_convert<float>, dim3(grid1(cnt * tot, 256, 1L << 30)), dim3(256), 0, c->stream,
                           cnt * tot, gall, op->Gb + i0 * tot);
        HIPCHK(hipGetLastError());
    }
    op->batch_k0(i0, cnt, k0s, c->stream);
}

// d_f of instance i (vie_fft/fft.h:405-588) for both operators from its
// permittivity e (device, V)
void batch_diag(vie_ctx* c, Op<double>* o64, Op<float>* o32, int i, double k0, const double2* e)
{
    const int* n = c->desc.n;
    const double* d = c->desc.d;
    double cim[3];
    for (int f = 0; f < 3; ++f) cim[f] = -(d[f] / 2.0) * kZ0 / k0;
    hipLaunchKernelGGL(k_diag<double>, dim3(grid1(c->V)), dim3(NT), 0, c->stream, n[0], n[1], n[2], cim[0], cim[1],
                       cim[2], e, o64->dgb + (long)i * 3 * c->V);
    hipLaunchKernelGGL(k_diag<float>, dim3(grid1(c->V)), dim3(NT), 0, c->stream, n[0], n[1], n[2], cim[0], cim[1],
                       cim[2], e, o32->dgb + (long)i * 3 * c->V);
    HIPCHK(hipGetLastError());
}

void batch_operators(vie_ctx* c, int i0, int cnt, const double* k0s, const double* eps_tables, int n_mat)
{
    require(c, c->medium_ready, "vie_set_medium has not been called");
    // both operators: the mixed solve's inner FP32 one and the FP64 one of
    // true residuals
    Op<double>* o64 = make_op<double>(c);
    Op<float>* o32 = make_op<float>(c);
    o64->batch_alloc(c->nbcap);
    o32->batch_alloc(c->nbcap);
    std::vector<double*> owned;  // device scratch, freed at the end (and on errors)
    auto scratch = [&](long nd) {
        owned.push_back(nullptr);
        owned.back() = dalloc<double>(nd);
        return owned.back();
    };
    double* cd_im = nullptr;
    try {
        double* kd = scratch(cnt);
        copy_sync(kd, k0s, cnt * sizeof(double), hipMemcpyHostToDevice, c->stream);
        // G~ octants of an operator's windowed-product grid l (tot entries per
        // instance), scaled like short_octant, for every instance.  The two
        // operators' grids differ when the tuner gave them different short
        // periods; then each gets its own.
        auto octants = [&](const int* l, long tot) -> const double2* {
            double2* w1 = reinterpret_cast<double2*>(scratch(2 * cnt * tot));
            double2* w2 = reinterpret_cast<double2*>(scratch(2 * cnt * tot));
            const int nmax = std::max(l[0], std::max(l[1], l[2]));
            std::vector<double> ctab(3 * 2 * nmax);
            for (int ax = 0; ax < 3; ++ax)
                for (int j = 0; j < 2 * l[ax]; ++j)
                    ctab[ax * 2 * nmax + j] = (double)cosl((long double)kPi * j / l[ax]);
            double* cd = scratch((long)ctab.size());
            copy_sync(cd, ctab.data(), ctab.size() * sizeof(double), hipMemcpyHostToDevice, c->stream);
            const double* d = c->desc.d;
            const dim3 gg(grid1(tot, 256, 1L << 30), cnt);
            hipLaunchKernelGGL(k_green_offsets, gg, dim3(256), 0, c->stream, l[0], l[1], l[2], d[0], d[1], d[2], 0.0,
                               kd, w1);
            HIPCHK(hipGetLastError());
            double2* src = w1;
            double2* dst = w2;
            for (int ax = 0; ax < 3; ++ax) {  // x: w1 -> w2, y: w2 -> w1, z: w1 -> w2
                hipLaunchKernelGGL(k_green_cos, gg, dim3(256), 0, c->stream, l[0] + 1, l[1] + 1, l[2] + 1, ax,
                                   cd + ax * 2 * nmax, src, dst);
                HIPCHK(hipGetLastError());
                std::swap(src, dst);
            }
            const int* n = c->desc.n;
            hipLaunchKernelGGL(k_scale_c64, dim3(grid1(cnt * tot, 256, 1L << 30)), dim3(256), 0, c->stream,
                               cnt * tot, (double)n[0] * n[1] * n[2] / ((double)l[0] * l[1] * l[2]), src);
            HIPCHK(hipGetLastError());
            return src;
        };
        const bool same = o64->bl[0] == o32->bl[0] && o64->bl[1] == o32->bl[1] && o64->bl[2] == o32->bl[2];
        const double2* g32 = octants(o32->bl, o32->gob);
        const double2* g64 = same ? g32 : octants(o64->bl, o64->gob);
        batch_tables<double>(c, o64, i0, cnt, k0s, g64);
        batch_tables<float>(c, o32, i0, cnt, k0s, g32);
        const double* d = c->desc.d;
        const int* n = c->desc.n;
        // per-instance permittivity from the tables (all uploaded at once),
        // one instance at a time through one V-sized buffer, and its d_f
        if (eps_tables) {
            require(c, c->mat_dev != nullptr, "permittivity tables need a medium with material ids");
            require(c, n_mat >= 1 && n_mat <= 256, "n_mat must be 1..256");
            for (long v = 0; v < 2L * cnt * n_mat; ++v)
                if (!std::isfinite(eps_tables[v])) throw Err(VIE_ERR_NAN, "non-finite permittivity");
            if (c->btabd_cap < c->nbcap) {
                dfree(c->btabd);
                c->btabd = nullptr;
                c->btabd_cap = 0;
                c->btabd = dalloc<double2>(256L * c->nbcap);
                c->btabd_cap = c->nbcap;
            }
            double2* tab_d = c->btabd + 256L * i0;  // kept for the per-instance outputs
            std::vector<double2> t(256L * cnt, make_double2(kEpsBackground, 0.0));
            for (int i = 0; i < cnt; ++i)
                for (int m = 0; m < n_mat; ++m)
                    t[256L * i + m] =
                        make_double2(eps_tables[2L * (i * n_mat + m)], eps_tables[2L * (i * n_mat + m) + 1]);
            copy_sync(tab_d, t.data(), t.size() * sizeof(double2), hipMemcpyHostToDevice, c->stream);
            // every instance's d_f (both precisions) in one launch, c_f as batch_diag forms it
            std::vector<double> cims(3L * cnt);
            for (int i = 0; i < cnt; ++i)
                for (int f = 0; f < 3; ++f) cims[3L * i + f] = -(d[f] / 2.0) * kZ0 / k0s[i];
            cd_im = dalloc<double>(cims.size());
            copy_sync(cd_im, cims.data(), cims.size() * sizeof(double), hipMemcpyHostToDevice, c->stream);
            hipLaunchKernelGGL(k_diag_batch, dim3(grid1(c->V), std::min(cnt, 65535)), dim3(NT), 0, c->stream, cnt,
                               n[0], n[1], n[2], cd_im, c->mat_dev, tab_d, o64->dgb + (long)i0 * 3 * c->V,
                               o32->dgb + (long)i0 * 3 * c->V, 3 * c->V);
            HIPCHK(hipGetLastError());
        } else {
            for (int i = 0; i < cnt; ++i) batch_diag(c, o64, o32, i0 + i, k0s[i], c->eps_dev);
        }
        HIPCHK(hipStreamSynchronize(c->stream));
        if (eps_tables) {  // kept for the per-instance absorption
            if ((int)c->btab.size() != 2 * c->nbcap * n_mat || c->bnmat != n_mat) {
                c->btab.assign(2L * c->nbcap * n_mat, 0.0);
                c->bnmat = n_mat;
            }
            std::copy(eps_tables, eps_tables + 2L * cnt * n_mat, c->btab.begin() + 2L * i0 * n_mat);
        }
    } catch (...) {
        for (double*& q : owned) dfree(q);
        dfree(cd_im);
        throw;
    }
    for (double*& q : owned) dfree(q);
    dfree(cd_im);
    for (int i = 0; i < cnt; ++i) {
        c->bk0[i0 + i] = k0s[i];
        c->bop[i0 + i] = eps_tables ? 2 : 1;
    }
    c->bmulti = true;
}

// Instance i's permittivity into c->beps (device, V): its table, or the
// context's permittivity.
const double2* batch_eps(vie_ctx* c, int i)
{
    if (!c->bmulti || c->bop[i] != 2) return c->eps_dev;
    if (!c->beps) c->beps = dalloc<double2>(c->V);
    hipLaunchKernelGGL(k_eps_table, dim3(grid1(c->V)), dim3(NT), 0, c->stream, c->V, c->mat_dev,
                       c->btabd + 256L * i, c->beps);
    HIPCHK(hipGetLastError());
    return c->beps;  // ready in stream order
}

// ||b_i - A64 x_i|| / ||b_i|| for the bn instances (FP64 operators), into res.
void true_relres_batch(vie_ctx* c, std::vector<double>& res)
{
    Op<double>* op = make_op<double>(c);
    const long n = 3 * c->V;
    KState* s = c->aux;
    batch_all(c);
    state_init(c, s, 0.0, 1, 0.0, false);
    LinComb<double> L{};
    L.n = n;
    L.x[0] = c->b64;
    L.dots = dots_with<double>(c, 1, {1, 1}, c->b64, nullptr);
    long nb = lincomb<double>(c, L, nullptr);
    scalar(c, s, SC_BNRM, nb, 1);
    // the solutions vanish on inactive faces: windowed products into a zeroed t64
    fill0<double>(c, c->t64, n);
    op->apply(lin_kw<double>(op->zr, c->x64), c->x64, c->t64, nodots<double>(), false, nullptr, c->stream, nullptr);
    LinComb<double> M{};
    M.n = n;
    M.x[0] = c->b64;
    M.x[1] = c->t64;
    M.hc[0][0] = 1.0;
    M.hc[1][0] = -1.0;
    M.dots = dots_with<double>(c, 1, {0, 0}, nullptr, nullptr);
    nb = lincomb<double>(c, M, nullptr);
    scalar(c, s, SC_RELRES, nb, 1);
    std::vector<KState> h;
    read_states(c, s, h);
    res.resize(c->bn);
    for (int i = 0; i < c->bn; ++i) res[i] = h[i].err;
}

// Leaves batch mode on every exit path of a batched solve.
struct BatchScope {
    vie_ctx* c;
    BatchScope(vie_ctx* cc, int nb) : c(cc)
    {
        c->bn = nb;
        c->in_batch = true;
        batch_all(c);
        if (c->op64) c->op64->bmulti = c->bmulti;
        if (c->op32) c->op32->bmulti = c->bmulti;
    }
    ~BatchScope()
    {
        c->bn = 1;
        c->in_batch = false;
        c->cur.nb = 1;
        c->cur.ilist = nullptr;
        if (c->op64) c->op64->bmulti = false;
        if (c->op32) c->op32->bmulti = false;
    }
};

// ====================================================================== ABI
extern "C" {

int vie_supported_n(int n) { return size_supported(n) ? 1 : 0; }

int vie_create(vie_ctx** out, const vie_desc* desc)
{
    if (!out || !desc) return VIE_ERR_ARG;
    *out = nullptr;
    // a length without compiled pass kernels runs embedded in the next
    // compiled length: the padded period 2n' >= 2n gives the same operator on
    // the caller's cells, and the extra cells are inactive background
    int nint[3];
    for (int a = 0; a < 3; ++a) {
        if (desc->n[a] < 1) return VIE_ERR_ARG;
        nint[a] = desc->n[a];
        while (nint[a] <= 256 && !size_supported(nint[a])) ++nint[a];
        if (nint[a] > 256) return VIE_ERR_UNSUPPORTED;
        if (!(desc->d[a] > 0) || !std::isfinite(desc->d[a])) return VIE_ERR_ARG;
    }
    if (!(desc->k0 > 0) || !std::isfinite(desc->k0)) return VIE_ERR_ARG;
    if (desc->precision != VIE_FP64 && desc->precision != VIE_FP32) return VIE_ERR_ARG;
    // the pass kernels index the 3 x 4V spectral array with 32-bit element
    // offsets, and the x pass with 32-bit byte offsets (16-byte FP64 entries)
    if (12L * nint[0] * nint[1] * nint[2] >= (1L << 31) || 12L * 16 * nint[0] * nint[1] * nint[2] >= (1L << 32))
        return VIE_ERR_UNSUPPORTED;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return VIE_ERR_DEVICE;
    vie_ctx* c = new vie_ctx();
    c->desc = *desc;
    c->cur.dstr = kStateInts;
    c->cur.cstr = kStateCx;
    for (int a = 0; a < 3; ++a) {
        c->nu[a] = desc->n[a];
        c->nr[a] = c->desc.n[a] = nint[a];
        c->dr[a] = desc->d[a];
        c->emb = c->emb || nint[a] != desc->n[a];
    }
    c->remap = c->emb;
    c->V = (long)nint[0] * nint[1] * nint[2];
    c->Vu = (long)desc->n[0] * desc->n[1] * desc->n[2];
    try {
        HIPCHK(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
        c->state = dalloc<KState>(1);
        c->aux = dalloc<KState>(1);
        c->hist = dalloc<double>(MAX_HIST);
        const int* ni = c->desc.n;
        ensure_partial(c, 3 * std::max<long>(std::max(zinv_blocks<double>(Geo{ni[0], ni[1], ni[2], c->V}),
                                                      zinv_blocks<float>(Geo{ni[0], ni[1], ni[2], c->V})),
                                             4096));
        c->b64 = dalloc<double2>(3 * c->V);
        c->x64 = dalloc<double2>(3 * c->V);
        c->t64 = dalloc<double2>(3 * c->V);
        c->r64 = dalloc<double2>(3 * c->V);
        if (c->desc.precision == VIE_FP64) make_op<double>(c); else make_op<float>(c);
        c->brhs.assign(1, 0);
        c->bop.assign(1, 0);
        c->bk0.assign(1, c->desc.k0);
        scratch_ptr();  // this device's copy of the far field's static scratch
    } catch (const Err& e) {
        int code = e.code;
        vie_destroy(c);
        return code;
    }
    *out = c;
    return VIE_OK;
}

int vie_destroy(vie_ctx* c)
{
    if (!c) return VIE_OK;
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    delete c->op32;   // shares buffers owned by op64: delete first
    delete c->op64;
    dfree(c->Goff);
    dfree(c->Gshort);
    dfree(c->gwork);
    dfree(c->gdct);
    dfree(c->eps_dev);
    dfree(c->mat_dev);
    dfree(c->state);
    dfree(c->aux);
    dfree(c->hist);
    dfree(c->partial);
    dfree(c->b64);
    dfree(c->x64);
    dfree(c->t64);
    dfree(c->r64);
    dfree(c->ilist);
    dfree(c->btol);
    dfree(c->bmaxit);
    dfree(c->beps);
    dfree(c->btabd);
    if (c->snap) (void)hipHostFree(c->snap);
    for (hipEvent_t e : c->snap_ev)
        if (e) (void)hipEventDestroy(e);
    for (void* p : c->vec64) if (p) raw_free(p);
    for (void* p : c->vec32) if (p) raw_free(p);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return VIE_OK;
}

const char* vie_last_error(const vie_ctx* c) { return c ? c->err.c_str() : "null context"; }

// eps (c->eps_h) must be finite, and 1/(eps - 1) finite on every face the
// operator uses (flags: c->flags_h)
static void check_eps(vie_ctx* c)
{
    for (long i = 0; i < 2 * c->V; ++i)
        if (!std::isfinite(c->eps_h[i])) throw Err(VIE_ERR_NAN, "non-finite permittivity");
    // 1/(eps - 1) must be finite on every face the operator uses
    for (long v = 0; v < c->V; ++v)
        if (c->eps_h[2 * v] == 1.0 && c->eps_h[2 * v + 1] == 0.0) {
            // eps = 1 only matters where a face term reads it
            const int* n = c->desc.n;
            const long st[3] = {(long)n[1] * n[2], n[2], 1};
            const int idx[3] = {(int)(v / st[0]), (int)((v / n[2]) % n[1]), (int)(v % n[2])};
            for (int f = 0; f < 3; ++f) {
                if (idx[f] < n[f] - 1 && c->flags_h[f * c->V + v])
                    throw Err(VIE_ERR_ARG, "eps = 1 on an active face (1/(eps-1) is infinite)");
                if (idx[f] > 0 && c->flags_h[f * c->V + v - st[f]])
                    throw Err(VIE_ERR_ARG, "eps = 1 on an active face (1/(eps-1) is infinite)");
            }
        }
}

// The diagonal term d_f of every operator from eps_dev (k_diag, the part of
// upload_medium that depends on eps and k0).
static void rebuild_diag(vie_ctx* c, bool new_medium = true)
{
    const int* n = c->desc.n;
    const double* d = c->desc.d;
    double cim[3];
    for (int f = 0; f < 3; ++f) cim[f] = -(d[f] / 2.0) * kZ0 / c->desc.k0;
    if (c->op64) {
        hipLaunchKernelGGL(k_diag<double>, dim3(grid1(c->V)), dim3(NT), 0, c->stream, n[0], n[1], n[2], cim[0], cim[1],
                           cim[2], c->eps_dev, c->op64->dg);
        ++c->op64->dgen;
    }
    if (c->op32) {
        hipLaunchKernelGGL(k_diag<float>, dim3(grid1(c->V)), dim3(NT), 0, c->stream, n[0], n[1], n[2], cim[0], cim[1],
                           cim[2], c->eps_dev, c->op32->dg);
        ++c->op32->dgen;
    }
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(c->stream));
    if (!new_medium) return;
    // per-instance operators and right-hand sides follow a new medium
    c->bop.assign(c->nbcap, 0);
    c->bmulti = false;
    c->brhs.assign(c->nbcap, 0);
}

static int set_medium_common(vie_ctx* c, const unsigned char* flags)
{
    c->flags_h.assign(flags, flags + 3 * c->V);
    check_eps(c);
    if (!c->eps_dev) c->eps_dev = dalloc<double2>(c->V);
    copy_sync(c->eps_dev, c->eps_h.data(), c->V * sizeof(double2), hipMemcpyHostToDevice, c->stream);
    c->medium_ready = true;
    // per-instance operators follow the medium's windows: set them again
    c->bop.assign(c->nbcap, 0);
    c->bmulti = false;
    c->brhs.assign(c->nbcap, 0);
    if (c->op64) upload_medium<double>(c, c->op64);
    if (c->op32) upload_medium<float>(c, c->op32);
    if (c->op64) tune_op<double>(c, c->op64);
    if (c->op32) tune_op<float>(c, c->op32);
    short_sync(c);
    return VIE_OK;
}

// Switches the internal axis order to perm.  The operators are rebuilt for the
// permuted grid (and G~ recomputed when it was computed); the medium,
// right-hand side and solution of the old order are dropped.
static void set_frame(vie_ctx* c, const int* perm)
{
    if (perm[0] == c->perm[0] && perm[1] == c->perm[1] && perm[2] == c->perm[2]) return;
    const bool had64 = c->op64 != nullptr, had32 = c->op32 != nullptr, green = c->green_ready;
    delete c->op32;  // shares buffers owned by op64: delete first
    delete c->op64;
    c->op32 = nullptr;
    c->op64 = nullptr;
    bool ident = true;
    for (int a = 0; a < 3; ++a) {
        c->perm[a] = perm[a];
        c->desc.n[a] = c->nr[perm[a]];
        c->desc.d[a] = c->dr[perm[a]];
        ident = ident && perm[a] == a;
    }
    c->remap = c->emb || !ident;
    c->green_ready = c->medium_ready = c->rhs_ready = c->sol_ready = c->x_held = c->guess = false;
    const int* ni = c->desc.n;
    ensure_partial(c, 3 * std::max<long>(std::max(zinv_blocks<double>(Geo{ni[0], ni[1], ni[2], c->V}),
                                                  zinv_blocks<float>(Geo{ni[0], ni[1], ni[2], c->V})),
                                         4096));
    if (had64) make_op<double>(c);
    if (had32) make_op<float>(c);
    if (green) green_compute(c);
}

// Checks the caller-layout flags of an embedded grid and picks the internal
// axis order for them.  A medium rejected later by set_medium_common leaves
// the context without a medium, as after any failed set_medium.
static void prepare_frame(vie_ctx* c, const unsigned char* flags)
{
    if (c->emb) check_last_planes(c, flags);
    const char* e = std::getenv("VIE_WINDOW");
    c->windows = !(e && e[0] == '0');
    int perm[3];
    choose_frame(c, flags, perm);
    set_frame(c, perm);
}

static int set_medium_int(vie_ctx* c, const unsigned char* material, const double* eps_table, int n_mat,
                          const unsigned char* flags)
{
    c->eps_h.resize(2 * c->V);
    for (long v = 0; v < c->V; ++v) {
        const int m = material[v];
        if (m >= n_mat) throw Err(VIE_ERR_ARG, "material id outside the eps table");
        c->eps_h[2 * v] = eps_table[2 * m];
        c->eps_h[2 * v + 1] = eps_table[2 * m + 1];
    }
    require(c, n_mat <= 256, "at most 256 materials");
    dfree(c->mat_dev);
    c->mat_dev = dalloc<uint8_t>(c->V);
    copy_sync(c->mat_dev, material, c->V, hipMemcpyHostToDevice, c->stream);
    return set_medium_common(c, flags);
}

// whether the caller's flags (and material ids) are those of the resident medium
static bool same_geometry(const vie_ctx* c, const unsigned char* material, const unsigned char* flags)
{
    if (!c->medium_ready || c->flags_in.size() != (size_t)(3 * c->Vu)) return false;
    if (material ? c->mat_in.size() != (size_t)c->Vu : !c->mat_in.empty()) return false;
    if (std::memcmp(c->flags_in.data(), flags, 3 * c->Vu) != 0) return false;
    return !material || std::memcmp(c->mat_in.data(), material, c->Vu) == 0;
}

static void keep_geometry(vie_ctx* c, const unsigned char* material, const unsigned char* flags)
{
    c->flags_in.assign(flags, flags + 3 * c->Vu);
    if (material) {
        c->mat_in.assign(material, material + c->Vu);
        c->mat_max = *std::max_element(material, material + c->Vu);
    } else {
        c->mat_in.clear();
    }
}

// A new eps table for the resident material ids (the frame, flags, windows,
// row ranges, plans and G~ octants stay): eps and d_f are rebuilt on the
// device.  Returns false when the full path must run (a table entry of
// exactly 1 needs the face check over every cell).
static bool set_eps_table_fast(vie_ctx* c, const double* eps_table, int n_mat)
{
    require(c, n_mat <= 256, "at most 256 materials");
    if (c->mat_max >= n_mat) throw Err(VIE_ERR_ARG, "material id outside the eps table");
    std::vector<double2> t(256, make_double2(kEpsBackground, 0.0));
    for (int m = 0; m < n_mat; ++m) {
        if (!std::isfinite(eps_table[2 * m]) || !std::isfinite(eps_table[2 * m + 1]))
            throw Err(VIE_ERR_NAN, "non-finite permittivity");
        if (eps_table[2 * m] == 1.0 && eps_table[2 * m + 1] == 0.0) return false;
        t[m] = make_double2(eps_table[2 * m], eps_table[2 * m + 1]);
    }
    double2* tab = dalloc<double2>(256);
    try {
        copy_sync(tab, t.data(), 256 * sizeof(double2), hipMemcpyHostToDevice, c->stream);
        hipLaunchKernelGGL(k_eps_table, dim3(grid1(c->V)), dim3(NT), 0, c->stream, c->V, c->mat_dev, tab, c->eps_dev);
        HIPCHK(hipGetLastError());
        rebuild_diag(c);
    } catch (...) {
        dfree(tab);
        throw;
    }
    dfree(tab);
    return true;
}

int vie_set_medium(vie_ctx* c, const unsigned char* material, const double* eps_table, int n_mat,
                   const unsigned char* flags)
{
    if (!c || !material || !eps_table || !flags || n_mat < 1) return VIE_ERR_ARG;
    try {
        if (same_geometry(c, material, flags) && set_eps_table_fast(c, eps_table, n_mat)) return VIE_OK;
        c->flags_in.clear();  // a failed set-up leaves no resident geometry
        prepare_frame(c, flags);
        int rc;
        if (c->remap) {  // background (id 0) and inactive faces on the embedding cells
            std::vector<unsigned char> m(c->V), f(3 * c->V);
            embed_h<unsigned char>(c, material, m.data(), 1, 0);
            embed_h<unsigned char>(c, flags, f.data(), 3, 0);
            rc = set_medium_int(c, m.data(), eps_table, n_mat, f.data());
        } else {
            rc = set_medium_int(c, material, eps_table, n_mat, flags);
        }
        if (rc == VIE_OK) keep_geometry(c, material, flags);
        return rc;
    } catch (const Err& e) {
        return fail(c, e);
    }
}

static int set_medium_eps_full(vie_ctx* c, const double* eps, const unsigned char* flags)
{
    prepare_frame(c, flags);
    if (c->remap) {  // dummy background permittivity (vie_fft/Model.h:82) and inactive faces outside
        std::vector<cd> e(c->V);
        std::vector<unsigned char> f(3 * c->V);
        embed_h<cd>(c, (const cd*)eps, e.data(), 1, cd(kEpsBackground, 0.0));
        embed_h<unsigned char>(c, flags, f.data(), 3, 0);
        c->eps_h.assign((const double*)e.data(), (const double*)e.data() + 2 * c->V);
        dfree(c->mat_dev);
        return set_medium_common(c, f.data());
    }
    c->eps_h.assign(eps, eps + 2 * c->V);
    dfree(c->mat_dev);
    return set_medium_common(c, flags);
}

int vie_set_medium_eps(vie_ctx* c, const double* eps, const unsigned char* flags)
{
    if (!c || !eps || !flags) return VIE_ERR_ARG;
    try {
        if (same_geometry(c, nullptr, flags)) {  // only eps changed: eps and d_f
            if (c->remap) {
                std::vector<cd> e(c->V);
                embed_h<cd>(c, (const cd*)eps, e.data(), 1, cd(kEpsBackground, 0.0));
                c->eps_h.assign((const double*)e.data(), (const double*)e.data() + 2 * c->V);
            } else {
                c->eps_h.assign(eps, eps + 2 * c->V);
            }
            check_eps(c);
            copy_sync(c->eps_dev, c->eps_h.data(), c->V * sizeof(double2), hipMemcpyHostToDevice, c->stream);
            rebuild_diag(c);
            return VIE_OK;
        }
        c->flags_in.clear();
        const int rc = set_medium_eps_full(c, eps, flags);
        if (rc == VIE_OK) keep_geometry(c, nullptr, flags);
        return rc;
    } catch (const Err& e) {
        c->flags_in.clear();
        return fail(c, e);
    }
}

// Material ids and face flags of analytic shapes, on the device, in the
// caller's layout (GetMaterial / GetFlag, vie_fft/Model.h:305-398).
static void voxelize(vie_ctx* c, int nshape, const vie_shape* shapes, unsigned char* material, unsigned char* flags)
{
    static_assert(sizeof(vie_shape) == sizeof(VoxShape), "vie_shape layout");
    require(c, nshape >= 0 && nshape <= 4096, "0..4096 shapes");
    for (int s = 0; s < nshape; ++s) {
        const vie_shape& q = shapes[s];
        require(c, q.kind == VIE_SPHERE || q.kind == VIE_ZBAND, "unknown shape kind");
        require(c, q.mat >= 0 && q.mat <= 255, "material id must be 0..255");
        for (double v : {q.c[0], q.c[1], q.c[2], q.r, q.z0, q.z1})
            if (std::isnan(v)) throw Err(VIE_ERR_NAN, "NaN in a shape");
    }
    const long V = c->Vu;
    const int* u = c->nu;
    VoxShape* sd = nullptr;
    uint8_t* md = nullptr;
    uint8_t* fd = nullptr;
    try {
        sd = dalloc<VoxShape>(std::max(nshape, 1));
        md = dalloc<uint8_t>(V);
        fd = dalloc<uint8_t>(3 * V);
        if (nshape) copy_sync(sd, shapes, nshape * sizeof(VoxShape), hipMemcpyHostToDevice, c->stream);
        hipLaunchKernelGGL(k_voxelize, dim3(grid1(V, NT, 65536)), dim3(NT), 0, c->stream, u[0], u[1], u[2], nshape, sd,
                           md, fd);
        HIPCHK(hipGetLastError());
        if (material) copy_sync(material, md, V, hipMemcpyDeviceToHost, c->stream);
        copy_sync(flags, fd, 3 * V, hipMemcpyDeviceToHost, c->stream);
    } catch (...) {
        dfree(sd);
        dfree(md);
        dfree(fd);
        throw;
    }
    dfree(sd);
    dfree(md);
    dfree(fd);
}

int vie_voxelize(vie_ctx* c, int nshape, const vie_shape* shapes, unsigned char* material, unsigned char* flags)
{
    if (!c || (nshape > 0 && !shapes) || !flags) return VIE_ERR_ARG;
    try {
        voxelize(c, nshape, shapes, material, flags);
        return VIE_OK;
    } catch (const Err& e) {
        return fail(c, e);
    }
}

int vie_set_medium_shapes(vie_ctx* c, int nshape, const vie_shape* shapes, const double* eps_table, int n_mat)
{
    if (!c || (nshape > 0 && !shapes) || !eps_table || n_mat < 1) return VIE_ERR_ARG;
    try {
        std::vector<unsigned char> m(c->Vu), f(3 * c->Vu);
        voxelize(c, nshape, shapes, m.data(), f.data());
        return vie_set_medium(c, m.data(), eps_table, n_mat, f.data());
    } catch (const Err& e) {
        return fail(c, e);
    }
}

int vie_set_k0(vie_ctx* c, double k0)
{
    if (!c) return VIE_ERR_ARG;
    try {
        require(c, k0 > 0 && std::isfinite(k0), "k0 must be positive");
        c->desc.k0 = k0;
        if (c->op64) c->op64->set_k0(k0, c->stream);
        if (c->op32) c->op32->set_k0(k0, c->stream);
        c->green_ready = false;
        c->rhs_ready = false;
        c->sol_ready = false;
        // the diagonal term's factor (Z0 / k0) changed: rebuild it
        if (c->medium_ready) rebuild_diag(c, false);
        return VIE_OK;
    } catch (const Err& e) {
        return fail(c, e);
    }
}

int vie_green_compute(vie_ctx* c)
{
    if (!c) return VIE_ERR_ARG;
    try {
        green_compute(c);
        return VIE_OK;
    } catch (const Err& e) {
        return fail(c, e);
    }
}

// Incident_Wave (vie_fft/GreenFunction.h:144-248): exp(-j k.r) on each
// component's face centre times the cell integral along its own axis.
int vie_incident(vie_ctx* c, double theta, double phi, double psi, double* E)
{
    if (!c || !E) return VIE_ERR_ARG;
    const int* n = c->nu;     // the caller's grid
    const double* d = c->dr;  // and cell sizes
    const double k0 = c->desc.k0;
    const double th = kPi * theta / 180.0, ph = kPi * phi / 180.0, ps = kPi * psi / 180.0;
    const double kv[3] = {k0 * std::sin(th) * std::cos(ph), k0 * std::sin(th) * std::sin(ph), k0 * std::cos(th)};
    const double pol[3] = {std::cos(th) * std::cos(ph) * std::cos(ps) - std::sin(ph) * std::sin(ps),
                           std::cos(th) * std::sin(ph) * std::cos(ps) + std::cos(ph) * std::sin(ps),
                           -std::sin(th) * std::cos(ps)};
    std::complex<double>* out = (std::complex<double>*)E;
    for (int cpt = 0; cpt < 3; ++cpt) {
        std::complex<double> seg;
        if (std::fabs(kv[cpt]) < kEps) {
            seg = d[cpt];
        } else {
            // integral of exp(-j k s) over a cell of length d centred on 0
            seg = 2.0 * std::sin(kv[cpt] * d[cpt] / 2.0) / kv[cpt];
        }
        seg *= pol[cpt];
        for (int i = 0; i < n[0]; ++i)
            for (int j = 0; j < n[1]; ++j)
                for (int k = 0; k < n[2]; ++k) {
                    double pos[3] = {(i + 0.5) * d[0], (j + 0.5) * d[1], (k + 0.5) * d[2]};
                    pos[cpt] += 0.5 * d[cpt];
                    const double a = kv[0] * pos[0] + kv[1] * pos[1] + kv[2] * pos[2];
                    out[cpt * c->Vu + ((long)i * n[1] + j) * n[2] + k] = std::polar(1.0, -a) * seg;
                }
    }
    return VIE_OK;
}

int vie_apply_async(vie_ctx* c, const void* x, void* y, int adjoint, void* stream)
{
    if (!c || !x || !y) return VIE_ERR_ARG;
    if (((uintptr_t)x | (uintptr_t)y) % 16) {
        c->err = "device buffers must be 16-byte aligned";
        return VIE_ERR_ARG;
    }
    try {
        check_ready(c);
        hipStream_t st = stream ? (hipStream_t)stream : c->stream;
        const Remap m = remap_of(c);
        const unsigned ge = grid1(3 * c->V), gc = grid1(3 * c->Vu);
        if (c->desc.precision == VIE_FP64) {
            if (!c->remap) {
                apply_dev<double>(c, (const double2*)x, (double2*)y, adjoint, st);
            } else {
                hipLaunchKernelGGL(k_embed<double2>, dim3(ge), dim3(NT), 0, st, m, 3, (const double2*)x, c->t64);
                apply_dev<double>(c, c->t64, c->r64, adjoint, st);
                hipLaunchKernelGGL(k_crop<double2>, dim3(gc), dim3(NT), 0, st, m, 3, c->r64, (double2*)y);
            }
        } else {
            if (!c->remap) {
                apply_dev<float>(c, (const float2*)x, (float2*)y, adjoint, st);
            } else {
                float2* xs = vec<float>(c, VT);
                float2* ys = vec<float>(c, VT2);
                hipLaunchKernelGGL(k_embed<float2>, dim3(ge), dim3(NT), 0, st, m, 3, (const float2*)x, xs);
                apply_dev<float>(c, xs, ys, adjoint, st);
                hipLaunchKernelGGL(k_crop<float2>, dim3(gc), dim3(NT), 0, st, m, 3, ys, (float2*)y);
            }
        }
        HIPCHK(hipGetLastError());
        return VIE_OK;
    } catch (const Err& e) {
        return fail(c, e);
    }
}

static int apply_host(vie_ctx* c, const double* x, double* y, int adjoint, bool mask)
{
    if (!c || !x || !y) return VIE_ERR_ARG;
    try {
        check_ready(c);
        const long n = 3 * c->V, nu3 = 3 * c->Vu;
        const Remap m = remap_of(c);
        if (c->remap) {  // caller layout staged in r64, internal layout in t64
            copy_sync(c->r64, x, nu3 * sizeof(double2), hipMemcpyHostToDevice, c->stream);
            hipLaunchKernelGGL(k_embed<double2>, dim3(grid1(n)), dim3(NT), 0, c->stream, m, 3, c->r64, c->t64);
            HIPCHK(hipGetLastError());
        } else {
            copy_sync(c->t64, x, n * sizeof(double2), hipMemcpyHostToDevice, c->stream);
        }
        if (c->desc.precision == VIE_FP64) {
            apply_dev<double>(c, c->t64, c->r64, adjoint, c->stream, mask);
        } else {
            float2* xs = vec<float>(c, VT);
            float2* ys = vec<float>(c, VT2);
            hipLaunchKernelGGL(k_convert<float>, dim3(grid1(n)), dim3(NT), 0, c->stream, n, c->t64, xs);
            apply_dev<float>(c, xs, ys, adjoint, c->stream, mask);
            hipLaunchKernelGGL(k_to_f64<float>, dim3(grid1(n)), dim3(NT), 0, c->stream, n, ys, c->r64);
        }
        if (c->remap) {
            hipLaunchKernelGGL(k_crop<double2>, dim3(grid1(nu3)), dim3(NT), 0, c->stream, m, 3, c->r64, c->t64);
            HIPCHK(hipGetLastError());
            copy_sync(y, c->t64, nu3 * sizeof(double2), hipMemcpyDeviceToHost, c->stream);
        } else {
            copy_sync(y, c->r64, n * sizeof(double2), hipMemcpyDeviceToHost, c->stream);
        }
        return VIE_OK;
    } catch (const Err& e) {
        return fail(c, e);
    }
}

int vie_apply(vie_ctx* c, const double* x, double* y, int adjoint) { return apply_host(c, x, y, adjoint, true); }

int vie_apply_raw(vie_ctx* c, const double* x, double* y, int adjoint) { return apply_host(c, x, y, adjoint, false); }

int vie_set_rhs_incident(vie_ctx* c, double theta, double phi, double psi)
{
    if (!c) return VIE_ERR_ARG;
    try {
        require(c, c->medium_ready, "vie_set_medium has not been called");
        rhs_incident(c, c->desc.k0, theta, phi, psi, c->b64);
        HIPCHK(hipStreamSynchronize(c->stream));
        c->rhs_ready = true;
        c->sol_ready = false;
        c->brhs[0] = 1;
        return VIE_OK;
    } catch (const Err& e) {
        return fail(c, e);
    }
}

int vie_set_rhs(vie_ctx* c, const double* b)
{
    if (!c || !b) return VIE_ERR_ARG;
    try {
        require(c, c->medium_ready, "vie_set_medium has not been called");
        rhs_host(c, b, c->b64);
        HIPCHK(hipStreamSynchronize(c->stream));
        c->rhs_ready = true;
        c->sol_ready = false;
        c->brhs[0] = 1;
        return VIE_OK;
    } catch (const Err& e) {
        return fail(c, e);
    }
}

int vie_solve_resident(vie_ctx* c, int method, double tol, int maxit, vie_info* info)
{
    if (!c) return VIE_ERR_ARG;
    try {
        check_ready(c);
        require(c, c->rhs_ready, "vie_set_rhs has not been called");
        require(c, tol > 0, "tol must be positive");
        if (maxit <= 0) maxit = (int)std::min<long>(c->V, 1L << 30);
        const long n = 3 * c->V;
        timing_begin(c);
        auto t0 = std::chrono::steady_clock::now();
        KState h;
        const bool warm = c->guess;
        double e0 = 1.0;  // ||b - A x0|| / ||b|| of a guess x0
        if (warm) {
            require(c, method != VIE_BICG, "a guess needs a residual-norm stop rule (not BiCG's r^T r)");
            e0 = warm_begin(c);
            if ((method == VIE_CG ? e0 * e0 : e0) < tol) {  // the guess meets the tolerance: no iteration
                warm_take_guess(c);
                h = KState{};
                h.converged = 1;
                h.err = method == VIE_CG ? e0 * e0 : e0;
                goto solved;
            }
            tol = method == VIE_CG ? tol / (e0 * e0) : tol / e0;
        }
        if (c->desc.precision == VIE_FP64) {
            h = solve_method<double>(c, method, c->b64, c->x64, tol, maxit);
        } else {
            float2* b32 = vec<float>(c, VB);
            float2* x32 = vec<float>(c, VX);
            hipLaunchKernelGGL(k_convert<float>, dim3(grid1(n)), dim3(NT), 0, c->stream, n, c->b64, b32);
            h = solve_method<float>(c, method, b32, x32, tol, maxit);
            hipLaunchKernelGGL(k_to_f64<float>, dim3(grid1(n)), dim3(NT), 0, c->stream, n, x32, c->x64);
        }
        if (warm) {
            warm_end(c);
            h.err *= method == VIE_CG ? e0 * e0 : e0;  // the stop quantity relative to b
        }
    solved:
        HIPCHK(hipStreamSynchronize(c->stream));
        auto t1 = std::chrono::steady_clock::now();
        timing_end(c);
        c->sol_ready = true;
        c->x_held = true;
        if (info) {
            info->iterations = h.iter;
            info->converged = h.converged;
            info->err = h.err;
            info->outer = 0;
            info->seconds = std::chrono::duration<double>(t1 - t0).count();
            info->true_relres = true_relres(c);
        }
        return status_of(c, h);
    } catch (const Err& e) {
        return fail(c, e);
    }
}

int vie_solve_mixed_resident(vie_ctx* c, int strategy, int inner_method, double tol, double inner_param,
                             int maxit, vie_info* info)
{
    if (!c) return VIE_ERR_ARG;
    if (strategy == VIE_MIXED_RELIABLE) {
        try {
            check_ready(c);
            require(c, c->rhs_ready, "vie_set_rhs has not been called");
            require(c, c->desc.precision == VIE_FP64, "mixed solve needs an FP64 context");
            require(c, inner_method == VIE_BICGSTAB, "reliable updates are implemented for BiCGSTAB");
            require(c, tol > 0, "tol must be positive");
            if (maxit <= 0) maxit = (int)std::min<long>(c->V, 1L << 30);
            const double delta = inner_param > 0 ? inner_param : 0.1;
            timing_begin(c);
            auto t0 = std::chrono::steady_clock::now();
            double terr = 1.0;
            int nrep = 0;
            KState h{};
            const bool warm = c->guess;
            const double e0 = warm ? warm_begin(c) : 1.0;
            if (warm && e0 < tol) {  // the guess meets the tolerance
                warm_take_guess(c);
                h.converged = 1;
                terr = e0;
            } else {
                h = solve_bicgstab_reliable(c, tol / e0, maxit, delta, &terr, &nrep);
                if (warm) {
                    warm_end(c);
                    terr *= e0;  // relative to b
                }
            }
            HIPCHK(hipStreamSynchronize(c->stream));
            auto t1 = std::chrono::steady_clock::now();
            timing_end(c);
            c->sol_ready = true;
            c->x_held = true;
            if (info) {
                info->iterations = h.iter;
                info->converged = h.converged;
                info->err = terr;
                info->outer = nrep;
                info->seconds = std::chrono::duration<double>(t1 - t0).count();
                info->true_relres = terr;
            }
            return status_of(c, h);
        } catch (const Err& e) {
            return fail(c, e);
        }
    }
    const double inner_tol = inner_param;
    try {
        require(c, strategy == VIE_MIXED_REFINE, "unknown mixed-precision strategy");
        require(c, inner_method == VIE_BICGSTAB || inner_method == VIE_CORS || inner_method == VIE_BICORSTAB,
                "inner method must be BiCGSTAB, CORS or BiCORSTAB");
        check_ready(c);
        require(c, c->rhs_ready, "vie_set_rhs has not been called");
        require(c, c->desc.precision == VIE_FP64, "mixed solve needs an FP64 context");
        require(c, tol > 0, "tol must be positive");
        if (maxit <= 0) maxit = (int)std::min<long>(c->V, 1L << 30);
        const double itol0 = inner_tol > 0 ? inner_tol : 1e-4;
        const long n = 3 * c->V;
        Op<double>* op64 = make_op<double>(c);
        make_op<float>(c);
        float2* b32 = vec<float>(c, VB);
        float2* e32 = vec<float>(c, VX);
        KState* s = c->aux;
        timing_begin(c);
        auto t0 = std::chrono::steady_clock::now();
        // bnrm2 and x = 0, or x = the guess (refinement starts from its residual)
        const bool fresh = !c->guess;
        if (!fresh) {
            HIPCHK(hipMemcpyAsync(c->x64, vec<double>(c, VW), n * sizeof(double2), hipMemcpyDeviceToDevice,
                                  c->stream));
            c->guess = false;
        }
        state_init(c, s, tol, 1);
        LinComb<double> L{};
        L.n = n;
        L.x[0] = c->b64;
        L.out2 = fresh ? c->x64 : nullptr;  // x = 0 * b
        L.dots = dots_with<double>(c, 1, {1, 1}, c->b64, nullptr);
        long nb = lincomb<double>(c, L, nullptr);
        scalar(c, s, SC_BNRM, nb, 1);
        int total = 0, outer = 0, status = VIE_OK;
        double err = 1.0, rnorm = 0.0, prev_err = 0.0, prev_scale = 0.0;
        KState last{};
        // Safeguards: one inner solve runs at most inner_cap iterations, and a
        // correction that raises the FP64 residual is taken back (x -= s e)
        // before the loop stops, so the result is never worse than before.
        const int inner_cap = std::max(200, (int)std::min<long>(maxit, 4000));
        for (;;) {
            // r64 = b - A64 x64 (FP64 residual), ||r64||
            if (outer == 0 && fresh) {
                HIPCHK(hipMemcpyAsync(c->r64, c->b64, n * sizeof(double2), hipMemcpyDeviceToDevice, c->stream));
                KState h = read_state(c, s);
                if (h.status) throw Err(h.status == KS_NAN ? VIE_ERR_NAN : VIE_BREAKDOWN, "zero or non-finite rhs");
                rnorm = std::sqrt(h.bnrm2);
                err = 1.0;
            } else {
                op64->apply(lin_k<double>(op64->zr, c->x64), c->x64, c->t64, nodots<double>(), false, nullptr, c->stream,
                            &c->timer);
                LinComb<double> M{};
                M.n = n;
                M.x[0] = c->b64;
                M.x[1] = c->t64;
                M.hc[0][0] = 1.0;
                M.hc[1][0] = -1.0;
                M.out = c->r64;
                M.dots = dots_with<double>(c, 1, {0, 0}, nullptr, nullptr);
                nb = lincomb<double>(c, M, nullptr);
                scalar(c, s, SC_RELRES, nb, 1);
                KState h = read_state(c, s);
                err = h.err;
                rnorm = std::sqrt(h.sums[0].x);
            }
            if (!std::isfinite(err) || (outer > 0 && err >= prev_err)) {
                // the last correction did not help: take it back
                hipLaunchKernelGGL(k_axpy_to64<float>, dim3(grid1(n)), dim3(NT), 0, c->stream, n, -prev_scale, e32,
                                   c->x64);
                HIPCHK(hipGetLastError());
                err = prev_err;
                status = VIE_NOT_CONVERGED;
                c->err = "FP32 inner solve stopped reducing the FP64 residual (" + std::to_string(err) + ")";
                break;
            }
            if (err < tol) break;
            if (total >= maxit) {
                status = VIE_NOT_CONVERGED;
                c->err = "maxit reached";
                break;
            }
            // inner FP32 solve of A e = r / ||r||
            const double sc = 1.0 / rnorm;
            hipLaunchKernelGGL(k_scale_from64<float>, dim3(grid1(n)), dim3(NT), 0, c->stream, n, sc, c->r64, b32);
            HIPCHK(hipGetLastError());
            const double itol = std::max(itol0, 0.5 * tol / err);
            last = solve_method<float>(c, inner_method, b32, e32, std::min(itol, 0.5),
                                       std::min(maxit - total, inner_cap));
            total += last.iter;
            ++outer;
            if (last.status == KS_NAN || (last.status == KS_BREAKDOWN && last.iter == 0)) {
                status = last.status == KS_NAN ? VIE_ERR_NAN : VIE_BREAKDOWN;
                c->err = "FP32 inner solve failed at its start";
                break;
            }
            hipLaunchKernelGGL(k_axpy_to64<float>, dim3(grid1(n)), dim3(NT), 0, c->stream, n, rnorm, e32, c->x64);
            HIPCHK(hipGetLastError());
            prev_err = err;
            prev_scale = rnorm;
        }
        HIPCHK(hipStreamSynchronize(c->stream));
        auto t1 = std::chrono::steady_clock::now();
        timing_end(c);
        c->sol_ready = true;
        c->x_held = true;
        if (info) {
            info->iterations = total;
            info->converged = status == VIE_OK;
            info->err = err;
            info->outer = outer;
            info->seconds = std::chrono::duration<double>(t1 - t0).count();
            info->true_relres = err;
        }
        return status;
    } catch (const Err& e) {
        return fail(c, e);
    }
}

int vie_set_guess(vie_ctx* c, const double* x0)
{
    if (!c) return VIE_ERR_ARG;
    try {
        require(c, c->medium_ready, "vie_set_medium has not been called");
        double2* w = vec<double>(c, VW);
        const long n = 3 * c->V;
        if (x0) {  // masked like the solution (it vanishes on inactive faces)
            rhs_host(c, x0, w);
        } else {
            require(c, c->x_held, "no solution on the device to start from");
            HIPCHK(hipMemcpyAsync(w, c->x64, n * sizeof(double2), hipMemcpyDeviceToDevice, c->stream));
        }
        HIPCHK(hipStreamSynchronize(c->stream));
        c->guess = true;
        return VIE_OK;
    } catch (const Err& e) {
        return fail(c, e);
    }
}

int vie_solve_x0(vie_ctx* c, int method, double tol, int maxit, const double* b, double* x, vie_info* info)
{
    if (!c || !b || !x) return VIE_ERR_ARG;
    int rc = vie_set_rhs(c, b);
    if (rc != VIE_OK) return rc;
    rc = vie_set_guess(c, x);
    if (rc != VIE_OK) return rc;
    rc = vie_solve_resident(c, method, tol, maxit, info);
    if (rc != VIE_OK && rc != VIE_NOT_CONVERGED && rc != VIE_BREAKDOWN) return rc;
    int rc2 = vie_get_solution(c, x);
    return rc2 != VIE_OK ? rc2 : rc;
}

int vie_get_solution(vie_ctx* c, double* x)
{
    if (!c || !x) return VIE_ERR_ARG;
    try {
        require(c, c->sol_ready, "no solution on the device");
        if (c->remap) {  // caller layout staged in t64
            const long nu3 = 3 * c->Vu;
            hipLaunchKernelGGL(k_crop<double2>, dim3(grid1(nu3)), dim3(NT), 0, c->stream, remap_of(c), 3, c->x64,
                               c->t64);
            HIPCHK(hipGetLastError());
            copy_sync(x, c->t64, nu3 * sizeof(double2), hipMemcpyDeviceToHost, c->stream);
        } else {
            copy_sync(x, c->x64, 3 * c->V * sizeof(double2), hipMemcpyDeviceToHost, c->stream);
        }
        return VIE_OK;
    } catch (const Err& e) {
        return fail(c, e);
    }
}

int vie_batch_reserve(vie_ctx* c, int nb)
{
    if (!c || nb < 1) return VIE_ERR_ARG;
    try {
        batch_reserve(c, nb);
        // back to one operator for every instance
        c->bop.assign(c->nbcap, 0);
        c->bmulti = false;
        return VIE_OK;
    } catch (const Err& e) {
        return fail(c, e);
    }
}

int vie_batch_capacity(vie_ctx* c, int* nb_max)
{
    if (!c || !nb_max) return VIE_ERR_ARG;
    try {
        size_t fr = 0, tot = 0;
        HIPCHK(hipMemGetInfo(&fr, &tot));
        fr += pool_kept_bytes();
        const double more = 0.8 * (double)fr / batch_bytes_per_instance(c);
        *nb_max = (int)std::min<double>(c->nbcap + std::max(0.0, more), 1 << 20);
        return VIE_OK;
    } catch (const Err& e) {
        return fail(c, e);
    }
}

int vie_batch_set_operators(vie_ctx* c, int i0, int count, const double* k0, const double* eps_tables, int n_mat)
{
    if (!c || i0 < 0 || count < 1 || !k0) return VIE_ERR_ARG;
    try {
        require(c, i0 + count <= c->nbcap, "instances beyond the reserved batch (vie_batch_reserve)");
        for (int i = 0; i < count; ++i) require(c, k0[i] > 0 && std::isfinite(k0[i]), "k0 must be positive");
        batch_operators(c, i0, count, k0, eps_tables, n_mat);
        for (int i = 0; i < count; ++i) c->brhs[i0 + i] = 0;
        return VIE_OK;
    } catch (const Err& e) {
        return fail(c, e);
    }
}

// the wavenumber instance i is solved at
static double batch_k0(const vie_ctx* c, int i) { return c->bmulti ? c->bk0[i] : c->desc.k0; }

int vie_batch_set_rhs_incident(vie_ctx* c, int i, double theta, double phi, double psi)
{
    if (!c || i < 0) return VIE_ERR_ARG;
    try {
        require(c, c->medium_ready, "vie_set_medium has not been called");
        require(c, i < c->nbcap, "instance beyond the reserved batch (vie_batch_reserve)");
        require(c, !c->bmulti || c->bop[i], "instance without an operator (vie_batch_set_operators)");
        // stream-ordered before every later use; nothing on the host to wait for
        rhs_incident(c, batch_k0(c, i), theta, phi, psi, c->b64 + 3 * c->V * i);
        c->brhs[i] = 1;
        if (i == 0) c->rhs_ready = true;
        c->sol_ready = false;
        return VIE_OK;
    } catch (const Err& e) {
        return fail(c, e);
    }
}

int vie_batch_set_rhs(vie_ctx* c, int i, const double* b)
{
    if (!c || i < 0 || !b) return VIE_ERR_ARG;
    try {
        require(c, c->medium_ready, "vie_set_medium has not been called");
        require(c, i < c->nbcap, "instance beyond the reserved batch (vie_batch_reserve)");
        rhs_host(c, b, c->b64 + 3 * c->V * i);
        HIPCHK(hipStreamSynchronize(c->stream));
        c->brhs[i] = 1;
        if (i == 0) c->rhs_ready = true;
        c->sol_ready = false;
        return VIE_OK;
    } catch (const Err& e) {
        return fail(c, e);
    }
}

int vie_batch_solve(vie_ctx* c, int nb, int strategy, int method, double tol, int maxit, const double* tols,
                    const int* maxits, vie_info* infos)
{
    if (!c || nb < 1) return VIE_ERR_ARG;
    try {
        check_ready(c);
        require(c, nb <= c->nbcap, "more instances than reserved (vie_batch_reserve)");
        require(c, strategy == VIE_BATCH_PLAIN || strategy == VIE_BATCH_MIXED, "unknown batch strategy");
        if (strategy == VIE_BATCH_MIXED) {
            require(c, c->desc.precision == VIE_FP64, "mixed solve needs an FP64 context");
        } else {
            require(c, method >= VIE_CG && method <= VIE_BICORSTAB, "unknown method");
        }
        const int mdef = (int)std::min<long>(c->V, 1L << 30);
        std::vector<double> tv(nb);
        std::vector<int> mv(nb);
        for (int i = 0; i < nb; ++i) {
            require(c, c->brhs[i] != 0, "instance without a right-hand side");
            require(c, !c->bmulti || c->bop[i], "instance without an operator (vie_batch_set_operators)");
            tv[i] = tols ? tols[i] : tol;
            mv[i] = maxits ? maxits[i] : maxit;
            if (mv[i] <= 0) mv[i] = mdef;
            require(c, tv[i] > 0, "tol must be positive");
        }
        copy_sync(c->btol, tv.data(), nb * sizeof(double), hipMemcpyHostToDevice, c->stream);
        copy_sync(c->bmaxit, mv.data(), nb * sizeof(int), hipMemcpyHostToDevice, c->stream);
        make_op<double>(c);  // the true FP64 residuals
        BatchScope scope(c, nb);
        const long n = 3 * c->V, nt = n * nb;
        auto t0 = std::chrono::steady_clock::now();
        std::vector<double> tres(nb);
        if (strategy == VIE_BATCH_MIXED) {
            solve_bicgstab_reliable_batch(c, 0.1);
            HIPCHK(hipStreamSynchronize(c->stream));
            for (int i = 0; i < nb; ++i) tres[i] = c->bstate[i].terr;
        } else {
            if (c->desc.precision == VIE_FP64) {
                solve_method<double>(c, method, c->b64, c->x64, tol, maxit);
            } else {
                float2* b32 = vec<float>(c, VB);
                float2* x32 = vec<float>(c, VX);
                hipLaunchKernelGGL(k_convert<float>, dim3(grid1(nt, NT, 16384)), dim3(NT), 0, c->stream, nt, c->b64,
                                   b32);
                solve_method<float>(c, method, b32, x32, tol, maxit);
                hipLaunchKernelGGL(k_to_f64<float>, dim3(grid1(nt, NT, 16384)), dim3(NT), 0, c->stream, nt, x32,
                                   c->x64);
            }
            HIPCHK(hipStreamSynchronize(c->stream));
        }
        auto t1 = std::chrono::steady_clock::now();
        const std::vector<KState> hs = c->bstate;
        if (strategy != VIE_BATCH_MIXED) true_relres_batch(c, tres);
        c->sol_ready = true;
        c->x_held = true;
        int worst = VIE_OK;
        std::string first_err;
        for (int i = 0; i < nb; ++i) {
            KState h = hs[i];
            if (strategy == VIE_BATCH_MIXED) {
                if (h.terr < tv[i]) {
                    h.converged = 1;
                    h.status = KS_OK;
                } else if (h.status == KS_OK) {
                    h.converged = 0;
                }
                h.err = h.terr;
            }
            if (infos) {
                infos[i].iterations = h.iter;
                infos[i].converged = h.converged;
                infos[i].err = h.err;
                infos[i].outer = strategy == VIE_BATCH_MIXED ? h.nrep : 0;
                infos[i].seconds = std::chrono::duration<double>(t1 - t0).count();
                infos[i].true_relres = tres[i];
            }
            const int rc = status_of(c, h);
            if (rc != VIE_OK && worst == VIE_OK) {
                worst = rc;
                first_err = "instance " + std::to_string(i) + ": " + c->err;
            }
        }
        if (worst != VIE_OK) c->err = first_err;
        return worst;
    } catch (const Err& e) {
        return fail(c, e);
    }
}

int vie_batch_get_solution(vie_ctx* c, int i, double* x)
{
    if (!c || i < 0 || !x) return VIE_ERR_ARG;
    try {
        require(c, c->sol_ready, "no solution on the device");
        require(c, i < c->nbcap, "instance beyond the reserved batch");
        const double2* xi = c->x64 + 3 * c->V * i;
        if (c->remap) {  // caller layout staged in t64
            const long nu3 = 3 * c->Vu;
            hipLaunchKernelGGL(k_crop<double2>, dim3(grid1(nu3)), dim3(NT), 0, c->stream, remap_of(c), 3, xi, c->t64);
            HIPCHK(hipGetLastError());
            copy_sync(x, c->t64, nu3 * sizeof(double2), hipMemcpyDeviceToHost, c->stream);
        } else {
            copy_sync(x, xi, 3 * c->V * sizeof(double2), hipMemcpyDeviceToHost, c->stream);
        }
        return VIE_OK;
    } catch (const Err& e) {
        return fail(c, e);
    }
}

int vie_batch_forward_scatter(vie_ctx* c, int i, double theta, double phi, double psi, double* value)
{
    if (!c || i < 0 || !value) return VIE_ERR_ARG;
    try {
        require(c, c->sol_ready, "no solution on the device");
        require(c, i < c->nbcap, "instance beyond the reserved batch");
        *value = forward_scatter_of(c, c->x64 + 3 * c->V * i, batch_k0(c, i), theta, phi, psi);
        return VIE_OK;
    } catch (const Err& e) {
        return fail(c, e);
    }
}

int vie_batch_absorbed(vie_ctx* c, int i, int mat, double* value)
{
    if (!c || i < 0 || !value) return VIE_ERR_ARG;
    try {
        require(c, c->sol_ready, "no solution on the device");
        require(c, i < c->nbcap, "instance beyond the reserved batch");
        require(c, c->mat_dev != nullptr, "vie_batch_absorbed needs material ids (vie_set_medium)");
        *value = absorbed_of(c, c->x64 + 3 * c->V * i, batch_eps(c, i), batch_k0(c, i), mat);
        return VIE_OK;
    } catch (const Err& e) {
        return fail(c, e);
    }
}

int vie_batch_outputs(vie_ctx* c, int nb, const double* theta, const double* phi, const double* psi, int mat,
                      double* forward, double* absorbed)
{
    if (!c || nb < 0 || (nb > 0 && (!theta || !phi || !psi || !forward))) return VIE_ERR_ARG;
    try {
        require(c, c->sol_ready, "no solution on the device");
        require(c, nb <= c->nbcap, "instances beyond the reserved batch");
        if (absorbed) require(c, c->mat_dev != nullptr, "absorbed power needs material ids (vie_set_medium)");
        if (nb == 0) return VIE_OK;
        const unsigned gsz = out_gsz(c);
        const long per = (absorbed ? 4L : 3L) * gsz;  // partial sums per instance
        ensure_partial(c, per * nb);
        for (int i = 0; i < nb; ++i) {
            const double2* J = c->x64 + 3 * c->V * i;
            far_launch(c, J, batch_k0(c, i), theta[i], phi[i], c->partial + per * i);
            if (absorbed) abs_launch(c, J, batch_eps(c, i), batch_k0(c, i), mat, c->partial + per * i + 3 * gsz);
        }
        std::vector<double2> h(per * nb);
        HIPCHK(hipMemcpyAsync(h.data(), c->partial, h.size() * sizeof(double2), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
        for (int i = 0; i < nb; ++i) {
            forward[i] = far_finish(c, &h[per * i], batch_k0(c, i), theta[i], phi[i], psi[i]);
            if (absorbed) absorbed[i] = abs_finish(c, &h[per * i + 3 * gsz]);
        }
        return VIE_OK;
    } catch (const Err& e) {
        return fail(c, e);
    }
}

int vie_solve_batched(vie_ctx* c, int nrhs, int strategy, int method, double tol, int maxit, const double* B,
                      double* X, vie_info* infos)
{
    if (!c || nrhs < 0 || (nrhs > 0 && (!B || !X))) return VIE_ERR_ARG;
    if (strategy != VIE_BATCH_PLAIN && strategy != VIE_BATCH_MIXED) {
        c->err = "unknown batch strategy";
        return VIE_ERR_ARG;
    }
    if (nrhs == 0) return VIE_OK;
    const long n2 = 6 * c->Vu;  // doubles per right-hand side (caller's grid)
    int worst = VIE_OK;
    std::string first_err;
    try {
        // as many right-hand sides per batch as the device holds
        int cap = 0;
        int rc = vie_batch_capacity(c, &cap);
        if (rc != VIE_OK) return rc;
        const int chunk = std::max(1, std::min(nrhs, cap));
        batch_reserve(c, chunk);
        c->bop.assign(c->nbcap, 0);
        c->bmulti = false;
        for (int k0 = 0; k0 < nrhs; k0 += chunk) {
            const int cnt = std::min(chunk, nrhs - k0);
            for (int i = 0; i < cnt; ++i) {
                rhs_host(c, B + (k0 + i) * n2, c->b64 + 3 * c->V * i);
                c->brhs[i] = 1;
            }
            HIPCHK(hipStreamSynchronize(c->stream));
            rc = vie_batch_solve(c, cnt, strategy, method, tol, maxit, nullptr, nullptr, infos ? infos + k0 : nullptr);
            if (rc != VIE_OK && rc != VIE_NOT_CONVERGED && rc != VIE_BREAKDOWN) return rc;  // hard error
            if (rc != VIE_OK && worst == VIE_OK) {
                worst = rc;
                first_err = c->err;
            }
            for (int i = 0; i < cnt; ++i) {
                const int rg = vie_batch_get_solution(c, i, X + (k0 + i) * n2);
                if (rg != VIE_OK) return rg;
            }
        }
        c->rhs_ready = true;  // instance 0 holds the last batch's first right-hand side
    } catch (const Err& e) {
        return fail(c, e);
    }
    if (worst != VIE_OK) c->err = first_err;
    return worst;
}

int vie_solve(vie_ctx* c, int method, double tol, int maxit, const double* b, double* x, vie_info* info)
{
    if (!c || !b || !x) return VIE_ERR_ARG;
    int rc = vie_set_rhs(c, b);
    if (rc != VIE_OK) return rc;
    rc = vie_solve_resident(c, method, tol, maxit, info);
    if (rc != VIE_OK && rc != VIE_NOT_CONVERGED && rc != VIE_BREAKDOWN) return rc;
    int rc2 = vie_get_solution(c, x);
    return rc2 != VIE_OK ? rc2 : rc;
}

int vie_forward_scatter(vie_ctx* c, double theta, double phi, double psi, double* value)
{
    if (!c || !value) return VIE_ERR_ARG;
    try {
        require(c, c->sol_ready, "no solution on the device");
        *value = forward_scatter_of(c, c->x64, c->desc.k0, theta, phi, psi);
        return VIE_OK;
    } catch (const Err& e) {
        return fail(c, e);
    }
}

// Far-field amplitudes S_c(v) = dV exp(j k0 delta_c v_c / 2) sum_cells J_c
// exp(j k0 r.v) of the resident solution (FarField1/FarField2 sums,
// vie_fft/NearFar.h:59-96, 321-380) for ndir directions; S holds ndir x 3
// complex.  Device buffers: vdev (3 ndir directions), part (nblk x 3 ndir
// per-block partials, nblk k_farfield workgroups), sums (3 ndir).
static void far_field_run(vie_ctx* c, int ndir, const double* theta_deg, const double* phi_deg,
                          std::complex<double>* S, double* vdev, double2* part, double2* sums, unsigned nblk,
                          double* vh, double2* h)
{
    const int* n = c->desc.n;
    const double* d = c->desc.d;
    const double k0 = c->desc.k0;
    for (int q = 0; q < ndir; ++q) {
        const double th = kPi * theta_deg[q] / 180.0, ph = kPi * phi_deg[q] / 180.0;
        const double vc[3] = {std::sin(th) * std::cos(ph), std::sin(th) * std::sin(ph), std::cos(th)};
        for (int a = 0; a < 3; ++a) vh[3 * q + a] = vc[c->perm[a]];  // internal axes
    }
    copy_sync(vdev, vh, 3 * ndir * sizeof(double), hipMemcpyHostToDevice, c->stream);
    hipLaunchKernelGGL(k_farfield, dim3(nblk), dim3(NT), 3 * n[2] * sizeof(double2), c->stream, n[0], n[1], n[2],
                       d[0], d[1], d[2], k0, ndir, vdev, c->x64, part);
    HIPCHK(hipGetLastError());
    hipLaunchKernelGGL(k_colsum, dim3((3 * ndir + NT - 1) / NT), dim3(NT), 0, c->stream, (long)nblk, 3 * ndir, part,
                       sums);
    HIPCHK(hipGetLastError());
    copy_sync(h, sums, 3 * ndir * sizeof(double2), hipMemcpyDeviceToHost, c->stream);
    const double dV = d[0] * d[1] * d[2];
    for (int q = 0; q < ndir; ++q)
        for (int cpt = 0; cpt < 3; ++cpt)
            S[3 * q + c->perm[cpt]] = std::polar(1.0, k0 * 0.5 * d[cpt] * vh[3 * q + cpt]) *
                                      std::complex<double>(h[cpt * ndir + q].x, h[cpt * ndir + q].y) * dV;
}

static void far_field_dev(vie_ctx* c, int ndir, const double* theta_deg, const double* phi_deg,
                          std::complex<double>* S)
{
    const long npen = (long)c->desc.n[0] * c->desc.n[1];
    const unsigned nblk = (unsigned)std::min<long>(npen, 1024);
    std::vector<double> vh(3 * ndir);
    std::vector<double2> h(3 * ndir);
    double* vdev = dalloc<double>(3 * ndir);
    double2* part = nullptr;
    double2* sums = nullptr;
    try {
        part = dalloc<double2>((long)nblk * 3 * ndir);
        sums = dalloc<double2>(3 * ndir);
        far_field_run(c, ndir, theta_deg, phi_deg, S, vdev, part, sums, nblk, vh.data(), h.data());
    } catch (...) {
        dfree(vdev);
        dfree(part);
        dfree(sums);
        throw;
    }
    dfree(vdev);
    dfree(part);
    dfree(sums);
}

int vie_far_field(vie_ctx* c, int ndir, const double* theta_deg, const double* phi_deg, double* S)
{
    if (!c || ndir < 0 || (ndir > 0 && (!theta_deg || !phi_deg || !S))) return VIE_ERR_ARG;
    try {
        require(c, c->sol_ready, "no solution on the device");
        require(c, c->desc.n[2] <= 1024, "z extent too large for the far-field kernel");
        for (int q0 = 0; q0 < ndir; q0 += FF_MAXD * NT) {
            const int m = std::min(ndir - q0, FF_MAXD * NT);
            far_field_dev(c, m, theta_deg + q0, phi_deg + q0, (std::complex<double>*)S + 3 * (long)q0);
        }
        return VIE_OK;
    } catch (const Err& e) {
        return fail(c, e);
    }
}

int vie_far_field_noalloc(vie_ctx* c, int ndir, const double* theta_deg, const double* phi_deg, double* S)
{
    if (!c || ndir < 0 || (ndir > 0 && (!theta_deg || !phi_deg || !S))) return VIE_ERR_ARG;
    if (ndir > FF_NA_MAXDIR) {
        c->err = "vie_far_field_noalloc: at most " + std::to_string(FF_NA_MAXDIR) + " directions per call";
        return VIE_ERR_ARG;
    }
    if (ndir == 0) return VIE_OK;
    try {
        require(c, c->sol_ready, "no solution on the device");
        require(c, c->desc.n[2] <= 1024, "z extent too large for the far-field kernel");
        double2* base = (double2*)scratch_ptr();
        double2* part = base;
        double2* sums = base + FF_NA_PART;
        double* vdev = (double*)(base + FF_NA_PART + 3 * FF_NA_MAXDIR);
        double vh[3 * FF_NA_MAXDIR];
        double2 h[3 * FF_NA_MAXDIR];
        const long npen = (long)c->desc.n[0] * c->desc.n[1];
        const unsigned nblk = (unsigned)std::min<long>(npen, FF_NA_BLOCKS);
        far_field_run(c, ndir, theta_deg, phi_deg, (std::complex<double>*)S, vdev, part, sums, nblk, vh, h);
        return VIE_OK;
    } catch (const Err& e) {
        return fail(c, e);
    }
}

int vie_rcs(vie_ctx* c, double phi_deg, double* rcs_theta, double* rcs_phi)
{
    if (!c || !rcs_theta || !rcs_phi) return VIE_ERR_ARG;
    try {
        require(c, c->sol_ready, "no solution on the device");
        std::vector<double> th(362), ph(362);
        for (int a = 0; a <= 180; ++a) {
            th[a] = th[181 + a] = a;
            ph[a] = phi_deg;          // E plane
            ph[181 + a] = phi_deg + 90.0;  // H plane
        }
        std::vector<std::complex<double>> S(3 * 362);
        far_field_dev(c, 362, th.data(), ph.data(), S.data());
        const double k0 = c->desc.k0, lam = 2.0 * kPi / k0;
        for (int a = 0; a <= 180; ++a) {
            const double t = kPi * a / 180.0;
            double p = kPi * phi_deg / 180.0;
            const std::complex<double>* e = &S[3 * a];
            const std::complex<double> mt = e[0] * (std::cos(t) * std::cos(p)) + e[1] * (std::cos(t) * std::sin(p)) -
                                            e[2] * std::sin(t);
            rcs_theta[a] = 10.0 * std::log10(k0 * k0 * kZ0 * kZ0 * std::norm(mt) / (4.0 * kPi) / lam / lam);
            p = kPi * (phi_deg + 90.0) / 180.0;
            const std::complex<double>* h = &S[3 * (181 + a)];
            const std::complex<double> mp = h[0] * std::sin(p) - h[1] * std::cos(p);
            rcs_phi[a] = 10.0 * std::log10(k0 * k0 * kZ0 * kZ0 * std::norm(mp) / (4.0 * kPi) / lam / lam);
        }
        return VIE_OK;
    } catch (const Err& e) {
        return fail(c, e);
    }
}

int vie_absorbed(vie_ctx* c, int mat, double* value)
{
    if (!c || !value) return VIE_ERR_ARG;
    try {
        require(c, c->sol_ready, "no solution on the device");
        require(c, c->mat_dev != nullptr, "vie_absorbed needs material ids (vie_set_medium)");
        *value = absorbed_of(c, c->x64, c->eps_dev, c->desc.k0, mat);
        return VIE_OK;
    } catch (const Err& e) {
        return fail(c, e);
    }
}

// Absorbed power over all active faces of the resident solution; with rhs
// non-null also the adjoint right-hand side delta_f w conj(J).
static double absorption_dev(vie_ctx* c, double2* rhs)
{
    const int* n = c->desc.n;
    const double* d = c->desc.d;
    const unsigned gsz = grid1(c->V, NT, 1024);
    ensure_partial(c, gsz);
    const double K = d[0] * d[1] * d[2] * kZ0 / (2.0 * c->desc.k0);
    hipLaunchKernelGGL(k_abs_rhs, dim3(gsz), dim3(NT), 0, c->stream, n[0], n[1], n[2], K, d[0], d[1], d[2],
                       c->x64, c->eps_dev, make_op<double>(c)->flags, rhs, c->partial);
    HIPCHK(hipGetLastError());
    std::vector<double2> h(gsz);
    HIPCHK(hipMemcpyAsync(h.data(), c->partial, h.size() * sizeof(double2), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    double acc = 0;
    for (unsigned b = 0; b < gsz; ++b) acc += h[b].x;
    return acc;
}

int vie_absorption(vie_ctx* c, double* value)
{
    if (!c || !value) return VIE_ERR_ARG;
    try {
        check_ready(c);
        require(c, c->sol_ready, "no solution on the device");
        *value = absorption_dev(c, nullptr);
        return VIE_OK;
    } catch (const Err& e) {
        return fail(c, e);
    }
}

// The adjoint gradient: method >= VIE_CG solves the adjoint system with that
// method in the context's precision; mixed runs it as the mixed-precision
// BiCGSTAB of vie_solve_mixed_resident (FP32 iterations, FP64 residual
// replacements and stop rule).
static int absorption_grad(vie_ctx* c, bool mixed, int method, double tol, int maxit, double* value,
                           double* grad_re, double* grad_im, vie_info* info)
{
    if (!c || !grad_re) return VIE_ERR_ARG;
    try {
        check_ready(c);
        require(c, c->sol_ready, "no solution on the device");
        if (mixed) {
            require(c, c->desc.precision == VIE_FP64, "mixed solve needs an FP64 context");
        } else {
            require(c, method >= VIE_CG && method <= VIE_BICORSTAB, "unknown method");
        }
        require(c, tol > 0, "tol must be positive");
        if (maxit <= 0) maxit = (int)std::min<long>(c->V, 1L << 30);
        const long n = 3 * c->V;
        double2* g = vec<double>(c, VG);
        double2* nu = vec<double>(c, VNU);
        auto t0 = std::chrono::steady_clock::now();
        const double P = absorption_dev(c, g);
        KState h;
        double terr = -1.0;
        int nrep = 0;
        if (mixed) {
            // the mixed solver works on the context's FP64 right-hand side and
            // solution: they point at g and nu during the adjoint solve (the
            // forward solution in x64 is needed again by the gradient kernel)
            struct Swap {
                vie_ctx* c;
                double2* b;
                double2* x;
                ~Swap()
                {
                    c->b64 = b;
                    c->x64 = x;
                }
            } keep{c, c->b64, c->x64};
            c->b64 = g;
            c->x64 = nu;
            h = solve_bicgstab_reliable(c, tol, maxit, 0.1, &terr, &nrep);
        } else if (c->desc.precision == VIE_FP64) {
            h = solve_method<double>(c, method, g, nu, tol, maxit);
        } else {
            float2* g32 = vec<float>(c, VB);
            float2* nu32 = vec<float>(c, VX);
            hipLaunchKernelGGL(k_convert<float>, dim3(grid1(n)), dim3(NT), 0, c->stream, n, g, g32);
            h = solve_method<float>(c, method, g32, nu32, tol, maxit);
            hipLaunchKernelGGL(k_to_f64<float>, dim3(grid1(n)), dim3(NT), 0, c->stream, n, nu32, nu);
        }
        double* gd = dalloc<double>(2 * c->V);
        const int* nn = c->desc.n;
        const double* d = c->desc.d;
        const double K = d[0] * d[1] * d[2] * kZ0 / (2.0 * c->desc.k0);
        hipLaunchKernelGGL(k_abs_grad, dim3(grid1(c->V)), dim3(NT), 0, c->stream, nn[0], nn[1], nn[2], K,
                           kZ0 / c->desc.k0, d[0], d[1], d[2], c->x64, nu, c->eps_dev, make_op<double>(c)->flags,
                           gd, gd + c->V);
        HIPCHK(hipGetLastError());
        try {
            std::vector<double> h(2 * c->V);
            copy_sync(h.data(), gd, 2 * c->V * sizeof(double), hipMemcpyDeviceToHost, c->stream);
            crop_h<double>(c, h.data(), grad_re, 1);
            if (grad_im) crop_h<double>(c, h.data() + c->V, grad_im, 1);
        } catch (...) {
            dfree(gd);
            throw;
        }
        dfree(gd);
        auto t1 = std::chrono::steady_clock::now();
        if (value) *value = P;
        if (info) {
            info->iterations = h.iter;
            info->converged = h.converged;
            info->err = mixed ? terr : h.err;
            info->outer = nrep;
            info->seconds = std::chrono::duration<double>(t1 - t0).count();
            info->true_relres = terr;
        }
        return status_of(c, h);
    } catch (const Err& e) {
        return fail(c, e);
    }
}

int vie_absorption_grad(vie_ctx* c, int method, double tol, int maxit, double* value, double* grad_re,
                        double* grad_im, vie_info* info)
{
    return absorption_grad(c, false, method, tol, maxit, value, grad_re, grad_im, info);
}

int vie_absorption_grad_mixed(vie_ctx* c, double tol, int maxit, double* value, double* grad_re, double* grad_im,
                              vie_info* info)
{
    return absorption_grad(c, true, VIE_BICGSTAB, tol, maxit, value, grad_re, grad_im, info);
}

int vie_tuning(const vie_ctx* c, int* out, double* us)
{
    if (!c || !out) return VIE_ERR_ARG;
    auto put = [&](const auto* op, int* o, double* u, double* ud) {
        o[0] = op ? (op->tuned ? 1 : 0) : -1;
        o[1] = op ? op->var : -1;
        o[2] = op ? (op->nxs ? op->nxs : c->desc.n[0]) : 0;
        o[3] = op ? (op->nys ? op->nys : c->desc.n[1]) : 0;
        o[4] = op ? (op->nzs ? op->nzs : c->desc.n[2]) : 0;
        if (u) *u = op ? op->tus : 0.0;
        if (ud) *ud = op ? op->tus_def : 0.0;
    };
    put(c->op64, out, us, us ? us + 2 : nullptr);
    put(c->op32, out + 5, us ? us + 1 : nullptr, us ? us + 3 : nullptr);
    return VIE_OK;
}

int vie_release_buffers(size_t* released)
{
    const size_t r = pool_release();
    if (released) *released = r;
    return VIE_OK;
}

int vie_set_preconditioner(vie_ctx* c, int kind)
{
    if (!c) return VIE_ERR_ARG;
    if (kind != VIE_PRECOND_NONE && kind != VIE_PRECOND_JACOBI) {
        c->err = "preconditioner must be VIE_PRECOND_NONE or VIE_PRECOND_JACOBI";
        return VIE_ERR_ARG;
    }
    c->precond = kind;
    return VIE_OK;
}

int vie_set_divergence_stop(vie_ctx* c, double factor)
{
    if (!c) return VIE_ERR_ARG;
    if (!(factor == 0.0 || (factor >= 1.0 && std::isfinite(factor)))) {
        c->err = "divergence factor must be 0 (off) or a finite value >= 1";
        return VIE_ERR_ARG;
    }
    c->dvg = factor;
    return VIE_OK;
}

int vie_history(vie_ctx* c, int max, double* err, int* count)
{
    if (!c || max < 0 || (max > 0 && !err) || !count) return VIE_ERR_ARG;
    try {
        require(c, c->sol_ready, "no solve has run");
        KState h = read_state(c, c->state);
        const int n = std::min(std::min(h.iter, MAX_HIST), max);
        if (n > 0) copy_sync(err, c->hist, n * sizeof(double), hipMemcpyDeviceToHost, c->stream);
        *count = std::min(h.iter, MAX_HIST);
        return VIE_OK;
    } catch (const Err& e) {
        return fail(c, e);
    }
}

int vie_time_kernel(vie_ctx* c, int kernel_class)
{
    if (!c || kernel_class < 0 || kernel_class > 17 || (kernel_class > 7 && kernel_class < 11)) return VIE_ERR_ARG;
    if (kernel_class == 17) kernel_class = 7;
    c->timer.cls = kernel_class;
    c->timed_launches = 0;
    c->timed_ms = 0;
    return VIE_OK;
}

int vie_kernel_time(vie_ctx* c, int* launches, double* mean_ms)
{
    if (!c) return VIE_ERR_ARG;
    if (launches) *launches = c->timed_launches;
    if (mean_ms) *mean_ms = c->timed_launches ? c->timed_ms / c->timed_launches : 0.0;
    return VIE_OK;
}

// Krylov-product model: 81 entries of the precision per cell of the x window
// (input and output vectors, diagonal, the padded W1/W2 arrays each pass
// reads and writes), the G~ octant of the x pass's pencil length and 4 bytes
// of flags and row ranges per window cell.
double vie_model_bytes_matvec(const vie_ctx* c)
{
    if (!c) return 0.0;
    int p[15];
    if (vie_plan(c, p) != VIE_OK) return 0.0;
    const double s = c->desc.precision == VIE_FP64 ? 16.0 : 8.0;
    const double Vw = (double)(p[4] - p[3]) * p[7] * p[8];
    const double oct = (double)(p[5] + 1) * (p[7] + 1) * (p[8] + 1);
    return 81.0 * Vw * s + oct * s + 4.0 * Vw;
}

int vie_plan(const vie_ctx* c, int* out)
{
    if (!c || !out) return VIE_ERR_ARG;
    const int* n = c->desc.n;
    int x0 = 0, x1 = n[0], y0 = 0, y1 = n[1], z0 = 0, z1 = n[2], ns = 0, nsy = 0, nsz = 0;
    auto take = [&](auto* o) {
        x0 = o->wx0;
        x1 = o->wx1;
        y0 = o->wy0;
        y1 = o->wy1;
        z0 = o->wz0;
        z1 = o->wz1;
        ns = o->nxs;
        nsy = o->nys;
        nsz = o->nzs;
    };
    if (c->op64)
        take(c->op64);
    else if (c->op32)
        take(c->op32);
    for (int a = 0; a < 3; ++a) {
        out[a] = c->perm[a];
        out[6 + a] = n[a];
    }
    out[3] = x0;
    out[4] = x1;
    out[5] = ns ? ns : n[0];
    out[9] = y0;
    out[10] = y1;
    out[11] = nsy ? nsy : n[1];
    out[12] = z0;
    out[13] = z1;
    out[14] = nsz ? nsz : n[2];
    return VIE_OK;
}

}  // extern "C"
