// Exercises the C ABI of include/viefft.h end to end on small grids so that
// the library's host code runs under AddressSanitizer: built against
// tests/asan/libviefft_asan.so (host code instrumented, device code as in
// the product library) by tests/asan/build.py and run by tests/test_gpu_asan.py.
// Every call's status is checked; the first unexpected one ends the run with
// exit code 1.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "viefft.h"

static int g_fail = 0;
static vie_ctx* g_ctx = nullptr;  // the context whose last error a failure prints
// A failed check ends the process at once (no static destructors: the HIP
// runtime's own teardown with contexts still alive is not what is checked)
#define EXPECT(cond)                                                             \
    do {                                                                         \
        if (!(cond)) {                                                           \
            std::printf("FAILED line %d: %s (%s)\nabi_driver FAILED\n", __LINE__, \
                        #cond, g_ctx ? vie_last_error(g_ctx) : "");              \
            std::fflush(stdout);                                                 \
            std::_Exit(1);                                                       \
        }                                                                        \
    } while (0)
#define OK(call) EXPECT((call) == VIE_OK)
#define SOLVED(call)                                                             \
    do {                                                                         \
        const int st_ = (call);                                                  \
        EXPECT(st_ == VIE_OK || st_ == VIE_NOT_CONVERGED);                       \
    } while (0)

static bool finite(const std::vector<double>& v)
{
    for (double a : v)
        if (!std::isfinite(a)) return false;
    return true;
}

static const double kPi = 3.14159265358979323846;

// a dielectric sphere of radius r cells centred in an n^3 grid, 4 nm cells, 400 nm
static void single_context(int n, vie_precision prec)
{
    vie_desc d{};
    d.n[0] = d.n[1] = d.n[2] = n;
    d.d[0] = d.d[1] = d.d[2] = 0.04;
    d.k0 = 2 * kPi / 4.0;
    d.precision = prec;
    vie_ctx* c = nullptr;
    OK(vie_create(&c, &d));
    g_ctx = c;
    const long V = (long)n * n * n;
    vie_shape sh{};
    sh.kind = VIE_SPHERE;
    sh.mat = 1;
    sh.c[0] = sh.c[1] = sh.c[2] = n / 2.0;
    sh.r = n / 3.0;
    std::vector<unsigned char> mat(V), flags(3 * V);
    OK(vie_voxelize(c, 1, &sh, mat.data(), flags.data()));
    // background 1.1 (the reference's dummy, vie_fft/Model.h:82): the unmasked
    // product below then has a finite diagonal term on every face
    const double eps[4] = {1.1, 0.0, 4.0, -0.5};
    OK(vie_set_medium(c, mat.data(), eps, 2, flags.data()));
    OK(vie_green_compute(c));
    std::vector<double> E(6 * V), y(6 * V), x(6 * V);
    OK(vie_incident(c, 20.0, 30.0, 10.0, E.data()));
    OK(vie_apply(c, E.data(), y.data(), 0));
    OK(vie_apply(c, E.data(), y.data(), 1));
    OK(vie_apply_raw(c, E.data(), y.data(), 0));
    EXPECT(finite(y));
    vie_info info{};
    for (int m = VIE_CG; m <= VIE_BICORSTAB; ++m) {
        SOLVED(vie_solve(c, m, 1e-6, 400, E.data(), x.data(), &info));
        EXPECT(finite(x));
    }
    OK(vie_set_rhs_incident(c, 0.0, 0.0, 0.0));
    SOLVED(vie_solve_resident(c, VIE_BICGSTAB, 1e-8, 0, &info));
    OK(vie_get_solution(c, x.data()));
    std::vector<double> hist(64);
    int cnt = 0;
    OK(vie_history(c, 64, hist.data(), &cnt));
    EXPECT(cnt > 0);
    if (prec == VIE_FP64) {
        SOLVED(vie_solve_mixed_resident(c, VIE_MIXED_RELIABLE, VIE_BICGSTAB, 1e-8, 0.0, 0, &info));
        SOLVED(vie_solve_mixed_resident(c, VIE_MIXED_REFINE, VIE_BICGSTAB, 1e-8, 1e-3, 0, &info));
        OK(vie_set_guess(c, x.data()));
        SOLVED(vie_solve_resident(c, VIE_BICGSTAB, 1e-10, 0, &info));
        SOLVED(vie_solve_x0(c, VIE_BICGSTAB, 1e-10, 0, E.data(), x.data(), &info));
        OK(vie_set_preconditioner(c, VIE_PRECOND_JACOBI));
        SOLVED(vie_solve_resident(c, VIE_BICGSTAB, 1e-8, 0, &info));
        SOLVED(vie_solve_mixed_resident(c, VIE_MIXED_RELIABLE, VIE_BICGSTAB, 1e-8, 0.0, 0, &info));
        OK(vie_set_preconditioner(c, VIE_PRECOND_NONE));
        OK(vie_set_divergence_stop(c, 1e3));
        const int st = vie_solve_resident(c, VIE_CORS, 1e-8, 300, &info);
        EXPECT(st == VIE_OK || st == VIE_NOT_CONVERGED || st == VIE_DIVERGED);
        OK(vie_set_divergence_stop(c, 0.0));
        SOLVED(vie_solve_resident(c, VIE_BICGSTAB, 1e-10, 0, &info));
    }
    double v = 0;
    OK(vie_forward_scatter(c, 0.0, 0.0, 0.0, &v));
    OK(vie_absorbed(c, 1, &v));
    OK(vie_absorption(c, &v));
    std::vector<double> rt(181), rp(181);
    OK(vie_rcs(c, 0.0, rt.data(), rp.data()));
    const int nd = 1500;
    std::vector<double> th(nd), ph(nd), S(6 * nd);
    for (int i = 0; i < nd; ++i) {
        th[i] = 180.0 * i / nd;
        ph[i] = 7.0 * i;
    }
    OK(vie_far_field(c, nd, th.data(), ph.data(), S.data()));
    OK(vie_far_field_noalloc(c, 512, th.data(), ph.data(), S.data()));
    EXPECT(vie_far_field_noalloc(c, 513, th.data(), ph.data(), S.data()) == VIE_ERR_ARG);
    std::vector<double> gre(V), gim(V);
    SOLVED(vie_absorption_grad(c, VIE_BICGSTAB, 1e-8, 0, &v, gre.data(), gim.data(), &info));
    if (prec == VIE_FP64) {
        SOLVED(vie_absorption_grad_mixed(c, 1e-8, 0, &v, gre.data(), gim.data(), &info));
    } else {
        EXPECT(vie_absorption_grad_mixed(c, 1e-8, 0, &v, gre.data(), gim.data(), &info) == VIE_ERR_ARG);
    }
    EXPECT(finite(gre) && finite(gim));
    int plan[16] = {0}, tun[10] = {0};
    double us[4] = {0};
    OK(vie_plan(c, plan));
    OK(vie_tuning(c, tun, us));
    EXPECT(vie_model_bytes_matvec(c) > 0);
    OK(vie_time_kernel(c, 13));
    SOLVED(vie_solve_resident(c, VIE_BICGSTAB, 1e-6, 0, &info));
    int launches = 0;
    double ms = 0;
    OK(vie_kernel_time(c, &launches, &ms));
    OK(vie_time_kernel(c, 0));
    // non-finite inputs are rejected
    std::vector<double> bad(E);
    bad[7] = std::nan("");
    EXPECT(vie_solve(c, VIE_BICGSTAB, 1e-6, 0, bad.data(), x.data(), &info) == VIE_ERR_NAN);
    // a wavelength step, per-voxel permittivity (same geometry), shapes
    OK(vie_set_k0(c, 2 * kPi / 4.5));
    OK(vie_green_compute(c));
    std::vector<double> ev(2 * V);
    for (long i = 0; i < V; ++i) {
        ev[2 * i] = mat[i] ? 3.5 : 1.0;
        ev[2 * i + 1] = mat[i] ? -0.3 : 0.0;
    }
    OK(vie_set_medium_eps(c, ev.data(), flags.data()));
    OK(vie_set_rhs_incident(c, 0.0, 0.0, 90.0));
    SOLVED(vie_solve_resident(c, VIE_BICGSTAB, 1e-8, 0, &info));
    OK(vie_set_medium_shapes(c, 1, &sh, eps, 2));
    OK(vie_green_compute(c));
    if (prec == VIE_FP64) {
        // batches: per-instance operators, mixed and plain solves, outputs
        int cap = 0;
        OK(vie_batch_capacity(c, &cap));
        EXPECT(cap >= 4);
        OK(vie_batch_reserve(c, 4));
        const double k0s[4] = {2 * kPi / 4.0, 2 * kPi / 5.0, 2 * kPi / 6.0, 2 * kPi / 7.0};
        const double tabs[16] = {1.0, 0.0, 4.0, -0.5, 1.0, 0.0, 3.0, -0.2, 1.0, 0.0, 2.5, -0.1, 1.0, 0.0, 5.0, -1.0};
        OK(vie_batch_set_operators(c, 0, 4, k0s, tabs, 2));
        for (int i = 0; i < 4; ++i) OK(vie_batch_set_rhs_incident(c, i, 0.0, 0.0, 45.0 * i));
        std::vector<vie_info> infos(4);
        SOLVED(vie_batch_solve(c, 4, VIE_BATCH_MIXED, VIE_BICGSTAB, 1e-8, 0, nullptr, nullptr, infos.data()));
        OK(vie_batch_get_solution(c, 3, x.data()));
        EXPECT(finite(x));
        const double bt[4] = {0, 0, 0, 0}, bp[4] = {0, 0, 0, 0}, bs[4] = {0, 45, 90, 135};
        double fw[4], ab[4];
        OK(vie_batch_outputs(c, 4, bt, bp, bs, 1, fw, ab));
        OK(vie_batch_forward_scatter(c, 2, 0.0, 0.0, 90.0, &v));
        OK(vie_batch_absorbed(c, 2, 1, &v));
        const double tols[4] = {1e-6, 1e-8, 1e-4, 1e-7};
        const int maxits[4] = {0, 0, 0, 5};
        const int st = vie_batch_solve(c, 4, VIE_BATCH_PLAIN, VIE_BICGSTAB, 1e-6, 0, tols, maxits, infos.data());
        EXPECT(st == VIE_OK || st == VIE_NOT_CONVERGED);
        EXPECT(vie_batch_solve(c, 5, VIE_BATCH_PLAIN, VIE_BICGSTAB, 1e-6, 0, nullptr, nullptr, nullptr) == VIE_ERR_ARG);
        std::vector<double> B(2 * 6 * V), X(2 * 6 * V);
        for (long i = 0; i < 6 * V; ++i) {
            B[i] = E[i];
            B[6 * V + i] = 0.5 * E[i];
        }
        std::vector<vie_info> in2(2);
        SOLVED(vie_solve_batched(c, 2, VIE_BATCH_MIXED, VIE_BICGSTAB, 1e-8, 0, B.data(), X.data(), in2.data()));
        EXPECT(finite(X));
    } else {
        // an FP32 context's batch: per-instance operators, plain solves
        OK(vie_batch_reserve(c, 3));
        const double k0s[3] = {2 * kPi / 4.0, 2 * kPi / 5.0, 2 * kPi / 6.0};
        const double tabs[12] = {1.1, 0.0, 4.0, -0.5, 1.1, 0.0, 3.0, -0.2, 1.1, 0.0, 2.5, -0.1};
        OK(vie_batch_set_operators(c, 0, 3, k0s, tabs, 2));
        for (int i = 0; i < 3; ++i) OK(vie_batch_set_rhs_incident(c, i, 10.0 * i, 0.0, 0.0));
        std::vector<vie_info> infos(3);
        SOLVED(vie_batch_solve(c, 3, VIE_BATCH_PLAIN, VIE_BICGSTAB, 1e-5, 0, nullptr, nullptr, infos.data()));
        OK(vie_batch_get_solution(c, 1, x.data()));
        EXPECT(finite(x));
    }
    EXPECT(vie_last_error(c) != nullptr);
    g_ctx = nullptr;
    OK(vie_destroy(c));
}

// a grid length without compiled kernels (embedded in the next one)
static void embedded_context()
{
    vie_desc d{};
    d.n[0] = 18;
    d.n[1] = 14;
    d.n[2] = 11;
    d.d[0] = d.d[1] = d.d[2] = 0.04;
    d.k0 = 2 * kPi / 4.0;
    d.precision = VIE_FP64;
    vie_ctx* c = nullptr;
    OK(vie_create(&c, &d));
    g_ctx = c;
    const long V = 18L * 14 * 11;
    vie_shape sh{};
    sh.kind = VIE_SPHERE;
    sh.mat = 1;
    sh.c[0] = 9;
    sh.c[1] = 7;
    sh.c[2] = 5.5;
    sh.r = 4.5;
    std::vector<unsigned char> mat(V), flags(3 * V);
    OK(vie_voxelize(c, 1, &sh, mat.data(), flags.data()));
    const double eps[4] = {1.0, 0.0, -4.0, -0.7};
    OK(vie_set_medium(c, mat.data(), eps, 2, flags.data()));
    OK(vie_green_compute(c));
    std::vector<double> E(6 * V), x(6 * V), y(6 * V);
    OK(vie_incident(c, 0.0, 0.0, 0.0, E.data()));
    OK(vie_apply(c, E.data(), y.data(), 0));
    vie_info info{};
    SOLVED(vie_solve(c, VIE_BICGSTAB, 1e-8, 0, E.data(), x.data(), &info));
    EXPECT(finite(x));
    std::vector<double> gre(V);
    SOLVED(vie_absorption_grad(c, VIE_BICGSTAB, 1e-8, 0, nullptr, gre.data(), nullptr, &info));
    g_ctx = nullptr;
    OK(vie_destroy(c));
}

// call sequences across features on a non-cubic, embedded grid: batches
// re-reserved at another size, a new geometry after a batch, a wavelength
// step after a batch, the preconditioner and the mixed adjoint
static void sequences()
{
    vie_desc d{};
    d.n[0] = 20;
    d.n[1] = 13;
    d.n[2] = 17;
    d.d[0] = 0.04;
    d.d[1] = 0.05;
    d.d[2] = 0.03;
    d.k0 = 2 * kPi / 5.0;
    d.precision = VIE_FP64;
    vie_ctx* c = nullptr;
    OK(vie_create(&c, &d));
    g_ctx = c;
    const long V = 20L * 13 * 17;
    vie_shape sh[2] = {};
    sh[0].kind = VIE_SPHERE;
    sh[0].mat = 1;
    sh[0].c[0] = 8;
    sh[0].c[1] = 6.5;
    sh[0].c[2] = 8;
    sh[0].r = 4;
    sh[1].kind = VIE_ZBAND;
    sh[1].mat = 2;
    sh[1].z0 = 11;
    sh[1].z1 = 15;
    const double eps[6] = {1.1, 0.0, 4.0, -0.5, 2.2, -0.1};
    OK(vie_set_medium_shapes(c, 2, sh, eps, 3));
    OK(vie_green_compute(c));
    vie_info info{};
    std::vector<double> x(6 * V), gre(V);
    for (int nb : {3, 5, 2}) {  // batches of different sizes one after the other
        OK(vie_batch_reserve(c, nb));
        std::vector<double> k0s(nb), tabs(6 * nb);
        for (int i = 0; i < nb; ++i) {
            k0s[i] = 2 * kPi / (4.0 + 0.5 * i);
            for (int m = 0; m < 6; ++m) tabs[6 * i + m] = eps[m] * (m >= 2 ? 1.0 + 0.05 * i : 1.0);
        }
        OK(vie_batch_set_operators(c, 0, nb, k0s.data(), tabs.data(), 3));
        for (int i = 0; i < nb; ++i) OK(vie_batch_set_rhs_incident(c, i, 15.0 * i, 30.0, 0.0));
        std::vector<vie_info> infos(nb);
        SOLVED(vie_batch_solve(c, nb, VIE_BATCH_MIXED, VIE_BICGSTAB, 1e-7, 0, nullptr, nullptr, infos.data()));
        OK(vie_batch_get_solution(c, nb - 1, x.data()));
        EXPECT(finite(x));
    }
    // a wavelength step after the batches, then single solves
    OK(vie_set_k0(c, 2 * kPi / 6.0));
    OK(vie_green_compute(c));
    OK(vie_set_rhs_incident(c, 0.0, 0.0, 0.0));
    SOLVED(vie_solve_resident(c, VIE_BICGSTAB, 1e-9, 0, &info));
    OK(vie_set_preconditioner(c, VIE_PRECOND_JACOBI));
    SOLVED(vie_solve_mixed_resident(c, VIE_MIXED_RELIABLE, VIE_BICGSTAB, 1e-9, 0.0, 0, &info));
    OK(vie_set_preconditioner(c, VIE_PRECOND_NONE));
    SOLVED(vie_absorption_grad_mixed(c, 1e-9, 0, nullptr, gre.data(), nullptr, &info));
    EXPECT(finite(gre));
    // a new geometry (other windows) after a batch, then a batch on it
    sh[0].c[0] = 12;
    sh[0].r = 5;
    OK(vie_set_medium_shapes(c, 2, sh, eps, 3));
    OK(vie_green_compute(c));
    OK(vie_batch_reserve(c, 2));
    const double k0b[2] = {2 * kPi / 4.0, 2 * kPi / 7.0};
    const double tb[12] = {1.1, 0.0, 3.0, -0.3, 2.0, -0.1, 1.1, 0.0, 5.0, -0.6, 2.5, -0.2};
    OK(vie_batch_set_operators(c, 0, 2, k0b, tb, 3));
    for (int i = 0; i < 2; ++i) OK(vie_batch_set_rhs_incident(c, i, 0.0, 0.0, 90.0 * i));
    std::vector<vie_info> infos(2);
    {  // CORS may break down or stagnate on this medium: any solver status, no error
        const int st = vie_batch_solve(c, 2, VIE_BATCH_PLAIN, VIE_CORS, 1e-6, 500, nullptr, nullptr, infos.data());
        EXPECT(st == VIE_OK || st == VIE_NOT_CONVERGED || st == VIE_BREAKDOWN || st == VIE_DIVERGED);
    }
    SOLVED(vie_batch_solve(c, 2, VIE_BATCH_MIXED, VIE_BICGSTAB, 1e-8, 0, nullptr, nullptr, infos.data()));
    double fw[2], ab[2];
    const double z2[2] = {0, 0}, ps[2] = {0, 90};
    OK(vie_batch_outputs(c, 2, z2, z2, ps, 2, fw, ab));
    EXPECT(std::isfinite(fw[0]) && std::isfinite(ab[1]));
    g_ctx = nullptr;
    OK(vie_destroy(c));
}

static void bad_arguments()
{
    vie_desc d{};
    d.n[0] = d.n[1] = d.n[2] = 8;
    d.d[0] = d.d[1] = d.d[2] = 0.04;
    d.k0 = 1.0;
    d.precision = VIE_FP64;
    vie_ctx* c = nullptr;
    EXPECT(vie_create(nullptr, &d) == VIE_ERR_ARG);
    d.n[1] = 300;
    EXPECT(vie_create(&c, &d) == VIE_ERR_UNSUPPORTED);
    d.n[1] = 8;
    d.k0 = -1.0;
    EXPECT(vie_create(&c, &d) == VIE_ERR_ARG);
    EXPECT(vie_supported_n(128) == 1);
    EXPECT(vie_last_error(nullptr) != nullptr);
    d.k0 = 1.0;
    OK(vie_create(&c, &d));
    std::vector<double> E(6 * 512), x(6 * 512);
    vie_info info{};
    EXPECT(vie_solve(c, VIE_BICGSTAB, 1e-6, 0, E.data(), x.data(), &info) == VIE_ERR_ARG);  // no medium yet
    EXPECT(vie_green_compute(nullptr) == VIE_ERR_ARG);
    OK(vie_destroy(c));
}

int main()
{
    bad_arguments();
    if (!g_fail) single_context(16, VIE_FP64);
    if (!g_fail) single_context(16, VIE_FP32);
    if (!g_fail) embedded_context();
    if (!g_fail) sequences();
    size_t rel = 0;
    if (!g_fail && vie_release_buffers(&rel) != VIE_OK) g_fail = 1;
    std::printf(g_fail ? "abi_driver FAILED\n" : "abi_driver ok\n");
    std::fflush(stdout);
    // Every context is destroyed by now.  Leave without the static teardown:
    // the HIP runtime's own destructors free memory through the sanitizer's
    // quarantine after the HSA runtime has gone, and the sanitizer's device
    // allocator then fails a CHECK of its own now and then.
    std::_Exit(g_fail);
}
