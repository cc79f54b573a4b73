/*
 * vie_compat.hpp -- source-level drop-in for the reference's GreenFunction.h,
 * CG_Iteration.h and fft.h (paths relative to the reference root).
 *
 * A driver written against the reference (vie_fft/CG_FFT_VIE.cpp:52-61) keeps
 * its calls
 *     GetMaterial(); GetFlag(); Green_Function_Compute(); Incident_Wave();
 *     FFT_Program(); FarField1(); FarField2(); NearField();
 * and code that applies the operator itself keeps X_FFT / Y_FFT / Z_FFT, the
 * _Conjugate set and Current_Reset (vie_fft/CG_Iteration.h:86-173,
 * vie_fft/fft.h:13-35) with the reference's signatures.
 * and replaces the three includes
 *     #include "GreenFunction.h"  #include "CG_Iteration.h"  (fft.h comes with it)
 * by
 *     #include "vie_compat.hpp"
 * placed after Model.h; link with -lviefft (libviefft.so, gfx950) and
 * -lamdhip64.  NearFar.h stays the reference's own: it reads Current_X/Y/Z,
 * which FFT_Program fills here.
 *
 * The header uses the reference's globals exactly as vie_fft/Model.h:13-36,
 * 44-46 and 77-83 declare them (Current_X/Y/Z, Green_Function, Material,
 * Ex_inc/Ey_inc/Ez_inc, Flag_X/Y/Z, k0, delta_x/y/z, Nx/Ny/Nz, Theta, Phi,
 * Psi, NUM_DIE, Real_EPR, Imag_EPR) and RMemory3D (vie_fft/common.h:97), and
 * requires the arrays to be contiguous from a[0][0] as Memory3D allocates them
 * (vie_fft/common.h:68-94).
 *
 * Behaviour kept: CG_THR and Iteration_Type as in vie_fft/CG_Iteration.h:46-47
 * (override with -DVIE_COMPAT_CG_THR=... / -DVIE_COMPAT_ITERATION_TYPE=...),
 * the stop rules of the five solvers, maxit = Nx*Ny*Nz, the "CG Error" /
 * "Steps" lines, and FFT_Program freeing Ex/Ey/Ez_inc, Green_Function and
 * Flag_X/Y/Z (vie_fft/CG_Iteration.h:769-775).  Green_Function is not filled:
 * G~ lives on the GPU.  Errors print the library's message and exit(1), as
 * the reference has no error path.
 */
#ifndef VIE_COMPAT_HPP
#define VIE_COMPAT_HPP

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "viefft.h"

#ifndef VIE_COMPAT_CG_THR
#define VIE_COMPAT_CG_THR 0.001
#endif
#ifndef VIE_COMPAT_ITERATION_TYPE
#define VIE_COMPAT_ITERATION_TYPE 3
#endif

const double CG_THR = VIE_COMPAT_CG_THR;              // vie_fft/CG_Iteration.h:46
const int Iteration_Type = VIE_COMPAT_ITERATION_TYPE;  // vie_fft/CG_Iteration.h:47

namespace vie_compat {

inline vie_ctx*& context()
{
    static vie_ctx* c = nullptr;
    return c;
}

inline void check(int st, const char* what)
{
    if (st == VIE_OK || st == VIE_NOT_CONVERGED) return;
    const char* msg = context() ? vie_last_error(context()) : "";
    std::fprintf(stderr, "%s failed (status %d): %s\n", what, st, msg ? msg : "");
    std::exit(1);
}

inline long volume() { return (long)Nx * Ny * Nz; }

// three N^3 complex arrays -> one 3V interleaved buffer (x | y | z)
template <class C>
inline void gather3(C*** a, C*** b, C*** c, std::vector<double>& out)
{
    const long V = volume();
    out.resize(6 * V);
    C*** src[3] = {a, b, c};
    for (int s = 0; s < 3; ++s) {
        const C* p = src[s][0][0];
        for (long v = 0; v < V; ++v) {
            out[2 * (s * V + v)] = p[v].real;
            out[2 * (s * V + v) + 1] = p[v].imag;
        }
    }
}

template <class C>
inline void scatter3(const std::vector<double>& in, C*** a, C*** b, C*** c)
{
    const long V = volume();
    C*** dst[3] = {a, b, c};
    for (int s = 0; s < 3; ++s) {
        C* p = dst[s][0][0];
        for (long v = 0; v < V; ++v) {
            p[v].real = in[2 * (s * V + v)];
            p[v].imag = in[2 * (s * V + v) + 1];
        }
    }
}

}  // namespace vie_compat

// Green_Function_Compute (vie_fft/GreenFunction.h:8-141): binds the context to
// this grid and wavenumber, uploads the medium (Material, the eps table and
// the three flag arrays) and computes G~ on the GPU.
inline void Green_Function_Compute()
{
    vie_ctx*& c = vie_compat::context();
    if (c) {
        vie_destroy(c);
        c = nullptr;
    }
    vie_desc desc;
    desc.n[0] = Nx;
    desc.n[1] = Ny;
    desc.n[2] = Nz;
    desc.d[0] = delta_x;
    desc.d[1] = delta_y;
    desc.d[2] = delta_z;
    desc.k0 = k0;
    desc.precision = VIE_FP64;
    vie_compat::check(vie_create(&c, &desc), "vie_create");
    const long V = vie_compat::volume();
    std::vector<double> eps(2 * NUM_DIE);
    for (int m = 0; m < NUM_DIE; ++m) {
        eps[2 * m] = Real_EPR[m];
        eps[2 * m + 1] = Imag_EPR[m];
    }
    std::vector<unsigned char> flags(3 * V);
    std::memcpy(&flags[0], Flag_X[0][0], V);
    std::memcpy(&flags[V], Flag_Y[0][0], V);
    std::memcpy(&flags[2 * V], Flag_Z[0][0], V);
    vie_compat::check(vie_set_medium(c, Material[0][0], eps.data(), NUM_DIE, flags.data()), "vie_set_medium");
    vie_compat::check(vie_green_compute(c), "vie_green_compute");
}

// Incident_Wave (vie_fft/GreenFunction.h:144-248): Ex_inc, Ey_inc, Ez_inc.
inline void Incident_Wave()
{
    std::vector<double> E(6 * vie_compat::volume());
    vie_compat::check(vie_incident(vie_compat::context(), Theta, Phi, Psi, E.data()), "vie_incident");
    vie_compat::scatter3(E, Ex_inc, Ey_inc, Ez_inc);
}

namespace vie_compat {

// Field += component f of A S or A^H S (no output mask), S = (Sx, Sy, Sz):
// the whole product runs on the GPU and component f is added on the host.
inline void add_component(int f, int adjoint, MKL_Complex16*** Field, MKL_Complex16*** const Sx,
                          MKL_Complex16*** const Sy, MKL_Complex16*** const Sz, const int* Dim)
{
    if (!context()) {
        std::fprintf(stderr, "X_FFT/Y_FFT/Z_FFT before Green_Function_Compute\n");
        std::exit(1);
    }
    if (Dim[0] != Nx || Dim[1] != Ny || Dim[2] != Nz) {
        std::fprintf(stderr, "Dim (%d, %d, %d) is not the grid (%d, %d, %d)\n", Dim[0], Dim[1], Dim[2], Nx, Ny, Nz);
        std::exit(1);
    }
    const long V = volume();
    std::vector<double> x, y(6 * V);
    gather3(Sx, Sy, Sz, x);
    check(vie_apply_raw(context(), x.data(), y.data(), adjoint), "vie_apply_raw");
    MKL_Complex16* p = Field[0][0];
    for (long v = 0; v < V; ++v) {
        p[v].real += y[2 * (f * V + v)];
        p[v].imag += y[2 * (f * V + v) + 1];
    }
}

}  // namespace vie_compat

// The operator calls of vie_fft/CG_Iteration.h:86-173 with their signatures:
// Field += component x, y or z of A S (X_FFT, Y_FFT, Z_FFT) or of A^H S (the
// _Conjugate set), S = (Sourx, Soury, Sourz), without Current_Reset, as in the
// reference.  temp and Green are not used: the padded transforms and G~ live
// on the GPU (Green_Function_Compute set them up).
inline void X_FFT(MKL_Complex16*** Field, MKL_Complex16*** const Sourx, MKL_Complex16*** const Soury,
                  MKL_Complex16*** const Sourz, MKL_Complex16*** temp, MKL_Complex16*** const Green, const int* Dim)
{
    (void)temp;
    (void)Green;
    vie_compat::add_component(0, 0, Field, Sourx, Soury, Sourz, Dim);
}
inline void Y_FFT(MKL_Complex16*** Field, MKL_Complex16*** const Sourx, MKL_Complex16*** const Soury,
                  MKL_Complex16*** const Sourz, MKL_Complex16*** temp, MKL_Complex16*** const Green, const int* Dim)
{
    (void)temp;
    (void)Green;
    vie_compat::add_component(1, 0, Field, Sourx, Soury, Sourz, Dim);
}
inline void Z_FFT(MKL_Complex16*** Field, MKL_Complex16*** const Sourx, MKL_Complex16*** const Soury,
                  MKL_Complex16*** const Sourz, MKL_Complex16*** temp, MKL_Complex16*** const Green, const int* Dim)
{
    (void)temp;
    (void)Green;
    vie_compat::add_component(2, 0, Field, Sourx, Soury, Sourz, Dim);
}
inline void X_FFT_Conjugate(MKL_Complex16*** Field, MKL_Complex16*** const Sourx, MKL_Complex16*** const Soury,
                            MKL_Complex16*** const Sourz, MKL_Complex16*** temp, MKL_Complex16*** const Green,
                            const int* Dim)
{
    (void)temp;
    (void)Green;
    vie_compat::add_component(0, 1, Field, Sourx, Soury, Sourz, Dim);
}
inline void Y_FFT_Conjugate(MKL_Complex16*** Field, MKL_Complex16*** const Sourx, MKL_Complex16*** const Soury,
                            MKL_Complex16*** const Sourz, MKL_Complex16*** temp, MKL_Complex16*** const Green,
                            const int* Dim)
{
    (void)temp;
    (void)Green;
    vie_compat::add_component(1, 1, Field, Sourx, Soury, Sourz, Dim);
}
inline void Z_FFT_Conjugate(MKL_Complex16*** Field, MKL_Complex16*** const Sourx, MKL_Complex16*** const Soury,
                            MKL_Complex16*** const Sourz, MKL_Complex16*** temp, MKL_Complex16*** const Green,
                            const int* Dim)
{
    (void)temp;
    (void)Green;
    vie_compat::add_component(2, 1, Field, Sourx, Soury, Sourz, Dim);
}

// Current_Reset (vie_fft/fft.h:13-35): component Flag of a field times its
// face flags (Flag_X, Flag_Y, Flag_Z).
inline void Current_Reset(MKL_Complex16*** Current, const int Flag, const int* Dim)
{
    unsigned char*** fl = Flag == 0 ? Flag_X : Flag == 1 ? Flag_Y : Flag_Z;
    const long V = (long)Dim[0] * Dim[1] * Dim[2];
    MKL_Complex16* p = Current[0][0];
    const unsigned char* f = fl[0][0];
    for (long v = 0; v < V; ++v)
        if (!f[v]) p[v].real = p[v].imag = 0.0;
}

// FFT_Program (vie_fft/CG_Iteration.h:1365-1388) with its solvers' output and
// frees (vie_fft/CG_Iteration.h:735-775).
inline void FFT_Program()
{
    std::vector<double> E, J(6 * vie_compat::volume(), 0.0);
    vie_compat::gather3(Ex_inc, Ey_inc, Ez_inc, E);
    vie_info info;
    const int st = vie_solve(vie_compat::context(), Iteration_Type, CG_THR, 0, E.data(), J.data(), &info);
    vie_compat::check(st, "vie_solve");
    std::printf("CG Error: %lf\n", info.err);
    std::printf("Steps: %d\n\n", info.iterations - 1);
    vie_compat::scatter3(J, Current_X, Current_Y, Current_Z);
    RMemory3D(Ex_inc);
    RMemory3D(Ey_inc);
    RMemory3D(Ez_inc);
    RMemory3D(Green_Function);
    RMemory3D(Flag_X);
    RMemory3D(Flag_Y);
    RMemory3D(Flag_Z);
}

#endif
