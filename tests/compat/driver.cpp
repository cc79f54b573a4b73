// Test driver for include/vie_compat.hpp: a minimal stand-in for the
// reference's program environment (the MKL complex type, the contiguous 3-D
// allocators of vie_fft/common.h:68-97 and the globals of vie_fft/Model.h),
// written for this test.  It reads a case file, runs the reference's call
// sequence (Green_Function_Compute, Incident_Wave, FFT_Program) and writes
// Current_X/Y/Z.
//
//   driver <case.bin> <out.bin>
//   driver op <case.bin> <x.bin> <out.bin>
// case.bin: int32 n[3]; float64 d[3], k0, theta, phi, psi; int32 nmat;
//           float64 eps[2 nmat]; uint8 material[V]; uint8 flags[3V]
// out.bin:  float64 J[6V] (Current_X | Current_Y | Current_Z, interleaved)
// The op mode reads a field S (x.bin: 3V complex) after Green_Function_Compute
// and writes, each 3V complex: X_FFT/Y_FFT/Z_FFT of S into zeroed fields,
// the same after Current_Reset, and the _Conjugate set of S.
#include <cstdio>
#include <cstdlib>
#include <cstring>

struct MKL_Complex16 {
    double real;
    double imag;
};

template <class T> T*** Memory3D(T*** a, const int m, const int n, const int l)
{
    a = (T***)std::malloc(sizeof(T**) * m);
    a[0] = (T**)std::malloc(sizeof(T*) * m * n);
    a[0][0] = (T*)std::calloc((size_t)m * n * l, sizeof(T));
    for (int i = 0; i < m; ++i) {
        a[i] = a[0] + (long)i * n;
        for (int j = 0; j < n; ++j) a[i][j] = a[0][0] + ((long)i * n + j) * l;
    }
    return a;
}

template <class T> T*** RMemory3D(T*** a)
{
    if (a) {
        std::free(a[0][0]);
        std::free(a[0]);
        std::free(a);
    }
    return nullptr;
}

MKL_Complex16*** Current_X;
MKL_Complex16*** Current_Y;
MKL_Complex16*** Current_Z;
MKL_Complex16*** Green_Function;
unsigned char*** Material;
MKL_Complex16*** Ex_inc;
MKL_Complex16*** Ey_inc;
MKL_Complex16*** Ez_inc;
unsigned char*** Flag_X;
unsigned char*** Flag_Y;
unsigned char*** Flag_Z;
double k0;
double delta_x, delta_y, delta_z;
int Nx, Ny, Nz;
double Theta = 0, Phi = 0, Psi = 0;
const int NUM_DIE = 8;
double Real_EPR[NUM_DIE];
double Imag_EPR[NUM_DIE];

#include "vie_compat.hpp"

static void rd(void* p, size_t n, FILE* f)
{
    if (std::fread(p, 1, n, f) != n) {
        std::fprintf(stderr, "short case file\n");
        std::exit(2);
    }
}

static void write3(FILE* o, MKL_Complex16*** a, MKL_Complex16*** b, MKL_Complex16*** c, long V)
{
    MKL_Complex16*** J[3] = {a, b, c};
    for (int s = 0; s < 3; ++s) std::fwrite(J[s][0][0], sizeof(MKL_Complex16), V, o);
}

int main(int argc, char** argv)
{
    const bool op = argc == 5 && std::strcmp(argv[1], "op") == 0;
    if (argc != 3 && !op) {
        std::fprintf(stderr, "usage: driver case.bin out.bin | driver op case.bin x.bin out.bin\n");
        return 2;
    }
    if (op) ++argv;
    FILE* f = std::fopen(argv[1], "rb");
    if (!f) return 2;
    int n[3], nmat;
    double hd[7];
    rd(n, sizeof n, f);
    rd(hd, sizeof hd, f);
    rd(&nmat, sizeof nmat, f);
    if (nmat < 1 || nmat > NUM_DIE) return 2;
    Nx = n[0];
    Ny = n[1];
    Nz = n[2];
    delta_x = hd[0];
    delta_y = hd[1];
    delta_z = hd[2];
    k0 = hd[3];
    Theta = hd[4];
    Phi = hd[5];
    Psi = hd[6];
    for (int m = 0; m < NUM_DIE; ++m) Real_EPR[m] = 1.1, Imag_EPR[m] = 0.0;
    for (int m = 0; m < nmat; ++m) {
        rd(&Real_EPR[m], sizeof(double), f);
        rd(&Imag_EPR[m], sizeof(double), f);
    }
    const long V = (long)Nx * Ny * Nz;
    // the allocations of Parameter_Input (vie_fft/Model.h:166-180)
    Current_X = Memory3D(Current_X, Nx, Ny, Nz);
    Current_Y = Memory3D(Current_Y, Nx, Ny, Nz);
    Current_Z = Memory3D(Current_Z, Nx, Ny, Nz);
    Green_Function = Memory3D(Green_Function, 2 * Nx, 2 * Ny, 2 * Nz);
    Material = Memory3D(Material, Nx, Ny, Nz);
    Ex_inc = Memory3D(Ex_inc, Nx, Ny, Nz);
    Ey_inc = Memory3D(Ey_inc, Nx, Ny, Nz);
    Ez_inc = Memory3D(Ez_inc, Nx, Ny, Nz);
    Flag_X = Memory3D(Flag_X, Nx, Ny, Nz);
    Flag_Y = Memory3D(Flag_Y, Nx, Ny, Nz);
    Flag_Z = Memory3D(Flag_Z, Nx, Ny, Nz);
    rd(Material[0][0], V, f);  // GetMaterial
    rd(Flag_X[0][0], V, f);    // GetFlag
    rd(Flag_Y[0][0], V, f);
    rd(Flag_Z[0][0], V, f);
    std::fclose(f);

    Green_Function_Compute();
    if (op) {
        // the operator calls of vie_fft/CG_Iteration.h:86-173 on a given field
        MKL_Complex16*** S[3] = {nullptr, nullptr, nullptr};
        MKL_Complex16*** F[3] = {nullptr, nullptr, nullptr};
        FILE* xf = std::fopen(argv[2], "rb");
        if (!xf) return 2;
        for (int s = 0; s < 3; ++s) {
            S[s] = Memory3D(S[s], Nx, Ny, Nz);
            F[s] = Memory3D(F[s], Nx, Ny, Nz);
            rd(S[s][0][0], sizeof(MKL_Complex16) * V, xf);
        }
        std::fclose(xf);
        const int Dim[3] = {Nx, Ny, Nz};
        MKL_Complex16*** temp = Memory3D(temp, 2 * Nx, 2 * Ny, 2 * Nz);
        X_FFT(F[0], S[0], S[1], S[2], temp, Green_Function, Dim);
        Y_FFT(F[1], S[0], S[1], S[2], temp, Green_Function, Dim);
        Z_FFT(F[2], S[0], S[1], S[2], temp, Green_Function, Dim);
        FILE* o = std::fopen(argv[3], "wb");
        if (!o) return 2;
        write3(o, F[0], F[1], F[2], V);
        for (int s = 0; s < 3; ++s) Current_Reset(F[s], s, Dim);
        write3(o, F[0], F[1], F[2], V);
        for (int s = 0; s < 3; ++s) std::memset(F[s][0][0], 0, sizeof(MKL_Complex16) * V);
        X_FFT_Conjugate(F[0], S[0], S[1], S[2], temp, Green_Function, Dim);
        Y_FFT_Conjugate(F[1], S[0], S[1], S[2], temp, Green_Function, Dim);
        Z_FFT_Conjugate(F[2], S[0], S[1], S[2], temp, Green_Function, Dim);
        write3(o, F[0], F[1], F[2], V);
        std::fclose(o);
        for (int s = 0; s < 3; ++s) {
            RMemory3D(S[s]);
            RMemory3D(F[s]);
        }
        RMemory3D(temp);
        vie_destroy(vie_compat::context());
        return 0;
    }
    Incident_Wave();
    FFT_Program();

    FILE* o = std::fopen(argv[2], "wb");
    if (!o) return 2;
    write3(o, Current_X, Current_Y, Current_Z, V);
    std::fclose(o);
    RMemory3D(Current_X);
    RMemory3D(Current_Y);
    RMemory3D(Current_Z);
    RMemory3D(Material);
    vie_destroy(vie_compat::context());
    return 0;
}
