// ref_harness.cpp — exposes the reference's own hot-path templates through a C ABI so the
// oracle restatement can be checked against them.  TEST INFRASTRUCTURE ONLY: compiled by
// `make ref` from the headers where they lie under the reference tree into _ref/ (git-
// ignored); nothing in the product loads it.  Only system headers are added here.
#include <cstddef>
#include <cstdint>
#include <cmath>
#include <cstdio>
#include <vector>
#include "GridDifference8th.h"
#include "GramSchmidt.h"
#include "GramSchmidt_mpi.h"

extern "C" {

// Laplasian8th<double> (GridDifference8th.h:177-185): out += coef * Lap8 p, periodic.
void ref_laplacian8(int nx, int ny, int nz, double dx, double dy, double dz, double coef,
                    const double* p, double* out) {
    GridRange g{0, 0, 0, nx, ny, nz};
    Laplasian8th<double>(g, out, p, coef, dx, dy, dz);
}

// Laplasian8th<float>: the template instantiated at float, as the reference would run it.
void ref_laplacian8_f32(int nx, int ny, int nz, double dx, double dy, double dz, double coef,
                        const float* p, float* out) {
    GridRange g{0, 0, 0, nx, ny, nz};
    Laplasian8th<float>(g, out, p, coef, dx, dy, dz);
}

// GramSchmidt (GramSchmidt.h:6-38): num bands of ng points, band-major, in place, with the
// reference's own wave-function wrappers referring to the caller's storage.
void ref_gram_schmidt(int ng, int num, double* psi, double dvol) {
    std::vector<RspaceFunc<double>> set(num);
    for (int n = 0; n < num; ++n) set[n].Refer(psi + (size_t)n * ng);
    GramSchmidt(ng, num, set.data(), dvol);
}

}
