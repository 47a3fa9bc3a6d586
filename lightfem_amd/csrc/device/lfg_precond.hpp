// Preconditioner applications shared by the PCG kernels (Jacobi and the 2x2 nodal
// block-Jacobi of precond.cpp with one node per subdomain).
#pragma once

#include "lfg_common.hpp"
#include "lightfem_gpu.h"

namespace lfg {

// inverse_diagonal_of (solver.cpp:656-669): 1 / d, or 1 for a diagonal of magnitude
// at most 1e-30.
__device__ __forceinline__ double jacobi_inverse(double d) { return fabs(d) > 1.0e-30 ? 1.0 / d : 1.0; }

// precond.cpp:266-287 for a size-2 block: forward then backward substitution with
// the packed factor (reciprocal pivots).
__device__ __forceinline__ void block2_solve(const double* f, double r0, double r1, double& z0, double& z1) {
    const double f0 = f[0], f1 = f[1], f2 = f[2];
    double s0 = r0 * f0;
    double s1 = r1;
    s1 -= f1 * s0;
    s1 = s1 * f2;
    s1 = s1 * f2;
    double t = s0;
    t -= f1 * s1;
    z0 = t * f0;
    z1 = s1;
}

// z = M^-1 r for node i (dofs 2i, 2i+1).
template <int Prec>
__device__ __forceinline__ void precond_node(const double* m, int node, double r0, double r1, double& z0, double& z1) {
    if (Prec == LFG_PRECOND_JACOBI) {
        const double2 d = reinterpret_cast<const double2*>(m)[node];
        z0 = d.x * r0;  // solver.cpp:502-505
        z1 = d.y * r1;
    } else {
        block2_solve(m + 3LL * node, r0, r1, z0, z1);
    }
}

// The preconditioner data of one node, loaded ahead of its use: the two Jacobi
// inverses, or the packed 2x2 factor.
template <int Prec>
struct NodePrecond {
    double a, b, c;
    __device__ __forceinline__ void load(const double* m, int node) {
        if (Prec == LFG_PRECOND_JACOBI) {
            const double2 d = reinterpret_cast<const double2*>(m)[node];
            a = d.x;
            b = d.y;
            c = 0.0;
        } else {
            a = m[3LL * node];
            b = m[3LL * node + 1];
            c = m[3LL * node + 2];
        }
    }
    // same arithmetic as precond_node
    __device__ __forceinline__ void apply(double r0, double r1, double& z0, double& z1) const {
        if (Prec == LFG_PRECOND_JACOBI) {
            z0 = a * r0;
            z1 = b * r1;
        } else {
            const double f[3] = {a, b, c};
            block2_solve(f, r0, r1, z0, z1);
        }
    }
};

}  // namespace lfg
