// Launch wrappers of the gfx950 kernels (host-callable, stream-ordered).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "lfg_common.hpp"

namespace lfg {

struct MeshView {
    int n_nodes;
    int n_elems;
    const double* xy;        // 2 per node
    const int* enodes;       // 4 node indices per element
    const int* elem_mat;     // material index per element
    const double* materials; // E, nu, t per material
    int plane_mode;
};

struct CsrView {
    int n;
    long long nnz;
    const int* row_ptr;
    const int* col;
    int max_row;  // longest row (entries)
};

// ---- element kernels (lfg_element.hip)
enum KeMode { KE_STORE = 0, KE_ATOMIC = 1, KE_COLORED = 2 };
// Ke for elements [0, count) of `elems` (or all when elems == nullptr). STORE writes
// ke_out[64e + 8r + c]; ATOMIC/COLORED scatter into values through emap.
void launch_ke(hipStream_t s, const MeshView& m, KeMode mode, const int* elems, int count, const int* emap,
               double* values, double* ke_out, int* jacobian_failures);
// values[k] = sum of ke[contrib[j]] for j in [cptr[k], cptr[k+1]) in element order.
void launch_gather(hipStream_t s, long long nnz, const int* cptr, const int* contrib, const double* ke,
                   double* values);
void launch_fill(hipStream_t s, double* p, long long n, double v);
void launch_fields(hipStream_t s, const MeshView& m, const double* u, double* centre, double* gp);
void launch_smooth(hipStream_t s, int n_nodes, const int* node_elem_ptr, const int* node_elem, const double* gp,
                   double* nodal);

// ---- solver kernels (lfg_solver.hip)
void launch_dirichlet(hipStream_t s, const CsrView& a, const double* values, const double* f, const int* constrained,
                      const double* prescribed, double* kc, double* fc);
void launch_spmv_serial_rows(hipStream_t s, const CsrView& a, const double* values, const double* x, const double* sub,
                             double* y);
void launch_extract_diag(hipStream_t s, const CsrView& a, const double* values, double* diag, int* status);
void launch_block2_factor(hipStream_t s, const CsrView& a, const double* values, double* factor, int* degenerate);
void launch_jacobi_inv(hipStream_t s, int n, const double* diag, double* inv, int* status);
void launch_precond_apply(hipStream_t s, int kind, int n, const double* m, const double* in, double* out);
void launch_to_f32(hipStream_t s, const double* in, float* out, long long n);

// PCG building blocks; every kernel returns at once when st->done is set.
struct PcgBuffers {
    int n;
    double* x;
    double* r;
    double* p;
    double* q;
    double* z;          // IC(0) output (others recompute z)
    const double* b;
    const double* m;    // Jacobi inverse diagonal (n) or 2x2 factors (3 per node)
    int precond;        // LFG_PRECOND_*
    PcgState* st;
    TicketSet* tickets; // 0 init, 1 SpMV, 2 update, 3 replacement
    double* partials;   // >= 2 * (max blocks + 8)
};

int spmv_blocks(int n, int variant);
void launch_pcg_init(hipStream_t s, const PcgBuffers& b, int stop_rule, double tol, long long max_iter, double drop,
                     double mixed_switch);
// q = A p; finishes p.q and alpha. Without vals32 the FP64 values are used; with
// both, the kernel reads the FP32 values once the state is in the FP32 phase.
void launch_pcg_spmv(hipStream_t s, const PcgBuffers& b, const CsrView& a, const double* vals64, const float* vals32,
                     int variant);
// x += alpha p, r -= alpha q, ||r||^2, r.z; convergence / replacement decision, beta.
void launch_pcg_update(hipStream_t s, const PcgBuffers& b, int mixed);
// mixed only: r = b - K x in FP64 when a replacement is due; then r.z and beta.
void launch_pcg_replace(hipStream_t s, const PcgBuffers& b, const CsrView& a, const double* vals64);
// p = z + beta p.
void launch_pcg_direction(hipStream_t s, const PcgBuffers& b);
// IC(0): z = (L L^T)^-1 r by level-scheduled sweeps, then r.z and beta (one block).
struct Ic0View {
    int n;
    const int* lrow;     // lower factor rows (CSR, diagonal last), strict part used
    const int* lcol;
    const double* lval;
    const double* inv_diag;
    const int* urow;     // transpose of the strict lower part, rows descending per column
    const int* ucol;
    const double* uval;
    const int* fwd_level_ptr;
    const int* fwd_rows;
    int fwd_levels;
    const int* bwd_level_ptr;
    const int* bwd_rows;
    int bwd_levels;
};
void launch_ic0_apply(hipStream_t s, const Ic0View& f, const double* in, double* out);
void launch_pcg_ic0_step(hipStream_t s, const PcgBuffers& b, const Ic0View& f);
// ||b - K x|| in FP64 (reporting).
void launch_residual_norm(hipStream_t s, const CsrView& a, const double* vals64, const double* x, const double* bvec,
                          double* partials, TicketSet* tickets, double* out2);

}  // namespace lfg
