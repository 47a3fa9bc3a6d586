/* Test infrastructure only: CPU restatement ("oracle") of the LightFEM Q4
 * assembly + PCG path. It is the checker for the HIP library in tests/, for
 * __graft_entry__.smoke() and for bench.py's CPU baseline; the product never
 * links it.
 *
 * Every function restates one reference stage literally, in the same loop and
 * summation order, so that its results are bit-identical to the compiled
 * reference (the golden vectors under tests/golden/ pin that). Indices are
 * int64 throughout; ids are the model's node ids.
 *
 * Pinning (what checks each function; see tests/test_oracle.py):
 *   lfo_q4_stiffness       golden ke (bit-exact), symmetry, rigid modes, patch test,
 *                          closed-form rectangle, linearity in E
 *   lfo_q4_fields          golden centre / Gauss-point fields (bit-exact), patch test
 *   lfo_pattern_*          golden row_ptr/col/emap (exact), nnz = 4(3n+1)^2
 *   lfo_rcm                golden RCM order (exact), 200x200 bandwidth/profile from README
 *   lfo_colors             golden colours (exact), conflict-freedom
 *   lfo_assemble_serial    golden K (bit-exact)
 *   lfo_dirichlet          golden Kc/fc (bit-exact), dense brute-force restatement
 *   lfo_pcg                golden u and iteration counts (bit-exact), scipy direct solve
 *   lfo_ic0_* / block      golden preconditioner applications (bit-exact)
 *   lfo_bsr_spmv           golden block-storage products (bit-exact) and golden
 *                          BSR-storage PCG solutions and iteration counts
 *   lfo_split_chunks       golden chunk subdomain maps (exact)
 *   lfo_block_*            golden subdomain block-Jacobi applications (bit-exact),
 *                          golden block-preconditioned solutions and iteration counts
 *   lfo_*_var, lfo_truss_*, lfo_assemble_flat, lfo_pin_zero_pivots, lfo_smooth_nodal_ex
 *                          golden bar-model patterns, maps, colours, K, u, R and fields
 *                          (bit-exact), bar statics (u = FL/EA, N = P/(2 sin 45))
 *   lfo_smooth_nodal       golden nodal fields (bit-exact)
 */
#ifndef LFO_H
#define LFO_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { LFO_PLANE_STRESS = 0, LFO_PLANE_STRAIN = 1 };
enum { LFO_PRECOND_JACOBI = 0, LFO_PRECOND_BLOCK1 = 1, LFO_PRECOND_IC0 = 2 };
enum { LFO_STOP_REFERENCE = 0, LFO_STOP_RELATIVE = 1 };
enum { LFO_PRECOND_BLOCK = 3 };  /* SubdomainBlockJacobi on a given dof -> subdomain map */
enum { LFO_STORAGE_CSR = 0, LFO_STORAGE_BSR = 1 };

/* 0 on success, 1 on a non-positive Jacobian. */
int lfo_q4_stiffness(const double xy[8], double young, double nu, double thickness, int plane_mode,
                     double ke[64]);
int lfo_q4_fields(const double xy[8], double young, double nu, double thickness, int plane_mode,
                  const double u[8], double centre[7], double gp[36]);

int64_t lfo_pattern_nnz(int64_t ndofs, int64_t nelem, const int64_t* edofs);
int lfo_pattern_build(int64_t ndofs, int64_t nelem, const int64_t* edofs, int64_t* row_ptr,
                      int64_t* col, int64_t* emap);

int lfo_rcm(int64_t nnodes, const int64_t* node_ids, int64_t nelem, const int64_t* enode_ids,
            int64_t* order_out);
int64_t lfo_colors(int64_t nelem, const int64_t* enode_ids, int64_t* color_out);
void lfo_csr_profile(int64_t n, const int64_t* row_ptr, const int64_t* col, int64_t out[3]);

void lfo_assemble_serial(int64_t nelem, const double* ke, const int64_t* emap, int64_t nnz,
                         double* values);
void lfo_dirichlet(int64_t n, const int64_t* row_ptr, const int64_t* col, double* values, double* rhs,
                   int64_t nc, const int64_t* cdofs, const double* cvals);
void lfo_spmv(int64_t n, const int64_t* row_ptr, const int64_t* col, const double* values,
              const double* x, double* y);

/* IC(0) factor of the lower triangle; returns breakdown count, -1 on error. */
int64_t lfo_ic0_factor_size(int64_t n, const int64_t* row_ptr, const int64_t* col);
int64_t lfo_ic0_factor(int64_t n, const int64_t* row_ptr, const int64_t* col, const double* values,
                       int64_t* lrow, int64_t* lcol, double* lval, double* inv_diag);
void lfo_ic0_apply(int64_t n, const int64_t* lrow, const int64_t* lcol, const double* lval,
                   const double* inv_diag, const double* in, double* out);
/* 2x2 nodal block-Jacobi (SubdomainBlockJacobi with one node per subdomain).
 * factor: 3 doubles per node [1/l00, l10, 1/l11]; returns degenerate count. */
int64_t lfo_block1_factor(int64_t n, const int64_t* row_ptr, const int64_t* col, const double* values,
                          double* factor);
void lfo_block1_apply(int64_t n, const double* factor, const double* in, double* out);

/* Bars (element.cpp:250-306, solver.cpp:71-83, 100-142, 753-871, 916-940). Element
 * lists of any length: element e owns edofs[eptr[e] .. eptr[e+1]) (Q4 8, bar 4) and
 * node ids enode_ids[nptr[e] .. nptr[e+1]) (Q4 4, bar 2), Q4 elements first. */
int64_t lfo_pattern_nnz_var(int64_t ndofs, int64_t nelem, const int64_t* eptr, const int64_t* edofs);
int lfo_pattern_build_var(int64_t ndofs, int64_t nelem, const int64_t* eptr, const int64_t* edofs,
                          int64_t* row_ptr, int64_t* col, int64_t* emap);
int64_t lfo_colors_var(int64_t nelem, const int64_t* nptr, const int64_t* enode_ids, int64_t* color_out);
void lfo_assemble_flat(int64_t total, const double* ke, const int64_t* emap, int64_t nnz, double* values);
int lfo_truss_stiffness(const double xy[4], double young, double area, double ke[16]);
int lfo_truss_fields(const double xy[4], double young, const double u[4], double out[7]);
int64_t lfo_pin_zero_pivots(int64_t n, const int64_t* row_ptr, const int64_t* col, double* values, double* rhs,
                            int64_t* bad);
void lfo_smooth_nodal_ex(int64_t nnodes, int64_t nelem, const int64_t* enode_idx, const double* gp, int64_t ntruss,
                         const int64_t* tnode_idx, const double* tstress, double* nodal);

/* Block storage product (BsrMatrix::from_csr + multiply_into, sparse.cpp:128-198). */
void lfo_bsr_spmv(int64_t n, const int64_t* row_ptr, const int64_t* col, const double* values, const double* x,
                  double* y);
int64_t lfo_bsr_blocks(int64_t n, const int64_t* row_ptr, const int64_t* col);

/* Subdomains: chunk split (precond.cpp:19-28); the coordinate bisection is the host
 * library's (pinned by the golden maps). SubdomainBlockJacobi on any dof map. */
int64_t lfo_split_chunks(int64_t n_nodes, int64_t nodes_per_subdomain, int64_t* node_subdomain);
typedef struct lfo_block lfo_block;
lfo_block* lfo_block_create(int64_t n, const int64_t* row_ptr, const int64_t* col, const double* values,
                            const int64_t* dof_subdomain, int64_t n_sub);
void lfo_block_apply(const lfo_block* b, const double* in, double* out);
void lfo_block_stats(const lfo_block* b, double out[3]); /* degenerate, factor entries, capture ratio */
void lfo_block_free(lfo_block* b);

typedef struct {
    int64_t iterations;
    int32_t converged;
    int32_t status; /* 0 ok, 1 non-positive initial rho, 2 non-positive curvature, 3 rho breakdown */
    double initial_residual;
    double final_residual;
} lfo_pcg_report;

int lfo_pcg(int64_t n, const int64_t* row_ptr, const int64_t* col, const double* values, const double* rhs,
            int precond, int stop_rule, double tolerance, int64_t max_iterations, double* x,
            lfo_pcg_report* report);
int lfo_pcg_ex(int64_t n, const int64_t* row_ptr, const int64_t* col, const double* values, const double* rhs,
               int precond, int storage, const int64_t* dof_subdomain, int64_t n_sub, int stop_rule,
               double tolerance, int64_t max_iterations, double* x, lfo_pcg_report* report);

/* Adjoint element sensitivities -lambda_e^T Ke(E=1) u_e * dE_e (no reference
 * counterpart; the derivative of q4_stiffness in E). props (E, nu, t) per element. */
int64_t lfo_sensitivity(int64_t nelem, const double* xy_nodes, const int64_t* enode_idx, const double* props,
                        int plane_mode, const double* u, const double* lam, const double* dE, double* grad);

/* Nodal smoothing (solver.cpp:807-871); node index = dof_x / 2. */
void lfo_smooth_nodal(int64_t nnodes, int64_t nelem, const int64_t* enode_idx, const double* gp,
                      double* nodal);

#ifdef __cplusplus
}
#endif

#endif
