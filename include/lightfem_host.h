/* lightfem_host.h - C ABI of the host-side stages that stay on the CPU by design
 * (their outputs are integer structures the device kernels consume). Each one is
 * bit-identical in value to the reference stage cited beside it; they are also
 * used by lightfem::LinearElasticSolver (liblightfem.so) and by ctypes callers.
 * Node indices are positions in the model's node order (node i owns dofs 2i,
 * 2i+1, src/model.cpp:74-92). Return 0 on success, non-zero on invalid input.
 */
#ifndef LIGHTFEM_HOST_H
#define LIGHTFEM_HOST_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Number of CSR nonzeros build_csr_assembly_pattern (sparse.cpp:15-58) produces
 * for Q4 elements given as 4 node indices each; -1 on invalid input. */
int64_t lfh_csr_pattern_nnz(int64_t n_nodes, int64_t n_elems, const int32_t* enodes);

/* The pattern itself: row_ptr (2*n_nodes+1), sorted columns (nnz) and the
 * element->value map (64 per element, local entry r*8+c), sparse.cpp:15-58. */
int lfh_csr_pattern(int64_t n_nodes, int64_t n_elems, const int32_t* enodes, int32_t* row_ptr, int32_t* col,
                    int32_t* emap);

/* The same with bars (2 node indices each in bnodes, after the Q4 elements in the
 * element order of solver.cpp:59-83): bmap gets 16 entries per bar (local r*4+c). */
int64_t lfh_csr_pattern_nnz_bars(int64_t n_nodes, int64_t n_elems, const int32_t* enodes, int64_t n_bars,
                                 const int32_t* bnodes);
int lfh_csr_pattern_bars(int64_t n_nodes, int64_t n_elems, const int32_t* enodes, int64_t n_bars,
                         const int32_t* bnodes, int32_t* row_ptr, int32_t* col, int32_t* emap, int32_t* bmap);

/* reverse_cuthill_mckee (reordering.cpp:11-62): node ids in RCM order. node_ids
 * lists the model's nodes; enode_ids the 4 node ids of each Q4 element. */
int lfh_rcm(int64_t n_nodes, const int32_t* node_ids, int64_t n_elems, const int32_t* enode_ids, int32_t* order_out);
/* ... with bars (2 node ids each): their node pairs join the graph (reordering.cpp:25-28). */
int lfh_rcm_bars(int64_t n_nodes, const int32_t* node_ids, int64_t n_elems, const int32_t* enode_ids,
                 int64_t n_bars, const int32_t* bnode_ids, int32_t* order_out);

/* element_colors (solver.cpp:100-133): greedy first-fit colour per element in
 * element order (elements sharing a node conflict). Returns the colour count via
 * n_colors_out; color_ptr/color_elems (n_colors+1 / n_elems, ascending elements
 * per colour) are filled when non-NULL and sized by the caller (color_ptr with
 * room for n_elems+1 entries suffices). */
int lfh_element_colors(int64_t n_nodes, int64_t n_elems, const int32_t* enodes, int32_t* color_out,
                       int32_t* n_colors_out, int32_t* color_ptr, int32_t* color_elems);
/* ... over the Q4 elements then the bars (element n_elems + t is bar t). */
int lfh_element_colors_bars(int64_t n_nodes, int64_t n_elems, const int32_t* enodes, int64_t n_bars,
                            const int32_t* bnodes, int32_t* color_out, int32_t* n_colors_out, int32_t* color_ptr,
                            int32_t* color_elems);

/* Subdomains of the `block` preconditioner: build_dof_subdomain's node split
 * (solver.cpp:631-654) into node_subdomain (n_nodes), with split 0 = "rcb" (recursive
 * coordinate bisection of the node positions xy[2i], xy[2i+1], precond.cpp:30-105)
 * or 1 = "chunk" (node index / nodes_per_subdomain, precond.cpp:19-28). The count
 * goes to n_subdomains_out. */
int lfh_subdomain_split(int64_t n_nodes, const double* xy, int32_t split, int64_t nodes_per_subdomain,
                        int32_t* node_subdomain, int64_t* n_subdomains_out);

/* csr_profile (sparse.cpp:227-242): out = {nonzeros, bandwidth, profile}. */
int lfh_csr_profile(int64_t n_rows, const int32_t* row_ptr, const int32_t* col, int64_t out[3]);

/* The node-level schedule of the device IC(0) sweeps (lfg_pcg with ic0 on a node-blocked
 * pattern) for the IC(0) factor of the CSR pattern row_ptr/col (n_dofs rows): no
 * counterpart in the reference (whose sweeps run row by row, precond.cpp:386-412); for
 * inspection and tests of the schedule the device waits on. backward 0: the forward
 * sweep over L's rows, 1: the backward sweep over its transposed columns. Fills
 * ord[n_dofs / 2] (position -> node), qptr[n_dofs / 2 + 1] and key[< key_cap] (each
 * position's dependency positions), parts[< n_dofs / 2 + 2] (part boundaries: levels
 * split into parts of at most 64 nodes; *n_parts boundaries written) and *reach (the
 * largest distance back a position reads). Returns 0; 1 for invalid arguments; 2 when
 * the factor's rows 2i, 2i+1 do not share their dependency nodes (the device then uses
 * the row form); 3 when key_cap is too small. */
int lfh_ic0_node_schedule(int64_t n_dofs, const int32_t* row_ptr, const int32_t* col, int32_t backward,
                          int32_t* ord, int32_t* qptr, int32_t* key, int64_t key_cap, int32_t* parts,
                          int64_t* n_parts, int32_t* reach);

#ifdef __cplusplus
}
#endif

#endif /* LIGHTFEM_HOST_H */
