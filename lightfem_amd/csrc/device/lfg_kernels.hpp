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
    const int* mat_mode;     // plane mode per material (batches mixing modes), or null: plane_mode
};

// Bars (truss members) after the Q4 elements: bar t is element n_q4 + t.
struct BarView {
    int n_bars;
    int n_q4;
    const int* nodes;    // 2 node indices per bar
    const int* mat;      // material index per bar (E from MeshView::materials)
    const double* area;  // cross-section area per bar
    const int* map;      // 16 value indices per bar (row-major over its 4 dofs)
};

struct CsrView {
    int n;
    long long nnz;
    const int* row_ptr;
    const int* col;
    int max_row;  // longest row (entries)
};

// Node-slice layout of a node-blocked CSR matrix (rows 2i and 2i+1 share one
// column list made of dof pairs 2j, 2j+1, as every two-dof-per-node pattern is):
// nodes are cut into slices of 64 (one wavefront), and slice s stores width_s x 64
// 2x2 blocks column-major, entry k of node 64s + l at block ptr[s] + 64k + l, in
// the row's column order. Each block holds a node column and two value pairs,
// upper = (K[2i][2j], K[2i][2j+1]) in the first plane and lower =
// (K[2i+1][2j], K[2i+1][2j+1]) in the second; padding blocks are zero and point
// at their own node. One int32 column per four values, and every load of a
// wavefront is 64 consecutive blocks.
// The FP32 copy of the planes holds the off-diagonal blocks only (the diagonal block
// is zero there) and diag32 the diagonal blocks in FP64 (upper pair at 2i, lower at
// 2i + 1), each minus the rounding error of its node's FP32 off-diagonal blocks, so
// every row keeps its FP64 sum (see StencilView).
struct SliceView {
    int n_slices;
    long long blocks;  // padded blocks, ptr[n_slices]
    const int* ptr;    // n_slices + 1 block offsets (multiples of 64)
    const int* col;    // node column per block
    int max_width;
    const double* diag32;  // compensated diagonal of the FP32 copy
    int pair_sum;          // 1: block storage summation order (add_block)
};

// Node-stencil layout of a node-blocked matrix whose blocks lie at a few fixed node
// offsets, as on a structured mesh numbered row by row (node i couples with nodes
// i + d, d in a short list). Plane k holds for every node i the 2x2 block
// (i, i + off[k]) as an upper pair (K[2i][2j], K[2i][2j+1]) and a lower pair
// (K[2i+1][2j], K[2i+1][2j+1]); a block the pattern lacks, or whose column falls
// outside the mesh, is zero. Storage is per 64-node chunk (one wavefront): chunk c
// holds, plane after plane, the 64 upper pairs then the 64 lower pairs of its
// nodes, so a wavefront streams one contiguous run per chunk. No column index is
// stored, and every load of a wavefront moves 64 consecutive pairs.
//  - exact form: every offset of the pattern, ascending, so a row accumulates in the
//    CSR column order (bit-identical to the CSR row loop);
//  - symmetric form: off[0] = 0 and the positive offsets; block (i, i - d) is read as
//    the transpose of block (i - d, i), which halves the values streamed. Its FP32
//    copy keeps the diagonal plane in FP64 and gives it the rounding error of the
//    node's off-diagonal FP32 blocks with the opposite sign, so every row keeps its
//    FP64 sum and the FP32 operator acts exactly on nodal translations.
constexpr int kMaxStencil = 17;
struct StencilView {
    int n_nodes;
    int n_off;              // planes (exact: all offsets; symmetric: 0 and the positive ones)
    int sym;
    int pair_sum;           // exact form: 1 = block storage summation order (add_block)
    int off[kMaxStencil];   // ascending
    // Symmetric form, fused PCG iteration: the offsets -off[k], 0, off[k] grouped into
    // clusters of nearby offsets; a 256-node tile needs the search direction on one
    // window per cluster, [i0 + lo, i0 + 255 + hi], kept in LDS at win_base.
    int n_clusters;
    int win_lo[kMaxStencil];
    int win_hi[kMaxStencil];
    int win_base[kMaxStencil];
    int win_total;          // LDS pairs of all windows (0: no fused iteration)
    int slot_zero;          // LDS offset (plus the node's tile position t) of node i = i0 + t
    int slot_up[kMaxStencil];  // ... of i + off[k]
    int slot_lo[kMaxStencil];  // ... of i - off[k]
    // Symmetric form whose positive offsets are exactly 1, L-1, L, L+1 (a Q4 mesh numbered
    // row by row with L nodes per row, L >= 64): row_len = L, and the fused iteration can
    // tile the mesh in two dimensions (variant 11); else 0.
    int row_len;
};
constexpr int kFusedTile = 256;         // nodes per tile (one per thread)
constexpr int kGridRows = 4;            // variant 11: mesh rows per tile (one wavefront each, 64 nodes)
constexpr int kFusedMaxWindow = 2048;   // LDS pairs of all windows of a tile

// ---- element kernels (lfg_element.hip)
enum KeMode { KE_STORE = 0, KE_ATOMIC = 1, KE_COLORED = 2 };
// Ke for elements [0, count) of `elems` (or all when elems == nullptr). STORE writes
// ke_out[64e + 8r + c]; ATOMIC/COLORED scatter into values through emap.
void launch_ke(hipStream_t s, const MeshView& m, KeMode mode, const int* elems, int count, const int* emap,
               double* values, double* ke_out, int* jacobian_failures);
// Q4 meshes whose rows hold at most kAssembleRowSlots values: writes every value of
// rows 2i, 2i+1 once, summing its contributions from 0.0 in the order of node i's
// (element << 2 | corner) list (element order: serial; colour order: coloring).
constexpr int kAssembleRowSlots = 32;
void launch_assemble_rows(hipStream_t s, const MeshView& m, int n_nodes, const int* row_ptr, const int* node_ptr,
                          const int* node_list, const int* emap, double* values, int* jacobian_failures);
// The same sums with each (element, corner) row pair formed once per block of 64 nodes
// and the value positions taken from the pattern's columns; for meshes whose every block
// of 64 consecutive nodes has at most kAssembleBlockSlots (element, corner) entries.
constexpr int kAssembleBlockNodes = 64;
constexpr int kAssembleBlockSlots = 272;
// Atomic assembly of a W x H element grid numbered row by row (element ey W + ex, nodes
// n, n + 1, n + W + 2, n + W + 1 with n = ey (W + 1) + ex; rows of at most
// kAssembleGridRow values), with K zeroed: 8 x 8-element tiles summed in LDS, stores for
// the rows a tile completes and FP64 atomics for the others.
constexpr int kAssembleGridRow = 18;
void launch_assemble_atomic_tiles(hipStream_t s, const MeshView& m, int grid_w, int grid_h, const int* row_ptr,
                                  const int* emap, double* values, int* jacobian_failures);
void launch_assemble_block_rows(hipStream_t s, const MeshView& m, int n_nodes, const int* row_ptr, const int* col,
                                const int* node_ptr, const int* node_list, double* values, int* jacobian_failures);
// values[k] = sum of ke[contrib[j]] for j in [cptr[k], cptr[k+1]) in element order.
void launch_gather(hipStream_t s, long long nnz, const int* cptr, const int* contrib, const double* ke,
                   double* values);
void launch_fill(hipStream_t s, double* p, long long n, double v);
void launch_fields(hipStream_t s, const MeshView& m, const double* u, double* centre, double* gp);
// Nodal smoothing over node -> (element << 2 | local node) lists in element order;
// elements n_q4 and up are bars (their stress tensor counts once, solver.cpp:841-851).
void launch_smooth(hipStream_t s, int n_nodes, int n_q4, const int* node_elem_ptr, const int* node_elem,
                   const double* gp, double* nodal);
// Bar matrices (bars == nullptr: all bars in order), as KE_STORE into ke_out (16 per
// bar), KE_ATOMIC or KE_COLORED into values; zero-length bars count in failures.
void launch_bar_ke(hipStream_t s, const MeshView& m, const BarView& b, KeMode mode, const int* bars, int count,
                   double* values, double* ke_out, int* failures);
void launch_bar_fields(hipStream_t s, const MeshView& m, const BarView& b, const double* u, double* centre,
                       double* gp);
// grad[e] = -lam_e^T Ke(E = 1) u_e * (dE ? dE[e] : 1).
void launch_sensitivity(hipStream_t s, const MeshView& m, const double* u, const double* lam, const double* dE,
                        double* grad);

// ---- solver kernels (lfg_solver.hip)
void launch_dirichlet(hipStream_t s, const CsrView& a, const double* values, const double* f, const int* constrained,
                      const double* prescribed, double* kc, double* fc);
// The zero-pivot sweep after the elimination (solver.cpp:916-940) on (values, f) in
// place. scale_bits must start at 0, counters at {0, INT_MAX}: counters[0] gets the
// pinned count, counters[1] the first loaded zero-pivot row (INT_MAX: none).
void launch_pin_zero_pivots(hipStream_t s, const CsrView& a, double* values, double* f,
                            unsigned long long* scale_bits, int* counters);
void launch_spmv_serial_rows(hipStream_t s, const CsrView& a, const double* values, const double* x, const double* sub,
                             double* y);
void launch_extract_diag(hipStream_t s, const CsrView& a, const double* values, double* diag, int* status);
void launch_block2_factor(hipStream_t s, const CsrView& a, const double* values, double* factor, int* degenerate);
void launch_jacobi_inv(hipStream_t s, int n, const double* diag, double* inv, int* status);
void launch_precond_apply(hipStream_t s, int kind, int n, const double* m, const double* in, double* out);
void launch_to_f32(hipStream_t s, const double* in, float* out, long long n);
// out[i] = constrained[i] ? 0 : in[i] (adjoint right-hand side).
void launch_mask_constrained(hipStream_t s, int n, const int* constrained, const double* in, double* out);

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
// fused = 1 prepares the fused stencil iteration: p and q zero, alpha = beta = 0 and
// the iteration count at -1, so the first fused launch forms p0 = M^-1 r0 and A p0.
void launch_pcg_init(hipStream_t s, const PcgBuffers& b, int stop_rule, double tol, long long max_iter, double drop,
                     double mixed_switch, int fused = 0);
// q = A p; finishes p.q and alpha. Without vals32 the FP64 values are used; with
// both, the kernel reads the FP32 values once the state is in the FP32 phase.
void launch_pcg_spmv(hipStream_t s, const PcgBuffers& b, const CsrView& a, const double* vals64, const float* vals32,
                     int variant);
// Packs the CSR `values` into the slice planes (out64: 4 * blocks doubles; out32:
// 4 * blocks floats, off-diagonal blocks only, with diag_out: 4 doubles per node, the
// compensated diagonal; any may be null) and, with col_out, the node columns.
void launch_pack_slices(hipStream_t s, const CsrView& a, const SliceView& v, const double* values, double* out64,
                        float* out32, int* col_out, double* diag_out = nullptr);
// q = A p on the slice layout; finishes p.q and alpha like launch_pcg_spmv. Each
// row is summed in column order, as the CSR row loop does.
void launch_pcg_spmv_slices(hipStream_t s, const PcgBuffers& b, const SliceView& v, const double* planes64,
                            const float* planes32);
int slice_spmv_blocks(int n_slices);
// Jacobi / 2x2 block only: the slice multiply of p = z + beta p_old, z = M^-1 r,
// formed per gathered column (p_old, r and the preconditioner read-only), p written
// for the own nodes into p_new; replaces launch_pcg_direction + the SpMV (variant 7:
// 83 us against 64 + 10 us at 1000x1000, so not the default).
void launch_pcg_spmv_slices_dir(hipStream_t s, const PcgBuffers& b, const SliceView& v, const double* planes64,
                                const float* planes32, const double* p_old, double* p_new);
// Packs the constrained CSR `values` into the stencil planes (C = ceil(n_nodes / 64)
// chunks): out64 (all planes, 256 * n_off * C doubles); out32 (symmetric form: the
// off-diagonal planes, 256 * (n_off - 1) * C floats) with diag_c (256 * C doubles: the
// diagonal blocks carrying the FP32 rounding compensation). Any output may be null.
void launch_pack_stencil(hipStream_t s, const CsrView& a, const StencilView& v, const double* values, double* out64,
                         float* out32, double* diag_c);
// q = A p on the stencil layout; finishes p.q and alpha like launch_pcg_spmv. With
// planes32 (and diag_c), the FP32 copy is read once the state is in the FP32 phase
// (always, without planes64).
void launch_pcg_spmv_stencil(hipStream_t s, const PcgBuffers& b, const StencilView& v, const double* planes64,
                             const float* planes32, const double* diag_c);
// One whole Jacobi / 2x2 block-Jacobi PCG iteration on the symmetric stencil form
// (requires v.win_total > 0). With alpha, beta of the previous launch: x += alpha p,
// r' = r - alpha q, z = M^-1 r', p' = z + beta p (own nodes written, the windows the
// multiply needs formed in LDS), q' = A p'; sums ||r'||^2, r'.z, p'.q', q'.z, q'.M^-1 q'.
// The last block applies the stop test to r', then alpha' = r'.z / p'.q' and
// beta' = rho'' / r'.z with rho'' = r'.z - 2 alpha' q'.z + alpha'^2 q'.M^-1 q', the r.z
// of the next residual, so no second reduction is needed. r, p, q alternate between
// the *_in and *_out buffers from one launch to the next (launch L, counted from 0
// after pcg_init, reads buffer L & 1). x is updated in place on even launches only,
// with both iterations' terms; launch_pcg_fused_flush completes it after the solve.
void launch_pcg_fused_stencil(hipStream_t s, const PcgBuffers& b, const StencilView& v, const double* planes64,
                              const float* planes32, const double* diag_c, const double* r_in, double* r_out,
                              const double* p_in, double* p_out, const double* q_in, double* q_out);
// After the last fused launch: the x term of an odd final launch (p_odd = its p_in,
// the buffer odd launches read); nothing after an even one.
void launch_pcg_fused_flush(hipStream_t s, const PcgBuffers& b, const double* p_odd);
// Variant 11: the iteration of launch_pcg_fused_stencil on 2D tiles of kGridRows mesh rows
// by 64 nodes (needs v.row_len > 0); same buffers, same state updates.
void launch_pcg_fused_grid(hipStream_t s, const PcgBuffers& b, const StencilView& v, const double* planes64,
                           const float* planes32, const double* diag_c, const double* r_in, double* r_out,
                           const double* p_in, double* p_out, const double* q_in, double* q_out);
// Variant 12: the same iteration on the node-slice layout (any node-blocked pattern); p' of
// the column nodes is formed from r_in, q_in, p_in as the owner forms it.
void launch_pcg_fused_slices(hipStream_t s, const PcgBuffers& b, const SliceView& v, const double* planes64,
                             const float* planes32, const double* r_in, double* r_out, const double* p_in,
                             double* p_out, const double* q_in, double* q_out);
// Variant 14's blocks: kFoldThreads threads, one node slice per wavefront, at most
// kFoldMaxBlocks blocks (1,024 slices); part_in / part_out hold 5 doubles per block.
constexpr int kFoldThreads = 256;
constexpr int kFoldMaxSlices = 1024;
constexpr int kFoldMaxBlocks = kFoldMaxSlices / (kFoldThreads / 64);
// Variant 14: the iteration of variant 12 with the reduction finished at the start of the
// next launch (same bits). The state and the partial sums alternate by launch parity: a
// launch reads st_in and part_in (5 doubles per block of the previous launch) and writes
// st_out and part_out. At most 1,024 slices (one slice per wavefront, one-level fold).
void launch_pcg_fold_slices(hipStream_t s, const PcgBuffers& b, const SliceView& v, const double* planes64,
                            const float* planes32, const double* r_in, double* r_out, const double* p_in,
                            double* p_out, const double* q_in, double* q_out, const PcgState* st_in,
                            PcgState* st_out, const double* part_in, double* part_out);
// Variant 13: the same iteration in two launches on the node slices (the vector update,
// then the multiply with the state update); same buffers.
void launch_pcg_split_slices(hipStream_t s, const PcgBuffers& b, const SliceView& v, const double* planes64,
                             const float* planes32, const double* r_in, double* r_out, const double* p_in,
                             double* p_out, const double* q_in, double* q_out);
// Symmetric node-slice layout (variant 15) of a node-blocked matrix whose pattern is
// symmetric: per node i, the "upper" blocks (i, j) with j >= i (the diagonal first) keep
// their values and node column, the "lower" blocks (j < i) keep their node column and
// the position of block (j, i) among the upper blocks, whose transpose they are read as.
// Both lists are stored per 64-node slice column-major, as in SliceView (entry k of node
// 64 s + l at ptr[s] + 64 k + l). The upper values are two pair planes (upper pairs, then
// lower pairs), each with a zero block at index ublocks that padded lower entries point
// to; padded upper entries are zero blocks. The FP32 copy holds the off-diagonal upper
// blocks in FP32 (the diagonal entry zero) and the diagonal block in FP64 (diag32), minus
// the rounding error of every FP32 block the node's rows read (its own upper blocks and
// the mirrored blocks of its lower ones), so each row keeps its FP64 sum and nodal
// translations multiply exactly, as in the other FP32 layouts.
struct SymSliceView {
    int n_slices;
    long long ublocks;  // padded upper entries, uptr[n_slices]
    const int* uptr;    // n_slices + 1 upper entry offsets (multiples of 64)
    const int* ucol;    // node column per upper entry
    const int* lptr;    // n_slices + 1 lower entry offsets (multiples of 64)
    const int* lcol;    // node column per lower entry
    const int* ltpos;   // position of the transposed upper block per lower entry
    int max_uwidth, max_lwidth;
    const double* diag32;  // FP32 copy: compensated FP64 diagonal block per node (two pairs)
};
// Copies every node's upper blocks from its two CSR rows into the FP64 planes
// (4 * (ublocks + 1) doubles) and / or the FP32 planes (as many floats; then the
// compensated diagonal into v.diag32), with the zero block.
void launch_pack_sym_slices(hipStream_t s, const CsrView& a, const SymSliceView& v, const double* values,
                            double* planes64, float* planes32);
// q = A p on the symmetric node slices (FP64 planes, or FP32 ones when planes64 is null or
// the state's phase says so), p.q and alpha (as launch_pcg_spmv_slices).
void launch_pcg_spmv_sym_slices(hipStream_t s, const PcgBuffers& b, const SymSliceView& v, const double* planes64,
                                const float* planes32);
// Variant 15: the two-launch iteration of variant 13 with the multiply on the symmetric
// node slices.
void launch_pcg_sym_split_slices(hipStream_t s, const PcgBuffers& b, const SymSliceView& v, const double* planes64,
                                 const float* planes32, const double* r_in, double* r_out, const double* p_in,
                                 double* p_out, const double* q_in, double* q_out);
// x += alpha p, r -= alpha q, ||r||^2, r.z; convergence / replacement decision, beta.
void launch_pcg_update(hipStream_t s, const PcgBuffers& b, int mixed);
// mixed only: r = b - K x in FP64 when a replacement is due; then r.z and beta.
void launch_pcg_replace(hipStream_t s, const PcgBuffers& b, const CsrView& a, const double* vals64);
// p = z + beta p.
void launch_pcg_direction(hipStream_t s, const PcgBuffers& b);
// One triangular sweep of IC(0) in level order: position p (rows sorted by level,
// stable) holds row ord[p], its dependencies as positions key[ptr[p] .. ptr[p+1]) with
// factor values val[...] in the reference's summation order, and the reciprocal pivot
// scale[p]; level l covers positions [level_ptr[l], level_ptr[l+1]). reach = max over
// rows of (position - smallest dependency position); widest = largest level.
struct Ic0Sweep {
    const int* ord;
    const int* ptr;
    const int* key;
    const double* val;
    const double* scale;
    const int* level_ptr;
    int levels;
    int reach;
    int widest;
};

// The same sweep by nodes, for factors whose rows 2i, 2i+1 share their dependency
// nodes (every two-dof-per-node pattern): position q holds node ord[q], whose two rows
// are computed by one lane, "first" then "second" (forward: 2i then 2i+1; backward:
// 2i+1 then 2i), so the dependency between them never crosses lanes and the levels
// are those of the node graph (half as many). Dependency k of position q is the node at
// position key[k]; v_first[2k], v_first[2k+1] and v_second[2k], v_second[2k+1] are the
// factor values its two dofs meet in the two rows, in the reference's order (forward:
// dof 2j then 2j+1; backward: 2j+1 then 2j); cross[q] couples the second row to the
// first and comes last; scale_first / scale_second are the reciprocal pivots.
struct Ic0NodeSweep {
    const int* ord;
    const int* ptr;
    const int* key;
    const double* v_first;
    const double* v_second;
    const double* cross;
    const double* scale_first;
    const double* scale_second;
    const int* level_ptr;
    int levels;
    int reach;
    int widest;
};

// IC(0): z = (L L^T)^-1 r, then r.z and beta (one block). Factors whose level-ordered
// dependencies reach back at most 4096 - 128 positions take the pipelined sweeps (the
// values of the last positions stay in an LDS window, levels run back to back with
// their data prefetched), by nodes where the factor allows it; others take the plain
// level-scheduled sweeps.
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
    Ic0Sweep fwd, bwd;   // level-ordered forms of the two sweeps
    Ic0NodeSweep nfwd, nbwd;  // by nodes (ord == nullptr when the factor has no node form)
    int* fault;          // set when a dependency wait exceeds its bound (never for a valid factor)
};
void launch_ic0_apply(hipStream_t s, const Ic0View& f, const double* in, double* out);
void launch_pcg_ic0_step(hipStream_t s, const PcgBuffers& b, const Ic0View& f);

// ---- subdomain block-Jacobi (lfg_subdomain.hip)
// Subdomain s owns dofs[dof_begin[s] .. dof_begin[s+1]) in matrix order and the packed
// lower Cholesky factor at factor + factor_begin[s] (row i at i(i+1)/2, reciprocal
// pivots on the diagonal), precond.cpp:121-232. max_size <= kMaxSubdomainDofs.
constexpr int kMaxSubdomainDofs = 512;
struct SubdomainView {
    int n_sub;
    int max_size;
    const int* dof_begin;
    const int* dofs;
    const long long* factor_begin;
    const double* factor;
};
void launch_subdomain_apply(hipStream_t s, const SubdomainView& v, const double* in, double* out);
// z = M^-1 r with rho = r.z and beta into the PCG state (uses ticket set 5).
void launch_pcg_subdomain_step(hipStream_t s, const PcgBuffers& b, const SubdomainView& v);
// ---- batched PCG (lfg_batch.hip): independent problems in whole 64-node slices,
// NC load cases per problem stored per node ([node][case][2]).
struct BatchSystem {      // one (problem, case)
    double rho, alpha, beta, rr, target;
    double rnorm0;        // ||r0|| = ||b|| (x starts at zero)
    double true_rr;       // ||b - K x||^2 in FP64 from the residual kernel
    long long iter;
    long long switch_iter;  // mixed: iteration at which the problem moved to FP32 values (-1: never)
    int status, converged, done, pad;
};
struct BatchGlobal {
    unsigned int problems_done;
    int done_all;
    int n_active;     // entries of the active slice list the iteration kernels walk
    int changed;      // a problem stopped since the list was last compacted
    int compact_now;  // the scan found changes: the fill kernel rewrites the list
    int pad[11];
};
struct BatchView {
    int n_problems, n_cases, n_slices;
    const int* slice_problem;  // problem of each slice
    const int* pslice_ptr;     // slices of problem p: [pslice_ptr[p], pslice_ptr[p+1])
    double* x;
    double* r;
    double* p;
    double* q;
    const double* b;
    const double* minv;        // Jacobi inverse per dof (0 on padding rows)
    BatchSystem* sys;
    int* problem_done;
    unsigned int* tickets;     // one per problem, 32 apart
    double* partials;          // slices x 2 * n_cases
    BatchGlobal* global;
    int* problem_phase32;      // mixed: 1 once problem p multiplies with FP32 values; null for FP64
    int* active;               // slices of the running problems (problem order), global->n_active of them
    int* active_first;         // compaction: position of problem p's first slice in the new list
    double switch_tau;         // mixed: relative residual that moves a problem to FP32 (< 0: from the start)
    double tol;
    int stop_rule;
    long long max_iter;
};
struct BatchShape {
    int block_threads;    // 64 .. 512
    int slices_per_wave;  // slices each wavefront walks per launch
    int span = 0;         // grid sized for this many slices (0: every slice); the iteration
                          // kernels walk the active list with any grid
};
unsigned int batch_grid(const BatchShape& shape, int n_slices);
void launch_batch_init(hipStream_t s, const BatchView& bv, const BatchShape& shape);
// q = K p with the per-problem p.q fold and alpha. planes32 non-null selects the mixed
// kernel (both plane sets, per-problem phase). sym non-null: the FP64 values are the
// symmetric node slices' (planes64 = their planes), the FP32 ones stay the node slices'.
void launch_batch_spmv(hipStream_t s, const BatchView& bv, const BatchShape& shape, const SliceView& v,
                       const double* planes64, const float* planes32, const SymSliceView* sym = nullptr);
void launch_batch_update(hipStream_t s, const BatchView& bv, const BatchShape& shape);
void launch_batch_direction(hipStream_t s, const BatchView& bv, const BatchShape& shape);
// Rebuilds the active slice list without the problems that stopped since the last
// compaction (two kernels; both return at once when nothing changed).
void launch_batch_compact(hipStream_t s, const BatchView& bv);
// res = b - K x (FP64 values), ||res||^2 per system into true_rr.
void launch_batch_residual(hipStream_t s, const BatchView& bv, const BatchShape& shape, const SliceView& v,
                           const double* planes64, double* res);
void launch_batch_jacobi(hipStream_t s, const CsrView& a, const double* values, double* minv, int* bad);
void launch_batch_scatter_case(hipStream_t s, int n, int n_cases, int c, const double* src, double* dst);
void launch_batch_gather_case(hipStream_t s, int n, int n_cases, int c, const double* src, double* dst);

// ---- allocation-free multiply (lfg_noalloc.hip, lfg_spmv_noalloc)
// One file-scope scratch array per device: the PcgState slot at 0, two TicketSets at
// kNoallocTicketsAt (the multiplies count on the second), and the per-block partials
// of the p.q reduction at kNoallocPartialsAt for grids of up to kNoallocMaxGrid blocks.
constexpr size_t kNoallocTicketsAt = 256;
constexpr size_t kNoallocPartialsAt = 2560;
constexpr int kNoallocMaxGrid = 8192;
constexpr size_t kNoallocScratchBytes = kNoallocPartialsAt + sizeof(double) * (kNoallocMaxGrid + 8);
static_assert(sizeof(PcgState) <= kNoallocTicketsAt, "solver state slot");
static_assert(kNoallocTicketsAt + 2 * sizeof(TicketSet) <= kNoallocPartialsAt, "ticket slots");
// With `device` current: the device address of that device's scratch array (once per
// device; later calls return at once).
hipError_t noalloc_scratch_init(int device);
// The address noalloc_scratch_init recorded for `device`, or null before it ran.
void* noalloc_scratch(int device);
// The grid of the multiply launch_stencil_spmv / launch_pcg_spmv_slices / launch_pcg_spmv
// issue for this layout (the blocks whose partials the reduction folds).
inline unsigned int stencil_spmv_blocks(int n_nodes) {
    const unsigned int need = (static_cast<unsigned int>(n_nodes) + 255u) / 256u;
    const unsigned int r8 = (need + 7u) & ~7u;
    return r8 < 4096u ? r8 : 4096u;
}

// The 2D-tiled stencil multiply (row-structured symmetric stencils, v.row_len > 0): one
// block per tile of kGridRows x 64 nodes, at most 4096 blocks (a block then takes several).
inline unsigned int stencil_grid_blocks(const StencilView& v) {
    const long long L = v.row_len;
    const long long tiles = ((v.n_nodes + L - 1) / L + kGridRows - 1) / kGridRows * ((L + 63) / 64);
    const unsigned int r8 = static_cast<unsigned int>((tiles + 7) & ~7LL);
    return r8 < 4096u ? r8 : 4096u;
}
inline bool stencil_grid_multiply(const StencilView& v) { return v.sym && v.n_off == 5 && v.row_len > 0; }

// ||b - K x||^2 in FP64 into out2 (reporting, iterative refinement); r_out, when not
// null, receives r = b - K x.
void launch_residual_norm(hipStream_t s, const CsrView& a, const double* vals64, const double* x, const double* bvec,
                          double* partials, TicketSet* tickets, double* out2, double* r_out = nullptr);
// y += a x
void launch_axpy(hipStream_t s, long long n, double a, const double* x, double* y);
// e = a + b + c + d over n double pairs (streaming-bandwidth calibration).
void launch_stream_probe(hipStream_t s, long long n, const double* a, const double* b, const double* c,
                         const double* d, double* e);

}  // namespace lfg
