/* lightfem_gpu.h - C ABI of the MI355X (gfx950) backend for LightFEM's Q4
 * assembly + PCG path.
 *
 * The reference has no FFI: its "operator API" is the C++ library, and every
 * stage below is a file-static function inside LinearElasticSolver::solve
 * (src/solver.cpp:875-960). Each entry point replaces one of those stages; the
 * host orchestrator (lightfem::LinearElasticSolver with the hip backend, or a
 * ctypes caller) calls them in the reference's order.
 *
 * Conventions
 *  - Status codes, never exceptions, cross the ABI; lfg_last_error() holds the
 *    message. The C++ wrapper rethrows each code with the reference's text.
 *  - The caller owns every host array. The library copies inputs into its own
 *    device buffers (owned until lfg_destroy) and writes results into
 *    caller-allocated host arrays. Calls are synchronous with respect to the
 *    host arrays.
 *  - One HIP stream per context; a context must not be shared between host
 *    threads without external locking.
 *  - Numbering is the reference's: node index i (after sort_and_index or RCM)
 *    owns dofs 2i and 2i+1 (src/model.cpp:74-92). Element node lists are node
 *    indices in the element's counter-clockwise order.
 */
#ifndef LIGHTFEM_GPU_H
#define LIGHTFEM_GPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
    LFG_OK = 0,
    LFG_ERR_ARG = 1,            /* std::invalid_argument / std::logic_error in the reference */
    LFG_ERR_JACOBIAN = 2,       /* element.cpp:52-54 */
    LFG_ERR_NOT_SPD = 3,        /* solver.cpp:591-593 non-positive curvature */
    LFG_ERR_NONFINITE = 4,      /* solver.cpp:579-581 initial rho, :616-618 rho breakdown */
    LFG_ERR_NOT_CONVERGED = 5,  /* solver.cpp:948-950 */
    LFG_ERR_HIP = 6,            /* HIP runtime error (device, memory) */
    LFG_ERR_STATE = 7,          /* call made before the stage it needs */
    LFG_ERR_UNSUPPORTED = 8,    /* option outside this backend's scope */
    LFG_ERR_MECHANISM = 9       /* a zero pivot carries a load (solver.cpp:922-934) */
} lfg_status;

typedef struct lfg_ctx lfg_ctx;

enum { LFG_PLANE_STRESS = 0, LFG_PLANE_STRAIN = 1 };
/* `assembly` setting (model.hpp:99, solver.cpp:275-299). */
enum { LFG_ASSEMBLY_SERIAL = 0, LFG_ASSEMBLY_ATOMIC = 1, LFG_ASSEMBLY_COLORING = 2, LFG_ASSEMBLY_THREAD_LOCAL = 3 };
/* `preconditioner` setting (model.hpp:105); BLOCK2 = `block` with subdomain_nodes 1
 * (2x2 nodal blocks, fused into the iteration); SUBDOMAIN = `block` on the subdomains
 * given to lfg_set_subdomains (any size, SubdomainBlockJacobi, precond.cpp:121-281). */
enum { LFG_PRECOND_JACOBI = 0, LFG_PRECOND_BLOCK2 = 1, LFG_PRECOND_IC0 = 2, LFG_PRECOND_SUBDOMAIN = 3 };
/* `storage` setting (model.hpp:103): which summation order the products follow, the CSR
 * row loop (sparse.cpp:104-120) or the 2x2 block storage product (sparse.cpp:128-198). */
enum { LFG_STORAGE_CSR = 0, LFG_STORAGE_BSR = 1 };
/* OR'ed into the variant of lfg_spmv / lfg_bench_spmv: the block storage order. */
enum { LFG_SPMV_BLOCK_ORDER = 256 };
enum { LFG_PRECISION_FP64 = 0, LFG_PRECISION_MIXED = 1 };
/* reference: ||r|| <= tol*max(1,||r0||) (solver.cpp:572); relative: ||r|| <= tol*||b||. */
enum { LFG_STOP_REFERENCE = 0, LFG_STOP_RELATIVE = 1 };

typedef struct {
    int32_t preconditioner;  /* LFG_PRECOND_* */
    int32_t precision;       /* LFG_PRECISION_* */
    int32_t stop_rule;       /* LFG_STOP_* */
    int32_t check_every;     /* iterations per captured graph between convergence reads; 0 = default */
    double tolerance;        /* SolverSettings::tolerance */
    int64_t max_iterations;  /* SolverSettings::max_iterations */
    double replace_drop;     /* mixed: replace the recurrence residual by the FP64 residual b - K x each time
                                it falls by this factor (kept while replacements halve the FP64 residual);
                                0 = default: no replacements (variant 10 has none) */
    double mixed_switch;     /* mixed: SpMVs use FP64 values until ||r|| <= mixed_switch*||r0||, then FP32
                                values; 0 = default 1e-2, negative = FP32 values from the first iteration */
    int32_t direct_launch;   /* 1: launch each kernel directly instead of replaying a captured hipGraph */
    int32_t spmv_variant;    /* 0 = default: for jacobi / block 14 on a node-blocked pattern of at
                                most 1024 slices (any numbering), else 11, else 10 where the node
                                stencil applies, else on a node-blocked pattern 15 (13 where the
                                pattern is not structurally symmetric); for ic0 /
                                subdomains 8 where the stencil applies, else 6; else 4. 1-3: node
                                groups of 4/8/16 lanes per row;
                                4, 5: row-segment streaming with 64 / 32 rows per wavefront; 6:
                                node-slice layout (one column per 2x2 block, 64 nodes per
                                wavefront); 7: node-slice layout with the Jacobi / block search
                                direction formed inside the multiply (two launches per iteration,
                                bit-identical to 6; slower on gfx950: the extra gathers cost more
                                than the launch they save); 8: node-stencil layout, symmetric form
                                (structured meshes numbered row by row: no column indices, the
                                diagonal and positive-offset blocks only, block (i, i-d) read as
                                the transpose of (i-d, i); FP32 copy with a rounding-compensated
                                FP64 diagonal); 9: node-stencil layout, every offset (bit-identical
                                to 4 and 6, FP64 only); 10: the layout of 8 with the whole Jacobi /
                                block iteration in one launch (update, search direction, multiply
                                and all sums; beta from the next residual's r.z recurrence, so one
                                reduction per iteration); 11: the iteration of 10 on 2D tiles of 4 mesh
                                rows by 64 nodes, for symmetric stencils with positive offsets 1, L-1,
                                L, L+1 (a Q4 mesh numbered row by row, L >= 64 nodes per row): a
                                140-position halo per 256-node tile instead of 518, every load issued
                                before the first use; 12: the iteration of 10 on the node-slice
                                layout (any node-blocked pattern, e.g. RCM numbering): p' of each
                                column node formed in the multiply from r, q, p of that node, one
                                launch per iteration; 13: the same iteration in two launches (the
                                vector update, then the multiply with the sums and state update),
                                cheaper than 12 where the column gathers outweigh a launch (large
                                systems); 14: the iteration of 12 with its sums folded at the start
                                of the next launch instead of by the last block (no completion
                                counter, no wait for the stores: bit-identical to 12, the stop seen
                                one launch later; up to 1024 slices); 15: the iteration of 13 on the
                                symmetric node slices (per node the blocks from the diagonal on hold
                                values, those before it are read as the transposes of the column
                                nodes' blocks: each mirrored pair streams once; CSR order; its FP32
                                copy keeps a rounding-compensated FP64 diagonal). 10-15 apply the
                                jacobi and block preconditioners only. */
    int32_t storage;         /* LFG_STORAGE_*: BSR sums each 2x2 block's products first (variants 6, 7
                                and 9 follow it bit for bit; 8 and 10 read the transposed blocks as
                                always; 1-5 are CSR only). Needs a node-blocked pattern. */
    int32_t refine_passes;   /* FP64 iterative refinement after the solve: up to this many passes of
                                r = b - K u with FP64 values, a correction solve K d = r (same precision
                                policy) to ||r_d|| <= max(refine_reduction ||r||, target) and u += d,
                                until ||b - K u|| meets the target or a pass fails to halve it (the FP64
                                residual floor; a pass that made it worse is undone). 0 = default: 4
                                for mixed precision, none for FP64 (the reference's recurrence rule
                                alone); negative = none. Mixed precision with passes left and the FP64
                                residual above the target and still falling reports NOT_CONVERGED. */
    double refine_reduction; /* residual reduction of each correction solve; 0 = default 1e-2 */
} lfg_pcg_options;

typedef struct {
    int32_t converged;
    int32_t status;               /* lfg_status of the solve */
    int64_t iterations;
    int64_t residual_replacements;  /* FP64 residual recomputations: replacements and refinement passes */
    double initial_residual;      /* ||r0|| = ||b|| (x0 = 0) */
    double final_residual;        /* ||r|| at stop: recurrence residual (FP64) or replaced residual (mixed) */
    double true_residual;         /* ||b - K u|| recomputed in FP64 after the solve */
    double solve_ms;              /* device time, first SpMV to convergence (hipEvent) */
    double setup_ms;              /* preconditioner / precision setup (hipEvent) */
    int64_t breakdowns;           /* IC(0) clamped pivots (precond.cpp:373-378) */
    int64_t degenerate_blocks;    /* 2x2 blocks replaced by identity (precond.cpp:218-223) */
    int64_t levels;               /* IC(0): forward + backward level count */
    int64_t refine_passes;        /* FP64 iterative-refinement correction solves run (their iterations
                                     are included in `iterations`) */
} lfg_pcg_report;

typedef struct {
    double element_stiffness_ms;  /* separate Ke kernel (PerformanceReport::element_stiffness_ms); 0 where
                                     the assembly kernel forms the element matrices itself */
    double assembly_ms;           /* assembly (PerformanceReport::assembly_ms) */
    int32_t coloring_count;
    int32_t jacobian_failures;
} lfg_assembly_report;

/* Context (one device, one stream). */
int lfg_create(int32_t device, lfg_ctx** out);
void lfg_destroy(lfg_ctx* ctx);
const char* lfg_last_error(const lfg_ctx* ctx);
const char* lfg_status_string(int32_t status);

/* Mesh and materials. Replaces q4_coordinates/find_material inside the Ke loop
 * (solver.cpp:48-56, 243-270). xy: 2*n_nodes; enodes: 4*n_elems node indices;
 * elem_material: n_elems material indices; materials: 3*n_materials (E, nu, t). A new
 * mesh drops the pattern, bars and colours of the previous one (lfg_set_pattern and
 * lfg_set_colors follow it). */
int lfg_set_mesh(lfg_ctx* ctx, int64_t n_nodes, const double* xy, int64_t n_elems, const int32_t* enodes,
                 const int32_t* elem_material, int64_t n_materials, const double* materials, int32_t plane_mode);

/* Replaces the material table (E, nu, t per material; same count as lfg_set_mesh)
 * without touching mesh or pattern; assembly must follow. Serves repeated solves
 * with changing moduli (the material rows parsed at input_parser.cpp:162-342),
 * e.g. the forward solves of a density-based gradient loop. */
int lfg_set_materials(lfg_ctx* ctx, int64_t n_materials, const double* materials);

/* Plane mode per material row (LFG_PLANE_*), overriding lfg_set_mesh's single mode, for
 * batches whose problems mix plane stress and plane strain (each problem has one mode in
 * the reference, model.hpp settings.plane_mode). NULL restores the single mode. */
int lfg_set_material_modes(lfg_ctx* ctx, const int32_t* modes);

/* CSR pattern and element->CSR-value map of build_csr_assembly_pattern
 * (sparse.cpp:15-58): row_ptr n_rows+1, col nnz, emap 64*n_elems (row-major local r*8+c). */
int lfg_set_pattern(lfg_ctx* ctx, int64_t n_rows, int64_t nnz, const int32_t* row_ptr, const int32_t* col,
                    const int32_t* emap);

/* Bars (truss members, element.cpp:250-306) after the Q4 elements: bar t is element
 * n_elems + t. bar_nodes 2 node indices per bar, bar_material the material index (E
 * from the materials table), area the cross-section, bar_map 16 value indices per bar
 * (lfh_csr_pattern_bars; the pattern given to lfg_set_pattern holds their couplings).
 * After lfg_set_pattern (which drops earlier bars) and before lfg_set_colors; serial
 * assembly adds them after the Q4 contributions of each value, and lfg_postprocess
 * returns n_elems + n_bars rows of element fields. */
int lfg_set_bars(lfg_ctx* ctx, int64_t n_bars, const int32_t* bar_nodes, const int32_t* bar_material,
                 const double* area, const int32_t* bar_map);

/* Colour groups of element_colors (solver.cpp:100-133): color_ptr n_colors+1, elements
 * ascending per colour (Q4 elements and bars, bar t as element n_elems + t). */
int lfg_set_colors(lfg_ctx* ctx, int32_t n_colors, const int32_t* color_ptr, const int32_t* color_elems);

/* Element stiffness alone (q4_stiffness, element.cpp:130-161) into ke_out (64*n_elems), for checks. */
int lfg_element_stiffness(lfg_ctx* ctx, double* ke_out);

/* Numeric assembly (solver.cpp:243-301). serial and thread_local sum every value's
 * contributions in element order (bit-identical to serial_assembly), coloring in colour
 * order (bit-identical to coloring_assembly), atomic adds with FP64 hardware atomics.
 * Q4 meshes with rows of at most 32 values run serial, thread_local and coloring as one
 * owner-computes pass (each node's rows formed and written by one lane group); meshes
 * with bars or longer rows store the element matrices and gather them (serial,
 * thread_local) or scatter them conflict-free colour by colour (coloring). */
int lfg_assemble(lfg_ctx* ctx, int32_t strategy, lfg_assembly_report* report);

/* Unconstrained stiffness values (AnalysisResult::stiffness, solver.cpp:906-909). */
int lfg_get_stiffness(lfg_ctx* ctx, double* values_out);
/* Overwrite the unconstrained values (tests, or a caller-assembled matrix). */
int lfg_set_stiffness(lfg_ctx* ctx, const double* values);

/* Load vector assembled on the host (solver.cpp:303-332). */
int lfg_set_load(lfg_ctx* ctx, const double* f);

/* Symmetric Dirichlet elimination (solver.cpp:355-413) into a constrained copy;
 * the unconstrained values and load stay for reactions. dofs ascending. */
int lfg_dirichlet(lfg_ctx* ctx, int64_t n_constraints, const int32_t* dofs, const double* values);

/* The zero-pivot sweep after the elimination (solver.cpp:916-940): a constrained-system
 * diagonal of magnitude at most 1e-30 (a bar end no other element stiffens) becomes 1
 * with a zero right-hand side, unless that right-hand side exceeds 1e-12 of the largest
 * |rhs|: then LFG_ERR_MECHANISM, with the first such dof, its load and diagonal in the
 * report (the reference's message names them and the dof's node). */
typedef struct {
    int64_t pinned;  /* dofs set to diagonal 1 */
    int64_t dof;     /* first loaded zero-pivot dof, -1 if none */
    double load;
    double diag;
} lfg_pivot_report;
int lfg_pin_zero_pivots(lfg_ctx* ctx, lfg_pivot_report* report);
int lfg_get_system(lfg_ctx* ctx, double* values_out, double* rhs_out);

/* PCG on the constrained system (solve_iterative + pcg_solve, solver.cpp:532-747). */
int lfg_pcg(lfg_ctx* ctx, const lfg_pcg_options* options, double* u_out, lfg_pcg_report* report);

/* R = K_unconstrained u - f (solver.cpp:953-954) for the last solution. */
int lfg_reactions(lfg_ctx* ctx, double* reaction_out);

/* Element fields (postprocess, solver.cpp:756-805: centre 7 doubles and 4 Gauss
 * points x 9 doubles [xi, eta, strain3, stress3, vm] per element) and nodal
 * smoothing (solver.cpp:807-871: 4 doubles per node). Any output may be NULL. */
int lfg_postprocess(lfg_ctx* ctx, double* centre, double* gauss_points, double* nodal);

/* Adjoint solve for a functional J(u) of the last lfg_pcg solution: K lambda = dJ/du
 * on the constrained system (the entries of dJdu at constrained dofs are ignored and
 * lambda is zero there), by the same PCG (solver.cpp:532-627) with the options'
 * preconditioner, which is reused when it matches the forward solve's. The reference
 * has no gradient code; this is the adjoint of its forward path. lambda_out may be NULL
 * (lambda stays on the device for lfg_sensitivity). */
int lfg_adjoint_solve(lfg_ctx* ctx, const lfg_pcg_options* opt, const double* dJdu, double* lambda_out,
                      lfg_pcg_report* report);

/* Element sensitivities grad_e = -lambda_e^T (dKe/dE_e) u_e * dE_drho[e] (dKe/dE = Ke
 * of q4_stiffness, element.cpp:130-161, at E = 1; dE_drho NULL means 1, giving dJ/dE_e),
 * with u the last lfg_pcg solution and lambda the given array (n dofs) or, when NULL,
 * the last lfg_adjoint_solve result. For compliance f^T u, lambda = u. */
int lfg_sensitivity(lfg_ctx* ctx, const double* lambda, const double* dE_drho, double* grad_out);

/* ---- Batches of independent problems (many meshes and load cases in one solve).
 *
 * The mesh given to lfg_set_mesh / lfg_set_pattern is the disjoint union of the
 * problems: problem p owns nodes [node_ptr[p], node_ptr[p+1]), every node_ptr entry a
 * multiple of 64 (problems are padded with nodes that no element uses), and no element
 * spans two problems. Each problem is a LightFEM model of its own (per-problem plane
 * modes through lfg_set_material_modes); its systems are solved by Jacobi-PCG
 * (solver.cpp:532-627) with per-system scalars and convergence. */
int lfg_batch_set_problems(lfg_ctx* ctx, int32_t n_problems, const int32_t* node_ptr);

/* After assembly: n_cases (1, 2 or 4) load cases, loads[c * n_dofs + i], and the Dirichlet
 * supports shared by all cases (apply_dirichlet, solver.cpp:355-413, per case). */
int lfg_batch_set_loads(lfg_ctx* ctx, int32_t n_cases, const double* loads, int64_t n_constraints,
                        const int32_t* dofs, const double* values);

typedef struct {
    int32_t precision;        /* LFG_PRECISION_FP64, or MIXED (see mixed_switch) */
    int32_t stop_rule;        /* LFG_STOP_* on each system's own ||b|| */
    int32_t check_every;      /* iterations per captured graph; 0 = 16 */
    int32_t tune;             /* 1: choose the launch shape by timing a few iterations of this batch per
                                 candidate shape (cached per context and precision) */
    double tolerance;
    int64_t max_iterations;   /* per system */
    double mixed_switch;      /* mixed: a problem multiplies with FP64 values until every running case's
                                 residual is at most mixed_switch * ||b||, then with FP32 off-diagonal
                                 values and an FP64 diagonal carrying their rounding (vectors and scalars
                                 stay FP64); 0 = 1e-3, negative = FP32 from the first iteration */
    int32_t block_threads;    /* launch shape when not tuning: 64-512 (0 = 256) */
    int32_t slices_per_wave;  /* slices each wavefront walks per launch (0 = 1) */
} lfg_batch_options;

typedef struct {
    int32_t converged;        /* the recurrence residual reached the target */
    int32_t status;           /* lfg_status of the system */
    int64_t iterations;
    int64_t switch_iteration; /* mixed: iteration after which FP32 values were used (-1: never) */
    double initial_residual;  /* ||b|| */
    double final_residual;    /* last recurrence residual */
    double true_residual;     /* ||b - K u|| in FP64 */
} lfg_batch_system_report;

typedef struct {
    double solve_ms;          /* init to the final FP64 residual, device time */
    double tune_ms;
    int64_t graph_launches;
    int32_t problems_switched; /* mixed: problems that moved to FP32 values */
    int32_t block_threads;
    int32_t slices_per_wave;
    int32_t systems_converged;
    int64_t max_iterations;   /* over systems */
    double multiply_ms;       /* the batch multiply (with its per-problem fold) at the tuned launch shape,
                                 one launch with every problem running, timed by the last tuning */
    int32_t sym_slices;       /* 1: FP64 products on the symmetric node slices (lower blocks read as the
                                 mirrored upper ones), 0: on the node slices */
} lfg_batch_summary;

/* Solves every (problem, case) system; u_out[c * n_dofs + i] (NULL to skip), reports
 * [p * n_cases + c] (NULL to skip). Returns LFG_ERR_NOT_CONVERGED when a system did not
 * converge (u_out and reports are still written). */
int lfg_batch_pcg(lfg_ctx* ctx, const lfg_batch_options* opt, double* u_out, lfg_batch_system_report* reports,
                  lfg_batch_summary* summary);

/* SpMV layouts the pattern set by lfg_set_pattern admits. The node-stencil layouts
 * (variants 8, 9) need every 2x2 block (i, j) at one of a few node offsets j - i
 * (a structured mesh numbered row by row: the CSR pattern of sparse.cpp:15-58 then
 * has the same column offsets in every row). */
typedef struct {
    int32_t node_blocked;          /* rows 2i, 2i+1 share one column list of dof pairs (variants 6, 7) */
    int32_t n_slices;              /* 64-node slices of the node-slice layout */
    int32_t stencil_exact_planes;  /* variant 9: offsets of the pattern (0 = not available) */
    int32_t stencil_sym_planes;    /* variant 8: 0 and the positive offsets (0 = not available) */
    int32_t fused_window_pairs;    /* variant 10: LDS pairs per 256-node tile (0 = not available) */
    int32_t stencil_offsets[17];   /* the exact form's offsets, ascending */
    int32_t grid_row_len;          /* variant 11: nodes per mesh row L (0 = not available) */
    int32_t sym_slices;            /* variant 15: node-blocked and structurally symmetric (0 = not available) */
} lfg_layout_info;
int lfg_layout_info_get(lfg_ctx* ctx, lfg_layout_info* info);

/* Subdomains of the `block` preconditioner: node_subdomain[i] in [0, n_subdomains) for
 * every node index i (both dofs of a node share its subdomain, build_dof_subdomain,
 * solver.cpp:631-654; lfh_subdomain_split computes the reference's split). The dense
 * Cholesky factors (precond.cpp:178-232) are formed at the next solve with
 * LFG_PRECOND_SUBDOMAIN; subdomains up to 4096 dofs. */
int lfg_set_subdomains(lfg_ctx* ctx, int64_t n_subdomains, const int32_t* node_subdomain);

/* Checks and measurement hooks on the constrained system. */
int lfg_precond_apply(lfg_ctx* ctx, int32_t preconditioner, const double* in, double* out);
/* y = Kc x with the PCG SpMV kernel of `variant` (values of `precision`; 10 and 11 multiply
 * as 8, 12-14 as 6, 15 on the symmetric node slices; variant | LFG_SPMV_BLOCK_ORDER
 * for 6, 7, 9: the block storage order). */
int lfg_spmv(lfg_ctx* ctx, int32_t precision, int32_t variant, const double* x, double* y);
/* Allocation-free form of lfg_spmv, for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops): no counterpart in the reference. y_dev =
 * Kc x_dev for caller-owned device vectors of n doubles, and x.y into the device double
 * xy_dev when it is not NULL. Allocates nothing: the layout of this precision and variant
 * must already exist (an lfg_spmv or lfg_pcg with them built it; else LFG_ERR_STATE), and
 * the in-kernel reduction's state, counters and per-block partials use one 68 KB scratch
 * array per device that holds partials for up to 8192 blocks (every multiply's grid is
 * capped at 4096, so no system size exceeds it; a larger grid is refused with
 * LFG_ERR_UNSUPPORTED). Calls on one device share that array: calls that overlap on one
 * device (several host threads, or streams of other contexts) race. Synchronous. */
int lfg_spmv_noalloc(lfg_ctx* ctx, int32_t precision, int32_t variant, const double* x_dev, double* y_dev,
                     double* xy_dev);
/* Average device time (ms) of `reps` SpMV launches of the given precision and
 * variant on resident data (variant 7: the multiply with the Jacobi direction formed
 * on the fly, as its PCG iteration runs it; variants 10-15: one whole fused Jacobi PCG
 * iteration, its launches included). */
int lfg_bench_spmv(lfg_ctx* ctx, int32_t precision, int32_t variant, int32_t reps, double* avg_ms);
/* Average device time (ms) of a device-to-device copy of `bytes` bytes (the runtime's
 * copy): the calibration dispatch for hardware traffic counters. */
int lfg_bench_copy(lfg_ctx* ctx, int64_t bytes, int32_t reps, double* avg_ms);
/* Average device time (ms) of a streaming kernel over five device buffers of `bytes`
 * bytes each: four read, one written, one 16-byte pair per thread (the read / write mix
 * of the fused PCG iterations, 5 x bytes moved per launch). The calibration a fused
 * iteration's achieved bandwidth is compared with at its own launch size. No counterpart
 * in the reference (a measurement helper, like lfg_bench_copy). */
int lfg_bench_stream(lfg_ctx* ctx, int64_t bytes, int32_t reps, double* avg_ms);

#ifdef __cplusplus
}
#endif

#endif /* LIGHTFEM_GPU_H */
