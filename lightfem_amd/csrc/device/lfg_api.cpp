// C ABI of the gfx950 backend (include/lightfem_gpu.h): context, device buffers,
// host-side set-up of the device layouts, and the graph-driven PCG loop.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

#include "lfg_ic0_plan.hpp"
#include "lfg_kernels.hpp"
#include "lightfem_gpu.h"

namespace {

struct LfgError : std::runtime_error {
    int code;
    LfgError(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

#define LFG_HIP(call)                                                                                   \
    do {                                                                                                \
        const hipError_t err_ = (call);                                                                 \
        if (err_ != hipSuccess)                                                                         \
            throw LfgError(LFG_ERR_HIP, std::string(#call) + " failed: " + hipGetErrorString(err_));     \
    } while (0)

// Owning device allocation.
template <typename T>
struct DevArray {
    T* p = nullptr;
    size_t n = 0;
    DevArray() = default;
    DevArray(const DevArray&) = delete;
    DevArray& operator=(const DevArray&) = delete;
    ~DevArray() { release(); }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    void alloc(size_t count) {
        if (count == n && p) return;
        release();
        if (count == 0) return;
        LFG_HIP(hipMalloc(&p, count * sizeof(T)));
        n = count;
    }
    void upload(const T* host, size_t count, hipStream_t s) {
        alloc(count);
        if (count) LFG_HIP(hipMemcpyAsync(p, host, count * sizeof(T), hipMemcpyHostToDevice, s));
    }
    void download(T* host, size_t count, hipStream_t s) const {
        if (count) LFG_HIP(hipMemcpyAsync(host, p, count * sizeof(T), hipMemcpyDeviceToHost, s));
    }
};

struct Ic0SweepDevice {
    DevArray<int> ord, ptr, key, level_ptr;
    DevArray<double> val, scale;
    int levels = 0, reach = 0, widest = 0;
};

struct Ic0NodeSweepDevice {
    DevArray<int> ord, ptr, key, level_ptr;
    DevArray<double> v_first, v_second, cross, scale_first, scale_second;
    int levels = 0, reach = 0, widest = 0;
    bool ready = false;
};

struct Ic0Device {
    DevArray<int> lrow, lcol, urow, ucol, fptr, frows, bptr, brows;
    DevArray<double> lval, uval, inv_diag;
    Ic0SweepDevice fwd, bwd;  // level-ordered sweeps
    Ic0NodeSweepDevice nfwd, nbwd;  // by nodes
    int fwd_levels = 0, bwd_levels = 0;
    long long breakdowns = 0;
    bool ready = false;
};

struct SubdomainDevice {
    DevArray<int> dof_begin, dofs;
    DevArray<long long> factor_begin;
    DevArray<double> factor;
    int n_sub = 0, max_size = 0;
    long long degenerate = 0;
};

}  // namespace

struct lfg_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev_start = nullptr, ev_stop = nullptr, ev_ring[2] = {nullptr, nullptr};
    std::string last_error;

    // mesh
    int n_nodes = 0, n_elems = 0, n_mats = 0, plane_mode = 0;
    DevArray<double> xy, materials;
    DevArray<int> mat_mode;  // per-material plane modes (lfg_set_material_modes), else empty
    DevArray<int> enodes, elem_mat, node_elem_ptr, node_elem;
    std::vector<int> h_enodes;  // Q4 node indices (node -> element lists when bars join)
    bool have_mesh = false;
    // bars (truss members), elements n_elems .. n_elems + n_bars - 1
    int n_bars = 0;
    DevArray<int> bar_nodes, bar_mat, bar_map;
    DevArray<double> bar_area;
    std::vector<int> h_bar_nodes;
    // zero-pivot sweep: max |f| bits and {pinned, first loaded zero-pivot row}
    DevArray<unsigned long long> pivot_scale;
    DevArray<int> pivot_counters;

    // pattern
    int n = 0;
    int max_row = 0;
    int max_block_slots = 0;  // most node-list entries of one 64-node block (block-row assembly)
    int grid_w = 0, grid_h = 0;  // elements per row and rows of a row-numbered element grid, else 0
    long long nnz = 0;
    std::vector<int> h_row_ptr, h_col;
    DevArray<int> row_ptr, col, emap, cptr, contrib;
    bool gather_valid = false;  // cptr / contrib match the current pattern and bars
    bool have_pattern = false;
    // node-slice copy of the constrained values (SpMV variant 6), when the pattern is node-blocked
    bool have_slices = false, slices64_valid = false, slices32_valid = false;
    int n_slices = 0, slice_max_width = 0;
    long long slice_blocks = 0;
    DevArray<int> slice_ptr, slice_col;
    DevArray<double> slices64, slice_diag;  // slice_diag: compensated FP64 diagonal of slices32
    DevArray<float> slices32;
    // symmetric node-slice copy (variant 15); its lists are a cache of the pattern, built on
    // first use (also from const contexts): sym_state 0 = not built yet, 1 = built, -1 = the
    // pattern is not structurally symmetric
    mutable int sym_state = 0;
    bool sym64_valid = false, sym32_valid = false;
    mutable int sym_max_uwidth = 0, sym_max_lwidth = 0;
    mutable long long sym_ublocks = 0;
    mutable DevArray<int> sym_uptr, sym_ucol, sym_lptr, sym_lcol, sym_ltpos;
    DevArray<double> sym64, sym_diag;  // sym_diag: compensated FP64 diagonal of sym32
    DevArray<float> sym32;
    // node-stencil copies of the constrained values (SpMV variants 8 and 9), when every
    // block column lies at one of a few fixed node offsets (structured meshes numbered
    // row by row)
    bool have_stencil_sym = false, have_stencil_exact = false;
    bool stencil64_valid = false, stencil32_valid = false, stencil_exact_valid = false;
    lfg::StencilView stencil_sym{}, stencil_exact{};
    DevArray<double> stencil64, stencil_exact64, stencil_diag;
    DevArray<float> stencil32;
    std::vector<int> h_color_ptr;      // Q4 elements of colour c: color_elems[h_color_ptr[c] ..)
    DevArray<int> color_elems;
    DevArray<int> color_node_ptr, color_node;  // Q4-only meshes: node -> (element, corner) in colour order
    std::vector<int> h_color_bar_ptr;  // bars of colour c: color_bars[h_color_bar_ptr[c] ..)
    DevArray<int> color_bars;

    // values
    DevArray<double> K, Kc, f, fc, ke_buf;
    DevArray<float> K32;
    bool have_K = false, have_f = false, have_system = false, k32_valid = false;
    DevArray<int> constrained;
    DevArray<double> prescribed;

    // solver
    DevArray<double> x, r, p, p2, q, z, minv, factor, diag, partials, scratch;
    DevArray<double> r_alt, q_alt;  // fused stencil iteration: second r and q buffers (p uses p2)
    DevArray<double> ir_r, ir_d;    // iterative refinement: FP64 residual b - K x and the correction
    DevArray<double> lam, adj_rhs, dE, grad;  // adjoint solution and right-hand side, sensitivities
    bool have_adjoint = false;
    DevArray<lfg::PcgState> state;
    DevArray<lfg::PcgState> state_alt;  // variant 14: the state of odd launches
    DevArray<double> fold_partials;     // variant 14: two sets of 5 sums per block
    DevArray<lfg::TicketSet> tickets;  // 0 init, 1 SpMV, 2 update, 3 replacement, 4 residual, 5 subdomains
    DevArray<int> flags;
    lfg::PcgState* h_state = nullptr;  // pinned, 2 slots
    bool have_solution = false;
    Ic0Device ic0;
    // subdomain block-Jacobi: node -> subdomain map (host) and the device factors
    std::vector<int> sub_node;
    int n_sub = 0;
    SubdomainDevice sub;
    int pair_sum = 0;  // products in the block storage order (LFG_STORAGE_BSR) while set
    int precond_ready = -1;  // which preconditioner the device factors belong to

    // fields
    DevArray<double> centre, gp, nodal;

    // batch of independent problems
    int batch_problems = 0, batch_cases = 0;
    bool have_batch = false, have_batch_loads = false;
    int batch_tuned_for = -1;  // precision the cached launch shape was timed with
    lfg::BatchShape batch_shape{256, 1};
    double batch_tune_ms = 0.0;
    double batch_multiply_ms = 0.0;  // the chosen launch shape's multiply, per launch
    DevArray<int> batch_slice_problem, batch_pslice_ptr, batch_problem_done, batch_phase32;
    DevArray<int> batch_active, batch_active_first;  // active slice list and its compaction offsets
    DevArray<unsigned int> batch_tickets;
    DevArray<double> batch_b, batch_x, batch_r, batch_p, batch_q, batch_res, batch_minv, batch_partials, batch_tmp;
    DevArray<lfg::BatchSystem> batch_sys;
    DevArray<lfg::BatchGlobal> batch_global;
    lfg::BatchGlobal* h_batch_global = nullptr;  // pinned, 2 slots
};

namespace {

template <typename F>
int guarded(lfg_ctx* ctx, F&& body) {
    if (!ctx) return LFG_ERR_ARG;
    try {
        ctx->last_error.clear();
        return body();
    } catch (const LfgError& e) {
        ctx->last_error = e.what();
        return e.code;
    } catch (const std::exception& e) {
        ctx->last_error = e.what();
        return LFG_ERR_ARG;
    }
}

void require(bool cond, int code, const char* msg) {
    if (!cond) throw LfgError(code, msg);
}

lfg::MeshView mesh_view(const lfg_ctx* c) {
    return lfg::MeshView{c->n_nodes,     c->n_elems,     c->xy.p,         c->enodes.p,
                         c->elem_mat.p,  c->materials.p, c->plane_mode,
                         c->mat_mode.p};
}

lfg::BarView bar_view(const lfg_ctx* c) {
    return lfg::BarView{c->n_bars, c->n_elems, c->bar_nodes.p, c->bar_mat.p, c->bar_area.p, c->bar_map.p};
}

// node -> (element << 2 | local node) in ascending element order (Q4, then bars), for
// smoothing
void build_node_elements(lfg_ctx* c) {
    const int n_nodes = c->n_nodes, n_q4 = c->n_elems, n_bars = c->n_bars;
    std::vector<int> ptr(static_cast<size_t>(n_nodes) + 1, 0);
    for (int k = 0; k < 4 * n_q4; ++k) ptr[static_cast<size_t>(c->h_enodes[k]) + 1]++;
    for (int k = 0; k < 2 * n_bars; ++k) ptr[static_cast<size_t>(c->h_bar_nodes[k]) + 1]++;
    for (int i = 0; i < n_nodes; ++i) ptr[i + 1] += ptr[i];
    std::vector<int> list(static_cast<size_t>(ptr[n_nodes]));
    std::vector<int> cur(ptr.begin(), ptr.end() - 1);
    for (int e = 0; e < n_q4; ++e)
        for (int l = 0; l < 4; ++l) list[cur[c->h_enodes[4 * e + l]]++] = (e << 2) | l;
    for (int t = 0; t < n_bars; ++t)
        for (int l = 0; l < 2; ++l) list[cur[c->h_bar_nodes[2 * t + l]]++] = ((n_q4 + t) << 2) | l;
    c->node_elem_ptr.upload(ptr.data(), ptr.size(), c->stream);
    c->node_elem.upload(list.data(), list.size(), c->stream);
    // the most (element, corner) entries of one block of the block-row assembly (the
    // colour-ordered lists hold the same entries per node)
    c->max_block_slots = 0;
    for (int n0 = 0; n0 < n_nodes; n0 += lfg::kAssembleBlockNodes) {
        const int n1 = std::min(n_nodes, n0 + lfg::kAssembleBlockNodes);
        c->max_block_slots = std::max(c->max_block_slots, ptr[n1] - ptr[n0]);
    }
}

// A Q4 mesh whose elements form a W x H grid numbered row by row with counter-clockwise
// corners from the lower left (element ey W + ex: nodes n, n + 1, n + W + 2, n + W + 1,
// n = ey (W + 1) + ex) sets grid_w / grid_h; any other mesh leaves them 0.
void detect_element_grid(lfg_ctx* c) {
    c->grid_w = c->grid_h = 0;
    const long long n_e = c->n_elems;
    if (n_e == 0) return;
    const int* en = c->h_enodes.data();
    const long long w = static_cast<long long>(en[3]) - en[0] - 1;
    if (w <= 0 || n_e % w != 0 || en[0] != 0) return;
    const long long h = n_e / w;
    if ((w + 1) * (h + 1) != c->n_nodes) return;
    for (long long e = 0; e < n_e; ++e) {
        const long long n = (e / w) * (w + 1) + e % w;
        if (en[4 * e] != n || en[4 * e + 1] != n + 1 || en[4 * e + 2] != n + w + 2 || en[4 * e + 3] != n + w + 1)
            return;
    }
    c->grid_w = static_cast<int>(w);
    c->grid_h = static_cast<int>(h);
}

lfg::CsrView csr_view(const lfg_ctx* c) { return lfg::CsrView{c->n, c->nnz, c->row_ptr.p, c->col.p, c->max_row}; }

lfg::SliceView slice_view(const lfg_ctx* c) {
    return lfg::SliceView{c->n_slices,       c->slice_blocks,  c->slice_ptr.p, c->slice_col.p,
                          c->slice_max_width, c->slice_diag.p, c->pair_sum};
}

lfg::SymSliceView sym_view(const lfg_ctx* c) {
    return lfg::SymSliceView{c->n_slices,   c->sym_ublocks,  c->sym_uptr.p,      c->sym_ucol.p,    c->sym_lptr.p,
                             c->sym_lcol.p, c->sym_ltpos.p,  c->sym_max_uwidth, c->sym_max_lwidth, c->sym_diag.p};
}

// Sets the block storage summation order for the products enqueued in its scope.
struct PairScope {
    lfg_ctx* c;
    PairScope(lfg_ctx* ctx, bool on) : c(ctx) { c->pair_sum = on ? 1 : 0; }
    ~PairScope() { c->pair_sum = 0; }
};

constexpr int kSliceVariant = 6;  // node-slice SpMV, separate direction kernel
constexpr int kSliceFused = 7;    // node-slice SpMV forming the Jacobi / block direction on the fly
constexpr int kStencilSym = 8;    // node-stencil planes, symmetric form (diagonal + positive offsets)
constexpr int kStencilExact = 9;  // node-stencil planes, every offset (CSR column order)
constexpr int kStencilFused = 10; // symmetric node stencil, whole Jacobi / block PCG iteration in one launch
constexpr int kStencilGrid = 11;  // the iteration of 10 on 2D tiles (row-structured stencils)
constexpr int kSliceIter = 12;    // node slices, whole Jacobi / block PCG iteration in one launch
constexpr int kSliceSplit = 13;   // the iteration of 12 in two launches (vector update, multiply)
constexpr int kSliceFold = 14;    // the iteration of 12, its reduction folded by the next launch
constexpr int kSliceSym = 15;     // the iteration of 13 on the symmetric node slices
constexpr int kMaxVariant = kSliceSym;
constexpr int kSliceIterMaxSlices = 1024;  // above this the split iteration (13) is the faster default
static_assert(kSliceIterMaxSlices <= lfg::kFoldMaxSlices, "variant 14 serves the one-launch default");
constexpr int kFoldPartials = 5 * lfg::kFoldMaxBlocks;  // one set of variant 14's partial sums

bool slice_layout(int variant) {
    return variant == kSliceVariant || variant == kSliceFused || variant == kSliceIter || variant == kSliceSplit ||
           variant == kSliceFold;
}
bool stencil_layout(int variant) {
    return variant == kStencilSym || variant == kStencilExact || variant == kStencilFused || variant == kStencilGrid;
}
// The whole Jacobi / block iteration in one launch (1D or 2D stencil tiles, node slices)
// or two (variants 13, 15).
bool fused_iteration(int variant) {
    return variant == kStencilFused || variant == kStencilGrid || variant == kSliceIter || variant == kSliceSplit ||
           variant == kSliceFold || variant == kSliceSym;
}
// The node-slice iterations of one or two launches per iteration (Jacobi and block only).
bool slice_iteration_variant(int variant) {
    return variant == kSliceIter || variant == kSliceSplit || variant == kSliceFold || variant == kSliceSym;
}
bool grid_stencil_ok(const lfg_ctx* c) { return c->have_stencil_sym && c->stencil_sym.row_len > 0; }
bool fused_stencil_ok(const lfg_ctx* c) { return c->have_stencil_sym && c->stencil_sym.win_total > 0; }
// The fused node-slice iteration for Jacobi / block on a node-blocked pattern: one launch
// per iteration, its sums folded by the next launch, for small and medium systems; two
// launches for large ones.
int slice_iteration(const lfg_ctx* c) { return c->n_slices <= kSliceIterMaxSlices ? kSliceFold : kSliceSplit; }

// The host lists of the symmetric node-slice layout (SymSliceView), built from the
// pattern on first use: per node the blocks before its diagonal ("lower") and from it on
// ("upper"), per slice as wide as its widest node, and for every lower block (i, j) the
// position of the upper block (j, i) it is read as. False (remembered) when the pattern
// is not node-blocked or some block has no mirror.
bool sym_slices_ok(const lfg_ctx* c) {
    if (c->sym_state != 0) return c->sym_state > 0;
    c->sym_state = -1;
    if (!c->have_slices) return false;
    const int n_nodes = c->n / 2;
    const std::vector<int>& rp = c->h_row_ptr;
    const std::vector<int>& col = c->h_col;
    std::vector<int> deg(static_cast<size_t>(n_nodes)), dlo(static_cast<size_t>(n_nodes), 0);
    for (int i = 0; i < n_nodes; ++i) {
        const int r0 = rp[2 * i], r1 = rp[2 * i + 1];
        if (r1 - r0 == 1) {  // a node in no element: its diagonal block alone (build_slices checked it)
            deg[i] = 1;
            continue;
        }
        deg[i] = (r1 - r0) / 2;
        int d = 0;
        while (d < deg[i] && col[r0 + 2 * d] < 2 * i) ++d;
        dlo[i] = d;
    }
    const int n_slices = c->n_slices;
    std::vector<int> uptr(static_cast<size_t>(n_slices) + 1, 0), lptr(static_cast<size_t>(n_slices) + 1, 0);
    long long ut = 0, lt = 0;
    int max_u = 0, max_l = 0;
    for (int s = 0; s < n_slices; ++s) {
        int uw = 0, lw = 0;
        for (int i = s * 64; i < std::min(n_nodes, s * 64 + 64); ++i) {
            uw = std::max(uw, deg[i] - dlo[i]);
            lw = std::max(lw, dlo[i]);
        }
        max_u = std::max(max_u, uw);
        max_l = std::max(max_l, lw);
        ut += 64LL * uw;
        lt += 64LL * lw;
        if (ut >= (1LL << 31) - 1 || lt >= (1LL << 31)) return false;  // positions are int32
        uptr[s + 1] = static_cast<int>(ut);
        lptr[s + 1] = static_cast<int>(lt);
    }
    std::vector<int> ucol(static_cast<size_t>(ut)), lcol(static_cast<size_t>(lt)), ltpos(static_cast<size_t>(lt));
    for (int s = 0; s < n_slices; ++s) {
        const int uw = (uptr[s + 1] - uptr[s]) / 64, lw = (lptr[s + 1] - lptr[s]) / 64;
        for (int lane = 0; lane < 64; ++lane) {
            const int node = s * 64 + lane;
            const int self = std::min(node, n_nodes - 1);
            const bool real = node < n_nodes;
            const int r0 = real ? rp[2 * node] : 0;
            const bool lone = real && rp[2 * node + 1] - r0 == 1;
            for (int k = 0; k < uw; ++k) {
                int j = self;  // padding: a zero block times the node's own pair
                if (real && k < deg[node] - dlo[node]) j = lone ? node : col[r0 + 2 * (dlo[node] + k)] >> 1;
                ucol[static_cast<size_t>(uptr[s]) + 64 * k + lane] = j;
            }
            for (int k = 0; k < lw; ++k) {
                const size_t at = static_cast<size_t>(lptr[s]) + 64 * k + lane;
                lcol[at] = self;
                ltpos[at] = static_cast<int>(ut);  // padding: the zero block
                if (!real || k >= dlo[node]) continue;
                const int j = col[r0 + 2 * k] >> 1;  // j < node
                // block (j, node) among node j's column pairs (ascending)
                int lo = 0, hi = deg[j];
                const int rj = rp[2 * j];
                while (lo < hi) {
                    const int mid = (lo + hi) / 2;
                    if (col[rj + 2 * mid] < 2 * node)
                        lo = mid + 1;
                    else
                        hi = mid;
                }
                if (lo >= deg[j] || rp[2 * j + 1] - rj == 1 || col[rj + 2 * lo] != 2 * node) return false;
                const int u = lo - dlo[j];  // >= 0: node > j
                lcol[at] = j;
                ltpos[at] = uptr[j / 64] + 64 * u + j % 64;
            }
        }
    }
    c->sym_uptr.upload(uptr.data(), uptr.size(), c->stream);
    c->sym_lptr.upload(lptr.data(), lptr.size(), c->stream);
    c->sym_ucol.upload(ucol.data(), ucol.size(), c->stream);
    c->sym_lcol.upload(lcol.data(), lcol.size(), c->stream);
    c->sym_ltpos.upload(ltpos.data(), ltpos.size(), c->stream);
    LFG_HIP(hipStreamSynchronize(c->stream));  // the host lists end with this scope
    c->sym_ublocks = ut;
    c->sym_max_uwidth = max_u;
    c->sym_max_lwidth = max_l;
    c->sym_state = 1;
    return true;
}

// 0 selects, in order: for the Jacobi and block preconditioners the folded node-slice
// iteration up to kSliceIterMaxSlices slices (the fastest at those sizes whatever the
// numbering: 7.7-9.9 us per iteration from 50x50 to 250x250 against 8.5-13 us for the
// stencil iterations), then the fused stencil iterations; the symmetric node-stencil
// multiply, the node-slice multiply, row-segment streaming.
// Preconditioners applied by a separate sweep kernel into z (not per node in the
// update / direction / fused kernels).
bool sweep_precond(int precond) { return precond == LFG_PRECOND_IC0 || precond == LFG_PRECOND_SUBDOMAIN; }

bool small_slices(const lfg_ctx* c) { return c->have_slices && c->n_slices <= kSliceIterMaxSlices; }

int resolve_variant(const lfg_ctx* c, int variant, int precond = LFG_PRECOND_JACOBI, bool block_order = false) {
    if (variant == 0 && small_slices(c) && !sweep_precond(precond)) return resolve_variant(c, kSliceFold, precond, block_order);
    if (block_order) {
        const int v = variant == 0 ? (!sweep_precond(precond) && grid_stencil_ok(c)
                                          ? kStencilGrid
                                          : (fused_stencil_ok(c) && !sweep_precond(precond)
                                                 ? kStencilFused
                                                 : (c->have_stencil_sym ? kStencilSym
                                                                        : (sweep_precond(precond) ? kSliceVariant
                                                                                                  : slice_iteration(c)))))
                                   : variant;
        require(v >= kSliceVariant && v <= kSliceFold, LFG_ERR_UNSUPPORTED,
                "block storage runs on the node-slice (6, 7, 12-14) and node-stencil (8-11) SpMVs");
        return resolve_variant(c, v, precond, false);
    }
    if (variant == 0) {
        if (grid_stencil_ok(c) && !sweep_precond(precond)) return kStencilGrid;
        if (fused_stencil_ok(c) && !sweep_precond(precond)) return kStencilFused;
        if (c->have_stencil_sym) return kStencilSym;
        if (c->have_slices) return sweep_precond(precond) ? kSliceVariant : slice_iteration(c);
        return 4;
    }
    require(variant >= 1 && variant <= kMaxVariant, LFG_ERR_ARG, "SpMV variant must be 0-15");
    require(!slice_iteration_variant(variant) || !sweep_precond(precond), LFG_ERR_UNSUPPORTED,
            "the fused node-slice iteration applies the jacobi and 2x2 block preconditioners, not ic0 or subdomains");
    require(variant != kStencilGrid || grid_stencil_ok(c), LFG_ERR_UNSUPPORTED,
            "the 2D-tiled fused iteration needs a symmetric node stencil with positive offsets 1, L-1, L, L+1 "
            "(L >= 64 nodes per mesh row)");
    require(variant != kStencilGrid || !sweep_precond(precond), LFG_ERR_UNSUPPORTED,
            "the fused stencil iteration applies the jacobi and 2x2 block preconditioners, not ic0 or subdomains");
    require(variant != kStencilFused || fused_stencil_ok(c), LFG_ERR_UNSUPPORTED,
            "the fused stencil iteration needs the symmetric node-stencil layout with offset windows that fit "
            "in LDS");
    require(variant != kStencilFused || !sweep_precond(precond), LFG_ERR_UNSUPPORTED,
            "the fused stencil iteration applies the jacobi and 2x2 block preconditioners, not ic0 or subdomains");
    require(!slice_layout(variant) || c->have_slices, LFG_ERR_UNSUPPORTED,
            "the node-slice SpMV needs rows 2i, 2i+1 sharing one column list of dof pairs");
    require(variant != kSliceFold || c->n_slices <= kSliceIterMaxSlices, LFG_ERR_UNSUPPORTED,
            "the folded node-slice iteration (14) takes at most 1,024 slices of 64 nodes; use 12 or 13");
    require(variant != kSliceSym || sym_slices_ok(c), LFG_ERR_UNSUPPORTED,
            "the symmetric node-slice layout (15) needs a node-blocked, structurally symmetric pattern");
    require(variant != kStencilSym || c->have_stencil_sym, LFG_ERR_UNSUPPORTED,
            "the symmetric node-stencil SpMV needs a symmetric node-blocked pattern with at most 8 positive node "
            "offsets");
    require(variant != kStencilExact || c->have_stencil_exact, LFG_ERR_UNSUPPORTED,
            "the node-stencil SpMV needs a node-blocked pattern with at most 17 node offsets");
    return variant;
}

// Node-blocked patterns (every two-dof-per-node pattern): builds the slice offsets
// on the host and the node columns on the device. A node in no element has only its
// two diagonal entries; it becomes one diagonal block.
void build_slices(lfg_ctx* c) {
    c->have_slices = false;
    c->slices64_valid = c->slices32_valid = false;
    c->sym_state = 0;
    c->sym64_valid = c->sym32_valid = false;
    const int n_nodes = c->n / 2;
    const std::vector<int>& rp = c->h_row_ptr;
    const std::vector<int>& col = c->h_col;
    std::vector<int> deg(static_cast<size_t>(n_nodes));
    for (int i = 0; i < n_nodes; ++i) {
        const int r0 = rp[2 * i], r1 = rp[2 * i + 1], r2 = rp[2 * i + 2];
        const int len = r1 - r0;
        if (len == 1 && r2 - r1 == 1 && col[r0] == 2 * i && col[r1] == 2 * i + 1) {  // node in no element
            deg[i] = 1;
            continue;
        }
        if (len != r2 - r1 || (len & 1)) return;
        for (int k = 0; k < len; k += 2) {
            const int c0 = col[r0 + k];
            if ((c0 & 1) || col[r0 + k + 1] != c0 + 1 || col[r1 + k] != c0 || col[r1 + k + 1] != c0 + 1) return;
        }
        deg[i] = len / 2;
    }
    const int n_slices = (n_nodes + 63) / 64;
    std::vector<int> ptr(static_cast<size_t>(n_slices) + 1, 0);
    long long total = 0;
    int max_width = 0;
    for (int s = 0; s < n_slices; ++s) {
        int w = 0;
        for (int i = s * 64; i < std::min(n_nodes, s * 64 + 64); ++i) w = std::max(w, deg[i]);
        max_width = std::max(max_width, w);
        total += 64LL * w;
        if (total >= (1LL << 31)) return;  // offsets are int32
        ptr[s + 1] = static_cast<int>(total);
    }
    c->n_slices = n_slices;
    c->slice_blocks = total;
    c->slice_max_width = max_width;
    c->slice_ptr.upload(ptr.data(), ptr.size(), c->stream);
    c->slice_col.alloc(static_cast<size_t>(total));
    lfg::launch_pack_slices(c->stream, csr_view(c), slice_view(c), nullptr, nullptr, nullptr, c->slice_col.p);
    c->have_slices = true;
}

// LDS windows of the fused stencil iteration: the offsets -off[k], 0, off[k] in
// clusters (a gap above 64 starts a new one), one window of 256 + (hi - lo) nodes per
// cluster. Leaves win_total = 0 when the windows exceed the LDS budget.
void plan_fused_windows(lfg::StencilView& v) {
    std::vector<int> all;
    for (int k = v.n_off - 1; k >= 1; --k) all.push_back(-v.off[k]);
    for (int k = 0; k < v.n_off; ++k) all.push_back(v.off[k]);
    v.n_clusters = 0;
    v.win_total = 0;
    int base = 0;
    for (size_t a = 0; a < all.size();) {
        size_t e = a;
        while (e + 1 < all.size() && all[e + 1] - all[e] <= 64) ++e;
        const int c = v.n_clusters++;
        v.win_lo[c] = all[a];
        v.win_hi[c] = all[e];
        v.win_base[c] = base;
        base += lfg::kFusedTile + all[e] - all[a];
        a = e + 1;
    }
    auto slot = [&](int d) {
        for (int c = 0; c < v.n_clusters; ++c)
            if (d >= v.win_lo[c] && d <= v.win_hi[c]) return v.win_base[c] + d - v.win_lo[c];
        return -1;
    };
    v.slot_zero = slot(0);
    for (int k = 0; k < v.n_off; ++k) {
        v.slot_up[k] = slot(v.off[k]);
        v.slot_lo[k] = slot(-v.off[k]);
    }
    if (base <= lfg::kFusedMaxWindow) v.win_total = base;
}

// After build_slices (node-blocked patterns): the node offsets j - i of all blocks
// (i, j). Up to kMaxStencil offsets give the exact stencil form; a symmetric pattern
// with at most 8 positive offsets also gives the symmetric form.
void build_stencil(lfg_ctx* c) {
    c->have_stencil_sym = c->have_stencil_exact = false;
    c->stencil64_valid = c->stencil32_valid = c->stencil_exact_valid = false;
    if (!c->have_slices) return;
    const int n_nodes = c->n / 2;
    const std::vector<int>& rp = c->h_row_ptr;
    const std::vector<int>& col = c->h_col;
    std::vector<int> offs;
    auto add = [&](int d) {
        if (std::find(offs.begin(), offs.end(), d) != offs.end()) return true;
        if (static_cast<int>(offs.size()) >= lfg::kMaxStencil) return false;
        offs.push_back(d);
        return true;
    };
    if (!add(0)) return;
    for (int i = 0; i < n_nodes; ++i) {
        const int r0 = rp[2 * i], r1 = rp[2 * i + 1];
        if (r1 - r0 == 1) continue;  // node in no element: diagonal only
        for (int k = r0; k < r1; k += 2)
            if (!add((col[k] >> 1) - i)) return;
    }
    std::sort(offs.begin(), offs.end());
    lfg::StencilView ex{};
    ex.n_nodes = n_nodes;
    ex.n_off = static_cast<int>(offs.size());
    ex.sym = 0;
    for (int k = 0; k < ex.n_off; ++k) ex.off[k] = offs[k];
    c->stencil_exact = ex;
    c->have_stencil_exact = true;
    // symmetric form: the pattern must hold block (j, i) for every block (i, j)
    std::vector<int> pos;
    for (int d : offs)
        if (d > 0) pos.push_back(d);
    if (pos.size() > 8) return;
    for (int d : pos)
        if (std::find(offs.begin(), offs.end(), -d) == offs.end()) return;
    auto has_block = [&](int i, int j) {
        const int r0 = rp[2 * i], r1 = rp[2 * i + 1];
        if (r1 - r0 == 1) return j == i;
        int lo = 0, hi = (r1 - r0) / 2;
        while (lo < hi) {
            const int mid = (lo + hi) / 2;
            if ((col[r0 + 2 * mid] >> 1) < j) lo = mid + 1;
            else hi = mid;
        }
        return lo < (r1 - r0) / 2 && (col[r0 + 2 * lo] >> 1) == j;
    };
    for (int i = 0; i < n_nodes; ++i) {
        const int r0 = rp[2 * i], r1 = rp[2 * i + 1];
        if (r1 - r0 == 1) continue;
        for (int k = r0; k < r1; k += 2) {
            const int j = col[k] >> 1;
            if (j > i && !has_block(j, i)) return;
        }
    }
    lfg::StencilView sy{};
    sy.n_nodes = n_nodes;
    sy.n_off = 1 + static_cast<int>(pos.size());
    sy.sym = 1;
    sy.off[0] = 0;
    for (size_t k = 0; k < pos.size(); ++k) sy.off[k + 1] = pos[k];
    plan_fused_windows(sy);
    if (sy.n_off == 5 && sy.off[1] == 1 && sy.off[3] >= 64 && sy.off[2] == sy.off[3] - 1 &&
        sy.off[4] == sy.off[3] + 1)
        sy.row_len = sy.off[3];
    c->stencil_sym = sy;
    c->have_stencil_sym = true;
}

// Packs the constrained values into the stencil planes that the next SpMVs read.
void ensure_stencil(lfg_ctx* c, int variant, bool need64, bool need32) {
    if (fused_iteration(variant)) variant = kStencilSym;
    const size_t nn = 64 * static_cast<size_t>((c->n / 2 + 63) / 64);  // whole 64-node chunks
    if (variant == kStencilExact) {
        if (!c->stencil_exact_valid) {
            c->stencil_exact64.alloc(4 * nn * c->stencil_exact.n_off);
            lfg::launch_pack_stencil(c->stream, csr_view(c), c->stencil_exact, c->Kc.p, c->stencil_exact64.p, nullptr,
                                     nullptr);
            c->stencil_exact_valid = true;
        }
        return;
    }
    if (need64 && !c->stencil64_valid) {
        c->stencil64.alloc(4 * nn * c->stencil_sym.n_off);
        lfg::launch_pack_stencil(c->stream, csr_view(c), c->stencil_sym, c->Kc.p, c->stencil64.p, nullptr, nullptr);
        c->stencil64_valid = true;
    }
    if (need32 && !c->stencil32_valid) {
        c->stencil32.alloc(4 * nn * std::max(1, c->stencil_sym.n_off - 1));
        c->stencil_diag.alloc(4 * nn);
        lfg::launch_pack_stencil(c->stream, csr_view(c), c->stencil_sym, c->Kc.p, nullptr, c->stencil32.p,
                                 c->stencil_diag.p);
        c->stencil32_valid = true;
    }
}

// The multiply of one PCG iteration for a stencil variant (q = A p, p.q, alpha). With
// mixed, the symmetric form switches to its FP32 copy in the FP32 phase.
void launch_stencil_spmv(lfg_ctx* c, const lfg::PcgBuffers& b, int variant, bool fp64, bool fp32) {
    if (fused_iteration(variant)) variant = kStencilSym;  // the plain multiply of the same planes
    if (variant == kStencilExact) {
        lfg::StencilView ex = c->stencil_exact;
        ex.pair_sum = c->pair_sum;
        lfg::launch_pcg_spmv_stencil(c->stream, b, ex, c->stencil_exact64.p, nullptr, nullptr);
    }
    else
        lfg::launch_pcg_spmv_stencil(c->stream, b, c->stencil_sym, fp64 ? c->stencil64.p : nullptr,
                                     fp32 ? c->stencil32.p : nullptr, fp32 ? c->stencil_diag.p : nullptr);
}

// Packs the constrained values into the slice planes that the next SpMVs read.
void ensure_slices(lfg_ctx* c, bool need64, bool need32) {
    if (need64 && !c->slices64_valid) {
        c->slices64.alloc(4 * static_cast<size_t>(c->slice_blocks));
        lfg::launch_pack_slices(c->stream, csr_view(c), slice_view(c), c->Kc.p, c->slices64.p, nullptr, nullptr);
        c->slices64_valid = true;
    }
    if (need32 && !c->slices32_valid) {
        c->slices32.alloc(4 * static_cast<size_t>(c->slice_blocks));
        c->slice_diag.alloc(2 * static_cast<size_t>(c->n));
        lfg::launch_pack_slices(c->stream, csr_view(c), slice_view(c), c->Kc.p, nullptr, c->slices32.p, nullptr,
                                c->slice_diag.p);
        c->slices32_valid = true;
    }
}

// Packs the constrained values into the symmetric slice planes (variant 15): FP64, and the
// FP32 copy with its compensated diagonal.
void ensure_sym_slices(lfg_ctx* c, bool need64, bool need32) {
    require(sym_slices_ok(c), LFG_ERR_UNSUPPORTED,
            "the symmetric node-slice layout (15) needs a node-blocked, structurally symmetric pattern");
    if (need64 && !c->sym64_valid) {
        c->sym64.alloc(4 * static_cast<size_t>(c->sym_ublocks + 1));
        lfg::launch_pack_sym_slices(c->stream, csr_view(c), sym_view(c), c->Kc.p, c->sym64.p, nullptr);
        c->sym64_valid = true;
    }
    if (need32 && !c->sym32_valid) {
        c->sym32.alloc(4 * static_cast<size_t>(c->sym_ublocks + 1));
        c->sym_diag.alloc(2 * static_cast<size_t>(c->n));
        lfg::launch_pack_sym_slices(c->stream, csr_view(c), sym_view(c), c->Kc.p, nullptr, c->sym32.p);
        c->sym32_valid = true;
    }
}

void sync(lfg_ctx* c) { LFG_HIP(hipStreamSynchronize(c->stream)); }

float elapsed(hipEvent_t a, hipEvent_t b) {
    float ms = 0.0f;
    LFG_HIP(hipEventElapsedTime(&ms, a, b));
    return ms;
}

// Stable counting sort of the element->value map by target: for every CSR value,
// the (element, local entry) pairs that add into it, in ascending element order.
// Contribution k < 64 n_elems is Q4 entry k; 64 n_elems + 16 t + j is entry j of bar t
// (the bar matrices follow the Q4 ones in the element-matrix buffer).
void build_gather_map(lfg_ctx* c, const int32_t* emap, const int32_t* bmap, int n_bars) {
    const long long q4 = 64LL * c->n_elems;
    const long long total = q4 + 16LL * n_bars;
    auto target = [&](long long k) { return static_cast<size_t>(k < q4 ? emap[k] : bmap[k - q4]); };
    std::vector<int> counts(static_cast<size_t>(c->nnz) + 1, 0);
    for (long long k = 0; k < total; ++k) counts[target(k) + 1]++;
    for (long long k = 0; k < c->nnz; ++k) counts[static_cast<size_t>(k) + 1] += counts[static_cast<size_t>(k)];
    std::vector<int> contrib(static_cast<size_t>(total));
    std::vector<int> cursor(counts.begin(), counts.end() - 1);
    for (long long k = 0; k < total; ++k) contrib[static_cast<size_t>(cursor[target(k)]++)] = static_cast<int>(k);
    c->cptr.upload(counts.data(), counts.size(), c->stream);
    c->contrib.upload(contrib.data(), contrib.size(), c->stream);
    sync(c);
    c->gather_valid = true;
}

// The gather map of the serial assembly with bars, built on first use from the maps on
// the device (meshes of Q4 elements only assemble row by row and never need it).
void ensure_gather_map(lfg_ctx* c) {
    if (c->gather_valid) return;
    std::vector<int32_t> emap(64 * static_cast<size_t>(c->n_elems)), bmap(16 * static_cast<size_t>(c->n_bars));
    c->emap.download(emap.data(), emap.size(), c->stream);
    c->bar_map.download(bmap.data(), bmap.size(), c->stream);
    sync(c);
    build_gather_map(c, emap.data(), bmap.data(), c->n_bars);
}

using lfg_ic0::split_levels;

// The node form of the two sweeps (Ic0NodeSweep, lfg_ic0::node_plan) when rows 2i, 2i+1
// of the factor have the same dependency nodes; otherwise the row form serves.
void build_ic0_nodes(lfg_ctx* c, const std::vector<int>& lrow, const std::vector<int>& lcol,
                     const std::vector<double>& lval, const std::vector<int>& urow, const std::vector<int>& ucol,
                     const std::vector<double>& uval, const std::vector<double>& inv_diag) {
    Ic0Device& d = c->ic0;
    d.nfwd.ready = d.nbwd.ready = false;
    const int nn = c->n / 2;
    auto node_form = [&](bool backward, const std::vector<int>& ptr, const std::vector<int>& col,
                         const std::vector<double>& val, Ic0NodeSweepDevice& out) {
        lfg_ic0::NodePlan plan;
        if (!lfg_ic0::node_plan(nn, backward, ptr.data(), col.data(), plan)) return false;
        std::vector<double> qvf, qvs, qcross(static_cast<size_t>(nn)), qsf(static_cast<size_t>(nn)),
            qss(static_cast<size_t>(nn));
        qvf.reserve(2 * plan.key.size());
        qvs.reserve(2 * plan.key.size());
        for (int q = 0; q < nn; ++q) {
            const int i = plan.ord[q];
            const int count = plan.qptr[q + 1] - plan.qptr[q];
            const int f = plan.first_at[q], s2 = plan.second_at[q];
            for (int u = 0; u < count; ++u) {
                qvf.push_back(val[f + 2 * u]);
                qvf.push_back(val[f + 2 * u + 1]);
                qvs.push_back(val[s2 + 2 * u]);
                qvs.push_back(val[s2 + 2 * u + 1]);
            }
            qcross[q] = val[s2 + 2 * count];
            qsf[q] = inv_diag[backward ? 2 * i + 1 : 2 * i];
            qss[q] = inv_diag[backward ? 2 * i : 2 * i + 1];
        }
        out.ord.upload(plan.ord.data(), plan.ord.size(), c->stream);
        out.ptr.upload(plan.qptr.data(), plan.qptr.size(), c->stream);
        out.key.upload(plan.key.data(), plan.key.size(), c->stream);
        out.v_first.upload(qvf.data(), qvf.size(), c->stream);
        out.v_second.upload(qvs.data(), qvs.size(), c->stream);
        out.cross.upload(qcross.data(), qcross.size(), c->stream);
        out.scale_first.upload(qsf.data(), qsf.size(), c->stream);
        out.scale_second.upload(qss.data(), qss.size(), c->stream);
        out.level_ptr.upload(plan.parts.data(), plan.parts.size(), c->stream);
        out.levels = static_cast<int>(plan.parts.size()) - 1;
        out.reach = plan.reach;
        out.widest = plan.widest;
        out.ready = true;
        return true;
    };
    if (nn == 0 || !node_form(false, lrow, lcol, lval, d.nfwd) || !node_form(true, urow, ucol, uval, d.nbwd))
        d.nfwd.ready = d.nbwd.ready = false;
}

// IC(0) of the lower triangle (precond.cpp:298-384), level sets for the device
// sweeps, and the transposed strict lower part with rows in descending order.
void build_ic0(lfg_ctx* c) {
    const int n = c->n;
    std::vector<double> kc(static_cast<size_t>(c->nnz));
    c->Kc.download(kc.data(), kc.size(), c->stream);
    sync(c);
    const lfg_ic0::Pattern pat = lfg_ic0::pattern(n, c->h_row_ptr.data(), c->h_col.data());
    const std::vector<int>& lrow = pat.lrow;
    const std::vector<int>& lcol = pat.lcol;
    const std::vector<int>& diag_index = pat.diag;
    std::vector<double> lval(pat.lcol.size()), inv_diag(static_cast<size_t>(n), 1.0);
    for (size_t k = 0; k < lval.size(); ++k) lval[k] = pat.lsrc[k] >= 0 ? kc[static_cast<size_t>(pat.lsrc[k])] : 1.0;
    double scale = 0.0;
    for (int row = 0; row < n; ++row) scale = std::max(scale, std::abs(lval[diag_index[row]]));
    const double floor_value = std::max(1.0e-30, 1.0e-14 * scale);
    long long breakdowns = 0;
    for (int row = 0; row < n; ++row) {
        const int begin = lrow[row], d = diag_index[row];
        for (int index = begin; index < d; ++index) {
            const int column = lcol[index];
            const int other_diag = diag_index[column];
            double sum = lval[index];
            int left = begin, right = lrow[column];
            while (left < index && right < other_diag) {
                const int lc = lcol[left], rc = lcol[right];
                if (lc < rc) ++left;
                else if (lc > rc) ++right;
                else {
                    sum -= lval[left] * lval[right];
                    ++left;
                    ++right;
                }
            }
            lval[index] = sum * inv_diag[column];
        }
        double sum = lval[d];
        for (int index = begin; index < d; ++index) sum -= lval[index] * lval[index];
        if (!(sum > 0.0) || !std::isfinite(sum)) {
            ++breakdowns;
            sum = floor_value;
        }
        const double pivot = std::sqrt(sum);
        lval[d] = pivot;
        inv_diag[row] = 1.0 / pivot;
    }
    // forward levels
    std::vector<int> level(static_cast<size_t>(n), 0);
    int fwd_levels = 0;
    for (int row = 0; row < n; ++row) {
        int lv = 0;
        for (int k = lrow[row]; k < lrow[row + 1] - 1; ++k) lv = std::max(lv, level[lcol[k]] + 1);
        level[row] = lv;
        fwd_levels = std::max(fwd_levels, lv + 1);
    }
    std::vector<int> fptr(static_cast<size_t>(fwd_levels) + 1, 0), frows(static_cast<size_t>(n));
    for (int row = 0; row < n; ++row) fptr[level[row] + 1]++;
    for (int l = 0; l < fwd_levels; ++l) fptr[l + 1] += fptr[l];
    {
        std::vector<int> cur(fptr.begin(), fptr.end() - 1);
        for (int row = 0; row < n; ++row) frows[cur[level[row]]++] = row;
    }
    // transpose of the strict lower part, rows descending per column
    const std::vector<int>& urow = pat.urow;
    const std::vector<int>& ucol = pat.ucol;
    std::vector<double> uval(pat.usrc.size());
    for (size_t k = 0; k < uval.size(); ++k) uval[k] = lval[static_cast<size_t>(pat.usrc[k])];
    // backward levels
    std::vector<int> blevel(static_cast<size_t>(n), 0);
    int bwd_levels = 0;
    for (int cc = n - 1; cc >= 0; --cc) {
        int lv = 0;
        for (int k = urow[cc]; k < urow[cc + 1]; ++k) lv = std::max(lv, blevel[ucol[k]] + 1);
        blevel[cc] = lv;
        bwd_levels = std::max(bwd_levels, lv + 1);
    }
    std::vector<int> bptr(static_cast<size_t>(bwd_levels) + 1, 0), brows(static_cast<size_t>(n));
    for (int cc = 0; cc < n; ++cc) bptr[blevel[cc] + 1]++;
    for (int l = 0; l < bwd_levels; ++l) bptr[l + 1] += bptr[l];
    {
        std::vector<int> cur(bptr.begin(), bptr.end() - 1);
        for (int cc = 0; cc < n; ++cc) brows[cur[blevel[cc]]++] = cc;
    }
    Ic0Device& d = c->ic0;
    d.lrow.upload(lrow.data(), lrow.size(), c->stream);
    d.lcol.upload(lcol.data(), lcol.size(), c->stream);
    d.lval.upload(lval.data(), lval.size(), c->stream);
    d.inv_diag.upload(inv_diag.data(), inv_diag.size(), c->stream);
    d.urow.upload(urow.data(), urow.size(), c->stream);
    d.ucol.upload(ucol.data(), ucol.size(), c->stream);
    d.uval.upload(uval.data(), uval.size(), c->stream);
    d.fptr.upload(fptr.data(), fptr.size(), c->stream);
    d.frows.upload(frows.data(), frows.size(), c->stream);
    d.bptr.upload(bptr.data(), bptr.size(), c->stream);
    d.brows.upload(brows.data(), brows.size(), c->stream);
    d.fwd_levels = fwd_levels;
    d.bwd_levels = bwd_levels;
    // level-ordered sweeps: positions in level order (rows of a level ascending, i.e. the
    // level lists above), dependencies rewritten as positions, in the reference's order
    auto level_order = [&](const std::vector<int>& rows_by_level, const std::vector<int>& level_ptr,
                           const std::vector<int>& dptr, const std::vector<int>& dcol,
                           const std::vector<double>& dval, bool strict_lower, Ic0SweepDevice& out) {
        std::vector<int> pos(static_cast<size_t>(n));
        for (int q = 0; q < n; ++q) pos[rows_by_level[q]] = q;
        std::vector<int> ptr(static_cast<size_t>(n) + 1, 0), key;
        std::vector<double> val, scale(static_cast<size_t>(n));
        key.reserve(dcol.size());
        val.reserve(dcol.size());
        int reach = 0;
        for (int q = 0; q < n; ++q) {
            const int r = rows_by_level[q];
            const int end = strict_lower ? dptr[r + 1] - 1 : dptr[r + 1];
            for (int k = dptr[r]; k < end; ++k) {
                key.push_back(pos[dcol[k]]);
                val.push_back(dval[k]);
                reach = std::max(reach, q - pos[dcol[k]]);
            }
            ptr[q + 1] = static_cast<int>(key.size());
            scale[q] = inv_diag[r];
        }
        out.ord.upload(rows_by_level.data(), rows_by_level.size(), c->stream);
        out.ptr.upload(ptr.data(), ptr.size(), c->stream);
        out.key.upload(key.data(), key.size(), c->stream);
        out.val.upload(val.data(), val.size(), c->stream);
        out.scale.upload(scale.data(), scale.size(), c->stream);
        const std::vector<int> parts = split_levels(level_ptr);
        out.level_ptr.upload(parts.data(), parts.size(), c->stream);
        out.levels = static_cast<int>(parts.size()) - 1;
        out.reach = reach;
        out.widest = 0;
        for (int l = 0; l < out.levels; ++l) out.widest = std::max(out.widest, parts[l + 1] - parts[l]);
    };
    level_order(frows, fptr, lrow, lcol, lval, true, d.fwd);
    level_order(brows, bptr, urow, ucol, uval, false, d.bwd);
    build_ic0_nodes(c, lrow, lcol, lval, urow, ucol, uval, inv_diag);
    d.breakdowns = breakdowns;
    d.ready = true;
    sync(c);
}

// SubdomainBlockJacobi set-up (precond.cpp:121-232) on the host, in the reference's
// order: each subdomain's dofs in matrix order, the dense block of its constrained
// values, an in-place row-by-row Cholesky (a failed block becomes the identity) and
// the packed lower factor with reciprocal pivots. Subdomains are independent, so
// host threads take them in interleaved order.
void build_subdomains(lfg_ctx* c) {
    require(c->n_sub > 0 && static_cast<int>(c->sub_node.size()) == c->n / 2, LFG_ERR_STATE,
            "lfg_set_subdomains must precede a subdomain block-Jacobi solve");
    const int n = c->n, n_sub = c->n_sub;
    std::vector<double> vals(static_cast<size_t>(c->nnz));
    c->Kc.download(vals.data(), vals.size(), c->stream);
    sync(c);
    const std::vector<int>& rp = c->h_row_ptr;
    const std::vector<int>& col = c->h_col;
    std::vector<int> owner(static_cast<size_t>(n));
    for (int d = 0; d < n; ++d) owner[d] = c->sub_node[d / 2];
    std::vector<int> dof_begin(static_cast<size_t>(n_sub) + 1, 0), dofs(static_cast<size_t>(n)),
        local(static_cast<size_t>(n));
    for (int d = 0; d < n; ++d) dof_begin[owner[d] + 1]++;
    for (int s = 0; s < n_sub; ++s) dof_begin[s + 1] += dof_begin[s];
    {
        std::vector<int> cur(dof_begin.begin(), dof_begin.end() - 1);
        for (int d = 0; d < n; ++d) {
            local[d] = cur[owner[d]] - dof_begin[owner[d]];
            dofs[cur[owner[d]]++] = d;
        }
    }
    std::vector<long long> fbeg(static_cast<size_t>(n_sub) + 1, 0);
    int max_size = 0;
    for (int s = 0; s < n_sub; ++s) {
        const long long size = dof_begin[s + 1] - dof_begin[s];
        fbeg[s + 1] = fbeg[s] + size * (size + 1) / 2;
        max_size = std::max(max_size, static_cast<int>(size));
    }
    require(max_size <= lfg::kMaxSubdomainDofs, LFG_ERR_UNSUPPORTED,
            "subdomains above 512 dofs (256 nodes) are outside this backend's block-Jacobi");
    std::vector<double> factor(static_cast<size_t>(fbeg[n_sub]) + 1, 0.0);
    std::vector<char> bad(static_cast<size_t>(n_sub), 0);
    auto factor_range = [&](int first, int step) {
        std::vector<double> dense(static_cast<size_t>(max_size) * max_size);
        for (int s = first; s < n_sub; s += step) {
            const int begin = dof_begin[s], size = dof_begin[s + 1] - begin;
            if (size == 0) continue;
            std::fill(dense.begin(), dense.begin() + static_cast<long long>(size) * size, 0.0);
            for (int i = 0; i < size; ++i) {
                const int dof = dofs[begin + i];
                for (int k = rp[dof]; k < rp[dof + 1]; ++k)
                    if (owner[col[k]] == s) dense[static_cast<size_t>(i) * size + local[col[k]]] = vals[k];
            }
            bool ok = true;
            for (int i = 0; i < size && ok; ++i) {
                for (int j = 0; j <= i; ++j) {
                    double sum = dense[static_cast<size_t>(i) * size + j];
                    for (int k = 0; k < j; ++k)
                        sum -= dense[static_cast<size_t>(i) * size + k] * dense[static_cast<size_t>(j) * size + k];
                    if (i == j) {
                        if (!(sum > 0.0) || !std::isfinite(sum)) {
                            ok = false;
                            break;
                        }
                        dense[static_cast<size_t>(i) * size + i] = std::sqrt(sum);
                    } else {
                        dense[static_cast<size_t>(i) * size + j] = sum / dense[static_cast<size_t>(j) * size + j];
                    }
                }
            }
            if (!ok) {
                bad[s] = 1;
                for (int i = 0; i < size; ++i)
                    for (int j = 0; j <= i; ++j) dense[static_cast<size_t>(i) * size + j] = i == j ? 1.0 : 0.0;
            }
            double* t = factor.data() + fbeg[s];
            for (int i = 0; i < size; ++i) {
                for (int j = 0; j <= i; ++j) t[i * (i + 1) / 2 + j] = dense[static_cast<size_t>(i) * size + j];
                t[i * (i + 1) / 2 + i] = 1.0 / t[i * (i + 1) / 2 + i];
            }
        }
    };
    const int workers = std::max(1, std::min<int>(16, static_cast<int>(std::thread::hardware_concurrency())));
    if (workers == 1 || n_sub < 64) {
        factor_range(0, 1);
    } else {
        std::vector<std::thread> pool;
        for (int w = 0; w < workers; ++w) pool.emplace_back(factor_range, w, workers);
        for (std::thread& t : pool) t.join();
    }
    SubdomainDevice& d = c->sub;
    d.n_sub = n_sub;
    d.max_size = max_size;
    d.degenerate = 0;
    for (char b : bad) d.degenerate += b;
    d.dof_begin.upload(dof_begin.data(), dof_begin.size(), c->stream);
    d.dofs.upload(dofs.data(), dofs.size(), c->stream);
    d.factor_begin.upload(fbeg.data(), fbeg.size(), c->stream);
    d.factor.upload(factor.data(), factor.size(), c->stream);
    sync(c);
}

lfg::SubdomainView subdomain_view(const lfg_ctx* c) {
    return lfg::SubdomainView{c->sub.n_sub, c->sub.max_size, c->sub.dof_begin.p, c->sub.dofs.p,
                              c->sub.factor_begin.p, c->sub.factor.p};
}

lfg::Ic0View ic0_view(const lfg_ctx* c) {
    const Ic0Device& d = c->ic0;
    const lfg::Ic0Sweep fwd{d.fwd.ord.p,       d.fwd.ptr.p,  d.fwd.key.p,   d.fwd.val.p,   d.fwd.scale.p,
                            d.fwd.level_ptr.p, d.fwd.levels, d.fwd.reach, d.fwd.widest};
    const lfg::Ic0Sweep bwd{d.bwd.ord.p,       d.bwd.ptr.p,  d.bwd.key.p,   d.bwd.val.p,   d.bwd.scale.p,
                            d.bwd.level_ptr.p, d.bwd.levels, d.bwd.reach, d.bwd.widest};
    auto node_view = [](const Ic0NodeSweepDevice& v) {
        return lfg::Ic0NodeSweep{v.ready ? v.ord.p : nullptr, v.ptr.p, v.key.p, v.v_first.p, v.v_second.p,
                                 v.cross.p, v.scale_first.p, v.scale_second.p, v.level_ptr.p, v.levels,
                                 v.reach, v.widest};
    };
    const lfg::Ic0NodeSweep nfwd = node_view(d.nfwd), nbwd = node_view(d.nbwd);
    return lfg::Ic0View{c->n,        d.lrow.p,  d.lcol.p,       d.lval.p,  d.inv_diag.p, d.urow.p,
                        d.ucol.p,    d.uval.p,  d.fptr.p,       d.frows.p, d.fwd_levels, d.bptr.p,
                        d.brows.p,   d.bwd_levels, fwd, bwd, nfwd, nbwd, c->flags.p + 2};
}

void ensure_vectors(lfg_ctx* c) {
    const size_t n = static_cast<size_t>(c->n);
    c->x.alloc(n);
    c->r.alloc(n);
    c->p.alloc(n);
    c->q.alloc(n);
    c->z.alloc(n);
    c->partials.alloc(65536);
    c->scratch.alloc(4096 + 8);
    if (!c->state.p) {
        c->state.alloc(1);
        c->state_alt.alloc(1);
        c->fold_partials.alloc(2 * kFoldPartials);
        LFG_HIP(hipMemsetAsync(c->state.p, 0, sizeof(lfg::PcgState), c->stream));
        c->tickets.alloc(6);
        LFG_HIP(hipMemsetAsync(c->tickets.p, 0, 6 * sizeof(lfg::TicketSet), c->stream));
        c->flags.alloc(4);
        LFG_HIP(hipMemsetAsync(c->scratch.p, 0, c->scratch.n * sizeof(double), c->stream));
    }
    if (!c->h_state) LFG_HIP(hipHostMalloc(&c->h_state, 2 * sizeof(lfg::PcgState), hipHostMallocDefault));
}

// Preconditioner factors on the constrained matrix. Returns set-up milliseconds.
void setup_preconditioner(lfg_ctx* c, int kind, long long* degenerate) {
    *degenerate = 0;
    if (c->precond_ready == kind) return;
    LFG_HIP(hipMemsetAsync(c->flags.p, 0, 4 * sizeof(int), c->stream));
    const lfg::CsrView a = csr_view(c);
    if (kind == LFG_PRECOND_JACOBI) {
        c->diag.alloc(static_cast<size_t>(c->n));
        c->minv.alloc(static_cast<size_t>(c->n));
        lfg::launch_extract_diag(c->stream, a, c->Kc.p, c->diag.p, c->flags.p);
        lfg::launch_jacobi_inv(c->stream, c->n, c->diag.p, c->minv.p, c->flags.p);
        int flags = 0;
        LFG_HIP(hipMemcpyAsync(&flags, c->flags.p, sizeof(int), hipMemcpyDeviceToHost, c->stream));
        sync(c);
        if (flags & 1) throw LfgError(LFG_ERR_ARG, "requested entry is absent from preallocated CSR pattern");
        if (flags & 2) throw LfgError(LFG_ERR_NONFINITE, "Jacobi preconditioner encountered a non-finite diagonal");
    } else if (kind == LFG_PRECOND_BLOCK2) {
        c->factor.alloc(3 * static_cast<size_t>(c->n / 2));
        lfg::launch_block2_factor(c->stream, a, c->Kc.p, c->factor.p, c->flags.p + 1);
        int deg = 0;
        LFG_HIP(hipMemcpyAsync(&deg, c->flags.p + 1, sizeof(int), hipMemcpyDeviceToHost, c->stream));
        sync(c);
        *degenerate = deg;
    } else if (kind == LFG_PRECOND_SUBDOMAIN) {
        build_subdomains(c);
        *degenerate = c->sub.degenerate;
    } else {
        build_ic0(c);
    }
    c->precond_ready = kind;
}

lfg::PcgBuffers pcg_buffers(lfg_ctx* c, int kind, const double* rhs, double* xs) {
    lfg::PcgBuffers b{};
    b.n = c->n;
    b.x = xs;
    b.r = c->r.p;
    b.p = c->p.p;
    b.q = c->q.p;
    b.z = c->z.p;
    b.b = rhs;
    b.m = kind == LFG_PRECOND_JACOBI ? c->minv.p : (kind == LFG_PRECOND_BLOCK2 ? c->factor.p : nullptr);
    b.precond = kind;
    b.st = c->state.p;
    b.tickets = c->tickets.p;
    b.partials = c->partials.p;
    return b;
}

// Whether an iteration forms its search direction inside the slice SpMV (two
// launches per iteration, p alternating between two buffers) instead of in a
// direction kernel after the update.
bool fused_direction(int variant, int precond) { return variant == kSliceFused && !sweep_precond(precond); }

// One PCG iteration. With the fused direction, iteration `parity` reads p from
// buffer parity (b.p or c->p2) and writes the new p into the other one.
void enqueue_iteration(lfg_ctx* c, const lfg::PcgBuffers& b, int mixed, int replace, int variant, int parity) {
    const lfg::CsrView a = csr_view(c);
    if (fused_iteration(variant)) {
        double* const rb[2] = {c->r.p, c->r_alt.p};
        double* const pb[2] = {c->p.p, c->p2.p};
        double* const qb[2] = {c->q.p, c->q_alt.p};
        if (variant == kSliceFold) {  // state and partials alternate with the launch parity
            lfg::PcgState* const sb[2] = {b.st, c->state_alt.p};
            double* const fb[2] = {c->fold_partials.p, c->fold_partials.p + kFoldPartials};
            lfg::launch_pcg_fold_slices(c->stream, b, slice_view(c), c->slices64.p, mixed ? c->slices32.p : nullptr,
                                        rb[parity], rb[parity ^ 1], pb[parity], pb[parity ^ 1], qb[parity],
                                        qb[parity ^ 1], sb[parity], sb[parity ^ 1], fb[parity ^ 1], fb[parity]);
            return;
        }
        if (variant == kSliceSym) {
            lfg::launch_pcg_sym_split_slices(c->stream, b, sym_view(c), c->sym64.p, mixed ? c->sym32.p : nullptr,
                                             rb[parity], rb[parity ^ 1],
                                             pb[parity], pb[parity ^ 1], qb[parity], qb[parity ^ 1]);
            return;
        }
        if (variant == kSliceIter || variant == kSliceSplit) {
            (variant == kSliceIter ? lfg::launch_pcg_fused_slices : lfg::launch_pcg_split_slices)(
                c->stream, b, slice_view(c), c->slices64.p, mixed ? c->slices32.p : nullptr, rb[parity],
                rb[parity ^ 1], pb[parity], pb[parity ^ 1], qb[parity], qb[parity ^ 1]);
            return;
        }
        (variant == kStencilGrid ? lfg::launch_pcg_fused_grid : lfg::launch_pcg_fused_stencil)(
            c->stream, b, c->stencil_sym, c->stencil64.p, mixed ? c->stencil32.p : nullptr,
            mixed ? c->stencil_diag.p : nullptr, rb[parity], rb[parity ^ 1], pb[parity], pb[parity ^ 1], qb[parity],
            qb[parity ^ 1]);
        return;
    }
    if (fused_direction(variant, b.precond)) {
        double* const pb[2] = {b.p, c->p2.p};
        lfg::PcgBuffers bi = b;
        bi.p = pb[parity ^ 1];
        lfg::launch_pcg_spmv_slices_dir(c->stream, b, slice_view(c), c->slices64.p, mixed ? c->slices32.p : nullptr,
                                        pb[parity], bi.p);
        lfg::launch_pcg_update(c->stream, bi, mixed);
        if (replace) lfg::launch_pcg_replace(c->stream, bi, a, c->Kc.p);
        return;
    }
    if (stencil_layout(variant))
        launch_stencil_spmv(c, b, variant, true, mixed != 0);
    else if (slice_layout(variant))
        lfg::launch_pcg_spmv_slices(c->stream, b, slice_view(c), c->slices64.p, mixed ? c->slices32.p : nullptr);
    else
        lfg::launch_pcg_spmv(c->stream, b, a, c->Kc.p, mixed ? c->K32.p : nullptr, variant);
    lfg::launch_pcg_update(c->stream, b, mixed);
    if (replace) lfg::launch_pcg_replace(c->stream, b, a, c->Kc.p);
    if (b.precond == LFG_PRECOND_IC0) lfg::launch_pcg_ic0_step(c->stream, b, ic0_view(c));
    if (b.precond == LFG_PRECOND_SUBDOMAIN) lfg::launch_pcg_subdomain_step(c->stream, b, subdomain_view(c));
    lfg::launch_pcg_direction(c->stream, b);
}

// The IC(0) sweeps bound every dependency wait; a factor whose rows never become
// ready (no valid factor does) ends the sweep with this error instead of a hang.
constexpr const char* kIc0Fault = "IC(0) triangular sweep exceeded its dependency wait bound";

const char* status_text(int s) {
    switch (s) {
        case LFG_OK: return "ok";
        case LFG_ERR_ARG: return "invalid argument";
        case LFG_ERR_JACOBIAN: return "Q4 stiffness evaluation received a non-positive Jacobian";
        case LFG_ERR_NOT_SPD: return "PCG encountered non-positive curvature; system is not SPD";
        case LFG_ERR_NONFINITE: return "PCG preconditioned residual became non-positive or non-finite";
        case LFG_ERR_NOT_CONVERGED: return "maximum iterations reached without convergence";
        case LFG_ERR_HIP: return "HIP runtime error";
        case LFG_ERR_STATE: return "call order violated";
        case LFG_ERR_UNSUPPORTED: return "unsupported option";
        case LFG_ERR_MECHANISM: return "the structure is a mechanism";
        default: return "unknown status";
    }
}

}  // namespace

extern "C" {

const char* lfg_status_string(int32_t status) { return status_text(status); }

int lfg_create(int32_t device, lfg_ctx** out) {
    if (!out) return LFG_ERR_ARG;
    *out = nullptr;
    auto ctx = std::make_unique<lfg_ctx>();
    ctx->device = device;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || device < 0 || device >= count) return LFG_ERR_HIP;
    if (hipSetDevice(device) != hipSuccess) return LFG_ERR_HIP;
    if (hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess) return LFG_ERR_HIP;
    // this device's address of the allocation-free multiply's scratch array (lfg_spmv_noalloc)
    if (lfg::noalloc_scratch_init(device) != hipSuccess) return LFG_ERR_HIP;
    if (hipEventCreate(&ctx->ev_start) != hipSuccess || hipEventCreate(&ctx->ev_stop) != hipSuccess ||
        hipEventCreate(&ctx->ev_ring[0]) != hipSuccess || hipEventCreate(&ctx->ev_ring[1]) != hipSuccess)
        return LFG_ERR_HIP;
    *out = ctx.release();
    return LFG_OK;
}

void lfg_destroy(lfg_ctx* ctx) {
    if (!ctx) return;
    (void)hipSetDevice(ctx->device);
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    if (ctx->h_state) (void)hipHostFree(ctx->h_state);
    if (ctx->h_batch_global) (void)hipHostFree(ctx->h_batch_global);
    for (hipEvent_t e : {ctx->ev_start, ctx->ev_stop, ctx->ev_ring[0], ctx->ev_ring[1]})
        if (e) (void)hipEventDestroy(e);
    hipStream_t s = ctx->stream;
    delete ctx;  // device buffers free here
    if (s) (void)hipStreamDestroy(s);
}

const char* lfg_last_error(const lfg_ctx* ctx) { return ctx ? ctx->last_error.c_str() : "null context"; }

int lfg_set_mesh(lfg_ctx* ctx, int64_t n_nodes, const double* xy, int64_t n_elems, const int32_t* enodes,
                 const int32_t* elem_material, int64_t n_materials, const double* materials, int32_t plane_mode) {
    return guarded(ctx, [&] {
        require(n_nodes > 0 && n_elems >= 0 && n_materials > 0 && xy && materials &&
                    (n_elems == 0 || (enodes && elem_material)),
                LFG_ERR_ARG, "mesh arrays must be non-empty");
        require(n_nodes < (1LL << 30) && n_elems < (1LL << 24), LFG_ERR_UNSUPPORTED,
                "mesh exceeds the int32 index range of this backend");
        require(plane_mode == LFG_PLANE_STRESS || plane_mode == LFG_PLANE_STRAIN, LFG_ERR_ARG, "unknown plane mode");
        for (int64_t k = 0; k < 4 * n_elems; ++k)
            require(enodes[k] >= 0 && enodes[k] < n_nodes, LFG_ERR_ARG, "validated element references an unknown node");
        for (int64_t e = 0; e < n_elems; ++e)
            require(elem_material[e] >= 0 && elem_material[e] < n_materials, LFG_ERR_ARG,
                    "validated element references an unknown material");
        LFG_HIP(hipSetDevice(ctx->device));
        ctx->n_nodes = static_cast<int>(n_nodes);
        ctx->n_elems = static_cast<int>(n_elems);
        ctx->n_mats = static_cast<int>(n_materials);
        ctx->plane_mode = plane_mode;
        ctx->xy.upload(xy, 2 * static_cast<size_t>(n_nodes), ctx->stream);
        ctx->enodes.upload(enodes, 4 * static_cast<size_t>(n_elems), ctx->stream);
        ctx->elem_mat.upload(elem_material, static_cast<size_t>(n_elems), ctx->stream);
        ctx->materials.upload(materials, 3 * static_cast<size_t>(n_materials), ctx->stream);
        ctx->mat_mode.release();
        ctx->h_enodes.assign(enodes, enodes + 4 * n_elems);
        detect_element_grid(ctx);
        ctx->n_bars = 0;
        ctx->h_bar_nodes.clear();
        build_node_elements(ctx);
        sync(ctx);
        // the pattern, its element map and the colour groups describe the previous mesh
        ctx->have_pattern = false;
        ctx->gather_valid = false;
        ctx->h_color_ptr.clear();
        ctx->h_color_bar_ptr.clear();
        ctx->have_mesh = true;
        ctx->have_K = ctx->have_system = ctx->have_solution = false;
        ctx->precond_ready = -1;
        return LFG_OK;
    });
}

int lfg_set_materials(lfg_ctx* ctx, int64_t n_materials, const double* materials) {
    return guarded(ctx, [&] {
        require(ctx->have_mesh, LFG_ERR_STATE, "lfg_set_mesh must precede lfg_set_materials");
        require(n_materials == ctx->n_mats && materials, LFG_ERR_ARG,
                "the material table must keep the size given to lfg_set_mesh");
        LFG_HIP(hipSetDevice(ctx->device));
        ctx->materials.upload(materials, 3 * static_cast<size_t>(n_materials), ctx->stream);
        sync(ctx);
        ctx->have_K = ctx->have_system = ctx->have_solution = ctx->have_adjoint = false;
        ctx->precond_ready = -1;
        return LFG_OK;
    });
}

int lfg_set_material_modes(lfg_ctx* ctx, const int32_t* modes) {
    return guarded(ctx, [&] {
        require(ctx->have_mesh, LFG_ERR_STATE, "lfg_set_mesh must precede lfg_set_material_modes");
        LFG_HIP(hipSetDevice(ctx->device));
        if (!modes) {
            ctx->mat_mode.release();
        } else {
            for (int k = 0; k < ctx->n_mats; ++k)
                require(modes[k] == LFG_PLANE_STRESS || modes[k] == LFG_PLANE_STRAIN, LFG_ERR_ARG, "unknown plane mode");
            ctx->mat_mode.upload(modes, static_cast<size_t>(ctx->n_mats), ctx->stream);
            sync(ctx);
        }
        ctx->have_K = ctx->have_system = ctx->have_solution = ctx->have_adjoint = false;
        ctx->precond_ready = -1;
        return LFG_OK;
    });
}

int lfg_set_pattern(lfg_ctx* ctx, int64_t n_rows, int64_t nnz, const int32_t* row_ptr, const int32_t* col,
                    const int32_t* emap) {
    return guarded(ctx, [&] {
        require(ctx->have_mesh, LFG_ERR_STATE, "lfg_set_mesh must precede lfg_set_pattern");
        require(n_rows == 2LL * ctx->n_nodes, LFG_ERR_ARG, "CSR dimension must be two dofs per node");
        require(nnz > 0 && nnz < (1LL << 31) && row_ptr && col && (emap || ctx->n_elems == 0), LFG_ERR_ARG,
                "invalid CSR pattern");
        require(row_ptr[0] == 0 && row_ptr[n_rows] == nnz, LFG_ERR_ARG, "row offsets do not span the nonzeros");
        for (int64_t k = 0; k < 64LL * ctx->n_elems; ++k)
            require(emap[k] >= 0 && emap[k] < nnz, LFG_ERR_ARG, "element map entry outside the CSR values");
        LFG_HIP(hipSetDevice(ctx->device));
        ctx->n = static_cast<int>(n_rows);
        ctx->nnz = nnz;
        ctx->h_row_ptr.assign(row_ptr, row_ptr + n_rows + 1);
        ctx->max_row = 0;
        for (int64_t r = 0; r < n_rows; ++r) {
            require(row_ptr[r + 1] >= row_ptr[r], LFG_ERR_ARG, "row offsets must be non-decreasing");
            ctx->max_row = std::max(ctx->max_row, row_ptr[r + 1] - row_ptr[r]);
        }
        ctx->h_col.assign(col, col + nnz);
        ctx->row_ptr.upload(row_ptr, static_cast<size_t>(n_rows) + 1, ctx->stream);
        ctx->col.upload(col, static_cast<size_t>(nnz), ctx->stream);
        ctx->emap.upload(emap, 64 * static_cast<size_t>(ctx->n_elems), ctx->stream);
        ctx->gather_valid = false;
        build_slices(ctx);
        build_stencil(ctx);
        ctx->K.alloc(static_cast<size_t>(nnz));
        ctx->have_pattern = true;
        ctx->have_K = ctx->have_system = ctx->have_solution = false;
        ctx->precond_ready = -1;
        ctx->sub_node.clear();  // subdomains belong to the node numbering of a pattern
        ctx->n_sub = 0;
        if (ctx->n_bars > 0) {  // bar maps index the previous pattern
            ctx->n_bars = 0;
            ctx->h_bar_nodes.clear();
            build_node_elements(ctx);
            sync(ctx);
        }
        ctx->h_color_ptr.clear();
        ctx->h_color_bar_ptr.clear();
        return LFG_OK;
    });
}

int lfg_set_bars(lfg_ctx* ctx, int64_t n_bars, const int32_t* bar_nodes, const int32_t* bar_material,
                 const double* area, const int32_t* bar_map) {
    return guarded(ctx, [&] {
        require(ctx->have_pattern, LFG_ERR_STATE, "lfg_set_pattern must precede lfg_set_bars");
        require(n_bars >= 0 && n_bars < (1LL << 24) && (n_bars == 0 || (bar_nodes && bar_material && area && bar_map)),
                LFG_ERR_ARG, "bar arrays must be given");
        for (int64_t k = 0; k < 2 * n_bars; ++k)
            require(bar_nodes[k] >= 0 && bar_nodes[k] < ctx->n_nodes, LFG_ERR_ARG,
                    "validated truss references an unknown node");
        for (int64_t t = 0; t < n_bars; ++t)
            require(bar_material[t] >= 0 && bar_material[t] < ctx->n_mats, LFG_ERR_ARG,
                    "validated truss references an unknown material");
        for (int64_t k = 0; k < 16 * n_bars; ++k)
            require(bar_map[k] >= 0 && bar_map[k] < ctx->nnz, LFG_ERR_ARG, "bar map entry is outside the pattern");
        LFG_HIP(hipSetDevice(ctx->device));
        ctx->n_bars = static_cast<int>(n_bars);
        ctx->h_bar_nodes.assign(bar_nodes, bar_nodes + 2 * n_bars);
        ctx->bar_nodes.upload(bar_nodes, 2 * static_cast<size_t>(n_bars), ctx->stream);
        ctx->bar_mat.upload(bar_material, static_cast<size_t>(n_bars), ctx->stream);
        ctx->bar_area.upload(area, static_cast<size_t>(n_bars), ctx->stream);
        ctx->bar_map.upload(bar_map, 16 * static_cast<size_t>(n_bars), ctx->stream);
        // the serial gather and the smoothing lists take the bars after the Q4 elements
        ctx->gather_valid = false;
        build_node_elements(ctx);
        ctx->h_color_ptr.clear();  // colour lists name the elements with the bars included
        ctx->h_color_bar_ptr.clear();
        sync(ctx);
        ctx->have_K = ctx->have_system = ctx->have_solution = false;
        ctx->precond_ready = -1;
        return LFG_OK;
    });
}

int lfg_pin_zero_pivots(lfg_ctx* ctx, lfg_pivot_report* report) {
    return guarded(ctx, [&] {
        require(ctx->have_system, LFG_ERR_STATE, "lfg_dirichlet must precede lfg_pin_zero_pivots");
        LFG_HIP(hipSetDevice(ctx->device));
        ctx->pivot_scale.alloc(1);
        ctx->pivot_counters.alloc(2);
        const int init[2] = {0, INT32_MAX};
        LFG_HIP(hipMemsetAsync(ctx->pivot_scale.p, 0, sizeof(unsigned long long), ctx->stream));
        LFG_HIP(hipMemcpyAsync(ctx->pivot_counters.p, init, sizeof(init), hipMemcpyHostToDevice, ctx->stream));
        lfg::launch_pin_zero_pivots(ctx->stream, csr_view(ctx), ctx->Kc.p, ctx->fc.p, ctx->pivot_scale.p,
                                    ctx->pivot_counters.p);
        int out[2] = {0, 0};
        LFG_HIP(hipMemcpyAsync(out, ctx->pivot_counters.p, sizeof(out), hipMemcpyDeviceToHost, ctx->stream));
        sync(ctx);
        lfg_pivot_report rep{};
        rep.pinned = out[0];
        rep.dof = out[1] == INT32_MAX ? -1 : out[1];
        if (rep.dof >= 0) {
            // its load and diagonal, untouched by the sweep
            const int row = static_cast<int>(rep.dof);
            LFG_HIP(hipMemcpy(&rep.load, ctx->fc.p + row, sizeof(double), hipMemcpyDeviceToHost));
            for (int k = ctx->h_row_ptr[row]; k < ctx->h_row_ptr[row + 1]; ++k)
                if (ctx->h_col[k] == row) LFG_HIP(hipMemcpy(&rep.diag, ctx->Kc.p + k, sizeof(double), hipMemcpyDeviceToHost));
        }
        if (report) *report = rep;
        if (rep.pinned > 0) {  // the packed copies and factors of Kc are stale
            ctx->k32_valid = false;
            ctx->slices64_valid = ctx->slices32_valid = ctx->sym64_valid = ctx->sym32_valid = false;
            ctx->stencil64_valid = ctx->stencil32_valid = ctx->stencil_exact_valid = false;
            ctx->precond_ready = -1;
        }
        if (rep.dof >= 0)
            throw LfgError(LFG_ERR_MECHANISM, "the structure is a mechanism: a bar end has no transverse stiffness but "
                                              "carries a transverse load");
        return LFG_OK;
    });
}

int lfg_set_colors(lfg_ctx* ctx, int32_t n_colors, const int32_t* color_ptr, const int32_t* color_elems) {
    return guarded(ctx, [&] {
        require(ctx->have_mesh, LFG_ERR_STATE, "lfg_set_mesh must precede lfg_set_colors");
        const int total = ctx->n_elems + ctx->n_bars;
        require(n_colors > 0 && color_ptr && color_elems && color_ptr[0] == 0 && color_ptr[n_colors] == total,
                LFG_ERR_ARG, "colour groups must partition the elements");
        for (int64_t k = 0; k < total; ++k)
            require(color_elems[k] >= 0 && color_elems[k] < total, LFG_ERR_ARG, "colour lists an unknown element");
        // each colour as its Q4 elements and its bars (bar t is element n_elems + t)
        std::vector<int> q4, bars;
        ctx->h_color_ptr.assign(1, 0);
        ctx->h_color_bar_ptr.assign(1, 0);
        for (int c = 0; c < n_colors; ++c) {
            for (int k = color_ptr[c]; k < color_ptr[c + 1]; ++k) {
                if (color_elems[k] < ctx->n_elems) q4.push_back(color_elems[k]);
                else bars.push_back(color_elems[k] - ctx->n_elems);
            }
            ctx->h_color_ptr.push_back(static_cast<int>(q4.size()));
            ctx->h_color_bar_ptr.push_back(static_cast<int>(bars.size()));
        }
        ctx->color_elems.upload(q4.data(), q4.size(), ctx->stream);
        ctx->color_bars.upload(bars.data(), bars.size(), ctx->stream);
        if (ctx->n_bars == 0) {
            // node -> (element << 2 | corner), colour by colour: the order in which the
            // coloured passes add into each value of the node's rows
            const int n_nodes = ctx->n_nodes;
            std::vector<int> ptr(static_cast<size_t>(n_nodes) + 1, 0);
            for (int k = 0; k < 4 * ctx->n_elems; ++k) ptr[static_cast<size_t>(ctx->h_enodes[k]) + 1]++;
            for (int i = 0; i < n_nodes; ++i) ptr[i + 1] += ptr[i];
            std::vector<int> list(static_cast<size_t>(ptr[n_nodes]));
            std::vector<int> cur(ptr.begin(), ptr.end() - 1);
            for (const int e : q4)
                for (int l = 0; l < 4; ++l) list[static_cast<size_t>(cur[ctx->h_enodes[4 * e + l]]++)] = (e << 2) | l;
            ctx->color_node_ptr.upload(ptr.data(), ptr.size(), ctx->stream);
            ctx->color_node.upload(list.data(), list.size(), ctx->stream);
        }
        sync(ctx);
        return LFG_OK;
    });
}

int lfg_element_stiffness(lfg_ctx* ctx, double* ke_out) {
    return guarded(ctx, [&] {
        require(ctx->have_mesh && ke_out, LFG_ERR_STATE, "lfg_set_mesh must precede lfg_element_stiffness");
        LFG_HIP(hipSetDevice(ctx->device));
        ensure_vectors(ctx);
        ctx->ke_buf.alloc(64 * static_cast<size_t>(ctx->n_elems));
        LFG_HIP(hipMemsetAsync(ctx->flags.p, 0, sizeof(int), ctx->stream));
        lfg::launch_ke(ctx->stream, mesh_view(ctx), lfg::KE_STORE, nullptr, ctx->n_elems, nullptr, nullptr,
                       ctx->ke_buf.p, ctx->flags.p);
        int failures = 0;
        LFG_HIP(hipMemcpyAsync(&failures, ctx->flags.p, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
        ctx->ke_buf.download(ke_out, ctx->ke_buf.n, ctx->stream);
        sync(ctx);
        if (failures) throw LfgError(LFG_ERR_JACOBIAN, "Q4 stiffness evaluation received a non-positive Jacobian");
        return LFG_OK;
    });
}

int lfg_assemble(lfg_ctx* ctx, int32_t strategy, lfg_assembly_report* report) {
    return guarded(ctx, [&] {
        require(ctx->have_pattern, LFG_ERR_STATE, "lfg_set_pattern must precede lfg_assemble");
        LFG_HIP(hipSetDevice(ctx->device));
        ensure_vectors(ctx);
        const lfg::MeshView m = mesh_view(ctx);
        const lfg::BarView bv = bar_view(ctx);
        LFG_HIP(hipMemsetAsync(ctx->flags.p, 0, 4 * sizeof(int), ctx->stream));
        lfg_assembly_report rep{};
        const bool rows = ctx->n_bars == 0 && strategy != LFG_ASSEMBLY_ATOMIC && ctx->max_row <= lfg::kAssembleRowSlots;
        if (rows) {
            // Q4 meshes, serial and coloured sums: one pass, each row owned by one thread
            // (element order for serial / thread_local, colour order for coloring)
            const bool colored = strategy == LFG_ASSEMBLY_COLORING;
            require(strategy == LFG_ASSEMBLY_SERIAL || strategy == LFG_ASSEMBLY_THREAD_LOCAL || colored, LFG_ERR_ARG,
                    "validated model has an unknown assembly strategy");
            require(!colored || !ctx->h_color_ptr.empty(), LFG_ERR_STATE,
                    "lfg_set_colors must precede coloring assembly");
            const int* node_ptr = colored ? ctx->color_node_ptr.p : ctx->node_elem_ptr.p;
            const int* node_list = colored ? ctx->color_node.p : ctx->node_elem.p;
            LFG_HIP(hipEventRecord(ctx->ev_start, ctx->stream));
            if (ctx->have_slices && ctx->max_block_slots <= lfg::kAssembleBlockSlots)
                lfg::launch_assemble_block_rows(ctx->stream, m, ctx->n_nodes, ctx->row_ptr.p, ctx->col.p, node_ptr,
                                                node_list, ctx->K.p, ctx->flags.p);
            else
                lfg::launch_assemble_rows(ctx->stream, m, ctx->n_nodes, ctx->row_ptr.p, node_ptr, node_list,
                                          ctx->emap.p, ctx->K.p, ctx->flags.p);
            LFG_HIP(hipEventRecord(ctx->ev_stop, ctx->stream));
            sync(ctx);
            rep.assembly_ms = elapsed(ctx->ev_start, ctx->ev_stop);  // element matrices formed in the pass
            if (colored) rep.coloring_count = static_cast<int>(ctx->h_color_ptr.size()) - 1;
        } else if (strategy == LFG_ASSEMBLY_SERIAL || strategy == LFG_ASSEMBLY_THREAD_LOCAL) {
            ensure_gather_map(ctx);
            const size_t q4 = 64 * static_cast<size_t>(ctx->n_elems);
            ctx->ke_buf.alloc(q4 + 16 * static_cast<size_t>(ctx->n_bars));
            LFG_HIP(hipEventRecord(ctx->ev_start, ctx->stream));
            lfg::launch_ke(ctx->stream, m, lfg::KE_STORE, nullptr, ctx->n_elems, nullptr, nullptr, ctx->ke_buf.p,
                           ctx->flags.p);
            lfg::launch_bar_ke(ctx->stream, m, bv, lfg::KE_STORE, nullptr, ctx->n_bars, nullptr, ctx->ke_buf.p + q4,
                               ctx->flags.p + 3);
            LFG_HIP(hipEventRecord(ctx->ev_ring[0], ctx->stream));
            lfg::launch_gather(ctx->stream, ctx->nnz, ctx->cptr.p, ctx->contrib.p, ctx->ke_buf.p, ctx->K.p);
            LFG_HIP(hipEventRecord(ctx->ev_stop, ctx->stream));
            sync(ctx);
            rep.element_stiffness_ms = elapsed(ctx->ev_start, ctx->ev_ring[0]);
            rep.assembly_ms = elapsed(ctx->ev_ring[0], ctx->ev_stop);
        } else if (strategy == LFG_ASSEMBLY_ATOMIC) {
            LFG_HIP(hipEventRecord(ctx->ev_start, ctx->stream));
            lfg::launch_fill(ctx->stream, ctx->K.p, ctx->nnz, 0.0);
            if (ctx->grid_w > 0 && ctx->n_bars == 0 && ctx->max_row <= lfg::kAssembleGridRow)
                lfg::launch_assemble_atomic_tiles(ctx->stream, m, ctx->grid_w, ctx->grid_h, ctx->row_ptr.p,
                                                  ctx->emap.p, ctx->K.p, ctx->flags.p);
            else
                lfg::launch_ke(ctx->stream, m, lfg::KE_ATOMIC, nullptr, ctx->n_elems, ctx->emap.p, ctx->K.p, nullptr,
                               ctx->flags.p);
            lfg::launch_bar_ke(ctx->stream, m, bv, lfg::KE_ATOMIC, nullptr, ctx->n_bars, ctx->K.p, nullptr,
                               ctx->flags.p + 3);
            LFG_HIP(hipEventRecord(ctx->ev_stop, ctx->stream));
            sync(ctx);
            rep.assembly_ms = elapsed(ctx->ev_start, ctx->ev_stop);  // Ke fused with the scatter
        } else if (strategy == LFG_ASSEMBLY_COLORING) {
            require(!ctx->h_color_ptr.empty(), LFG_ERR_STATE, "lfg_set_colors must precede coloring assembly");
            LFG_HIP(hipEventRecord(ctx->ev_start, ctx->stream));
            lfg::launch_fill(ctx->stream, ctx->K.p, ctx->nnz, 0.0);
            const int colors = static_cast<int>(ctx->h_color_ptr.size()) - 1;
            for (int cidx = 0; cidx < colors; ++cidx) {
                const int begin = ctx->h_color_ptr[cidx];
                lfg::launch_ke(ctx->stream, m, lfg::KE_COLORED, ctx->color_elems.p + begin,
                               ctx->h_color_ptr[cidx + 1] - begin, ctx->emap.p, ctx->K.p, nullptr, ctx->flags.p);
                const int bbegin = ctx->h_color_bar_ptr[cidx];
                lfg::launch_bar_ke(ctx->stream, m, bv, lfg::KE_COLORED, ctx->color_bars.p + bbegin,
                                   ctx->h_color_bar_ptr[cidx + 1] - bbegin, ctx->K.p, nullptr, ctx->flags.p + 3);
            }
            LFG_HIP(hipEventRecord(ctx->ev_stop, ctx->stream));
            sync(ctx);
            rep.assembly_ms = elapsed(ctx->ev_start, ctx->ev_stop);
            rep.coloring_count = colors;
        } else {
            throw LfgError(LFG_ERR_ARG, "validated model has an unknown assembly strategy");
        }
        int fl[4] = {0, 0, 0, 0};
        LFG_HIP(hipMemcpy(fl, ctx->flags.p, 4 * sizeof(int), hipMemcpyDeviceToHost));
        const int failures = fl[0];
        rep.jacobian_failures = failures;
        if (report) *report = rep;
        if (failures) throw LfgError(LFG_ERR_JACOBIAN, "Q4 stiffness evaluation received a non-positive Jacobian");
        if (fl[3]) throw LfgError(LFG_ERR_JACOBIAN, "truss stiffness evaluation received a zero-length member");
        ctx->have_K = true;
        ctx->have_system = ctx->have_solution = false;
        ctx->precond_ready = -1;
        return LFG_OK;
    });
}

int lfg_get_stiffness(lfg_ctx* ctx, double* values_out) {
    return guarded(ctx, [&] {
        require(ctx->have_K && values_out, LFG_ERR_STATE, "no assembled stiffness");
        ctx->K.download(values_out, static_cast<size_t>(ctx->nnz), ctx->stream);
        sync(ctx);
        return LFG_OK;
    });
}

int lfg_set_stiffness(lfg_ctx* ctx, const double* values) {
    return guarded(ctx, [&] {
        require(ctx->have_pattern && values, LFG_ERR_STATE, "lfg_set_pattern must precede lfg_set_stiffness");
        ctx->K.upload(values, static_cast<size_t>(ctx->nnz), ctx->stream);
        sync(ctx);
        ctx->have_K = true;
        ctx->have_system = ctx->have_solution = false;
        ctx->precond_ready = -1;
        return LFG_OK;
    });
}

int lfg_set_load(lfg_ctx* ctx, const double* f) {
    return guarded(ctx, [&] {
        require(ctx->have_pattern && f, LFG_ERR_STATE, "lfg_set_pattern must precede lfg_set_load");
        ctx->f.upload(f, static_cast<size_t>(ctx->n), ctx->stream);
        sync(ctx);
        ctx->have_f = true;
        ctx->have_system = ctx->have_solution = false;
        return LFG_OK;
    });
}

int lfg_dirichlet(lfg_ctx* ctx, int64_t n_constraints, const int32_t* dofs, const double* values) {
    return guarded(ctx, [&] {
        require(ctx->have_K && ctx->have_f, LFG_ERR_STATE, "assembly and load must precede lfg_dirichlet");
        require(n_constraints >= 0 && (n_constraints == 0 || (dofs && values)), LFG_ERR_ARG, "invalid constraints");
        std::vector<int> flag(static_cast<size_t>(ctx->n), 0);
        std::vector<double> presc(static_cast<size_t>(ctx->n), 0.0);
        for (int64_t k = 0; k < n_constraints; ++k) {
            require(dofs[k] >= 0 && dofs[k] < ctx->n, LFG_ERR_ARG, "validated boundary references an unknown node");
            flag[dofs[k]] = 1;
            presc[dofs[k]] = values[k];
        }
        LFG_HIP(hipSetDevice(ctx->device));
        ctx->constrained.upload(flag.data(), flag.size(), ctx->stream);
        ctx->prescribed.upload(presc.data(), presc.size(), ctx->stream);
        ctx->Kc.alloc(static_cast<size_t>(ctx->nnz));
        ctx->fc.alloc(static_cast<size_t>(ctx->n));
        if (n_constraints == 0) {  // solver.cpp:358: nothing to eliminate
            LFG_HIP(hipMemcpyAsync(ctx->Kc.p, ctx->K.p, ctx->nnz * sizeof(double), hipMemcpyDeviceToDevice, ctx->stream));
            LFG_HIP(hipMemcpyAsync(ctx->fc.p, ctx->f.p, ctx->n * sizeof(double), hipMemcpyDeviceToDevice, ctx->stream));
        } else {
            lfg::launch_dirichlet(ctx->stream, csr_view(ctx), ctx->K.p, ctx->f.p, ctx->constrained.p,
                                  ctx->prescribed.p, ctx->Kc.p, ctx->fc.p);
        }
        sync(ctx);
        ctx->have_system = true;
        ctx->have_solution = false;
        ctx->k32_valid = false;
        ctx->slices64_valid = ctx->slices32_valid = ctx->sym64_valid = ctx->sym32_valid = false;
        ctx->stencil64_valid = ctx->stencil32_valid = ctx->stencil_exact_valid = false;
        ctx->precond_ready = -1;
        return LFG_OK;
    });
}

int lfg_get_system(lfg_ctx* ctx, double* values_out, double* rhs_out) {
    return guarded(ctx, [&] {
        require(ctx->have_system, LFG_ERR_STATE, "lfg_dirichlet must precede lfg_get_system");
        if (values_out) ctx->Kc.download(values_out, static_cast<size_t>(ctx->nnz), ctx->stream);
        if (rhs_out) ctx->fc.download(rhs_out, static_cast<size_t>(ctx->n), ctx->stream);
        sync(ctx);
        return LFG_OK;
    });
}

}  // extern "C"

namespace {

// Preconditioned CG on the constrained system for right-hand side `rhs` (device),
// solution into `xs` (device) and, when `x_out` is given, the host. Shared by the
// forward solve and the adjoint solve, which reuses the forward's preconditioner.
void run_pcg(lfg_ctx* ctx, const lfg_pcg_options* opt, const double* rhs, double* xs, double* x_out,
         lfg_pcg_report* report) {
    require(ctx->have_system, LFG_ERR_STATE, "lfg_dirichlet must precede lfg_pcg");
    require(opt != nullptr, LFG_ERR_ARG, "missing PCG options");
    const int kind = opt->preconditioner;
    require(kind == LFG_PRECOND_JACOBI || kind == LFG_PRECOND_BLOCK2 || kind == LFG_PRECOND_IC0 ||
                kind == LFG_PRECOND_SUBDOMAIN,
            LFG_ERR_ARG, "preconditioner must be jacobi, block, or ic0");
    require(opt->storage == LFG_STORAGE_CSR || opt->storage == LFG_STORAGE_BSR, LFG_ERR_ARG,
            "storage must be csr or bsr");
    const int mixed = opt->precision == LFG_PRECISION_MIXED;
    require(!(mixed && sweep_precond(kind)), LFG_ERR_UNSUPPORTED,
            "mixed precision is implemented for the jacobi and 2x2 block preconditioners");
    require(opt->tolerance > 0.0 && opt->tolerance < 1.0, LFG_ERR_ARG, "solver tolerance must satisfy 0 < tolerance < 1");
    require(opt->max_iterations > 0, LFG_ERR_ARG, "max_iterations must be greater than zero");
    LFG_HIP(hipSetDevice(ctx->device));
    ensure_vectors(ctx);
    lfg_pcg_report rep{};
    // set-up: preconditioner factors and FP32 values
    LFG_HIP(hipEventRecord(ctx->ev_start, ctx->stream));
    long long degenerate = 0;
    const int previous = ctx->precond_ready;
    setup_preconditioner(ctx, kind, &degenerate);
    if (previous == kind && kind == LFG_PRECOND_BLOCK2) degenerate = -1;  // factor reused; count unknown here
    if (kind == LFG_PRECOND_SUBDOMAIN) degenerate = ctx->sub.degenerate;
    const bool block_order = opt->storage == LFG_STORAGE_BSR;
    int variant = resolve_variant(ctx, opt->spmv_variant, kind, block_order);
    // a large pattern without a stencil: the two-launch iteration on the symmetric node
    // slices, which streams each mirrored pair of blocks once
    if (opt->spmv_variant == 0 && variant == kSliceSplit && sym_slices_ok(ctx)) variant = kSliceSym;
    const PairScope pair_scope(ctx, block_order);
    require(!(mixed && variant == kStencilExact), LFG_ERR_UNSUPPORTED,
            "mixed precision uses the symmetric stencil form (variant 8), not the exact one");
    if (variant == kSliceSym) {
        ensure_sym_slices(ctx, true, mixed != 0);
    } else if (stencil_layout(variant)) {
        ensure_stencil(ctx, variant, true, mixed != 0);
    } else if (slice_layout(variant)) {
        ensure_slices(ctx, true, mixed != 0);
    } else if (mixed && !ctx->k32_valid) {
        ctx->K32.alloc(static_cast<size_t>(ctx->nnz));
        lfg::launch_to_f32(ctx->stream, ctx->Kc.p, ctx->K32.p, ctx->nnz);
        ctx->k32_valid = true;
    }
    LFG_HIP(hipEventRecord(ctx->ev_stop, ctx->stream));
    sync(ctx);
    rep.setup_ms = elapsed(ctx->ev_start, ctx->ev_stop);
    rep.degenerate_blocks = degenerate;
    if (kind == LFG_PRECOND_IC0) {
        rep.breakdowns = ctx->ic0.breakdowns;
        rep.levels = ctx->ic0.fwd_levels + ctx->ic0.bwd_levels;
    }

    const bool stencil_fused = fused_iteration(variant);
    const bool fused = fused_direction(variant, kind) || stencil_fused;
    if (fused) ctx->p2.alloc(static_cast<size_t>(ctx->n));
    if (stencil_fused) {
        ctx->r_alt.alloc(static_cast<size_t>(ctx->n));
        ctx->q_alt.alloc(static_cast<size_t>(ctx->n));
    }
    // a captured graph starts every launch on buffer 0, so it holds an even count when fused
    int per_graph = opt->check_every > 0 ? opt->check_every : 32;
    if (fused && (per_graph & 1)) ++per_graph;
    // mixed: FP32 values once ||r|| <= 1e-2 ||r0|| by default; FP64 residual replacements
    // only on request (the fused stencil iteration has none)
    const double drop = opt->replace_drop > 0.0 && !stencil_fused ? opt->replace_drop : 0.0;
    const double switch_tau = !mixed ? 0.0 : (opt->mixed_switch != 0.0 ? opt->mixed_switch : 1.0e-2);
    const int replace = mixed && drop > 0.0;

    // One PCG solve of K x = b_vec from x = 0 into x_vec (solver.cpp:532-627) with the
    // stop rule (rule, tol); returns the final device state and the device milliseconds.
    auto pcg_pass = [&](const double* b_vec, double* x_vec, int rule, double tol, long long max_iter,
                        float* ms) -> lfg::PcgState {
        const lfg::PcgBuffers b = pcg_buffers(ctx, kind, b_vec, x_vec);
        hipGraph_t graph = nullptr;
        hipGraphExec_t exec = nullptr;
        const bool use_graph = opt->direct_launch == 0;
        if (use_graph) {  // capture `per_graph` iterations once
            LFG_HIP(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
            for (int i = 0; i < per_graph; ++i) enqueue_iteration(ctx, b, mixed, replace, variant, i & 1);
            LFG_HIP(hipStreamEndCapture(ctx->stream, &graph));
            LFG_HIP(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
        }
        struct GraphGuard {
            hipGraph_t g;
            hipGraphExec_t e;
            ~GraphGuard() {
                if (e) (void)hipGraphExecDestroy(e);
                if (g) (void)hipGraphDestroy(g);
            }
        } guard{graph, exec};

        LFG_HIP(hipEventRecord(ctx->ev_start, ctx->stream));
        lfg::launch_pcg_init(ctx->stream, b, rule, tol, max_iter, drop, switch_tau, stencil_fused ? 1 : 0);
        if (kind == LFG_PRECOND_IC0) {
            lfg::launch_pcg_ic0_step(ctx->stream, b, ic0_view(ctx));
            lfg::launch_pcg_direction(ctx->stream, b);
        } else if (kind == LFG_PRECOND_SUBDOMAIN) {
            lfg::launch_pcg_subdomain_step(ctx->stream, b, subdomain_view(ctx));
            lfg::launch_pcg_direction(ctx->stream, b);
        }
        // every iteration the cap allows fits in this many graph launches (a fused iteration
        // spends its first launch on p0 and A p0; variant 14 folds the last sums one launch
        // later)
        const long long launches_needed = max_iter + (stencil_fused ? 1 : 0) + (variant == kSliceFold ? 1 : 0);
        const long long launch_cap = std::max<long long>(1, (launches_needed + per_graph - 1) / per_graph);
        long long launched = 0;
        auto launch_one = [&](int slot) {
            if (use_graph) {
                LFG_HIP(hipGraphLaunch(exec, ctx->stream));
            } else {
                for (int i = 0; i < per_graph; ++i) enqueue_iteration(ctx, b, mixed, replace, variant, i & 1);
            }
            LFG_HIP(hipMemcpyAsync(ctx->h_state + slot, ctx->state.p, sizeof(lfg::PcgState), hipMemcpyDeviceToHost,
                                   ctx->stream));
            LFG_HIP(hipEventRecord(ctx->ev_ring[slot], ctx->stream));
            ++launched;
        };
        launch_one(0);
        long long k = 0;
        while (true) {
            const int slot = static_cast<int>(k & 1);
            if (launched < launch_cap) launch_one(1 - slot);
            LFG_HIP(hipEventSynchronize(ctx->ev_ring[slot]));
            if (ctx->h_state[slot].done || launched >= launch_cap) {
                *ms = elapsed(ctx->ev_start, ctx->ev_ring[slot]);
                break;
            }
            ++k;
        }
        if (stencil_fused) lfg::launch_pcg_fused_flush(ctx->stream, b, ctx->p2.p);  // pb[1] of enqueue_iteration
        sync(ctx);
        lfg::PcgState st{};
        LFG_HIP(hipMemcpy(&st, ctx->state.p, sizeof(st), hipMemcpyDeviceToHost));
        return st;
    };
    // ||b - K x|| with FP64 values; r_vec, when given, receives b - K x
    auto true_residual = [&](const double* b_vec, const double* x_vec, double* r_vec) {
        double r2 = 0.0;
        lfg::launch_residual_norm(ctx->stream, csr_view(ctx), ctx->Kc.p, x_vec, b_vec, ctx->scratch.p,
                                  ctx->tickets.p + 4, ctx->scratch.p + 4096, r_vec);
        LFG_HIP(hipMemcpyAsync(&r2, ctx->scratch.p + 4096, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
        sync(ctx);
        return std::sqrt(std::max(0.0, r2));
    };

    float pass_ms = 0.0f;
    lfg::PcgState st = pcg_pass(rhs, xs, opt->stop_rule, opt->tolerance, opt->max_iterations, &pass_ms);
    double solve_ms = pass_ms;
    long long iterations = st.iter;
    const double target = st.target;
    // FP64 iterative refinement: r = b - K x with FP64 values, a correction solve K d = r
    // (same precision policy) to ||r_d|| <= max(eta ||r||, target), x += d; until the FP64
    // residual meets the target, the passes run out, or a pass fails to halve it (the
    // FP64 residual floor, where the iterate is as accurate as FP64 values make it: a
    // pass that made it worse is undone).
    const int passes_allowed = opt->refine_passes > 0 ? opt->refine_passes : (opt->refine_passes == 0 && mixed ? 4 : 0);
    const double eta = opt->refine_reduction > 0.0 && opt->refine_reduction < 1.0 ? opt->refine_reduction : 1.0e-2;
    const bool refine = passes_allowed > 0 && st.status == LFG_OK && st.converged;
    if (refine) {
        ctx->ir_r.alloc(static_cast<size_t>(ctx->n));
        ctx->ir_d.alloc(static_cast<size_t>(ctx->n));
    }
    double rtrue = true_residual(rhs, xs, refine ? ctx->ir_r.p : nullptr);
    int passes = 0;
    bool floor_reached = false;
    while (refine && rtrue > target && passes < passes_allowed) {
        const double inner_tol = std::min(0.5, std::max(eta, target / rtrue));
        const lfg::PcgState sd = pcg_pass(ctx->ir_r.p, ctx->ir_d.p, LFG_STOP_RELATIVE, inner_tol,
                                          opt->max_iterations, &pass_ms);
        solve_ms += pass_ms;
        iterations += sd.iter;
        ++passes;
        if (sd.status != LFG_OK || !sd.converged) {  // the correction solve failed: keep x as it is
            st.status = sd.status != LFG_OK ? sd.status : LFG_ERR_NOT_CONVERGED;
            st.converged = 0;
            break;
        }
        lfg::launch_axpy(ctx->stream, ctx->n, 1.0, ctx->ir_d.p, xs);
        const double rnew = true_residual(rhs, xs, ctx->ir_r.p);
        st.rr = sd.rr;
        if (rnew > 0.5 * rtrue) {
            floor_reached = true;
            if (rnew > rtrue) {  // undo a correction that made the FP64 residual worse
                lfg::launch_axpy(ctx->stream, ctx->n, -1.0, ctx->ir_d.p, xs);
                sync(ctx);
                break;
            }
            rtrue = rnew;
            break;
        }
        rtrue = rnew;
    }
    if (refine && st.status == LFG_OK && rtrue > target && !floor_reached) {
        st.converged = 0;  // the passes ran out while each still halved the FP64 residual
    }
    if (x_out) LFG_HIP(hipMemcpyAsync(x_out, xs, ctx->n * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    sync(ctx);
    rep.converged = st.converged;
    rep.iterations = iterations;
    rep.residual_replacements = st.replacements + passes;
    rep.initial_residual = st.rnorm0;
    rep.final_residual = std::sqrt(std::max(0.0, st.rr));
    rep.true_residual = rtrue;
    rep.solve_ms = solve_ms;
    rep.refine_passes = passes;
    rep.status = st.status != LFG_OK ? st.status : (st.converged ? LFG_OK : LFG_ERR_NOT_CONVERGED);
    if (report) *report = rep;
    if (st.status == LFG_ERR_HIP && kind == LFG_PRECOND_IC0) throw LfgError(st.status, kIc0Fault);
    if (st.status != LFG_OK) throw LfgError(st.status, status_text(st.status));
    if (!st.converged) throw LfgError(LFG_ERR_NOT_CONVERGED, "maximum iterations reached without convergence");
}

}  // namespace

extern "C" {


int lfg_pcg(lfg_ctx* ctx, const lfg_pcg_options* opt, double* u_out, lfg_pcg_report* report) {
    return guarded(ctx, [&] {
        require(ctx->have_system, LFG_ERR_STATE, "lfg_dirichlet must precede lfg_pcg");
        ctx->have_solution = false;
        ctx->have_adjoint = false;
        ensure_vectors(ctx);
        try {
            run_pcg(ctx, opt, ctx->fc.p, ctx->x.p, u_out, report);
        } catch (const LfgError& e) {
            // a failed or unconverged solve still leaves its last iterate for reactions/fields
            ctx->have_solution = e.code == LFG_ERR_NOT_CONVERGED;
            throw;
        }
        ctx->have_solution = true;
        return LFG_OK;
    });
}

int lfg_adjoint_solve(lfg_ctx* ctx, const lfg_pcg_options* opt, const double* dJdu, double* lambda_out,
                      lfg_pcg_report* report) {
    return guarded(ctx, [&] {
        require(ctx->have_solution, LFG_ERR_STATE, "lfg_pcg must precede lfg_adjoint_solve");
        require(dJdu != nullptr, LFG_ERR_ARG, "missing adjoint right-hand side");
        LFG_HIP(hipSetDevice(ctx->device));
        ctx->have_adjoint = false;
        ctx->lam.alloc(static_cast<size_t>(ctx->n));
        ctx->adj_rhs.upload(dJdu, static_cast<size_t>(ctx->n), ctx->stream);
        lfg::launch_mask_constrained(ctx->stream, ctx->n, ctx->constrained.p, ctx->adj_rhs.p, ctx->adj_rhs.p);
        run_pcg(ctx, opt, ctx->adj_rhs.p, ctx->lam.p, lambda_out, report);
        ctx->have_adjoint = true;
        return LFG_OK;
    });
}

int lfg_sensitivity(lfg_ctx* ctx, const double* lambda, const double* dE_drho, double* grad_out) {
    return guarded(ctx, [&] {
        require(ctx->have_solution, LFG_ERR_STATE, "lfg_pcg must precede lfg_sensitivity");
        require(lambda != nullptr || ctx->have_adjoint, LFG_ERR_STATE,
                "lfg_adjoint_solve or an explicit lambda must precede lfg_sensitivity");
        require(grad_out != nullptr, LFG_ERR_ARG, "missing gradient output");
        LFG_HIP(hipSetDevice(ctx->device));
        if (lambda) {
            ctx->lam.upload(lambda, static_cast<size_t>(ctx->n), ctx->stream);
            ctx->have_adjoint = true;
        }
        const double* dE = nullptr;
        if (dE_drho) {
            ctx->dE.upload(dE_drho, static_cast<size_t>(ctx->n_elems), ctx->stream);
            dE = ctx->dE.p;
        }
        ctx->grad.alloc(static_cast<size_t>(ctx->n_elems));
        lfg::launch_sensitivity(ctx->stream, mesh_view(ctx), ctx->x.p, ctx->lam.p, dE, ctx->grad.p);
        ctx->grad.download(grad_out, static_cast<size_t>(ctx->n_elems), ctx->stream);
        sync(ctx);
        return LFG_OK;
    });
}

int lfg_reactions(lfg_ctx* ctx, double* reaction_out) {
    return guarded(ctx, [&] {
        require(ctx->have_solution && reaction_out, LFG_ERR_STATE, "lfg_pcg must precede lfg_reactions");
        LFG_HIP(hipSetDevice(ctx->device));
        lfg::launch_spmv_serial_rows(ctx->stream, csr_view(ctx), ctx->K.p, ctx->x.p, ctx->f.p, ctx->q.p);
        ctx->q.download(reaction_out, static_cast<size_t>(ctx->n), ctx->stream);
        sync(ctx);
        return LFG_OK;
    });
}

int lfg_postprocess(lfg_ctx* ctx, double* centre, double* gauss_points, double* nodal) {
    return guarded(ctx, [&] {
        require(ctx->have_solution, LFG_ERR_STATE, "lfg_pcg must precede lfg_postprocess");
        LFG_HIP(hipSetDevice(ctx->device));
        const size_t e = static_cast<size_t>(ctx->n_elems) + static_cast<size_t>(ctx->n_bars);
        ctx->centre.alloc(7 * e);
        ctx->gp.alloc(36 * e);
        ctx->nodal.alloc(4 * static_cast<size_t>(ctx->n_nodes));
        lfg::launch_fields(ctx->stream, mesh_view(ctx), ctx->x.p, ctx->centre.p, ctx->gp.p);
        lfg::launch_bar_fields(ctx->stream, mesh_view(ctx), bar_view(ctx), ctx->x.p, ctx->centre.p, ctx->gp.p);
        lfg::launch_smooth(ctx->stream, ctx->n_nodes, ctx->n_elems, ctx->node_elem_ptr.p, ctx->node_elem.p, ctx->gp.p,
                           ctx->nodal.p);
        if (centre) ctx->centre.download(centre, 7 * e, ctx->stream);
        if (gauss_points) ctx->gp.download(gauss_points, 36 * e, ctx->stream);
        if (nodal) ctx->nodal.download(nodal, 4 * static_cast<size_t>(ctx->n_nodes), ctx->stream);
        sync(ctx);
        return LFG_OK;
    });
}

int lfg_set_subdomains(lfg_ctx* ctx, int64_t n_subdomains, const int32_t* node_subdomain) {
    return guarded(ctx, [&] {
        require(ctx->have_pattern, LFG_ERR_STATE, "lfg_set_pattern must precede lfg_set_subdomains");
        require(n_subdomains > 0 && n_subdomains <= (1LL << 30) && node_subdomain, LFG_ERR_ARG,
                "subdomain count must be positive");
        const int n_nodes = ctx->n / 2;
        std::vector<int> map(static_cast<size_t>(n_nodes));
        for (int i = 0; i < n_nodes; ++i) {
            require(node_subdomain[i] >= 0 && node_subdomain[i] < n_subdomains, LFG_ERR_ARG, "subdomain id out of range");
            map[i] = node_subdomain[i];
        }
        ctx->sub_node.swap(map);
        ctx->n_sub = static_cast<int>(n_subdomains);
        if (ctx->precond_ready == LFG_PRECOND_SUBDOMAIN) ctx->precond_ready = -1;
        return LFG_OK;
    });
}

int lfg_precond_apply(lfg_ctx* ctx, int32_t preconditioner, const double* in, double* out) {
    return guarded(ctx, [&] {
        require(ctx->have_system && in && out, LFG_ERR_STATE, "lfg_dirichlet must precede lfg_precond_apply");
        LFG_HIP(hipSetDevice(ctx->device));
        ensure_vectors(ctx);
        long long degenerate = 0;
        setup_preconditioner(ctx, preconditioner, &degenerate);
        LFG_HIP(hipMemcpyAsync(ctx->r.p, in, ctx->n * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
        if (preconditioner == LFG_PRECOND_IC0) {
            LFG_HIP(hipMemsetAsync(ctx->flags.p + 2, 0, sizeof(int), ctx->stream));
            lfg::launch_ic0_apply(ctx->stream, ic0_view(ctx), ctx->r.p, ctx->z.p);
            int fault = 0;
            LFG_HIP(hipMemcpyAsync(&fault, ctx->flags.p + 2, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
            sync(ctx);
            require(fault == 0, LFG_ERR_HIP, kIc0Fault);
        } else if (preconditioner == LFG_PRECOND_SUBDOMAIN) {
            lfg::launch_subdomain_apply(ctx->stream, subdomain_view(ctx), ctx->r.p, ctx->z.p);
        } else {
            lfg::launch_precond_apply(ctx->stream, preconditioner, ctx->n,
                                      preconditioner == LFG_PRECOND_JACOBI ? ctx->minv.p : ctx->factor.p, ctx->r.p,
                                      ctx->z.p);
        }
        ctx->z.download(out, static_cast<size_t>(ctx->n), ctx->stream);
        sync(ctx);
        ctx->have_solution = false;
        return LFG_OK;
    });
}

// One SpMV launch of `variant`; with fused_dir (timing only) variant 7 runs the
// multiply its PCG iteration uses, forming p = M^-1 r + beta p_old (Jacobi) on the fly.
static void run_spmv(lfg_ctx* ctx, int precision, int variant, bool fused_dir = false) {
    lfg::PcgBuffers b = pcg_buffers(ctx, LFG_PRECOND_JACOBI, ctx->fc.p, ctx->x.p);
    b.st = ctx->state.p;
    const bool f32 = precision == LFG_PRECISION_MIXED;  // one value array per launch
    if (fused_dir && variant == kSliceFused)
        lfg::launch_pcg_spmv_slices_dir(ctx->stream, b, slice_view(ctx), f32 ? nullptr : ctx->slices64.p,
                                        f32 ? ctx->slices32.p : nullptr, ctx->p.p, ctx->p2.p);
    else if (variant == kSliceSym)
        lfg::launch_pcg_spmv_sym_slices(ctx->stream, b, sym_view(ctx), f32 ? nullptr : ctx->sym64.p,
                                        f32 ? ctx->sym32.p : nullptr);
    else if (stencil_layout(variant))
        launch_stencil_spmv(ctx, b, variant, !f32, f32);
    else if (slice_layout(variant))
        lfg::launch_pcg_spmv_slices(ctx->stream, b, slice_view(ctx), f32 ? nullptr : ctx->slices64.p,
                                    f32 ? ctx->slices32.p : nullptr);
    else
        lfg::launch_pcg_spmv(ctx->stream, b, csr_view(ctx), f32 ? nullptr : ctx->Kc.p, f32 ? ctx->K32.p : nullptr,
                             variant);
}

static void prepare_spmv(lfg_ctx* ctx, int precision, int variant) {
    ensure_vectors(ctx);
    lfg::PcgState zero{};
    LFG_HIP(hipMemcpyAsync(ctx->state.p, &zero, sizeof(zero), hipMemcpyHostToDevice, ctx->stream));
    if (variant == kSliceSym) {
        ensure_sym_slices(ctx, precision != LFG_PRECISION_MIXED, precision == LFG_PRECISION_MIXED);
    } else if (stencil_layout(variant)) {
        require(!(variant == kStencilExact && precision == LFG_PRECISION_MIXED), LFG_ERR_UNSUPPORTED,
                "the exact stencil form has no FP32 copy");
        ensure_stencil(ctx, variant, precision != LFG_PRECISION_MIXED, precision == LFG_PRECISION_MIXED);
    } else if (slice_layout(variant)) {
        ensure_slices(ctx, precision != LFG_PRECISION_MIXED, precision == LFG_PRECISION_MIXED);
    } else if (precision == LFG_PRECISION_MIXED && !ctx->k32_valid) {
        ctx->K32.alloc(static_cast<size_t>(ctx->nnz));
        lfg::launch_to_f32(ctx->stream, ctx->Kc.p, ctx->K32.p, ctx->nnz);
        ctx->k32_valid = true;
    }
}

// lfg_spmv on caller-owned device vectors without any allocation: the layout must exist
// (lfg_spmv or lfg_pcg with this precision and variant built it), and the p.q
// reduction's state, counters and partials live in the device's file-scope scratch
// array (lfg_noalloc.hip), cleared by a fill kernel at every call.
int lfg_spmv_noalloc(lfg_ctx* ctx, int32_t precision, int32_t variant, const double* x_dev, double* y_dev,
                     double* xy_dev) {
    return guarded(ctx, [&] {
        require(ctx->have_system, LFG_ERR_STATE, "lfg_dirichlet must precede lfg_spmv_noalloc");
        require(x_dev != nullptr && y_dev != nullptr, LFG_ERR_ARG, "missing device vectors");
        require(precision == LFG_PRECISION_FP64 || precision == LFG_PRECISION_MIXED, LFG_ERR_ARG,
                "precision must be fp64 or mixed");
        LFG_HIP(hipSetDevice(ctx->device));
        const bool block_order = (variant & LFG_SPMV_BLOCK_ORDER) != 0;
        const int v = resolve_variant(ctx, variant & ~LFG_SPMV_BLOCK_ORDER, LFG_PRECOND_JACOBI, block_order);
        const bool f32 = precision == LFG_PRECISION_MIXED;
        bool ready = false;
        unsigned int grid = 0;
        if (v == kSliceSym) {
            ready = f32 ? ctx->sym32_valid : ctx->sym64_valid;
            grid = static_cast<unsigned int>(lfg::slice_spmv_blocks(ctx->n_slices));
        } else if (stencil_layout(v)) {
            ready = v == kStencilExact ? (!f32 && ctx->stencil_exact_valid)
                                       : (f32 ? ctx->stencil32_valid : ctx->stencil64_valid);
            grid = v != kStencilExact && lfg::stencil_grid_multiply(ctx->stencil_sym)
                       ? lfg::stencil_grid_blocks(ctx->stencil_sym)
                       : lfg::stencil_spmv_blocks(ctx->n / 2);
        } else if (slice_layout(v)) {
            ready = f32 ? ctx->slices32_valid : ctx->slices64_valid;
            grid = static_cast<unsigned int>(lfg::slice_spmv_blocks(ctx->n_slices));
        } else {
            ready = !f32 || ctx->k32_valid;
            grid = static_cast<unsigned int>(lfg::spmv_blocks(ctx->n, v));
        }
        require(ready, LFG_ERR_STATE,
                "lfg_spmv_noalloc allocates nothing: run lfg_spmv or lfg_pcg with this precision and variant once "
                "first, which builds its value layout");
        require(grid <= static_cast<unsigned int>(lfg::kNoallocMaxGrid), LFG_ERR_UNSUPPORTED,
                "the multiply's grid exceeds the 8192 block partials of the allocation-free scratch");
        unsigned char* base = static_cast<unsigned char*>(lfg::noalloc_scratch(ctx->device));
        require(base != nullptr, LFG_ERR_STATE, "no scratch address for this device");
        const PairScope pair_scope(ctx, block_order);
        // solver state and counters start at zero (a fill kernel, no memset node)
        lfg::launch_fill(ctx->stream, reinterpret_cast<double*>(base),
                         static_cast<long long>(lfg::kNoallocPartialsAt / sizeof(double)), 0.0);
        lfg::PcgBuffers b{};
        b.n = ctx->n;
        b.p = const_cast<double*>(x_dev);
        b.q = y_dev;
        b.precond = LFG_PRECOND_JACOBI;
        b.st = reinterpret_cast<lfg::PcgState*>(base);
        b.tickets = reinterpret_cast<lfg::TicketSet*>(base + lfg::kNoallocTicketsAt);
        b.partials = reinterpret_cast<double*>(base + lfg::kNoallocPartialsAt);
        if (v == kSliceSym)
            lfg::launch_pcg_spmv_sym_slices(ctx->stream, b, sym_view(ctx), f32 ? nullptr : ctx->sym64.p,
                                            f32 ? ctx->sym32.p : nullptr);
        else if (stencil_layout(v))
            launch_stencil_spmv(ctx, b, v, !f32, f32);
        else if (slice_layout(v))
            lfg::launch_pcg_spmv_slices(ctx->stream, b, slice_view(ctx), f32 ? nullptr : ctx->slices64.p,
                                        f32 ? ctx->slices32.p : nullptr);
        else
            lfg::launch_pcg_spmv(ctx->stream, b, csr_view(ctx), f32 ? nullptr : ctx->Kc.p, f32 ? ctx->K32.p : nullptr,
                                 v);
        if (xy_dev)
            LFG_HIP(hipMemcpyAsync(xy_dev, &b.st->pq, sizeof(double), hipMemcpyDeviceToDevice, ctx->stream));
        sync(ctx);
        return LFG_OK;
    });
}

int lfg_layout_info_get(lfg_ctx* ctx, lfg_layout_info* info) {
    return guarded(ctx, [&] {
        require(ctx->have_pattern && info, LFG_ERR_STATE, "lfg_set_pattern must precede lfg_layout_info_get");
        lfg_layout_info out{};
        out.node_blocked = ctx->have_slices ? 1 : 0;
        out.n_slices = ctx->have_slices ? ctx->n_slices : 0;
        out.stencil_exact_planes = ctx->have_stencil_exact ? ctx->stencil_exact.n_off : 0;
        out.stencil_sym_planes = ctx->have_stencil_sym ? ctx->stencil_sym.n_off : 0;
        out.fused_window_pairs = ctx->have_stencil_sym ? ctx->stencil_sym.win_total : 0;
        for (int k = 0; k < out.stencil_exact_planes; ++k) out.stencil_offsets[k] = ctx->stencil_exact.off[k];
        out.grid_row_len = ctx->have_stencil_sym ? ctx->stencil_sym.row_len : 0;
        out.sym_slices = sym_slices_ok(ctx) ? 1 : 0;
        *info = out;
        return LFG_OK;
    });
}

int lfg_spmv(lfg_ctx* ctx, int32_t precision, int32_t variant, const double* x, double* y) {
    return guarded(ctx, [&] {
        require(ctx->have_system && x && y, LFG_ERR_STATE, "lfg_dirichlet must precede lfg_spmv");
        LFG_HIP(hipSetDevice(ctx->device));
        const bool block_order = (variant & LFG_SPMV_BLOCK_ORDER) != 0;
        const int v = resolve_variant(ctx, variant & ~LFG_SPMV_BLOCK_ORDER, LFG_PRECOND_JACOBI, block_order);
        const PairScope pair_scope(ctx, block_order);
        prepare_spmv(ctx, precision, v);
        LFG_HIP(hipMemcpyAsync(ctx->p.p, x, ctx->n * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
        run_spmv(ctx, precision, v);
        ctx->q.download(y, static_cast<size_t>(ctx->n), ctx->stream);
        sync(ctx);
        ctx->have_solution = false;
        return LFG_OK;
    });
}

int lfg_bench_spmv(lfg_ctx* ctx, int32_t precision, int32_t variant, int32_t reps, double* avg_ms) {
    return guarded(ctx, [&] {
        require(ctx->have_system && reps > 0 && avg_ms, LFG_ERR_STATE, "lfg_dirichlet must precede lfg_bench_spmv");
        LFG_HIP(hipSetDevice(ctx->device));
        const bool block_order = (variant & LFG_SPMV_BLOCK_ORDER) != 0;
        const int v = resolve_variant(ctx, variant & ~LFG_SPMV_BLOCK_ORDER, LFG_PRECOND_JACOBI, block_order);
        const PairScope pair_scope(ctx, block_order);
        if (fused_iteration(v)) {
            // whole Jacobi PCG iterations (the fused kernel does the multiply and the vector
            // work), from x = 0 with a stop threshold no residual reaches
            long long degenerate = 0;
            ensure_vectors(ctx);
            if (ctx->precond_ready != LFG_PRECOND_JACOBI) setup_preconditioner(ctx, LFG_PRECOND_JACOBI, &degenerate);
            const bool f32 = precision == LFG_PRECISION_MIXED;
            if (v == kSliceSym) {
                ensure_sym_slices(ctx, !f32, f32);
            } else if (slice_layout(v)) {
                ensure_slices(ctx, !f32, f32);
            } else {
                ensure_stencil(ctx, v, !f32, f32);
            }
            ctx->p2.alloc(static_cast<size_t>(ctx->n));
            ctx->r_alt.alloc(static_cast<size_t>(ctx->n));
            ctx->q_alt.alloc(static_cast<size_t>(ctx->n));
            const lfg::PcgBuffers b = pcg_buffers(ctx, LFG_PRECOND_JACOBI, ctx->fc.p, ctx->x.p);
            lfg::launch_pcg_init(ctx->stream, b, LFG_STOP_RELATIVE, 1e-300, 1LL << 60, 0.0, f32 ? -1.0 : 0.0, 1);
            enqueue_iteration(ctx, b, f32, 0, v, 0);  // warm-up: p0, A p0
            enqueue_iteration(ctx, b, f32, 0, v, 1);
            LFG_HIP(hipEventRecord(ctx->ev_start, ctx->stream));
            for (int i = 0; i < reps; ++i) enqueue_iteration(ctx, b, f32, 0, v, i & 1);
            LFG_HIP(hipEventRecord(ctx->ev_stop, ctx->stream));
            sync(ctx);
            *avg_ms = elapsed(ctx->ev_start, ctx->ev_stop) / reps;
            ctx->have_solution = false;
            return LFG_OK;
        }
        const bool fused = v == kSliceFused;  // the multiply its Jacobi / block PCG iteration runs
        if (fused) {
            long long degenerate = 0;
            if (ctx->precond_ready != LFG_PRECOND_JACOBI) setup_preconditioner(ctx, LFG_PRECOND_JACOBI, &degenerate);
            ctx->p2.alloc(static_cast<size_t>(ctx->n));
        }
        prepare_spmv(ctx, precision, v);
        lfg::launch_fill(ctx->stream, ctx->p.p, ctx->n, 1.0);
        lfg::launch_fill(ctx->stream, ctx->r.p, ctx->n, 1.0);
        run_spmv(ctx, precision, v, fused);  // warm-up
        LFG_HIP(hipEventRecord(ctx->ev_start, ctx->stream));
        for (int i = 0; i < reps; ++i) run_spmv(ctx, precision, v, fused);
        LFG_HIP(hipEventRecord(ctx->ev_stop, ctx->stream));
        sync(ctx);
        *avg_ms = elapsed(ctx->ev_start, ctx->ev_stop) / reps;
        ctx->have_solution = false;
        return LFG_OK;
    });
}

int lfg_batch_set_problems(lfg_ctx* ctx, int32_t n_problems, const int32_t* node_ptr) {
    return guarded(ctx, [&] {
        require(ctx->have_pattern, LFG_ERR_STATE, "lfg_set_pattern must precede lfg_batch_set_problems");
        require(ctx->have_slices, LFG_ERR_UNSUPPORTED, "batches need a node-blocked pattern");
        require(n_problems > 0 && node_ptr && node_ptr[0] == 0 && node_ptr[n_problems] == ctx->n_nodes, LFG_ERR_ARG,
                "problem node ranges must start at 0 and cover every node");
        const int n_slices = ctx->n_slices;
        std::vector<int> slice_problem(static_cast<size_t>(n_slices));
        std::vector<int> pslice(static_cast<size_t>(n_problems) + 1);
        for (int p = 0; p < n_problems; ++p) {
            require(node_ptr[p + 1] > node_ptr[p] && node_ptr[p] % 64 == 0 && node_ptr[p + 1] % 64 == 0, LFG_ERR_ARG,
                    "problem node ranges must be non-empty multiples of 64");
            pslice[p] = node_ptr[p] / 64;
            for (int sl = node_ptr[p] / 64; sl < node_ptr[p + 1] / 64; ++sl) slice_problem[sl] = p;
            // block diagonal: rows of problem p only reference problem p's nodes
            const long long lo = 2LL * node_ptr[p], hi = 2LL * node_ptr[p + 1];
            for (long long r = lo; r < hi; ++r)
                for (int k = ctx->h_row_ptr[r]; k < ctx->h_row_ptr[r + 1]; ++k)
                    require(ctx->h_col[k] >= lo && ctx->h_col[k] < hi, LFG_ERR_ARG, "an element spans two problems");
        }
        pslice[n_problems] = n_slices;
        LFG_HIP(hipSetDevice(ctx->device));
        ctx->batch_problems = n_problems;
        ctx->batch_slice_problem.upload(slice_problem.data(), slice_problem.size(), ctx->stream);
        ctx->batch_pslice_ptr.upload(pslice.data(), pslice.size(), ctx->stream);
        ctx->batch_problem_done.alloc(static_cast<size_t>(n_problems));
        ctx->batch_phase32.alloc(static_cast<size_t>(n_problems));
        ctx->batch_active.alloc(static_cast<size_t>(n_slices));
        ctx->batch_active_first.alloc(static_cast<size_t>(n_problems));
        ctx->batch_tickets.alloc(32 * static_cast<size_t>(n_problems));
        LFG_HIP(hipMemsetAsync(ctx->batch_tickets.p, 0, ctx->batch_tickets.n * sizeof(unsigned int), ctx->stream));
        ctx->batch_partials.alloc(8 * static_cast<size_t>(n_slices));
        ctx->batch_global.alloc(1);
        if (!ctx->h_batch_global)
            LFG_HIP(hipHostMalloc(&ctx->h_batch_global, 2 * sizeof(lfg::BatchGlobal), hipHostMallocDefault));
        sync(ctx);
        ctx->have_batch = true;
        ctx->have_batch_loads = false;
        ctx->batch_tuned_for = -1;
        return LFG_OK;
    });
}

int lfg_batch_set_loads(lfg_ctx* ctx, int32_t n_cases, const double* loads, int64_t n_constraints,
                        const int32_t* dofs, const double* values) {
    return guarded(ctx, [&] {
        require(ctx->have_batch && ctx->have_K, LFG_ERR_STATE,
                "lfg_batch_set_problems and assembly must precede lfg_batch_set_loads");
        require((n_cases == 1 || n_cases == 2 || n_cases == 4) && loads, LFG_ERR_ARG,
                "a batch takes 1, 2 or 4 load cases");
        require(n_constraints >= 0 && (n_constraints == 0 || (dofs && values)), LFG_ERR_ARG, "invalid constraints");
        const size_t n = static_cast<size_t>(ctx->n);
        std::vector<int> flag(n, 0);
        std::vector<double> presc(n, 0.0);
        for (int64_t k = 0; k < n_constraints; ++k) {
            require(dofs[k] >= 0 && dofs[k] < ctx->n, LFG_ERR_ARG, "validated boundary references an unknown node");
            flag[dofs[k]] = 1;
            presc[dofs[k]] = values[k];
        }
        LFG_HIP(hipSetDevice(ctx->device));
        ctx->constrained.upload(flag.data(), n, ctx->stream);
        ctx->prescribed.upload(presc.data(), n, ctx->stream);
        ctx->Kc.alloc(static_cast<size_t>(ctx->nnz));
        ctx->fc.alloc(n);
        const size_t vec = n * static_cast<size_t>(n_cases);
        ctx->batch_b.alloc(vec);
        for (int c = 0; c < n_cases; ++c) {
            ctx->f.upload(loads + static_cast<size_t>(c) * n, n, ctx->stream);
            lfg::launch_dirichlet(ctx->stream, csr_view(ctx), ctx->K.p, ctx->f.p, ctx->constrained.p, ctx->prescribed.p,
                                  ctx->Kc.p, ctx->fc.p);
            lfg::launch_batch_scatter_case(ctx->stream, ctx->n, n_cases, c, ctx->fc.p, ctx->batch_b.p);
        }
        ctx->batch_minv.alloc(n);
        ensure_vectors(ctx);
        LFG_HIP(hipMemsetAsync(ctx->flags.p, 0, sizeof(int), ctx->stream));
        lfg::launch_batch_jacobi(ctx->stream, csr_view(ctx), ctx->Kc.p, ctx->batch_minv.p, ctx->flags.p);
        for (DevArray<double>* v : {&ctx->batch_x, &ctx->batch_r, &ctx->batch_p, &ctx->batch_q, &ctx->batch_res})
            v->alloc(vec);
        ctx->batch_tmp.alloc(n);
        int bad = 0;
        LFG_HIP(hipMemcpyAsync(&bad, ctx->flags.p, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
        sync(ctx);
        require(bad == 0, LFG_ERR_NONFINITE, "Jacobi preconditioner encountered a non-finite diagonal");
        ctx->batch_cases = n_cases;
        ctx->batch_sys.alloc(static_cast<size_t>(ctx->batch_problems) * n_cases);
        ctx->have_system = true;
        ctx->have_solution = ctx->have_adjoint = false;
        ctx->k32_valid = false;
        ctx->slices64_valid = ctx->slices32_valid = ctx->sym64_valid = ctx->sym32_valid = false;
        ctx->stencil64_valid = ctx->stencil32_valid = ctx->stencil_exact_valid = false;
        ctx->precond_ready = -1;
        ctx->have_batch_loads = true;
        return LFG_OK;
    });
}

}  // extern "C"

namespace {

lfg::BatchView batch_view(lfg_ctx* c, const lfg_batch_options* opt) {
    lfg::BatchView bv{};
    bv.n_problems = c->batch_problems;
    bv.n_cases = c->batch_cases;
    bv.n_slices = c->n_slices;
    bv.slice_problem = c->batch_slice_problem.p;
    bv.pslice_ptr = c->batch_pslice_ptr.p;
    bv.x = c->batch_x.p;
    bv.r = c->batch_r.p;
    bv.p = c->batch_p.p;
    bv.q = c->batch_q.p;
    bv.b = c->batch_b.p;
    bv.minv = c->batch_minv.p;
    bv.sys = c->batch_sys.p;
    bv.problem_done = c->batch_problem_done.p;
    bv.tickets = c->batch_tickets.p;
    bv.partials = c->batch_partials.p;
    bv.global = c->batch_global.p;
    const bool mixed = opt->precision == LFG_PRECISION_MIXED;
    bv.problem_phase32 = mixed ? c->batch_phase32.p : nullptr;
    bv.active = c->batch_active.p;
    bv.active_first = c->batch_active_first.p;
    // default 1e-3: a later switch than the single solve's, which refines afterwards; the batch
    // has one pass, and FP32 values from 1e-2 cost its slowest problems 16% more iterations
    bv.switch_tau = opt->mixed_switch != 0.0 ? opt->mixed_switch : 1.0e-3;
    bv.tol = opt->tolerance;
    bv.stop_rule = opt->stop_rule;
    bv.max_iter = opt->max_iterations;
    return bv;
}

// Launch-shape autotuner: for each block size and slices-per-wavefront candidate, runs
// a few iterations of this batch from its initial state (which the solve then restores)
// and keeps the fastest shape; the multiply alone is a poor proxy, since the vector
// kernels share the shape and a shape's register and cache pressure differs per kernel.
// Mixed precision times the FP32-value iteration, where the mixed solve spends most of its
// iterations. Then times the multiply with its per-problem fold at the chosen shape.
void batch_autotune(lfg_ctx* c, const lfg::BatchView& bv_in, bool mixed, bool sym) {
    const int threads[] = {64, 128, 256, 512};
    const int per_wave[] = {1, 2, 4};
    constexpr int kReps = 4;
    lfg::BatchView bv = bv_in;
    if (mixed) bv.switch_tau = -1.0;  // every problem on FP32 values from the first iteration
    const lfg::SliceView v = slice_view(c);
    const lfg::SymSliceView sv = sym_view(c);
    const double* p64 = sym ? c->sym64.p : c->slices64.p;
    const float* p32 = mixed ? c->slices32.p : nullptr;
    const lfg::SymSliceView* svp = sym ? &sv : nullptr;
    lfg::BatchGlobal start{};
    start.n_active = c->n_slices;
    auto iteration = [&](const lfg::BatchShape& shape) {
        lfg::launch_batch_spmv(c->stream, bv, shape, v, p64, p32, svp);
        lfg::launch_batch_update(c->stream, bv, shape);
        lfg::launch_batch_direction(c->stream, bv, shape);
    };
    float best = 0.0f;
    lfg::BatchShape chosen{256, 1};
    LFG_HIP(hipEventRecord(c->ev_ring[0], c->stream));
    for (int t : threads)
        for (int w : per_wave) {
            const lfg::BatchShape shape{t, w};
            LFG_HIP(hipMemcpyAsync(c->batch_global.p, &start, sizeof(start), hipMemcpyHostToDevice, c->stream));
            lfg::launch_batch_init(c->stream, bv, shape);
            iteration(shape);  // warm-up
            LFG_HIP(hipEventRecord(c->ev_start, c->stream));
            for (int i = 0; i < kReps; ++i) iteration(shape);
            LFG_HIP(hipEventRecord(c->ev_stop, c->stream));
            sync(c);
            const float ms = elapsed(c->ev_start, c->ev_stop);
            if (best == 0.0f || ms < best) {
                best = ms;
                chosen = shape;
            }
        }
    LFG_HIP(hipEventRecord(c->ev_ring[1], c->stream));
    sync(c);
    c->batch_tune_ms = elapsed(c->ev_ring[0], c->ev_ring[1]);
    // the multiply and its fold at the chosen shape from the initial state (every problem
    // running; the repeated products recompute the same q and alpha)
    LFG_HIP(hipMemcpyAsync(c->batch_global.p, &start, sizeof(start), hipMemcpyHostToDevice, c->stream));
    lfg::launch_batch_init(c->stream, bv, chosen);
    lfg::launch_batch_spmv(c->stream, bv, chosen, v, p64, p32, svp);
    LFG_HIP(hipEventRecord(c->ev_start, c->stream));
    for (int i = 0; i < 5; ++i) lfg::launch_batch_spmv(c->stream, bv, chosen, v, p64, p32, svp);
    LFG_HIP(hipEventRecord(c->ev_stop, c->stream));
    sync(c);
    c->batch_multiply_ms = elapsed(c->ev_start, c->ev_stop) / 5.0;
    c->batch_shape = chosen;
    c->batch_tuned_for = mixed ? LFG_PRECISION_MIXED : LFG_PRECISION_FP64;
}

// PCG over all systems (right-hand sides batch_b, solutions in batch_x); returns the
// number of graph launches.
long long batch_solve(lfg_ctx* c, const lfg_batch_options* opt, const lfg::BatchView& bv, bool mixed, bool sym) {
    const lfg::SliceView v = slice_view(c);
    const lfg::SymSliceView sv = sym_view(c);
    const lfg::BatchShape shape = c->batch_shape;
    // five kernels per iteration (two multiply-and-fold, the direction): 16 iterations and
    // the compaction per captured graph
    const int per_graph = opt->check_every > 0 ? opt->check_every : 16;
    // every slice active at the start (the init kernel writes the identity list)
    lfg::BatchGlobal start{};
    start.n_active = c->n_slices;
    LFG_HIP(hipMemcpyAsync(c->batch_global.p, &start, sizeof(start), hipMemcpyHostToDevice, c->stream));
    lfg::launch_batch_init(c->stream, bv, shape);
    auto iteration = [&](const lfg::BatchShape& sh) {
        lfg::launch_batch_spmv(c->stream, bv, sh, v, sym ? c->sym64.p : c->slices64.p,
                               mixed ? c->slices32.p : nullptr, sym ? &sv : nullptr);
        lfg::launch_batch_update(c->stream, bv, sh);
        lfg::launch_batch_direction(c->stream, bv, sh);
    };
    // One graph per grid span: every slice, then a half, a quarter, ... (down to 256
    // slices). The host launches the smallest whose grid still gives each slice of the
    // active list a wavefront, by the active count it last read (the list only shrinks, so
    // the count when the graph runs is never larger): a launch in the tail then dispatches
    // blocks for the running problems' slices, not for every slice of the batch.
    std::vector<int> spans{c->n_slices};
    for (int sp = c->n_slices / 2; sp >= 256; sp /= 2) spans.push_back(sp);
    struct Graphs {
        std::vector<hipGraph_t> g;
        std::vector<hipGraphExec_t> e;
        ~Graphs() {
            for (hipGraphExec_t x : e) (void)hipGraphExecDestroy(x);
            for (hipGraph_t x : g) (void)hipGraphDestroy(x);
        }
    } graphs;
    for (int sp : spans) {
        lfg::BatchShape sh = shape;
        sh.span = sp;
        hipGraph_t graph = nullptr;
        LFG_HIP(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
        for (int i = 0; i < per_graph; ++i) iteration(sh);
        // drop the problems that stopped during these iterations from the active list (until
        // then every kernel skips them by their done flag)
        lfg::launch_batch_compact(c->stream, bv);
        LFG_HIP(hipStreamEndCapture(c->stream, &graph));
        graphs.g.push_back(graph);
        hipGraphExec_t exec = nullptr;
        LFG_HIP(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
        graphs.e.push_back(exec);
    }
    const long long launch_cap = std::max<long long>(1, (opt->max_iterations + per_graph - 1) / per_graph);
    long long launched = 0;
    int active = c->n_slices;  // the active slice count the host read last
    auto launch_one = [&](int slot) {
        size_t k = 0;
        while (k + 1 < spans.size() && spans[k + 1] >= active) ++k;
        LFG_HIP(hipGraphLaunch(graphs.e[k], c->stream));
        LFG_HIP(hipMemcpyAsync(c->h_batch_global + slot, c->batch_global.p, sizeof(lfg::BatchGlobal),
                               hipMemcpyDeviceToHost, c->stream));
        LFG_HIP(hipEventRecord(c->ev_ring[slot], c->stream));
        ++launched;
    };
    launch_one(0);
    for (long long k = 0;; ++k) {
        const int slot = static_cast<int>(k & 1);
        if (launched < launch_cap) launch_one(1 - slot);
        LFG_HIP(hipEventSynchronize(c->ev_ring[slot]));
        if (c->h_batch_global[slot].done_all || launched >= launch_cap) break;
        active = c->h_batch_global[slot].n_active;
    }
    sync(c);
    return launched;
}

}  // namespace

extern "C" {

int lfg_batch_pcg(lfg_ctx* ctx, const lfg_batch_options* opt, double* u_out, lfg_batch_system_report* reports,
                  lfg_batch_summary* summary) {
    return guarded(ctx, [&] {
        require(ctx->have_batch_loads, LFG_ERR_STATE, "lfg_batch_set_loads must precede lfg_batch_pcg");
        require(opt != nullptr, LFG_ERR_ARG, "missing batch options");
        require(opt->tolerance > 0.0 && opt->tolerance < 1.0, LFG_ERR_ARG, "solver tolerance must satisfy 0 < tolerance < 1");
        require(opt->max_iterations > 0, LFG_ERR_ARG, "max_iterations must be greater than zero");
        require(opt->mixed_switch < 1.0, LFG_ERR_ARG,
                "mixed_switch must be below 1 (negative: FP32 values from the first iteration)");
        const bool mixed = opt->precision == LFG_PRECISION_MIXED;
        LFG_HIP(hipSetDevice(ctx->device));
        ensure_slices(ctx, true, mixed);
        // FP64 products on the symmetric node slices where the pattern admits them (each
        // mirrored pair of blocks streamed once); the FP32 ones and the final FP64 residual
        // on the node slices
        const bool sym = sym_slices_ok(ctx);
        if (sym) ensure_sym_slices(ctx, true, false);
        lfg::BatchView bv = batch_view(ctx, opt);
        if (opt->tune) {
            if (ctx->batch_tuned_for != opt->precision) batch_autotune(ctx, bv, mixed, sym);
        } else {
            const int t = opt->block_threads > 0 ? opt->block_threads : 256;
            const int w = opt->slices_per_wave > 0 ? opt->slices_per_wave : 1;
            require(t % 64 == 0 && t >= 64 && t <= 512 && w >= 1, LFG_ERR_ARG,
                    "block_threads must be 64-512 in steps of 64, slices_per_wave positive");
            ctx->batch_shape = lfg::BatchShape{t, w};
        }
        const lfg::SliceView v = slice_view(ctx);
        const size_t n_sys = static_cast<size_t>(ctx->batch_problems) * ctx->batch_cases;
        std::vector<lfg::BatchSystem> sys(n_sys);

        LFG_HIP(hipEventRecord(ctx->ev_start, ctx->stream));
        const long long launches = batch_solve(ctx, opt, bv, mixed, sym);
        // FP64 residual of every system's solution
        lfg::launch_batch_residual(ctx->stream, bv, ctx->batch_shape, v, ctx->slices64.p, ctx->batch_res.p);
        LFG_HIP(hipEventRecord(ctx->ev_stop, ctx->stream));
        LFG_HIP(hipMemcpyAsync(sys.data(), ctx->batch_sys.p, n_sys * sizeof(lfg::BatchSystem), hipMemcpyDeviceToHost,
                               ctx->stream));
        sync(ctx);
        const double solve_ms = elapsed(ctx->ev_start, ctx->ev_stop);
        if (u_out) {
            for (int c = 0; c < ctx->batch_cases; ++c) {
                lfg::launch_batch_gather_case(ctx->stream, ctx->n, ctx->batch_cases, c, ctx->batch_x.p,
                                              ctx->batch_tmp.p);
                ctx->batch_tmp.download(u_out + static_cast<size_t>(c) * ctx->n, static_cast<size_t>(ctx->n),
                                        ctx->stream);
            }
            sync(ctx);
        }
        int converged = 0, switched = 0;
        long long max_iter = 0;
        int first_status = LFG_OK;
        for (size_t k = 0; k < n_sys; ++k) {
            const lfg::BatchSystem& st = sys[k];
            const int ok = st.converged;
            converged += ok;
            if (k % static_cast<size_t>(ctx->batch_cases) == 0 && st.switch_iter >= 0) ++switched;
            max_iter = std::max(max_iter, st.iter);
            if (first_status == LFG_OK && st.status != LFG_OK) first_status = st.status;
            if (reports) {
                lfg_batch_system_report& r = reports[k];
                r.converged = ok;
                r.status = st.status != LFG_OK ? st.status : (ok ? LFG_OK : LFG_ERR_NOT_CONVERGED);
                r.iterations = st.iter;
                r.switch_iteration = st.switch_iter;
                r.initial_residual = st.rnorm0;
                r.final_residual = std::sqrt(std::max(0.0, st.rr));
                r.true_residual = std::sqrt(std::max(0.0, st.true_rr));
            }
        }
        if (summary) {
            summary->solve_ms = solve_ms;
            summary->tune_ms = ctx->batch_tune_ms;
            summary->graph_launches = launches;
            summary->problems_switched = switched;
            summary->block_threads = ctx->batch_shape.block_threads;
            summary->slices_per_wave = ctx->batch_shape.slices_per_wave;
            summary->systems_converged = converged;
            summary->max_iterations = max_iter;
            summary->multiply_ms = ctx->batch_multiply_ms;
            summary->sym_slices = sym ? 1 : 0;
        }
        if (first_status != LFG_OK) throw LfgError(first_status, status_text(first_status));
        if (converged != static_cast<int>(n_sys))
            throw LfgError(LFG_ERR_NOT_CONVERGED, "maximum iterations reached without convergence");
        return LFG_OK;
    });
}

int lfg_bench_copy(lfg_ctx* ctx, int64_t bytes, int32_t reps, double* avg_ms) {
    return guarded(ctx, [&] {
        require(bytes > 0 && reps > 0 && avg_ms, LFG_ERR_ARG, "copy size and repetitions must be positive");
        LFG_HIP(hipSetDevice(ctx->device));
        void* src = nullptr;
        void* dst = nullptr;
        LFG_HIP(hipMalloc(&src, static_cast<size_t>(bytes)));
        struct Free {
            void* p;
            ~Free() {
                if (p) (void)hipFree(p);
            }
        } free_src{src};
        LFG_HIP(hipMalloc(&dst, static_cast<size_t>(bytes)));
        Free free_dst{dst};
        LFG_HIP(hipMemsetAsync(src, 1, static_cast<size_t>(bytes), ctx->stream));
        LFG_HIP(hipMemcpyAsync(dst, src, static_cast<size_t>(bytes), hipMemcpyDeviceToDevice, ctx->stream));
        LFG_HIP(hipEventRecord(ctx->ev_start, ctx->stream));
        for (int i = 0; i < reps; ++i)
            LFG_HIP(hipMemcpyAsync(dst, src, static_cast<size_t>(bytes), hipMemcpyDeviceToDevice, ctx->stream));
        LFG_HIP(hipEventRecord(ctx->ev_stop, ctx->stream));
        sync(ctx);
        *avg_ms = elapsed(ctx->ev_start, ctx->ev_stop) / reps;
        return LFG_OK;
    });
}

int lfg_bench_stream(lfg_ctx* ctx, int64_t bytes, int32_t reps, double* avg_ms) {
    return guarded(ctx, [&] {
        require(bytes >= 16 && reps > 0 && avg_ms, LFG_ERR_ARG, "stream size and repetitions must be positive");
        LFG_HIP(hipSetDevice(ctx->device));
        const long long pairs = bytes / 16;
        DevArray<double> buf[5];
        for (auto& a : buf) {
            a.alloc(static_cast<size_t>(2 * pairs));
            LFG_HIP(hipMemsetAsync(a.p, 0, static_cast<size_t>(16 * pairs), ctx->stream));
        }
        lfg::launch_stream_probe(ctx->stream, pairs, buf[0].p, buf[1].p, buf[2].p, buf[3].p, buf[4].p);
        LFG_HIP(hipEventRecord(ctx->ev_start, ctx->stream));
        for (int i = 0; i < reps; ++i)
            lfg::launch_stream_probe(ctx->stream, pairs, buf[0].p, buf[1].p, buf[2].p, buf[3].p, buf[4].p);
        LFG_HIP(hipEventRecord(ctx->ev_stop, ctx->stream));
        sync(ctx);
        *avg_ms = elapsed(ctx->ev_start, ctx->ev_stop) / reps;
        return LFG_OK;
    });
}

}  // extern "C"
