// LinearElasticSolver::solve on the gfx950 backend. The stage order, results and
// exception texts are those of src/solver.cpp:875-960; host stages (numbering,
// RCM, pattern + map, colouring, loads) run here and every numeric stage runs on
// the GPU through include/lightfem_gpu.h. There is no CPU fallback: a missing
// device is an error.
#include <algorithm>
#include <chrono>
#include <climits>
#include <cmath>
#include <cstdint>
#include <memory>
#include <stdexcept>
#include <string>

#include <sys/mman.h>

#include "lightfem/lightfem.hpp"
#include "lightfem_gpu.h"
#include "lightfem_host.h"

namespace lightfem {

namespace {

double ms_since(std::chrono::steady_clock::time_point t) {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t).count();
}

struct DeviceContext {
    lfg_ctx* ctx = nullptr;
    DeviceContext() {
        const int rc = lfg_create(0, &ctx);
        if (rc != LFG_OK) throw std::runtime_error("no usable gfx950 device: lfg_create failed");
    }
    ~DeviceContext() { lfg_destroy(ctx); }
    DeviceContext(const DeviceContext&) = delete;
    DeviceContext& operator=(const DeviceContext&) = delete;

    // Rethrows a failed stage with the reference's exception type and text.
    void check(int rc) const {
        if (rc == LFG_OK) return;
        const std::string message = lfg_last_error(ctx);
        switch (rc) {
            case LFG_ERR_ARG: throw std::invalid_argument(message);
            case LFG_ERR_STATE: throw std::logic_error(message);
            default: throw std::runtime_error(message);
        }
    }
};

struct HostArrays {
    std::vector<double> xy;
    std::vector<int32_t> enodes;      // node indices
    std::vector<int32_t> enode_ids;   // node ids (for RCM)
    std::vector<int32_t> elem_mat;
    std::vector<double> materials;
    std::vector<int32_t> bar_nodes;     // 2 node indices per bar
    std::vector<int32_t> bar_node_ids;  // 2 node ids per bar (for RCM)
    std::vector<int32_t> bar_mat;
    std::vector<double> bar_area;
};

HostArrays host_arrays(const Model& model) {
    HostArrays h;
    h.xy.resize(2 * model.nodes.size());
    for (std::size_t i = 0; i < model.nodes.size(); ++i) {
        h.xy[2 * i] = model.nodes[i].position.x;
        h.xy[2 * i + 1] = model.nodes[i].position.y;
    }
    h.materials.reserve(3 * model.materials.size());
    for (const Material& m : model.materials) {
        h.materials.push_back(m.young_modulus);
        h.materials.push_back(m.poisson_ratio);
        h.materials.push_back(m.thickness);
    }
    h.enodes.reserve(4 * model.elements.size());
    h.elem_mat.reserve(model.elements.size());
    for (const Q4Element& e : model.elements) {
        for (const int id : e.node_ids) {
            const Node* node = model.find_node(id);
            if (!node) throw std::logic_error("validated element references an unknown node");
            h.enodes.push_back(static_cast<int32_t>(node->dof_x / 2));
            h.enode_ids.push_back(id);
        }
        const Material* m = model.find_material(e.material_id);
        if (!m) throw std::logic_error("validated element references an unknown material");
        h.elem_mat.push_back(static_cast<int32_t>(m - model.materials.data()));
    }
    for (const TrussElement& t : model.trusses) {
        for (const int id : t.node_ids) {
            const Node* node = model.find_node(id);
            if (!node) throw std::logic_error("validated truss references an unknown node");
            h.bar_nodes.push_back(static_cast<int32_t>(node->dof_x / 2));
            h.bar_node_ids.push_back(id);
        }
        const Material* m = model.find_material(t.material_id);
        if (!m) throw std::logic_error("validated truss references an unknown material");
        h.bar_mat.push_back(static_cast<int32_t>(m - model.materials.data()));
        h.bar_area.push_back(t.area);
    }
    return h;
}

// Sizes a fresh vector of n value-initialised elements, advising 2 MB pages for its
// storage first: the first touch of hundreds of MB in 4 KB pages (the pattern and map,
// K values, element results at 1000x1000) costs several times the work that fills them.
template <typename T>
void size_fresh(std::vector<T>& v, std::size_t n) {
    constexpr uintptr_t kHuge = uintptr_t{2} << 20;
    v.reserve(n);
    const uintptr_t lo = (reinterpret_cast<uintptr_t>(v.data()) + kHuge - 1) & ~(kHuge - 1);
    const uintptr_t hi = reinterpret_cast<uintptr_t>(v.data() + n) & ~(kHuge - 1);
    if (hi > lo) (void)madvise(reinterpret_cast<void*>(lo), hi - lo, MADV_HUGEPAGE);  // advice only
    v.resize(n);
}

struct Pattern {
    std::vector<int32_t> row_ptr, col, emap, bmap;
};

Pattern build_pattern(const Model& model, const HostArrays& h) {
    Pattern p;
    const int64_t n_nodes = static_cast<int64_t>(model.nodes.size());
    const int64_t n_elems = static_cast<int64_t>(model.elements.size());
    const int64_t n_bars = static_cast<int64_t>(model.trusses.size());
    const int64_t nnz = lfh_csr_pattern_nnz_bars(n_nodes, n_elems, h.enodes.data(), n_bars, h.bar_nodes.data());
    if (nnz < 0) throw std::out_of_range("element DOF is outside CSR matrix dimension");
    p.row_ptr.resize(static_cast<std::size_t>(2 * n_nodes + 1));
    size_fresh(p.col, static_cast<std::size_t>(nnz));
    size_fresh(p.emap, static_cast<std::size_t>(64 * n_elems));
    p.bmap.resize(static_cast<std::size_t>(16 * n_bars));
    if (lfh_csr_pattern_bars(n_nodes, n_elems, h.enodes.data(), n_bars, h.bar_nodes.data(), p.row_ptr.data(),
                             p.col.data(), p.emap.data(), p.bmap.data()) != 0)
        throw std::overflow_error("CSR pattern exceeds the int32 index range of the gfx950 backend");
    return p;
}

CsrProfile profile_of(const Pattern& p) {
    int64_t out[3] = {0, 0, 0};
    lfh_csr_profile(static_cast<int64_t>(p.row_ptr.size()) - 1, p.row_ptr.data(), p.col.data(), out);
    CsrProfile c;
    c.dimension = p.row_ptr.size() - 1;
    c.nonzeros = static_cast<std::size_t>(out[0]);
    c.bandwidth = static_cast<std::size_t>(out[1]);
    c.profile = static_cast<std::size_t>(out[2]);
    return c;
}

std::array<Vec2, 4> corners(const Model& model, const Q4Element& e) {
    std::array<Vec2, 4> c{};
    for (int i = 0; i < 4; ++i) {
        const Node* node = model.find_node(e.node_ids[i]);
        if (!node) throw std::logic_error("validated element references an unknown node");
        c[i] = node->position;
    }
    return c;
}

// solver.cpp:303-332: nodal loads, then body forces, then edge loads, in model order.
std::vector<double> load_vector(const Model& model) {
    std::vector<double> f(model.summary().dofs, 0.0);
    for (const NodalLoad& load : model.loads) {
        const Node* node = model.find_node(load.node_id);
        if (!node) throw std::logic_error("validated load references an unknown node");
        f[node->dof_x] += load.fx;
        f[node->dof_y] += load.fy;
    }
    const auto add = [&](const Q4Element& e, const std::array<double, 8>& local) {
        const auto dofs = q4_dof_indices(model, e);
        for (int i = 0; i < 8; ++i) f[dofs[i]] += local[i];
    };
    for (const BodyForce& force : model.body_forces) {
        const Q4Element* e = model.find_element(force.element_id);
        if (!e) throw std::logic_error("validated body force references an unknown element");
        const Material* m = model.find_material(e->material_id);
        if (!m) throw std::logic_error("validated element references an unknown material");
        add(*e, q4_body_force(corners(model, *e), *m, force.bx, force.by));
    }
    for (const EdgeLoad& load : model.edge_loads) {
        const Q4Element* e = model.find_element(load.element_id);
        if (!e) throw std::logic_error("validated edge load references an unknown element");
        const Material* m = model.find_material(e->material_id);
        if (!m) throw std::logic_error("validated element references an unknown material");
        const auto c = corners(model, *e);
        add(*e, load.type == EdgeLoadType::Traction ? q4_edge_traction(c, *m, load.edge, load.first, load.second)
                                                    : q4_edge_pressure(c, *m, load.edge, load.first));
    }
    return f;
}

// The int32 row offsets / columns widened to the CsrMatrix index type (parallel: 36 M
// columns at 1000x1000).
std::vector<std::size_t> widened(const std::vector<int32_t>& v) {
    std::vector<std::size_t> out;
    size_fresh(out, v.size());
    const int64_t count = static_cast<int64_t>(v.size());
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < count; ++i) out[static_cast<std::size_t>(i)] = static_cast<std::size_t>(v[i]);
    return out;
}

int assembly_code(const std::string& name) {
    if (name == "serial" || name == "scan") return LFG_ASSEMBLY_SERIAL;  // scan computes the same sums
    if (name == "atomic") return LFG_ASSEMBLY_ATOMIC;
    if (name == "coloring") return LFG_ASSEMBLY_COLORING;
    if (name == "thread_local") return LFG_ASSEMBLY_THREAD_LOCAL;
    throw std::logic_error("validated model has an unknown assembly strategy");
}

}  // namespace

std::vector<int> reverse_cuthill_mckee(const Model& model) {
    if (!model.is_indexed()) throw std::logic_error("model must be indexed before RCM ordering");
    std::vector<int32_t> ids(model.nodes.size());
    for (std::size_t i = 0; i < model.nodes.size(); ++i) ids[i] = model.nodes[i].id;
    std::vector<int32_t> enode_ids, bar_ids;
    enode_ids.reserve(4 * model.elements.size());
    for (const Q4Element& e : model.elements)
        for (const int id : e.node_ids) enode_ids.push_back(id);
    for (const TrussElement& t : model.trusses)
        for (const int id : t.node_ids) bar_ids.push_back(id);
    std::vector<int32_t> order(ids.size());
    if (lfh_rcm_bars(static_cast<int64_t>(ids.size()), ids.data(), static_cast<int64_t>(model.elements.size()),
                     enode_ids.data(), static_cast<int64_t>(model.trusses.size()), bar_ids.data(), order.data()) != 0)
        throw std::logic_error("validated element references an unknown node");
    return std::vector<int>(order.begin(), order.end());
}

AnalysisResult LinearElasticSolver::solve(Model& model) const {
    if (!model.is_indexed()) throw std::logic_error("model must pass validation and sort_and_index before solving");
    const SolverSettings& s = model.settings;
    const auto solve_start = std::chrono::steady_clock::now();
    AnalysisResult result;
    result.performance.reordering = s.reordering;
    const bool rcm = s.reordering == "rcm";
    const auto profile_start = std::chrono::steady_clock::now();
    if (rcm) {
        result.performance.before = profile_of(build_pattern(model, host_arrays(model)));
        model.index_with_node_order(reverse_cuthill_mckee(model));
    }
    result.performance.profile_ms = ms_since(profile_start);

    DeviceContext dev;
    const auto pattern_start = std::chrono::steady_clock::now();
    const HostArrays h = host_arrays(model);
    const Pattern pat = build_pattern(model, h);
    result.performance.pattern_ms = ms_since(pattern_start);
    const int64_t n_nodes = static_cast<int64_t>(model.nodes.size());
    const int64_t n_elems = static_cast<int64_t>(model.elements.size());
    const int64_t n = 2 * n_nodes;
    const int64_t nnz = static_cast<int64_t>(pat.col.size());
    dev.check(lfg_set_mesh(dev.ctx, n_nodes, h.xy.data(), n_elems, h.enodes.data(), h.elem_mat.data(),
                           static_cast<int64_t>(model.materials.size()), h.materials.data(),
                           s.plane_mode == PlaneMode::PlaneStrain ? LFG_PLANE_STRAIN : LFG_PLANE_STRESS));
    dev.check(lfg_set_pattern(dev.ctx, n, nnz, pat.row_ptr.data(), pat.col.data(), pat.emap.data()));
    const int64_t n_bars = static_cast<int64_t>(model.trusses.size());
    if (n_bars > 0)
        dev.check(lfg_set_bars(dev.ctx, n_bars, h.bar_nodes.data(), h.bar_mat.data(), h.bar_area.data(),
                               pat.bmap.data()));
    const int strategy = assembly_code(s.assembly);
    if (strategy == LFG_ASSEMBLY_COLORING) {
        const std::size_t total = static_cast<std::size_t>(n_elems + n_bars);
        std::vector<int32_t> color(total), ptr(total + 2), elems(total);
        int32_t count = 0;
        lfh_element_colors_bars(n_nodes, n_elems, h.enodes.data(), n_bars, h.bar_nodes.data(), color.data(), &count,
                                ptr.data(), elems.data());
        dev.check(lfg_set_colors(dev.ctx, count, ptr.data(), elems.data()));
    }
    lfg_assembly_report asm_rep{};
    dev.check(lfg_assemble(dev.ctx, strategy, &asm_rep));
    const std::vector<double> load = load_vector(model);
    dev.check(lfg_set_load(dev.ctx, load.data()));

    const auto retain_start = std::chrono::steady_clock::now();
    std::vector<double> values;
    size_fresh(values, static_cast<std::size_t>(nnz));
    dev.check(lfg_get_stiffness(dev.ctx, values.data()));
    result.stiffness = CsrMatrix(static_cast<std::size_t>(n), widened(pat.row_ptr), widened(pat.col), std::move(values));
    result.load = load;
    result.performance.retain_matrix_ms = ms_since(retain_start);
    const auto after_start = std::chrono::steady_clock::now();
    result.performance.after = profile_of(pat);
    result.performance.profile_ms += ms_since(after_start);
    if (!rcm) result.performance.before = result.performance.after;
    result.performance.assembly = s.assembly;
    result.performance.requested_threads = s.threads;
    result.performance.threads_used = 1;
    result.performance.coloring_count = static_cast<std::size_t>(asm_rep.coloring_count);
    result.performance.element_stiffness_ms = asm_rep.element_stiffness_ms;
    result.performance.assembly_ms = asm_rep.assembly_ms;

    // solver.cpp:340-351, 355-413
    const auto boundary_start = std::chrono::steady_clock::now();
    std::vector<std::pair<int32_t, double>> constraints;
    for (const DirichletBC& bc : model.boundaries) {
        const Node* node = model.find_node(bc.node_id);
        if (!node) throw std::logic_error("validated boundary references an unknown node");
        if (bc.ux) constraints.emplace_back(static_cast<int32_t>(node->dof_x), *bc.ux);
        if (bc.uy) constraints.emplace_back(static_cast<int32_t>(node->dof_y), *bc.uy);
    }
    std::sort(constraints.begin(), constraints.end(), [](const auto& a, const auto& b) { return a.first < b.first; });
    std::vector<int32_t> cdofs;
    std::vector<double> cvals;
    for (const auto& c : constraints) {
        cdofs.push_back(c.first);
        cvals.push_back(c.second);
    }
    dev.check(lfg_dirichlet(dev.ctx, static_cast<int64_t>(cdofs.size()), cdofs.data(), cvals.data()));
    // solver.cpp:916-940: a bar end no other element stiffens leaves a zero pivot
    lfg_pivot_report pivots{};
    const int pivot_rc = lfg_pin_zero_pivots(dev.ctx, &pivots);
    if (pivot_rc == LFG_ERR_MECHANISM) {
        const std::size_t i = static_cast<std::size_t>(pivots.dof);
        const Node& node = model.nodes[i / 2];
        throw std::runtime_error(
            "the structure is a mechanism: a bar end has no transverse stiffness but carries a transverse load; add "
            "a support there or load the bar along its axis [dof " + std::to_string(i) + " load " +
            std::to_string(pivots.load) + " diag " + std::to_string(pivots.diag) + " node " +
            std::to_string(node.id) + " xy " + std::to_string(node.position.x) + "," +
            std::to_string(node.position.y) + "]");
    }
    dev.check(pivot_rc);
    result.performance.boundary_ms = ms_since(boundary_start);

    // block: subdomains of build_dof_subdomain (solver.cpp:631-654); one node per
    // subdomain is the 2x2 nodal form fused into the iteration
    const bool subdomains = s.preconditioner == "block" && s.subdomain_nodes != 1;
    std::vector<int32_t> node_sub;
    int64_t n_sub = 0;
    if (subdomains) {
        node_sub.resize(static_cast<std::size_t>(n_nodes));
        if (lfh_subdomain_split(n_nodes, h.xy.data(), s.subdomain_split == "chunk" ? 1 : 0,
                                static_cast<int64_t>(s.subdomain_nodes), node_sub.data(), &n_sub) != 0)
            throw std::invalid_argument("invalid subdomain split");
        dev.check(lfg_set_subdomains(dev.ctx, n_sub, node_sub.data()));
    }
    lfg_pcg_options opt{};
    opt.preconditioner = s.preconditioner == "ic0"     ? LFG_PRECOND_IC0
                         : subdomains                  ? LFG_PRECOND_SUBDOMAIN
                         : s.preconditioner == "block" ? LFG_PRECOND_BLOCK2
                                                       : LFG_PRECOND_JACOBI;
    opt.storage = s.storage == "bsr" ? LFG_STORAGE_BSR : LFG_STORAGE_CSR;
    opt.precision = s.precision == "mixed" ? LFG_PRECISION_MIXED : LFG_PRECISION_FP64;
    opt.stop_rule = s.stop_rule == "relative" ? LFG_STOP_RELATIVE : LFG_STOP_REFERENCE;
    opt.tolerance = s.tolerance;
    opt.max_iterations = static_cast<int64_t>(s.max_iterations);
    lfg_pcg_report rep{};
    result.displacement.assign(static_cast<std::size_t>(n), 0.0);
    const int rc = lfg_pcg(dev.ctx, &opt, result.displacement.data(), &rep);
    LinearSolverReport& sr = result.solver;
    sr.converged = rep.converged != 0;
    sr.iterations = static_cast<std::size_t>(rep.iterations);
    sr.initial_residual = rep.initial_residual;
    sr.final_residual = rep.final_residual;
    sr.solve_ms = rep.solve_ms;
    sr.matvec_calls = static_cast<std::size_t>(rep.iterations);
    sr.preconditioner = s.preconditioner;
    sr.storage = s.storage;
    sr.preconditioner_setup_ms = rep.setup_ms;
    sr.precision = s.precision;
    sr.stop_rule = s.stop_rule;
    sr.true_residual = rep.true_residual;
    sr.residual_replacements = static_cast<std::size_t>(rep.residual_replacements);
    if (s.storage == "bsr") {
        // 2x2 blocks of the block storage (BsrMatrix::from_csr, sparse.cpp:128-174): the
        // union of the block columns of rows 2I and 2I+1, 4 doubles and one uint32 each
        std::size_t blocks = 0;
        for (int64_t br = 0; br < n / 2; ++br) {
            int32_t p0 = pat.row_ptr[2 * br], p1 = pat.row_ptr[2 * br + 1];
            const int32_t e0 = pat.row_ptr[2 * br + 1], e1 = pat.row_ptr[2 * br + 2];
            while (p0 < e0 || p1 < e1) {
                const int32_t c0 = p0 < e0 ? pat.col[p0] / 2 : INT32_MAX, c1 = p1 < e1 ? pat.col[p1] / 2 : INT32_MAX;
                const int32_t bc = std::min(c0, c1);
                while (p0 < e0 && pat.col[p0] / 2 == bc) ++p0;
                while (p1 < e1 && pat.col[p1] / 2 == bc) ++p1;
                ++blocks;
            }
        }
        sr.matrix_bytes_per_matvec = blocks * (4 * sizeof(double) + sizeof(uint32_t));
    } else {
        // the report's figure for CSR storage (solver.cpp:683-684, 743-745): a double value
        // and a size_t column per nonzero, whatever layout the device multiplies with
        sr.matrix_bytes_per_matvec = static_cast<std::size_t>(nnz) * (sizeof(double) + sizeof(std::size_t));
    }
    if (s.preconditioner == "block") {
        // subdomain_capture_ratio (precond.cpp:107-119): the share of nonzeros whose row
        // and column lie in one subdomain
        sr.subdomains = subdomains ? static_cast<std::size_t>(n_sub) : static_cast<std::size_t>(n_nodes);
        std::size_t inside = 0;
        for (int64_t row = 0; row < n; ++row)
            for (int32_t k = pat.row_ptr[row]; k < pat.row_ptr[row + 1]; ++k) {
                const int64_t a = row / 2, b = pat.col[k] / 2;
                if (subdomains ? node_sub[a] == node_sub[b] : a == b) ++inside;
            }
        sr.capture_ratio = nnz ? static_cast<double>(inside) / static_cast<double>(nnz) : 0.0;
    }
    if (rc == LFG_ERR_NOT_SPD) throw std::runtime_error("PCG encountered non-positive curvature; system is not SPD");
    if (rc == LFG_ERR_NONFINITE) {
        throw std::runtime_error(rep.iterations == 0
                                     ? "PCG failed to form a positive initial preconditioned residual"
                                     : "PCG preconditioned residual became non-positive or non-finite");
    }
    if (rc == LFG_ERR_NOT_CONVERGED || !sr.converged) {
        sr.message = "maximum iterations reached without convergence";
        throw std::runtime_error("PCG did not converge: " + sr.message);
    }
    dev.check(rc);
    sr.message = rep.initial_residual <= (s.stop_rule == "relative" ? s.tolerance * rep.initial_residual
                                                                     : s.tolerance * std::max(1.0, rep.initial_residual))
                     ? "initial residual satisfies tolerance"
                     : "converged";

    const auto post_start = std::chrono::steady_clock::now();
    result.reaction.clear();
    size_fresh(result.reaction, static_cast<std::size_t>(n));
    dev.check(lfg_reactions(dev.ctx, result.reaction.data()));
    const int64_t n_results = n_elems + n_bars;  // Q4 elements, then bars (solver.cpp:753-805)
    std::vector<double> centre, gp, nodal;
    size_fresh(centre, 7 * static_cast<std::size_t>(n_results));
    size_fresh(gp, 36 * static_cast<std::size_t>(n_results));
    size_fresh(nodal, 4 * static_cast<std::size_t>(n_nodes));
    dev.check(lfg_postprocess(dev.ctx, centre.data(), gp.data(), nodal.data()));
    result.element_results.clear();
    size_fresh(result.element_results, static_cast<std::size_t>(n_results));
#pragma omp parallel for schedule(static)
    for (int64_t e = 0; e < n_results; ++e) {
        ElementResult& er = result.element_results[static_cast<std::size_t>(e)];
        er.element_id = e < n_elems ? model.elements[static_cast<std::size_t>(e)].id
                                    : model.trusses[static_cast<std::size_t>(e - n_elems)].id;
        const double* c = centre.data() + 7 * e;
        er.field.strain = {c[0], c[1], c[2]};
        er.field.stress = {c[3], c[4], c[5]};
        er.field.von_mises = c[6];
        for (int q = 0; q < 4; ++q) {
            const double* g = gp.data() + 36 * e + 9 * q;
            er.gauss_points[q].xi = g[0];
            er.gauss_points[q].eta = g[1];
            er.gauss_points[q].field.strain = {g[2], g[3], g[4]};
            er.gauss_points[q].field.stress = {g[5], g[6], g[7]};
            er.gauss_points[q].field.von_mises = g[8];
        }
    }
    result.node_results.clear();
    size_fresh(result.node_results, static_cast<std::size_t>(n_nodes));
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n_nodes; ++i) {
        NodeResult& nr = result.node_results[static_cast<std::size_t>(i)];
        nr.node_id = model.nodes[static_cast<std::size_t>(i)].id;
        nr.stress = {nodal[4 * i], nodal[4 * i + 1], nodal[4 * i + 2]};
        nr.von_mises = nodal[4 * i + 3];
    }
    result.performance.postprocess_ms = ms_since(post_start);
    result.performance.total_ms = ms_since(solve_start);
    return result;
}

}  // namespace lightfem
