// Test infrastructure only: a harness around the reference LightFEM sources.
//
// This translation unit is compiled together with the reference's own source
// files, where they lie (oracle/Makefile). It includes the reference solver
// translation unit so the file-static stages (element_colors, assemble,
// apply_dirichlet, pcg_solve, ...) become callable, and it never ships: its
// outputs are the golden vectors under tests/golden/ and the "reference" CPU
// baseline timed by bench.py.
//
// Modes
//   dump  <model.fem> <outdir> [key=value ...]   every intermediate of one solve as .npy
//   bench <nx> <ny> <young.bin> <plane_mode> <nu> <iterations> <threads> <preconditioner>
//         [<relative_tolerance> <u_out.bin>]
//         PCG on a structured cantilever, one JSON line on stdout: exactly
//         <iterations> steps, or with <relative_tolerance> to ||r|| <= tol ||r0|| (the
//         reference's rule tol' max(1, ||r0||) with tol' = tol ||r0|| / max(1, ||r0||))
//         capped at <iterations>, the displacements written as raw float64
#include "src/solver.cpp"  // reference solver TU (anonymous-namespace stages)

#include "lightfem/input_parser.hpp"
#include "lightfem/validation.hpp"

#include <cstdio>
#include <cstring>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <sstream>
#include <string>

namespace {

using namespace lightfem;

// Minimal NPY v1.0 writer (little-endian, C order).
void write_npy(const std::string& path, const char* descr, const std::vector<std::size_t>& shape,
               const void* data, std::size_t bytes) {
    std::string header = "{'descr': '";
    header += descr;
    header += "', 'fortran_order': False, 'shape': (";
    for (std::size_t i = 0; i < shape.size(); ++i) {
        header += std::to_string(shape[i]);
        if (shape.size() == 1 || i + 1 < shape.size()) header += ",";
        if (i + 1 < shape.size()) header += " ";
    }
    header += "), }";
    const std::size_t prefix = 10;
    std::size_t total = prefix + header.size() + 1;
    const std::size_t pad = (64 - total % 64) % 64;
    header.append(pad, ' ');
    header += '\n';
    std::ofstream out(path, std::ios::binary);
    if (!out) throw std::runtime_error("cannot write " + path);
    const unsigned char magic[8] = {0x93, 'N', 'U', 'M', 'P', 'Y', 1, 0};
    out.write(reinterpret_cast<const char*>(magic), 8);
    const unsigned short hlen = static_cast<unsigned short>(header.size());
    out.write(reinterpret_cast<const char*>(&hlen), 2);
    out.write(header.data(), static_cast<std::streamsize>(header.size()));
    out.write(static_cast<const char*>(data), static_cast<std::streamsize>(bytes));
}

void npy_f8(const std::string& dir, const std::string& name, const std::vector<double>& v,
            std::vector<std::size_t> shape = {}) {
    if (shape.empty()) shape = {v.size()};
    write_npy(dir + "/" + name + ".npy", "<f8", shape, v.data(), v.size() * sizeof(double));
}

void npy_i8(const std::string& dir, const std::string& name, const std::vector<long long>& v,
            std::vector<std::size_t> shape = {}) {
    if (shape.empty()) shape = {v.size()};
    write_npy(dir + "/" + name + ".npy", "<i8", shape, v.data(), v.size() * sizeof(long long));
}

template <typename T>
std::vector<long long> as_i8(const std::vector<T>& v) {
    return std::vector<long long>(v.begin(), v.end());
}

void apply_override(SolverSettings& s, const std::string& kv) {
    const auto eq = kv.find('=');
    if (eq == std::string::npos) throw std::invalid_argument("override must be key=value: " + kv);
    const std::string key = kv.substr(0, eq);
    const std::string value = kv.substr(eq + 1);
    if (key == "assembly") s.assembly = value;
    else if (key == "reordering") s.reordering = value;
    else if (key == "preconditioner") s.preconditioner = value;
    else if (key == "storage") s.storage = value;
    else if (key == "subdomain_split") s.subdomain_split = value;
    else if (key == "subdomain_nodes") s.subdomain_nodes = std::stoul(value);
    else if (key == "threads") s.threads = std::stoul(value);
    else if (key == "tolerance") s.tolerance = std::stod(value);
    else if (key == "max_iterations") s.max_iterations = std::stoul(value);
    else if (key == "plane_mode") s.plane_mode = value == "plane_strain" ? PlaneMode::PlaneStrain : PlaneMode::PlaneStress;
    else throw std::invalid_argument("unknown override key: " + key);
}

std::string json_profile(const CsrProfile& p) {
    std::ostringstream o;
    o << "{\"dimension\": " << p.dimension << ", \"nonzeros\": " << p.nonzeros << ", \"bandwidth\": "
      << p.bandwidth << ", \"profile\": " << p.profile << "}";
    return o.str();
}

int dump(const std::string& path, const std::string& dir, const std::vector<std::string>& overrides) {
    Model model = InputParser{}.parse_file(path);
    for (const auto& kv : overrides) apply_override(model.settings, kv);
    const ValidationReport report = ModelValidator{}.validate(model);
    if (!report.ok()) {
        std::cerr << "model failed validation\n";
        return 3;
    }
    model.sort_and_index();
    Model solved = model;  // the full solve() reorders its own copy

    // Stage-by-stage restatement of LinearElasticSolver::solve (solver.cpp:875-960)
    // so every intermediate is observable.
    const bool rcm = model.settings.reordering == "rcm";
    CsrProfile before;
    if (rcm) {
        before = profile_for_model(model);
        const std::vector<int> order = reverse_cuthill_mckee(model);
        std::vector<long long> order_i8(order.begin(), order.end());
        npy_i8(dir, "rcm_order", order_i8);
        model.index_with_node_order(order);
    }
    std::vector<long long> node_ids;
    std::vector<double> xy;
    for (const Node& n : model.nodes) {
        node_ids.push_back(n.id);
        xy.push_back(n.position.x);
        xy.push_back(n.position.y);
    }
    npy_i8(dir, "node_ids", node_ids);
    npy_f8(dir, "xy", xy, {model.nodes.size(), 2});

    const std::size_t dofs = model.summary().dofs;
    const ElementDofList element_dofs = element_dof_lists(model);
    const CsrAssemblyPattern pattern = build_csr_assembly_pattern(dofs, element_dofs);
    npy_i8(dir, "row_ptr", as_i8(pattern.matrix.row_offsets()));
    npy_i8(dir, "col", as_i8(pattern.matrix.column_indices()));
    {
        std::vector<long long> ed, em;
        for (const auto& d : element_dofs) for (std::size_t v : d) ed.push_back(static_cast<long long>(v));
        for (const auto& m : pattern.element_value_indices) for (std::size_t v : m) em.push_back(static_cast<long long>(v));
        if (model.trusses.empty()) {
            npy_i8(dir, "edofs", ed, {element_dofs.size(), 8});
            npy_i8(dir, "emap", em, {element_dofs.size(), 64});
        } else {
            npy_i8(dir, "edofs_flat", ed);
            npy_i8(dir, "emap_flat", em);
        }
    }
    {
        const auto groups = element_colors(model);
        std::vector<long long> color(element_dofs.size(), -1);
        for (std::size_t c = 0; c < groups.size(); ++c)
            for (std::size_t e : groups[c]) color[e] = static_cast<long long>(c);
        npy_i8(dir, "colors", color);
    }
    {
        std::vector<double> ke;
        std::vector<double> mat;  // E nu t per element
        for (const Q4Element& element : model.elements) {
            const Material* material = model.find_material(element.material_id);
            const ElementMatrix m = q4_stiffness(q4_coordinates(model, element), *material, model.settings.plane_mode);
            ke.insert(ke.end(), m.begin(), m.end());
            mat.push_back(material->young_modulus);
            mat.push_back(material->poisson_ratio);
            mat.push_back(material->thickness);
        }
        npy_f8(dir, "ke", ke, {model.elements.size(), 64});
        npy_f8(dir, "elem_material", mat, {model.elements.size(), 3});
    }
    if (!model.trusses.empty()) {
        // bars: dofs, element matrices (element.cpp:266-286) and (E, area)
        std::vector<long long> td;
        std::vector<double> kt, tp;
        for (const TrussElement& truss : model.trusses) {
            const auto d = truss_dof_indices(model, truss);
            for (std::size_t v : d) td.push_back(static_cast<long long>(v));
            const Material* material = model.find_material(truss.material_id);
            std::array<Vec2, 2> coordinates{};
            for (std::size_t i = 0; i < 2; ++i) coordinates[i] = model.find_node(truss.node_ids[i])->position;
            const auto k = truss_stiffness(coordinates, truss.area, material->young_modulus);
            kt.insert(kt.end(), k.begin(), k.end());
            tp.push_back(material->young_modulus);
            tp.push_back(truss.area);
        }
        npy_i8(dir, "tdofs", td, {model.trusses.size(), 4});
        npy_f8(dir, "ke_truss", kt, {model.trusses.size(), 16});
        npy_f8(dir, "truss_props", tp, {model.trusses.size(), 2});
    }

    AssembledSystem assembled = assemble(model);
    npy_f8(dir, "K", assembled.stiffness.values());
    npy_f8(dir, "f", assembled.load);
    const CsrProfile after = csr_profile(assembled.stiffness);
    if (!rcm) before = after;

    const auto constraints = collect_constraints(model);
    {
        std::vector<long long> cd;
        std::vector<double> cv;
        for (const auto& c : constraints) {
            cd.push_back(static_cast<long long>(c.first));
            cv.push_back(c.second);
        }
        npy_i8(dir, "cdofs", cd);
        npy_f8(dir, "cvals", cv);
    }
    CsrMatrix kc = assembled.stiffness;
    std::vector<double> fc = assembled.load;
    apply_dirichlet(kc, fc, constraints);
    npy_f8(dir, "Kc", kc.values());
    npy_f8(dir, "fc", fc);

    // Preconditioner applications on the constrained system, probed with fc and
    // with a fixed non-trivial vector.
    {
        std::vector<double> probe(dofs);
        for (std::size_t i = 0; i < dofs; ++i) probe[i] = std::sin(0.37 * static_cast<double>(i) + 0.11) + 0.05;
        npy_f8(dir, "probe", probe);
        npy_f8(dir, "jacobi_invdiag", inverse_diagonal_of(kc.diagonal()));
        IncompleteCholesky0 ic0(kc);
        std::vector<double> z(dofs);
        ic0.apply(probe, z);
        npy_f8(dir, "ic0_apply_probe", z);
        std::size_t count = 0;
        std::vector<std::uint32_t> dsub = build_dof_subdomain(model, "chunk", 1, count);
        SubdomainBlockJacobi block(kc, dsub, count);
        block.apply(probe, z);
        npy_f8(dir, "block1_apply_probe", z);
        // the block storage product (sparse.cpp:128-198) on the same probe
        const BsrMatrix bsr = BsrMatrix::from_csr(kc);
        npy_f8(dir, "bsr_probe", bsr.multiply(probe));
        std::ostringstream o;
        o << "{\"breakdowns\": " << ic0.stats().breakdowns << ", \"factor_entries\": "
          << ic0.stats().factor_entries << ", \"block1_capture_ratio\": " << block.stats().capture_ratio
          << ", \"block1_degenerate\": " << block.stats().degenerate_subdomains
          << ", \"bsr_block_nonzeros\": " << bsr.block_nonzeros();
        // the model's own subdomain block-Jacobi (solver.cpp:631-654, precond.cpp:19-291)
        if (model.settings.preconditioner == "block") {
            std::size_t n_sub = 0;
            const std::vector<std::uint32_t> own =
                build_dof_subdomain(model, model.settings.subdomain_split, model.settings.subdomain_nodes, n_sub);
            std::vector<long long> own64(own.begin(), own.end());
            npy_i8(dir, "dof_subdomain", own64);
            SubdomainBlockJacobi sb(kc, own, n_sub);
            sb.apply(probe, z);
            npy_f8(dir, "block_apply_probe", z);
            o << ", \"block_subdomains\": " << sb.stats().subdomains << ", \"block_capture_ratio\": "
              << std::setprecision(17) << sb.stats().capture_ratio << ", \"block_degenerate\": "
              << sb.stats().degenerate_subdomains << ", \"block_factor_nonzeros\": " << sb.stats().factor_nonzeros;
        }
        o << "}";
        std::ofstream(dir + "/precond.json") << o.str() << "\n";
    }

    // The full library call on the untouched copy: displacement, reactions and fields.
    const AnalysisResult result = LinearElasticSolver{}.solve(solved);
    npy_f8(dir, "u", result.displacement);
    npy_f8(dir, "R", result.reaction);
    {
        std::vector<double> centre, gp, nodal;
        for (const ElementResult& er : result.element_results) {
            for (double v : er.field.strain) centre.push_back(v);
            for (double v : er.field.stress) centre.push_back(v);
            centre.push_back(er.field.von_mises);
            for (const GaussPointField& g : er.gauss_points) {
                gp.push_back(g.xi);
                gp.push_back(g.eta);
                for (double v : g.field.strain) gp.push_back(v);
                for (double v : g.field.stress) gp.push_back(v);
                gp.push_back(g.field.von_mises);
            }
        }
        for (const NodeResult& nr : result.node_results) {
            for (double v : nr.stress) nodal.push_back(v);
            nodal.push_back(nr.von_mises);
        }
        npy_f8(dir, "elem_centre", centre, {result.element_results.size(), 7});
        npy_f8(dir, "elem_gp", gp, {result.element_results.size(), 4, 9});
        npy_f8(dir, "node_fields", nodal, {result.node_results.size(), 4});
    }
    std::ofstream meta(dir + "/meta.json");
    meta << "{\"dofs\": " << dofs << ", \"nnz\": " << assembled.stiffness.nonzeros()
         << ", \"elements\": " << model.elements.size() << ", \"nodes\": " << model.nodes.size()
         << ", \"iterations\": " << result.solver.iterations
         << ", \"converged\": " << (result.solver.converged ? "true" : "false")
         << ", \"initial_residual\": " << std::setprecision(17) << result.solver.initial_residual
         << ", \"final_residual\": " << result.solver.final_residual
         << ", \"coloring_count\": " << result.performance.coloring_count
         << ", \"assembly\": \"" << model.settings.assembly << "\""
         << ", \"reordering\": \"" << model.settings.reordering << "\""
         << ", \"preconditioner\": \"" << model.settings.preconditioner << "\""
         << ", \"storage\": \"" << model.settings.storage << "\""
         << ", \"tolerance\": " << model.settings.tolerance
         << ", \"plane_mode\": \"" << to_string(model.settings.plane_mode) << "\""
         << ", \"before\": " << json_profile(result.performance.before)
         << ", \"after\": " << json_profile(result.performance.after)
         << ", \"subdomains\": " << result.solver.subdomains
         << ", \"matrix_bytes_per_matvec\": " << result.solver.matrix_bytes_per_matvec
         << ", \"subdomain_split\": \"" << model.settings.subdomain_split << "\""
         << ", \"subdomain_nodes\": " << model.settings.subdomain_nodes
         << ", \"capture_ratio\": " << result.solver.capture_ratio << "}\n";
    return 0;
}

// Structured nx x ny cantilever on [0,1]x[0,1] with the numbering of
// tests/test_solver.cpp:371-396: node j*(nx+1)+i+1, CCW elements, left edge
// fixed, right-edge tributary loads totalling -1.
Model structured_model(int nx, int ny, const std::vector<double>& young, PlaneMode mode, double nu) {
    Model model;
    model.settings.plane_mode = mode;
    const auto node_id = [nx](int i, int j) { return j * (nx + 1) + i + 1; };
    for (int j = 0; j <= ny; ++j)
        for (int i = 0; i <= nx; ++i)
            model.nodes.push_back(Node{node_id(i, j), Vec2{static_cast<double>(i) / nx, static_cast<double>(j) / ny}});
    const std::size_t n_elem = static_cast<std::size_t>(nx) * ny;
    if (young.empty()) {
        model.materials.push_back(Material{1, 1.0, nu, 1.0, 0});
    } else {
        for (std::size_t e = 0; e < n_elem; ++e)
            model.materials.push_back(Material{static_cast<int>(e + 1), young[e], nu, 1.0, 0});
    }
    int eid = 1;
    for (int j = 0; j < ny; ++j)
        for (int i = 0; i < nx; ++i) {
            model.elements.push_back(Q4Element{eid, {{node_id(i, j), node_id(i + 1, j), node_id(i + 1, j + 1), node_id(i, j + 1)}},
                                               young.empty() ? 1 : eid, 0});
            ++eid;
        }
    for (int j = 0; j <= ny; ++j) model.boundaries.push_back(DirichletBC{node_id(0, j), 0.0, 0.0, 0});
    for (int j = 0; j <= ny; ++j) {
        const double fy = (j == 0 || j == ny) ? -0.5 / ny : -1.0 / ny;
        model.loads.push_back(NodalLoad{node_id(nx, j), 0.0, fy, 0});
    }
    return model;
}

int bench(int argc, char** argv) {
    if (argc != 10 && argc != 12) {
        std::cerr << "bench <nx> <ny> <young.bin|-> <plane_mode> <nu> <iterations> <threads> <preconditioner> "
                     "[<relative_tolerance> <u_out.bin>]\n";
        return 1;
    }
    const double rel_tol = argc == 12 ? std::atof(argv[10]) : 0.0;
    const int nx = std::atoi(argv[2]);
    const int ny = std::atoi(argv[3]);
    std::vector<double> young;
    if (std::strcmp(argv[4], "-") != 0) {
        young.resize(static_cast<std::size_t>(nx) * ny);
        std::ifstream in(argv[4], std::ios::binary);
        in.read(reinterpret_cast<char*>(young.data()), static_cast<std::streamsize>(young.size() * sizeof(double)));
        if (!in) throw std::runtime_error("cannot read young modulus file");
    }
    const PlaneMode mode = std::string(argv[5]) == "plane_strain" ? PlaneMode::PlaneStrain : PlaneMode::PlaneStress;
    const double nu = std::atof(argv[6]);
    const std::size_t iterations = std::strtoul(argv[7], nullptr, 10);
    const std::size_t threads = std::strtoul(argv[8], nullptr, 10);
    Model model = structured_model(nx, ny, young, mode, nu);
    model.settings.threads = threads;
    model.settings.preconditioner = argv[9];
    model.settings.max_iterations = iterations;
    model.settings.tolerance = 1e-300;  // run exactly `iterations` steps
    model.sort_and_index();

    const auto t0 = std::chrono::steady_clock::now();
    AssembledSystem assembled = assemble(model);
    const double setup_ms = elapsed_ms(t0);
    CsrMatrix kc = std::move(assembled.stiffness);
    std::vector<double> fc = assembled.load;
    const auto tb = std::chrono::steady_clock::now();
    apply_dirichlet(kc, fc, collect_constraints(model));
    const double boundary_ms = elapsed_ms(tb);
    if (rel_tol > 0.0) {
        double r0 = 0.0;
        for (double v : fc) r0 += v * v;
        r0 = std::sqrt(r0);
        model.settings.tolerance = rel_tol * r0 / std::max(1.0, r0);
    }
    std::vector<double> u;
    const LinearSolverReport rep = solve_iterative(model, kc, fc, u);
    if (argc == 12) {
        std::ofstream out(argv[11], std::ios::binary);
        out.write(reinterpret_cast<const char*>(u.data()), static_cast<std::streamsize>(u.size() * sizeof(double)));
        if (!out) throw std::runtime_error("cannot write displacement file");
    }
    std::cout << std::setprecision(17) << "{\"dofs\": " << kc.rows() << ", \"nnz\": " << kc.nonzeros()
              << ", \"threads_used\": " << assembled.threads_used << ", \"iterations\": " << rep.iterations
              << ", \"pcg_ms\": " << rep.solve_ms << ", \"matvec_ms\": " << rep.matvec_ms
              << ", \"matvec_calls\": " << rep.matvec_calls
              << ", \"preconditioner_apply_ms\": " << rep.preconditioner_apply_ms
              << ", \"element_stiffness_ms\": " << assembled.element_stiffness_ms
              << ", \"assembly_ms\": " << assembled.assembly_ms << ", \"pattern_ms\": " << assembled.pattern_ms
              << ", \"setup_ms\": " << setup_ms << ", \"boundary_ms\": " << boundary_ms
              << ", \"initial_residual\": " << rep.initial_residual << ", \"final_residual\": " << rep.final_residual
              << ", \"converged\": " << (rep.converged ? "true" : "false") << ", \"tolerance\": "
              << model.settings.tolerance << "}" << std::endl;
    return 0;
}

}  // namespace

int main(int argc, char** argv) {
    try {
        if (argc >= 4 && std::string(argv[1]) == "dump") {
            std::vector<std::string> overrides(argv + 4, argv + argc);
            return dump(argv[2], argv[3], overrides);
        }
        if (argc >= 2 && std::string(argv[1]) == "bench") return bench(argc, argv);
        std::cerr << "usage: lfref dump <model.fem> <outdir> [key=value ...] | lfref bench ...\n";
        return 1;
    } catch (const std::exception& error) {
        std::cerr << "ERROR: " << error.what() << "\n";
        return 4;
    }
}
