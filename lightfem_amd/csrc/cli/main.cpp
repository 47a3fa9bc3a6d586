// `lightfem` command line on the gfx950 backend: `check` and `solve` with the
// reference's usage text, summary / diagnostics / SOLVED / PROFILE / PARALLEL /
// TIMING / ITERATIVE lines and exit codes (src/main.cpp:19-166): 0 ok, 1 usage,
// 2 parse error, 3 invalid model, 4 any other failure.
#include <chrono>
#include <exception>
#include <filesystem>
#include <iostream>
#include <string>

#include "lightfem/lightfem.hpp"

namespace {

double ms_since(std::chrono::steady_clock::time_point t) {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t).count();
}

void usage() {
    std::cout << "LightFEM 0.8.0 - 2D Q4 + truss linear elastic solver\n\n"
              << "Usage:\n"
              << "  lightfem check <model.fem>\n"
              << "  lightfem solve <model.fem> [result.vtk]\n"
              << "  lightfem --help\n"
              << "  lightfem --version\n\n"
              << "solve uses 2x2 Gauss Q4 stiffness, preallocated CSR assembly,\n"
              << "symmetric Dirichlet elimination and Jacobi-preconditioned CG.\n";
}

void summary(const lightfem::Model& m) {
    const lightfem::ModelSummary s = m.summary();
    std::cout << "Model: " << m.source_name << '\n'
              << "  nodes=" << s.nodes << ", elements=" << s.elements << ", trusses=" << s.trusses
              << ", materials=" << s.materials << ", boundary_records=" << s.boundary_records
              << ", nodal_loads=" << s.loads << ", body_forces=" << s.body_forces << ", edge_loads=" << s.edge_loads
              << '\n'
              << "  plane_mode=" << lightfem::to_string(m.settings.plane_mode) << ", solver=" << m.settings.solver
              << ", tolerance=" << m.settings.tolerance << ", max_iterations=" << m.settings.max_iterations
              << ", reordering=" << m.settings.reordering << ", assembly=" << m.settings.assembly
              << ", threads=" << m.settings.threads << ", deformation_scale=" << m.settings.deformation_scale << '\n';
}

}  // namespace

int main(int argc, char* argv[]) {
    const std::string first = argc > 1 ? argv[1] : "";
    if (argc == 2 && first == "--version") {
        std::cout << "LightFEM 0.8.0\n";
        return 0;
    }
    if (argc == 2 && (first == "--help" || first == "-h")) {
        usage();
        return 0;
    }
    if (argc < 3 || argc > 4 || (first != "check" && first != "solve") || (first == "check" && argc != 3)) {
        usage();
        return 1;
    }
    try {
        const auto process_start = std::chrono::steady_clock::now();
        const auto parse_start = std::chrono::steady_clock::now();
        lightfem::Model model = lightfem::InputParser{}.parse_file(argv[2]);
        const double parse_only_ms = ms_since(parse_start);
        const lightfem::ValidationReport report = lightfem::ModelValidator{}.validate(model);
        const double parse_ms = ms_since(parse_start);
        summary(model);
        for (const lightfem::Diagnostic& d : report.diagnostics()) {
            std::cout << lightfem::to_string(d.severity) << " [" << d.code << "]";
            if (d.source_line != 0) std::cout << " line " << d.source_line;
            std::cout << ": " << d.message << '\n';
        }
        if (!report.ok()) {
            std::cout << "INVALID: " << report.error_count() << " error(s), " << report.warning_count()
                      << " warning(s)\n";
            return 3;
        }
        model.sort_and_index();
        if (first == "check") {
            std::cout << "VALID: " << model.summary().dofs << " displacement DOFs, " << report.warning_count()
                      << " warning(s)\n";
            return 0;
        }
        const std::string output = argc == 4 ? argv[3] : model.settings.output_file;
        const lightfem::AnalysisResult r = lightfem::LinearElasticSolver{}.solve(model);
        const auto io_start = std::chrono::steady_clock::now();
        lightfem::write_vtk(model, r, output);
        const std::string prefix = model.settings.csv_prefix.empty()
                                       ? std::filesystem::path(output).replace_extension().string()
                                       : model.settings.csv_prefix;
        const lightfem::CsvExportPaths csv = lightfem::write_csv_results(model, r, prefix);
        const double export_ms = ms_since(io_start);
        const auto& p = r.performance;
        std::cout << "SOLVED: dofs=" << r.displacement.size() << ", nnz=" << r.stiffness.nonzeros()
                  << ", pcg_iterations=" << r.solver.iterations << ", initial_residual=" << r.solver.initial_residual
                  << ", final_residual=" << r.solver.final_residual << '\n'
                  << "PROFILE: ordering=" << p.reordering << ", before(nnz=" << p.before.nonzeros
                  << ", bandwidth=" << p.before.bandwidth << ", profile=" << p.before.profile << ')'
                  << ", after(nnz=" << p.after.nonzeros << ", bandwidth=" << p.after.bandwidth
                  << ", profile=" << p.after.profile << ")\n"
                  << "PARALLEL: assembly=" << p.assembly << ", requested_threads=" << p.requested_threads
                  << ", threads_used=" << p.threads_used << ", colors=" << p.coloring_count
                  << ", element_stiffness_ms=" << p.element_stiffness_ms << ", assembly_ms=" << p.assembly_ms << '\n'
                  << "VTK: " << output << " (deformation_scale=" << model.settings.deformation_scale << ")\n"
                  << "CSV: " << csv.nodes << ", " << csv.gauss_points << '\n'
                  << "TIMING: parse_only_ms=" << parse_only_ms << ", validate_ms=" << (parse_ms - parse_only_ms)
                  << ", profile_ms=" << p.profile_ms << ", pattern_ms=" << p.pattern_ms
                  << ", element_stiffness_ms=" << p.element_stiffness_ms << ", assembly_ms=" << p.assembly_ms
                  << ", boundary_ms=" << p.boundary_ms << ", pcg_ms=" << r.solver.solve_ms
                  << ", matvec_ms=" << r.solver.matvec_ms << ", matvec_calls=" << r.solver.matvec_calls
                  << ", postprocess_ms=" << p.postprocess_ms << ", retain_matrix_ms=" << p.retain_matrix_ms
                  << ", export_ms=" << export_ms << ", solve_total_ms=" << p.total_ms
                  << ", process_ms=" << ms_since(process_start) << '\n';
        std::cout << "ITERATIVE: storage=" << r.solver.storage << ", preconditioner=" << r.solver.preconditioner
                  << ", subdomains=" << r.solver.subdomains << ", capture_ratio=" << r.solver.capture_ratio
                  << ", preconditioner_setup_ms=" << r.solver.preconditioner_setup_ms
                  << ", preconditioner_apply_ms=" << r.solver.preconditioner_apply_ms
                  << ", storage_convert_ms=" << r.solver.storage_convert_ms << ", matrix_mb_per_matvec="
                  << (static_cast<double>(r.solver.matrix_bytes_per_matvec) / (1024.0 * 1024.0)) << '\n';
        std::cout << "DEVICE: backend=gfx950, precision=" << r.solver.precision << ", stop_rule="
                  << r.solver.stop_rule << ", true_residual=" << r.solver.true_residual
                  << ", residual_replacements=" << r.solver.residual_replacements << '\n';
        return 0;
    } catch (const lightfem::ParseError& error) {
        std::cerr << "PARSE ERROR: " << error.what() << '\n';
        return 2;
    } catch (const std::exception& error) {
        std::cerr << "FATAL ERROR: " << error.what() << '\n';
        return 4;
    }
}
