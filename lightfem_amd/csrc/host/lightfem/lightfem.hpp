// The rest of the LightFEM library surface on the gfx950 backend: input parser,
// validator, element loads, sparse results, RCM, the solver and the writers.
// Names and signatures follow the reference headers (include/lightfem/*.hpp) so
// existing callers compile and behave the same; the numeric stages run on the
// GPU through include/lightfem_gpu.h.
#pragma once

#include <array>
#include <cstddef>
#include <istream>
#include <stdexcept>
#include <string>
#include <vector>

#include "lightfem/model.hpp"

namespace lightfem {

// ---- input (input_parser.hpp)
class ParseError : public std::runtime_error {
public:
    ParseError(std::string source, std::size_t line, std::string message);
    [[nodiscard]] const std::string& source() const noexcept { return source_; }
    [[nodiscard]] std::size_t line() const noexcept { return line_; }

private:
    std::string source_;
    std::size_t line_{0};
};

class InputParser {
public:
    [[nodiscard]] Model parse_file(const std::string& path) const;
    [[nodiscard]] Model parse(std::istream& input, std::string source_name = "<stream>") const;
};

// ---- validation (validation.hpp)
enum class Severity { Error, Warning };

struct Diagnostic {
    Severity severity{Severity::Error};
    std::string code;
    std::string message;
    std::size_t source_line{0};
};

class ValidationReport {
public:
    void error(std::string code, std::string message, std::size_t source_line = 0);
    void warning(std::string code, std::string message, std::size_t source_line = 0);
    [[nodiscard]] bool ok() const noexcept;
    [[nodiscard]] std::size_t error_count() const noexcept;
    [[nodiscard]] std::size_t warning_count() const noexcept;
    [[nodiscard]] const std::vector<Diagnostic>& diagnostics() const noexcept { return diagnostics_; }

private:
    std::vector<Diagnostic> diagnostics_;
};

class ModelValidator {
public:
    [[nodiscard]] ValidationReport validate(const Model& model) const;
};

[[nodiscard]] const char* to_string(Severity severity) noexcept;

// ---- element (element.hpp)
using ElementMatrix = std::array<double, 64>;

struct ElementField {
    std::array<double, 3> strain{{0.0, 0.0, 0.0}};
    std::array<double, 3> stress{{0.0, 0.0, 0.0}};
    double von_mises{0.0};
};

struct GaussPointField {
    double xi{0.0};
    double eta{0.0};
    ElementField field{};
};

[[nodiscard]] std::array<std::size_t, 8> q4_dof_indices(const Model& model, const Q4Element& element);
[[nodiscard]] std::array<double, 8> q4_body_force(const std::array<Vec2, 4>& coordinates, const Material& material,
                                                  double bx, double by);
[[nodiscard]] std::array<double, 8> q4_edge_traction(const std::array<Vec2, 4>& coordinates, const Material& material,
                                                     int edge, double tx, double ty);
[[nodiscard]] std::array<double, 8> q4_edge_pressure(const std::array<Vec2, 4>& coordinates, const Material& material,
                                                     int edge, double pressure);

// ---- sparse (sparse.hpp): the assembled matrix as returned to callers
class CsrMatrix {
public:
    CsrMatrix() = default;
    CsrMatrix(std::size_t rows, std::vector<std::size_t> offsets, std::vector<std::size_t> columns,
              std::vector<double> values);
    [[nodiscard]] std::size_t rows() const noexcept { return rows_; }
    [[nodiscard]] std::size_t nonzeros() const noexcept { return values_.size(); }
    [[nodiscard]] const std::vector<std::size_t>& row_offsets() const noexcept { return row_offsets_; }
    [[nodiscard]] const std::vector<std::size_t>& column_indices() const noexcept { return column_indices_; }
    [[nodiscard]] const std::vector<double>& values() const noexcept { return values_; }
    [[nodiscard]] std::vector<double>& values() noexcept { return values_; }
    [[nodiscard]] std::size_t value_index(std::size_t row, std::size_t column) const;
    [[nodiscard]] double at(std::size_t row, std::size_t column) const;
    [[nodiscard]] bool contains(std::size_t row, std::size_t column) const;
    [[nodiscard]] std::vector<double> multiply(const std::vector<double>& vector) const;

private:
    std::size_t rows_{0};
    std::vector<std::size_t> row_offsets_{0};
    std::vector<std::size_t> column_indices_;
    std::vector<double> values_;
};

struct CsrProfile {
    std::size_t dimension{0};
    std::size_t nonzeros{0};
    std::size_t bandwidth{0};
    std::size_t profile{0};
};

[[nodiscard]] CsrProfile csr_profile(const CsrMatrix& matrix);

// ---- reordering (reordering.hpp)
[[nodiscard]] std::vector<int> reverse_cuthill_mckee(const Model& model);

// ---- solver (solver.hpp)
struct LinearSolverReport {
    bool converged{false};
    std::size_t iterations{0};
    double initial_residual{0.0};
    double final_residual{0.0};
    std::string message;
    std::string preconditioner{"jacobi"};
    double solve_ms{0.0};
    double matvec_ms{0.0};
    std::size_t matvec_calls{0};
    std::string storage{"csr"};
    std::size_t matrix_bytes_per_matvec{0};
    double storage_convert_ms{0.0};
    double preconditioner_setup_ms{0.0};
    double preconditioner_apply_ms{0.0};
    std::size_t subdomains{0};
    double capture_ratio{0.0};
    // gfx950 backend additions
    std::string precision{"fp64"};
    std::string stop_rule{"reference"};
    double true_residual{0.0};
    std::size_t residual_replacements{0};
};

struct ElementResult {
    int element_id{0};
    ElementField field{};
    std::array<GaussPointField, 4> gauss_points{};
};

struct NodeResult {
    int node_id{0};
    std::array<double, 3> stress{{0.0, 0.0, 0.0}};
    double von_mises{0.0};
};

struct PerformanceReport {
    std::string reordering{"none"};
    CsrProfile before;
    CsrProfile after;
    std::string assembly{"serial"};
    std::size_t requested_threads{0};
    std::size_t threads_used{1};
    std::size_t coloring_count{0};
    double profile_ms{0.0};
    double pattern_ms{0.0};
    double element_stiffness_ms{0.0};
    double assembly_ms{0.0};
    double boundary_ms{0.0};
    double postprocess_ms{0.0};
    double retain_matrix_ms{0.0};
    double total_ms{0.0};
};

struct AnalysisResult {
    CsrMatrix stiffness;
    std::vector<double> load;
    std::vector<double> displacement;
    std::vector<double> reaction;
    std::vector<ElementResult> element_results;
    std::vector<NodeResult> node_results;
    LinearSolverReport solver;
    PerformanceReport performance;
};

class LinearElasticSolver {
public:
    [[nodiscard]] AnalysisResult solve(Model& model) const;
};

// ---- output (vtk_writer.hpp)
void write_vtk(const Model& model, const AnalysisResult& result, const std::string& path);

struct CsvExportPaths {
    std::string nodes;
    std::string gauss_points;
};

[[nodiscard]] CsvExportPaths write_csv_results(const Model& model, const AnalysisResult& result,
                                               const std::string& prefix);

}  // namespace lightfem
