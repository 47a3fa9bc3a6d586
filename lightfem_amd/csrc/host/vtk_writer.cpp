// Result export with the reference's file layout (src/vtk_writer.cpp:110-337):
// legacy ASCII VTK unstructured grid (deformed points, nodal and cell fields,
// per-Gauss-point stresses) and the two CSV tables. Numbers are written with
// std::to_chars, the shortest text that round-trips, through a buffered sink.
#include <charconv>
#include <cmath>
#include <fstream>
#include <stdexcept>
#include <string>

#include "lightfem/lightfem.hpp"

namespace lightfem {

namespace {

class Sink {
public:
    explicit Sink(const std::string& path) : path_(path), out_(path, std::ios::binary) {
        if (!out_) throw std::runtime_error("cannot open output file: " + path);
        buf_.reserve(kChunk + 256);
    }
    Sink& operator<<(const char* text) {
        buf_ += text;
        return spill();
    }
    Sink& operator<<(char c) {
        buf_ += c;
        return spill();
    }
    Sink& integer(long long v) {
        char tmp[24];
        const auto r = std::to_chars(tmp, tmp + sizeof tmp, v);
        buf_.append(tmp, r.ptr);
        return spill();
    }
    Sink& real(double v) {
        char tmp[40];
        const auto r = std::to_chars(tmp, tmp + sizeof tmp, v);
        if (r.ec != std::errc{}) throw std::runtime_error("failed to format a number for " + path_);
        buf_.append(tmp, r.ptr);
        return spill();
    }
    void close() {
        flush();
        out_.close();
        if (!out_) throw std::runtime_error("failed while closing " + path_);
    }

private:
    static constexpr std::size_t kChunk = std::size_t{1} << 20;
    Sink& spill() {
        if (buf_.size() >= kChunk) flush();
        return *this;
    }
    void flush() {
        out_.write(buf_.data(), static_cast<std::streamsize>(buf_.size()));
        if (!out_) throw std::runtime_error("failed while writing " + path_);
        buf_.clear();
    }
    std::string path_;
    std::ofstream out_;
    std::string buf_;
};

void require_dofs(const Model& model, const AnalysisResult& result, const char* format) {
    const std::size_t dofs = model.summary().dofs;
    if (result.displacement.size() != dofs)
        throw std::invalid_argument(std::string(format) + " displacement size does not match model DOFs");
    if (result.reaction.size() != dofs)
        throw std::invalid_argument(std::string(format) + " reaction size does not match model DOFs");
}

void scalar_block(Sink& out, const char* name, const std::vector<ElementResult>& cells,
                  double (*get)(const ElementResult&)) {
    out << "SCALARS " << name << " double 1\nLOOKUP_TABLE default\n";
    for (const ElementResult& cell : cells) out.real(get(cell)) << '\n';
}

}  // namespace

void write_vtk(const Model& model, const AnalysisResult& result, const std::string& path) {
    if (!model.is_indexed()) throw std::logic_error("model must be indexed before VTK export");
    require_dofs(model, result, "VTK");
    const std::size_t cells = model.elements.size() + model.trusses.size();
    if (result.element_results.size() != cells)
        throw std::invalid_argument("VTK element result count does not match model elements");
    if (result.node_results.size() != model.nodes.size())
        throw std::invalid_argument("VTK nodal result count does not match model nodes");
    const double scale = model.settings.deformation_scale;
    const auto& u = result.displacement;
    Sink out(path);
    out << "# vtk DataFile Version 3.0\nLightFEM linear static Q4 result\nASCII\nDATASET UNSTRUCTURED_GRID\nPOINTS ";
    out.integer(static_cast<long long>(model.nodes.size())) << " double\n";
    for (const Node& node : model.nodes) {
        out.real(node.position.x + scale * u[node.dof_x]) << ' ';
        out.real(node.position.y + scale * u[node.dof_y]) << " 0\n";
    }
    out << "CELLS ";
    out.integer(static_cast<long long>(cells)) << ' ';
    out.integer(static_cast<long long>(5 * model.elements.size() + 3 * model.trusses.size())) << '\n';
    for (const Q4Element& e : model.elements) {
        out << '4';
        for (const int id : e.node_ids) {
            const Node* node = model.find_node(id);
            if (!node) throw std::logic_error("validated element references an unknown node");
            out << ' ';
            out.integer(static_cast<long long>(node->dof_x / 2));
        }
        out << '\n';
    }
    for (const TrussElement& t : model.trusses) {
        out << '2';
        for (const int id : t.node_ids) {
            const Node* node = model.find_node(id);
            if (!node) throw std::logic_error("validated truss references an unknown node");
            out << ' ';
            out.integer(static_cast<long long>(node->dof_x / 2));
        }
        out << '\n';
    }
    out << "CELL_TYPES ";
    out.integer(static_cast<long long>(cells)) << '\n';
    for (std::size_t i = 0; i < model.elements.size(); ++i) out << "9\n";
    for (std::size_t i = 0; i < model.trusses.size(); ++i) out << "3\n";
    out << "POINT_DATA ";
    out.integer(static_cast<long long>(model.nodes.size())) << "\nVECTORS displacement double\n";
    for (const Node& node : model.nodes) {
        out.real(u[node.dof_x]) << ' ';
        out.real(u[node.dof_y]) << " 0\n";
    }
    const char* nodal_names[3] = {"nodal_stress_xx", "nodal_stress_yy", "nodal_stress_xy"};
    for (int c = 0; c < 3; ++c) {
        out << "SCALARS " << nodal_names[c] << " double 1\nLOOKUP_TABLE default\n";
        for (const NodeResult& nr : result.node_results) out.real(nr.stress[c]) << '\n';
    }
    out << "SCALARS nodal_von_mises_in_plane double 1\nLOOKUP_TABLE default\n";
    for (const NodeResult& nr : result.node_results) out.real(nr.von_mises) << '\n';
    out << "VECTORS reaction double\n";
    for (const Node& node : model.nodes) {
        out.real(result.reaction[node.dof_x]) << ' ';
        out.real(result.reaction[node.dof_y]) << " 0\n";
    }
    out << "SCALARS displacement_magnitude double 1\nLOOKUP_TABLE default\n";
    for (const Node& node : model.nodes) out.real(std::hypot(u[node.dof_x], u[node.dof_y])) << '\n';
    out << "CELL_DATA ";
    out.integer(static_cast<long long>(cells)) << "\nSCALARS element_id int 1\nLOOKUP_TABLE default\n";
    for (const ElementResult& er : result.element_results) out.integer(er.element_id) << '\n';
    const auto& cellsv = result.element_results;
    scalar_block(out, "strain_xx", cellsv, [](const ElementResult& e) { return e.field.strain[0]; });
    scalar_block(out, "strain_yy", cellsv, [](const ElementResult& e) { return e.field.strain[1]; });
    scalar_block(out, "strain_xy", cellsv, [](const ElementResult& e) { return e.field.strain[2]; });
    scalar_block(out, "stress_xx", cellsv, [](const ElementResult& e) { return e.field.stress[0]; });
    scalar_block(out, "stress_yy", cellsv, [](const ElementResult& e) { return e.field.stress[1]; });
    scalar_block(out, "stress_xy", cellsv, [](const ElementResult& e) { return e.field.stress[2]; });
    scalar_block(out, "von_mises_in_plane", cellsv, [](const ElementResult& e) { return e.field.von_mises; });
    const char* stress_names[3] = {"stress_xx", "stress_yy", "stress_xy"};
    for (int g = 0; g < 4; ++g) {
        for (int c = 0; c < 3; ++c) {
            out << "SCALARS " << stress_names[c] << "_gp";
            out.integer(g + 1) << " double 1\nLOOKUP_TABLE default\n";
            for (const ElementResult& er : cellsv) out.real(er.gauss_points[g].field.stress[c]) << '\n';
        }
        out << "SCALARS von_mises_gp";
        out.integer(g + 1) << " double 1\nLOOKUP_TABLE default\n";
        for (const ElementResult& er : cellsv) out.real(er.gauss_points[g].field.von_mises) << '\n';
    }
    out.close();
}

CsvExportPaths write_csv_results(const Model& model, const AnalysisResult& result, const std::string& prefix) {
    if (!model.is_indexed()) throw std::logic_error("model must be indexed before CSV export");
    if (prefix.empty()) throw std::invalid_argument("CSV prefix must not be empty");
    require_dofs(model, result, "VTK");
    if (result.node_results.size() != model.nodes.size() ||
        result.element_results.size() != model.elements.size() + model.trusses.size())
        throw std::invalid_argument("CSV result count does not match model");
    CsvExportPaths paths{prefix + "_nodes.csv", prefix + "_gauss_points.csv"};
    const double scale = model.settings.deformation_scale;
    const auto& u = result.displacement;
    Sink nodes(paths.nodes);
    nodes << "node_id,x,y,deformed_x,deformed_y,ux,uy,reaction_x,reaction_y,stress_xx,stress_yy,stress_xy,von_mises\n";
    for (std::size_t i = 0; i < model.nodes.size(); ++i) {
        const Node& node = model.nodes[i];
        const NodeResult& nr = result.node_results[i];
        nodes.integer(node.id) << ',';
        nodes.real(node.position.x) << ',';
        nodes.real(node.position.y) << ',';
        nodes.real(node.position.x + scale * u[node.dof_x]) << ',';
        nodes.real(node.position.y + scale * u[node.dof_y]) << ',';
        nodes.real(u[node.dof_x]) << ',';
        nodes.real(u[node.dof_y]) << ',';
        nodes.real(result.reaction[node.dof_x]) << ',';
        nodes.real(result.reaction[node.dof_y]);
        for (int c = 0; c < 3; ++c) {
            nodes << ',';
            nodes.real(nr.stress[c]);
        }
        nodes << ',';
        nodes.real(nr.von_mises) << '\n';
    }
    nodes.close();
    Sink gps(paths.gauss_points);
    gps << "element_id,gauss_point,xi,eta,strain_xx,strain_yy,strain_xy,stress_xx,stress_yy,stress_xy,von_mises\n";
    for (const ElementResult& er : result.element_results) {
        for (int g = 0; g < 4; ++g) {
            const GaussPointField& p = er.gauss_points[g];
            gps.integer(er.element_id) << ',';
            gps.integer(g + 1) << ',';
            gps.real(p.xi) << ',';
            gps.real(p.eta);
            for (int c = 0; c < 3; ++c) {
                gps << ',';
                gps.real(p.field.strain[c]);
            }
            for (int c = 0; c < 3; ++c) {
                gps << ',';
                gps.real(p.field.stress[c]);
            }
            gps << ',';
            gps.real(p.field.von_mises) << '\n';
        }
    }
    gps.close();
    return paths;
}

}  // namespace lightfem
