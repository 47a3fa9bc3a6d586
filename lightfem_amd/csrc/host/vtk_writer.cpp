// Result export with the reference's file layout (src/vtk_writer.cpp:110-337):
// legacy ASCII VTK unstructured grid (deformed points, nodal and cell fields,
// per-Gauss-point stresses) and the two CSV tables. Numbers are written with
// std::to_chars, the shortest text that round-trips; large tables are formatted in
// parallel chunks (OpenMP) and written in order, so the bytes do not depend on the
// thread count.
#include <algorithm>
#include <charconv>
#include <cmath>
#include <fstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "lightfem/lightfem.hpp"

namespace lightfem {

namespace {

// Appends numbers to a text buffer with std::to_chars.
struct Text {
    std::string s;
    Text& operator<<(const char* text) {
        s += text;
        return *this;
    }
    Text& operator<<(char c) {
        s += c;
        return *this;
    }
    Text& integer(long long v) {
        char tmp[24];
        const auto r = std::to_chars(tmp, tmp + sizeof tmp, v);
        s.append(tmp, r.ptr);
        return *this;
    }
    Text& real(double v) {
        char tmp[40];
        const auto r = std::to_chars(tmp, tmp + sizeof tmp, v);
        if (r.ec != std::errc{}) throw std::runtime_error("failed to format a number");
        s.append(tmp, r.ptr);
        return *this;
    }
};

class Sink : public Text {
public:
    explicit Sink(const std::string& path) : path_(path), out_(path, std::ios::binary) {
        if (!out_) throw std::runtime_error("cannot open output file: " + path);
        s.reserve(kChunk + 256);
    }
    Sink& operator<<(const char* text) {
        Text::operator<<(text);
        return spill();
    }
    Sink& operator<<(char c) {
        Text::operator<<(c);
        return spill();
    }
    Sink& integer(long long v) {
        Text::integer(v);
        return spill();
    }
    Sink& real(double v) {
        Text::real(v);
        return spill();
    }
    // Rows [0, n) formatted by row(i, Text&) in parallel chunks, written in order.
    template <typename Row>
    Sink& rows(std::size_t n, Row&& row) {
        constexpr std::size_t kRows = 8192;
        const std::size_t chunks = (n + kRows - 1) / kRows;
        const std::size_t wave = 256;  // chunks formatted per round (bounded memory)
        std::vector<Text> parts;
        for (std::size_t c0 = 0; c0 < chunks; c0 += wave) {
            const std::size_t c1 = std::min(chunks, c0 + wave);
            parts.assign(c1 - c0, Text{});
            bool failed = false;
#pragma omp parallel for schedule(dynamic, 1)
            for (std::size_t c = c0; c < c1; ++c) {
                Text& t = parts[c - c0];
                try {
                    for (std::size_t i = c * kRows; i < std::min(n, (c + 1) * kRows); ++i) row(i, t);
                } catch (...) {
#pragma omp atomic write
                    failed = true;
                }
            }
            if (failed) throw std::runtime_error("failed to format a number for " + path_);
            flush();
            for (const Text& t : parts) write(t.s);
        }
        return *this;
    }
    void close() {
        flush();
        out_.close();
        if (!out_) throw std::runtime_error("failed while closing " + path_);
    }

private:
    static constexpr std::size_t kChunk = std::size_t{1} << 20;
    Sink& spill() {
        if (s.size() >= kChunk) flush();
        return *this;
    }
    void write(const std::string& text) {
        out_.write(text.data(), static_cast<std::streamsize>(text.size()));
        if (!out_) throw std::runtime_error("failed while writing " + path_);
    }
    void flush() {
        write(s);
        s.clear();
    }
    std::string path_;
    std::ofstream out_;
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
    out.rows(cells.size(), [&](std::size_t i, Text& t) { t.real(get(cells[i])) << '\n'; });
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
    out.rows(model.nodes.size(), [&](std::size_t i, Text& t) {
        const Node& node = model.nodes[i];
        t.real(node.position.x + scale * u[node.dof_x]) << ' ';
        t.real(node.position.y + scale * u[node.dof_y]) << " 0\n";
    });
    out << "CELLS ";
    out.integer(static_cast<long long>(cells)) << ' ';
    out.integer(static_cast<long long>(5 * model.elements.size() + 3 * model.trusses.size())) << '\n';
    for (const Q4Element& e : model.elements)
        for (const int id : e.node_ids)
            if (!model.find_node(id)) throw std::logic_error("validated element references an unknown node");
    out.rows(model.elements.size(), [&](std::size_t i, Text& t) {
        t << '4';
        for (const int id : model.elements[i].node_ids) {
            t << ' ';
            t.integer(static_cast<long long>(model.find_node(id)->dof_x / 2));
        }
        t << '\n';
    });
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
    out.rows(model.elements.size(), [](std::size_t, Text& t) { t << "9\n"; });
    for (std::size_t i = 0; i < model.trusses.size(); ++i) out << "3\n";
    out << "POINT_DATA ";
    out.integer(static_cast<long long>(model.nodes.size())) << "\nVECTORS displacement double\n";
    out.rows(model.nodes.size(), [&](std::size_t i, Text& t) {
        const Node& node = model.nodes[i];
        t.real(u[node.dof_x]) << ' ';
        t.real(u[node.dof_y]) << " 0\n";
    });
    const char* nodal_names[3] = {"nodal_stress_xx", "nodal_stress_yy", "nodal_stress_xy"};
    for (int c = 0; c < 3; ++c) {
        out << "SCALARS " << nodal_names[c] << " double 1\nLOOKUP_TABLE default\n";
        out.rows(result.node_results.size(),
                 [&](std::size_t i, Text& t) { t.real(result.node_results[i].stress[c]) << '\n'; });
    }
    out << "SCALARS nodal_von_mises_in_plane double 1\nLOOKUP_TABLE default\n";
    out.rows(result.node_results.size(), [&](std::size_t i, Text& t) { t.real(result.node_results[i].von_mises) << '\n'; });
    out << "VECTORS reaction double\n";
    out.rows(model.nodes.size(), [&](std::size_t i, Text& t) {
        const Node& node = model.nodes[i];
        t.real(result.reaction[node.dof_x]) << ' ';
        t.real(result.reaction[node.dof_y]) << " 0\n";
    });
    out << "SCALARS displacement_magnitude double 1\nLOOKUP_TABLE default\n";
    out.rows(model.nodes.size(), [&](std::size_t i, Text& t) {
        const Node& node = model.nodes[i];
        t.real(std::hypot(u[node.dof_x], u[node.dof_y])) << '\n';
    });
    out << "CELL_DATA ";
    out.integer(static_cast<long long>(cells)) << "\nSCALARS element_id int 1\nLOOKUP_TABLE default\n";
    out.rows(result.element_results.size(),
             [&](std::size_t i, Text& t) { t.integer(result.element_results[i].element_id) << '\n'; });
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
            out.rows(cellsv.size(), [&](std::size_t i, Text& t) { t.real(cellsv[i].gauss_points[g].field.stress[c]) << '\n'; });
        }
        out << "SCALARS von_mises_gp";
        out.integer(g + 1) << " double 1\nLOOKUP_TABLE default\n";
        out.rows(cellsv.size(), [&](std::size_t i, Text& t) { t.real(cellsv[i].gauss_points[g].field.von_mises) << '\n'; });
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
    nodes.rows(model.nodes.size(), [&](std::size_t i, Text& t) {
        const Node& node = model.nodes[i];
        const NodeResult& nr = result.node_results[i];
        t.integer(node.id) << ',';
        t.real(node.position.x) << ',';
        t.real(node.position.y) << ',';
        t.real(node.position.x + scale * u[node.dof_x]) << ',';
        t.real(node.position.y + scale * u[node.dof_y]) << ',';
        t.real(u[node.dof_x]) << ',';
        t.real(u[node.dof_y]) << ',';
        t.real(result.reaction[node.dof_x]) << ',';
        t.real(result.reaction[node.dof_y]);
        for (int c = 0; c < 3; ++c) {
            t << ',';
            t.real(nr.stress[c]);
        }
        t << ',';
        t.real(nr.von_mises) << '\n';
    });
    nodes.close();
    Sink gps(paths.gauss_points);
    gps << "element_id,gauss_point,xi,eta,strain_xx,strain_yy,strain_xy,stress_xx,stress_yy,stress_xy,von_mises\n";
    gps.rows(result.element_results.size(), [&](std::size_t i, Text& t) {
        const ElementResult& er = result.element_results[i];
        for (int g = 0; g < 4; ++g) {
            const GaussPointField& p = er.gauss_points[g];
            t.integer(er.element_id) << ',';
            t.integer(g + 1) << ',';
            t.real(p.xi) << ',';
            t.real(p.eta);
            for (int c = 0; c < 3; ++c) {
                t << ',';
                t.real(p.field.strain[c]);
            }
            for (int c = 0; c < 3; ++c) {
                t << ',';
                t.real(p.field.stress[c]);
            }
            t << ',';
            t.real(p.field.von_mises) << '\n';
        }
    });
    gps.close();
    return paths;
}

}  // namespace lightfem
