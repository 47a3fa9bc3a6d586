// Consistent Q4 loads assembled on the host (src/element.cpp:114-128, 163-210):
// body force by 2x2 Gauss quadrature of N^T b det J t, uniform edge traction
// split evenly over the edge's end nodes, pressure as an inward normal traction.
// The stiffness and field kernels themselves run on the GPU.
#include <algorithm>
#include <cmath>
#include <stdexcept>
#include <utility>

#include "lightfem/lightfem.hpp"

namespace lightfem {

namespace {

constexpr double kGauss = 0.57735026918962576451;

double jacobian(const std::array<Vec2, 4>& c, double xi, double eta) {
    const double dxi[4] = {-0.25 * (1.0 - eta), +0.25 * (1.0 - eta), +0.25 * (1.0 + eta), -0.25 * (1.0 + eta)};
    const double deta[4] = {-0.25 * (1.0 - xi), -0.25 * (1.0 + xi), +0.25 * (1.0 + xi), +0.25 * (1.0 - xi)};
    double x_xi = 0.0, x_eta = 0.0, y_xi = 0.0, y_eta = 0.0;
    for (int i = 0; i < 4; ++i) {
        x_xi += dxi[i] * c[i].x;
        x_eta += deta[i] * c[i].x;
        y_xi += dxi[i] * c[i].y;
        y_eta += deta[i] * c[i].y;
    }
    const double det = x_xi * y_eta - x_eta * y_xi;
    if (!(det > 1.0e-14) || !std::isfinite(det))
        throw std::runtime_error("Q4 stiffness evaluation received a non-positive Jacobian");
    return det;
}

std::pair<int, int> edge_ends(int edge) {
    if (edge < 1 || edge > 4) throw std::invalid_argument("Q4 edge must be in [1, 4]");
    return {edge - 1, edge % 4};
}

}  // namespace

std::array<std::size_t, 8> q4_dof_indices(const Model& model, const Q4Element& element) {
    if (!model.is_indexed()) throw std::logic_error("model must be indexed before Q4 DOFs are requested");
    std::array<std::size_t, 8> dofs{};
    for (int i = 0; i < 4; ++i) {
        const Node* node = model.find_node(element.node_ids[i]);
        if (!node) throw std::logic_error("validated Q4 element references an unknown node");
        dofs[2 * i] = node->dof_x;
        dofs[2 * i + 1] = node->dof_y;
    }
    return dofs;
}

std::array<double, 8> q4_body_force(const std::array<Vec2, 4>& c, const Material& material, double bx, double by) {
    const double pts[4][2] = {{-kGauss, -kGauss}, {kGauss, -kGauss}, {kGauss, kGauss}, {-kGauss, kGauss}};
    std::array<double, 8> out{};
    for (const auto& p : pts) {
        const double xi = p[0], eta = p[1];
        const double w = jacobian(c, xi, eta) * material.thickness;
        const double shape[4] = {0.25 * (1.0 - xi) * (1.0 - eta), 0.25 * (1.0 + xi) * (1.0 - eta),
                                 0.25 * (1.0 + xi) * (1.0 + eta), 0.25 * (1.0 - xi) * (1.0 + eta)};
        for (int n = 0; n < 4; ++n) {
            out[2 * n] += shape[n] * bx * w;
            out[2 * n + 1] += shape[n] * by * w;
        }
    }
    return out;
}

std::array<double, 8> q4_edge_traction(const std::array<Vec2, 4>& c, const Material& material, int edge, double tx,
                                       double ty) {
    const auto [a, b] = edge_ends(edge);
    const double length = std::hypot(c[b].x - c[a].x, c[b].y - c[a].y);
    if (!(length > 1.0e-14)) throw std::runtime_error("Q4 edge traction received a zero-length edge");
    const double half = 0.5 * length * material.thickness;
    std::array<double, 8> out{};
    out[2 * a] = half * tx;
    out[2 * a + 1] = half * ty;
    out[2 * b] = half * tx;
    out[2 * b + 1] = half * ty;
    return out;
}

std::array<double, 8> q4_edge_pressure(const std::array<Vec2, 4>& c, const Material& material, int edge,
                                       double pressure) {
    const auto [a, b] = edge_ends(edge);
    const double dx = c[b].x - c[a].x;
    const double dy = c[b].y - c[a].y;
    const double length = std::hypot(dx, dy);
    if (!(length > 1.0e-14)) throw std::runtime_error("Q4 edge pressure received a zero-length edge");
    // outward normal of a counter-clockwise edge is (dy, -dx)/L; positive pressure pushes inward
    return q4_edge_traction(c, material, edge, -pressure * dy / length, pressure * dx / length);
}

CsrMatrix::CsrMatrix(std::size_t rows, std::vector<std::size_t> offsets, std::vector<std::size_t> columns,
                     std::vector<double> values)
    : rows_(rows), row_offsets_(std::move(offsets)), column_indices_(std::move(columns)), values_(std::move(values)) {}

std::size_t CsrMatrix::value_index(std::size_t row, std::size_t column) const {
    if (row >= rows_ || column >= rows_) throw std::out_of_range("CSR matrix index is out of range");
    const auto begin = column_indices_.begin() + static_cast<std::ptrdiff_t>(row_offsets_[row]);
    const auto end = column_indices_.begin() + static_cast<std::ptrdiff_t>(row_offsets_[row + 1]);
    const auto it = std::lower_bound(begin, end, column);
    if (it == end || *it != column) throw std::logic_error("requested entry is absent from preallocated CSR pattern");
    return static_cast<std::size_t>(it - column_indices_.begin());
}

double CsrMatrix::at(std::size_t row, std::size_t column) const { return values_[value_index(row, column)]; }

bool CsrMatrix::contains(std::size_t row, std::size_t column) const {
    if (row >= rows_ || column >= rows_) return false;
    const auto begin = column_indices_.begin() + static_cast<std::ptrdiff_t>(row_offsets_[row]);
    const auto end = column_indices_.begin() + static_cast<std::ptrdiff_t>(row_offsets_[row + 1]);
    const auto it = std::lower_bound(begin, end, column);
    return it != end && *it == column;
}

std::vector<double> CsrMatrix::multiply(const std::vector<double>& x) const {
    if (x.size() != rows_) throw std::invalid_argument("CSR multiply vector size mismatch");
    std::vector<double> y(rows_, 0.0);
    for (std::size_t row = 0; row < rows_; ++row) {
        double sum = 0.0;
        for (std::size_t k = row_offsets_[row]; k < row_offsets_[row + 1]; ++k) sum += values_[k] * x[column_indices_[k]];
        y[row] = sum;
    }
    return y;
}

CsrProfile csr_profile(const CsrMatrix& matrix) {
    CsrProfile p;
    p.dimension = matrix.rows();
    p.nonzeros = matrix.nonzeros();
    const auto& offsets = matrix.row_offsets();
    const auto& cols = matrix.column_indices();
    for (std::size_t row = 0; row < matrix.rows(); ++row) {
        std::size_t lowest = row;
        for (std::size_t k = offsets[row]; k < offsets[row + 1]; ++k) {
            const std::size_t c = cols[k];
            p.bandwidth = std::max(p.bandwidth, row > c ? row - c : c - row);
            if (c <= row) lowest = std::min(lowest, c);
        }
        p.profile += row - lowest;
    }
    return p;
}

}  // namespace lightfem
