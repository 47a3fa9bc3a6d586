// Semantic model checks with the reference's diagnostic codes and texts
// (src/validation.cpp:111-476): ids, coordinates, coincident nodes (x-sorted
// sweep), materials, Q4 connectivity and Jacobian sign at the centre and the
// four Gauss points, trusses, unused entities, single connected mesh, duplicate
// constraints, rigid-body modes (rank of the constraint rows), loads and the
// setting values, including the backend's precision and stop_rule. Ids resolve
// through direct tables, the per-element geometry runs in parallel, and every
// diagnostic is emitted in the reference's order.
#include <algorithm>
#include <array>
#include <cmath>
#include <filesystem>
#include <queue>
#include <set>
#include <sstream>
#include <unordered_map>
#include <unordered_set>

#include "lightfem/lightfem.hpp"

namespace lightfem {

namespace {

// id -> position of its first record: a direct table for positive ids of moderate
// range (the usual numbering), a hash map otherwise.
class IdIndex {
public:
    template <typename T>
    explicit IdIndex(const std::vector<T>& items) {
        long long max_id = 0;
        bool positive = true;
        for (const T& item : items) {
            positive = positive && item.id > 0;
            max_id = std::max<long long>(max_id, item.id);
        }
        dense_ = positive && max_id <= 4LL * static_cast<long long>(items.size()) + 1024;
        if (dense_) {
            table_.assign(static_cast<std::size_t>(max_id) + 1, -1);
            for (std::size_t i = 0; i < items.size(); ++i)
                if (table_[items[i].id] < 0) table_[items[i].id] = static_cast<long long>(i);
        } else {
            map_.reserve(items.size());
            for (std::size_t i = 0; i < items.size(); ++i) map_.emplace(items[i].id, static_cast<long long>(i));
        }
    }
    long long find(int id) const {
        if (dense_) return (id > 0 && static_cast<std::size_t>(id) < table_.size()) ? table_[id] : -1;
        const auto it = map_.find(id);
        return it == map_.end() ? -1 : it->second;
    }

private:
    bool dense_ = false;
    std::vector<long long> table_;
    std::unordered_map<int, long long> map_;
};

// A record is a duplicate when an earlier record has its id (validation.cpp:111-130).
template <typename T>
void check_ids(const std::vector<T>& items, const IdIndex& index, const char* what, const char* duplicate_code,
               ValidationReport& r) {
    for (std::size_t i = 0; i < items.size(); ++i) {
        const T& item = items[i];
        if (item.id <= 0)
            r.error("E_NONPOSITIVE_ID", std::string(what) + " ID must be positive: " + std::to_string(item.id),
                    item.source_line);
        if (index.find(item.id) != static_cast<long long>(i))
            r.error(duplicate_code, std::string("duplicate ") + what + " ID: " + std::to_string(item.id),
                    item.source_line);
    }
}

// Whether any two nodes lie within 1e-14 in both x and y. In x order, such a pair is
// in one run of nodes whose consecutive x differ by at most 1e-14, and then the
// pair's y values are at most 1e-14 apart, so some two y-neighbours of that run are
// too. False means the reference's pairwise sweep would report nothing; true sends
// the model through that sweep, which reports the pairs in the reference's order.
bool may_have_coincident_nodes(const std::vector<Node>& nodes) {
    const std::size_t n = nodes.size();
    std::vector<std::pair<double, double>> xy(n);
    for (std::size_t i = 0; i < n; ++i) xy[i] = {nodes[i].position.x, nodes[i].position.y};
    for (const auto& p : xy)
        if (!std::isfinite(p.first) || !std::isfinite(p.second)) return true;  // the reference's sweep decides
    std::sort(xy.begin(), xy.end());
    std::vector<std::size_t> run_start;
    for (std::size_t i = 0; i < n; ++i)
        if (i == 0 || !(xy[i].first - xy[i - 1].first <= 1.0e-14)) run_start.push_back(i);
    run_start.push_back(n);
    int found = 0;
    const std::size_t runs = run_start.size() - 1;
#pragma omp parallel for schedule(dynamic, 64) reduction(| : found)
    for (std::size_t k = 0; k < runs; ++k) {
        const std::size_t a = run_start[k], b = run_start[k + 1];
        if (b - a < 2) continue;
        std::vector<double> ys(b - a);
        for (std::size_t i = a; i < b; ++i) ys[i - a] = xy[i].second;
        std::sort(ys.begin(), ys.end());
        for (std::size_t i = 1; i < ys.size(); ++i)
            if (ys[i] - ys[i - 1] <= 1.0e-14) {
                found |= 1;
                break;
            }
    }
    return found != 0;
}

// Connected components over node positions.
struct UnionFind {
    std::vector<long long> parent;
    explicit UnionFind(std::size_t n) : parent(n) {
        for (std::size_t i = 0; i < n; ++i) parent[i] = static_cast<long long>(i);
    }
    long long root(long long a) {
        while (parent[a] != a) {
            parent[a] = parent[parent[a]];
            a = parent[a];
        }
        return a;
    }
    void join(long long a, long long b) {
        a = root(a);
        b = root(b);
        if (a != b) parent[std::max(a, b)] = std::min(a, b);
    }
};

// det J of the bilinear map at (xi, eta); the 1/4 factors are applied per term.
double jacobian_det(const std::array<const Node*, 4>& n, double xi, double eta) {
    const double sxi[4] = {-(1.0 - eta), 1.0 - eta, 1.0 + eta, -(1.0 + eta)};
    const double seta[4] = {-(1.0 - xi), -(1.0 + xi), 1.0 + xi, 1.0 - xi};
    double a = 0.0, b = 0.0, c = 0.0, d = 0.0;  // dx/dxi, dy/dxi, dx/deta, dy/deta
    for (int i = 0; i < 4; ++i) {
        a += 0.25 * sxi[i] * n[i]->position.x;
        b += 0.25 * sxi[i] * n[i]->position.y;
        c += 0.25 * seta[i] * n[i]->position.x;
        d += 0.25 * seta[i] * n[i]->position.y;
    }
    return a * d - c * b;
}

// Rank of the 3-column rigid-mode constraint rows by Gauss-Jordan elimination
// with partial pivoting (tolerance 1e-12).
int constraint_rank(std::vector<std::array<double, 3>> rows) {
    int rank = 0;
    const int m = static_cast<int>(rows.size());
    for (int col = 0; col < 3 && rank < m; ++col) {
        int pivot = rank;
        for (int i = rank + 1; i < m; ++i)
            if (std::abs(rows[i][col]) > std::abs(rows[pivot][col])) pivot = i;
        if (std::abs(rows[pivot][col]) <= 1.0e-12) continue;
        std::swap(rows[pivot], rows[rank]);
        const double s = rows[rank][col];
        for (int j = col; j < 3; ++j) rows[rank][j] /= s;
        for (int i = 0; i < m; ++i) {
            if (i == rank) continue;
            const double factor = rows[i][col];
            for (int j = col; j < 3; ++j) rows[i][j] -= factor * rows[rank][j];
        }
        ++rank;
    }
    return rank;
}

bool one_of(const std::string& v, std::initializer_list<const char*> options) {
    for (const char* o : options)
        if (v == o) return true;
    return false;
}

}  // namespace

void ValidationReport::error(std::string code, std::string message, std::size_t line) {
    diagnostics_.push_back(Diagnostic{Severity::Error, std::move(code), std::move(message), line});
}

void ValidationReport::warning(std::string code, std::string message, std::size_t line) {
    diagnostics_.push_back(Diagnostic{Severity::Warning, std::move(code), std::move(message), line});
}

bool ValidationReport::ok() const noexcept { return error_count() == 0; }

std::size_t ValidationReport::error_count() const noexcept {
    return static_cast<std::size_t>(std::count_if(diagnostics_.begin(), diagnostics_.end(),
                                                  [](const Diagnostic& d) { return d.severity == Severity::Error; }));
}

std::size_t ValidationReport::warning_count() const noexcept {
    return diagnostics_.size() - error_count();
}

const char* to_string(Severity severity) noexcept { return severity == Severity::Error ? "ERROR" : "WARNING"; }

ValidationReport ModelValidator::validate(const Model& model) const {
    ValidationReport r;
    if (model.nodes.size() < 2) r.error("E_TOO_FEW_NODES", "at least two nodes are required");
    if (model.materials.empty()) r.error("E_NO_MATERIAL", "at least one material is required");
    if (model.elements.empty() && model.trusses.empty())
        r.error("E_NO_ELEMENT", "at least one Q4 element or truss member is required");
    const IdIndex node_index(model.nodes), material_index(model.materials), element_index(model.elements),
        truss_index(model.trusses);
    check_ids(model.nodes, node_index, "node", "E_DUPLICATE_NODE_ID", r);
    check_ids(model.materials, material_index, "material", "E_DUPLICATE_MATERIAL_ID", r);
    check_ids(model.elements, element_index, "element", "E_DUPLICATE_ELEMENT_ID", r);
    check_ids(model.trusses, truss_index, "truss", "E_DUPLICATE_TRUSS_ID", r);

    for (const Node& node : model.nodes)
        if (!std::isfinite(node.position.x) || !std::isfinite(node.position.y))
            r.error("E_INVALID_COORDINATE", "node coordinates must be finite", node.source_line);
    if (may_have_coincident_nodes(model.nodes)) {  // coincident nodes: sweep in x order within the x tolerance
        std::vector<std::size_t> order(model.nodes.size());
        for (std::size_t i = 0; i < order.size(); ++i) order[i] = i;
        std::sort(order.begin(), order.end(), [&](std::size_t a, std::size_t b) {
            return model.nodes[a].position.x < model.nodes[b].position.x;
        });
        for (std::size_t i = 0; i < order.size(); ++i) {
            const Node& left = model.nodes[order[i]];
            for (std::size_t j = i + 1; j < order.size(); ++j) {
                const Node& right = model.nodes[order[j]];
                if (right.position.x - left.position.x > 1.0e-14) break;
                if (std::abs(left.position.y - right.position.y) <= 1.0e-14)
                    r.error("E_COINCIDENT_NODES", "nodes " + std::to_string(left.id) + " and " +
                                                      std::to_string(right.id) + " have coincident coordinates",
                            right.source_line);
            }
        }
    }
    for (const Material& m : model.materials) {
        if (!std::isfinite(m.young_modulus) || m.young_modulus <= 0.0)
            r.error("E_INVALID_YOUNG_MODULUS", "Young's modulus must be finite and positive", m.source_line);
        if (!std::isfinite(m.poisson_ratio) || m.poisson_ratio <= -1.0 || m.poisson_ratio >= 0.5)
            r.error("E_INVALID_POISSON_RATIO", "Poisson's ratio must satisfy -1 < nu < 0.5", m.source_line);
        if (!std::isfinite(m.thickness) || m.thickness <= 0.0)
            r.error("E_INVALID_THICKNESS", "material thickness must be finite and positive", m.source_line);
    }

    // used flags per first record of an id; components by union-find over those records
    std::vector<char> node_used(model.nodes.size(), 0), material_used(model.materials.size(), 0);
    UnionFind components_of(model.nodes.size());
    const double g = 0.57735026918962576451;
    const double points[5][2] = {{0.0, 0.0}, {-g, -g}, {g, -g}, {g, g}, {-g, g}};
    // per-element facts in parallel, diagnostics and bookkeeping in element order
    const std::size_t n_elems = model.elements.size();
    std::vector<std::array<long long, 4>> corner_of(n_elems);
    std::vector<unsigned char> facts(n_elems);  // bit 0: repeated node, bit 1: bad Jacobian
#pragma omp parallel for schedule(static)
    for (std::size_t k = 0; k < n_elems; ++k) {
        const Q4Element& e = model.elements[k];
        const auto& id = e.node_ids;
        const bool repeated = id[0] == id[1] || id[0] == id[2] || id[0] == id[3] || id[1] == id[2] ||
                              id[1] == id[3] || id[2] == id[3];
        bool known = true;
        std::array<const Node*, 4> corner{{nullptr, nullptr, nullptr, nullptr}};
        for (int i = 0; i < 4; ++i) {
            corner_of[k][i] = node_index.find(id[i]);
            if (corner_of[k][i] < 0) known = false;
            else corner[i] = &model.nodes[static_cast<std::size_t>(corner_of[k][i])];
        }
        bool bad_jacobian = false;
        if (known && !repeated)
            for (const auto& p : points)
                if (!(jacobian_det(corner, p[0], p[1]) > 1.0e-14)) {
                    bad_jacobian = true;
                    break;
                }
        facts[k] = static_cast<unsigned char>((repeated ? 1 : 0) | (bad_jacobian ? 2 : 0));
    }
    for (std::size_t k = 0; k < n_elems; ++k) {
        const Q4Element& e = model.elements[k];
        const bool repeated = facts[k] & 1;
        if (repeated) {
            std::ostringstream nodes;
            nodes << '[' << e.node_ids[0] << ", " << e.node_ids[1] << ", " << e.node_ids[2] << ", " << e.node_ids[3]
                  << ']';
            r.error("E_REPEATED_ELEMENT_NODE",
                    "Q4 element " + std::to_string(e.id) + " must reference four distinct nodes: " + nodes.str(),
                    e.source_line);
        }
        bool known = true;
        for (int i = 0; i < 4; ++i) {
            if (corner_of[k][i] < 0) {
                r.error("E_UNKNOWN_ELEMENT_NODE", "element " + std::to_string(e.id) + " references unknown node " +
                                                      std::to_string(e.node_ids[i]),
                        e.source_line);
                known = false;
            } else {
                node_used[static_cast<std::size_t>(corner_of[k][i])] = 1;
            }
        }
        const long long mat = material_index.find(e.material_id);
        if (mat < 0) {
            r.error("E_UNKNOWN_MATERIAL", "element " + std::to_string(e.id) + " references unknown material " +
                                              std::to_string(e.material_id),
                    e.source_line);
        } else {
            material_used[static_cast<std::size_t>(mat)] = 1;
        }
        if (known && !repeated) {
            if (facts[k] & 2)
                r.error("E_NONPOSITIVE_JACOBIAN",
                        "element " + std::to_string(e.id) +
                            " has non-positive or near-zero Jacobian; use counter-clockwise, "
                            "non-self-intersecting Q4 connectivity",
                        e.source_line);
            for (int i = 0; i < 4; ++i) components_of.join(corner_of[k][i], corner_of[k][(i + 1) % 4]);
        }
    }
    for (const TrussElement& t : model.trusses) {
        if (t.node_ids[0] == t.node_ids[1])
            r.error("E_REPEATED_TRUSS_NODE", "truss " + std::to_string(t.id) + " must reference two distinct nodes: [" +
                                                 std::to_string(t.node_ids[0]) + ", " + std::to_string(t.node_ids[1]) +
                                                 "]",
                    t.source_line);
        std::array<long long, 2> ends{{-1, -1}};
        bool known = true;
        for (int i = 0; i < 2; ++i) {
            ends[i] = node_index.find(t.node_ids[i]);
            if (ends[i] < 0) {
                r.error("E_UNKNOWN_TRUSS_NODE", "truss " + std::to_string(t.id) + " references unknown node " +
                                                    std::to_string(t.node_ids[i]),
                        t.source_line);
                known = false;
            } else {
                node_used[static_cast<std::size_t>(ends[i])] = 1;
            }
        }
        const long long mat = material_index.find(t.material_id);
        if (mat < 0) {
            r.error("E_UNKNOWN_MATERIAL", "truss " + std::to_string(t.id) + " references unknown material " +
                                              std::to_string(t.material_id),
                    t.source_line);
        } else {
            material_used[static_cast<std::size_t>(mat)] = 1;
        }
        if (!std::isfinite(t.area) || t.area <= 0.0)
            r.error("E_INVALID_TRUSS_AREA", "truss cross-section area must be finite and positive", t.source_line);
        if (known && t.node_ids[0] != t.node_ids[1]) {
            const Node& a = model.nodes[static_cast<std::size_t>(ends[0])];
            const Node& b = model.nodes[static_cast<std::size_t>(ends[1])];
            const double dx = b.position.x - a.position.x;
            const double dy = b.position.y - a.position.y;
            if (!(std::sqrt(dx * dx + dy * dy) > 1.0e-12)) {
                r.error("E_ZERO_LENGTH_TRUSS", "truss " + std::to_string(t.id) + " has zero length", t.source_line);
            } else {
                components_of.join(ends[0], ends[1]);
            }
        }
    }
    for (const Node& node : model.nodes)
        if (!node_used[static_cast<std::size_t>(node_index.find(node.id))])
            r.warning("W_UNUSED_NODE", "node " + std::to_string(node.id) + " is not used by any element",
                      node.source_line);
    for (const Material& m : model.materials)
        if (!material_used[static_cast<std::size_t>(material_index.find(m.id))])
            r.warning("W_UNUSED_MATERIAL", "material " + std::to_string(m.id) + " is not used by any element",
                      m.source_line);
    {  // connected components over element edges, among the used nodes
        std::size_t components = 0;
        for (std::size_t i = 0; i < model.nodes.size(); ++i)
            if (node_used[i] && components_of.root(static_cast<long long>(i)) == static_cast<long long>(i)) ++components;
        if (components > 1)
            r.error("E_DISCONNECTED_MESH",
                    "the current solver scope requires one connected mesh; found " + std::to_string(components) +
                        " components");
    }

    std::set<std::pair<int, char>> constrained;
    std::vector<std::array<double, 3>> rigid;
    bool prescribed_motion = false;
    for (const DirichletBC& bc : model.boundaries) {
        const long long at = node_index.find(bc.node_id);
        if (at < 0) {
            r.error("E_UNKNOWN_BOUNDARY_NODE", "boundary condition references unknown node " + std::to_string(bc.node_id),
                    bc.source_line);
            continue;
        }
        if (!bc.ux && !bc.uy) r.warning("W_EMPTY_BOUNDARY", "boundary record constrains neither ux nor uy", bc.source_line);
        const Vec2 p = model.nodes[static_cast<std::size_t>(at)].position;
        if (bc.ux) {
            if (!constrained.emplace(bc.node_id, 'x').second)
                r.error("E_DUPLICATE_CONSTRAINT", "ux is constrained more than once at node " + std::to_string(bc.node_id),
                        bc.source_line);
            rigid.push_back({{1.0, 0.0, -p.y}});
            prescribed_motion |= std::abs(*bc.ux) > 0.0;
        }
        if (bc.uy) {
            if (!constrained.emplace(bc.node_id, 'y').second)
                r.error("E_DUPLICATE_CONSTRAINT", "uy is constrained more than once at node " + std::to_string(bc.node_id),
                        bc.source_line);
            rigid.push_back({{0.0, 1.0, p.x}});
            prescribed_motion |= std::abs(*bc.uy) > 0.0;
        }
    }
    if (constraint_rank(rigid) < 3)
        r.error("E_RIGID_BODY_MODES", "displacement constraints do not remove all three 2D rigid-body modes");

    std::unordered_map<int, std::size_t> loads_at;
    bool excited = false;
    for (const NodalLoad& load : model.loads) {
        if (node_index.find(load.node_id) < 0) {
            r.error("E_UNKNOWN_LOAD_NODE", "load references unknown node " + std::to_string(load.node_id),
                    load.source_line);
            continue;
        }
        if (++loads_at[load.node_id] > 1)
            r.warning("W_REPEATED_LOAD_NODE", "multiple load records at node " + std::to_string(load.node_id) +
                                                  " will be accumulated by the solver",
                      load.source_line);
        excited |= std::abs(load.fx) > 0.0 || std::abs(load.fy) > 0.0;
        if (std::abs(load.fx) > 0.0 && constrained.count({load.node_id, 'x'}))
            r.warning("W_LOAD_ON_CONSTRAINED_DOF",
                      "nonzero fx is applied to constrained ux at node " + std::to_string(load.node_id), load.source_line);
        if (std::abs(load.fy) > 0.0 && constrained.count({load.node_id, 'y'}))
            r.warning("W_LOAD_ON_CONSTRAINED_DOF",
                      "nonzero fy is applied to constrained uy at node " + std::to_string(load.node_id), load.source_line);
    }
    for (const BodyForce& f : model.body_forces) {
        if (element_index.find(f.element_id) < 0) {
            r.error("E_UNKNOWN_BODY_FORCE_ELEMENT", "body force references unknown element " + std::to_string(f.element_id),
                    f.source_line);
            continue;
        }
        excited |= std::abs(f.bx) > 0.0 || std::abs(f.by) > 0.0;
    }
    for (const EdgeLoad& load : model.edge_loads) {
        if (element_index.find(load.element_id) < 0) {
            r.error("E_UNKNOWN_EDGE_LOAD_ELEMENT", "edge load references unknown element " +
                                                       std::to_string(load.element_id),
                    load.source_line);
            continue;
        }
        if (load.edge < 1 || load.edge > 4)
            r.error("E_INVALID_EDGE_LOAD_EDGE", "Q4 edge load edge must be in [1, 4]", load.source_line);
        excited |= std::abs(load.first) > 0.0 || std::abs(load.second) > 0.0;
    }
    if (!excited && !prescribed_motion)
        r.warning("W_NO_EXCITATION", "model has no nonzero load or prescribed displacement");

    const SolverSettings& s = model.settings;
    if (!std::isfinite(s.tolerance) || s.tolerance <= 0.0 || s.tolerance >= 1.0)
        r.error("E_INVALID_TOLERANCE", "solver tolerance must satisfy 0 < tolerance < 1");
    if (s.max_iterations == 0) r.error("E_INVALID_MAX_ITERATIONS", "max_iterations must be greater than zero");
    if (!std::isfinite(s.deformation_scale) || s.deformation_scale < 0.0)
        r.error("E_INVALID_DEFORMATION_SCALE", "deformation_scale must be finite and non-negative");
    if (s.solver != "pcg") r.error("E_UNSUPPORTED_SOLVER", "only solver=pcg is supported in the current project scope");
    if (!one_of(s.reordering, {"none", "rcm"})) r.error("E_UNSUPPORTED_REORDERING", "reordering must be none or rcm");
    if (!one_of(s.assembly, {"serial", "atomic", "coloring", "thread_local", "scan"}))
        r.error("E_UNSUPPORTED_ASSEMBLY", "assembly must be serial, atomic, coloring, thread_local, or scan");
    if (!one_of(s.storage, {"csr", "bsr"})) r.error("E_UNSUPPORTED_STORAGE", "storage must be csr or bsr");
    if (!one_of(s.preconditioner, {"jacobi", "block", "ic0"}))
        r.error("E_UNSUPPORTED_PRECONDITIONER", "preconditioner must be jacobi, block, or ic0");
    if (!one_of(s.subdomain_split, {"rcb", "chunk"}))
        r.error("E_UNSUPPORTED_SUBDOMAIN_SPLIT", "subdomain_split must be rcb or chunk");
    if (s.subdomain_nodes == 0) r.error("E_INVALID_SUBDOMAIN_SIZE", "subdomain_nodes must be at least 1");
    if (!one_of(s.precision, {"fp64", "mixed"})) r.error("E_UNSUPPORTED_PRECISION", "precision must be fp64 or mixed");
    if (!one_of(s.stop_rule, {"reference", "relative"}))
        r.error("E_UNSUPPORTED_STOP_RULE", "stop_rule must be reference or relative");
    if (s.output_file.empty()) {
        r.error("E_EMPTY_OUTPUT", "output file name must not be empty");
    } else if (std::filesystem::path(s.output_file).extension() != ".vtk") {
        r.warning("W_OUTPUT_EXTENSION", "VTK output file normally uses the .vtk extension");
    }
    return r;
}

}  // namespace lightfem
