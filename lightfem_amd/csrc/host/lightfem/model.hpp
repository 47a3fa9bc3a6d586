// LightFEM model surface (same names and fields as include/lightfem/model.hpp of
// the reference, so callers compile unchanged). Two settings are added for the
// gfx950 backend: `precision` (fp64 | mixed) and `stop_rule` (reference | relative).
#pragma once

#include <array>
#include <cstddef>
#include <limits>
#include <optional>
#include <string>
#include <unordered_map>
#include <vector>

namespace lightfem {

enum class PlaneMode { PlaneStress, PlaneStrain };

struct Vec2 {
    double x{0.0};
    double y{0.0};
};

struct Node {
    int id{0};
    Vec2 position{};
    std::size_t dof_x{std::numeric_limits<std::size_t>::max()};
    std::size_t dof_y{std::numeric_limits<std::size_t>::max()};
    std::size_t source_line{0};
};

struct Material {
    int id{0};
    double young_modulus{0.0};
    double poisson_ratio{0.0};
    double thickness{0.0};
    std::size_t source_line{0};
};

struct Q4Element {
    int id{0};
    std::array<int, 4> node_ids{{0, 0, 0, 0}};
    int material_id{0};
    std::size_t source_line{0};
};

struct TrussElement {
    int id{0};
    std::array<int, 2> node_ids{{0, 0}};
    int material_id{0};
    double area{0.0};
    std::size_t source_line{0};
};

struct DirichletBC {
    int node_id{0};
    std::optional<double> ux{};
    std::optional<double> uy{};
    std::size_t source_line{0};
};

struct NodalLoad {
    int node_id{0};
    double fx{0.0};
    double fy{0.0};
    std::size_t source_line{0};
};

struct BodyForce {
    int element_id{0};
    double bx{0.0};
    double by{0.0};
    std::size_t source_line{0};
};

enum class EdgeLoadType { Traction, Pressure };

struct EdgeLoad {
    int element_id{0};
    int edge{0};
    EdgeLoadType type{EdgeLoadType::Traction};
    double first{0.0};
    double second{0.0};
    std::size_t source_line{0};
};

struct SolverSettings {
    PlaneMode plane_mode{PlaneMode::PlaneStress};
    std::string solver{"pcg"};
    double tolerance{1.0e-10};
    std::size_t max_iterations{5000};
    std::string reordering{"none"};
    std::string assembly{"serial"};
    std::size_t threads{0};
    std::string storage{"csr"};
    std::string preconditioner{"jacobi"};
    std::string subdomain_split{"rcb"};
    std::size_t subdomain_nodes{64};
    std::string output_file{"results.vtk"};
    double deformation_scale{1.0};
    std::string csv_prefix{};
    // gfx950 backend
    std::string precision{"fp64"};       // fp64 | mixed (FP32 matrix values, FP64 residual replacement)
    std::string stop_rule{"reference"};  // reference: ||r|| <= tol*max(1,||r0||); relative: ||r|| <= tol*||b||
};

struct ModelSummary {
    std::size_t nodes{0};
    std::size_t elements{0};
    std::size_t materials{0};
    std::size_t boundary_records{0};
    std::size_t loads{0};
    std::size_t body_forces{0};
    std::size_t edge_loads{0};
    std::size_t dofs{0};
    std::size_t trusses{0};
};

class Model {
public:
    std::vector<Node> nodes;
    std::vector<Material> materials;
    std::vector<Q4Element> elements;
    std::vector<TrussElement> trusses;
    std::vector<DirichletBC> boundaries;
    std::vector<NodalLoad> loads;
    std::vector<BodyForce> body_forces;
    std::vector<EdgeLoad> edge_loads;
    SolverSettings settings;
    std::string source_name;

    // Sorts entities by id and numbers node i (in id order) with dofs 2i, 2i+1.
    void sort_and_index();
    // Numbers nodes in the given id order (e.g. RCM); other entities stay id-sorted.
    void index_with_node_order(const std::vector<int>& node_order);

    [[nodiscard]] const Node* find_node(int id) const noexcept;
    [[nodiscard]] const Material* find_material(int id) const noexcept;
    [[nodiscard]] const Q4Element* find_element(int id) const noexcept;
    [[nodiscard]] ModelSummary summary() const noexcept;
    [[nodiscard]] bool is_indexed() const noexcept { return indexed_; }

private:
    std::unordered_map<int, std::size_t> node_index_;
    std::unordered_map<int, std::size_t> material_index_;
    std::unordered_map<int, std::size_t> element_index_;
    bool indexed_{false};

    void rebuild_indices();
};

[[nodiscard]] const char* to_string(PlaneMode mode) noexcept;

}  // namespace lightfem
