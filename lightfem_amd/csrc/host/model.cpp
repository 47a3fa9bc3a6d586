// Model storage and dof numbering (behaviour of src/model.cpp:24-138): entities
// sorted by id; node i in the numbering order owns dofs 2i and 2i+1.
#include <algorithm>
#include <stdexcept>
#include <string>

#include "lightfem/model.hpp"

namespace lightfem {

namespace {

template <typename T>
void sort_by_id(std::vector<T>& items) {
    std::sort(items.begin(), items.end(), [](const T& a, const T& b) { return a.id < b.id; });
}

template <typename T>
void require_unique(const std::vector<T>& sorted, const char* what) {
    const auto dup = std::adjacent_find(sorted.begin(), sorted.end(),
                                        [](const T& a, const T& b) { return a.id == b.id; });
    if (dup != sorted.end()) throw std::logic_error(std::string("duplicate ") + what + " ID: " + std::to_string(dup->id));
}

}  // namespace

void Model::sort_and_index() {
    sort_by_id(nodes);
    sort_by_id(materials);
    sort_by_id(elements);
    sort_by_id(trusses);
    require_unique(nodes, "node");
    require_unique(materials, "material");
    require_unique(elements, "element");
    require_unique(trusses, "truss");
    rebuild_indices();
}

void Model::index_with_node_order(const std::vector<int>& node_order) {
    if (node_order.size() != nodes.size()) {
        throw std::invalid_argument("node reordering must contain every model node exactly once");
    }
    std::vector<Node> by_id = nodes;
    sort_by_id(by_id);
    require_unique(by_id, "node");
    sort_by_id(materials);
    sort_by_id(elements);
    sort_by_id(trusses);
    require_unique(materials, "material");
    require_unique(elements, "element");
    require_unique(trusses, "truss");
    std::vector<char> taken(by_id.size(), 0);
    std::vector<Node> reordered;
    reordered.reserve(by_id.size());
    for (const int id : node_order) {
        const auto it = std::lower_bound(by_id.begin(), by_id.end(), id,
                                         [](const Node& n, int value) { return n.id < value; });
        const std::size_t pos = static_cast<std::size_t>(it - by_id.begin());
        if (it == by_id.end() || it->id != id || taken[pos]) {
            throw std::invalid_argument("node reordering references an unknown or duplicate node ID");
        }
        taken[pos] = 1;
        reordered.push_back(*it);
    }
    nodes = std::move(reordered);
    rebuild_indices();
}

void Model::rebuild_indices() {
    node_index_.clear();
    material_index_.clear();
    element_index_.clear();
    node_index_.reserve(nodes.size());
    for (std::size_t i = 0; i < nodes.size(); ++i) {
        nodes[i].dof_x = 2 * i;
        nodes[i].dof_y = 2 * i + 1;
        node_index_.emplace(nodes[i].id, i);
    }
    for (std::size_t i = 0; i < materials.size(); ++i) material_index_.emplace(materials[i].id, i);
    for (std::size_t i = 0; i < elements.size(); ++i) element_index_.emplace(elements[i].id, i);
    indexed_ = true;
}

namespace {
template <typename T>
const T* lookup(bool indexed, const std::unordered_map<int, std::size_t>& index, const std::vector<T>& items, int id) {
    if (indexed) {
        const auto it = index.find(id);
        return it == index.end() ? nullptr : &items[it->second];
    }
    for (const T& item : items)
        if (item.id == id) return &item;
    return nullptr;
}
}  // namespace

const Node* Model::find_node(int id) const noexcept { return lookup(indexed_, node_index_, nodes, id); }
const Material* Model::find_material(int id) const noexcept { return lookup(indexed_, material_index_, materials, id); }
const Q4Element* Model::find_element(int id) const noexcept { return lookup(indexed_, element_index_, elements, id); }

ModelSummary Model::summary() const noexcept {
    ModelSummary s;
    s.nodes = nodes.size();
    s.elements = elements.size();
    s.materials = materials.size();
    s.boundary_records = boundaries.size();
    s.loads = loads.size();
    s.body_forces = body_forces.size();
    s.edge_loads = edge_loads.size();
    s.dofs = 2 * nodes.size();
    s.trusses = trusses.size();
    return s;
}

const char* to_string(PlaneMode mode) noexcept {
    return mode == PlaneMode::PlaneStrain ? "plane_strain" : "plane_stress";
}

}  // namespace lightfem
