// Host stages that produce the integer structures of the path: CSR pattern and
// element->value map, RCM ordering, element colouring and the CSR profile.
//
// Each result is identical to the reference stage it replaces. The pattern is
// built from node adjacency instead of one hash set per dof row: both dofs of a
// node couple to exactly the dofs of the nodes sharing an element with it, so a
// sorted, unique neighbour-node list per node yields every row in ascending
// column order, and a map entry is a short binary search in that list. This is
// O(nnz) with OpenMP over nodes, where the reference's unordered_set build takes
// seconds at a million elements.
#include "lightfem_host.h"

#include "../device/lfg_ic0_plan.hpp"

#include <algorithm>
#include <cstdint>
#include <queue>
#include <vector>

namespace {

// Q4 elements (4 node indices each) followed by bars (2 each), the element order of
// solver.cpp:59-83: element e < n_q4 is a Q4, e >= n_q4 is bar e - n_q4.
struct Elems {
    int64_t n_q4 = 0, n_bars = 0;
    const int32_t* q4 = nullptr;
    const int32_t* bars = nullptr;
    int64_t count() const { return n_q4 + n_bars; }
    int nodes(int64_t e) const { return e < n_q4 ? 4 : 2; }
    int node(int64_t e, int l) const { return e < n_q4 ? q4[4 * e + l] : bars[2 * (e - n_q4) + l]; }
};

struct NodeElements {
    std::vector<int> ptr;
    std::vector<int> elems;  // ascending element index per node
};

NodeElements node_elements(int64_t n_nodes, const Elems& el) {
    NodeElements ne;
    ne.ptr.assign(static_cast<size_t>(n_nodes) + 1, 0);
    for (int64_t e = 0; e < el.count(); ++e)
        for (int l = 0; l < el.nodes(e); ++l) ne.ptr[static_cast<size_t>(el.node(e, l)) + 1]++;
    for (int64_t i = 0; i < n_nodes; ++i) ne.ptr[i + 1] += ne.ptr[i];
    ne.elems.resize(static_cast<size_t>(ne.ptr[n_nodes]));
    std::vector<int> cur(ne.ptr.begin(), ne.ptr.end() - 1);
    for (int64_t e = 0; e < el.count(); ++e)
        for (int l = 0; l < el.nodes(e); ++l) ne.elems[cur[el.node(e, l)]++] = static_cast<int>(e);
    return ne;
}

bool valid_elements(int64_t n_nodes, const Elems& el) {
    if (n_nodes <= 0 || el.n_q4 < 0 || el.n_bars < 0 || (el.n_q4 > 0 && !el.q4) || (el.n_bars > 0 && !el.bars))
        return false;
    for (int64_t e = 0; e < el.count(); ++e)
        for (int l = 0; l < el.nodes(e); ++l)
            if (el.node(e, l) < 0 || el.node(e, l) >= n_nodes) return false;
    return true;
}

// Sorted unique neighbour nodes (including the node itself when it has elements).
void neighbour_lists(int64_t n_nodes, const Elems& el, const NodeElements& ne, std::vector<int>& ptr,
                     std::vector<int>& list) {
    ptr.assign(static_cast<size_t>(n_nodes) + 1, 0);
    std::vector<std::vector<int>> tmp(static_cast<size_t>(n_nodes));
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n_nodes; ++i) {
        std::vector<int>& v = tmp[static_cast<size_t>(i)];
        for (int j = ne.ptr[i]; j < ne.ptr[i + 1]; ++j) {
            const int e = ne.elems[j];
            for (int l = 0; l < el.nodes(e); ++l) v.push_back(el.node(e, l));
        }
        std::sort(v.begin(), v.end());
        v.erase(std::unique(v.begin(), v.end()), v.end());
    }
    for (int64_t i = 0; i < n_nodes; ++i) ptr[i + 1] = ptr[i] + static_cast<int>(tmp[static_cast<size_t>(i)].size());
    list.resize(static_cast<size_t>(ptr[n_nodes]));
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n_nodes; ++i)
        std::copy(tmp[static_cast<size_t>(i)].begin(), tmp[static_cast<size_t>(i)].end(), list.begin() + ptr[i]);
}

int64_t row_length(const std::vector<int>& nptr, int64_t node) {
    const int64_t m = nptr[node + 1] - nptr[node];
    return m == 0 ? 1 : 2 * m;  // an unused node keeps only its diagonal
}

// The map of one element's (2n)^2 entries (row-major over its dofs) into the pattern.
void element_map(const Elems& el, int64_t e, const std::vector<int>& nptr, const std::vector<int>& nlist,
                 const int32_t* row_ptr, int32_t* out) {
    const int n = el.nodes(e), d = 2 * n;
    for (int r = 0; r < d; ++r) {
        const int a = el.node(e, r / 2);
        const int* begin = nlist.data() + nptr[a];
        const int* end = nlist.data() + nptr[a + 1];
        const int32_t row_start = row_ptr[2 * a + (r & 1)];
        for (int c = 0; c < d; ++c) {
            const int b = el.node(e, c / 2);
            const int pos = static_cast<int>(std::lower_bound(begin, end, b) - begin);
            out[d * r + c] = row_start + 2 * pos + (c & 1);
        }
    }
}

int64_t pattern_nnz(int64_t n_nodes, const Elems& el) {
    if (!valid_elements(n_nodes, el)) return -1;
    const NodeElements ne = node_elements(n_nodes, el);
    std::vector<int> nptr, nlist;
    neighbour_lists(n_nodes, el, ne, nptr, nlist);
    int64_t nnz = 0;
    for (int64_t i = 0; i < n_nodes; ++i) nnz += 2 * row_length(nptr, i);
    return nnz;
}

int pattern(int64_t n_nodes, const Elems& el, int32_t* row_ptr, int32_t* col, int32_t* emap, int32_t* bmap) {
    if (!valid_elements(n_nodes, el) || !row_ptr || !col || (el.n_q4 > 0 && !emap) || (el.n_bars > 0 && !bmap))
        return 1;
    const NodeElements ne = node_elements(n_nodes, el);
    std::vector<int> nptr, nlist;
    neighbour_lists(n_nodes, el, ne, nptr, nlist);
    int64_t total = 0;
    row_ptr[0] = 0;
    for (int64_t i = 0; i < n_nodes; ++i) {
        const int64_t len = row_length(nptr, i);
        if (total + 2 * len > INT32_MAX) return 2;
        row_ptr[2 * i + 1] = static_cast<int32_t>(total + len);
        row_ptr[2 * i + 2] = static_cast<int32_t>(total + 2 * len);
        total += 2 * len;
    }
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n_nodes; ++i) {
        for (int comp = 0; comp < 2; ++comp) {
            int32_t* out = col + row_ptr[2 * i + comp];
            if (nptr[i + 1] == nptr[i]) {
                out[0] = static_cast<int32_t>(2 * i + comp);
                continue;
            }
            for (int k = nptr[i]; k < nptr[i + 1]; ++k) {
                *out++ = 2 * nlist[k];
                *out++ = 2 * nlist[k] + 1;
            }
        }
    }
#pragma omp parallel for schedule(static)
    for (int64_t e = 0; e < el.count(); ++e)
        element_map(el, e, nptr, nlist, row_ptr, e < el.n_q4 ? emap + 64 * e : bmap + 16 * (e - el.n_q4));
    return 0;
}

int rcm(int64_t n_nodes, const int32_t* node_ids, const Elems& id_el, int32_t* order_out) {
    if (n_nodes <= 0 || !node_ids || !order_out || (id_el.n_q4 > 0 && !id_el.q4) ||
        (id_el.n_bars > 0 && !id_el.bars))
        return 1;
    // node id -> index in node_ids
    std::vector<int> by_id(static_cast<size_t>(n_nodes));
    for (int64_t i = 0; i < n_nodes; ++i) by_id[static_cast<size_t>(i)] = static_cast<int>(i);
    std::sort(by_id.begin(), by_id.end(), [&](int a, int b) { return node_ids[a] < node_ids[b]; });
    std::vector<int> sorted_ids(static_cast<size_t>(n_nodes));
    for (int64_t i = 0; i < n_nodes; ++i) sorted_ids[static_cast<size_t>(i)] = node_ids[by_id[static_cast<size_t>(i)]];
    auto index_of = [&](int id) -> int {
        const auto it = std::lower_bound(sorted_ids.begin(), sorted_ids.end(), id);
        if (it == sorted_ids.end() || *it != id) return -1;
        return by_id[static_cast<size_t>(it - sorted_ids.begin())];
    };
    std::vector<int32_t> q4(static_cast<size_t>(4 * id_el.n_q4)), bars(static_cast<size_t>(2 * id_el.n_bars));
    for (int64_t k = 0; k < 4 * id_el.n_q4; ++k)
        if ((q4[static_cast<size_t>(k)] = index_of(id_el.q4[k])) < 0) return 2;
    for (int64_t k = 0; k < 2 * id_el.n_bars; ++k)
        if ((bars[static_cast<size_t>(k)] = index_of(id_el.bars[k])) < 0) return 2;
    const Elems el{id_el.n_q4, id_el.n_bars, q4.data(), bars.data()};
    // distinct neighbours over the 6 node pairs of each Q4 and the pair of each bar
    // (reordering.cpp:17-28)
    const NodeElements ne = node_elements(n_nodes, el);
    std::vector<int> nptr, nlist;
    neighbour_lists(n_nodes, el, ne, nptr, nlist);
    std::vector<int> degree(static_cast<size_t>(n_nodes));
    std::vector<int> adj_ptr(static_cast<size_t>(n_nodes) + 1, 0), adj;
    adj.reserve(nlist.size());
    for (int64_t i = 0; i < n_nodes; ++i) {
        for (int k = nptr[i]; k < nptr[i + 1]; ++k)
            if (nlist[k] != i) adj.push_back(nlist[k]);
        adj_ptr[i + 1] = static_cast<int>(adj.size());
        degree[static_cast<size_t>(i)] = adj_ptr[i + 1] - adj_ptr[i];
    }
    const auto key_less = [&](int a, int b) {  // (degree, id), reordering.cpp:30-34
        if (degree[a] != degree[b]) return degree[a] < degree[b];
        return node_ids[a] < node_ids[b];
    };
    std::vector<int> seeds(static_cast<size_t>(n_nodes));
    for (int64_t i = 0; i < n_nodes; ++i) seeds[static_cast<size_t>(i)] = static_cast<int>(i);
    std::sort(seeds.begin(), seeds.end(), key_less);
    std::vector<char> visited(static_cast<size_t>(n_nodes), 0);
    std::vector<int> queue(static_cast<size_t>(n_nodes)), neigh;
    int64_t out = 0;
    for (const int seed : seeds) {
        if (visited[seed]) continue;
        visited[seed] = 1;
        size_t head = 0, tail = 0;
        queue[tail++] = seed;
        while (head < tail) {
            const int current = queue[head++];
            order_out[out++] = node_ids[current];
            neigh.assign(adj.begin() + adj_ptr[current], adj.begin() + adj_ptr[current + 1]);
            std::sort(neigh.begin(), neigh.end(), key_less);
            for (const int nb : neigh)
                if (!visited[nb]) {
                    visited[nb] = 1;
                    queue[tail++] = nb;
                }
        }
    }
    std::reverse(order_out, order_out + out);
    return 0;
}

int colors(int64_t n_nodes, const Elems& el, int32_t* color_out, int32_t* n_colors_out, int32_t* color_ptr,
           int32_t* color_elems) {
    if (!valid_elements(n_nodes, el) || !color_out || !n_colors_out) return 1;
    const int64_t n_elems = el.count();
    const NodeElements ne = node_elements(n_nodes, el);
    std::vector<int64_t> stamp_of(static_cast<size_t>(n_elems) + 1, 0);
    int max_color = -1;
    for (int64_t e = 0; e < n_elems; ++e) color_out[e] = -1;
    for (int64_t e = 0; e < n_elems; ++e) {
        const int64_t stamp = e + 1;
        for (int l = 0; l < el.nodes(e); ++l) {
            const int node = el.node(e, l);
            for (int j = ne.ptr[node]; j < ne.ptr[node + 1]; ++j) {
                const int other = ne.elems[j];
                if (color_out[other] >= 0) stamp_of[static_cast<size_t>(color_out[other])] = stamp;
            }
        }
        int color = 0;
        while (stamp_of[static_cast<size_t>(color)] == stamp) ++color;
        color_out[e] = color;
        max_color = std::max(max_color, color);
    }
    *n_colors_out = max_color + 1;
    if (color_ptr && color_elems) {
        std::fill(color_ptr, color_ptr + max_color + 2, 0);
        for (int64_t e = 0; e < n_elems; ++e) color_ptr[color_out[e] + 1]++;
        for (int c = 0; c <= max_color; ++c) color_ptr[c + 1] += color_ptr[c];
        std::vector<int32_t> cur(color_ptr, color_ptr + max_color + 1);
        for (int64_t e = 0; e < n_elems; ++e) color_elems[cur[color_out[e]]++] = static_cast<int32_t>(e);
    }
    return 0;
}

}  // namespace

extern "C" {

int64_t lfh_csr_pattern_nnz(int64_t n_nodes, int64_t n_elems, const int32_t* enodes) {
    return pattern_nnz(n_nodes, Elems{n_elems, 0, enodes, nullptr});
}

int64_t lfh_csr_pattern_nnz_bars(int64_t n_nodes, int64_t n_elems, const int32_t* enodes, int64_t n_bars,
                                 const int32_t* bnodes) {
    return pattern_nnz(n_nodes, Elems{n_elems, n_bars, enodes, bnodes});
}

int lfh_csr_pattern(int64_t n_nodes, int64_t n_elems, const int32_t* enodes, int32_t* row_ptr, int32_t* col,
                    int32_t* emap) {
    return pattern(n_nodes, Elems{n_elems, 0, enodes, nullptr}, row_ptr, col, emap, nullptr);
}

int lfh_csr_pattern_bars(int64_t n_nodes, int64_t n_elems, const int32_t* enodes, int64_t n_bars,
                         const int32_t* bnodes, int32_t* row_ptr, int32_t* col, int32_t* emap, int32_t* bmap) {
    return pattern(n_nodes, Elems{n_elems, n_bars, enodes, bnodes}, row_ptr, col, emap, bmap);
}

int lfh_rcm(int64_t n_nodes, const int32_t* node_ids, int64_t n_elems, const int32_t* enode_ids, int32_t* order_out) {
    return rcm(n_nodes, node_ids, Elems{n_elems, 0, enode_ids, nullptr}, order_out);
}

int lfh_rcm_bars(int64_t n_nodes, const int32_t* node_ids, int64_t n_elems, const int32_t* enode_ids,
                 int64_t n_bars, const int32_t* bnode_ids, int32_t* order_out) {
    return rcm(n_nodes, node_ids, Elems{n_elems, n_bars, enode_ids, bnode_ids}, order_out);
}

int lfh_element_colors(int64_t n_nodes, int64_t n_elems, const int32_t* enodes, int32_t* color_out,
                       int32_t* n_colors_out, int32_t* color_ptr, int32_t* color_elems) {
    return colors(n_nodes, Elems{n_elems, 0, enodes, nullptr}, color_out, n_colors_out, color_ptr, color_elems);
}

int lfh_element_colors_bars(int64_t n_nodes, int64_t n_elems, const int32_t* enodes, int64_t n_bars,
                            const int32_t* bnodes, int32_t* color_out, int32_t* n_colors_out, int32_t* color_ptr,
                            int32_t* color_elems) {
    return colors(n_nodes, Elems{n_elems, n_bars, enodes, bnodes}, color_out, n_colors_out, color_ptr, color_elems);
}

int lfh_subdomain_split(int64_t n_nodes, const double* xy, int32_t split, int64_t nodes_per_subdomain,
                        int32_t* node_subdomain, int64_t* n_subdomains_out) {
    if (n_nodes < 0 || (n_nodes > 0 && (!xy || !node_subdomain)) || !n_subdomains_out || (split != 0 && split != 1))
        return 1;
    const int64_t per = nodes_per_subdomain > 1 ? nodes_per_subdomain : 1;
    if (split == 1) {  // chunk: consecutive node indices
        for (int64_t i = 0; i < n_nodes; ++i) node_subdomain[i] = static_cast<int32_t>(i / per);
        *n_subdomains_out = n_nodes == 0 ? 0 : (n_nodes + per - 1) / per;
        return 0;
    }
    *n_subdomains_out = 0;
    if (n_nodes == 0) return 0;
    // Recursive coordinate bisection to exactly ceil(n / per) parts: the largest
    // remaining range of the node permutation is cut at its median along the longer
    // side of its bounding box. Which of several nodes with equal coordinates lands
    // on each side, and the final numbering (ranges in heap storage order), follow
    // from std::nth_element and the std heap operations applied in this order, so
    // the split is the reference's node for node.
    struct Span {
        std::size_t lo, hi;
    };
    auto smaller = [](const Span& a, const Span& b) { return (a.hi - a.lo) < (b.hi - b.lo); };
    std::vector<std::size_t> perm(static_cast<std::size_t>(n_nodes));
    for (std::size_t i = 0; i < perm.size(); ++i) perm[i] = i;
    const std::size_t parts = static_cast<std::size_t>((n_nodes + per - 1) / per);
    std::vector<Span> spans{Span{0, perm.size()}};
    for (std::size_t made = 1; made < parts; ++made) {
        std::pop_heap(spans.begin(), spans.end(), smaller);
        const Span cut = spans.back();
        spans.pop_back();
        const std::size_t len = cut.hi - cut.lo;
        if (len < 2) {  // only single nodes are left
            spans.push_back(cut);
            std::push_heap(spans.begin(), spans.end(), smaller);
            break;
        }
        double x0 = xy[2 * perm[cut.lo]], x1 = x0, y0 = xy[2 * perm[cut.lo] + 1], y1 = y0;
        for (std::size_t k = cut.lo + 1; k < cut.hi; ++k) {
            const double x = xy[2 * perm[k]], y = xy[2 * perm[k] + 1];
            x0 = std::min(x0, x);
            x1 = std::max(x1, x);
            y0 = std::min(y0, y);
            y1 = std::max(y1, y);
        }
        const int axis = (x1 - x0) >= (y1 - y0) ? 0 : 1;
        const std::size_t mid = cut.lo + len / 2;
        std::nth_element(perm.begin() + static_cast<std::ptrdiff_t>(cut.lo),
                         perm.begin() + static_cast<std::ptrdiff_t>(mid),
                         perm.begin() + static_cast<std::ptrdiff_t>(cut.hi),
                         [&](std::size_t a, std::size_t b) { return xy[2 * a + axis] < xy[2 * b + axis]; });
        spans.push_back(Span{cut.lo, mid});
        std::push_heap(spans.begin(), spans.end(), smaller);
        spans.push_back(Span{mid, cut.hi});
        std::push_heap(spans.begin(), spans.end(), smaller);
    }
    int32_t id = 0;
    for (const Span& sp : spans) {
        for (std::size_t k = sp.lo; k < sp.hi; ++k) node_subdomain[perm[k]] = id;
        ++id;
    }
    *n_subdomains_out = id;
    return 0;
}

int lfh_csr_profile(int64_t n_rows, const int32_t* row_ptr, const int32_t* col, int64_t out[3]) {
    if (n_rows < 0 || !row_ptr || !col || !out) return 1;
    int64_t bandwidth = 0, profile = 0;
    for (int64_t row = 0; row < n_rows; ++row) {
        int64_t minimum = row;
        for (int32_t k = row_ptr[row]; k < row_ptr[row + 1]; ++k) {
            const int64_t c = col[k];
            bandwidth = std::max(bandwidth, row > c ? row - c : c - row);
            if (c <= row) minimum = std::min(minimum, c);
        }
        profile += row - minimum;
    }
    out[0] = row_ptr[n_rows];
    out[1] = bandwidth;
    out[2] = profile;
    return 0;
}

int lfh_ic0_node_schedule(int64_t n_dofs, const int32_t* row_ptr, const int32_t* col, int32_t backward,
                          int32_t* ord, int32_t* qptr, int32_t* key, int64_t key_cap, int32_t* parts,
                          int64_t* n_parts, int32_t* reach) {
    if (n_dofs < 2 || n_dofs % 2 != 0 || n_dofs >= (int64_t{1} << 30) || !row_ptr || !col || !ord || !qptr ||
        !key || !parts || !n_parts || !reach)
        return 1;
    const int n = static_cast<int>(n_dofs);
    const lfg_ic0::Pattern pat = lfg_ic0::pattern(n, row_ptr, col);
    lfg_ic0::NodePlan plan;
    const bool ok = backward ? lfg_ic0::node_plan(n / 2, true, pat.urow.data(), pat.ucol.data(), plan)
                             : lfg_ic0::node_plan(n / 2, false, pat.lrow.data(), pat.lcol.data(), plan);
    if (!ok) return 2;
    if (static_cast<int64_t>(plan.key.size()) > key_cap) return 3;
    std::copy(plan.ord.begin(), plan.ord.end(), ord);
    std::copy(plan.qptr.begin(), plan.qptr.end(), qptr);
    std::copy(plan.key.begin(), plan.key.end(), key);
    std::copy(plan.parts.begin(), plan.parts.end(), parts);
    *n_parts = static_cast<int64_t>(plan.parts.size());
    *reach = plan.reach;
    return 0;
}

}  // extern "C"
