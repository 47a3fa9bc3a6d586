// Host-side structure of the IC(0) sweeps (precond.cpp:298-412): the factor's pattern,
// its transpose, and the level schedule the device sweeps wait on. Plain C++, shared by
// the device library (lfg_api.cpp builds the sweeps from it) and the host library
// (lfh_ic0_node_schedule exposes the schedule for inspection and tests).
#pragma once

#include <algorithm>
#include <vector>

namespace lfg_ic0 {

// L's pattern: row r holds the entries of the CSR row with column <= r in ascending
// order, the diagonal last (a row without one gets a unit diagonal: src -1); src gives
// each entry's index in the CSR arrays. The transpose of the strict lower part lists,
// per column c, the rows below it in descending order; usrc gives each entry's index in
// L's arrays.
struct Pattern {
    std::vector<int> lrow, lcol, lsrc, diag;
    std::vector<int> urow, ucol, usrc;
};

inline Pattern pattern(int n, const int* row_ptr, const int* col) {
    Pattern p;
    p.lrow.assign(static_cast<size_t>(n) + 1, 0);
    p.diag.assign(static_cast<size_t>(n), 0);
    for (int row = 0; row < n; ++row) {
        bool found = false;
        for (int k = row_ptr[row]; k < row_ptr[row + 1]; ++k) {
            const int ck = col[k];
            if (ck > row) break;
            p.lcol.push_back(ck);
            p.lsrc.push_back(k);
            if (ck == row) {
                p.diag[row] = static_cast<int>(p.lcol.size()) - 1;
                found = true;
            }
        }
        if (!found) {
            p.lcol.push_back(row);
            p.lsrc.push_back(-1);
            p.diag[row] = static_cast<int>(p.lcol.size()) - 1;
        }
        p.lrow[row + 1] = static_cast<int>(p.lcol.size());
    }
    p.urow.assign(static_cast<size_t>(n) + 1, 0);
    for (int row = 0; row < n; ++row)
        for (int k = p.lrow[row]; k < p.lrow[row + 1] - 1; ++k) p.urow[p.lcol[k] + 1]++;
    for (int i = 0; i < n; ++i) p.urow[i + 1] += p.urow[i];
    p.ucol.resize(static_cast<size_t>(p.urow[n]));
    p.usrc.resize(static_cast<size_t>(p.urow[n]));
    std::vector<int> cur(p.urow.begin(), p.urow.end() - 1);
    for (int row = n - 1; row >= 0; --row)
        for (int k = p.lrow[row]; k < p.lrow[row + 1] - 1; ++k) {
            const int c = p.lcol[k];
            p.ucol[cur[c]] = row;
            p.usrc[cur[c]] = k;
            ++cur[c];
        }
    return p;
}

// Level parts of at most 64 positions (rows of one level are independent, so any split
// keeps the order valid): the units one wavefront of a sweep takes at a time.
inline std::vector<int> split_levels(const std::vector<int>& level_ptr) {
    std::vector<int> parts(1, 0);
    for (size_t l = 0; l + 1 < level_ptr.size(); ++l) {
        const int b = level_ptr[l], e = level_ptr[l + 1];
        const int pieces = (e - b + 63) / 64;
        for (int k = 1; k <= pieces; ++k)
            parts.push_back(b + static_cast<int>((static_cast<long long>(e - b) * k) / pieces));
    }
    return parts;
}

// The node form of one sweep, when rows 2i, 2i+1 of the factor have the same dependency
// nodes. Forward (L's rows): row 2i's strict lower part is the pairs (2j, 2j+1) of its
// nodes j in ascending order and row 2i+1's is the same followed by 2i; backward (the
// transposed columns): column 2i+1 holds the pairs (2j+1, 2j) in descending j, column 2i
// the same followed by row 2i+1. Nodes are sorted by level (a node's level is one more
// than its dependencies' highest), levels split into parts of at most 64 nodes; every
// dependency becomes the position of its node in that order.
struct NodePlan {
    std::vector<int> ord;        // position -> node
    std::vector<int> qptr, key;  // position q's dependency positions: key[qptr[q] .. qptr[q+1])
    std::vector<int> first_at;   // position q: its first row's entries start at ptr[first]
    std::vector<int> second_at;  // ... and its second row's
    std::vector<int> parts;      // part boundaries (positions)
    int reach = 0;               // largest q - key: how far back a position reads
    int widest = 0;              // largest part
};

// ptr/col: the rows (forward, strict lower part plus the diagonal last) or columns
// (backward, strict part only) of the sweep. Returns false when the pattern is not in
// node form. Dependency u of position q has its values at first_at[q] + 2u, + 2u + 1 and
// second_at[q] + 2u, + 2u + 1; the second row's cross entry at second_at[q] + 2 * count.
inline bool node_plan(int nn, bool backward, const int* ptr, const int* col, NodePlan& out) {
    std::vector<int> deps_ptr(static_cast<size_t>(nn) + 1, 0), deps, fat(static_cast<size_t>(nn)),
        sat(static_cast<size_t>(nn));
    for (int i = 0; i < nn; ++i) {
        const int first = backward ? 2 * i + 1 : 2 * i, second = backward ? 2 * i : 2 * i + 1;
        const int bf = ptr[first], ef = backward ? ptr[first + 1] : ptr[first + 1] - 1;
        const int bs = ptr[second], es = backward ? ptr[second + 1] : ptr[second + 1] - 1;
        if ((ef - bf) % 2 != 0 || es - bs != ef - bf + 1 || col[es - 1] != first) return false;
        for (int k = 0; k < ef - bf; k += 2) {
            const int c0 = col[bf + k], c1 = col[bf + k + 1];
            const int j = backward ? c1 / 2 : c0 / 2;
            const bool pair_ok = backward ? (c0 == 2 * j + 1 && c1 == 2 * j) : (c0 == 2 * j && c1 == 2 * j + 1);
            if (!pair_ok || col[bs + k] != c0 || col[bs + k + 1] != c1) return false;
            deps.push_back(j);
        }
        deps_ptr[i + 1] = static_cast<int>(deps.size());
        fat[i] = bf;
        sat[i] = bs;
    }
    std::vector<int> level(static_cast<size_t>(nn), 0);
    int n_levels = 0;
    for (int t = 0; t < nn; ++t) {
        const int i = backward ? nn - 1 - t : t;
        int lv = 0;
        for (int k = deps_ptr[i]; k < deps_ptr[i + 1]; ++k) lv = std::max(lv, level[deps[k]] + 1);
        level[i] = lv;
        n_levels = std::max(n_levels, lv + 1);
    }
    std::vector<int> lptr(static_cast<size_t>(n_levels) + 1, 0);
    out.ord.assign(static_cast<size_t>(nn), 0);
    for (int i = 0; i < nn; ++i) lptr[level[i] + 1]++;
    for (int l = 0; l < n_levels; ++l) lptr[l + 1] += lptr[l];
    {
        std::vector<int> cur(lptr.begin(), lptr.end() - 1);
        for (int i = 0; i < nn; ++i) out.ord[cur[level[i]]++] = i;
    }
    std::vector<int> pos(static_cast<size_t>(nn));
    for (int q = 0; q < nn; ++q) pos[out.ord[q]] = q;
    out.qptr.assign(static_cast<size_t>(nn) + 1, 0);
    out.key.clear();
    out.first_at.assign(static_cast<size_t>(nn), 0);
    out.second_at.assign(static_cast<size_t>(nn), 0);
    out.reach = 0;
    for (int q = 0; q < nn; ++q) {
        const int i = out.ord[q];
        for (int k = deps_ptr[i]; k < deps_ptr[i + 1]; ++k) {
            out.key.push_back(pos[deps[k]]);
            out.reach = std::max(out.reach, q - pos[deps[k]]);
        }
        out.qptr[q + 1] = static_cast<int>(out.key.size());
        out.first_at[q] = fat[i];
        out.second_at[q] = sat[i];
    }
    out.parts = split_levels(lptr);
    out.widest = 0;
    for (size_t l = 0; l + 1 < out.parts.size(); ++l) out.widest = std::max(out.widest, out.parts[l + 1] - out.parts[l]);
    return true;
}

}  // namespace lfg_ic0
