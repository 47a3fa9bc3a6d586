/* Test infrastructure only: CPU restatement ("oracle") of the LightFEM Q4
 * assembly + PCG path. Only tests/, __graft_entry__.smoke() and bench.py's
 * CPU-baseline leg use it, and only as the checker. Each function restates the
 * reference stage cited beside it, in the same loop and summation order, built
 * with -ffp-contract=off so no multiply-add is fused (the reference is built for
 * plain x86-64, which has no FMA either). */
#include "lfo.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

static const double kGauss = 0.57735026918962576451; /* element.cpp:132 */

static double max0(double v) { return (0.0 < v) ? v : 0.0; } /* std::max(0.0, v) */

/* element.cpp:16-32 */
static void constitutive(double e, double nu, int plane_mode, double d[9]) {
    if (plane_mode == LFO_PLANE_STRESS) {
        const double factor = e / (1.0 - nu * nu);
        d[0] = factor; d[1] = factor * nu; d[2] = 0.0;
        d[3] = factor * nu; d[4] = factor; d[5] = 0.0;
        d[6] = 0.0; d[7] = 0.0; d[8] = factor * (1.0 - nu) * 0.5;
    } else {
        const double factor = e / ((1.0 + nu) * (1.0 - 2.0 * nu));
        d[0] = factor * (1.0 - nu); d[1] = factor * nu; d[2] = 0.0;
        d[3] = factor * nu; d[4] = factor * (1.0 - nu); d[5] = 0.0;
        d[6] = 0.0; d[7] = 0.0; d[8] = factor * (1.0 - 2.0 * nu) * 0.5;
    }
}

/* element.cpp:34-73; returns 0, or 1 on a non-positive Jacobian. */
static int kinematics(const double xy[8], double xi, double eta, double b[24], double* det_out) {
    const double dxi[4] = {-0.25 * (1.0 - eta), +0.25 * (1.0 - eta), +0.25 * (1.0 + eta), -0.25 * (1.0 + eta)};
    const double deta[4] = {-0.25 * (1.0 - xi), -0.25 * (1.0 + xi), +0.25 * (1.0 + xi), +0.25 * (1.0 - xi)};
    double dx_dxi = 0.0, dx_deta = 0.0, dy_dxi = 0.0, dy_deta = 0.0;
    for (int i = 0; i < 4; ++i) {
        dx_dxi += dxi[i] * xy[2 * i];
        dx_deta += deta[i] * xy[2 * i];
        dy_dxi += dxi[i] * xy[2 * i + 1];
        dy_deta += deta[i] * xy[2 * i + 1];
    }
    const double det = dx_dxi * dy_deta - dx_deta * dy_dxi;
    if (!(det > 1.0e-14) || !isfinite(det)) return 1;
    const double inv00 = dy_deta / det;
    const double inv01 = -dx_deta / det;
    const double inv10 = -dy_dxi / det;
    const double inv11 = dx_dxi / det;
    for (int i = 0; i < 4; ++i) {
        const double dnx = inv00 * dxi[i] + inv10 * deta[i];
        const double dny = inv01 * dxi[i] + inv11 * deta[i];
        b[2 * i] = dnx;
        b[2 * i + 1] = 0.0;
        b[8 + 2 * i] = 0.0;
        b[8 + 2 * i + 1] = dny;
        b[16 + 2 * i] = dny;
        b[16 + 2 * i + 1] = dnx;
    }
    *det_out = det;
    return 0;
}

/* element.cpp:130-161 */
int lfo_q4_stiffness(const double xy[8], double young, double nu, double thickness, int plane_mode,
                     double ke[64]) {
    const double points[4][2] = {{-kGauss, -kGauss}, {kGauss, -kGauss}, {kGauss, kGauss}, {-kGauss, kGauss}};
    double d[9];
    constitutive(young, nu, plane_mode, d);
    for (int k = 0; k < 64; ++k) ke[k] = 0.0;
    for (int p = 0; p < 4; ++p) {
        double b[24], det;
        if (kinematics(xy, points[p][0], points[p][1], b, &det)) return 1;
        double db[24];
        for (int k = 0; k < 24; ++k) db[k] = 0.0;
        for (int row = 0; row < 3; ++row)
            for (int column = 0; column < 8; ++column)
                for (int inner = 0; inner < 3; ++inner) db[8 * row + column] += d[3 * row + inner] * b[8 * inner + column];
        const double weight = det * thickness;
        for (int row = 0; row < 8; ++row)
            for (int column = 0; column < 8; ++column) {
                double contribution = 0.0;
                for (int s = 0; s < 3; ++s) contribution += b[8 * s + row] * db[8 * s + column];
                ke[8 * row + column] += contribution * weight;
            }
    }
    return 0;
}

/* Adjoint element sensitivity. The reference has no gradient code: this is the
 * derivative of its forward path. Ke is linear in E (E enters only through D,
 * element.cpp:16-32), so dKe/dE is q4_stiffness (element.cpp:130-161) at E = 1.
 * grad_e = -(lambda_e^T Ke(1) u_e) * dE_e: per row r, (Ke(1) u_e)_r sums columns in
 * order, t_r = lambda_e[r] * that, and the eight t_r are summed pairwise,
 * ((t0+t1)+(t2+t3)) + ((t4+t5)+(t6+t7)). props: (E, nu, t) per element (E unused);
 * dE NULL means 1. Returns the number of elements with a non-positive Jacobian. */
int64_t lfo_sensitivity(int64_t nelem, const double* xy_nodes, const int64_t* enode_idx, const double* props,
                        int plane_mode, const double* u, const double* lam, const double* dE, double* grad) {
    int64_t failures = 0;
    for (int64_t e = 0; e < nelem; ++e) {
        double xy[8], ke[64], ue[8], le[8], t[8];
        for (int i = 0; i < 4; ++i) {
            const int64_t node = enode_idx[4 * e + i];
            xy[2 * i] = xy_nodes[2 * node];
            xy[2 * i + 1] = xy_nodes[2 * node + 1];
            ue[2 * i] = u[2 * node];
            ue[2 * i + 1] = u[2 * node + 1];
            le[2 * i] = lam[2 * node];
            le[2 * i + 1] = lam[2 * node + 1];
        }
        if (lfo_q4_stiffness(xy, 1.0, props[3 * e + 1], props[3 * e + 2], plane_mode, ke)) {
            ++failures;
            grad[e] = 0.0;
            continue;
        }
        for (int r = 0; r < 8; ++r) {
            double ku = 0.0;
            for (int c = 0; c < 8; ++c) ku += ke[8 * r + c] * ue[c];
            t[r] = le[r] * ku;
        }
        const double s = ((t[0] + t[1]) + (t[2] + t[3])) + ((t[4] + t[5]) + (t[6] + t[7]));
        grad[e] = -s * (dE ? dE[e] : 1.0);
    }
    return failures;
}

/* element.cpp:75-95 and 212-248 */
static void field_at(const double b[24], const double d[9], const double u[8], double out[7]) {
    double strain[3] = {0.0, 0.0, 0.0}, stress[3] = {0.0, 0.0, 0.0};
    for (int row = 0; row < 3; ++row)
        for (int column = 0; column < 8; ++column) strain[row] += b[8 * row + column] * u[column];
    for (int row = 0; row < 3; ++row)
        for (int column = 0; column < 3; ++column) stress[row] += d[3 * row + column] * strain[column];
    const double sx = stress[0], sy = stress[1], txy = stress[2];
    out[0] = strain[0]; out[1] = strain[1]; out[2] = strain[2];
    out[3] = sx; out[4] = sy; out[5] = txy;
    out[6] = sqrt(max0(sx * sx - sx * sy + sy * sy + 3.0 * txy * txy));
}

int lfo_q4_fields(const double xy[8], double young, double nu, double thickness, int plane_mode,
                  const double u[8], double centre[7], double gp[36]) {
    (void)thickness;
    const double points[4][2] = {{-kGauss, -kGauss}, {kGauss, -kGauss}, {kGauss, kGauss}, {-kGauss, kGauss}};
    double d[9], b[24], det;
    constitutive(young, nu, plane_mode, d);
    if (kinematics(xy, 0.0, 0.0, b, &det)) return 1;
    field_at(b, d, u, centre);
    for (int p = 0; p < 4; ++p) {
        if (kinematics(xy, points[p][0], points[p][1], b, &det)) return 1;
        gp[9 * p] = points[p][0];
        gp[9 * p + 1] = points[p][1];
        field_at(b, d, u, gp + 9 * p + 2);
    }
    return 0;
}

/* ---------------------------------------------------------------- pattern */

static int cmp_i64(const void* a, const void* b) {
    const int64_t x = *(const int64_t*)a, y = *(const int64_t*)b;
    return (x > y) - (x < y);
}

/* sparse.cpp:15-41: per-row column set (diagonal always inserted), sorted. Element e
 * has the dofs edofs[eptr[e] .. eptr[e+1]) (Q4: 8, bar: 4). */
static int64_t build_rows_var(int64_t ndofs, int64_t nelem, const int64_t* eptr, const int64_t* edofs,
                              int64_t* row_ptr, int64_t* col) {
    int64_t* count = calloc((size_t)ndofs + 1, sizeof(int64_t));
    for (int64_t r = 0; r < ndofs; ++r) count[r] = 1;
    for (int64_t e = 0; e < nelem; ++e)
        for (int64_t a = eptr[e]; a < eptr[e + 1]; ++a) count[edofs[a]] += eptr[e + 1] - eptr[e];
    int64_t* start = malloc(((size_t)ndofs + 1) * sizeof(int64_t));
    start[0] = 0;
    for (int64_t r = 0; r < ndofs; ++r) start[r + 1] = start[r] + count[r];
    int64_t* raw = malloc((size_t)start[ndofs] * sizeof(int64_t));
    int64_t* fill = malloc((size_t)ndofs * sizeof(int64_t));
    for (int64_t r = 0; r < ndofs; ++r) {
        fill[r] = start[r];
        raw[fill[r]++] = r;
    }
    for (int64_t e = 0; e < nelem; ++e)
        for (int64_t a = eptr[e]; a < eptr[e + 1]; ++a) {
            const int64_t row = edofs[a];
            for (int64_t c = eptr[e]; c < eptr[e + 1]; ++c) raw[fill[row]++] = edofs[c];
        }
    int64_t nnz = 0;
    if (row_ptr) row_ptr[0] = 0;
    for (int64_t r = 0; r < ndofs; ++r) {
        int64_t* seg = raw + start[r];
        const int64_t len = start[r + 1] - start[r];
        qsort(seg, (size_t)len, sizeof(int64_t), cmp_i64);
        for (int64_t i = 0; i < len; ++i) {
            if (i > 0 && seg[i] == seg[i - 1]) continue;
            if (col) col[nnz] = seg[i];
            ++nnz;
        }
        if (row_ptr) row_ptr[r + 1] = nnz;
    }
    free(count); free(start); free(raw); free(fill);
    return nnz;
}

static int64_t* q4_eptr(int64_t nelem) {
    int64_t* eptr = malloc(((size_t)nelem + 1) * sizeof(int64_t));
    for (int64_t e = 0; e <= nelem; ++e) eptr[e] = 8 * e;
    return eptr;
}

static int64_t build_rows(int64_t ndofs, int64_t nelem, const int64_t* edofs, int64_t* row_ptr, int64_t* col) {
    int64_t* eptr = q4_eptr(nelem);
    const int64_t nnz = build_rows_var(ndofs, nelem, eptr, edofs, row_ptr, col);
    free(eptr);
    return nnz;
}

int64_t lfo_pattern_nnz_var(int64_t ndofs, int64_t nelem, const int64_t* eptr, const int64_t* edofs) {
    return build_rows_var(ndofs, nelem, eptr, edofs, NULL, NULL);
}

int64_t lfo_pattern_nnz(int64_t ndofs, int64_t nelem, const int64_t* edofs) {
    return build_rows(ndofs, nelem, edofs, NULL, NULL);
}

/* sparse.cpp:65-74 value_index (lower_bound in the row) */
static int64_t value_index(const int64_t* row_ptr, const int64_t* col, int64_t row, int64_t column) {
    int64_t lo = row_ptr[row], hi = row_ptr[row + 1];
    while (lo < hi) {
        const int64_t mid = lo + (hi - lo) / 2;
        if (col[mid] < column) lo = mid + 1; else hi = mid;
    }
    if (lo == row_ptr[row + 1] || col[lo] != column) return -1;
    return lo;
}

/* sparse.cpp:15-58 */
int lfo_pattern_build(int64_t ndofs, int64_t nelem, const int64_t* edofs, int64_t* row_ptr, int64_t* col,
                      int64_t* emap) {
    build_rows(ndofs, nelem, edofs, row_ptr, col);
    for (int64_t e = 0; e < nelem; ++e)
        for (int r = 0; r < 8; ++r)
            for (int c = 0; c < 8; ++c) {
                const int64_t idx = value_index(row_ptr, col, edofs[8 * e + r], edofs[8 * e + c]);
                if (idx < 0) return 1;
                emap[64 * e + 8 * r + c] = idx;
            }
    return 0;
}

/* sparse.cpp:15-58 with element dof lists of any length (Q4 elements, then bars,
 * solver.cpp:71-83); emap gets n_e^2 entries per element, row-major, in element order. */
int lfo_pattern_build_var(int64_t ndofs, int64_t nelem, const int64_t* eptr, const int64_t* edofs,
                          int64_t* row_ptr, int64_t* col, int64_t* emap) {
    build_rows_var(ndofs, nelem, eptr, edofs, row_ptr, col);
    int64_t m = 0;
    for (int64_t e = 0; e < nelem; ++e)
        for (int64_t r = eptr[e]; r < eptr[e + 1]; ++r)
            for (int64_t c = eptr[e]; c < eptr[e + 1]; ++c) {
                const int64_t idx = value_index(row_ptr, col, edofs[r], edofs[c]);
                if (idx < 0) return 1;
                emap[m++] = idx;
            }
    return 0;
}

/* sparse.cpp:227-242 */
void lfo_csr_profile(int64_t n, const int64_t* row_ptr, const int64_t* col, int64_t out[3]) {
    int64_t bandwidth = 0, profile = 0;
    for (int64_t row = 0; row < n; ++row) {
        int64_t minimum = row;
        for (int64_t i = row_ptr[row]; i < row_ptr[row + 1]; ++i) {
            const int64_t c = col[i];
            const int64_t dist = row > c ? row - c : c - row;
            if (bandwidth < dist) bandwidth = dist;
            if (c <= row && c < minimum) minimum = c;
        }
        profile += row - minimum;
    }
    out[0] = row_ptr[n];
    out[1] = bandwidth;
    out[2] = profile;
}

/* --------------------------------------------------------------- ordering */

typedef struct {
    int64_t nnodes;
    const int64_t* ids;    /* node ids, any order */
    int64_t* sorted_ids;   /* ascending */
    int64_t* sorted_pos;   /* position in ids[] */
} id_index;

static id_index* g_idx_for_sort;
static int cmp_pos_by_id(const void* a, const void* b) {
    const int64_t x = g_idx_for_sort->ids[*(const int64_t*)a], y = g_idx_for_sort->ids[*(const int64_t*)b];
    return (x > y) - (x < y);
}

static void id_index_build(id_index* ix, int64_t nnodes, const int64_t* ids) {
    ix->nnodes = nnodes;
    ix->ids = ids;
    ix->sorted_pos = malloc((size_t)nnodes * sizeof(int64_t));
    ix->sorted_ids = malloc((size_t)nnodes * sizeof(int64_t));
    for (int64_t i = 0; i < nnodes; ++i) ix->sorted_pos[i] = i;
    g_idx_for_sort = ix;
    qsort(ix->sorted_pos, (size_t)nnodes, sizeof(int64_t), cmp_pos_by_id);
    for (int64_t i = 0; i < nnodes; ++i) ix->sorted_ids[i] = ids[ix->sorted_pos[i]];
}

static int64_t id_lookup(const id_index* ix, int64_t id) {
    int64_t lo = 0, hi = ix->nnodes;
    while (lo < hi) {
        const int64_t mid = lo + (hi - lo) / 2;
        if (ix->sorted_ids[mid] < id) lo = mid + 1; else hi = mid;
    }
    return (lo < ix->nnodes && ix->sorted_ids[lo] == id) ? ix->sorted_pos[lo] : -1;
}

static void id_index_free(id_index* ix) {
    free(ix->sorted_pos);
    free(ix->sorted_ids);
}

/* RCM key (degree, id): reordering.cpp:274-278 */
static const int64_t* g_degree;
static const int64_t* g_ids;
static int cmp_degree_id(const void* a, const void* b) {
    const int64_t x = *(const int64_t*)a, y = *(const int64_t*)b; /* positions */
    if (g_degree[x] != g_degree[y]) return g_degree[x] < g_degree[y] ? -1 : 1;
    return (g_ids[x] > g_ids[y]) - (g_ids[x] < g_ids[y]);
}

/* reordering.cpp:11-62; order_out receives node ids. */
int lfo_rcm(int64_t nnodes, const int64_t* node_ids, int64_t nelem, const int64_t* enode_ids, int64_t* order_out) {
    id_index ix;
    id_index_build(&ix, nnodes, node_ids);
    /* adjacency as sorted-unique neighbour position lists (set semantics) */
    int64_t* cnt = calloc((size_t)nnodes + 1, sizeof(int64_t));
    int64_t* en = malloc((size_t)(4 * nelem > 0 ? 4 * nelem : 1) * sizeof(int64_t));
    for (int64_t k = 0; k < 4 * nelem; ++k) {
        en[k] = id_lookup(&ix, enode_ids[k]);
        if (en[k] < 0) { id_index_free(&ix); free(cnt); free(en); return 1; }
    }
    for (int64_t e = 0; e < nelem; ++e)
        for (int i = 0; i < 4; ++i)
            for (int j = i + 1; j < 4; ++j) { cnt[en[4 * e + i]]++; cnt[en[4 * e + j]]++; }
    int64_t* start = malloc(((size_t)nnodes + 1) * sizeof(int64_t));
    start[0] = 0;
    for (int64_t i = 0; i < nnodes; ++i) start[i + 1] = start[i] + cnt[i];
    int64_t* nb = malloc((size_t)(start[nnodes] > 0 ? start[nnodes] : 1) * sizeof(int64_t));
    int64_t* fill = malloc((size_t)nnodes * sizeof(int64_t));
    for (int64_t i = 0; i < nnodes; ++i) fill[i] = start[i];
    for (int64_t e = 0; e < nelem; ++e)
        for (int i = 0; i < 4; ++i)
            for (int j = i + 1; j < 4; ++j) {
                const int64_t a = en[4 * e + i], b = en[4 * e + j];
                nb[fill[a]++] = b;
                nb[fill[b]++] = a;
            }
    int64_t* degree = malloc((size_t)nnodes * sizeof(int64_t));
    int64_t* ustart = malloc(((size_t)nnodes + 1) * sizeof(int64_t));
    ustart[0] = 0;
    int64_t w = 0;
    for (int64_t i = 0; i < nnodes; ++i) {
        int64_t* seg = nb + start[i];
        const int64_t len = start[i + 1] - start[i];
        qsort(seg, (size_t)len, sizeof(int64_t), cmp_i64);
        int64_t d = 0;
        for (int64_t k = 0; k < len; ++k) {
            if (k > 0 && seg[k] == seg[k - 1]) continue;
            nb[w++] = seg[k];
            ++d;
        }
        degree[i] = d;
        ustart[i + 1] = w;
    }
    g_degree = degree;
    g_ids = node_ids;
    int64_t* seeds = malloc((size_t)(nnodes > 0 ? nnodes : 1) * sizeof(int64_t));
    for (int64_t i = 0; i < nnodes; ++i) seeds[i] = i; /* positions in model.nodes order */
    qsort(seeds, (size_t)nnodes, sizeof(int64_t), cmp_degree_id);
    char* visited = calloc((size_t)nnodes + 1, 1);
    int64_t* queue = malloc((size_t)(nnodes > 0 ? nnodes : 1) * sizeof(int64_t));
    int64_t* tmp = malloc((size_t)(start[nnodes] > 0 ? start[nnodes] : 1) * sizeof(int64_t));
    int64_t out = 0;
    for (int64_t s = 0; s < nnodes; ++s) {
        const int64_t seed = seeds[s];
        if (visited[seed]) continue;
        visited[seed] = 1;
        int64_t head = 0, tail = 0;
        queue[tail++] = seed;
        while (head < tail) {
            const int64_t current = queue[head++];
            order_out[out++] = node_ids[current];
            const int64_t len = ustart[current + 1] - ustart[current];
            memcpy(tmp, nb + ustart[current], (size_t)len * sizeof(int64_t));
            qsort(tmp, (size_t)len, sizeof(int64_t), cmp_degree_id);
            for (int64_t k = 0; k < len; ++k)
                if (!visited[tmp[k]]) { visited[tmp[k]] = 1; queue[tail++] = tmp[k]; }
        }
    }
    for (int64_t i = 0; i < out / 2; ++i) {
        const int64_t t = order_out[i];
        order_out[i] = order_out[out - 1 - i];
        order_out[out - 1 - i] = t;
    }
    free(cnt); free(en); free(start); free(nb); free(fill); free(degree); free(ustart);
    free(seeds); free(visited); free(queue); free(tmp);
    id_index_free(&ix);
    return 0;
}

/* solver.cpp:100-133 greedy first-fit colouring in element order; elements sharing
 * a node conflict. Element e has the node ids enode_ids[nptr[e] .. nptr[e+1]) (Q4: 4,
 * bar: 2). Returns the number of colours. */
int64_t lfo_colors_var(int64_t nelem, const int64_t* nptr, const int64_t* enode_ids, int64_t* color_out) {
    if (nelem == 0) return 0;
    const int64_t total = nptr[nelem];
    /* node id -> element list (elements appended in ascending order) */
    id_index ix;
    int64_t* ids = malloc((size_t)total * sizeof(int64_t));
    memcpy(ids, enode_ids, (size_t)total * sizeof(int64_t));
    qsort(ids, (size_t)total, sizeof(int64_t), cmp_i64);
    int64_t nu = 0;
    for (int64_t k = 0; k < total; ++k)
        if (k == 0 || ids[k] != ids[k - 1]) ids[nu++] = ids[k];
    id_index_build(&ix, nu, ids);
    int64_t* cnt = calloc((size_t)nu + 1, sizeof(int64_t));
    int64_t* en = malloc((size_t)total * sizeof(int64_t));
    for (int64_t k = 0; k < total; ++k) { en[k] = id_lookup(&ix, enode_ids[k]); cnt[en[k]]++; }
    int64_t* start = malloc(((size_t)nu + 1) * sizeof(int64_t));
    start[0] = 0;
    for (int64_t i = 0; i < nu; ++i) start[i + 1] = start[i] + cnt[i];
    int64_t* list = malloc((size_t)start[nu] * sizeof(int64_t));
    int64_t* fill = malloc((size_t)nu * sizeof(int64_t));
    for (int64_t i = 0; i < nu; ++i) fill[i] = start[i];
    for (int64_t e = 0; e < nelem; ++e)
        for (int64_t k = nptr[e]; k < nptr[e + 1]; ++k) list[fill[en[k]]++] = e;
    int64_t* stamp_of = calloc((size_t)nelem, sizeof(int64_t));
    int64_t stamp = 0, maximum = -1;
    for (int64_t e = 0; e < nelem; ++e) color_out[e] = -1;
    for (int64_t e = 0; e < nelem; ++e) {
        ++stamp;
        for (int64_t k = nptr[e]; k < nptr[e + 1]; ++k) {
            const int64_t node = en[k];
            for (int64_t q = start[node]; q < start[node + 1]; ++q) {
                const int64_t nbr = list[q];
                if (color_out[nbr] >= 0) stamp_of[color_out[nbr]] = stamp;
            }
        }
        int64_t color = 0;
        while (stamp_of[color] == stamp) ++color;
        color_out[e] = color;
        if (maximum < color) maximum = color;
    }
    free(ids); free(cnt); free(en); free(start); free(list); free(fill); free(stamp_of);
    id_index_free(&ix);
    return maximum + 1;
}

int64_t lfo_colors(int64_t nelem, const int64_t* enode_ids, int64_t* color_out) {
    int64_t* nptr = malloc(((size_t)nelem + 1) * sizeof(int64_t));
    for (int64_t e = 0; e <= nelem; ++e) nptr[e] = 4 * e;
    const int64_t count = lfo_colors_var(nelem, nptr, enode_ids, color_out);
    free(nptr);
    return count;
}

/* -------------------------------------------------------------- assembly */

/* solver.cpp:135-142 */
void lfo_assemble_serial(int64_t nelem, const double* ke, const int64_t* emap, int64_t nnz, double* values) {
    for (int64_t k = 0; k < nnz; ++k) values[k] = 0.0;
    for (int64_t e = 0; e < nelem; ++e)
        for (int k = 0; k < 64; ++k) values[emap[64 * e + k]] += ke[64 * e + k];
}

/* solver.cpp:135-142 over flat element matrices (Q4 64 entries each, then bars 16). */
void lfo_assemble_flat(int64_t total, const double* ke, const int64_t* emap, int64_t nnz, double* values) {
    for (int64_t k = 0; k < nnz; ++k) values[k] = 0.0;
    for (int64_t k = 0; k < total; ++k) values[emap[k]] += ke[k];
}

/* element.cpp:266-286: bar stiffness EA/L [c^2 cs -c^2 -cs; ...], row-major 4x4. xy =
 * (x1, y1, x2, y2). Returns 1 for a zero-length bar. */
int lfo_truss_stiffness(const double xy[4], double young, double area, double ke[16]) {
    const double dx = xy[2] - xy[0];
    const double dy = xy[3] - xy[1];
    const double length = sqrt(dx * dx + dy * dy);
    if (!(length > 1.0e-12) || !isfinite(length)) return 1;
    const double c = dx / length, s = dy / length;
    const double scale = young * area / length;
    const double cc = c * c, ss = s * s, cs = c * s;
    const double v[16] = {scale * cc,  scale * cs,  -scale * cc, -scale * cs, scale * cs,  scale * ss,
                          -scale * cs, -scale * ss, -scale * cc, -scale * cs, scale * cc,  scale * cs,
                          -scale * cs, -scale * ss, scale * cs,  scale * ss};
    memcpy(ke, v, sizeof(v));
    return 0;
}

/* element.cpp:288-306 and solver.cpp:753-803 for one bar: out = strain (axial, 0, 0),
 * stress sigma * (c^2, s^2, cs), von Mises |sigma|. Returns 1 for a zero-length bar. */
int lfo_truss_fields(const double xy[4], double young, const double u[4], double out[7]) {
    const double dx = xy[2] - xy[0];
    const double dy = xy[3] - xy[1];
    const double length = sqrt(dx * dx + dy * dy);
    if (!(length > 1.0e-12) || !isfinite(length)) return 1;
    const double c = dx / length, s = dy / length;
    const double du = u[2] - u[0], dv = u[3] - u[1];
    const double strain = (du * c + dv * s) / length;
    const double sigma = young * strain;
    out[0] = strain;
    out[1] = 0.0;
    out[2] = 0.0;
    out[3] = sigma * c * c;
    out[4] = sigma * s * s;
    out[5] = sigma * c * s;
    out[6] = fabs(sigma);
    return 0;
}

/* solver.cpp:916-940: after the Dirichlet elimination, a dof whose diagonal is at most
 * 1e-30 in magnitude is pinned (diagonal 1, right-hand side 0) unless it carries a load
 * above 1e-12 of the largest |rhs| (1 when all are zero): then it returns -1 with the
 * first such dof in *bad. Otherwise returns the number of pinned dofs. */
int64_t lfo_pin_zero_pivots(int64_t n, const int64_t* row_ptr, const int64_t* col, double* values, double* rhs,
                            int64_t* bad) {
    double scale = 0.0;
    for (int64_t i = 0; i < n; ++i) scale = scale > fabs(rhs[i]) ? scale : fabs(rhs[i]);
    if (scale <= 0.0) scale = 1.0;
    int64_t pinned = 0;
    for (int64_t i = 0; i < n; ++i) {
        const int64_t d = value_index(row_ptr, col, i, i);
        const double diag = d >= 0 ? values[d] : 0.0;
        if (fabs(diag) > 1.0e-30) continue;
        if (fabs(rhs[i]) > 1.0e-12 * scale) {
            *bad = i;
            return -1;
        }
        values[d] = 1.0;
        rhs[i] = 0.0;
        ++pinned;
    }
    return pinned;
}

/* solver.cpp:355-413 (the O(nnz) sweep). */
void lfo_dirichlet(int64_t n, const int64_t* row_ptr, const int64_t* col, double* values, double* rhs, int64_t nc,
                   const int64_t* cdofs, const double* cvals) {
    if (nc == 0) return;
    char* constrained = calloc((size_t)n, 1);
    double* prescribed = calloc((size_t)n, sizeof(double));
    for (int64_t k = 0; k < nc; ++k) {
        constrained[cdofs[k]] = 1;
        prescribed[cdofs[k]] = cvals[k];
    }
    for (int64_t row = 0; row < n; ++row) {
        const int64_t begin = row_ptr[row], end = row_ptr[row + 1];
        if (constrained[row]) {
            for (int64_t i = begin; i < end; ++i) values[i] = 0.0;
            values[value_index(row_ptr, col, row, row)] = 1.0;
            rhs[row] = prescribed[row];
            continue;
        }
        for (int64_t i = begin; i < end; ++i) {
            const int64_t c = col[i];
            if (constrained[c] == 0 || values[i] == 0.0) continue;
            rhs[row] -= values[i] * prescribed[c];
            values[i] = 0.0;
        }
    }
    free(constrained);
    free(prescribed);
}

/* sparse.cpp:104-120 (row order is irrelevant; each row is summed serially). */
void lfo_spmv(int64_t n, const int64_t* row_ptr, const int64_t* col, const double* values, const double* x,
              double* y) {
    for (int64_t row = 0; row < n; ++row) {
        double sum = 0.0;
        for (int64_t i = row_ptr[row]; i < row_ptr[row + 1]; ++i) sum += values[i] * x[col[i]];
        y[row] = sum;
    }
}

/* sparse.cpp:128-198: the block storage product. Block row I merges the sorted
 * columns of rows 2I and 2I+1 into 2x2 blocks (zero where a row lacks the column)
 * and sums each block's two products before adding them to the row total. */
void lfo_bsr_spmv(int64_t n, const int64_t* row_ptr, const int64_t* col, const double* values, const double* x,
                  double* y) {
    for (int64_t br = 0; br < n / 2; ++br) {
        const int64_t r0 = 2 * br, r1 = r0 + 1;
        int64_t p0 = row_ptr[r0], p1 = row_ptr[r1];
        const int64_t e0 = row_ptr[r0 + 1], e1 = row_ptr[r1 + 1];
        double sum0 = 0.0, sum1 = 0.0;
        while (p0 < e0 || p1 < e1) {
            const int64_t c0 = p0 < e0 ? col[p0] : INT64_MAX;
            const int64_t c1 = p1 < e1 ? col[p1] : INT64_MAX;
            const int64_t bc = (c0 < c1 ? c0 : c1) / 2;
            double blk[4] = {0.0, 0.0, 0.0, 0.0};
            for (; p0 < e0 && col[p0] / 2 == bc; ++p0) blk[col[p0] - 2 * bc] = values[p0];
            for (; p1 < e1 && col[p1] / 2 == bc; ++p1) blk[2 + col[p1] - 2 * bc] = values[p1];
            const double x0 = x[2 * bc], x1 = x[2 * bc + 1];
            sum0 += blk[0] * x0 + blk[1] * x1;
            sum1 += blk[2] * x0 + blk[3] * x1;
        }
        y[r0] = sum0;
        y[r1] = sum1;
    }
}

/* Block count of lfo_bsr_spmv's storage (the reference's block_nonzeros()). */
int64_t lfo_bsr_blocks(int64_t n, const int64_t* row_ptr, const int64_t* col) {
    int64_t blocks = 0;
    for (int64_t br = 0; br < n / 2; ++br) {
        int64_t p0 = row_ptr[2 * br], p1 = row_ptr[2 * br + 1];
        const int64_t e0 = row_ptr[2 * br + 1], e1 = row_ptr[2 * br + 2];
        while (p0 < e0 || p1 < e1) {
            const int64_t c0 = p0 < e0 ? col[p0] : INT64_MAX;
            const int64_t c1 = p1 < e1 ? col[p1] : INT64_MAX;
            const int64_t bc = (c0 < c1 ? c0 : c1) / 2;
            while (p0 < e0 && col[p0] / 2 == bc) ++p0;
            while (p1 < e1 && col[p1] / 2 == bc) ++p1;
            ++blocks;
        }
    }
    return blocks;
}

/* -------------------------------------------------------- preconditioners */

int64_t lfo_ic0_factor_size(int64_t n, const int64_t* row_ptr, const int64_t* col) {
    int64_t size = 0;
    for (int64_t row = 0; row < n; ++row) {
        int found = 0;
        for (int64_t i = row_ptr[row]; i < row_ptr[row + 1]; ++i) {
            if (col[i] > row) break;
            ++size;
            if (col[i] == row) found = 1;
        }
        if (!found) ++size;
    }
    return size;
}

/* precond.cpp:298-384 */
int64_t lfo_ic0_factor(int64_t n, const int64_t* row_ptr, const int64_t* col, const double* values, int64_t* lrow,
                       int64_t* lcol, double* lval, double* inv_diag) {
    int64_t* diag = malloc((size_t)(n > 0 ? n : 1) * sizeof(int64_t));
    int64_t m = 0;
    lrow[0] = 0;
    for (int64_t row = 0; row < n; ++row) {
        int found = 0;
        for (int64_t i = row_ptr[row]; i < row_ptr[row + 1]; ++i) {
            const int64_t c = col[i];
            if (c > row) break;
            lcol[m] = c;
            lval[m] = values[i];
            ++m;
            if (c == row) { diag[row] = m - 1; found = 1; }
        }
        if (!found) { lcol[m] = row; lval[m] = 1.0; ++m; diag[row] = m - 1; }
        lrow[row + 1] = m;
        inv_diag[row] = 1.0;
    }
    double scale = 0.0;
    for (int64_t row = 0; row < n; ++row) {
        const double a = fabs(lval[diag[row]]);
        if (scale < a) scale = a;
    }
    const double floor_value = (1.0e-30 < 1.0e-14 * scale) ? 1.0e-14 * scale : 1.0e-30;
    int64_t breakdowns = 0;
    for (int64_t row = 0; row < n; ++row) {
        const int64_t begin = lrow[row], d = diag[row];
        for (int64_t index = begin; index < d; ++index) {
            const int64_t column = lcol[index];
            const int64_t other_begin = lrow[column], other_diag = diag[column];
            double sum = lval[index];
            int64_t left = begin, right = other_begin;
            while (left < index && right < other_diag) {
                const int64_t lc = lcol[left], rc = lcol[right];
                if (lc < rc) ++left;
                else if (lc > rc) ++right;
                else { sum -= lval[left] * lval[right]; ++left; ++right; }
            }
            lval[index] = sum * inv_diag[column];
        }
        double sum = lval[d];
        for (int64_t index = begin; index < d; ++index) sum -= lval[index] * lval[index];
        if (!(sum > 0.0) || !isfinite(sum)) { ++breakdowns; sum = floor_value; }
        const double pivot = sqrt(sum);
        lval[d] = pivot;
        inv_diag[row] = 1.0 / pivot;
    }
    free(diag);
    return breakdowns;
}

/* precond.cpp:386-412 */
void lfo_ic0_apply(int64_t n, const int64_t* lrow, const int64_t* lcol, const double* lval, const double* inv_diag,
                   const double* in, double* out) {
    if (out != in) memcpy(out, in, (size_t)n * sizeof(double));
    for (int64_t row = 0; row < n; ++row) {
        double sum = out[row];
        for (int64_t i = lrow[row]; i + 1 < lrow[row + 1]; ++i) sum -= lval[i] * out[lcol[i]];
        out[row] = sum * inv_diag[row];
    }
    for (int64_t row = n; row-- > 0;) {
        out[row] *= inv_diag[row];
        const double value = out[row];
        for (int64_t i = lrow[row]; i + 1 < lrow[row + 1]; ++i) out[lcol[i]] -= lval[i] * value;
    }
}

/* precond.cpp:178-232 for one-node subdomains (dofs 2I, 2I+1 in matrix order). */
int64_t lfo_block1_factor(int64_t n, const int64_t* row_ptr, const int64_t* col, const double* values,
                          double* factor) {
    int64_t degenerate = 0;
    for (int64_t node = 0; node < n / 2; ++node) {
        double dense[4] = {0.0, 0.0, 0.0, 0.0};
        for (int i = 0; i < 2; ++i) {
            const int64_t dof = 2 * node + i;
            for (int64_t k = row_ptr[dof]; k < row_ptr[dof + 1]; ++k) {
                const int64_t c = col[k];
                if (c / 2 != node) continue;
                dense[i * 2 + (c - 2 * node)] = values[k];
            }
        }
        int factorized = 1;
        for (int i = 0; i < 2 && factorized; ++i)
            for (int j = 0; j <= i; ++j) {
                double sum = dense[i * 2 + j];
                for (int k = 0; k < j; ++k) sum -= dense[i * 2 + k] * dense[j * 2 + k];
                if (i == j) {
                    if (!(sum > 0.0) || !isfinite(sum)) { factorized = 0; break; }
                    dense[i * 2 + i] = sqrt(sum);
                } else {
                    dense[i * 2 + j] = sum / dense[j * 2 + j];
                }
            }
        if (!factorized) {
            ++degenerate;
            dense[0] = 1.0; dense[2] = 0.0; dense[3] = 1.0;
        }
        double* t = factor + 3 * node;
        t[0] = dense[0];
        t[0] = 1.0 / t[0];
        t[1] = dense[2];
        t[2] = dense[3];
        t[2] = 1.0 / t[2];
    }
    return degenerate;
}

/* precond.cpp:242-291 for size-2 blocks. */
void lfo_block1_apply(int64_t n, const double* factor, const double* in, double* out) {
    for (int64_t node = 0; node < n / 2; ++node) {
        const double* f = factor + 3 * node;
        double s0 = in[2 * node], s1 = in[2 * node + 1];
        /* forward */
        s0 = s0 * f[0];
        {
            double sum = s1;
            sum -= f[1] * s0;
            s1 = sum * f[2];
        }
        /* backward */
        s1 = s1 * f[2];
        {
            double sum = s0;
            sum -= f[1] * s1;
            s0 = sum * f[0];
        }
        out[2 * node] = s0;
        out[2 * node + 1] = s1;
    }
}

/* precond.cpp:19-28: node i belongs to subdomain i / size. Returns the count. */
int64_t lfo_split_chunks(int64_t n_nodes, int64_t nodes_per_subdomain, int64_t* node_subdomain) {
    const int64_t size = nodes_per_subdomain > 1 ? nodes_per_subdomain : 1;
    for (int64_t i = 0; i < n_nodes; ++i) node_subdomain[i] = i / size;
    return n_nodes == 0 ? 0 : (n_nodes + size - 1) / size;
}

/* SubdomainBlockJacobi (precond.cpp:121-232): dofs of each subdomain in matrix
 * order, one dense Cholesky per subdomain (row by row, pivots stored as their
 * reciprocals in a packed lower triangle; a failed block becomes the identity). */
struct lfo_block {
    int64_t n, n_sub, largest;
    int64_t* dof_begin;    /* n_sub + 1 */
    int64_t* dofs;         /* n */
    int64_t* factor_begin; /* n_sub + 1 */
    double* factor;
    int64_t degenerate;
    double capture_ratio;
};

lfo_block* lfo_block_create(int64_t n, const int64_t* row_ptr, const int64_t* col, const double* values,
                            const int64_t* dof_subdomain, int64_t n_sub) {
    lfo_block* b = calloc(1, sizeof(lfo_block));
    b->n = n;
    b->n_sub = n_sub;
    b->dof_begin = calloc((size_t)n_sub + 1, sizeof(int64_t));
    b->dofs = malloc((size_t)(n > 0 ? n : 1) * sizeof(int64_t));
    b->factor_begin = calloc((size_t)n_sub + 1, sizeof(int64_t));
    int64_t interior = 0;
    for (int64_t row = 0; row < n; ++row)
        for (int64_t k = row_ptr[row]; k < row_ptr[row + 1]; ++k)
            if (dof_subdomain[col[k]] == dof_subdomain[row]) ++interior;
    b->capture_ratio = row_ptr[n] ? (double)interior / (double)row_ptr[n] : 0.0; /* precond.cpp:107-119 */
    for (int64_t d = 0; d < n; ++d) b->dof_begin[dof_subdomain[d] + 1]++;
    for (int64_t s = 0; s < n_sub; ++s) b->dof_begin[s + 1] += b->dof_begin[s];
    int64_t* cursor = malloc((size_t)(n_sub > 0 ? n_sub : 1) * sizeof(int64_t));
    for (int64_t s = 0; s < n_sub; ++s) cursor[s] = b->dof_begin[s];
    for (int64_t d = 0; d < n; ++d) b->dofs[cursor[dof_subdomain[d]]++] = d;
    free(cursor);
    int64_t* local = malloc((size_t)(n > 0 ? n : 1) * sizeof(int64_t));
    for (int64_t s = 0; s < n_sub; ++s) {
        const int64_t size = b->dof_begin[s + 1] - b->dof_begin[s];
        for (int64_t i = 0; i < size; ++i) local[b->dofs[b->dof_begin[s] + i]] = i;
        b->factor_begin[s + 1] = b->factor_begin[s] + size * (size + 1) / 2;
        if (size > b->largest) b->largest = size;
    }
    b->factor = calloc((size_t)b->factor_begin[n_sub] + 1, sizeof(double));
    double* dense = malloc((size_t)(b->largest * b->largest + 1) * sizeof(double));
    for (int64_t s = 0; s < n_sub; ++s) {
        const int64_t begin = b->dof_begin[s], size = b->dof_begin[s + 1] - begin;
        if (size == 0) continue;
        memset(dense, 0, (size_t)(size * size) * sizeof(double));
        for (int64_t i = 0; i < size; ++i) {
            const int64_t dof = b->dofs[begin + i];
            for (int64_t k = row_ptr[dof]; k < row_ptr[dof + 1]; ++k)
                if (dof_subdomain[col[k]] == s) dense[i * size + local[col[k]]] = values[k];
        }
        int ok = 1;
        for (int64_t i = 0; i < size && ok; ++i)
            for (int64_t j = 0; j <= i; ++j) {
                double sum = dense[i * size + j];
                for (int64_t k = 0; k < j; ++k) sum -= dense[i * size + k] * dense[j * size + k];
                if (i == j) {
                    if (!(sum > 0.0) || !isfinite(sum)) { ok = 0; break; }
                    dense[i * size + i] = sqrt(sum);
                } else {
                    dense[i * size + j] = sum / dense[j * size + j];
                }
            }
        if (!ok) {
            ++b->degenerate;
            for (int64_t i = 0; i < size; ++i)
                for (int64_t j = 0; j <= i; ++j) dense[i * size + j] = i == j ? 1.0 : 0.0;
        }
        double* t = b->factor + b->factor_begin[s];
        for (int64_t i = 0; i < size; ++i) {
            for (int64_t j = 0; j <= i; ++j) t[i * (i + 1) / 2 + j] = dense[i * size + j];
            t[i * (i + 1) / 2 + i] = 1.0 / t[i * (i + 1) / 2 + i];
        }
    }
    free(dense);
    free(local);
    return b;
}

/* precond.cpp:234-281: gather, forward substitution with L, back substitution with
 * L^T (each sum in the reference's order), scatter. */
void lfo_block_apply(const lfo_block* b, const double* in, double* out) {
    double* w = malloc((size_t)(b->largest + 1) * sizeof(double));
    for (int64_t s = 0; s < b->n_sub; ++s) {
        const int64_t begin = b->dof_begin[s], size = b->dof_begin[s + 1] - begin;
        if (size == 0) continue;
        const double* f = b->factor + b->factor_begin[s];
        for (int64_t i = 0; i < size; ++i) w[i] = in[b->dofs[begin + i]];
        for (int64_t i = 0, base = 0; i < size; base += i + 1, ++i) {
            double sum = w[i];
            for (int64_t k = 0; k < i; ++k) sum -= f[base + k] * w[k];
            w[i] = sum * f[base + i];
        }
        for (int64_t i = size; i-- > 0;) {
            double sum = w[i];
            for (int64_t k = size - 1; k > i; --k) sum -= f[k * (k + 1) / 2 + i] * w[k];
            w[i] = sum * f[i * (i + 1) / 2 + i];
        }
        for (int64_t i = 0; i < size; ++i) out[b->dofs[begin + i]] = w[i];
    }
    free(w);
}

void lfo_block_stats(const lfo_block* b, double out[3]) {
    out[0] = (double)b->degenerate;
    out[1] = (double)b->factor_begin[b->n_sub];
    out[2] = b->capture_ratio;
}

void lfo_block_free(lfo_block* b) {
    if (!b) return;
    free(b->dof_begin); free(b->dofs); free(b->factor_begin); free(b->factor);
    free(b);
}

/* ------------------------------------------------------------------- PCG */

typedef struct {
    int kind;
    int64_t n;
    const double* inv_diag; /* jacobi */
    double* factor;         /* block1 */
    int64_t* lrow; int64_t* lcol; double* lval; double* ic_inv; /* ic0 */
    lfo_block* block;       /* subdomain block-Jacobi */
} precond_t;

static void precond_apply(const precond_t* m, const double* in, double* out) {
    if (m->kind == LFO_PRECOND_JACOBI) {
        for (int64_t i = 0; i < m->n; ++i) out[i] = m->inv_diag[i] * in[i]; /* solver.cpp:502-505 */
    } else if (m->kind == LFO_PRECOND_BLOCK1) {
        lfo_block1_apply(m->n, m->factor, in, out);
    } else if (m->kind == LFO_PRECOND_BLOCK) {
        lfo_block_apply(m->block, in, out);
    } else {
        lfo_ic0_apply(m->n, m->lrow, m->lcol, m->lval, m->ic_inv, in, out);
    }
}

int lfo_pcg(int64_t n, const int64_t* row_ptr, const int64_t* col, const double* values, const double* rhs,
            int precond, int stop_rule, double tolerance, int64_t max_iterations, double* x,
            lfo_pcg_report* report) {
    return lfo_pcg_ex(n, row_ptr, col, values, rhs, precond, LFO_STORAGE_CSR, NULL, 0, stop_rule, tolerance,
                      max_iterations, x, report);
}

/* solver.cpp:532-627 with the preconditioners and storage of solve_iterative
 * (:675-747); dof_subdomain / n_sub for LFO_PRECOND_BLOCK. */
int lfo_pcg_ex(int64_t n, const int64_t* row_ptr, const int64_t* col, const double* values, const double* rhs,
               int precond, int storage, const int64_t* dof_subdomain, int64_t n_sub, int stop_rule,
               double tolerance, int64_t max_iterations, double* x, lfo_pcg_report* report) {
    memset(report, 0, sizeof(*report));
    precond_t m;
    memset(&m, 0, sizeof(m));
    m.kind = precond;
    m.n = n;
    double* inv_diag = NULL;
    if (precond == LFO_PRECOND_JACOBI) {
        inv_diag = malloc((size_t)n * sizeof(double));
        for (int64_t row = 0; row < n; ++row) { /* solver.cpp:656-669 */
            const double d = values[value_index(row_ptr, col, row, row)];
            inv_diag[row] = fabs(d) > 1.0e-30 ? 1.0 / d : 1.0;
        }
        m.inv_diag = inv_diag;
    } else if (precond == LFO_PRECOND_BLOCK1) {
        m.factor = malloc((size_t)(3 * (n / 2) + 1) * sizeof(double));
        lfo_block1_factor(n, row_ptr, col, values, m.factor);
    } else if (precond == LFO_PRECOND_BLOCK) {
        m.block = lfo_block_create(n, row_ptr, col, values, dof_subdomain, n_sub);
    } else {
        const int64_t size = lfo_ic0_factor_size(n, row_ptr, col);
        m.lrow = malloc(((size_t)n + 1) * sizeof(int64_t));
        m.lcol = malloc((size_t)size * sizeof(int64_t));
        m.lval = malloc((size_t)size * sizeof(double));
        m.ic_inv = malloc((size_t)n * sizeof(double));
        lfo_ic0_factor(n, row_ptr, col, values, m.lrow, m.lcol, m.lval, m.ic_inv);
    }
    double* r = malloc((size_t)n * sizeof(double));
    double* z = malloc((size_t)n * sizeof(double));
    double* p = malloc((size_t)n * sizeof(double));
    double* q = malloc((size_t)n * sizeof(double));
    for (int64_t i = 0; i < n; ++i) { x[i] = 0.0; r[i] = rhs[i]; }
    precond_apply(&m, r, z);
    double rr = 0.0, rho = 0.0;
    for (int64_t i = 0; i < n; ++i) {
        const double v = r[i];
        rr += v * v;
        p[i] = z[i];
        rho += v * z[i];
    }
    report->initial_residual = sqrt(max0(rr));
    const double base = stop_rule == LFO_STOP_REFERENCE
                            ? ((1.0 < report->initial_residual) ? report->initial_residual : 1.0)
                            : report->initial_residual;
    const double target = tolerance * base;
    int status = 0;
    if (report->initial_residual <= target) {
        report->converged = 1;
        report->final_residual = report->initial_residual;
        goto done;
    }
    if (!(rho > 0.0) || !isfinite(rho)) { status = 1; goto done; }
    for (int64_t it = 1; it <= max_iterations; ++it) {
        if (storage == LFO_STORAGE_BSR) lfo_bsr_spmv(n, row_ptr, col, values, p, q);
        else lfo_spmv(n, row_ptr, col, values, p, q);
        double curvature = 0.0;
        for (int64_t i = 0; i < n; ++i) curvature += p[i] * q[i];
        if (!(curvature > 0.0) || !isfinite(curvature)) { status = 2; goto done; }
        const double alpha = rho / curvature;
        rr = 0.0;
        for (int64_t i = 0; i < n; ++i) {
            x[i] += alpha * p[i];
            const double v = r[i] - alpha * q[i];
            r[i] = v;
            rr += v * v;
        }
        report->iterations = it;
        report->final_residual = sqrt(max0(rr));
        if (report->final_residual <= target) { report->converged = 1; goto done; }
        precond_apply(&m, r, z);
        double next = 0.0;
        for (int64_t i = 0; i < n; ++i) next += r[i] * z[i];
        if (!(next > 0.0) || !isfinite(next)) { status = 3; goto done; }
        const double beta = next / rho;
        for (int64_t i = 0; i < n; ++i) p[i] = z[i] + beta * p[i];
        rho = next;
    }
    report->final_residual = sqrt(max0(rr));
done:
    report->status = status;
    free(r); free(z); free(p); free(q); free(inv_diag);
    free(m.factor); free(m.lrow); free(m.lcol); free(m.lval); free(m.ic_inv);
    lfo_block_free(m.block);
    return status;
}

/* ------------------------------------------------------------ smoothing */

/* solver.cpp:807-871 (Q4 part). gp layout per element: 4 x [xi, eta, e3, s3, vm]. */
void lfo_smooth_nodal(int64_t nnodes, int64_t nelem, const int64_t* enode_idx, const double* gp, double* nodal) {
    lfo_smooth_nodal_ex(nnodes, nelem, enode_idx, gp, 0, NULL, NULL, nodal);
}

/* solver.cpp:807-871 with bars: after the Q4 extrapolations, each bar adds its stress
 * tensor (3 per bar in tstress) to both of its nodes (tnode_idx, 2 per bar). */
void lfo_smooth_nodal_ex(int64_t nnodes, int64_t nelem, const int64_t* enode_idx, const double* gp, int64_t ntruss,
                         const int64_t* tnode_idx, const double* tstress, double* nodal) {
    const double root_three = 1.7320508075688772935;
    const double alpha = 0.5 * (1.0 + root_three);
    const double beta = 0.5 * (1.0 - root_three);
    const double ex[4][4] = {
        {alpha * alpha, alpha * beta, beta * beta, alpha * beta},
        {alpha * beta, alpha * alpha, alpha * beta, beta * beta},
        {beta * beta, alpha * beta, alpha * alpha, alpha * beta},
        {alpha * beta, beta * beta, alpha * beta, alpha * alpha},
    };
    double* sums = calloc((size_t)(3 * nnodes), sizeof(double));
    int64_t* counts = calloc((size_t)nnodes, sizeof(int64_t));
    for (int64_t e = 0; e < nelem; ++e)
        for (int l = 0; l < 4; ++l) {
            const int64_t node = enode_idx[4 * e + l];
            for (int c = 0; c < 3; ++c)
                for (int g = 0; g < 4; ++g) sums[3 * node + c] += ex[l][g] * gp[36 * e + 9 * g + 5 + c];
            counts[node]++;
        }
    for (int64_t t = 0; t < ntruss; ++t)
        for (int l = 0; l < 2; ++l) {
            const int64_t node = tnode_idx[2 * t + l];
            for (int c = 0; c < 3; ++c) sums[3 * node + c] += tstress[3 * t + c];
            counts[node]++;
        }
    for (int64_t node = 0; node < nnodes; ++node) {
        double* out = nodal + 4 * node;
        if (counts[node] == 0) { out[0] = out[1] = out[2] = out[3] = 0.0; continue; }
        for (int c = 0; c < 3; ++c) out[c] = sums[3 * node + c] / (double)counts[node];
        const double sx = out[0], sy = out[1], txy = out[2];
        out[3] = sqrt(max0(sx * sx - sx * sy + sy * sy + 3.0 * txy * txy));
    }
    free(sums);
    free(counts);
}
