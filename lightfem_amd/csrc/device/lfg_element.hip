// Q4 element kernels for gfx950: element stiffness with the three assembly
// strategies, and centre / Gauss-point field recovery with nodal smoothing.
//
// Arithmetic follows the reference's operation order term by term
// (element.cpp:16-161, 212-248; solver.cpp:135-179, 807-871) and the library is
// built with -ffp-contract=off, so Ke, serial/coloured K and the fields are
// bit-identical to the CPU reference. Ke is ~4.6 flop per byte written, so the
// kernels are memory-bound: one element is spread over 8 lanes (lane r owns Ke
// row r), which keeps registers low and makes each wave write 8 contiguous
// 512-byte Ke blocks or 512 scattered map entries per step.
#include "lfg_kernels.hpp"

namespace lfg {
namespace {

constexpr double kGauss = 0.57735026918962576451;  // element.cpp:132

struct Coords {
    double x[4];
    double y[4];
};

__host__ __device__ __forceinline__ void constitutive(double e, double nu, int plane_mode, double d[9]) {
    // element.cpp:16-32
    if (plane_mode == 0) {
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

// element.cpp:34-73: the shape-function gradients dN_i/dx, dN_i/dy and det J at (xi, eta);
// returns false on a non-positive Jacobian.
__host__ __device__ __forceinline__ bool shape_gradients(const Coords& c, double xi, double eta, double dnx[4],
                                                         double dny[4], double& det) {
    const double dxi[4] = {-0.25 * (1.0 - eta), +0.25 * (1.0 - eta), +0.25 * (1.0 + eta), -0.25 * (1.0 + eta)};
    const double deta[4] = {-0.25 * (1.0 - xi), -0.25 * (1.0 + xi), +0.25 * (1.0 + xi), +0.25 * (1.0 - xi)};
    double dx_dxi = 0.0, dx_deta = 0.0, dy_dxi = 0.0, dy_deta = 0.0;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        dx_dxi += dxi[i] * c.x[i];
        dx_deta += deta[i] * c.x[i];
        dy_dxi += dxi[i] * c.y[i];
        dy_deta += deta[i] * c.y[i];
    }
    det = dx_dxi * dy_deta - dx_deta * dy_dxi;
    const bool ok = det > 1.0e-14 && __builtin_isfinite(det);
    const double inv00 = dy_deta / det;
    const double inv01 = -dx_deta / det;
    const double inv10 = -dy_dxi / det;
    const double inv11 = dx_dxi / det;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        dnx[i] = inv00 * dxi[i] + inv10 * deta[i];
        dny[i] = inv01 * dxi[i] + inv11 * deta[i];
    }
    return ok;
}

// B (3x8 row-major) at (xi, eta): columns 2i = (dN_i/dx, 0, dN_i/dy), 2i+1 = (0, dN_i/dy, dN_i/dx).
__host__ __device__ __forceinline__ bool kinematics(const Coords& c, double xi, double eta, double b[24],
                                                    double& det) {
    double dnx[4], dny[4];
    const bool ok = shape_gradients(c, xi, eta, dnx, dny, det);
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        b[2 * i] = dnx[i];
        b[2 * i + 1] = 0.0;
        b[8 + 2 * i] = 0.0;
        b[8 + 2 * i + 1] = dny[i];
        b[16 + 2 * i] = dny[i];
        b[16 + 2 * i + 1] = dnx[i];
    }
    return ok;
}

__device__ __forceinline__ void load_element(const MeshView& m, int e, Coords& c, double& young, double& nu,
                                             double& thick, int& mode) {
    const int4 nodes = reinterpret_cast<const int4*>(m.enodes)[e];
    const int ids[4] = {nodes.x, nodes.y, nodes.z, nodes.w};
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const double2 p = reinterpret_cast<const double2*>(m.xy)[ids[i]];
        c.x[i] = p.x;
        c.y[i] = p.y;
    }
    const int mat = m.elem_mat[e];
    young = m.materials[3 * mat];
    nu = m.materials[3 * mat + 1];
    thick = m.materials[3 * mat + 2];
    mode = m.mat_mode ? m.mat_mode[mat] : m.plane_mode;
}

// Rows row0 .. row0 + NR - 1 of Ke (element.cpp:130-161):
// ke[j][c] = sum_gp (sum_s b[s][row0 + j] db[s][c]) * detJ * t; every row with the same
// operations whatever NR (rows of one Gauss point share its B and DB).
template <int NR>
__host__ __device__ __forceinline__ bool ke_rows(const Coords& c, double young, double nu, double thick, int plane_mode,
                                        int row0, double (&ke)[NR][8]) {
    double d[9];
    constitutive(young, nu, plane_mode, d);
#pragma unroll
    for (int j = 0; j < NR; ++j)
#pragma unroll
        for (int k = 0; k < 8; ++k) ke[j][k] = 0.0;
    bool ok = true;
    // Gauss points (-g,-g), (g,-g), (g,g), (-g,g) one at a time: unrolled, the four points'
    // B and DB do not fit the register file and spill
#pragma unroll 1
    for (int p = 0; p < 4; ++p) {
        const double xi = (p == 1 || p == 2) ? kGauss : -kGauss;
        const double eta = p >= 2 ? kGauss : -kGauss;
        double b[24], det;
        ok = kinematics(c, xi, eta, b, det) && ok;
        double db[24];
#pragma unroll
        for (int r = 0; r < 3; ++r)
#pragma unroll
            for (int col = 0; col < 8; ++col) {
                double acc = 0.0;
#pragma unroll
                for (int inner = 0; inner < 3; ++inner) acc += d[3 * r + inner] * b[8 * inner + col];
                db[8 * r + col] = acc;
            }
        const double weight = det * thick;
#pragma unroll
        for (int j = 0; j < NR; ++j) {
            // column `row` of B: selects over constant indices (an index by the lane's row
            // would move b to scratch memory)
            const int row = row0 + j;
            double b0 = b[0], b1 = b[8], b2 = b[16];
#pragma unroll
            for (int k = 1; k < 8; ++k) {
                b0 = row == k ? b[k] : b0;
                b1 = row == k ? b[8 + k] : b1;
                b2 = row == k ? b[16 + k] : b2;
            }
#pragma unroll
            for (int col = 0; col < 8; ++col) {
                double contribution = 0.0;
                contribution += b0 * db[col];
                contribution += b1 * db[8 + col];
                contribution += b2 * db[16 + col];
                ke[j][col] += contribution * weight;
            }
        }
    }
    return ok;
}

// Row `row` of Ke.
__device__ __forceinline__ bool ke_row(const Coords& c, double young, double nu, double thick, int plane_mode,
                                       int row, double ke[8]) {
    double one[1][8];
    const bool ok = ke_rows<1>(c, young, nu, thick, plane_mode, row, one);
#pragma unroll
    for (int k = 0; k < 8; ++k) ke[k] = one[0][k];
    return ok;
}

// Rows 2a, 2a + 1 of Ke with the operations of ke_rows<2>(..., 2a, ...) minus the products
// with the structural zeros of B and D (D's (0,2), (1,2), (2,0), (2,1) entries; B's zero in
// each column) and minus the additions of those sums' starting +0.0. An omitted product is
// an exact +-0 (its other factor is finite) and adding +-0 to a nonzero value returns it,
// so every intermediate value here equals ke_rows's, except that a zero may carry the other
// sign. Zero signs cannot reach the rows: each entry is a sum that starts from +0.0, and an
// IEEE sum is -0.0 only when both addends are -0.0, so an entry is never -0.0 and adding a
// zero of either sign leaves it unchanged. The rows are bit-identical to ke_rows's, with
// less than half its floating-point operations.
__host__ __device__ __forceinline__ bool ke_corner_rows(const Coords& c, double young, double nu, double thick,
                                                        int plane_mode, int a, double (&ke)[2][8]) {
    double d[9];
    constitutive(young, nu, plane_mode, d);
#pragma unroll
    for (int k = 0; k < 8; ++k) ke[0][k] = ke[1][k] = 0.0;
    bool ok = true;
#pragma unroll 1
    for (int p = 0; p < 4; ++p) {
        const double xi = (p == 1 || p == 2) ? kGauss : -kGauss;
        const double eta = p >= 2 ? kGauss : -kGauss;
        double dnx[4], dny[4], det;
        ok = shape_gradients(c, xi, eta, dnx, dny, det) && ok;
        const double weight = det * thick;
        // B column 2a = (ax, 0, ay), column 2a + 1 = (0, ay, ax)
        double ax = dnx[0], ay = dny[0];
#pragma unroll
        for (int k = 1; k < 4; ++k) {
            ax = a == k ? dnx[k] : ax;
            ay = a == k ? dny[k] : ay;
        }
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            // DB columns 2i and 2i + 1 (D rows (d0 d1 0), (d3 d4 0), (0 0 d8))
            const double db0x = d[0] * dnx[i];
            const double db1x = d[3] * dnx[i];
            const double db2x = d[8] * dny[i];
            const double db0y = d[1] * dny[i];
            const double db1y = d[4] * dny[i];
            const double db2y = d[8] * dnx[i];
            ke[0][2 * i] += (ax * db0x + ay * db2x) * weight;
            ke[0][2 * i + 1] += (ax * db0y + ay * db2y) * weight;
            ke[1][2 * i] += (ay * db1x + ax * db2x) * weight;
            ke[1][2 * i + 1] += (ay * db1y + ax * db2y) * weight;
        }
    }
    return ok;
}

template <int Mode>
__global__ __launch_bounds__(256) void ke_kernel(MeshView m, const int* __restrict__ elems, int count,
                                                 const int* __restrict__ emap, double* __restrict__ values,
                                                 double* __restrict__ ke_out, int* __restrict__ failures) {
    const long long tid = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
    const int slot = static_cast<int>(tid >> 3);
    const int row = static_cast<int>(tid & 7);
    if (slot >= count) return;
    const int e = elems ? elems[slot] : slot;
    Coords c;
    double young, nu, thick;
    int mode;
    load_element(m, e, c, young, nu, thick, mode);
    double ke[8];
    const bool ok = ke_row(c, young, nu, thick, mode, row, ke);
    if (!ok) {
        if (row == 0) atomicAdd(failures, 1);
#pragma unroll
        for (int k = 0; k < 8; ++k) ke[k] = 0.0;
    }
    const long long base = 64LL * e + 8 * row;
    if (Mode == KE_STORE) {
        double2* out = reinterpret_cast<double2*>(ke_out + base);
#pragma unroll
        for (int k = 0; k < 4; ++k) out[k] = make_double2(ke[2 * k], ke[2 * k + 1]);
    } else {
        const int4* map4 = reinterpret_cast<const int4*>(emap + base);
        const int4 m0 = map4[0], m1 = map4[1];
        const int idx[8] = {m0.x, m0.y, m0.z, m0.w, m1.x, m1.y, m1.z, m1.w};
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            if (Mode == KE_ATOMIC) {
                unsafeAtomicAdd(values + idx[k], ke[k]);  // global_atomic_add_f64
            } else {
                values[idx[k]] += ke[k];  // conflict-free within one colour
            }
        }
    }
}

// Owner-computes assembly of a Q4 mesh. A 4-lane group owns node i's rows 2i, 2i+1;
// its node list holds (element << 2 | corner a) in the order the values receive their
// contributions: element order is serial_assembly (solver.cpp:135-142), colour order is
// coloring_assembly (:163-179; at most one contribution per value and colour). Lane
// `slot` forms rows 2a, 2a + 1 of the element in list position 4j + slot (ke_rows: the
// bits of ke_kernel, one B and DB per Gauss point for both rows). Lane h < 2 of the group
// owns row 2i + h: it takes the four slots' rows with __shfl and adds them, in list
// order, into its own LDS copy of the row (touched by no other lane until the block
// barrier), onto 0.0. After the barrier the group writes both rows once, coalesced: no
// element matrix is stored, K needs no zeroing and no value is read back. Rows of up to
// kRowSlots values (positions are clamped into the row copy, so a map entry outside its
// row cannot write elsewhere).
constexpr int kRowSlots = kAssembleRowSlots;
constexpr int kAsmNodes = 256 / 4;

__global__ __launch_bounds__(256) void assemble_rows_kernel(MeshView m, int n_nodes, const int* __restrict__ rp,
                                                            const int* __restrict__ node_ptr,
                                                            const int* __restrict__ node_list,
                                                            const int* __restrict__ emap, double* __restrict__ values,
                                                            int* __restrict__ failures) {
    __shared__ double acc[kAsmNodes * 2 * kRowSlots];
    const int g = static_cast<int>(threadIdx.x) >> 2;
    const int slot = static_cast<int>(threadIdx.x) & 3;
    const int group_lane0 = static_cast<int>(threadIdx.x) & (kWave - 4);  // within the wavefront
    const int node = static_cast<int>(blockIdx.x) * kAsmNodes + g;
    const bool active = node < n_nodes;  // no early exit: every thread reaches the barrier
    const bool owner = slot < 2;         // lane h owns row 2i + h
    double* const mine = acc + (g * 2 + (slot & 1)) * kRowSlots;
    if (owner) {
        for (int k = 0; k < kRowSlots; ++k) mine[k] = 0.0;
    }
    int begin = 0, end = 0, row_begin0 = 0, row_begin1 = 0;
    if (active) {
        begin = node_ptr[node];
        end = node_ptr[node + 1];
        row_begin0 = rp[2 * node];
        row_begin1 = rp[2 * node + 1];
    }
    for (int s0 = begin; s0 < end; s0 += 4) {  // the same trip count for the 4 lanes of a group
        const int s = s0 + slot;
        double ke[2][8] = {{0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0}, {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0}};
        int pos0[8] = {0, 0, 0, 0, 0, 0, 0, 0}, pos1[8] = {0, 0, 0, 0, 0, 0, 0, 0};
        if (s < end) {
            const int code = node_list[s];
            const int e = code >> 2;
            const int a = code & 3;
            Coords c;
            double young, nu, thick;
            int mode;
            load_element(m, e, c, young, nu, thick, mode);
            if (!ke_rows<2>(c, young, nu, thick, mode, 2 * a, ke)) {
                if (a == 0) atomicAdd(failures, 1);  // once per element
#pragma unroll
                for (int k = 0; k < 8; ++k) ke[0][k] = ke[1][k] = 0.0;
            }
            const int4* map4 = reinterpret_cast<const int4*>(emap + 64LL * e + 16 * a);  // rows 2a, 2a + 1
            const int4 m0 = map4[0], m1 = map4[1], m2 = map4[2], m3 = map4[3];
            const int i0[8] = {m0.x, m0.y, m0.z, m0.w, m1.x, m1.y, m1.z, m1.w};
            const int i1[8] = {m2.x, m2.y, m2.z, m2.w, m3.x, m3.y, m3.z, m3.w};
#pragma unroll
            for (int k = 0; k < 8; ++k) {
                pos0[k] = min(max(i0[k] - row_begin0, 0), kRowSlots - 1);
                pos1[k] = min(max(i1[k] - row_begin1, 0), kRowSlots - 1);
            }
        }
#pragma unroll
        for (int q = 0; q < 4; ++q) {  // list order
            const int src = group_lane0 + q;
            double v0[8], v1[8];
            int p0[8], p1[8];
#pragma unroll
            for (int k = 0; k < 8; ++k) {
                v0[k] = __shfl(ke[0][k], src);
                v1[k] = __shfl(ke[1][k], src);
                p0[k] = __shfl(pos0[k], src);
                p1[k] = __shfl(pos1[k], src);
            }
            if (owner && s0 + q < end) {
#pragma unroll
                for (int k = 0; k < 8; ++k) mine[slot ? p1[k] : p0[k]] += slot ? v1[k] : v0[k];
            }
        }
    }
    __syncthreads();
    if (active) {
#pragma unroll
        for (int hh = 0; hh < 2; ++hh) {
            const int b = rp[2 * node + hh], len = rp[2 * node + hh + 1] - b;
            const double* row = acc + (g * 2 + hh) * kRowSlots;
            for (int k = slot; k < len; k += 4) values[b + k] = row[k];
        }
    }
}

// Owner-computes assembly with each (element, corner) row pair formed once per block: a
// block owns nodes n0 .. n0 + 63 and so rows 2n0 .. 2n0 + 127, whose (element, corner)
// slots are one contiguous range of the node lists (the same element and colour orders as
// above). Phase 1: one thread per slot forms rows 2a, 2a + 1 of its element's Ke
// (ke_corner_rows, the bits of ke_kernel) into LDS, column by column with the two rows'
// entries side by side, and keeps the element's four nodes. Phase 2: a half-wavefront per
// node, one lane per column dof of the node's rows (node-blocked pattern: rows 2n and
// 2n + 1 hold the same columns): the lane reads its column from the pattern and sums, for
// both rows at once, over the node's slots in list order, the entry of the corner whose
// node is the column node: the order in which serial_assembly (solver.cpp:135-142) and
// coloring_assembly (:163-179) add into the value, from 0.0. No element map is read (a
// value's position is where the pattern holds its column), every value is written once,
// coalesced, and the launch needs no zeroed K. The host uses it for node-blocked
// patterns whose rows hold at most kAssembleRowSlots values and whose every block fits
// kBlockSlots slots.
constexpr int kBlockNodes = kAssembleBlockNodes;
constexpr int kBlockSlots = kAssembleBlockSlots;
constexpr int kBlockStride = kBlockSlots + 1;  // LDS column stride: odd, so a node's columns differ in bank

// Phase 2's sum for column node j (column dof cp) of a node with cnt <= N slots from k0:
// each slot adds the entries of the corner whose node is j (+0.0 where none is, and for the
// unused places of a node with fewer slots: adding +0.0 to a sum that started from +0.0
// changes nothing); branch-free, so the reads of all the slots are in flight together.
// `repeated` is set when an element holds j at two corners (the caller's general loop).
template <int N>
__device__ __forceinline__ double2 block_row_sum(const int4* en_lds, const double2* ke_lds, int k0, int cnt, int j,
                                                 int cp, bool& repeated) {
    int at[N];
#pragma unroll
    for (int r = 0; r < N; ++r) {
        const int k = r < cnt ? k0 + r : kBlockSlots;
        const int4 en = en_lds[k];
        const int mx = en.x == j, my = en.y == j, mz = en.z == j, mw = en.w == j;
        repeated |= mx + my + mz + mw > 1;
        const int bsel = mx ? 0 : (my ? 1 : (mz ? 2 : 3));
        at[r] = (mx | my | mz | mw) ? (2 * bsel + cp) * kBlockStride + k : kBlockSlots;
    }
    double2 sum = make_double2(0.0, 0.0);
#pragma unroll
    for (int r = 0; r < N; ++r) {
        const double2 v = ke_lds[at[r]];
        sum.x += v.x;
        sum.y += v.y;
    }
    return sum;
}

__global__ __launch_bounds__(256) void assemble_block_rows_kernel(MeshView m, int n_nodes,
                                                                  const int* __restrict__ rp,
                                                                  const int* __restrict__ col,
                                                                  const int* __restrict__ node_ptr,
                                                                  const int* __restrict__ node_list,
                                                                  double* __restrict__ values,
                                                                  int* __restrict__ failures) {
    constexpr int kHalf = kWave / 2;                       // lanes per node (columns of a row)
    constexpr int kNodesPerRound = 256 / kHalf;            // nodes the block handles at once
    constexpr int kRounds = kBlockNodes / kNodesPerRound;  // per half-wavefront
    // column c of slot k at ke_lds[c * kStride + k]: a wavefront's phase-1 stores are
    // consecutive, and the odd stride puts a node's columns in different LDS banks
    constexpr int kStride = kBlockStride;
    constexpr int kZero = kBlockSlots;                     // (+0.0, +0.0) in column 0's spare place
    static_assert(kAssembleRowSlots <= kHalf, "a row's columns fit a half-wavefront");
    __shared__ double2 ke_lds[8 * kStride];          // slot k, column c: (row 2a, row 2a + 1)
    __shared__ int4 en_lds[kBlockSlots + 1];         // slot k: its element's nodes; a sentinel
    __shared__ int sp[kBlockNodes + 1];              // node list offsets of the block's nodes
    __shared__ int rps[2 * kBlockNodes + 1];         // row offsets of the block's rows
    const int t = static_cast<int>(threadIdx.x);
    const int n0 = static_cast<int>(blockIdx.x) * kBlockNodes;
    const int nb = min(kBlockNodes, n_nodes - n0);
    if (t <= nb) sp[t] = node_ptr[n0 + t];
    if (t <= 2 * nb) rps[t] = rp[2 * n0 + t];
    if (t == 0) {
        ke_lds[kZero] = make_double2(0.0, 0.0);
        en_lds[kBlockSlots] = make_int4(-2, -2, -2, -2);  // matches no column node
    }
    __syncthreads();
    // node g = t / kHalf + kNodesPerRound * q, column lane l = t % kHalf: the columns are
    // loaded now so their latency overlaps phase 1
    const int gl = t / kHalf;
    const int l = t % kHalf;
    int cols[kRounds];
#pragma unroll
    for (int q = 0; q < kRounds; ++q) {
        const int g = gl + kNodesPerRound * q;
        cols[q] = -1;
        if (g < nb && l < rps[2 * g + 1] - rps[2 * g]) cols[q] = __builtin_nontemporal_load(col + rps[2 * g] + l);
    }
    const int s0 = sp[0];
    const int ns = min(sp[nb] - s0, kBlockSlots);  // the host guarantees sp[nb] - s0 <= kBlockSlots
    for (int k = t; k < ns; k += 256) {
        const int code = node_list[s0 + k];
        const int e = code >> 2;
        const int a = code & 3;
        const int4 nodes = reinterpret_cast<const int4*>(m.enodes)[e];
        const int ids[4] = {nodes.x, nodes.y, nodes.z, nodes.w};
        Coords c;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const double2 p = reinterpret_cast<const double2*>(m.xy)[ids[i]];
            c.x[i] = p.x;
            c.y[i] = p.y;
        }
        const int mat = m.elem_mat[e];
        const double young = m.materials[3 * mat];
        const double nu = m.materials[3 * mat + 1];
        const double thick = m.materials[3 * mat + 2];
        const int mode = m.mat_mode ? m.mat_mode[mat] : m.plane_mode;
        double ke[2][8];
        if (!ke_corner_rows(c, young, nu, thick, mode, a, ke)) {
            if (a == 0) atomicAdd(failures, 1);  // once per element
#pragma unroll
            for (int j = 0; j < 8; ++j) ke[0][j] = ke[1][j] = 0.0;
        }
#pragma unroll
        for (int j = 0; j < 8; ++j) ke_lds[j * kStride + k] = make_double2(ke[0][j], ke[1][j]);
        en_lds[k] = nodes;
    }
    __syncthreads();
#pragma unroll 1
    for (int q = 0; q < kRounds; ++q) {
        const int g = gl + kNodesPerRound * q;
        if (g >= nb) break;  // the same for every lane of a half-wavefront; later rounds too
        const int cv = cols[q];
        const int k0 = sp[g] - s0;
        const int cnt = min(sp[g + 1] - s0, ns) - k0;
        const int j = cv >> 1;
        const int cp = cv & 1;
        // up to 4 slots (a structured mesh's nodes) or 6 branch-free; a node with more
        // slots, or an element that repeats a corner node, takes the general loop below
        constexpr int kFast = 6;
        bool repeated = false;
        double2 sum = make_double2(0.0, 0.0);
        if (cnt <= 4)
            sum = block_row_sum<4>(en_lds, ke_lds, k0, cnt, j, cp, repeated);
        else if (cnt <= kFast)
            sum = block_row_sum<kFast>(en_lds, ke_lds, k0, cnt, j, cp, repeated);
        if (cnt > kFast || repeated) {
            sum = make_double2(0.0, 0.0);
            for (int k = k0; k < k0 + cnt; ++k) {
                const int4 en = en_lds[k];
                const int ids[4] = {en.x, en.y, en.z, en.w};
#pragma unroll
                for (int bb = 0; bb < 4; ++bb) {
                    if (ids[bb] == j) {
                        const double2 v = ke_lds[(2 * bb + cp) * kStride + k];
                        sum.x += v.x;
                        sum.y += v.y;
                    }
                }
            }
        }
        if (cv >= 0) {
            values[rps[2 * g] + l] = sum.x;
            values[rps[2 * g + 1] + l] = sum.y;
        }
    }
}

// Atomic assembly (solver.cpp:144-161) of a mesh whose elements form a W x H grid numbered
// row by row (element e = ey W + ex with nodes n, n + 1, n + W + 2, n + W + 1 for
// n = ey (W + 1) + ex): a block takes a tile of 8 x 8 elements, one thread per (element,
// corner) forms rows 2a, 2a + 1 of Ke (ke_corner_rows) and adds them into an LDS copy of
// the rows of the tile's 9 x 9 nodes (LDS atomics, any order). A node whose every element
// lies in the tile (the 7 x 7 inside, more at the mesh edge) then has complete row sums and
// the block stores them; the rows of the tile's other nodes are added into K with one FP64
// atomic per touched value. That is about 11 atomics per element instead of 64; the sums
// differ from the serial order only by rounding, as the reference's atomic sums do.
constexpr int kTileE = 8;                  // elements per tile side
constexpr int kTileN = kTileE + 1;         // nodes per tile side
constexpr int kTileRow = kAssembleGridRow; // values per row (a structured Q4 row holds at most 18)

__global__ __launch_bounds__(256) void assemble_atomic_tiles_kernel(MeshView m, int grid_w, int grid_h,
                                                                    const int* __restrict__ rp,
                                                                    const int* __restrict__ emap,
                                                                    double* __restrict__ values,
                                                                    int* __restrict__ failures) {
    constexpr int kEntries = kTileN * kTileN * 2 * kTileRow;
    __shared__ double acc[kEntries];  // [local node][row h][position]
    const int t = static_cast<int>(threadIdx.x);
    const int tiles_x = (grid_w + kTileE - 1) / kTileE;
    const int tx0 = (static_cast<int>(blockIdx.x) % tiles_x) * kTileE;
    const int ty0 = (static_cast<int>(blockIdx.x) / tiles_x) * kTileE;
    const int row_len = grid_w + 1;
    for (int k = t; k < kEntries; k += 256) acc[k] = 0.0;
    __syncthreads();
    {
        const int el = t >> 2;
        const int a = t & 3;
        const int lx = el % kTileE, ly = el / kTileE;
        const int ex = tx0 + lx, ey = ty0 + ly;
        if (ex < grid_w && ey < grid_h) {
            const int e = ey * grid_w + ex;
            const int4 nodes = reinterpret_cast<const int4*>(m.enodes)[e];
            const int ids[4] = {nodes.x, nodes.y, nodes.z, nodes.w};
            Coords c;
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const double2 p = reinterpret_cast<const double2*>(m.xy)[ids[i]];
                c.x[i] = p.x;
                c.y[i] = p.y;
            }
            const int mat = m.elem_mat[e];
            const int mode = m.mat_mode ? m.mat_mode[mat] : m.plane_mode;
            double ke[2][8];
            if (!ke_corner_rows(c, m.materials[3 * mat], m.materials[3 * mat + 1], m.materials[3 * mat + 2], mode, a,
                                ke)) {
                if (a == 0) atomicAdd(failures, 1);  // once per element
#pragma unroll
                for (int j = 0; j < 8; ++j) ke[0][j] = ke[1][j] = 0.0;
            }
            const int dx = (a == 1 || a == 2) ? 1 : 0, dy = a >= 2 ? 1 : 0;
            const int ln = (ly + dy) * kTileN + lx + dx;
            const int node = ids[a];
            const int r0 = rp[2 * node], r1 = rp[2 * node + 1];
            const int4* map4 = reinterpret_cast<const int4*>(emap + 64LL * e + 16 * a);  // rows 2a, 2a + 1
            const int4 m0 = map4[0], m1 = map4[1], m2 = map4[2], m3 = map4[3];
            const int i0[8] = {m0.x, m0.y, m0.z, m0.w, m1.x, m1.y, m1.z, m1.w};
            const int i1[8] = {m2.x, m2.y, m2.z, m2.w, m3.x, m3.y, m3.z, m3.w};
            double* const row0 = acc + (2 * ln) * kTileRow;
            double* const row1 = row0 + kTileRow;
#pragma unroll
            for (int k = 0; k < 8; ++k) {
                atomicAdd(row0 + min(max(i0[k] - r0, 0), kTileRow - 1), ke[0][k]);
                atomicAdd(row1 + min(max(i1[k] - r1, 0), kTileRow - 1), ke[1][k]);
            }
        }
    }
    __syncthreads();
    for (int k = t; k < kEntries; k += 256) {
        const int ln = k / (2 * kTileRow);
        const int h = (k / kTileRow) & 1;
        const int pos = k % kTileRow;
        const int gx = tx0 + ln % kTileN, gy = ty0 + ln / kTileN;
        if (gx > grid_w || gy > grid_h) continue;
        const int node = gy * row_len + gx;
        const int rb = rp[2 * node + h];
        if (pos >= rp[2 * node + h + 1] - rb) continue;
        // every element of the node lies in this tile: its rows are complete here
        const bool own = max(gx - 1, 0) >= tx0 && min(gx, grid_w - 1) < tx0 + kTileE && max(gy - 1, 0) >= ty0 &&
                         min(gy, grid_h - 1) < ty0 + kTileE;
        const double v = acc[k];
        if (own) values[rb + pos] = v;
        else if (v != 0.0) unsafeAtomicAdd(values + rb + pos, v);
    }
}

// ---- bars (element.cpp:266-306)

// Bar t: end coordinates, E of its material and its area; the axial direction and
// length as the reference forms them. Returns false for a zero-length bar.
__device__ __forceinline__ bool bar_geometry(const MeshView& m, const BarView& b, int t, double& c, double& s,
                                             double& length, double& young, double& area) {
    const int n0 = b.nodes[2 * t], n1 = b.nodes[2 * t + 1];
    const double dx = m.xy[2 * n1] - m.xy[2 * n0];
    const double dy = m.xy[2 * n1 + 1] - m.xy[2 * n0 + 1];
    length = sqrt(dx * dx + dy * dy);
    young = m.materials[3 * b.mat[t]];
    area = b.area[t];
    if (!(length > 1.0e-12) || !isfinite(length)) return false;
    c = dx / length;
    s = dy / length;
    return true;
}

// EA/L [c^2 cs -c^2 -cs; cs s^2 -cs -s^2; ...] (element.cpp:266-286), one thread per bar.
template <int Mode>
__global__ __launch_bounds__(256) void bar_ke_kernel(MeshView m, BarView b, const int* __restrict__ bars, int count,
                                                     double* __restrict__ values, double* __restrict__ ke_out,
                                                     int* __restrict__ failures) {
    const int slot = blockIdx.x * blockDim.x + threadIdx.x;
    if (slot >= count) return;
    const int t = bars ? bars[slot] : slot;
    double c, sn, length, young, area;
    double k[16];
    if (!bar_geometry(m, b, t, c, sn, length, young, area)) {
        atomicAdd(failures, 1);
#pragma unroll
        for (int i = 0; i < 16; ++i) k[i] = 0.0;
    } else {
        const double scale = young * area / length;
        const double cc = c * c, ss = sn * sn, cs = c * sn;
        const double v[16] = {scale * cc,  scale * cs,  -scale * cc, -scale * cs, scale * cs,  scale * ss,
                              -scale * cs, -scale * ss, -scale * cc, -scale * cs, scale * cc,  scale * cs,
                              -scale * cs, -scale * ss, scale * cs,  scale * ss};
#pragma unroll
        for (int i = 0; i < 16; ++i) k[i] = v[i];
    }
    if (Mode == KE_STORE) {
#pragma unroll
        for (int i = 0; i < 16; ++i) ke_out[16LL * t + i] = k[i];
    } else {
#pragma unroll
        for (int i = 0; i < 16; ++i) {
            const int at = b.map[16 * t + i];
            if (Mode == KE_ATOMIC) unsafeAtomicAdd(values + at, k[i]);
            else values[at] += k[i];  // conflict-free within one colour
        }
    }
}

// Bar fields (element.cpp:288-306, solver.cpp:753-803): axial strain, the stress tensor
// sigma (c^2, s^2, cs), von Mises |sigma|; every Gauss point carries the same state at
// xi = eta = 0. Rows n_q4 + t of the centre / Gauss-point arrays.
__global__ __launch_bounds__(256) void bar_fields_kernel(MeshView m, BarView b, const double* __restrict__ u,
                                                         double* __restrict__ centre, double* __restrict__ gp) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= b.n_bars) return;
    double c, sn, length, young, area;
    double f[7] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    if (bar_geometry(m, b, t, c, sn, length, young, area)) {
        const int n0 = b.nodes[2 * t], n1 = b.nodes[2 * t + 1];
        const double du = u[2 * n1] - u[2 * n0];
        const double dv = u[2 * n1 + 1] - u[2 * n0 + 1];
        const double strain = (du * c + dv * sn) / length;
        const double sigma = young * strain;
        f[0] = strain;
        f[3] = sigma * c * c;
        f[4] = sigma * sn * sn;
        f[5] = sigma * c * sn;
        f[6] = fabs(sigma);
    }
    const long long e = static_cast<long long>(b.n_q4) + t;
#pragma unroll
    for (int i = 0; i < 7; ++i) centre[7 * e + i] = f[i];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        double* g = gp + 36 * e + 9 * q;
        g[0] = 0.0;
        g[1] = 0.0;
#pragma unroll
        for (int i = 0; i < 7; ++i) g[2 + i] = f[i];
    }
}

// Serial-order assembly as a gather (solver.cpp:135-142): each CSR value sums its
// contributions in ascending element order starting from 0.0.
__global__ __launch_bounds__(256) void gather_kernel(long long nnz, const int* __restrict__ cptr,
                                                     const int* __restrict__ contrib, const double* __restrict__ ke,
                                                     double* __restrict__ values) {
    const long long k = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (k >= nnz) return;
    double sum = 0.0;
    for (int j = cptr[k]; j < cptr[k + 1]; ++j) sum += ke[contrib[j]];
    values[k] = sum;
}

__global__ void fill_kernel(double* p, long long n, double v) {
    const long long i = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// element.cpp:75-95, 212-248: strain = B u, stress = D strain, in-plane von Mises.
__device__ __forceinline__ void field_at(const double b[24], const double d[9], const double u[8], double* out) {
    double strain[3] = {0.0, 0.0, 0.0};
    double stress[3] = {0.0, 0.0, 0.0};
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int c = 0; c < 8; ++c) strain[r] += b[8 * r + c] * u[c];
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int c = 0; c < 3; ++c) stress[r] += d[3 * r + c] * strain[c];
    const double sx = stress[0], sy = stress[1], txy = stress[2];
    out[0] = strain[0];
    out[1] = strain[1];
    out[2] = strain[2];
    out[3] = sx;
    out[4] = sy;
    out[5] = txy;
    out[6] = sqrt(max0(sx * sx - sx * sy + sy * sy + 3.0 * txy * txy));
}

// One thread per element: centre field (7 doubles) and 4 Gauss points (9 doubles each).
__global__ __launch_bounds__(128) void fields_kernel(MeshView m, const double* __restrict__ u,
                                                     double* __restrict__ centre, double* __restrict__ gp) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= m.n_elems) return;
    Coords c;
    double young, nu, thick;
    int mode;
    load_element(m, e, c, young, nu, thick, mode);
    const int4 nodes = reinterpret_cast<const int4*>(m.enodes)[e];
    const int ids[4] = {nodes.x, nodes.y, nodes.z, nodes.w};
    double ue[8];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const double2 v = reinterpret_cast<const double2*>(u)[ids[i]];
        ue[2 * i] = v.x;
        ue[2 * i + 1] = v.y;
    }
    double d[9], b[24], det;
    constitutive(young, nu, mode, d);
    double out[7];
    kinematics(c, 0.0, 0.0, b, det);
    field_at(b, d, ue, out);
#pragma unroll
    for (int k = 0; k < 7; ++k) centre[7LL * e + k] = out[k];
    const double points[4][2] = {{-kGauss, -kGauss}, {kGauss, -kGauss}, {kGauss, kGauss}, {-kGauss, kGauss}};
#pragma unroll
    for (int p = 0; p < 4; ++p) {
        kinematics(c, points[p][0], points[p][1], b, det);
        double* g = gp + 36LL * e + 9 * p;
        g[0] = points[p][0];
        g[1] = points[p][1];
        field_at(b, d, ue, out);
#pragma unroll
        for (int k = 0; k < 7; ++k) g[2 + k] = out[k];
    }
}

// solver.cpp:807-871 as a gather: node i sums extrapolated Gauss-point stresses of
// its elements in ascending element order (the order the reference's element
// loop adds them), then averages and forms the nodal von Mises stress.
__global__ __launch_bounds__(256) void smooth_kernel(int n_nodes, int n_q4, const int* __restrict__ ptr,
                                                     const int* __restrict__ list, const double* __restrict__ gp,
                                                     double* __restrict__ nodal) {
    const int node = blockIdx.x * blockDim.x + threadIdx.x;
    if (node >= n_nodes) return;
    const double root_three = 1.7320508075688772935;
    const double alpha = 0.5 * (1.0 + root_three);
    const double beta = 0.5 * (1.0 - root_three);
    const double ex[4][4] = {
        {alpha * alpha, alpha * beta, beta * beta, alpha * beta},
        {alpha * beta, alpha * alpha, alpha * beta, beta * beta},
        {beta * beta, alpha * beta, alpha * alpha, alpha * beta},
        {alpha * beta, beta * beta, alpha * beta, alpha * alpha},
    };
    double sums[3] = {0.0, 0.0, 0.0};
    const int begin = ptr[node], end = ptr[node + 1];
    for (int j = begin; j < end; ++j) {
        const int entry = list[j];
        const int e = entry >> 2;
        const int local = entry & 3;
        const double* g = gp + 36LL * e;
        if (e >= n_q4) {  // a bar adds its stress tensor once (solver.cpp:841-851)
#pragma unroll
            for (int comp = 0; comp < 3; ++comp) sums[comp] += g[5 + comp];
            continue;
        }
#pragma unroll
        for (int comp = 0; comp < 3; ++comp)
#pragma unroll
            for (int q = 0; q < 4; ++q) sums[comp] += ex[local][q] * g[9 * q + 5 + comp];
    }
    double* out = nodal + 4LL * node;
    const int count = end - begin;
    if (count == 0) {
        out[0] = out[1] = out[2] = out[3] = 0.0;
        return;
    }
    for (int comp = 0; comp < 3; ++comp) out[comp] = sums[comp] / static_cast<double>(count);
    const double sx = out[0], sy = out[1], txy = out[2];
    out[3] = sqrt(max0(sx * sx - sx * sy + sy * sy + 3.0 * txy * txy));
}

// Element sensitivities of J(u(E)) with K(E) u = f: dJ/dE_e = -lambda_e^T (dKe/dE) u_e,
// where K lambda = dJ/du (lambda zero on constrained dofs). Ke is linear in E
// (E enters only through D, element.cpp:16-32), so dKe/dE = Ke evaluated with
// E = 1. Lane r of an element's 8-lane group forms row r of Ke(1) (ke_row, the
// reference's operation order), dots it with u_e in column order and weights it
// by lambda_e[r]; the group then sums its 8 terms pairwise (lanes r^1, r^2, r^4),
// grad_e = -sum * dE_e (dE_e = dE/drho_e, or 1 without dE).
__global__ __launch_bounds__(256) void sensitivity_kernel(MeshView m, const double* __restrict__ u,
                                                          const double* __restrict__ lam,
                                                          const double* __restrict__ dE, double* __restrict__ grad) {
    const long long tid = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
    const int e = static_cast<int>(tid >> 3);
    const int row = static_cast<int>(tid & 7);
    if (e >= m.n_elems) return;  // whole 8-lane groups leave together
    Coords c;
    double young, nu, thick;
    int mode;
    load_element(m, e, c, young, nu, thick, mode);
    double ke[8];
    ke_row(c, 1.0, nu, thick, mode, row, ke);
    const int4 nodes = reinterpret_cast<const int4*>(m.enodes)[e];
    const int ids[4] = {nodes.x, nodes.y, nodes.z, nodes.w};
    double ku = 0.0;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const double2 v = reinterpret_cast<const double2*>(u)[ids[i]];
        ku += ke[2 * i] * v.x;
        ku += ke[2 * i + 1] * v.y;
    }
    const int my_node = ids[row >> 1];
    double t = lam[2 * my_node + (row & 1)] * ku;
    t = t + __shfl_xor(t, 1, 8);
    t = t + __shfl_xor(t, 2, 8);
    t = t + __shfl_xor(t, 4, 8);
    if (row == 0) grad[e] = -t * (dE ? dE[e] : 1.0);
}

inline unsigned int blocks_for(long long work, int per_block) {
    return static_cast<unsigned int>((work + per_block - 1) / per_block);
}

}  // namespace

void launch_ke(hipStream_t s, const MeshView& m, KeMode mode, const int* elems, int count, const int* emap,
               double* values, double* ke_out, int* failures) {
    if (count <= 0) return;
    const unsigned int grid = blocks_for(8LL * count, 256);
    switch (mode) {
        case KE_STORE:
            ke_kernel<KE_STORE><<<grid, 256, 0, s>>>(m, elems, count, emap, values, ke_out, failures);
            break;
        case KE_ATOMIC:
            ke_kernel<KE_ATOMIC><<<grid, 256, 0, s>>>(m, elems, count, emap, values, ke_out, failures);
            break;
        case KE_COLORED:
            ke_kernel<KE_COLORED><<<grid, 256, 0, s>>>(m, elems, count, emap, values, ke_out, failures);
            break;
    }
}

void launch_assemble_rows(hipStream_t s, const MeshView& m, int n_nodes, const int* row_ptr, const int* node_ptr,
                          const int* node_list, const int* emap, double* values, int* jacobian_failures) {
    if (n_nodes <= 0) return;
    assemble_rows_kernel<<<blocks_for(n_nodes, kAsmNodes), 256, 0, s>>>(m, n_nodes, row_ptr, node_ptr, node_list,
                                                                        emap, values, jacobian_failures);
}

void launch_assemble_block_rows(hipStream_t s, const MeshView& m, int n_nodes, const int* row_ptr, const int* col,
                                const int* node_ptr, const int* node_list, double* values, int* jacobian_failures) {
    if (n_nodes <= 0) return;
    assemble_block_rows_kernel<<<blocks_for(n_nodes, kBlockNodes), 256, 0, s>>>(m, n_nodes, row_ptr, col, node_ptr,
                                                                               node_list, values, jacobian_failures);
}

void launch_assemble_atomic_tiles(hipStream_t s, const MeshView& m, int grid_w, int grid_h, const int* row_ptr,
                                  const int* emap, double* values, int* jacobian_failures) {
    if (grid_w <= 0 || grid_h <= 0) return;
    const unsigned int tiles = static_cast<unsigned int>(((grid_w + kTileE - 1) / kTileE) *
                                                         ((grid_h + kTileE - 1) / kTileE));
    assemble_atomic_tiles_kernel<<<tiles, 256, 0, s>>>(m, grid_w, grid_h, row_ptr, emap, values, jacobian_failures);
}

void launch_gather(hipStream_t s, long long nnz, const int* cptr, const int* contrib, const double* ke,
                   double* values) {
    if (nnz <= 0) return;
    gather_kernel<<<blocks_for(nnz, 256), 256, 0, s>>>(nnz, cptr, contrib, ke, values);
}

void launch_fill(hipStream_t s, double* p, long long n, double v) {
    if (n <= 0) return;
    fill_kernel<<<blocks_for(n, 256), 256, 0, s>>>(p, n, v);
}

void launch_fields(hipStream_t s, const MeshView& m, const double* u, double* centre, double* gp) {
    if (m.n_elems <= 0) return;
    fields_kernel<<<blocks_for(m.n_elems, 128), 128, 0, s>>>(m, u, centre, gp);
}

void launch_sensitivity(hipStream_t s, const MeshView& m, const double* u, const double* lam, const double* dE,
                        double* grad) {
    if (m.n_elems <= 0) return;
    sensitivity_kernel<<<blocks_for(8LL * m.n_elems, 256), 256, 0, s>>>(m, u, lam, dE, grad);
}

void launch_smooth(hipStream_t s, int n_nodes, int n_q4, const int* ptr, const int* list, const double* gp,
                   double* nodal) {
    if (n_nodes <= 0) return;
    smooth_kernel<<<blocks_for(n_nodes, 256), 256, 0, s>>>(n_nodes, n_q4, ptr, list, gp, nodal);
}

void launch_bar_ke(hipStream_t s, const MeshView& m, const BarView& b, KeMode mode, const int* bars, int count,
                   double* values, double* ke_out, int* failures) {
    if (count <= 0) return;
    const unsigned int grid = blocks_for(count, 256);
    switch (mode) {
        case KE_STORE: bar_ke_kernel<KE_STORE><<<grid, 256, 0, s>>>(m, b, bars, count, values, ke_out, failures); break;
        case KE_ATOMIC: bar_ke_kernel<KE_ATOMIC><<<grid, 256, 0, s>>>(m, b, bars, count, values, ke_out, failures); break;
        case KE_COLORED:
            bar_ke_kernel<KE_COLORED><<<grid, 256, 0, s>>>(m, b, bars, count, values, ke_out, failures);
            break;
    }
}

void launch_bar_fields(hipStream_t s, const MeshView& m, const BarView& b, const double* u, double* centre,
                       double* gp) {
    if (b.n_bars <= 0) return;
    bar_fields_kernel<<<blocks_for(b.n_bars, 256), 256, 0, s>>>(m, b, u, centre, gp);
}

}  // namespace lfg
