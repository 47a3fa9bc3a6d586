// Node-stencil SpMV of the PCG iteration on gfx950 (see StencilView).
//
// On a structured mesh numbered row by row, node i couples only with nodes i + d for
// a handful of offsets d, so the multiply needs no column indices: plane k of the
// layout holds block (i, i + off[k]) for every node i, a wavefront's 64 lanes own 64
// consecutive nodes, and every value load and every p load of the wavefront moves
// 64 consecutive pairs. The symmetric form stores only the diagonal and the positive
// offsets and reads block (i, i - d) as the transpose of block (i - d, i): the
// second read of a plane trails the first by d nodes and is served from the XCD's
// L2 (each XCD walks one contiguous node range), so HBM sees each value once.
#include <stdexcept>

#include "lfg_kernels.hpp"
#include "lfg_precond.hpp"
#include "lfg_slices.hpp"

#include "lightfem_gpu.h"

namespace lfg {
namespace {

constexpr int kThreads = 256;

inline unsigned int blocks_for(long long work, int per_block) {
    return static_cast<unsigned int>((work + per_block - 1) / per_block);
}

inline unsigned int round8(unsigned int b) { return (b + 7u) & ~7u; }

// Block (i, j) of the CSR matrix: rows 2i and 2i+1 share one column list of dof pairs
// (sorted), so the pair of column 2j sits at the same offset in both rows. A node in
// no element has only its two diagonal entries. Returns false when the pattern has
// no such block.
struct Block {
    double a, b, c, d;  // (K[2i][2j], K[2i][2j+1]; K[2i+1][2j], K[2i+1][2j+1])
};

__device__ bool load_block(const int* __restrict__ rp, const int* __restrict__ col, const double* __restrict__ vals,
                           int i, int j, Block& out) {
    const int r0 = rp[2 * i], r1 = rp[2 * i + 1];
    if (r1 - r0 == 1) {
        if (j != i) return false;
        out = Block{vals[r0], 0.0, 0.0, vals[r1]};
        return true;
    }
    int lo = 0, hi = (r1 - r0) >> 1;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if ((col[r0 + 2 * mid] >> 1) < j) lo = mid + 1;
        else hi = mid;
    }
    if (lo >= ((r1 - r0) >> 1) || (col[r0 + 2 * lo] >> 1) != j) return false;
    const int t = 2 * lo;
    out = Block{vals[r0 + t], vals[r0 + t + 1], vals[r1 + t], vals[r1 + t + 1]};
    return true;
}

__device__ __forceinline__ float f32(double v) { return static_cast<float>(v); }

// Pair index of half h (0 upper, 1 lower) of plane k for node i in an array of np
// planes, stored per 64-node chunk: chunk c holds plane 0 upper, plane 0 lower,
// plane 1 upper, ... as 64 consecutive pairs each, so a wavefront's loads for one
// chunk walk one contiguous 128 * np-pair run.
__device__ __forceinline__ long long plane_at(int i, int k, int h, int np) {
    return ((static_cast<long long>(i >> 6) * np + k) * 2 + h) * kWave + (i & (kWave - 1));
}

typedef float f4 __attribute__((ext_vector_type(4)));

// FP32 planes hold one 16-byte record per node and plane, the upper pair then the lower
// pair, stored per 64-node chunk plane after plane: one load brings a node's block and a
// wavefront's load walks 64 consecutive records. Record index of plane k, node i:
__device__ __forceinline__ long long plane32_at(int i, int k, int np) {
    return (static_cast<long long>(i >> 6) * np + k) * kWave + (i & (kWave - 1));
}

// Upper and lower pairs of plane k at node i (FP64: two 16-byte loads; FP32: one);
// kNT marks the load as the data's last use.
template <bool kNT>
__device__ __forceinline__ void plane_block(const double* planes, int i, int k, int np, dpair& u, dpair& l) {
    const dpair* pl = reinterpret_cast<const dpair*>(planes);
    if (kNT) {
        u = __builtin_nontemporal_load(pl + plane_at(i, k, 0, np));
        l = __builtin_nontemporal_load(pl + plane_at(i, k, 1, np));
    } else {
        u = pl[plane_at(i, k, 0, np)];
        l = pl[plane_at(i, k, 1, np)];
    }
}

template <bool kNT>
__device__ __forceinline__ void plane_block(const float* planes, int i, int k, int np, fpair& u, fpair& l) {
    const f4* pl = reinterpret_cast<const f4*>(planes);
    const f4 v = kNT ? __builtin_nontemporal_load(pl + plane32_at(i, k, np)) : pl[plane32_at(i, k, np)];
    u = fpair{v.x, v.y};
    l = fpair{v.z, v.w};
}

// One thread per node: the node's block in every plane; with diag_c, the diagonal
// block minus the rounding error (FP32 value - FP64 value) of each off-diagonal
// block the FP32 multiply uses for this node's rows, in the row's column order:
// blocks to the right are the node's own rounded values, blocks to the left are the
// rounded transposes of the neighbours' blocks.
__global__ __launch_bounds__(256) void pack_stencil_kernel(int n_nodes, const int* __restrict__ rp,
                                                           const int* __restrict__ col,
                                                           const double* __restrict__ vals, StencilView v,
                                                           double* __restrict__ out64, float* __restrict__ out32,
                                                           double* __restrict__ diag_c) {
    const int i = static_cast<int>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n_nodes) return;
    const int np64 = v.n_off, np32 = v.n_off - 1;
    for (int k = 0; k < v.n_off; ++k) {
        const int j = i + v.off[k];
        Block blk{0.0, 0.0, 0.0, 0.0};
        if (j >= 0 && j < n_nodes) load_block(rp, col, vals, i, j, blk);
        if (out64) {
            dpair* o = reinterpret_cast<dpair*>(out64);
            o[plane_at(i, k, 0, np64)] = dpair{blk.a, blk.b};
            o[plane_at(i, k, 1, np64)] = dpair{blk.c, blk.d};
        }
        if (out32 && k >= 1)
            reinterpret_cast<f4*>(out32)[plane32_at(i, k - 1, np32)] = f4{f32(blk.a), f32(blk.b), f32(blk.c), f32(blk.d)};
    }
    if (!diag_c) return;
    Block d{0.0, 0.0, 0.0, 0.0};
    load_block(rp, col, vals, i, i, d);
    const int r0 = rp[2 * i], r1 = rp[2 * i + 1];
    if (r1 - r0 > 1) {
        for (int t = 0; t < r1 - r0; t += 2) {
            const int j = col[r0 + t] >> 1;
            if (j == i) continue;
            const double ea = vals[r0 + t], eb = vals[r0 + t + 1], ec = vals[r1 + t], ed = vals[r1 + t + 1];
            double ua, ub, uc, ud;  // the block the FP32 multiply applies at (i, j)
            if (j > i) {
                ua = f32(ea);
                ub = f32(eb);
                uc = f32(ec);
                ud = f32(ed);
            } else {
                Block tr{0.0, 0.0, 0.0, 0.0};
                load_block(rp, col, vals, j, i, tr);
                ua = f32(tr.a);
                ub = f32(tr.c);
                uc = f32(tr.b);
                ud = f32(tr.d);
            }
            d.a -= ua - ea;
            d.b -= ub - eb;
            d.c -= uc - ec;
            d.d -= ud - ed;
        }
    }
    dpair* o = reinterpret_cast<dpair*>(diag_c);
    o[plane_at(i, 0, 0, 1)] = dpair{d.a, d.b};
    o[plane_at(i, 0, 1, 1)] = dpair{d.c, d.d};
}

// Block (i, i - d) = transpose of block (i - d, i), whose upper / lower pairs are lu, ll.
template <typename P>
__device__ __forceinline__ void add_lower(P lu, P ll, dpair x, double& a0, double& a1) {
    a0 += static_cast<double>(lu.x) * x.x;
    a0 += static_cast<double>(ll.x) * x.y;
    a1 += static_cast<double>(lu.y) * x.x;
    a1 += static_cast<double>(ll.y) * x.y;
}

template <typename P>
__device__ __forceinline__ void add_upper(P uu, P ul, dpair x, double& a0, double& a1) {
    a0 += static_cast<double>(uu.x) * x.x;
    a0 += static_cast<double>(uu.y) * x.y;
    a1 += static_cast<double>(ul.x) * x.x;
    a1 += static_cast<double>(ul.y) * x.y;
}

// Symmetric form, node i: rows 2i and 2i+1 accumulate the transposed blocks to the
// left (farthest first), the diagonal block, then the blocks to the right, i.e. in
// column order. The diagonal block is plane 0 of `diag` (np_d planes); the block of
// offset off[k + 1] is plane kb + k of `planes` (np planes). kM > 0: exactly kM
// positive offsets, every load issued before the arithmetic; kM = 0: any count.
// kLds: the multiplied vector comes from a tile's LDS windows (node i = i0 + t):
// node i + d at win[slot + t] (see StencilView); otherwise from p2[i + d].
template <typename VT, int kM, bool kLds>
__device__ __forceinline__ void stencil_sym_node(const StencilView& v, const dpair* __restrict__ diag, int np_d,
                                                 const VT* __restrict__ planes, int np, int kb,
                                                 const dpair* __restrict__ p2, int i, double& a0, double& a1,
                                                 const dpair* win, int t) {
    using P = typename PairOf<VT>::type;
    const dpair du = __builtin_nontemporal_load(diag + plane_at(i, 0, 0, np_d));
    const dpair dl = __builtin_nontemporal_load(diag + plane_at(i, 0, 1, np_d));
    const dpair xi = kLds ? win[v.slot_zero + t] : p2[i];
    a0 = 0.0;
    a1 = 0.0;
    if (kM > 0) {
        constexpr int M = kM > 0 ? kM : 1;
        P lu[M], ll[M], uu[M], ul[M];
        dpair xl[M], xu[M];
        bool okl[M];
#pragma unroll
        for (int k = 0; k < M; ++k) {
            const int d = v.off[k + 1];
            const int jl = i - d;
            okl[k] = jl >= 0;
            const int jlc = okl[k] ? jl : i;
            const int ju = min(i + d, v.n_nodes - 1);
            plane_block<false>(planes, i, kb + k, np, uu[k], ul[k]);
            plane_block<true>(planes, jlc, kb + k, np, lu[k], ll[k]);
            xl[k] = kLds ? win[v.slot_lo[k + 1] + t] : p2[jlc];
            xu[k] = kLds ? win[v.slot_up[k + 1] + t] : p2[ju];
        }
#pragma unroll
        for (int k = M - 1; k >= 0; --k)
            if (okl[k]) add_lower(lu[k], ll[k], xl[k], a0, a1);
        add_upper(du, dl, xi, a0, a1);
#pragma unroll
        for (int k = 0; k < M; ++k) add_upper(uu[k], ul[k], xu[k], a0, a1);
    } else {
        const int m = v.n_off - 1;
        for (int k = m - 1; k >= 0; --k) {
            const int jl = i - v.off[k + 1];
            if (jl < 0) continue;
            const dpair x = kLds ? win[v.slot_lo[k + 1] + t] : p2[jl];
            P bu, bl;
            plane_block<false>(planes, jl, kb + k, np, bu, bl);
            add_lower(bu, bl, x, a0, a1);
        }
        add_upper(du, dl, xi, a0, a1);
        for (int k = 0; k < m; ++k) {
            const int ju = min(i + v.off[k + 1], v.n_nodes - 1);
            const dpair x = kLds ? win[v.slot_up[k + 1] + t] : p2[ju];
            P bu, bl;
            plane_block<false>(planes, i, kb + k, np, bu, bl);
            add_upper(bu, bl, x, a0, a1);
        }
    }
}

// Exact form, node i: every plane in ascending offset order (the CSR column order,
// or the block storage order with v.pair_sum); the zero blocks of absent couplings
// add nothing. kN > 0: exactly kN planes, loads
// first; kN = 0: any count.
template <int kN>
__device__ __forceinline__ void stencil_exact_node(const StencilView& v, const double* __restrict__ planes,
                                                   const dpair* __restrict__ p2, int i, double& a0, double& a1) {
    const dpair* __restrict__ pl = reinterpret_cast<const dpair*>(planes);
    const bool pair = v.pair_sum != 0;
    auto term = [&](dpair u, dpair l, dpair x) { add_block(a0, a1, u.x, u.y, l.x, l.y, x, pair); };
    a0 = 0.0;
    a1 = 0.0;
    if (kN > 0) {
        constexpr int N = kN > 0 ? kN : 1;
        dpair u[N], l[N], x[N];
#pragma unroll
        for (int k = 0; k < N; ++k) {
            const int j = min(max(i + v.off[k], 0), v.n_nodes - 1);
            u[k] = __builtin_nontemporal_load(pl + plane_at(i, k, 0, N));
            l[k] = __builtin_nontemporal_load(pl + plane_at(i, k, 1, N));
            x[k] = p2[j];
        }
#pragma unroll
        for (int k = 0; k < N; ++k) term(u[k], l[k], x[k]);
    } else {
        for (int k = 0; k < v.n_off; ++k) {
            const int j = min(max(i + v.off[k], 0), v.n_nodes - 1);
            term(pl[plane_at(i, k, 0, v.n_off)], pl[plane_at(i, k, 1, v.n_off)], p2[j]);
        }
    }
}

// q = A p over the stencil planes, fused p.q and alpha = rho / p.q (solver.cpp:589-597).
// Each block owns a contiguous node range (XCD-remapped), one node per thread and step.
// kSym: symmetric form with kM positive offsets; otherwise the exact form with kM planes.
template <int kM, bool kSym>
__global__ __launch_bounds__(kThreads) void pcg_spmv_stencil_kernel(PcgBuffers b, StencilView v,
                                                                    const double* __restrict__ planes64,
                                                                    const float* __restrict__ planes32,
                                                                    const double* __restrict__ diag_c) {
    if (b.st->done) return;
    __shared__ double sh[kThreads / kWave];
    const bool use32 = kSym && planes32 != nullptr && (planes64 == nullptr || b.st->phase32 != 0);
    const int lb = xcd_block(blockIdx.x, gridDim.x);
    const int per_block = ((v.n_nodes + gridDim.x - 1) / gridDim.x + kWave - 1) / kWave * kWave;
    const int start = lb * per_block;
    const int stop = min(start + per_block, v.n_nodes);
    const dpair* __restrict__ p2 = reinterpret_cast<const dpair*>(b.p);
    dpair* __restrict__ q2 = reinterpret_cast<dpair*>(b.q);
    double pq = 0.0;
    for (int i = start + static_cast<int>(threadIdx.x); i < stop; i += kThreads) {
        double a0, a1;
        if (kSym) {
            if (use32)
                stencil_sym_node<float, kM, false>(v, reinterpret_cast<const dpair*>(diag_c), 1, planes32,
                                                   v.n_off - 1, 0, p2, i, a0, a1, nullptr, 0);
            else
                stencil_sym_node<double, kM, false>(v, reinterpret_cast<const dpair*>(planes64), v.n_off, planes64,
                                                    v.n_off, 1, p2, i, a0, a1, nullptr, 0);
        } else {
            stencil_exact_node<kM>(v, planes64, p2, i, a0, a1);
        }
        q2[i] = dpair{a0, a1};
        const dpair pv = p2[i];
        pq += pv.x * a0 + pv.y * a1;
    }
    const double vals[1] = {block_sum<kThreads>(pq, sh)};
    double tot[1];
    if (!finish_reduction<kThreads, 1>(b.tickets + 1, b.partials, vals, tot, sh)) return;
    if (threadIdx.x == 0) {
        const double tpq = tot[0];
        PcgState* st = b.st;
        st->pq = tpq;
        if (!(tpq > 0.0) || !isfinite(tpq)) {  // solver.cpp:591-593
            st->status = LFG_ERR_NOT_SPD;
            st->done = 1;
        } else {
            st->alpha = st->rho / tpq;
        }
    }
}

// K sums over a block of NT threads in one LDS round; totals in v of thread 0.
template <int NT, int K>
__device__ __forceinline__ void block_sums(double (&v)[K], double* sh) {
    constexpr int W = NT / kWave;
#pragma unroll
    for (int k = 0; k < K; ++k) v[k] = wave_sum(v[k]);
    if ((threadIdx.x & (kWave - 1)) == 0) {
#pragma unroll
        for (int k = 0; k < K; ++k) sh[k * W + threadIdx.x / kWave] = v[k];
    }
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < K; ++k) {
            double acc = 0.0;
            for (int w = 0; w < W; ++w) acc += sh[k * W + w];
            v[k] = acc;
        }
    }
    __syncthreads();
}

// The matrix values one node's rows need in the symmetric form (kM positive offsets).
// They are loaded after the tile's vector phase: loading them first keeps ~70 more
// registers live through it, and the lower occupancy costs more (83 against 65 us per
// iteration at 1000x1000) than the overlap gains.
template <typename VT, int kM>
struct SymRows {
    using P = typename PairOf<VT>::type;
    P lu[kM], ll[kM], uu[kM], ul[kM];
    bool okl[kM];
    dpair du, dl;
};

template <typename VT, int kM>
__device__ __forceinline__ void load_sym_rows(const StencilView& v, const dpair* __restrict__ diag, int np_d,
                                              const VT* __restrict__ planes, int np, int kb, int i,
                                              SymRows<VT, kM>& a) {
#pragma unroll
    for (int k = 0; k < kM; ++k) {
        const int jl = i - v.off[k + 1];
        a.okl[k] = jl >= 0;
        const int jlc = a.okl[k] ? jl : i;
        plane_block<false>(planes, i, kb + k, np, a.uu[k], a.ul[k]);
        plane_block<true>(planes, jlc, kb + k, np, a.lu[k], a.ll[k]);
    }
    a.du = __builtin_nontemporal_load(diag + plane_at(i, 0, 0, np_d));
    a.dl = __builtin_nontemporal_load(diag + plane_at(i, 0, 1, np_d));
}

// The two halves of load_sym_rows: the node's own blocks (upper and diagonal, first read
// from HBM) and the transposed lower blocks (the neighbours' planes, mostly in L2).
template <typename VT, int kM>
__device__ __forceinline__ void load_sym_upper(const dpair* __restrict__ diag, int np_d, const VT* __restrict__ planes,
                                               int np, int kb, int i, SymRows<VT, kM>& a) {
#pragma unroll
    for (int k = 0; k < kM; ++k) plane_block<false>(planes, i, kb + k, np, a.uu[k], a.ul[k]);
    a.du = __builtin_nontemporal_load(diag + plane_at(i, 0, 0, np_d));
    a.dl = __builtin_nontemporal_load(diag + plane_at(i, 0, 1, np_d));
}

template <typename VT, int kM>
__device__ __forceinline__ void load_sym_lower(const StencilView& v, const VT* __restrict__ planes, int np, int kb,
                                               int i, SymRows<VT, kM>& a) {
#pragma unroll
    for (int k = 0; k < kM; ++k) {
        const int jl = i - v.off[k + 1];
        a.okl[k] = jl >= 0;
        const int jlc = a.okl[k] ? jl : i;
        plane_block<true>(planes, jlc, kb + k, np, a.lu[k], a.ll[k]);
    }
}

// Rows 2i, 2i+1 times the tile's LDS windows, in column order (as stencil_sym_node).
template <typename VT, int kM>
__device__ __forceinline__ void apply_sym_rows(const StencilView& v, const SymRows<VT, kM>& a,
                                               const dpair* __restrict__ win, int t, double& a0, double& a1) {
    a0 = 0.0;
    a1 = 0.0;
#pragma unroll
    for (int k = kM - 1; k >= 0; --k)
        if (a.okl[k]) add_lower(a.lu[k], a.ll[k], win[v.slot_lo[k + 1] + t], a0, a1);
    add_upper(a.du, a.dl, win[v.slot_zero + t], a0, a1);
#pragma unroll
    for (int k = 0; k < kM; ++k) add_upper(a.uu[k], a.ul[k], win[v.slot_up[k + 1] + t], a0, a1);
}

constexpr int kHalo = 3;  // halo window positions per thread and round

// The tiles of one block for values of type VT (see pcg_fused_stencil_kernel).
// x_mode: 0 leaves x alone, 1 adds alpha p_in, 2 adds alpha_prev p_out (read before
// p_out is overwritten) and then alpha p_in. The Jacobi inverses are read, not formed
// from the diagonal block: every node's inverse is first needed by another tile's
// halo, so computing it there saves no HBM bytes and adds two divisions per halo node
// (measured 1% slower at 1000x1000).
template <typename VT, int kM, int Prec>
__device__ __forceinline__ void fused_tiles(const PcgBuffers& b, const StencilView& v, const dpair* __restrict__ diag,
                                            int np_d, const VT* __restrict__ planes, int np, int kb,
                                            const double* __restrict__ r_in, double* __restrict__ r_out,
                                            const double* __restrict__ p_in, double* __restrict__ p_out,
                                            const double* __restrict__ q_in, double* __restrict__ q_out,
                                            double alpha, double beta, double alpha_prev, int x_mode,
                                            int tile_begin, int tile_end, dpair* win, double (&acc)[5]) {
    const int n = v.n_nodes;
    const int t = threadIdx.x;
    const dpair* __restrict__ r2 = reinterpret_cast<const dpair*>(r_in);
    const dpair* __restrict__ p2 = reinterpret_cast<const dpair*>(p_in);
    const dpair* __restrict__ q2 = reinterpret_cast<const dpair*>(q_in);
    dpair* __restrict__ x2 = reinterpret_cast<dpair*>(b.x);
    for (int tile = tile_begin; tile < tile_end; ++tile) {
        const int i0 = tile * kFusedTile;
        const int i = i0 + t;
        const bool own = i < n;
        const int ic = own ? i : n - 1;
        constexpr int M = kM > 0 ? kM : 1;
        SymRows<VT, M> rows;
        dpair pn = {0.0, 0.0}, zn = {0.0, 0.0};
        if (own) {
            const dpair ro = r2[i], qo = q2[i], po = p2[i];
            if (x_mode != 0) {  // solver.cpp:599-601, two iterations' terms in their order
                dpair x = x2[i];
                if (x_mode == 2) {
                    const dpair pp = reinterpret_cast<const dpair*>(p_out)[i];
                    x.x += alpha_prev * pp.x;
                    x.y += alpha_prev * pp.y;
                }
                x.x += alpha * po.x;
                x.y += alpha * po.y;
                x2[i] = x;
            }
            const dpair rn = {ro.x - alpha * qo.x, ro.y - alpha * qo.y};
            double z0, z1;
            precond_node<Prec>(b.m, i, rn.x, rn.y, z0, z1);
            zn = dpair{z0, z1};
            pn = dpair{z0 + beta * po.x, z1 + beta * po.y};
            reinterpret_cast<dpair*>(r_out)[i] = rn;
            reinterpret_cast<dpair*>(p_out)[i] = pn;
            acc[0] += rn.x * rn.x + rn.y * rn.y;
            acc[1] += rn.x * z0 + rn.y * z1;
        }
        win[v.slot_zero + t] = pn;
        // halo positions h of the windows (every slot but the own ones), kHalo per thread
        // and round with all their loads issued before any is used
        const int halo = v.win_total - kFusedTile;
        for (int h0 = 0; h0 < halo; h0 += kHalo * kFusedTile) {
            int hw[kHalo], hj[kHalo];
            dpair hr[kHalo], hq[kHalo], hp[kHalo];
            NodePrecond<Prec> hm[kHalo];
#pragma unroll
            for (int u = 0; u < kHalo; ++u) {
                const int h = h0 + u * kFusedTile + t;
                const int w = h < v.slot_zero ? h : h + kFusedTile;
                int c = 0;
                while (c + 1 < v.n_clusters && w >= v.win_base[c + 1]) ++c;
                const int j = i0 + v.win_lo[c] + (w - v.win_base[c]);
                hw[u] = h < halo ? w : -1;
                hj[u] = (h < halo && j >= 0 && j < n) ? j : -1;
                const int jc = hj[u] >= 0 ? hj[u] : 0;
                hr[u] = r2[jc];
                hq[u] = q2[jc];
                hp[u] = p2[jc];
                hm[u].load(b.m, jc);
            }
#pragma unroll
            for (int u = 0; u < kHalo; ++u) {
                if (hw[u] < 0) continue;
                dpair pj = {0.0, 0.0};
                if (hj[u] >= 0) {
                    double z0, z1;
                    hm[u].apply(hr[u].x - alpha * hq[u].x, hr[u].y - alpha * hq[u].y, z0, z1);
                    pj = dpair{z0 + beta * hp[u].x, z1 + beta * hp[u].y};
                }
                win[hw[u]] = pj;
            }
        }
        __syncthreads();
        if (own) {
            double a0, a1;
            if (kM > 0) {
                load_sym_rows<VT, M>(v, diag, np_d, planes, np, kb, ic, rows);
                apply_sym_rows<VT, M>(v, rows, win, t, a0, a1);
            }
            else
                stencil_sym_node<VT, 0, true>(v, diag, np_d, planes, np, kb, nullptr, i, a0, a1, win, t);
            reinterpret_cast<dpair*>(q_out)[i] = dpair{a0, a1};
            double m0, m1;
            precond_node<Prec>(b.m, i, a0, a1, m0, m1);
            acc[2] += pn.x * a0 + pn.y * a1;
            acc[3] += a0 * zn.x + a1 * zn.y;
            acc[4] += a0 * m0 + a1 * m1;
        }
        __syncthreads();  // the next tile overwrites the windows
    }
}

// The last block's state update of a fused iteration launch (tot: r'.r', r'.z, p'.q',
// q'.z, q'.M^-1 q'; alpha: the alpha the launch applied): the reference's stop test on r'
// (solver.cpp:606-611), its breakdown checks, alpha' = r'.z / p'.q' and beta' from the
// r.z recurrence of the next residual.
__device__ __forceinline__ void fused_state_update(PcgState* st, const double (&tot)[5], double alpha) {
    const double rr = tot[0], rho = tot[1], pq = tot[2], qz = tot[3], qmq = tot[4];
    st->iter += 1;  // -1 -> 0 on the first launch, which only forms p0 and A p0
    st->rr = rr;
    if (st->iter > 0 && sqrt(max0(rr)) <= st->target) {  // solver.cpp:606-611
        st->converged = 1;
        st->done = 1;
        return;
    }
    if (!(rho > 0.0) || !isfinite(rho)) {  // solver.cpp:579-581, 614-618
        st->status = LFG_ERR_NONFINITE;
        st->done = 1;
        return;
    }
    if (st->iter >= st->max_iter) {
        st->done = 1;
        return;
    }
    if (!(pq > 0.0) || !isfinite(pq)) {  // solver.cpp:591-593 of the next iteration
        st->status = LFG_ERR_NOT_SPD;
        st->done = 1;
        return;
    }
    const double a = rho / pq;
    // r''.M^-1 r'' for r'' = r' - a q' (M^-1 symmetric); a cancellation below zero
    // means r'' is negligible next to r', where beta = 0 is the exact limit
    const double rho_next = rho - 2.0 * a * qz + a * a * qmq;
    st->pq = pq;
    st->rho = rho;
    st->alpha_prev = alpha;
    st->alpha = a;
    st->beta = max0(rho_next) / rho;
    if (!st->phase32 && rr <= st->switch_rr) st->phase32 = 1;  // mixed: FP32 values from here on
}

// One PCG iteration per launch on the symmetric stencil form (launch_pcg_fused_stencil).
// Each block walks a contiguous range of 256-node tiles (XCD-remapped). Per tile:
//   1. every thread updates its own node: x += alpha p, r' = r - alpha q, z = M^-1 r',
//      p' = z + beta p, written to x, r_out, p_out and to the tile's own LDS window slot;
//   2. the remaining window positions (the halo the multiply reads, up to the largest
//      offset on either side) get p' formed the same way from the neighbours' r, q, p;
//   3. q' = A p' for the own node from the rows and the LDS windows, written to q_out.
// Inputs and outputs of r, p, q are different buffers, so a neighbour's old values are
// never overwritten while another block still reads them.
template <int kM, int Prec>
__global__ __launch_bounds__(kFusedTile) void pcg_fused_stencil_kernel(
    PcgBuffers b, StencilView v, const double* __restrict__ planes64, const float* __restrict__ planes32,
    const double* __restrict__ diag_c, const double* __restrict__ r_in, double* __restrict__ r_out,
    const double* __restrict__ p_in, double* __restrict__ p_out, const double* __restrict__ q_in,
    double* __restrict__ q_out) {
    if (b.st->done) return;
    extern __shared__ double win_lds[];
    dpair* const win = reinterpret_cast<dpair*>(win_lds);
    __shared__ double sh[5 * (kFusedTile / kWave)];
    const double alpha = b.st->alpha, beta = b.st->beta, alpha_prev = b.st->alpha_prev;
    // x += alpha p is deferred every second launch and done for two iterations in the
    // next (launch L reads p_L-1 from p_in and still finds p_L-2 in p_out), which saves
    // one read and one write of x per two iterations; launch_pcg_fused_flush adds a term
    // left over at the end. The launch index counts from 0 after pcg_init (iter = -1).
    const long long launch = b.st->iter + 1;
    const int x_mode = (launch & 1) ? 0 : (launch >= 2 ? 2 : 1);
    const bool use32 = planes32 != nullptr && (planes64 == nullptr || b.st->phase32 != 0);
    const int n_tiles = (v.n_nodes + kFusedTile - 1) / kFusedTile;
    const int lb = xcd_block(blockIdx.x, gridDim.x);
    const int per_block = (n_tiles + gridDim.x - 1) / gridDim.x;
    const int tile_begin = lb * per_block;
    const int tile_end = min(tile_begin + per_block, n_tiles);
    double acc[5] = {0.0, 0.0, 0.0, 0.0, 0.0};  // r'.r', r'.z, p'.q', q'.z, q'.M^-1 q'
    if (use32)
        fused_tiles<float, kM, Prec>(b, v, reinterpret_cast<const dpair*>(diag_c), 1, planes32, v.n_off - 1, 0, r_in,
                                     r_out, p_in, p_out, q_in, q_out, alpha, beta, alpha_prev, x_mode, tile_begin,
                                     tile_end, win, acc);
    else
        fused_tiles<double, kM, Prec>(b, v, reinterpret_cast<const dpair*>(planes64), v.n_off, planes64, v.n_off, 1,
                                      r_in, r_out, p_in, p_out, q_in, q_out, alpha, beta, alpha_prev, x_mode,
                                      tile_begin, tile_end, win, acc);
    block_sums<kFusedTile, 5>(acc, sh);
    double tot[5];
    if (!finish_reduction<kFusedTile, 5>(b.tickets + 1, b.partials, acc, tot, sh)) return;
    if (threadIdx.x == 0) fused_state_update(b.st, tot, alpha);
}

// ---- the fused iteration on 2D tiles (variant 11)
//
// A row-structured symmetric stencil (positive offsets 1, L-1, L, L+1) couples node i with
// i +- 1 in its mesh row and i +- L + {-1, 0, 1} in the rows beside it. A tile of kGridRows
// row segments of 64 nodes (node i = i0 + r L + l: wavefront r, lane l) then needs p' on a
// window of (kGridRows + 2) x 66 positions: its own 256 nodes and a ring of 140 around them,
// against two whole 258-node windows of the 1D tile's neighbouring rows (518 positions).
// A tile issues its loads before it uses any: the own r, q, p, M^-1 (and x, older p), one halo
// position in each of the first 35 lanes of every wavefront, and the node's own plane pairs
// (upper and diagonal blocks, 10 x 16 B), in that order, so the vector phase waits only for
// its own loads while the planes stream in from HBM. The transposed lower blocks (8 x 16 B,
// mostly L2 hits) are issued once the vector phase has freed its registers: 117 VGPRs, four
// wavefronts per SIMD. At 1000 x 1000: 52.4 us per FP64 launch (5.66 TB/s of algorithmic
// bytes), 42.5 us with FP32 values.
constexpr int kGridWinCols = kWave + 2;
template <int R>
constexpr int grid_win() { return (R + 2) * kGridWinCols; }
template <int R>
constexpr int grid_halo() { return grid_win<R>() - R * kWave; }  // 2 * 66 + 2 R

// Window slot of (row r, column l), r in [-1, R], l in [-1, 64]; the offsets 1, L-1, L,
// L+1 move the slot by 1, 65, 66, 67.
__device__ __forceinline__ int grid_slot(int r, int l) { return (r + 1) * kGridWinCols + l + 1; }

// One tile of the 2D iteration (values of type VT); adds the tile's sums to acc.
template <typename VT, int Prec, int R>
__device__ __forceinline__ void grid_tile(const PcgBuffers& b, const StencilView& v, const VT* __restrict__ planes,
                                          const dpair* __restrict__ diag, int np_d, int np, int kb,
                                          const double* __restrict__ r_in, double* __restrict__ r_out,
                                          const double* __restrict__ p_in, double* __restrict__ p_out,
                                          const double* __restrict__ q_in, double* __restrict__ q_out, double alpha,
                                          double beta, double alpha_prev, int x_mode, int tile, dpair* win,
                                          double (&acc)[5], int& qi, dpair& qv) {
    const int n = v.n_nodes, L = v.row_len;
    const int col_tiles = (L + kWave - 1) / kWave;
    const int row_tiles = ((n + L - 1) / L + R - 1) / R;
    if (tile < row_tiles * col_tiles) {
        const int rt = tile / col_tiles;
        const int c0 = (tile - rt * col_tiles) * kWave;
        const int r = static_cast<int>(threadIdx.x) / kWave, l = static_cast<int>(threadIdx.x) & (kWave - 1);
        const int i0 = rt * R * L + c0;
        const int i = i0 + r * L + l;
        const bool own = c0 + l < L && i < n;
        const int ic = own ? i : 0;
        const dpair* __restrict__ r2 = reinterpret_cast<const dpair*>(r_in);
        const dpair* __restrict__ p2 = reinterpret_cast<const dpair*>(p_in);
        const dpair* __restrict__ q2 = reinterpret_cast<const dpair*>(q_in);
        dpair* __restrict__ x2 = reinterpret_cast<dpair*>(b.x);
        // own vectors
        const dpair ro = r2[ic], qo = q2[ic], po = p2[ic];
        NodePrecond<Prec> mo;
        mo.load(b.m, ic);
        dpair xo = {0.0, 0.0}, pp = {0.0, 0.0};
        if (x_mode != 0) {
            xo = x2[ic];
            if (x_mode == 2) pp = reinterpret_cast<const dpair*>(p_out)[ic];
        }
        // halo position of this thread: row -1, row R, then columns -1 and 64, spread evenly
        // over the wavefronts (kHaloW positions each, in order) so that none waits at the
        // barrier for another's extra loads
        constexpr int kHaloN = grid_halo<R>();
        constexpr int kHaloW = (kHaloN + R - 1) / R;
        static_assert(kHaloW <= kWave, "one halo position per thread");
        const int h = (l < kHaloW) ? r * kHaloW + l : kHaloN;
        int hr = 0, hc = 0, hg = -1;
        if (h < kHaloN) {
            if (h < kGridWinCols) {
                hr = -1;
                hc = h - 1;
            } else if (h < 2 * kGridWinCols) {
                hr = R;
                hc = h - kGridWinCols - 1;
            } else {
                hr = (h - 2 * kGridWinCols) >> 1;
                hc = ((h - 2 * kGridWinCols) & 1) ? kWave : -1;
            }
            const int g = i0 + hr * L + hc;
            hg = (g >= 0 && g < n) ? g : -1;
        }
        const int hgc = hg >= 0 ? hg : 0;
        dpair hrv = {0.0, 0.0}, hqv = {0.0, 0.0}, hpv = {0.0, 0.0};
        NodePrecond<Prec> hm;
        if (h < kHaloN) {
            hrv = r2[hgc];
            hqv = q2[hgc];
            hpv = p2[hgc];
            hm.load(b.m, hgc);
        }
        // the node's own planes (upper and diagonal blocks)
        SymRows<VT, 4> rows;
        load_sym_upper<VT, 4>(diag, np_d, planes, np, kb, ic, rows);
        // 1. own node: x += alpha p (deferred), r' = r - alpha q, z = M^-1 r', p' = z + beta p
        dpair pn = {0.0, 0.0}, zn = {0.0, 0.0};
        if (own) {
            if (x_mode != 0) {  // solver.cpp:599-601, two iterations' terms in their order
                if (x_mode == 2) {
                    xo.x += alpha_prev * pp.x;
                    xo.y += alpha_prev * pp.y;
                }
                xo.x += alpha * po.x;
                xo.y += alpha * po.y;
                x2[i] = xo;
            }
            const dpair rn = {ro.x - alpha * qo.x, ro.y - alpha * qo.y};
            double z0, z1;
            mo.apply(rn.x, rn.y, z0, z1);
            zn = dpair{z0, z1};
            pn = dpair{z0 + beta * po.x, z1 + beta * po.y};
            reinterpret_cast<dpair*>(r_out)[i] = rn;
            reinterpret_cast<dpair*>(p_out)[i] = pn;
            acc[0] += rn.x * rn.x + rn.y * rn.y;
            acc[1] += rn.x * z0 + rn.y * z1;
        }
        win[grid_slot(r, l)] = pn;
        // the transposed lower blocks, once the own vectors' registers are free
        __builtin_amdgcn_sched_barrier(0);
        load_sym_lower<VT, 4>(v, planes, np, kb, ic, rows);
        // 2. the halo ring
        if (h < kHaloN) {
            dpair pj = {0.0, 0.0};
            if (hg >= 0) {
                double z0, z1;
                hm.apply(hrv.x - alpha * hqv.x, hrv.y - alpha * hqv.y, z0, z1);
                pj = dpair{z0 + beta * hpv.x, z1 + beta * hpv.y};
            }
            win[grid_slot(hr, hc)] = pj;
        }
        __syncthreads();
        // 3. q' = A p' in column order (as stencil_sym_node): transposed lower blocks from
        // the farthest, the diagonal block, the upper blocks
        if (own) {
            const int c = grid_slot(r, l);
            const int dw[4] = {1, kGridWinCols - 1, kGridWinCols, kGridWinCols + 1};
            double a0 = 0.0, a1 = 0.0;
#pragma unroll
            for (int k = 3; k >= 0; --k)
                if (rows.okl[k]) add_lower(rows.lu[k], rows.ll[k], win[c - dw[k]], a0, a1);
            add_upper(rows.du, rows.dl, win[c], a0, a1);
#pragma unroll
            for (int k = 0; k < 4; ++k) add_upper(rows.uu[k], rows.ul[k], win[c + dw[k]], a0, a1);
            qi = i;  // q' is stored after the reduction's ticket (pcg_fused_grid_kernel)
            qv = dpair{a0, a1};
            double m0, m1;
            mo.apply(a0, a1, m0, m1);
            acc[2] += pn.x * a0 + pn.y * a1;
            acc[3] += a0 * zn.x + a1 * zn.y;
            acc[4] += a0 * m0 + a1 * m1;
        }
    }
}

// One fused PCG iteration per launch on 2D tiles (launch_pcg_fused_grid): the FP32 or FP64
// planes by the state's phase, then the iteration's sums and state update as in
// pcg_fused_stencil_kernel.
template <int Prec, int R>
__global__ __launch_bounds__(R * kWave) void pcg_fused_grid_kernel(
    PcgBuffers b, StencilView v, const double* __restrict__ planes64, const float* __restrict__ planes32,
    const double* __restrict__ diag_c, const double* __restrict__ r_in, double* __restrict__ r_out,
    const double* __restrict__ p_in, double* __restrict__ p_out, const double* __restrict__ q_in,
    double* __restrict__ q_out) {
    if (b.st->done) return;
    __shared__ dpair win[grid_win<R>()];
    __shared__ double sh[5 * R];
    const double alpha = b.st->alpha, beta = b.st->beta, alpha_prev = b.st->alpha_prev;
    const long long launch = b.st->iter + 1;  // x update deferred as in pcg_fused_stencil_kernel
    const int x_mode = (launch & 1) ? 0 : (launch >= 2 ? 2 : 1);
    const bool use32 = planes32 != nullptr && (planes64 == nullptr || b.st->phase32 != 0);
    const int tb = xcd_block(blockIdx.x, gridDim.x);  // row bands stay on one XCD
    // Odd launches walk the tiles in reverse, so a launch starts with the tiles whose planes
    // and vectors the previous launch touched last (still in the L2 and the Infinity Cache):
    // 51.3 against 51.8 us per launch at 1000 x 1000. The block partials are folded in block
    // order either way, so the sums are run-to-run identical.
    const int tile = (launch & 1) ? static_cast<int>(gridDim.x) - 1 - tb : tb;
    double acc[5] = {0.0, 0.0, 0.0, 0.0, 0.0};  // r'.r', r'.z, p'.q', q'.z, q'.M^-1 q'
    int qi = -1;
    dpair qv = {0.0, 0.0};
    if (use32)
        grid_tile<float, Prec, R>(b, v, planes32, reinterpret_cast<const dpair*>(diag_c), 1, v.n_off - 1, 0, r_in, r_out,
                                  p_in, p_out, q_in, q_out, alpha, beta, alpha_prev, x_mode, tile, win, acc, qi, qv);
    else
        grid_tile<double, Prec, R>(b, v, planes64, reinterpret_cast<const dpair*>(planes64), v.n_off, v.n_off, 1, r_in,
                                   r_out, p_in, p_out, q_in, q_out, alpha, beta, alpha_prev, x_mode, tile, win, acc, qi,
                                   qv);
    block_sums<R * kWave, 5>(acc, sh);
    double tot[5];
    // q' is stored after the block's ticket: the ticket waits for the partial's store, and
    // would otherwise also wait for q''s (the next launch reads q' after this one ends)
    const bool last = finish_reduction<R * kWave, 5>(b.tickets + 1, b.partials, acc, tot, sh);
    if (qi >= 0) reinterpret_cast<dpair*>(q_out)[qi] = qv;
    if (!last) return;
    if (threadIdx.x == 0) fused_state_update(b.st, tot, alpha);
}

// ---- the fused iteration on the node-slice layout (variants 12 and 14)
//
// The iteration of pcg_fused_stencil_kernel for any node-blocked pattern (RCM or other
// numberings, where the blocks lie at no fixed node offsets). A wavefront owns one slice of
// 64 nodes (lane l: node 64 s + l). The lane updates its node (x deferred, r' = r - alpha q,
// z = M^-1 r', p' = z + beta p), then forms q' = A p' for the node's two rows over the
// slice's blocks in column order, as slice_rows does: p' of each column node j is formed
// from the launch's inputs r, q, p and the preconditioner of j (no block writes them during
// the launch) with the owner's arithmetic, so it has the bits the owner writes and no
// exchange between blocks is needed. The column gathers hit the XCD's L2: slices are
// XCD-remapped in contiguous ranges and a node's columns lie within the numbering's
// bandwidth. issue() starts every load that does not need the launch's scalars; finish()
// waits for them and does the arithmetic, so the scalars can arrive in between.
template <typename VT, int kW, int Prec>
struct SliceStep {
    using P = typename PairOf<VT>::type;
    int node, ic, begin, width;
    bool own, unrolled;
    dpair ro, qo, po, xo, pp;
    NodePrecond<Prec> mo;
    P u[kW > 0 ? kW : 1], l[kW > 0 ? kW : 1];
    int c[kW > 0 ? kW : 1];

    // The own node's r, q, p and preconditioner, the slice's blocks (up to kW: indices
    // clamped to the last block), and x and the older p when the launch updates x.
    __device__ __forceinline__ void issue(const PcgBuffers& b, const SliceView& v, const VT* __restrict__ planes,
                                          int s, const double* __restrict__ r_in, const double* __restrict__ p_in,
                                          const double* __restrict__ p_out, const double* __restrict__ q_in,
                                          int x_mode) {
        const int lane = static_cast<int>(threadIdx.x) & (kWave - 1);
        node = s * kWave + lane;
        own = node < (b.n >> 1);
        ic = own ? node : 0;
        ro = reinterpret_cast<const dpair*>(r_in)[ic];
        qo = reinterpret_cast<const dpair*>(q_in)[ic];
        po = reinterpret_cast<const dpair*>(p_in)[ic];
        mo.load(b.m, ic);
        begin = v.ptr[s];
        width = (v.ptr[s + 1] - begin) / kWave;
        const P* __restrict__ upper = reinterpret_cast<const P*>(planes);
        const P* __restrict__ lower = upper + v.blocks;
        unrolled = kW > 0 && width > 0 && width <= kW;
        if (unrolled) {
#pragma unroll
            for (int k = 0; k < kW; ++k) {
                const int at = begin + min(k, width - 1) * kWave + lane;
                u[k] = __builtin_nontemporal_load(upper + at);
                l[k] = __builtin_nontemporal_load(lower + at);
                c[k] = __builtin_nontemporal_load(v.col + at);
            }
        }
        xo = dpair{0.0, 0.0};
        pp = dpair{0.0, 0.0};
        if (x_mode != 0) {
            xo = reinterpret_cast<const dpair*>(b.x)[ic];
            if (x_mode == 2) pp = reinterpret_cast<const dpair*>(p_out)[ic];
        }
    }

    // The own node's update (stores r', p', x) and q' = A p' with the sums into acc; q' is
    // kept in (qi, qv) and stored by the caller, an earlier slice's here.
    __device__ __forceinline__ void finish(const PcgBuffers& b, const SliceView& v, const VT* __restrict__ planes,
                                           const double* __restrict__ r_in, double* __restrict__ r_out,
                                           const double* __restrict__ p_in, double* __restrict__ p_out,
                                           const double* __restrict__ q_in, double alpha, double beta,
                                           double alpha_prev, int x_mode, double (&acc)[5], int& qi, dpair& qv) {
        const int lane = static_cast<int>(threadIdx.x) & (kWave - 1);
        // p' of column node j from the launch's inputs, with the owner's arithmetic
        auto column = [&](int j) {
            const dpair rj = reinterpret_cast<const dpair*>(r_in)[j], qj = reinterpret_cast<const dpair*>(q_in)[j],
                        pj = reinterpret_cast<const dpair*>(p_in)[j];
            double z0, z1;
            precond_node<Prec>(b.m, j, rj.x - alpha * qj.x, rj.y - alpha * qj.y, z0, z1);
            return dpair{z0 + beta * pj.x, z1 + beta * pj.y};
        };
        // 1. own node (solver.cpp:599-604, 612-620), x two iterations at a time
        dpair pn = {0.0, 0.0}, zn = {0.0, 0.0};
        if (own) {
            if (x_mode != 0) {
                if (x_mode == 2) {
                    xo.x += alpha_prev * pp.x;
                    xo.y += alpha_prev * pp.y;
                }
                xo.x += alpha * po.x;
                xo.y += alpha * po.y;
                reinterpret_cast<dpair*>(b.x)[node] = xo;
            }
            const dpair rn = {ro.x - alpha * qo.x, ro.y - alpha * qo.y};
            double z0, z1;
            mo.apply(rn.x, rn.y, z0, z1);
            zn = dpair{z0, z1};
            pn = dpair{z0 + beta * po.x, z1 + beta * po.y};
            reinterpret_cast<dpair*>(r_out)[node] = rn;
            reinterpret_cast<dpair*>(p_out)[node] = pn;
            acc[0] += rn.x * rn.x + rn.y * rn.y;
            acc[1] += rn.x * z0 + rn.y * z1;
        }
        // 2. q' = A p' in column order (sparse.cpp:104-120; block order with v.pair_sum)
        double a0 = 0.0, a1 = 0.0;
        if (unrolled) {
#pragma unroll
            for (int k = 0; k < kW; ++k) {
                if (k < width) {
                    const dpair x = column(c[k]);
                    add_block(a0, a1, static_cast<double>(u[k].x), static_cast<double>(u[k].y),
                              static_cast<double>(l[k].x), static_cast<double>(l[k].y), x, v.pair_sum != 0);
                }
            }
        } else {
            const P* __restrict__ upper = reinterpret_cast<const P*>(planes);
            const P* __restrict__ lower = upper + v.blocks;
            for (int k = 0; k < width; ++k) {
                const int at = begin + k * kWave + lane;
                const P uk = __builtin_nontemporal_load(upper + at);
                const P lk = __builtin_nontemporal_load(lower + at);
                const dpair x = column(__builtin_nontemporal_load(v.col + at));
                add_block(a0, a1, static_cast<double>(uk.x), static_cast<double>(uk.y), static_cast<double>(lk.x),
                          static_cast<double>(lk.y), x, v.pair_sum != 0);
            }
        }
        if (own) {
            if (sizeof(VT) == sizeof(float)) add_diag32(v, node, pn, a0, a1);
            if (qi >= 0) reinterpret_cast<dpair*>(b.q)[qi] = qv;  // an earlier slice of this wavefront
            qi = node;
            qv = dpair{a0, a1};
            double m0, m1;
            mo.apply(a0, a1, m0, m1);
            acc[2] += pn.x * a0 + pn.y * a1;
            acc[3] += a0 * zn.x + a1 * zn.y;
            acc[4] += a0 * m0 + a1 * m1;
        }
    }
};

// One slice of variant 12: the scalars are read at the kernel's start; the launch parity
// (x_mode) is known before the loads, and a finished solve's launch stops after issuing
// them, before any store.
template <typename VT, int kW, int Prec>
__device__ __forceinline__ void fused_slice(const PcgBuffers& b, const SliceView& v, const VT* __restrict__ planes,
                                            int s, const double* __restrict__ r_in, double* __restrict__ r_out,
                                            const double* __restrict__ p_in, double* __restrict__ p_out,
                                            const double* __restrict__ q_in, double alpha, double beta,
                                            double alpha_prev, int x_mode, int done, double (&acc)[5], int& qi,
                                            dpair& qv) {
    SliceStep<VT, kW, Prec> step;
    step.issue(b, v, planes, s, r_in, p_in, p_out, q_in, x_mode);
    if (done) return;
    step.finish(b, v, planes, r_in, r_out, p_in, p_out, q_in, alpha, beta, alpha_prev, x_mode, acc, qi, qv);
}

// One fused PCG iteration per launch on the node slices (launch_pcg_fused_slices): each
// block takes a contiguous, XCD-remapped range of slices, one per wavefront and round; the
// sums, state update and q' store as in pcg_fused_grid_kernel (b.q is q_out).
template <int kW, int Prec>
__global__ __launch_bounds__(kThreads) void pcg_fused_slices_kernel(
    PcgBuffers b, SliceView v, const double* __restrict__ planes64, const float* __restrict__ planes32,
    const double* __restrict__ r_in, double* __restrict__ r_out, const double* __restrict__ p_in,
    double* __restrict__ p_out, const double* __restrict__ q_in) {
    __shared__ double sh[5 * (kThreads / kWave)];
    const int done = b.st->done;  // checked once the slice's loads are issued
    const double alpha = b.st->alpha, beta = b.st->beta, alpha_prev = b.st->alpha_prev;
    const long long launch = b.st->iter + 1;  // x update deferred as in pcg_fused_stencil_kernel
    const int x_mode = (launch & 1) ? 0 : (launch >= 2 ? 2 : 1);
    const bool use32 = planes32 != nullptr && (planes64 == nullptr || b.st->phase32 != 0);
    constexpr int kWaves = kThreads / kWave;
    const int lb = xcd_block(blockIdx.x, gridDim.x);
    const int per_block = (v.n_slices + gridDim.x - 1) / gridDim.x;
    const int start = lb * per_block;
    const int stop = min(start + per_block, v.n_slices);
    double acc[5] = {0.0, 0.0, 0.0, 0.0, 0.0};  // r'.r', r'.z, p'.q', q'.z, q'.M^-1 q'
    int qi = -1;
    dpair qv = {0.0, 0.0};
    for (int s = start + static_cast<int>(threadIdx.x) / kWave; s < stop; s += kWaves) {
        if (use32)
            fused_slice<float, kW, Prec>(b, v, planes32, s, r_in, r_out, p_in, p_out, q_in, alpha, beta, alpha_prev,
                                         x_mode, done, acc, qi, qv);
        else
            fused_slice<double, kW, Prec>(b, v, planes64, s, r_in, r_out, p_in, p_out, q_in, alpha, beta, alpha_prev,
                                          x_mode, done, acc, qi, qv);
    }
    if (done) return;
    block_sums<kThreads, 5>(acc, sh);
    double tot[5];
    const bool last = finish_reduction<kThreads, 5>(b.tickets + 1, b.partials, acc, tot, sh);
    if (qi >= 0) reinterpret_cast<dpair*>(b.q)[qi] = qv;
    if (!last) return;
    if (threadIdx.x == 0) fused_state_update(b.st, tot, alpha);
}

// ---- variant 14: the iteration of variant 12 with its reduction finished at the start of
// the next launch.
//
// Launch L stores its blocks' partial sums and ends: no ticket, no wait for the stores, no
// last-block fold, which is the chain of dependent memory round trips a small system's
// iteration waits for. Every block of launch L + 1 reads those partials while its slice
// loads are in flight, folds them in the order of finish_reduction's one-level path (so the
// totals, and every iterate, have the bits of variant 12), and applies the state update of
// variant 12 itself; block 0 writes the updated state for the host and the next launch. The
// state and the partials alternate between two buffers by launch parity (st_in / st_out,
// part_in / part_out): no block overwrites what another block of the same launch reads.
// `pending` in the state says whether part_in holds sums still to fold. The host sees a
// finished solve one launch later than with variant 12. Blocks of kFoldThreads threads, one
// slice per wavefront, at most kFoldMaxBlocks blocks (launch_pcg_fold_slices).
constexpr int kFoldParts = (kFoldMaxBlocks + kFoldThreads - 1) / kFoldThreads;  // partials per thread

// Folds the previous launch's partials into `cur` (the state that launch used, in
// registers): every thread gets the totals (its wave's sums through LDS, added in wave
// order as block_sums adds them for thread 0) and applies the state update itself, so no
// broadcast is needed. part[i] holds block t + i kFoldThreads's five sums (thread t; n_part
// of them exist); `sh` holds 5 doubles per wave. Block 0 publishes the result.
__device__ __forceinline__ void fold_previous(PcgState& cur, PcgState* __restrict__ st_out,
                                              const double (&part)[kFoldParts][5], int n_part, double* sh) {
    constexpr int W = kFoldThreads / kWave;
    if (cur.pending) {
        double tot[5];
#pragma unroll
        for (int k = 0; k < 5; ++k) {
            double acc = 0.0;  // finish_reduction's order: blocks t, t + NT, ...
#pragma unroll
            for (int i = 0; i < kFoldParts; ++i)
                if (i < n_part) acc += part[i][k];
            tot[k] = wave_sum(acc);
        }
        if ((threadIdx.x & (kWave - 1)) == 0) {
#pragma unroll
            for (int k = 0; k < 5; ++k) sh[k * W + threadIdx.x / kWave] = tot[k];
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < 5; ++k) {
            double acc = 0.0;
#pragma unroll
            for (int w = 0; w < W; ++w) acc += sh[k * W + w];
            tot[k] = acc;
        }
        fused_state_update(&cur, tot, cur.alpha);
    }
    cur.pending = cur.done ? 0 : 1;
    if (blockIdx.x == 0 && threadIdx.x == 0) {  // word by word: a struct copy keeps `cur` out of registers
        static_assert(sizeof(PcgState) % 8 == 0, "state words");
        const long long* src = reinterpret_cast<const long long*>(&cur);
        long long* dst = reinterpret_cast<long long*>(st_out);
#pragma unroll
        for (int i = 0; i < static_cast<int>(sizeof(PcgState) / 8); ++i) dst[i] = src[i];
    }
}

template <int kW, int Prec>
__global__ __launch_bounds__(kFoldThreads) void pcg_fold_slices_kernel(
    PcgBuffers b, SliceView v, const double* __restrict__ planes64, const float* __restrict__ planes32,
    const double* __restrict__ r_in, double* __restrict__ r_out, const double* __restrict__ p_in,
    double* __restrict__ p_out, const double* __restrict__ q_in, const PcgState* __restrict__ st_in,
    PcgState* __restrict__ st_out, const double* __restrict__ part_in, double* __restrict__ part_out) {
    __shared__ double sh[5 * (kFoldThreads / kWave)];
    __shared__ double sh_fold[5 * (kFoldThreads / kWave)];
    const int t = static_cast<int>(threadIdx.x);
    const int nb = static_cast<int>(gridDim.x);
    // the previous launch's partials, read whether or not they are pending: no wait on the state
    const int n_part = t < nb ? (nb - 1 - t) / kFoldThreads + 1 : 0;
    double part[kFoldParts][5];
#pragma unroll
    for (int i = 0; i < kFoldParts; ++i) {
#pragma unroll
        for (int k = 0; k < 5; ++k) part[i][k] = i < n_part ? part_in[(t + i * kFoldThreads) * 5 + k] : 0.0;
    }
    PcgState cur = *st_in;
    const long long launch = cur.iter + cur.pending + 1;  // the iteration count after the fold, plus one
    const int x_mode = (launch & 1) ? 0 : (launch >= 2 ? 2 : 1);
    const bool use32_in = planes32 != nullptr && (planes64 == nullptr || cur.phase32 != 0);
    const int lb = xcd_block(blockIdx.x, nb);
    const int per_block = (v.n_slices + nb - 1) / nb;  // at most one slice per wavefront
    const int s = lb * per_block + t / kWave;
    const bool has = t / kWave < per_block && s < v.n_slices;
    double acc[5] = {0.0, 0.0, 0.0, 0.0, 0.0};  // r'.r', r'.z, p'.q', q'.z, q'.M^-1 q'
    int qi = -1;
    dpair qv = {0.0, 0.0};
    if (use32_in) {  // the FP32 phase never ends within a solve
        SliceStep<float, kW, Prec> step;
        if (has) step.issue(b, v, planes32, s, r_in, p_in, p_out, q_in, x_mode);
        fold_previous(cur, st_out, part, n_part, sh_fold);
        if (has && !cur.done)
            step.finish(b, v, planes32, r_in, r_out, p_in, p_out, q_in, cur.alpha, cur.beta, cur.alpha_prev, x_mode,
                        acc, qi, qv);
    } else {
        SliceStep<double, kW, Prec> step;
        if (has) step.issue(b, v, planes64, s, r_in, p_in, p_out, q_in, x_mode);
        fold_previous(cur, st_out, part, n_part, sh_fold);
        if (has && !cur.done) {
            if (planes32 == nullptr || cur.phase32 == 0) {
                step.finish(b, v, planes64, r_in, r_out, p_in, p_out, q_in, cur.alpha, cur.beta, cur.alpha_prev,
                            x_mode, acc, qi, qv);
            } else {  // the fold switched to FP32 values: this launch's blocks again, as FP32
                SliceStep<float, kW, Prec> step32;
                step32.issue(b, v, planes32, s, r_in, p_in, p_out, q_in, x_mode);
                step32.finish(b, v, planes32, r_in, r_out, p_in, p_out, q_in, cur.alpha, cur.beta, cur.alpha_prev,
                              x_mode, acc, qi, qv);
            }
        }
    }
    if (cur.done) return;
    block_sums<kFoldThreads, 5>(acc, sh);
    if (t == 0) {
#pragma unroll
        for (int k = 0; k < 5; ++k) part_out[blockIdx.x * 5 + k] = acc[k];
    }
    if (qi >= 0) reinterpret_cast<dpair*>(b.q)[qi] = qv;
}

// ---- the same iteration in two launches (variant 13): for node slices of large systems,
// where forming every column's p' in the multiply (variant 12) costs more L2 traffic than
// one more launch.
//
// Vector launch: x (deferred), r' = r - alpha q, z = M^-1 r', p' = z + beta p for every node,
// ||r'||^2 and r'.z into the state. Its blocks take the nodes of the multiply launch's blocks
// (256 per block, the same XCD remapping), so a block's r', p' stay in its XCD's L2.
template <int Prec>
__global__ __launch_bounds__(kThreads) void pcg_slices_vector_kernel(PcgBuffers b, const double* __restrict__ r_in,
                                                                     double* __restrict__ r_out,
                                                                     const double* __restrict__ p_in,
                                                                     double* __restrict__ p_out,
                                                                     const double* __restrict__ q_in) {
    if (b.st->done) return;
    __shared__ double sh[2 * (kThreads / kWave)];
    const double alpha = b.st->alpha, beta = b.st->beta, alpha_prev = b.st->alpha_prev;
    const long long launch = b.st->iter + 1;  // x update deferred as in pcg_fused_stencil_kernel
    const int x_mode = (launch & 1) ? 0 : (launch >= 2 ? 2 : 1);
    const int n_nodes = b.n >> 1;
    const int lb = xcd_block(blockIdx.x, gridDim.x);
    const int per_block = ((n_nodes + kThreads - 1) / kThreads + gridDim.x - 1) / gridDim.x * kThreads;
    const int stop = min((lb + 1) * per_block, n_nodes);
    double acc[2] = {0.0, 0.0};
    for (int node = lb * per_block + static_cast<int>(threadIdx.x); node < stop; node += kThreads) {
        const dpair ro = reinterpret_cast<const dpair*>(r_in)[node];
        const dpair qo = reinterpret_cast<const dpair*>(q_in)[node];
        const dpair po = reinterpret_cast<const dpair*>(p_in)[node];
        NodePrecond<Prec> mo;
        mo.load(b.m, node);
        if (x_mode != 0) {  // solver.cpp:599-601, two iterations' terms in their order
            dpair xo = reinterpret_cast<const dpair*>(b.x)[node];
            if (x_mode == 2) {
                const dpair pp = reinterpret_cast<const dpair*>(p_out)[node];
                xo.x += alpha_prev * pp.x;
                xo.y += alpha_prev * pp.y;
            }
            xo.x += alpha * po.x;
            xo.y += alpha * po.y;
            reinterpret_cast<dpair*>(b.x)[node] = xo;
        }
        const dpair rn = {ro.x - alpha * qo.x, ro.y - alpha * qo.y};
        double z0, z1;
        mo.apply(rn.x, rn.y, z0, z1);
        reinterpret_cast<dpair*>(r_out)[node] = rn;
        reinterpret_cast<dpair*>(p_out)[node] = dpair{z0 + beta * po.x, z1 + beta * po.y};
        acc[0] += rn.x * rn.x + rn.y * rn.y;
        acc[1] += rn.x * z0 + rn.y * z1;
    }
    block_sums<kThreads, 2>(acc, sh);
    double tot[2];
    if (!finish_reduction<kThreads, 2>(b.tickets + 2, b.partials, acc, tot, sh)) return;
    if (threadIdx.x == 0) {
        b.st->rr_new = tot[0];
        b.st->rz_new = tot[1];
    }
}

// Multiply launch: q' = A p' over the node slices (the rows of slice_rows, p' gathered from
// p_out), p'.q', q'.z and q'.M^-1 q' with z = M^-1 r' formed again from r' (the vector
// launch's arithmetic), then the fused iteration's state update.
template <typename VT, int kW, int Prec>
__device__ __forceinline__ void split_slice(const PcgBuffers& b, const SliceView& v, const VT* __restrict__ planes,
                                            int s, const double* __restrict__ r_new,
                                            const double* __restrict__ p_new, double (&acc)[3], int& qi, dpair& qv) {
    using P = typename PairOf<VT>::type;
    const int lane = static_cast<int>(threadIdx.x) & (kWave - 1);
    const int node = s * kWave + lane;
    const bool own = node < (b.n >> 1);
    const int ic = own ? node : 0;
    const dpair* __restrict__ p2 = reinterpret_cast<const dpair*>(p_new);
    const dpair pn = p2[ic];
    const dpair rn = reinterpret_cast<const dpair*>(r_new)[ic];
    NodePrecond<Prec> mo;
    mo.load(b.m, ic);
    const int begin = v.ptr[s];
    const int width = (v.ptr[s + 1] - begin) / kWave;
    const P* __restrict__ upper = reinterpret_cast<const P*>(planes);
    const P* __restrict__ lower = upper + v.blocks;
    double a0 = 0.0, a1 = 0.0;
    if (kW > 0 && width > 0 && width <= kW) {
        P u[kW > 0 ? kW : 1], l[kW > 0 ? kW : 1];
        int c[kW > 0 ? kW : 1];
#pragma unroll
        for (int k = 0; k < kW; ++k) {
            const int at = begin + min(k, width - 1) * kWave + lane;
            u[k] = __builtin_nontemporal_load(upper + at);
            l[k] = __builtin_nontemporal_load(lower + at);
            c[k] = __builtin_nontemporal_load(v.col + at);
        }
        dpair x[kW > 0 ? kW : 1];
#pragma unroll
        for (int k = 0; k < kW; ++k) x[k] = p2[c[k]];
#pragma unroll
        for (int k = 0; k < kW; ++k) {
            if (k < width) {
                add_block(a0, a1, static_cast<double>(u[k].x), static_cast<double>(u[k].y),
                          static_cast<double>(l[k].x), static_cast<double>(l[k].y), x[k], v.pair_sum != 0);
            }
        }
    } else {
        for (int k = 0; k < width; ++k) {
            const int at = begin + k * kWave + lane;
            const P uk = __builtin_nontemporal_load(upper + at);
            const P lk = __builtin_nontemporal_load(lower + at);
            const dpair xk = p2[__builtin_nontemporal_load(v.col + at)];
            add_block(a0, a1, static_cast<double>(uk.x), static_cast<double>(uk.y), static_cast<double>(lk.x),
                      static_cast<double>(lk.y), xk, v.pair_sum != 0);
        }
    }
    if (own) {
        if (sizeof(VT) == sizeof(float)) add_diag32(v, node, pn, a0, a1);
        if (qi >= 0) reinterpret_cast<dpair*>(b.q)[qi] = qv;  // an earlier slice of this wavefront
        qi = node;
        qv = dpair{a0, a1};
        double z0, z1, m0, m1;
        mo.apply(rn.x, rn.y, z0, z1);
        mo.apply(a0, a1, m0, m1);
        acc[0] += pn.x * a0 + pn.y * a1;
        acc[1] += a0 * z0 + a1 * z1;
        acc[2] += a0 * m0 + a1 * m1;
    }
}

template <int kW, int Prec>
__global__ __launch_bounds__(kThreads) void pcg_slices_multiply_kernel(PcgBuffers b, SliceView v,
                                                                       const double* __restrict__ planes64,
                                                                       const float* __restrict__ planes32,
                                                                       const double* __restrict__ r_new,
                                                                       const double* __restrict__ p_new) {
    if (b.st->done) return;
    __shared__ double sh[3 * (kThreads / kWave)];
    const double alpha = b.st->alpha;
    const bool use32 = planes32 != nullptr && (planes64 == nullptr || b.st->phase32 != 0);
    constexpr int kWaves = kThreads / kWave;
    const int lb = xcd_block(blockIdx.x, gridDim.x);
    const int per_block = (v.n_slices + gridDim.x - 1) / gridDim.x;
    const int start = lb * per_block;
    const int stop = min(start + per_block, v.n_slices);
    double acc[3] = {0.0, 0.0, 0.0};  // p'.q', q'.z, q'.M^-1 q'
    int qi = -1;
    dpair qv = {0.0, 0.0};
    for (int s = start + static_cast<int>(threadIdx.x) / kWave; s < stop; s += kWaves) {
        if (use32)
            split_slice<float, kW, Prec>(b, v, planes32, s, r_new, p_new, acc, qi, qv);
        else
            split_slice<double, kW, Prec>(b, v, planes64, s, r_new, p_new, acc, qi, qv);
    }
    block_sums<kThreads, 3>(acc, sh);
    double tot[3];
    const bool last = finish_reduction<kThreads, 3>(b.tickets + 1, b.partials, acc, tot, sh);
    if (qi >= 0) reinterpret_cast<dpair*>(b.q)[qi] = qv;
    if (!last) return;
    if (threadIdx.x == 0) {
        const double all[5] = {b.st->rr_new, b.st->rz_new, tot[0], tot[1], tot[2]};
        fused_state_update(b.st, all, alpha);
    }
}

// ---- the stencil multiply alone on 2D tiles (variant 8 where the stencil is row-structured)
//
// q = A p for one tile of the 2D iteration's shape: p of the tile's 256 nodes and of its
// 140-position ring into the LDS window (one ring position per thread, as grid_tile), the
// node's planes (upper and diagonal blocks, then the transposed lower blocks), and the rows
// in column order from the window: the sums of stencil_sym_node, so q is bit-identical to
// the 1D multiply. p is read once per tile instead of once per row block.
template <typename VT, int R>
__device__ __forceinline__ void grid_spmv_tile(const PcgBuffers& b, const StencilView& v, const VT* __restrict__ planes,
                                               const dpair* __restrict__ diag, int np_d, int np, int kb, int tile,
                                               dpair* win, double& pq, int& qi, dpair& qv) {
    const int n = v.n_nodes, L = v.row_len;
    const int col_tiles = (L + kWave - 1) / kWave;
    const int rt = tile / col_tiles;
    const int c0 = (tile - rt * col_tiles) * kWave;
    const int r = static_cast<int>(threadIdx.x) / kWave, l = static_cast<int>(threadIdx.x) & (kWave - 1);
    const int i0 = rt * R * L + c0;
    const int i = i0 + r * L + l;
    const bool own = c0 + l < L && i < n;
    const int ic = own ? i : 0;
    const dpair* __restrict__ p2 = reinterpret_cast<const dpair*>(b.p);
    const dpair po = p2[ic];
    constexpr int kHaloN = grid_halo<R>();
    constexpr int kHaloW = (kHaloN + R - 1) / R;
    static_assert(kHaloW <= kWave, "one halo position per thread");
    const int h = (l < kHaloW) ? r * kHaloW + l : kHaloN;
    int hr = 0, hc = 0, hg = -1;
    if (h < kHaloN) {
        if (h < kGridWinCols) {
            hr = -1;
            hc = h - 1;
        } else if (h < 2 * kGridWinCols) {
            hr = R;
            hc = h - kGridWinCols - 1;
        } else {
            hr = (h - 2 * kGridWinCols) >> 1;
            hc = ((h - 2 * kGridWinCols) & 1) ? kWave : -1;
        }
        const int g = i0 + hr * L + hc;
        hg = (g >= 0 && g < n) ? g : -1;
    }
    const dpair hp = hg >= 0 ? p2[hg] : dpair{0.0, 0.0};
    SymRows<VT, 4> rows;
    load_sym_upper<VT, 4>(diag, np_d, planes, np, kb, ic, rows);
    win[grid_slot(r, l)] = own ? po : dpair{0.0, 0.0};
    __builtin_amdgcn_sched_barrier(0);
    load_sym_lower<VT, 4>(v, planes, np, kb, ic, rows);
    if (h < kHaloN) win[grid_slot(hr, hc)] = hp;
    __syncthreads();
    if (own) {
        const int c = grid_slot(r, l);
        const int dw[4] = {1, kGridWinCols - 1, kGridWinCols, kGridWinCols + 1};
        double a0 = 0.0, a1 = 0.0;
#pragma unroll
        for (int k = 3; k >= 0; --k)
            if (rows.okl[k]) add_lower(rows.lu[k], rows.ll[k], win[c - dw[k]], a0, a1);
        add_upper(rows.du, rows.dl, win[c], a0, a1);
#pragma unroll
        for (int k = 0; k < 4; ++k) add_upper(rows.uu[k], rows.ul[k], win[c + dw[k]], a0, a1);
        if (qi >= 0) reinterpret_cast<dpair*>(b.q)[qi] = qv;  // an earlier tile of this block
        qi = i;
        qv = dpair{a0, a1};
        pq += po.x * a0 + po.y * a1;
    }
    __syncthreads();  // the window is reused by the block's next tile
}

// q = A p, p.q and alpha = rho / p.q (solver.cpp:589-597) on 2D tiles; q is stored after
// the block's reduction ticket, as in pcg_fused_grid_kernel.
template <int R>
__global__ __launch_bounds__(R * kWave) void pcg_spmv_grid_kernel(PcgBuffers b, StencilView v,
                                                                  const double* __restrict__ planes64,
                                                                  const float* __restrict__ planes32,
                                                                  const double* __restrict__ diag_c) {
    if (b.st->done) return;
    __shared__ dpair win[grid_win<R>()];
    __shared__ double sh[R];
    const bool use32 = planes32 != nullptr && (planes64 == nullptr || b.st->phase32 != 0);
    const int L = v.row_len;
    const int n_tiles = ((v.n_nodes + L - 1) / L + R - 1) / R * ((L + kWave - 1) / kWave);
    double pq = 0.0;
    int qi = -1;
    dpair qv = {0.0, 0.0};
    for (int tile = xcd_block(blockIdx.x, gridDim.x); tile < n_tiles; tile += gridDim.x) {
        if (use32)
            grid_spmv_tile<float, R>(b, v, planes32, reinterpret_cast<const dpair*>(diag_c), 1, v.n_off - 1, 0, tile,
                                     win, pq, qi, qv);
        else
            grid_spmv_tile<double, R>(b, v, planes64, reinterpret_cast<const dpair*>(planes64), v.n_off, v.n_off, 1,
                                      tile, win, pq, qi, qv);
    }
    const double vals[1] = {block_sum<R * kWave>(pq, sh)};
    double tot[1];
    const bool last = finish_reduction<R * kWave, 1>(b.tickets + 1, b.partials, vals, tot, sh);
    if (qi >= 0) reinterpret_cast<dpair*>(b.q)[qi] = qv;
    if (!last) return;
    if (threadIdx.x == 0) {
        const double tpq = tot[0];
        PcgState* st = b.st;
        st->pq = tpq;
        if (!(tpq > 0.0) || !isfinite(tpq)) {  // solver.cpp:591-593
            st->status = LFG_ERR_NOT_SPD;
            st->done = 1;
        } else {
            st->alpha = st->rho / tpq;
        }
    }
}

// The x term an odd final launch deferred: x += alpha p_in of that launch. A solve
// that stopped there (done) left alpha unchanged; otherwise the launch moved it to
// alpha_prev.
__global__ __launch_bounds__(kThreads) void fused_flush_kernel(PcgBuffers b, const double* __restrict__ p_odd) {
    const PcgState* st = b.st;
    if (st->iter < 1 || (st->iter & 1) == 0) return;
    const double a = st->done ? st->alpha : st->alpha_prev;
    const int n_nodes = b.n >> 1;
    dpair* __restrict__ x2 = reinterpret_cast<dpair*>(b.x);
    const dpair* __restrict__ p2 = reinterpret_cast<const dpair*>(p_odd);
    for (int i = blockIdx.x * kThreads + threadIdx.x; i < n_nodes; i += gridDim.x * kThreads) {
        dpair x = x2[i];
        const dpair p = p2[i];
        x.x += a * p.x;
        x.y += a * p.y;
        x2[i] = x;
    }
}

// ---- the symmetric node-slice layout (variant 15)
//
// A node's rows read its upper blocks (j >= i) as stored and its lower blocks (j < i) as
// the transposes of the column nodes' upper blocks (j, i), so each pair of mirrored blocks
// streams its values once: at 1000 x 1000 about 5 blocks of 32 B and 5 columns plus 4
// (column, position) pairs per node instead of 9 blocks and 9 columns. A transposed block
// was streamed by its own node's wavefront shortly before (the slices of a launch are
// XCD-remapped in contiguous ranges, and a numbering's bandwidth is small next to a
// range), so it is read from the L2; the upper planes are therefore read with ordinary
// loads, not nontemporal ones. Each row sums in column order (lower blocks ascending, the
// diagonal block, the upper blocks ascending), as the CSR row loop; a transposed block's
// values are block (j, i)'s, which assembly makes equal to block (i, j)'s up to rounding
// (Ke(a, b) and Ke(b, a) are formed by different operation orders).

// One thread per (slice, lane): node 64 s + l's upper blocks from its two CSR rows (the
// blocks from the first column pair >= 2i on), zero padding; thread 0 also writes the zero
// block the padded lower entries point to. VT = float: the FP32 copy, whose diagonal entry
// is zero (the compensated FP64 diagonal block is added separately).
template <typename VT>
__global__ __launch_bounds__(256) void pack_sym_slices_kernel(int n_nodes, const int* __restrict__ rp,
                                                              const int* __restrict__ col, SymSliceView v,
                                                              const double* __restrict__ values,
                                                              VT* __restrict__ planes) {
    const long long t = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
    const int s = static_cast<int>(t >> 6);
    const int lane = static_cast<int>(t & (kWave - 1));
    if (t == 0) {
        for (int h = 0; h < 2; ++h) {
            planes[2 * (h * (v.ublocks + 1) + v.ublocks)] = VT(0);
            planes[2 * (h * (v.ublocks + 1) + v.ublocks) + 1] = VT(0);
        }
    }
    if (s >= v.n_slices) return;
    const int node = s * kWave + lane;
    const int begin = v.uptr[s];
    const int width = (v.uptr[s + 1] - begin) / kWave;
    int r0 = 0, r1 = 0, deg = 0, dlo = 0;
    bool lone = false;  // a node in no element: rows hold only their diagonals
    if (node < n_nodes) {
        r0 = rp[2 * node];
        r1 = rp[2 * node + 1];
        lone = r1 - r0 == 1;
        deg = lone ? 1 : (r1 - r0) >> 1;
        if (!lone)
            while (dlo < deg && col[r0 + 2 * dlo] < 2 * node) ++dlo;
    }
    constexpr bool kF32 = sizeof(VT) == sizeof(float);
    for (int k = 0; k < width; ++k) {
        const long long at = begin + static_cast<long long>(k) * kWave + lane;
        double u0 = 0.0, u1 = 0.0, l0 = 0.0, l1 = 0.0;
        const int kk = dlo + k;
        if (kF32 && k == 0) {
            // the diagonal block: FP64 in diag32
        } else if (lone && k == 0) {
            u0 = values[r0];
            l1 = values[r1];
        } else if (!lone && kk < deg) {
            u0 = values[r0 + 2 * kk];
            u1 = values[r0 + 2 * kk + 1];
            l0 = values[r1 + 2 * kk];
            l1 = values[r1 + 2 * kk + 1];
        }
        planes[2 * at] = static_cast<VT>(u0);
        planes[2 * at + 1] = static_cast<VT>(u1);
        planes[2 * (v.ublocks + 1 + at)] = static_cast<VT>(l0);
        planes[2 * (v.ublocks + 1 + at) + 1] = static_cast<VT>(l1);
    }
}

// The FP32 copy's diagonal block per node, after its planes are packed: the FP64 block
// minus, entry by entry, the rounding error of every FP32 value the node's rows read —
// its upper blocks' own values, and for a lower block (i, j) the mirrored FP32 block (j, i)
// in place of the FP64 values of block (i, j) — so every row keeps its FP64 sum.
__global__ __launch_bounds__(256) void sym_diag32_kernel(int n_nodes, const int* __restrict__ rp,
                                                         const int* __restrict__ col, SymSliceView v,
                                                         const double* __restrict__ values,
                                                         const float* __restrict__ planes32,
                                                         double* __restrict__ diag32) {
    const int node = static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x);
    if (node >= n_nodes) return;
    const int r0 = rp[2 * node], r1 = rp[2 * node + 1];
    double d0 = 0.0, d1 = 0.0, d2 = 0.0, d3 = 0.0;
    if (r1 - r0 == 1) {  // a node in no element
        d0 = values[r0];
        d3 = values[r1];
    } else {
        const int deg = (r1 - r0) >> 1;
        int dlo = 0;
        while (dlo < deg && col[r0 + 2 * dlo] < 2 * node) ++dlo;
        const fpair* __restrict__ up = reinterpret_cast<const fpair*>(planes32);
        const fpair* __restrict__ lo = up + (v.ublocks + 1);
        const int s = node / kWave, lane = node % kWave;
        for (int k = 0; k < deg; ++k) {
            const double u0 = values[r0 + 2 * k], u1 = values[r0 + 2 * k + 1];
            const double l0 = values[r1 + 2 * k], l1 = values[r1 + 2 * k + 1];
            if (k == dlo) {  // the diagonal block
                d0 += u0;
                d1 += u1;
                d2 += l0;
                d3 += l1;
            } else if (k > dlo) {  // an upper block: its own FP32 values
                d0 -= static_cast<double>(static_cast<float>(u0)) - u0;
                d1 -= static_cast<double>(static_cast<float>(u1)) - u1;
                d2 -= static_cast<double>(static_cast<float>(l0)) - l0;
                d3 -= static_cast<double>(static_cast<float>(l1)) - l1;
            } else {  // a lower block: read as the transpose of the mirrored FP32 block
                const int tp = v.ltpos[v.lptr[s] + k * kWave + lane];
                const fpair tu = up[tp], tl = lo[tp];
                d0 -= static_cast<double>(tu.x) - u0;
                d1 -= static_cast<double>(tl.x) - u1;
                d2 -= static_cast<double>(tu.y) - l0;
                d3 -= static_cast<double>(tl.y) - l1;
            }
        }
    }
    reinterpret_cast<dpair*>(diag32)[2 * node] = dpair{d0, d1};
    reinterpret_cast<dpair*>(diag32)[2 * node + 1] = dpair{d2, d3};
}

// Rows 2i, 2i+1 of node i = 64 s + lane over both lists, p from p2 (values of type VT; the
// FP32 copy adds its compensated FP64 diagonal block last). kWL / kWU > 0: lists up to that
// wide take unrolled paths whose loads are unconditional (entries clamped to the last one)
// and whose sums skip the clamped entries.
template <typename VT, int kWL, int kWU>
__device__ __forceinline__ void sym_slice_rows(const SymSliceView& v, const VT* __restrict__ planes,
                                               const dpair* __restrict__ p2, int s, double& a0, double& a1) {
    using P = typename PairOf<VT>::type;
    const int lane = static_cast<int>(threadIdx.x) & (kWave - 1);
    const P* __restrict__ up = reinterpret_cast<const P*>(planes);
    const P* __restrict__ lo = up + (v.ublocks + 1);
    const int lb = v.lptr[s];
    const int lw = (v.lptr[s + 1] - lb) / kWave;
    const int ub = v.uptr[s];
    const int uw = (v.uptr[s + 1] - ub) / kWave;
    auto add = [&](double ux, double uy, double lx, double ly, dpair x) { add_block(a0, a1, ux, uy, lx, ly, x, false); };
    if (kWL > 0 && kWU > 0 && lw > 0 && lw <= kWL && uw > 0 && uw <= kWU) {
        int lc[kWL > 0 ? kWL : 1], lt[kWL > 0 ? kWL : 1], uc[kWU > 0 ? kWU : 1];
#pragma unroll
        for (int k = 0; k < kWL; ++k) {
            const int at = lb + min(k, lw - 1) * kWave + lane;
            lc[k] = __builtin_nontemporal_load(v.lcol + at);
            lt[k] = __builtin_nontemporal_load(v.ltpos + at);
        }
        P uu[kWU > 0 ? kWU : 1], ul[kWU > 0 ? kWU : 1];
#pragma unroll
        for (int k = 0; k < kWU; ++k) {
            const int at = ub + min(k, uw - 1) * kWave + lane;
            uc[k] = __builtin_nontemporal_load(v.ucol + at);
            uu[k] = up[at];
            ul[k] = lo[at];
        }
        P tu[kWL > 0 ? kWL : 1], tl[kWL > 0 ? kWL : 1];
        dpair xl[kWL > 0 ? kWL : 1];
#pragma unroll
        for (int k = 0; k < kWL; ++k) {
            tu[k] = up[lt[k]];
            tl[k] = lo[lt[k]];
            xl[k] = p2[lc[k]];
        }
        dpair xu[kWU > 0 ? kWU : 1];
#pragma unroll
        for (int k = 0; k < kWU; ++k) xu[k] = p2[uc[k]];
#pragma unroll
        for (int k = 0; k < kWL; ++k)
            if (k < lw) add(tu[k].x, tl[k].x, tu[k].y, tl[k].y, xl[k]);
#pragma unroll
        for (int k = 0; k < kWU; ++k)
            if (k < uw) add(uu[k].x, uu[k].y, ul[k].x, ul[k].y, xu[k]);
    } else {
        for (int k = 0; k < lw; ++k) {
            const int at = lb + k * kWave + lane;
            const int tp = __builtin_nontemporal_load(v.ltpos + at);
            const P tu = up[tp], tl = lo[tp];
            add(tu.x, tl.x, tu.y, tl.y, p2[__builtin_nontemporal_load(v.lcol + at)]);
        }
        for (int k = 0; k < uw; ++k) {
            const int at = ub + k * kWave + lane;
            const P uu = up[at], ul = lo[at];
            add(uu.x, uu.y, ul.x, ul.y, p2[__builtin_nontemporal_load(v.ucol + at)]);
        }
    }
    if (sizeof(VT) == sizeof(float)) {
        const int node = s * kWave + lane;
        const dpair du = reinterpret_cast<const dpair*>(v.diag32)[2 * node];
        const dpair dl = reinterpret_cast<const dpair*>(v.diag32)[2 * node + 1];
        const dpair x = p2[node];
        a0 += du.x * x.x;
        a0 += du.y * x.y;
        a1 += dl.x * x.x;
        a1 += dl.y * x.y;
    }
}

// q = A p over the symmetric node slices, fused p.q and alpha = rho / p.q
// (as pcg_spmv_slices_kernel).
template <int kWL, int kWU>
__global__ __launch_bounds__(kThreads) void pcg_spmv_sym_slices_kernel(PcgBuffers b, SymSliceView v,
                                                                       const double* __restrict__ planes64,
                                                                       const float* __restrict__ planes32) {
    if (b.st->done) return;
    __shared__ double sh[kThreads / kWave];
    const bool use32 = planes32 != nullptr && (planes64 == nullptr || b.st->phase32 != 0);
    const int lb = xcd_block(blockIdx.x, gridDim.x);
    const int per_block = (v.n_slices + gridDim.x - 1) / gridDim.x;
    const int start = lb * per_block;
    const int stop = min(start + per_block, v.n_slices);
    const dpair* __restrict__ p2 = reinterpret_cast<const dpair*>(b.p);
    double pq = 0.0;
    for (int s = start + static_cast<int>(threadIdx.x) / kWave; s < stop; s += kThreads / kWave) {
        double a0 = 0.0, a1 = 0.0;
        if (use32)
            sym_slice_rows<float, kWL, kWU>(v, planes32, p2, s, a0, a1);
        else
            sym_slice_rows<double, kWL, kWU>(v, planes64, p2, s, a0, a1);
        const int node = s * kWave + (static_cast<int>(threadIdx.x) & (kWave - 1));
        if (node < (b.n >> 1)) {
            const dpair pv = p2[node];
            reinterpret_cast<dpair*>(b.q)[node] = dpair{a0, a1};
            pq += pv.x * a0 + pv.y * a1;
        }
    }
    const double vals[1] = {block_sum<kThreads>(pq, sh)};
    double tot[1];
    if (!finish_reduction<kThreads, 1>(b.tickets + 1, b.partials, vals, tot, sh)) return;
    if (threadIdx.x == 0) {
        const double tpq = tot[0];
        PcgState* st = b.st;
        st->pq = tpq;
        if (!(tpq > 0.0) || !isfinite(tpq)) {  // solver.cpp:591-593
            st->status = LFG_ERR_NOT_SPD;
            st->done = 1;
        } else {
            st->alpha = st->rho / tpq;
        }
    }
}

// The multiply launch of variant 15: pcg_slices_multiply_kernel's sums and state update
// with q' = A p' on the symmetric node slices. kMixed: the FP32 copy after the switch (an
// FP64 solve launches the instantiation without that path, which holds fewer registers).
template <int kWL, int kWU, int Prec, bool kMixed>
__global__ __launch_bounds__(kThreads) void pcg_sym_slices_multiply_kernel(PcgBuffers b, SymSliceView v,
                                                                           const double* __restrict__ planes64,
                                                                           const float* __restrict__ planes32,
                                                                           const double* __restrict__ r_new,
                                                                           const double* __restrict__ p_new) {
    if (b.st->done) return;
    __shared__ double sh[3 * (kThreads / kWave)];
    const double alpha = b.st->alpha;
    const bool use32 = kMixed && planes32 != nullptr && (planes64 == nullptr || b.st->phase32 != 0);
    constexpr int kWaves = kThreads / kWave;
    const int lb = xcd_block(blockIdx.x, gridDim.x);
    const int per_block = (v.n_slices + gridDim.x - 1) / gridDim.x;
    const int start = lb * per_block;
    const int stop = min(start + per_block, v.n_slices);
    const dpair* __restrict__ p2 = reinterpret_cast<const dpair*>(p_new);
    double acc[3] = {0.0, 0.0, 0.0};  // p'.q', q'.z, q'.M^-1 q'
    int qi = -1;
    dpair qv = {0.0, 0.0};
    for (int s = start + static_cast<int>(threadIdx.x) / kWave; s < stop; s += kWaves) {
        const int node = s * kWave + (static_cast<int>(threadIdx.x) & (kWave - 1));
        const bool own = node < (b.n >> 1);
        const int ic = own ? node : 0;
        const dpair pn = p2[ic];
        const dpair rn = reinterpret_cast<const dpair*>(r_new)[ic];
        NodePrecond<Prec> mo;
        mo.load(b.m, ic);
        double a0 = 0.0, a1 = 0.0;
        if (use32)
            sym_slice_rows<float, kWL, kWU>(v, planes32, p2, s, a0, a1);
        else
            sym_slice_rows<double, kWL, kWU>(v, planes64, p2, s, a0, a1);
        if (own) {
            if (qi >= 0) reinterpret_cast<dpair*>(b.q)[qi] = qv;  // an earlier slice of this wavefront
            qi = node;
            qv = dpair{a0, a1};
            double z0, z1, m0, m1;
            mo.apply(rn.x, rn.y, z0, z1);
            mo.apply(a0, a1, m0, m1);
            acc[0] += pn.x * a0 + pn.y * a1;
            acc[1] += a0 * z0 + a1 * z1;
            acc[2] += a0 * m0 + a1 * m1;
        }
    }
    block_sums<kThreads, 3>(acc, sh);
    double tot[3];
    const bool last = finish_reduction<kThreads, 3>(b.tickets + 1, b.partials, acc, tot, sh);
    if (qi >= 0) reinterpret_cast<dpair*>(b.q)[qi] = qv;
    if (!last) return;
    if (threadIdx.x == 0) {
        const double all[5] = {b.st->rr_new, b.st->rz_new, tot[0], tot[1], tot[2]};
        fused_state_update(b.st, all, alpha);
    }
}

}  // namespace

void launch_pcg_fused_flush(hipStream_t s, const PcgBuffers& b, const double* p_odd) {
    const unsigned int need = blocks_for(b.n >> 1, kThreads);
    fused_flush_kernel<<<need < 4096u ? (need > 0 ? need : 1u) : 4096u, kThreads, 0, s>>>(b, p_odd);
}

void launch_pcg_fused_grid(hipStream_t s, const PcgBuffers& b, const StencilView& v, const double* planes64,
                           const float* planes32, const double* diag_c, const double* r_in, double* r_out,
                           const double* p_in, double* p_out, const double* q_in, double* q_out) {
    constexpr int R = kGridRows;
    const int L = v.row_len;
    const unsigned int tiles = static_cast<unsigned int>(((v.n_nodes + L - 1) / L + R - 1) / R) *
                               static_cast<unsigned int>((L + kWave - 1) / kWave);
    const unsigned int grid = round8(tiles);
    if (b.precond == LFG_PRECOND_BLOCK2)
        pcg_fused_grid_kernel<LFG_PRECOND_BLOCK2, R><<<grid, R * kWave, 0, s>>>(b, v, planes64, planes32, diag_c,
                                                                                r_in, r_out, p_in, p_out, q_in, q_out);
    else
        pcg_fused_grid_kernel<LFG_PRECOND_JACOBI, R><<<grid, R * kWave, 0, s>>>(b, v, planes64, planes32, diag_c,
                                                                                r_in, r_out, p_in, p_out, q_in, q_out);
}

void launch_pcg_fused_slices(hipStream_t s, const PcgBuffers& b, const SliceView& v, const double* planes64,
                             const float* planes32, const double* r_in, double* r_out, const double* p_in,
                             double* p_out, const double* q_in, double* q_out) {
    PcgBuffers bq = b;
    bq.q = q_out;
    const unsigned int need = blocks_for(v.n_slices, kThreads / kWave);
    const unsigned int grid = round8(need < 8192u ? need : 8192u);
    const bool narrow = v.max_width <= 9;
    if (b.precond == LFG_PRECOND_BLOCK2) {
        if (narrow)
            pcg_fused_slices_kernel<9, LFG_PRECOND_BLOCK2><<<grid, kThreads, 0, s>>>(bq, v, planes64, planes32, r_in,
                                                                                   r_out, p_in, p_out, q_in);
        else
            pcg_fused_slices_kernel<0, LFG_PRECOND_BLOCK2><<<grid, kThreads, 0, s>>>(bq, v, planes64, planes32, r_in,
                                                                                   r_out, p_in, p_out, q_in);
    } else {
        if (narrow)
            pcg_fused_slices_kernel<9, LFG_PRECOND_JACOBI><<<grid, kThreads, 0, s>>>(bq, v, planes64, planes32, r_in,
                                                                                   r_out, p_in, p_out, q_in);
        else
            pcg_fused_slices_kernel<0, LFG_PRECOND_JACOBI><<<grid, kThreads, 0, s>>>(bq, v, planes64, planes32, r_in,
                                                                                   r_out, p_in, p_out, q_in);
    }
}

void launch_pcg_fold_slices(hipStream_t s, const PcgBuffers& b, const SliceView& v, const double* planes64,
                            const float* planes32, const double* r_in, double* r_out, const double* p_in,
                            double* p_out, const double* q_in, double* q_out, const PcgState* st_in,
                            PcgState* st_out, const double* part_in, double* part_out) {
    PcgBuffers bq = b;
    bq.q = q_out;
    const unsigned int grid = round8(blocks_for(v.n_slices, kFoldThreads / kWave));  // one slice per wavefront
    if (grid > static_cast<unsigned int>(kFoldMaxBlocks))
        throw std::runtime_error("the folded node-slice iteration takes at most 1,024 slices");
    const bool narrow = v.max_width <= 9;
    if (b.precond == LFG_PRECOND_BLOCK2) {
        if (narrow)
            pcg_fold_slices_kernel<9, LFG_PRECOND_BLOCK2><<<grid, kFoldThreads, 0, s>>>(
                bq, v, planes64, planes32, r_in, r_out, p_in, p_out, q_in, st_in, st_out, part_in, part_out);
        else
            pcg_fold_slices_kernel<0, LFG_PRECOND_BLOCK2><<<grid, kFoldThreads, 0, s>>>(
                bq, v, planes64, planes32, r_in, r_out, p_in, p_out, q_in, st_in, st_out, part_in, part_out);
    } else {
        if (narrow)
            pcg_fold_slices_kernel<9, LFG_PRECOND_JACOBI><<<grid, kFoldThreads, 0, s>>>(
                bq, v, planes64, planes32, r_in, r_out, p_in, p_out, q_in, st_in, st_out, part_in, part_out);
        else
            pcg_fold_slices_kernel<0, LFG_PRECOND_JACOBI><<<grid, kFoldThreads, 0, s>>>(
                bq, v, planes64, planes32, r_in, r_out, p_in, p_out, q_in, st_in, st_out, part_in, part_out);
    }
}

void launch_pcg_split_slices(hipStream_t s, const PcgBuffers& b, const SliceView& v, const double* planes64,
                             const float* planes32, const double* r_in, double* r_out, const double* p_in,
                             double* p_out, const double* q_in, double* q_out) {
    PcgBuffers bq = b;
    bq.q = q_out;
    const unsigned int need = blocks_for(v.n_slices, kThreads / kWave);
    const unsigned int grid = round8(need < 8192u ? need : 8192u);
    const bool narrow = v.max_width <= 9;
    if (b.precond == LFG_PRECOND_BLOCK2) {
        pcg_slices_vector_kernel<LFG_PRECOND_BLOCK2><<<grid, kThreads, 0, s>>>(b, r_in, r_out, p_in, p_out, q_in);
        if (narrow)
            pcg_slices_multiply_kernel<9, LFG_PRECOND_BLOCK2><<<grid, kThreads, 0, s>>>(bq, v, planes64, planes32,
                                                                                      r_out, p_out);
        else
            pcg_slices_multiply_kernel<0, LFG_PRECOND_BLOCK2><<<grid, kThreads, 0, s>>>(bq, v, planes64, planes32,
                                                                                      r_out, p_out);
    } else {
        pcg_slices_vector_kernel<LFG_PRECOND_JACOBI><<<grid, kThreads, 0, s>>>(b, r_in, r_out, p_in, p_out, q_in);
        if (narrow)
            pcg_slices_multiply_kernel<9, LFG_PRECOND_JACOBI><<<grid, kThreads, 0, s>>>(bq, v, planes64, planes32,
                                                                                      r_out, p_out);
        else
            pcg_slices_multiply_kernel<0, LFG_PRECOND_JACOBI><<<grid, kThreads, 0, s>>>(bq, v, planes64, planes32,
                                                                                      r_out, p_out);
    }
}

void launch_pack_sym_slices(hipStream_t s, const CsrView& a, const SymSliceView& v, const double* values,
                            double* planes64, float* planes32) {
    const long long threads = static_cast<long long>(v.n_slices) * kWave;
    const unsigned int grid = blocks_for(threads > 0 ? threads : 1, 256);
    if (planes64)
        pack_sym_slices_kernel<double><<<grid, 256, 0, s>>>(a.n >> 1, a.row_ptr, a.col, v, values, planes64);
    if (planes32) {
        pack_sym_slices_kernel<float><<<grid, 256, 0, s>>>(a.n >> 1, a.row_ptr, a.col, v, values, planes32);
        if (a.n > 0)
            sym_diag32_kernel<<<blocks_for(a.n >> 1, 256), 256, 0, s>>>(a.n >> 1, a.row_ptr, a.col, v, values,
                                                                        planes32, const_cast<double*>(v.diag32));
    }
}

void launch_pcg_spmv_sym_slices(hipStream_t s, const PcgBuffers& b, const SymSliceView& v, const double* planes64,
                                const float* planes32) {
    const unsigned int grid = static_cast<unsigned int>(slice_spmv_blocks(v.n_slices));
    if (v.max_lwidth <= 4 && v.max_uwidth <= 5)
        pcg_spmv_sym_slices_kernel<4, 5><<<grid, kThreads, 0, s>>>(b, v, planes64, planes32);
    else if (v.max_lwidth <= 8 && v.max_uwidth <= 9)
        pcg_spmv_sym_slices_kernel<8, 9><<<grid, kThreads, 0, s>>>(b, v, planes64, planes32);
    else
        pcg_spmv_sym_slices_kernel<0, 0><<<grid, kThreads, 0, s>>>(b, v, planes64, planes32);
}

template <int Prec, bool kMixed>
static void launch_sym_multiply(hipStream_t s, unsigned int grid, const PcgBuffers& bq, const SymSliceView& v,
                                const double* planes64, const float* planes32, const double* r_out,
                                const double* p_out) {
    if (v.max_lwidth <= 4 && v.max_uwidth <= 5)
        pcg_sym_slices_multiply_kernel<4, 5, Prec, kMixed><<<grid, kThreads, 0, s>>>(bq, v, planes64, planes32, r_out,
                                                                                     p_out);
    else if (v.max_lwidth <= 8 && v.max_uwidth <= 9)
        pcg_sym_slices_multiply_kernel<8, 9, Prec, kMixed><<<grid, kThreads, 0, s>>>(bq, v, planes64, planes32, r_out,
                                                                                     p_out);
    else
        pcg_sym_slices_multiply_kernel<0, 0, Prec, kMixed><<<grid, kThreads, 0, s>>>(bq, v, planes64, planes32, r_out,
                                                                                     p_out);
}

template <int Prec>
static void launch_sym_split(hipStream_t s, unsigned int grid, const PcgBuffers& b, const PcgBuffers& bq,
                             const SymSliceView& v, const double* planes64, const float* planes32,
                             const double* r_in, double* r_out, const double* p_in, double* p_out,
                             const double* q_in) {
    pcg_slices_vector_kernel<Prec><<<grid, kThreads, 0, s>>>(b, r_in, r_out, p_in, p_out, q_in);
    if (planes32)
        launch_sym_multiply<Prec, true>(s, grid, bq, v, planes64, planes32, r_out, p_out);
    else
        launch_sym_multiply<Prec, false>(s, grid, bq, v, planes64, planes32, r_out, p_out);
}

void launch_pcg_sym_split_slices(hipStream_t s, const PcgBuffers& b, const SymSliceView& v, const double* planes64,
                                 const float* planes32, const double* r_in, double* r_out, const double* p_in,
                                 double* p_out, const double* q_in, double* q_out) {
    PcgBuffers bq = b;
    bq.q = q_out;
    // the vector launch's 256-node ranges are the multiply's 4-slice ranges (same grid)
    const unsigned int need = blocks_for(v.n_slices, kThreads / kWave);
    const unsigned int grid = round8(need < 8192u ? need : 8192u);
    if (b.precond == LFG_PRECOND_BLOCK2)
        launch_sym_split<LFG_PRECOND_BLOCK2>(s, grid, b, bq, v, planes64, planes32, r_in, r_out, p_in, p_out, q_in);
    else
        launch_sym_split<LFG_PRECOND_JACOBI>(s, grid, b, bq, v, planes64, planes32, r_in, r_out, p_in, p_out, q_in);
}

void launch_pack_stencil(hipStream_t s, const CsrView& a, const StencilView& v, const double* values, double* out64,
                         float* out32, double* diag_c) {
    if (v.n_nodes <= 0) return;
    pack_stencil_kernel<<<blocks_for(v.n_nodes, 256), 256, 0, s>>>(v.n_nodes, a.row_ptr, a.col, values, v, out64,
                                                                   out32, diag_c);
}

void launch_pcg_spmv_stencil(hipStream_t s, const PcgBuffers& b, const StencilView& v, const double* planes64,
                             const float* planes32, const double* diag_c) {
    if (stencil_grid_multiply(v)) {
        pcg_spmv_grid_kernel<kGridRows><<<stencil_grid_blocks(v), kGridRows * kWave, 0, s>>>(b, v, planes64, planes32,
                                                                                          diag_c);
        return;
    }
    const unsigned int grid = stencil_spmv_blocks(v.n_nodes);
    if (v.sym) {
        // Q4 meshes numbered row by row have 4 positive offsets; other counts take the
        // general loop
        if (v.n_off == 5)
            pcg_spmv_stencil_kernel<4, true><<<grid, kThreads, 0, s>>>(b, v, planes64, planes32, diag_c);
        else
            pcg_spmv_stencil_kernel<0, true><<<grid, kThreads, 0, s>>>(b, v, planes64, planes32, diag_c);
    } else {
        if (v.n_off == 9)
            pcg_spmv_stencil_kernel<9, false><<<grid, kThreads, 0, s>>>(b, v, planes64, planes32, diag_c);
        else
            pcg_spmv_stencil_kernel<0, false><<<grid, kThreads, 0, s>>>(b, v, planes64, planes32, diag_c);
    }
}

void launch_pcg_fused_stencil(hipStream_t s, const PcgBuffers& b, const StencilView& v, const double* planes64,
                              const float* planes32, const double* diag_c, const double* r_in, double* r_out,
                              const double* p_in, double* p_out, const double* q_in, double* q_out) {
    const unsigned int tiles = blocks_for(v.n_nodes, kFusedTile);
    const unsigned int grid = round8(tiles < 8192u ? tiles : 8192u);
    const size_t lds = static_cast<size_t>(v.win_total) * sizeof(dpair);
    if (v.n_off == 5) {
        if (b.precond == LFG_PRECOND_BLOCK2)
            pcg_fused_stencil_kernel<4, LFG_PRECOND_BLOCK2><<<grid, kFusedTile, lds, s>>>(
                b, v, planes64, planes32, diag_c, r_in, r_out, p_in, p_out, q_in, q_out);
        else
            pcg_fused_stencil_kernel<4, LFG_PRECOND_JACOBI><<<grid, kFusedTile, lds, s>>>(
                b, v, planes64, planes32, diag_c, r_in, r_out, p_in, p_out, q_in, q_out);
    } else {
        if (b.precond == LFG_PRECOND_BLOCK2)
            pcg_fused_stencil_kernel<0, LFG_PRECOND_BLOCK2><<<grid, kFusedTile, lds, s>>>(
                b, v, planes64, planes32, diag_c, r_in, r_out, p_in, p_out, q_in, q_out);
        else
            pcg_fused_stencil_kernel<0, LFG_PRECOND_JACOBI><<<grid, kFusedTile, lds, s>>>(
                b, v, planes64, planes32, diag_c, r_in, r_out, p_in, p_out, q_in, q_out);
    }
}

}  // namespace lfg
