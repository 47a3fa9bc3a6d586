// Batched Jacobi-PCG for many independent problems, each with NC load cases, on gfx950.
//
// The problems are one disjoint mesh whose nodes are grouped per problem in whole
// 64-node slices (the node-slice layout of the single-problem SpMV), so a wavefront
// always works inside one problem. Vectors hold the NC cases of a node together
// (x[node][case][2]): one load of a 2x2 block of K serves all NC right-hand sides, and
// the p entries the block multiplies arrive as NC consecutive pairs. Every system
// (problem, case) has its own CG scalars and convergence flag; per-problem sums
// (p.Ap, r.r, r.z) are written as one partial per slice and folded in slice order by
// small finish kernels (one wavefront per problem), so results are run-to-run identical.
// A problem whose cases have all stopped is skipped by every later kernel, and every
// kernel returns at once when all problems have stopped, so a captured graph of many
// iterations costs almost nothing after convergence.
//
// Mixed precision switches per problem: a problem multiplies with FP64 values until
// every running case's recurrence residual has fallen to switch_tau of its initial
// value, then with the FP32 off-diagonal blocks plus the FP64 compensated diagonal
// block (the rounding of each row's FP32 entries subtracted from its diagonal, so
// rigid translations still multiply to zero). Vectors and scalars stay FP64.
#include "lfg_kernels.hpp"

#include "lightfem_gpu.h"

namespace lfg {
namespace {

__device__ __forceinline__ int lane_id() { return static_cast<int>(threadIdx.x) & (kWave - 1); }

// Vector kernels map lane l of sub-step j to the (node, case) pair t = 64 j + l of the
// slice, so each load moves 64 consecutive pairs; the case of a lane is l % NC in every
// sub-step. Sum over the lanes of one case (NC a power of two); every lane gets its
// case's total.
template <int NC>
__device__ __forceinline__ double case_sum(double v) {
#pragma unroll
    for (int o = kWave / 2; o >= NC; o >>= 1) v += __shfl_xor(v, o, kWave);
    return v;
}

// Contiguous slice range of this block (XCD-remapped), waves taking every wpb-th slice.
struct SliceRange {
    int first, stop, step;
};

// This block's share of the active slice list (XCD-remapped contiguous ranges, as
// block_slices): the iteration kernels walk only the running problems' slices, so a
// launch costs what the remaining work costs, however many problems have stopped.
__device__ __forceinline__ SliceRange active_range(const BatchView& bv) {
    const int na = bv.global->n_active;
    const int lb = xcd_block(blockIdx.x, gridDim.x);
    const int per_block = (na + gridDim.x - 1) / gridDim.x;
    const int start = lb * per_block;
    const int stop = min(start + per_block, na);
    return SliceRange{start + static_cast<int>(threadIdx.x) / kWave, stop, static_cast<int>(blockDim.x) / kWave};
}

__device__ __forceinline__ SliceRange block_slices(int n_slices) {
    const int lb = xcd_block(blockIdx.x, gridDim.x);
    const int per_block = (n_slices + gridDim.x - 1) / gridDim.x;
    const int start = lb * per_block;
    const int stop = min(start + per_block, n_slices);
    return SliceRange{start + static_cast<int>(threadIdx.x) / kWave, stop, static_cast<int>(blockDim.x) / kWave};
}

// Per-problem reduction of K values, finished by the problem's last wavefront. Lane 0
// of every wave holds the slice's values in vals; returns true (in every lane of the
// wave) in the last wave, with the totals in out (all lanes).
template <int K>
__device__ __forceinline__ bool finish_problem(const BatchView& bv, int p, int s, const double* vals, double* out) {
    const int first = bv.pslice_ptr[p];
    const int count = bv.pslice_ptr[p + 1] - first;
    double* part = bv.partials + static_cast<long long>(first) * K;
    unsigned int* ticket = bv.tickets + 32 * p;
    int last = 0;
    if (lane_id() == 0) {
#pragma unroll
        for (int k = 0; k < K; ++k) publish_partial(part + (s - first) * K + k, vals[k]);
        stores_complete();
        const unsigned int a = __hip_atomic_fetch_add(ticket, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        last = a == static_cast<unsigned int>(count - 1);
    }
    last = __shfl(last, 0, kWave);
    if (!last) return false;
#pragma unroll
    for (int k = 0; k < K; ++k) {
        double acc = 0.0;
        for (int i = lane_id(); i < count; i += kWave) acc += read_partial(part + i * K + k);
        out[k] = wave_sum(acc);
    }
    if (lane_id() == 0) __hip_atomic_store(ticket, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    return true;
}

// Marks problem p stopped once all its cases are; the last problem to stop ends the batch.
__device__ __forceinline__ void close_problem_if_done(const BatchView& bv, int p) {
    for (int c = 0; c < bv.n_cases; ++c)
        if (!bv.sys[p * bv.n_cases + c].done) return;
    bv.problem_done[p] = 1;
    __hip_atomic_store(&bv.global->changed, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const unsigned int before =
        __hip_atomic_fetch_add(&bv.global->problems_done, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (before + 1 == static_cast<unsigned int>(bv.n_problems))
        __hip_atomic_store(&bv.global->done_all, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// ------------------------------------------------------------ init

// x = 0, r = b, p = z = M^-1 r; per system ||r0||, target, rho (solver.cpp:560-583 per case).
template <int NC>
__global__ __launch_bounds__(512) void batch_init_kernel(BatchView bv) {
    const SliceRange range = block_slices(bv.n_slices);
    const dpair* __restrict__ b2 = reinterpret_cast<const dpair*>(bv.b);
    const dpair* __restrict__ m2 = reinterpret_cast<const dpair*>(bv.minv);
    for (int s = range.first; s < range.stop; s += range.step) {
        const int p = bv.slice_problem[s];
        if (lane_id() == 0) bv.active[s] = s;  // every problem starts running
        double rr = 0.0, rz = 0.0;
#pragma unroll
        for (int j = 0; j < NC; ++j) {
            const int t = j * kWave + lane_id();
            const long long pair = static_cast<long long>(s) * kWave * NC + t;
            const dpair m = m2[s * kWave + t / NC];
            const dpair r = b2[pair];
            const dpair z = {m.x * r.x, m.y * r.y};
            const dpair zero = {0.0, 0.0};
            reinterpret_cast<dpair*>(bv.x)[pair] = zero;
            reinterpret_cast<dpair*>(bv.r)[pair] = r;
            reinterpret_cast<dpair*>(bv.p)[pair] = z;
            rr += r.x * r.x + r.y * r.y;
            rz += r.x * z.x + r.y * z.y;
        }
        rr = case_sum<NC>(rr);
        rz = case_sum<NC>(rz);
        double vals[2 * NC];
#pragma unroll
        for (int k = 0; k < NC; ++k) {
            vals[2 * k] = __shfl(rr, k, kWave);
            vals[2 * k + 1] = __shfl(rz, k, kWave);
        }
        double tot[2 * NC];
        if (!finish_problem<2 * NC>(bv, p, s, vals, tot)) continue;
        if (lane_id() == 0) {
            for (int c = 0; c < NC; ++c) {
                BatchSystem* st = bv.sys + p * NC + c;
                const double rr = tot[2 * c], rz = tot[2 * c + 1];
                const double rnorm0 = sqrt(max0(rr));
                const double base = bv.stop_rule == LFG_STOP_REFERENCE ? ((1.0 < rnorm0) ? rnorm0 : 1.0) : rnorm0;
                const double target = bv.tol * base;
                st->rnorm0 = rnorm0;
                st->target = target;
                st->switch_iter = (bv.problem_phase32 && bv.switch_tau < 0.0) ? 0 : -1;
                st->rr = rr;
                st->rho = rz;
                st->alpha = 0.0;
                st->beta = 0.0;
                st->iter = 0;
                st->status = LFG_OK;
                st->converged = 0;
                st->done = 0;
                if (rnorm0 <= target) {
                    st->converged = 1;
                    st->done = 1;
                } else if (!(rz > 0.0) || !isfinite(rz)) {
                    st->status = LFG_ERR_NONFINITE;
                    st->done = 1;
                }
            }
            bv.problem_done[p] = 0;
            if (bv.problem_phase32) bv.problem_phase32[p] = bv.switch_tau < 0.0 ? 1 : 0;
            close_problem_if_done(bv, p);
        }
    }
}

// ------------------------------------------------------------ SpMV

// q = K p for the NC cases of the slice's nodes; rows summed in column order from
// 0.0 as the CSR row loop (sparse.cpp:104-120). Returns the per-case p.q of the lane.
template <int NC, typename VT, int kW>
__device__ __forceinline__ void slice_multiply(const BatchView& bv, const SliceView& v, const VT* __restrict__ planes,
                                               int s, double* pq, dpair* qout) {
    using P = typename PairOf<VT>::type;
    const int node = s * kWave + lane_id();
    const int begin = v.ptr[s];
    const int width = (v.ptr[s + 1] - begin) / kWave;
    const P* __restrict__ upper = reinterpret_cast<const P*>(planes);
    const P* __restrict__ lower = upper + v.blocks;
    const dpair* __restrict__ p2 = reinterpret_cast<const dpair*>(bv.p);
    double a0[NC], a1[NC];
#pragma unroll
    for (int c = 0; c < NC; ++c) a0[c] = a1[c] = 0.0;
    if (kW > 0 && width > 0 && width <= kW) {
        P u[kW > 0 ? kW : 1], l[kW > 0 ? kW : 1];
        int col[kW > 0 ? kW : 1];
#pragma unroll
        for (int k = 0; k < kW; ++k) {
            const int at = begin + min(k, width - 1) * kWave + lane_id();
            u[k] = __builtin_nontemporal_load(upper + at);
            l[k] = __builtin_nontemporal_load(lower + at);
            col[k] = __builtin_nontemporal_load(v.col + at);
        }
#pragma unroll
        for (int k = 0; k < kW; ++k) {
            if (k < width) {
                const dpair* pc = p2 + static_cast<long long>(col[k]) * NC;
#pragma unroll
                for (int c = 0; c < NC; ++c) {
                    const dpair x = pc[c];
                    a0[c] += static_cast<double>(u[k].x) * x.x;
                    a0[c] += static_cast<double>(u[k].y) * x.y;
                    a1[c] += static_cast<double>(l[k].x) * x.x;
                    a1[c] += static_cast<double>(l[k].y) * x.y;
                }
            }
        }
    } else {
        for (int k = 0; k < width; ++k) {
            const int at = begin + k * kWave + lane_id();
            const P uk = __builtin_nontemporal_load(upper + at);
            const P lk = __builtin_nontemporal_load(lower + at);
            const dpair* pc = p2 + static_cast<long long>(__builtin_nontemporal_load(v.col + at)) * NC;
#pragma unroll
            for (int c = 0; c < NC; ++c) {
                const dpair x = pc[c];
                a0[c] += static_cast<double>(uk.x) * x.x;
                a0[c] += static_cast<double>(uk.y) * x.y;
                a1[c] += static_cast<double>(lk.x) * x.x;
                a1[c] += static_cast<double>(lk.y) * x.y;
            }
        }
    }
    const dpair* pn = p2 + static_cast<long long>(node) * NC;
    if (sizeof(VT) == sizeof(float)) {
        // FP32 planes hold a zero diagonal block: add the FP64 compensated one
        const dpair du = reinterpret_cast<const dpair*>(v.diag32)[2 * node];
        const dpair dl = reinterpret_cast<const dpair*>(v.diag32)[2 * node + 1];
#pragma unroll
        for (int c = 0; c < NC; ++c) {
            const dpair x = pn[c];
            a0[c] += du.x * x.x;
            a0[c] += du.y * x.y;
            a1[c] += dl.x * x.x;
            a1[c] += dl.y * x.y;
        }
    }
#pragma unroll
    for (int c = 0; c < NC; ++c) {
        qout[c] = dpair{a0[c], a1[c]};
        const dpair pv = pn[c];
        pq[c] = pv.x * a0[c] + pv.y * a1[c];
    }
}

// q of slice s (the NC cases of each node) into dst.
template <int NC>
__device__ __forceinline__ void store_slice(double* dst, int s, const dpair* qv) {
    dpair* d = reinterpret_cast<dpair*>(dst) + static_cast<long long>(s * kWave + lane_id()) * NC;
#pragma unroll
    for (int c = 0; c < NC; ++c) d[c] = qv[c];
}

// The same product on the symmetric node slices (SymSliceView, FP64 values): each row
// adds its lower blocks (the transposes of the column nodes' upper blocks, ascending
// columns) and then its upper blocks, every block applied to the NC case pairs of its
// column node; within rounding of slice_multiply (a mirrored block differs only by Ke's
// (a, b) / (b, a) rounding), with each pair of mirrored blocks streamed once. kWL, kWU > 0:
// lists up to that wide take an unrolled path with unconditional (clamped) loads.
template <int NC, int kWL, int kWU>
__device__ __forceinline__ void sym_slice_multiply(const BatchView& bv, const SymSliceView& v,
                                                   const double* __restrict__ planes, int s, double* pq,
                                                   dpair* qout) {
    const int node = s * kWave + lane_id();
    const dpair* __restrict__ up = reinterpret_cast<const dpair*>(planes);
    const dpair* __restrict__ lo = up + (v.ublocks + 1);
    const dpair* __restrict__ p2 = reinterpret_cast<const dpair*>(bv.p);
    double a0[NC], a1[NC];
#pragma unroll
    for (int c = 0; c < NC; ++c) a0[c] = a1[c] = 0.0;
    // block (rows u, l) times the NC pairs of node col
    auto apply = [&](double ux, double uy, double lx, double ly, int col) {
        const dpair* pc = p2 + static_cast<long long>(col) * NC;
#pragma unroll
        for (int c = 0; c < NC; ++c) {
            const dpair x = pc[c];
            a0[c] += ux * x.x;
            a0[c] += uy * x.y;
            a1[c] += lx * x.x;
            a1[c] += ly * x.y;
        }
    };
    const int lb = v.lptr[s];
    const int lw = (v.lptr[s + 1] - lb) / kWave;
    const int ub = v.uptr[s];
    const int uw = (v.uptr[s + 1] - ub) / kWave;
    if (kWL > 0 && kWU > 0 && lw > 0 && lw <= kWL && uw > 0 && uw <= kWU) {
        int lc[kWL > 0 ? kWL : 1], lt[kWL > 0 ? kWL : 1], uc[kWU > 0 ? kWU : 1];
        dpair tu[kWL > 0 ? kWL : 1], tl[kWL > 0 ? kWL : 1], uu[kWU > 0 ? kWU : 1], ul[kWU > 0 ? kWU : 1];
#pragma unroll
        for (int k = 0; k < kWL; ++k) {
            const int at = lb + min(k, lw - 1) * kWave + lane_id();
            lc[k] = __builtin_nontemporal_load(v.lcol + at);
            lt[k] = __builtin_nontemporal_load(v.ltpos + at);
        }
#pragma unroll
        for (int k = 0; k < kWU; ++k) {
            const int at = ub + min(k, uw - 1) * kWave + lane_id();
            uc[k] = __builtin_nontemporal_load(v.ucol + at);
            uu[k] = up[at];
            ul[k] = lo[at];
        }
#pragma unroll
        for (int k = 0; k < kWL; ++k) {
            tu[k] = up[lt[k]];
            tl[k] = lo[lt[k]];
        }
#pragma unroll
        for (int k = 0; k < kWL; ++k)
            if (k < lw) apply(tu[k].x, tl[k].x, tu[k].y, tl[k].y, lc[k]);
#pragma unroll
        for (int k = 0; k < kWU; ++k)
            if (k < uw) apply(uu[k].x, uu[k].y, ul[k].x, ul[k].y, uc[k]);
    } else {
        for (int k = 0; k < lw; ++k) {
            const int at = lb + k * kWave + lane_id();
            const int tp = __builtin_nontemporal_load(v.ltpos + at);
            const dpair tu = up[tp], tl = lo[tp];
            apply(tu.x, tl.x, tu.y, tl.y, __builtin_nontemporal_load(v.lcol + at));
        }
        for (int k = 0; k < uw; ++k) {
            const int at = ub + k * kWave + lane_id();
            const dpair uu = up[at], ul = lo[at];
            apply(uu.x, uu.y, ul.x, ul.y, __builtin_nontemporal_load(v.ucol + at));
        }
    }
    const dpair* pn = p2 + static_cast<long long>(node) * NC;
#pragma unroll
    for (int c = 0; c < NC; ++c) {
        qout[c] = dpair{a0[c], a1[c]};
        const dpair pv = pn[c];
        pq[c] = pv.x * a0[c] + pv.y * a1[c];
    }
}

// The FP64 product of one slice: the symmetric node slices where given (kSym: 1 = the
// unrolled lists of a Q4 mesh, up to 4 lower and 5 upper blocks; 2 = any widths), else
// the node slices.
template <int NC, int kW, int kSym>
__device__ __forceinline__ void fp64_multiply(const BatchView& bv, const SliceView& v, const SymSliceView& sv,
                                              const double* __restrict__ planes64, int s, double* pq, dpair* qv) {
    if (kSym == 1) sym_slice_multiply<NC, 4, 5>(bv, sv, planes64, s, pq, qv);
    else if (kSym == 2) sym_slice_multiply<NC, 0, 0>(bv, sv, planes64, s, pq, qv);
    else slice_multiply<NC, double, kW>(bv, v, planes64, s, pq, qv);
}

// q = K p and the slice's per-case p.q partials (slot s of the partials); the problem's
// sums and alpha follow in batch_spmv_finish_kernel. kMixed: each problem's phase picks
// its values (FP64 planes, or FP32 planes + compensated diagonal); kSym > 0: the FP64
// values are the symmetric node slices' (planes64 = their planes).
template <int NC, int kW, bool kMixed, int kSym>
__global__ __launch_bounds__(512) void batch_spmv_kernel(BatchView bv, SliceView v, SymSliceView sv,
                                                         const double* __restrict__ planes64,
                                                         const float* __restrict__ planes32) {
    if (bv.global->done_all) return;
    const SliceRange range = active_range(bv);
    for (int a = range.first; a < range.stop; a += range.step) {
        const int s = bv.active[a];
        const int p = bv.slice_problem[s];
        if (bv.problem_done[p]) continue;
        double pq[NC];
        dpair qv[NC];
        if (kMixed && bv.problem_phase32[p]) slice_multiply<NC, float, kW>(bv, v, planes32, s, pq, qv);
        else fp64_multiply<NC, kW, kSym>(bv, v, sv, planes64, s, pq, qv);
        store_slice<NC>(bv.q, s, qv);
        double vals[NC];
#pragma unroll
        for (int c = 0; c < NC; ++c) vals[c] = wave_sum(pq[c]);
        if (lane_id() == 0) {
#pragma unroll
            for (int c = 0; c < NC; ++c) bv.partials[static_cast<long long>(s) * NC + c] = vals[c];
        }
    }
}

// The sum over problem p's slices of K values per slice (slots s * K + k of the partials),
// folded in slice order by one wavefront (lane l takes slices l, l + 64, ...), the fold
// of finish_problem: run-to-run and launch-shape identical.
template <int K>
__device__ __forceinline__ void fold_problem(const BatchView& bv, int p, double* out) {
    const int first = bv.pslice_ptr[p];
    const int count = bv.pslice_ptr[p + 1] - first;
    const double* part = bv.partials + static_cast<long long>(first) * K;
#pragma unroll
    for (int k = 0; k < K; ++k) {
        double acc = 0.0;
        for (int i = lane_id(); i < count; i += kWave) acc += part[i * K + k];
        out[k] = wave_sum(acc);
    }
}

// Per system alpha = rho / p.q (solver.cpp:585-595), one wavefront per running problem.
template <int NC>
__global__ __launch_bounds__(256) void batch_spmv_finish_kernel(BatchView bv) {
    if (bv.global->done_all) return;
    const int p = static_cast<int>((blockIdx.x * blockDim.x + threadIdx.x) / kWave);
    if (p >= bv.n_problems || bv.problem_done[p]) return;
    double tot[NC];
    fold_problem<NC>(bv, p, tot);
    if (lane_id() != 0) return;
    for (int c = 0; c < NC; ++c) {
        BatchSystem* st = bv.sys + p * NC + c;
        if (st->done) continue;
        const double tpq = tot[c];
        if (!(tpq > 0.0) || !isfinite(tpq)) {
            st->status = LFG_ERR_NOT_SPD;
            st->done = 1;
        } else {
            st->alpha = st->rho / tpq;
        }
    }
    close_problem_if_done(bv, p);
}

// ------------------------------------------------------------ update and direction

// x += alpha p, r -= alpha q, z = M^-1 r; the slice's per-case ||r||^2 and r.z partials
// (slots 2 NC s .. of the partials); the stop test and beta follow in
// batch_update_finish_kernel (solver.cpp:597-621).
template <int NC>
__global__ __launch_bounds__(512) void batch_update_kernel(BatchView bv) {
    if (bv.global->done_all) return;
    const SliceRange range = active_range(bv);
    const int my_case = lane_id() % NC;
    const dpair* __restrict__ m2 = reinterpret_cast<const dpair*>(bv.minv);
    dpair* __restrict__ x2 = reinterpret_cast<dpair*>(bv.x);
    dpair* __restrict__ r2 = reinterpret_cast<dpair*>(bv.r);
    const dpair* __restrict__ p2 = reinterpret_cast<const dpair*>(bv.p);
    const dpair* __restrict__ q2 = reinterpret_cast<const dpair*>(bv.q);
    for (int a = range.first; a < range.stop; a += range.step) {
        const int s = bv.active[a];
        const int p = bv.slice_problem[s];
        if (bv.problem_done[p]) continue;
        const BatchSystem* st = bv.sys + p * NC + my_case;
        const bool running = !st->done;
        const double alpha = st->alpha;
        double rr = 0.0, rz = 0.0;
        if (running) {
#pragma unroll
            for (int j = 0; j < NC; ++j) {
                const int t = j * kWave + lane_id();
                const long long pair = static_cast<long long>(s) * kWave * NC + t;
                const dpair m = m2[s * kWave + t / NC];
                const dpair pv = p2[pair], qv = q2[pair];
                dpair xv = x2[pair], rv = r2[pair];
                xv.x += alpha * pv.x;
                xv.y += alpha * pv.y;
                rv.x = rv.x - alpha * qv.x;
                rv.y = rv.y - alpha * qv.y;
                x2[pair] = xv;
                r2[pair] = rv;
                rr += rv.x * rv.x + rv.y * rv.y;
                rz += rv.x * (m.x * rv.x) + rv.y * (m.y * rv.y);
            }
        }
        rr = case_sum<NC>(rr);
        rz = case_sum<NC>(rz);
        double vals[2 * NC];
#pragma unroll
        for (int k = 0; k < NC; ++k) {
            vals[2 * k] = __shfl(rr, k, kWave);
            vals[2 * k + 1] = __shfl(rz, k, kWave);
        }
        if (lane_id() == 0) {
#pragma unroll
            for (int k = 0; k < 2 * NC; ++k) bv.partials[static_cast<long long>(s) * 2 * NC + k] = vals[k];
        }
    }
}

// Per system the stop test on the recurrence residual and beta, the mixed-precision
// phase switch, one wavefront per running problem.
template <int NC>
__global__ __launch_bounds__(256) void batch_update_finish_kernel(BatchView bv) {
    if (bv.global->done_all) return;
    const int p = static_cast<int>((blockIdx.x * blockDim.x + threadIdx.x) / kWave);
    if (p >= bv.n_problems || bv.problem_done[p]) return;
    double tot[2 * NC];
    fold_problem<2 * NC>(bv, p, tot);
    if (lane_id() != 0) return;
    for (int c = 0; c < NC; ++c) {
        BatchSystem* st = bv.sys + p * NC + c;
        if (st->done) continue;
        st->iter += 1;
        st->rr = tot[2 * c];
        const double rz = tot[2 * c + 1];
        if (sqrt(max0(st->rr)) <= st->target) {
            st->converged = 1;
            st->done = 1;
        } else if (!(rz > 0.0) || !isfinite(rz)) {
            st->status = LFG_ERR_NONFINITE;
            st->done = 1;
        } else {
            st->beta = rz / st->rho;
            st->rho = rz;
            if (st->iter >= bv.max_iter) st->done = 1;
        }
    }
    if (bv.problem_phase32 && !bv.problem_phase32[p]) {
        bool reached = true;
        for (int c = 0; c < NC; ++c) {
            const BatchSystem* st = bv.sys + p * NC + c;
            reached = reached && (st->done || sqrt(max0(st->rr)) <= bv.switch_tau * st->rnorm0);
        }
        if (reached) {
            bv.problem_phase32[p] = 1;
            for (int c = 0; c < NC; ++c) bv.sys[p * NC + c].switch_iter = bv.sys[p * NC + c].iter;
        }
    }
    close_problem_if_done(bv, p);
}

// p = z + beta p for the systems still running.
template <int NC>
__global__ __launch_bounds__(512) void batch_direction_kernel(BatchView bv) {
    if (bv.global->done_all) return;
    const SliceRange range = active_range(bv);
    const int my_case = lane_id() % NC;
    const dpair* __restrict__ m2 = reinterpret_cast<const dpair*>(bv.minv);
    const dpair* __restrict__ r2 = reinterpret_cast<const dpair*>(bv.r);
    dpair* __restrict__ p2 = reinterpret_cast<dpair*>(bv.p);
    for (int a = range.first; a < range.stop; a += range.step) {
        const int s = bv.active[a];
        const int p = bv.slice_problem[s];
        if (bv.problem_done[p]) continue;
        const BatchSystem* st = bv.sys + p * NC + my_case;
        if (st->done) continue;
        const double beta = st->beta;
#pragma unroll
        for (int j = 0; j < NC; ++j) {
            const int t = j * kWave + lane_id();
            const long long pair = static_cast<long long>(s) * kWave * NC + t;
            const dpair m = m2[s * kWave + t / NC];
            const dpair rv = r2[pair];
            dpair pv = p2[pair];
            pv.x = m.x * rv.x + beta * pv.x;
            pv.y = m.y * rv.y + beta * pv.y;
            p2[pair] = pv;
        }
    }
}

// ------------------------------------------------------------ active slice list

constexpr int kScanThreads = 256;

// Positions of the running problems' slices in the new list (an exclusive scan of their
// slice counts in problem order) and its length; nothing when no problem stopped since
// the last compaction.
__global__ __launch_bounds__(kScanThreads) void batch_compact_scan_kernel(BatchView bv) {
    BatchGlobal* g = bv.global;
    const int t = static_cast<int>(threadIdx.x);
    if (g->done_all || !g->changed) {
        if (t == 0) g->compact_now = 0;
        return;
    }
    __shared__ int sh[kScanThreads];
    int total = 0;
    for (int base = 0; base < bv.n_problems; base += kScanThreads) {
        const int p = base + t;
        const int cnt = (p < bv.n_problems && !bv.problem_done[p]) ? bv.pslice_ptr[p + 1] - bv.pslice_ptr[p] : 0;
        sh[t] = cnt;
        __syncthreads();
        for (int o = 1; o < kScanThreads; o <<= 1) {  // inclusive scan (Hillis-Steele)
            const int add = t >= o ? sh[t - o] : 0;
            __syncthreads();
            sh[t] += add;
            __syncthreads();
        }
        if (p < bv.n_problems) bv.active_first[p] = total + sh[t] - cnt;
        total += sh[kScanThreads - 1];
        __syncthreads();
    }
    if (t == 0) {
        g->n_active = total;
        g->changed = 0;
        g->compact_now = 1;
    }
}

// Writes the new list: slice s of running problem p at active_first[p] + (s - first slice of p).
__global__ __launch_bounds__(256) void batch_compact_fill_kernel(BatchView bv) {
    if (!bv.global->compact_now) return;
    for (int s = blockIdx.x * blockDim.x + threadIdx.x; s < bv.n_slices; s += gridDim.x * blockDim.x) {
        const int p = bv.slice_problem[s];
        if (!bv.problem_done[p]) bv.active[bv.active_first[p] + s - bv.pslice_ptr[p]] = s;
    }
}

// ------------------------------------------------------------ FP64 residual

// res = b - K x with FP64 values, per system ||res||^2 into true_rr.
template <int NC, int kW>
__global__ __launch_bounds__(512) void batch_residual_kernel(BatchView bv, SliceView v, const double* __restrict__ planes,
                                                            double* __restrict__ res) {
    const SliceRange range = block_slices(bv.n_slices);
    for (int s = range.first; s < range.stop; s += range.step) {
        const int p = bv.slice_problem[s];
        const int node = s * kWave + lane_id();
        // K x through the same multiply as the solver, with x as the multiplied vector
        BatchView tmp = bv;
        tmp.p = bv.x;
        tmp.q = res;
        double unused[NC];
        dpair kx[NC];
        slice_multiply<NC, double, kW>(tmp, v, planes, s, unused, kx);
        store_slice<NC>(res, s, kx);
        double vals[NC];
#pragma unroll
        for (int c = 0; c < NC; ++c) {
            const long long at = (static_cast<long long>(node) * NC + c) * 2;
            const double r0 = bv.b[at] - res[at];
            const double r1 = bv.b[at + 1] - res[at + 1];
            res[at] = r0;
            res[at + 1] = r1;
            vals[c] = wave_sum(r0 * r0 + r1 * r1);
        }
        double tot[NC];
        if (!finish_problem<NC>(bv, p, s, vals, tot)) continue;
        if (lane_id() == 0)
            for (int c = 0; c < NC; ++c) bv.sys[p * NC + c].true_rr = tot[c];
    }
}

// Jacobi inverse of the constrained matrix; 0 on the rows of padding nodes (in no
// element: the row is its diagonal alone, and it is zero).
__global__ void batch_jacobi_kernel(int n, const int* __restrict__ rp, const int* __restrict__ col,
                                    const double* __restrict__ values, double* __restrict__ minv, int* __restrict__ bad) {
    const int row = blockIdx.x * blockDim.x + threadIdx.x;
    if (row >= n) return;
    const int begin = rp[row], end = rp[row + 1];
    if (begin == end || (end - begin == 1 && col[begin] == row && values[begin] == 0.0)) {
        minv[row] = 0.0;
        return;
    }
    double d = 0.0;
    bool found = false;
    for (int k = begin; k < end; ++k)
        if (col[k] == row) {
            d = values[k];
            found = true;
        }
    if (!found || !(d != 0.0) || !isfinite(d)) {  // solver.cpp:656-669
        atomicAdd(bad, 1);
        minv[row] = 0.0;
        return;
    }
    minv[row] = 1.0 / d;
}

// dst[node][case][2] = src (one case, dof order).
__global__ void batch_scatter_case_kernel(int n, int n_cases, int c, const double* __restrict__ src,
                                          double* __restrict__ dst) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    dst[(static_cast<long long>(i >> 1) * n_cases + c) * 2 + (i & 1)] = src[i];
}

__global__ void batch_gather_case_kernel(int n, int n_cases, int c, const double* __restrict__ src,
                                         double* __restrict__ dst) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    dst[i] = src[(static_cast<long long>(i >> 1) * n_cases + c) * 2 + (i & 1)];
}

inline unsigned int blocks_for(long long work, int per_block) {
    return static_cast<unsigned int>((work + per_block - 1) / per_block);
}

}  // namespace

// ------------------------------------------------------------ launchers

unsigned int batch_grid(const BatchShape& shape, int n_slices) {
    const int waves = shape.block_threads / kWave;
    if (shape.span > 0 && shape.span < n_slices) n_slices = shape.span;
    const long long need = (static_cast<long long>(n_slices) + waves * shape.slices_per_wave - 1) /
                           (static_cast<long long>(waves) * shape.slices_per_wave);
    const unsigned int g = static_cast<unsigned int>(need < 1 ? 1 : need);
    return (g + 7u) & ~7u;
}

#define LFG_BATCH_NC(nc, body)                                 \
    switch (nc) {                                              \
        case 1: { constexpr int NC = 1; body; } break;         \
        case 2: { constexpr int NC = 2; body; } break;         \
        default: { constexpr int NC = 4; body; } break;        \
    }

void launch_batch_init(hipStream_t s, const BatchView& bv, const BatchShape& shape) {
    const unsigned int grid = batch_grid(shape, bv.n_slices);
    LFG_BATCH_NC(bv.n_cases, (batch_init_kernel<NC><<<grid, shape.block_threads, 0, s>>>(bv)));
}

template <int NC, int kW, int kSym>
static void spmv_nc(hipStream_t s, const BatchView& bv, const SliceView& v, const SymSliceView& sv,
                    const double* planes64, const float* planes32, unsigned int grid, int threads) {
    if (planes32) batch_spmv_kernel<NC, kW, true, kSym><<<grid, threads, 0, s>>>(bv, v, sv, planes64, planes32);
    else batch_spmv_kernel<NC, kW, false, kSym><<<grid, threads, 0, s>>>(bv, v, sv, planes64, nullptr);
    const unsigned int fin = blocks_for(static_cast<long long>(bv.n_problems) * kWave, 256);
    batch_spmv_finish_kernel<NC><<<fin, 256, 0, s>>>(bv);
}

void launch_batch_spmv(hipStream_t s, const BatchView& bv, const BatchShape& shape, const SliceView& v,
                       const double* planes64, const float* planes32, const SymSliceView* sym) {
    const unsigned int grid = batch_grid(shape, bv.n_slices);
    const int t = shape.block_threads;
    const SymSliceView sv = sym ? *sym : SymSliceView{};
    const int ksym = !sym ? 0 : (sym->max_lwidth <= 4 && sym->max_uwidth <= 5 ? 1 : 2);
    // kW is the FP32 node slices' unrolled width (mixed), whatever the FP64 layout
    if (ksym == 1 && v.max_width <= 9) {
        LFG_BATCH_NC(bv.n_cases, (spmv_nc<NC, 9, 1>(s, bv, v, sv, planes64, planes32, grid, t)));
    } else if (ksym == 1) {
        LFG_BATCH_NC(bv.n_cases, (spmv_nc<NC, 0, 1>(s, bv, v, sv, planes64, planes32, grid, t)));
    } else if (ksym == 2) {
        LFG_BATCH_NC(bv.n_cases, (spmv_nc<NC, 0, 2>(s, bv, v, sv, planes64, planes32, grid, t)));
    } else if (v.max_width <= 9) {
        LFG_BATCH_NC(bv.n_cases, (spmv_nc<NC, 9, 0>(s, bv, v, sv, planes64, planes32, grid, t)));
    } else {
        LFG_BATCH_NC(bv.n_cases, (spmv_nc<NC, 0, 0>(s, bv, v, sv, planes64, planes32, grid, t)));
    }
}

void launch_batch_update(hipStream_t s, const BatchView& bv, const BatchShape& shape) {
    const unsigned int grid = batch_grid(shape, bv.n_slices);
    const unsigned int fin = blocks_for(static_cast<long long>(bv.n_problems) * kWave, 256);
    LFG_BATCH_NC(bv.n_cases, (batch_update_kernel<NC><<<grid, shape.block_threads, 0, s>>>(bv),
                              batch_update_finish_kernel<NC><<<fin, 256, 0, s>>>(bv)));
}

void launch_batch_direction(hipStream_t s, const BatchView& bv, const BatchShape& shape) {
    const unsigned int grid = batch_grid(shape, bv.n_slices);
    LFG_BATCH_NC(bv.n_cases, (batch_direction_kernel<NC><<<grid, shape.block_threads, 0, s>>>(bv)));
}

void launch_batch_compact(hipStream_t s, const BatchView& bv) {
    batch_compact_scan_kernel<<<1, kScanThreads, 0, s>>>(bv);
    const unsigned int need = blocks_for(bv.n_slices, 256);
    batch_compact_fill_kernel<<<need < 1024u ? (need > 0 ? need : 1u) : 1024u, 256, 0, s>>>(bv);
}

void launch_batch_residual(hipStream_t s, const BatchView& bv, const BatchShape& shape, const SliceView& v,
                           const double* planes64, double* res) {
    const unsigned int grid = batch_grid(shape, bv.n_slices);
    if (v.max_width <= 9) {
        LFG_BATCH_NC(bv.n_cases,
                     (batch_residual_kernel<NC, 9><<<grid, shape.block_threads, 0, s>>>(bv, v, planes64, res)));
    } else {
        LFG_BATCH_NC(bv.n_cases,
                     (batch_residual_kernel<NC, 0><<<grid, shape.block_threads, 0, s>>>(bv, v, planes64, res)));
    }
}

void launch_batch_jacobi(hipStream_t s, const CsrView& a, const double* values, double* minv, int* bad) {
    if (a.n <= 0) return;
    batch_jacobi_kernel<<<blocks_for(a.n, 256), 256, 0, s>>>(a.n, a.row_ptr, a.col, values, minv, bad);
}

void launch_batch_scatter_case(hipStream_t s, int n, int n_cases, int c, const double* src, double* dst) {
    if (n <= 0) return;
    batch_scatter_case_kernel<<<blocks_for(n, 256), 256, 0, s>>>(n, n_cases, c, src, dst);
}

void launch_batch_gather_case(hipStream_t s, int n, int n_cases, int c, const double* src, double* dst) {
    if (n <= 0) return;
    batch_gather_case_kernel<<<blocks_for(n, 256), 256, 0, s>>>(n, n_cases, c, src, dst);
}

}  // namespace lfg
