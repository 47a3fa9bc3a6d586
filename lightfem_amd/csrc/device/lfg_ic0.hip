// IC(0) preconditioner applications on gfx950 (precond.cpp:386-412): z = L^-T L^-1 r
// with the factor computed on the host.
//
// Each triangular sweep is a dependency chain as long as the factor's level count
// (1,600 per sweep at 200x200 with RCM) with ~50 independent rows per level, so it
// runs in one workgroup and its time is the chain's latency. The host sorts the rows
// of each sweep by level and rewrites every dependency as a level-order position. The
// sweep then runs level after level, one wavefront per level in rotation: the
// wavefront that owns level l loaded its rows' data (factor values, dependency
// positions, inputs) while the 15 levels before it ran, waits for the LDS counter of
// finished levels to reach l, reads its dependencies from an LDS window of the last
// 4096 positions and advances the counter. Every wait has a bound: past it the sweep
// stops and sets the fault flag. Factors whose dependencies reach back further than
// the window holds take plain level-scheduled sweeps with one barrier per level.
//
// Each row sums its terms in the reference's order (forward: strict lower row in
// ascending columns; backward: the transposed factor's column in descending rows,
// equal to the reference's column scatter), so z is bit-identical to precond.cpp.
#include "lfg_kernels.hpp"
#include "lfg_precond.hpp"

#include "lightfem_gpu.h"

namespace lfg {
namespace {

constexpr int kIcThreads = 256;  // 4 wavefronts: 1024 measured 17% slower (more waiting waves contend)
constexpr int kIcWaves = kIcThreads / kWave;
constexpr int kIcWin = 4096;      // LDS window (positions)
constexpr int kIcDeps = 12;       // dependencies per row held in registers
constexpr long long kIcSpinLimit = 1LL << 22;

constexpr int kIcNodeWin = kIcWin / 2;  // LDS window of the node form (node positions)
constexpr int kIcNodeDeps = 8;          // dependency nodes per node held in registers

struct Ic0Lds {
    union {
        double z[kIcWin];
        dpair zn[kIcNodeWin];
    };
    int done_pos;  // every position below it is finished (levels finish in order)
    int fault;
};

// A lane's row of the level it processes next, loaded a round of levels ahead.
struct Ic0Row {
    int row, cnt;
    double in, scale;
    int key[kIcDeps];
    double val[kIcDeps];
};

__device__ __forceinline__ void load_ic0_row(const Ic0Sweep& sw, const double* in, int p, Ic0Row& r) {
    r.row = sw.ord[p];
    const int kb = sw.ptr[p];
    r.cnt = sw.ptr[p + 1] - kb;
    r.scale = sw.scale[p];
#pragma unroll
    for (int u = 0; u < kIcDeps; ++u) {
        const bool has = u < r.cnt;
        r.key[u] = has ? sw.key[kb + u] : 0;
        r.val[u] = has ? sw.val[kb + u] : 0.0;
    }
    r.in = in[r.row];
}

// The reference's row: sum = in - val_0 z_0 - val_1 z_1 - ... (dependency order), times
// the reciprocal pivot.
__device__ __forceinline__ double ic0_row_value(const Ic0Sweep& sw, const Ic0Lds& s, const Ic0Row& r, int p) {
    double sum = r.in;
#pragma unroll
    for (int u = 0; u < kIcDeps; ++u)
        if (u < r.cnt) sum -= r.val[u] * s.z[r.key[u] & (kIcWin - 1)];
    if (r.cnt > kIcDeps) {  // rows with more dependencies than registers
        const int kb = sw.ptr[p];
        for (int u = kIcDeps; u < r.cnt; ++u) sum -= sw.val[kb + u] * s.z[sw.key[kb + u] & (kIcWin - 1)];
    }
    return sum * r.scale;
}

// One sweep: wavefront w takes levels w, w + 16, w + 32, ... (the host splits levels
// into parts of at most 64 rows, one row per lane); it waits until every earlier level
// is done (done_pos has reached its first position), computes its rows from the data
// it prefetched, writes each value into the window slot of its position, advances
// done_pos to its level's end, then stores the values to `out` and prefetches its
// next level. A level reads earlier levels only, within `reach` positions, and a slot
// is overwritten only by a position more than reach + 64 later, so every value read
// is still in the window. LDS serves each wavefront's accesses in issue order: values
// stored before the counter are visible to any wavefront that has seen it
// (compiler-only fences keep that order in the code; a hardware fence would also wait
// for the sweep's global loads and stores). Adding the already finished terms before
// the wait was tried and measured slower (6.0 against 3.8 ms per application at
// 200x200): the waiting wavefronts then compete with the one on the critical path.
__device__ void ic0_sweep_pipelined(const Ic0Sweep& sw, Ic0Lds& s, const double* in, double* out) {
    const int lane = threadIdx.x & (kWave - 1);
    const int w = threadIdx.x / kWave;
    if (threadIdx.x == 0) s.done_pos = 0;
    __syncthreads();
    Ic0Row next;
    int next_begin = 0, next_end = 0;
    if (w < sw.levels) {
        next_begin = sw.level_ptr[w];
        next_end = sw.level_ptr[w + 1];
        if (next_begin + lane < next_end) load_ic0_row(sw, in, next_begin + lane, next);
    }
    bool faulted = false;
    for (int l = w; l < sw.levels; l += kIcWaves) {
        const Ic0Row cur = next;
        const int begin = next_begin, end = next_end;
        const int p = begin + lane;
        const bool mine = p < end;
        long long spins = 0;
        while (__hip_atomic_load(&s.done_pos, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP) < begin) {
            if (++spins > kIcSpinLimit) {
                faulted = true;
                break;
            }
            __builtin_amdgcn_s_sleep(1);
        }
        if (faulted) break;
        __atomic_signal_fence(__ATOMIC_SEQ_CST);
        double z = 0.0;
        if (mine) {
            z = ic0_row_value(sw, s, cur, p);
            s.z[p & (kIcWin - 1)] = z;
        }
        __atomic_signal_fence(__ATOMIC_SEQ_CST);
        if (lane == 0) __hip_atomic_store(&s.done_pos, end, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        if (mine) out[cur.row] = z;
        const int ln = l + kIcWaves;  // this wavefront's next level, loaded while the next 15 run
        if (ln < sw.levels) {
            next_begin = sw.level_ptr[ln];
            next_end = sw.level_ptr[ln + 1];
            if (next_begin + lane < next_end) load_ic0_row(sw, in, next_begin + lane, next);
        }
    }
    if (faulted) s.fault = 1;
    __syncthreads();
}

// Node form (Ic0NodeSweep): a lane computes both rows of its node, the second after the
// first, from dependency nodes read as value pairs from the window.
struct Ic0NodeRow {
    int node, cnt;
    double in_first, in_second, cross, scale_first, scale_second;
    int key[kIcNodeDeps];
    double vf[2 * kIcNodeDeps], vs[2 * kIcNodeDeps];
};

template <bool kBackward>
__device__ __forceinline__ void load_ic0_node(const Ic0NodeSweep& sw, const double* in, int p, Ic0NodeRow& r) {
    r.node = sw.ord[p];
    const int kb = sw.ptr[p];
    r.cnt = sw.ptr[p + 1] - kb;
    r.cross = sw.cross[p];
    r.scale_first = sw.scale_first[p];
    r.scale_second = sw.scale_second[p];
#pragma unroll
    for (int u = 0; u < kIcNodeDeps; ++u) {
        const bool has = u < r.cnt;
        r.key[u] = has ? sw.key[kb + u] : 0;
        r.vf[2 * u] = has ? sw.v_first[2 * (kb + u)] : 0.0;
        r.vf[2 * u + 1] = has ? sw.v_first[2 * (kb + u) + 1] : 0.0;
        r.vs[2 * u] = has ? sw.v_second[2 * (kb + u)] : 0.0;
        r.vs[2 * u + 1] = has ? sw.v_second[2 * (kb + u) + 1] : 0.0;
    }
    const dpair x = reinterpret_cast<const dpair*>(in)[r.node];
    r.in_first = kBackward ? x.y : x.x;
    r.in_second = kBackward ? x.x : x.y;
}

// Forward: first = 2i, second = 2i+1, a dependency pair in the order (2j, 2j+1);
// backward: first = 2i+1, second = 2i, the order (2j+1, 2j). The second row's cross
// term (its coupling to the first) comes last, as in the reference's rows / columns.
template <bool kBackward>
__device__ __forceinline__ dpair ic0_node_value(const Ic0NodeSweep& sw, const Ic0Lds& s, const Ic0NodeRow& r,
                                                int p) {
    double a = r.in_first, b = r.in_second;
    auto term = [&](int key, double f0, double f1, double s0, double s1) {
        const dpair zp = s.zn[key & (kIcNodeWin - 1)];
        const double z0 = kBackward ? zp.y : zp.x, z1 = kBackward ? zp.x : zp.y;
        a -= f0 * z0;
        a -= f1 * z1;
        b -= s0 * z0;
        b -= s1 * z1;
    };
#pragma unroll
    for (int u = 0; u < kIcNodeDeps; ++u)
        if (u < r.cnt) term(r.key[u], r.vf[2 * u], r.vf[2 * u + 1], r.vs[2 * u], r.vs[2 * u + 1]);
    if (r.cnt > kIcNodeDeps) {  // nodes with more dependency nodes than registers
        const int kb = sw.ptr[p];
        for (int u = kIcNodeDeps; u < r.cnt; ++u) {
            const int k = kb + u;
            term(sw.key[k], sw.v_first[2 * k], sw.v_first[2 * k + 1], sw.v_second[2 * k], sw.v_second[2 * k + 1]);
        }
    }
    const double zf = a * r.scale_first;
    b -= r.cross * zf;
    const double zs = b * r.scale_second;
    return kBackward ? dpair{zs, zf} : dpair{zf, zs};
}

// The node form of ic0_sweep_pipelined (same schedule, window and counter).
template <bool kBackward>
__device__ void ic0_node_sweep(const Ic0NodeSweep& sw, Ic0Lds& s, const double* in, double* out) {
    const int lane = threadIdx.x & (kWave - 1);
    const int w = threadIdx.x / kWave;
    if (threadIdx.x == 0) s.done_pos = 0;
    __syncthreads();
    Ic0NodeRow next;
    int next_begin = 0, next_end = 0;
    if (w < sw.levels) {
        next_begin = sw.level_ptr[w];
        next_end = sw.level_ptr[w + 1];
        if (next_begin + lane < next_end) load_ic0_node<kBackward>(sw, in, next_begin + lane, next);
    }
    bool faulted = false;
    for (int l = w; l < sw.levels; l += kIcWaves) {
        const Ic0NodeRow cur = next;
        const int begin = next_begin, end = next_end;
        const int p = begin + lane;
        const bool mine = p < end;
        long long spins = 0;
        while (__hip_atomic_load(&s.done_pos, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP) < begin) {
            if (++spins > kIcSpinLimit) {
                faulted = true;
                break;
            }
            __builtin_amdgcn_s_sleep(1);
        }
        if (faulted) break;
        __atomic_signal_fence(__ATOMIC_SEQ_CST);
        dpair z = {0.0, 0.0};
        if (mine) {
            z = ic0_node_value<kBackward>(sw, s, cur, p);
            s.zn[p & (kIcNodeWin - 1)] = z;
        }
        __atomic_signal_fence(__ATOMIC_SEQ_CST);
        if (lane == 0) __hip_atomic_store(&s.done_pos, end, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        if (mine) reinterpret_cast<dpair*>(out)[cur.node] = z;
        const int ln = l + kIcWaves;
        if (ln < sw.levels) {
            next_begin = sw.level_ptr[ln];
            next_end = sw.level_ptr[ln + 1];
            if (next_begin + lane < next_end) load_ic0_node<kBackward>(sw, in, next_begin + lane, next);
        }
    }
    if (faulted) s.fault = 1;
    __syncthreads();
}

__device__ __forceinline__ bool ic0_nodes_fit(const Ic0View& f) {
    return f.nfwd.ord != nullptr && f.nbwd.ord != nullptr && f.nfwd.widest <= kWave && f.nbwd.widest <= kWave &&
           f.nfwd.reach + 2 * kWave <= kIcNodeWin && f.nbwd.reach + 2 * kWave <= kIcNodeWin;
}

// The window must hold every position a level may read: the reach back from the
// level's last position, with a margin.
__device__ __forceinline__ bool ic0_pipelined_fits(const Ic0View& f) {
    return f.fwd.ord != nullptr && f.fwd.widest <= kWave && f.bwd.widest <= kWave &&
           f.fwd.reach + 2 * kWave <= kIcWin && f.bwd.reach + 2 * kWave <= kIcWin;
}

// Level-scheduled sweeps for factors whose dependencies reach further back than the
// LDS window holds: rows of a level gather over their strict lower row (forward) or
// the transposed column in descending rows (backward), one barrier per level.
__device__ void ic0_sweeps_levels(const Ic0View& f, const double* __restrict__ in, double* __restrict__ z) {
    for (int i = threadIdx.x; i < f.n; i += kIcThreads) z[i] = in[i];
    __syncthreads();
    for (int level = 0; level < f.fwd_levels; ++level) {
        const int begin = f.fwd_level_ptr[level], end = f.fwd_level_ptr[level + 1];
        for (int idx = begin + threadIdx.x; idx < end; idx += kIcThreads) {
            const int row = f.fwd_rows[idx];
            double sum = z[row];
            const int last = f.lrow[row + 1] - 1;
            for (int k = f.lrow[row]; k < last; ++k) sum -= f.lval[k] * z[f.lcol[k]];
            z[row] = sum * f.inv_diag[row];
        }
        __syncthreads();
    }
    for (int level = 0; level < f.bwd_levels; ++level) {
        const int begin = f.bwd_level_ptr[level], end = f.bwd_level_ptr[level + 1];
        for (int idx = begin + threadIdx.x; idx < end; idx += kIcThreads) {
            const int c = f.bwd_rows[idx];
            double sum = z[c];
            for (int k = f.urow[c]; k < f.urow[c + 1]; ++k) sum -= f.uval[k] * z[f.ucol[k]];
            z[c] = sum * f.inv_diag[c];
        }
        __syncthreads();
    }
}

// z = (L L^T)^-1 in; returns false (in every thread) when a dependency wait ran out.
__device__ bool ic0_sweeps(const Ic0View& f, Ic0Lds& s, const double* __restrict__ in, double* __restrict__ z) {
    const bool nodes = ic0_nodes_fit(f);
    if (!nodes && !ic0_pipelined_fits(f)) {
        ic0_sweeps_levels(f, in, z);
        return true;
    }
    if (threadIdx.x == 0) s.fault = 0;
    __syncthreads();
    if (nodes) {
        ic0_node_sweep<false>(f.nfwd, s, in, z);
        if (s.fault == 0) ic0_node_sweep<true>(f.nbwd, s, z, z);
    } else {
        ic0_sweep_pipelined(f.fwd, s, in, z);
        if (s.fault == 0) ic0_sweep_pipelined(f.bwd, s, z, z);
    }
    const bool ok = s.fault == 0;
    __syncthreads();
    return ok;
}

__global__ __launch_bounds__(kIcThreads) void ic0_apply_kernel(Ic0View f, const double* __restrict__ in,
                                                               double* __restrict__ out) {
    __shared__ Ic0Lds s;
    if (!ic0_sweeps(f, s, in, out) && threadIdx.x == 0) *f.fault = 1;
}

// z = IC0(r), r.z, then rho/beta (first call: rho = r.z, beta = 0).
__global__ __launch_bounds__(kIcThreads) void pcg_ic0_step_kernel(PcgBuffers b, Ic0View f) {
    if (b.st->done) return;
    __shared__ Ic0Lds s;
    __shared__ double sh[kIcThreads / kWave];
    const bool ok = ic0_sweeps(f, s, b.r, b.z);
    double rz = 0.0;
    for (int i = threadIdx.x; i < f.n; i += kIcThreads) rz += b.r[i] * b.z[i];
    const double trz = block_sum<kIcThreads>(rz, sh);
    if (threadIdx.x == 0) {
        PcgState* st = b.st;
        if (!ok) {
            *f.fault = 1;
            st->status = LFG_ERR_HIP;
            st->done = 1;
        } else if (st->iter == 0) {
            if (!(trz > 0.0) || !isfinite(trz)) {  // solver.cpp:579-581
                st->status = LFG_ERR_NONFINITE;
                st->done = 1;
            } else {
                st->rho = trz;
                st->beta = 0.0;
            }
        } else if (!(trz > 0.0) || !isfinite(trz)) {  // :616-618
            st->status = LFG_ERR_NONFINITE;
            st->done = 1;
        } else {
            st->beta = trz / st->rho;
            st->rho = trz;
        }
    }
}

}  // namespace

void launch_ic0_apply(hipStream_t s, const Ic0View& f, const double* in, double* out) {
    ic0_apply_kernel<<<1, kIcThreads, 0, s>>>(f, in, out);
}

void launch_pcg_ic0_step(hipStream_t s, const PcgBuffers& b, const Ic0View& f) {
    pcg_ic0_step_kernel<<<1, kIcThreads, 0, s>>>(b, f);
}

}  // namespace lfg
