// IC(0) preconditioner applications on gfx950 (precond.cpp:386-412): z = L^-T L^-1 r
// with the factor computed on the host.
//
// Each triangular sweep is a dependency chain as long as the factor's level count
// (1,600 per sweep at 200x200 with RCM) with ~50 independent rows per level, so it
// runs in one workgroup and its time is the chain's latency. The host sorts the rows
// of each sweep by level and rewrites every dependency as a level-order position. The
// sweep then runs level after level, one wavefront per level in rotation: the
// wavefront that owns level l loaded its rows' data (factor values, dependency
// positions, inputs) during the rotations before, waits for the LDS counter of
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

// The window; its last element is a zero slot that padded dependencies read.
struct Ic0Lds {
    union {
        double z[kIcWin + 2];
        dpair zn[kIcNodeWin + 1];
    };
    int done_pos;  // every position below it is finished (levels finish in order)
    int fault;
};

// Loads of a level run a pipeline of three rotations (a rotation: the four wavefronts
// each running one level). After running its level, a wavefront issues the body
// (dependency positions, factor values, input) of its next level, whose head it loaded
// one rotation earlier, the head (row or node index, factor offset and count, pivots)
// of the level after that, whose bounds it loaded one rotation earlier, and the bounds
// of the level after that, so none of the three dependent loads waits for another on
// the wavefront's way to its next level. Timestamps of a 200x200 RCM sweep (1,088 node
// levels) put the per-level time at ~1,100 cycles: ~250 from the counter's store to the
// next wavefront seeing it, ~580 for the window reads, ~160 for the arithmetic, ~120 to
// store the values and advance the counter; the loads have arrived when the wait ends.

// The positions [begin, end) of level l.
struct Ic0Bounds {
    int begin, end;
};

__device__ __forceinline__ Ic0Bounds ic0_bounds(const int* level_ptr, int levels, int l) {
    if (l >= levels) return Ic0Bounds{0, 0};
    return Ic0Bounds{level_ptr[l], level_ptr[l + 1]};
}

// kend - kb dependencies at kb: the count is formed where it is used, so issuing the
// head's loads never waits for one of them.
struct Ic0RowHead {
    int row, kb, kend;
    double scale;
};

struct Ic0RowBody {
    double in;
    int key[kIcDeps];
    double val[kIcDeps];
};

__device__ __forceinline__ void load_ic0_row_head(const Ic0Sweep& sw, int p, Ic0RowHead& h) {
    h.row = sw.ord[p];
    h.kb = sw.ptr[p];
    h.kend = sw.ptr[p + 1];
    h.scale = sw.scale[p];
}

__device__ __forceinline__ void load_ic0_row_body(const Ic0Sweep& sw, const double* in, const Ic0RowHead& h,
                                                  Ic0RowBody& r) {
    const int cnt = h.kend - h.kb;
#pragma unroll
    for (int u = 0; u < kIcDeps; ++u) {
        const bool has = u < cnt;
        r.key[u] = has ? sw.key[h.kb + u] : 0;
        r.val[u] = has ? sw.val[h.kb + u] : 0.0;
    }
    r.in = in[h.row];
}

// The reference's row: sum = in - val_0 z_0 - val_1 z_1 - ... (dependency order), times
// the reciprocal pivot. Every dependency's value is read from the window before the
// first is used, so a row costs one LDS round trip instead of one per dependency, and
// the terms run straight through all kIcDeps slots: a padded dependency has a zero
// value and reads the zero slot, and sum - (+0 * +0) is sum in every case (-0
// included), so each row's result is the reference's.
__device__ __forceinline__ double ic0_row_value(const Ic0Sweep& sw, const Ic0Lds& s, const Ic0RowHead& h,
                                                const Ic0RowBody& r) {
    const int cnt = h.kend - h.kb;
    double zd[kIcDeps];
#pragma unroll
    for (int u = 0; u < kIcDeps; ++u) zd[u] = s.z[u < cnt ? (r.key[u] & (kIcWin - 1)) : kIcWin];
    double sum = r.in;
#pragma unroll
    for (int u = 0; u < kIcDeps; ++u) sum -= r.val[u] * zd[u];
    if (cnt > kIcDeps) {  // rows with more dependencies than registers
        for (int u = kIcDeps; u < cnt; ++u) sum -= sw.val[h.kb + u] * s.z[sw.key[h.kb + u] & (kIcWin - 1)];
    }
    return sum * h.scale;
}

// Waits until every position before `pos` is finished; false when the wait ran out.
__device__ __forceinline__ bool ic0_wait(const Ic0Lds& s, int pos) {
    long long spins = 0;
    while (__hip_atomic_load(&s.done_pos, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP) < pos) {
        if (++spins > kIcSpinLimit) return false;
        __builtin_amdgcn_s_sleep(1);
    }
    return true;
}

// One sweep: wavefront w takes levels w, w + 4, w + 8, ... (the host splits levels
// into parts of at most 64 rows, one row per lane); it waits until every earlier level
// is done (done_pos has reached its first position), computes its rows from the data
// it prefetched, writes each value into the window slot of its position, advances
// done_pos to its level's end, then stores the values to `out` and issues the loads
// of its next levels. A level reads earlier levels only, within `reach` positions, and
// a slot is overwritten only by a position more than reach + 64 later, so every value
// read is still in the window. LDS serves each wavefront's accesses in issue order:
// values stored before the counter are visible to any wavefront that has seen it
// (compiler-only fences keep that order in the code; a hardware fence would also wait
// for the sweep's global loads and stores). A row's input is read one rotation before
// the row runs; in the backward sweep (in == out) no other row writes it before then.
// Adding the already finished terms before the wait was tried and measured slower
// (6.0 against 3.8 ms per application at 200x200): the waiting wavefronts then compete
// with the one on the critical path.
__device__ void ic0_sweep_pipelined(const Ic0Sweep& sw, Ic0Lds& s, const double* in, double* out) {
    const int lane = threadIdx.x & (kWave - 1);
    const int w = threadIdx.x / kWave;
    if (threadIdx.x == 0) {
        s.done_pos = 0;
        s.z[kIcWin] = 0.0;
    }
    __syncthreads();
    Ic0Bounds bc = ic0_bounds(sw.level_ptr, sw.levels, w);
    Ic0RowHead hc{0, 0, 0, 0.0}, hn{0, 0, 0, 0.0};
    Ic0RowBody cur;
    if (bc.begin + lane < bc.end) {
        load_ic0_row_head(sw, bc.begin + lane, hc);
        load_ic0_row_body(sw, in, hc, cur);
    }
    Ic0Bounds bn = ic0_bounds(sw.level_ptr, sw.levels, w + kIcWaves);
    if (bn.begin + lane < bn.end) load_ic0_row_head(sw, bn.begin + lane, hn);
    Ic0Bounds bnn = ic0_bounds(sw.level_ptr, sw.levels, w + 2 * kIcWaves);
    bool faulted = false;
    for (int l = w; l < sw.levels; l += kIcWaves) {
        const int p = bc.begin + lane;
        const bool mine = p < bc.end;
        if (!ic0_wait(s, bc.begin)) {
            faulted = true;
            break;
        }
        __atomic_signal_fence(__ATOMIC_SEQ_CST);
        double z = 0.0;
        if (mine) {
            z = ic0_row_value(sw, s, hc, cur);
            s.z[p & (kIcWin - 1)] = z;
        }
        __atomic_signal_fence(__ATOMIC_SEQ_CST);
        if (lane == 0) __hip_atomic_store(&s.done_pos, bc.end, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        if (mine) out[hc.row] = z;
        const bool next_mine = bn.begin + lane < bn.end;
        if (next_mine) load_ic0_row_body(sw, in, hn, cur);
        hc = hn;
        bc = bn;
        if (bnn.begin + lane < bnn.end) load_ic0_row_head(sw, bnn.begin + lane, hn);
        bn = bnn;
        bnn = ic0_bounds(sw.level_ptr, sw.levels, l + 3 * kIcWaves);
    }
    if (faulted) s.fault = 1;
    __syncthreads();
}

// Node form (Ic0NodeSweep): a lane computes both rows of its node, the second after the
// first, from dependency nodes read as value pairs from the window.
struct Ic0NodeHead {
    int node, kb, kend;
    double cross, scale_first, scale_second;
};

struct Ic0NodeBody {
    double in_first, in_second;
    int key[kIcNodeDeps];
    double vf[2 * kIcNodeDeps], vs[2 * kIcNodeDeps];
};

__device__ __forceinline__ void load_ic0_node_head(const Ic0NodeSweep& sw, int p, Ic0NodeHead& h) {
    h.node = sw.ord[p];
    h.kb = sw.ptr[p];
    h.kend = sw.ptr[p + 1];
    h.cross = sw.cross[p];
    h.scale_first = sw.scale_first[p];
    h.scale_second = sw.scale_second[p];
}

template <bool kBackward>
__device__ __forceinline__ void load_ic0_node_body(const Ic0NodeSweep& sw, const double* in, const Ic0NodeHead& h,
                                                   Ic0NodeBody& r) {
    const int cnt = h.kend - h.kb;
#pragma unroll
    for (int u = 0; u < kIcNodeDeps; ++u) {
        const bool has = u < cnt;
        const int k = h.kb + u;
        r.key[u] = has ? sw.key[k] : 0;
        r.vf[2 * u] = has ? sw.v_first[2 * k] : 0.0;
        r.vf[2 * u + 1] = has ? sw.v_first[2 * k + 1] : 0.0;
        r.vs[2 * u] = has ? sw.v_second[2 * k] : 0.0;
        r.vs[2 * u + 1] = has ? sw.v_second[2 * k + 1] : 0.0;
    }
    const dpair x = reinterpret_cast<const dpair*>(in)[h.node];
    r.in_first = kBackward ? x.y : x.x;
    r.in_second = kBackward ? x.x : x.y;
}

// Forward: first = 2i, second = 2i+1, a dependency pair in the order (2j, 2j+1);
// backward: first = 2i+1, second = 2i, the order (2j+1, 2j). The second row's cross
// term (its coupling to the first) comes last, as in the reference's rows / columns.
// Every dependency pair is read before the first is used (one LDS round trip per node)
// and the terms run straight through all kIcNodeDeps slots, as in ic0_row_value: a
// padded dependency has zero values and reads the zero slot's pair.
template <bool kBackward>
__device__ __forceinline__ dpair ic0_node_value(const Ic0NodeSweep& sw, const Ic0Lds& s, const Ic0NodeHead& h,
                                                const Ic0NodeBody& r) {
    const int cnt = h.kend - h.kb;
    dpair zd[kIcNodeDeps];
#pragma unroll
    for (int u = 0; u < kIcNodeDeps; ++u) zd[u] = s.zn[u < cnt ? (r.key[u] & (kIcNodeWin - 1)) : kIcNodeWin];
    double a = r.in_first, b = r.in_second;
#pragma unroll
    for (int u = 0; u < kIcNodeDeps; ++u) {
        const double z0 = kBackward ? zd[u].y : zd[u].x, z1 = kBackward ? zd[u].x : zd[u].y;
        a -= r.vf[2 * u] * z0;
        a -= r.vf[2 * u + 1] * z1;
        b -= r.vs[2 * u] * z0;
        b -= r.vs[2 * u + 1] * z1;
    }
    if (cnt > kIcNodeDeps) {  // nodes with more dependency nodes than registers
        for (int u = kIcNodeDeps; u < cnt; ++u) {
            const int k = h.kb + u;
            const dpair zp = s.zn[sw.key[k] & (kIcNodeWin - 1)];
            const double z0 = kBackward ? zp.y : zp.x, z1 = kBackward ? zp.x : zp.y;
            a -= sw.v_first[2 * k] * z0;
            a -= sw.v_first[2 * k + 1] * z1;
            b -= sw.v_second[2 * k] * z0;
            b -= sw.v_second[2 * k + 1] * z1;
        }
    }
    const double zf = a * h.scale_first;
    b -= h.cross * zf;
    const double zs = b * h.scale_second;
    return kBackward ? dpair{zs, zf} : dpair{zf, zs};
}

// The node form of ic0_sweep_pipelined (same schedule, window, counter and loads).
template <bool kBackward>
__device__ void ic0_node_sweep(const Ic0NodeSweep& sw, Ic0Lds& s, const double* in, double* out) {
    const int lane = threadIdx.x & (kWave - 1);
    const int w = threadIdx.x / kWave;
    if (threadIdx.x == 0) {
        s.done_pos = 0;
        s.zn[kIcNodeWin] = dpair{0.0, 0.0};
    }
    __syncthreads();
    Ic0Bounds bc = ic0_bounds(sw.level_ptr, sw.levels, w);
    Ic0NodeHead hc{0, 0, 0, 0.0, 0.0, 0.0}, hn{0, 0, 0, 0.0, 0.0, 0.0};
    Ic0NodeBody cur;
    if (bc.begin + lane < bc.end) {
        load_ic0_node_head(sw, bc.begin + lane, hc);
        load_ic0_node_body<kBackward>(sw, in, hc, cur);
    }
    Ic0Bounds bn = ic0_bounds(sw.level_ptr, sw.levels, w + kIcWaves);
    if (bn.begin + lane < bn.end) load_ic0_node_head(sw, bn.begin + lane, hn);
    Ic0Bounds bnn = ic0_bounds(sw.level_ptr, sw.levels, w + 2 * kIcWaves);
    bool faulted = false;
    for (int l = w; l < sw.levels; l += kIcWaves) {
        const int p = bc.begin + lane;
        const bool mine = p < bc.end;
        if (!ic0_wait(s, bc.begin)) {
            faulted = true;
            break;
        }
        __atomic_signal_fence(__ATOMIC_SEQ_CST);
        dpair z = {0.0, 0.0};
        if (mine) {
            z = ic0_node_value<kBackward>(sw, s, hc, cur);
            s.zn[p & (kIcNodeWin - 1)] = z;
        }
        __atomic_signal_fence(__ATOMIC_SEQ_CST);
        if (lane == 0) __hip_atomic_store(&s.done_pos, bc.end, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        if (mine) reinterpret_cast<dpair*>(out)[hc.node] = z;
        const bool next_mine = bn.begin + lane < bn.end;
        if (next_mine) load_ic0_node_body<kBackward>(sw, in, hn, cur);
        hc = hn;
        bc = bn;
        if (bnn.begin + lane < bnn.end) load_ic0_node_head(sw, bnn.begin + lane, hn);
        bn = bnn;
        bnn = ic0_bounds(sw.level_ptr, sw.levels, l + 3 * kIcWaves);
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
