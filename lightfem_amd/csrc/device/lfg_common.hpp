// Shared device helpers and the device-resident solver state of the
// LightFEM gfx950 backend. Wave size is 64 (CDNA4); blocks are multiples of it.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace lfg {

constexpr int kWave = 64;

// Device-resident Krylov state. One instance per solve lives in device memory;
// the final block of each reducing kernel updates it, and the host reads it
// back once per captured graph launch.
struct PcgState {
    double rho;        // r.z of the current iterate
    double rr;         // ||r||^2 of the current (recurrence or replaced) residual
    double pq;         // p.Ap
    double alpha;
    double beta;
    double target;     // stopping threshold on ||r||
    double rnorm0;     // ||r0||
    double maxr;       // mixed: residual norm at the last replacement
    double drop;       // mixed: replace when ||r|| < drop * maxr
    double rz_new;     // r.z computed by the update kernel
    double switch_rr;  // mixed: FP32 values once ||r||^2 <= switch_rr
    double true_rr;    // mixed: ||r||^2 at the last replacement
    double alpha_prev; // fused stencil iteration: the alpha before `alpha` (deferred x update)
    double rr_new;     // split node-slice iteration: ||r'||^2 of the vector launch (its r'.z in rz_new)
    long long iter;
    long long max_iter;
    long long replacements;
    int done;          // 1 once converged, failed or out of iterations
    int status;        // lfg_status
    int converged;
    int need_replace;  // mixed: the replacement kernel runs this iteration
    int phase32;       // mixed: SpMVs read the FP32 values
    int replaced;      // mixed: this iteration's residual is the FP64 one in z (accepted replacement)
    int pending;       // variant 14: the last launch's partial sums are still to be folded
};

// Completion counters of one reducing kernel: one per XCD and one on top, each on
// its own 128-byte line, so the blocks of an XCD contend only with each other.
struct TicketSet {
    unsigned int xcd[8][32];
    unsigned int top[32];
};

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int o = kWave / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, kWave);
    return v;
}

// Sum over a block of NT threads; the result is valid in thread 0. `sh` holds
// NT/64 doubles of LDS and is reusable after return.
template <int NT>
__device__ __forceinline__ double block_sum(double v, double* sh) {
    v = wave_sum(v);
    const int w = threadIdx.x / kWave;
    if ((threadIdx.x & (kWave - 1)) == 0) sh[w] = v;
    __syncthreads();
    double t = 0.0;
    if (threadIdx.x == 0) {
#pragma unroll
        for (int i = 0; i < NT / kWave; ++i) t += sh[i];
    }
    __syncthreads();
    return t;
}

// Partials cross blocks (and XCDs) through relaxed agent-scope atomic stores and
// loads, which the hardware keeps coherent across the XCDs' L2s by themselves.
// What orders a partial before the ticket that announces it is an explicit wait
// for the thread's outstanding vector-memory operations (stores_complete): a store
// and an atomic to different addresses go to different L2 channels and may
// complete in either order, and a workgroup-scope release fence emits no wait on
// gfx950 outside threadgroup-split mode. Agent-scope acquire/release would also
// order them, but writes back and invalidates the whole L2 of the XCD in every
// block. Everything else a kernel writes reaches the next kernel through the
// kernel boundary.
__device__ __forceinline__ void publish_partial(double* slot, double v) {
    __hip_atomic_store(slot, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Returns once every global store this thread issued has been acknowledged.
__device__ __forceinline__ void stores_complete() { asm volatile("s_waitcnt vmcnt(0)" ::: "memory"); }

__device__ __forceinline__ double read_partial(const double* slot) {
    return __hip_atomic_load(const_cast<double*>(slot), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Two-level in-kernel reduction of K values per block (grid a multiple of 8).
// Thread 0 holds the block's values in `vals`. Blocks are grouped as the
// hardware dispatches them (block b on XCD b % 8): the last block of each group
// folds its group's partials in a fixed order and publishes one group partial,
// and the last of the eight group finishers folds those. Returns true in every
// thread of that final block, with the totals in `out` of thread 0. The fold
// order is fixed, so results are bit-reproducible run to run. `partials` holds
// (gridDim.x + 8) * K doubles. A grid of at most kOneLevelGrid blocks takes one level
// (one counter, the last block folds every partial): three dependent memory round trips
// fewer, which is what a small system's iteration waits for, and contention on one
// counter that so few blocks cannot make costly.
constexpr int kOneLevelGrid = 256;
template <int NT, int K>
__device__ __forceinline__ bool finish_reduction(TicketSet* t, double* partials, const double* vals, double* out,
                                                 double* sh) {
    __shared__ int flag;
    const int nb = gridDim.x;
    if (nb <= kOneLevelGrid) {
        if (threadIdx.x == 0) {
#pragma unroll
            for (int k = 0; k < K; ++k) publish_partial(partials + blockIdx.x * K + k, vals[k]);
            stores_complete();
            const unsigned int a = __hip_atomic_fetch_add(&t->top[0], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            flag = (a == static_cast<unsigned int>(nb - 1)) ? 1 : 0;
        }
        __syncthreads();
        if (!flag) return false;
        double s[K];
#pragma unroll
        for (int k = 0; k < K; ++k) {
            double acc = 0.0;
            for (int i = threadIdx.x; i < nb; i += NT) acc += read_partial(partials + i * K + k);
            s[k] = block_sum<NT>(acc, sh);
        }
        if (threadIdx.x == 0) {
#pragma unroll
            for (int k = 0; k < K; ++k) out[k] = s[k];
            __hip_atomic_store(&t->top[0], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        return true;
    }
    const int x = blockIdx.x & 7;
    const int per = nb >> 3;
    double* group = partials + static_cast<long long>(nb) * K;
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < K; ++k) publish_partial(partials + blockIdx.x * K + k, vals[k]);
        stores_complete();
        const unsigned int a = __hip_atomic_fetch_add(&t->xcd[x][0], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        flag = (a == static_cast<unsigned int>(per - 1)) ? 1 : 0;
    }
    __syncthreads();
    if (!flag) return false;
    double s[K];
#pragma unroll
    for (int k = 0; k < K; ++k) {
        double acc = 0.0;
        for (int i = threadIdx.x; i < per; i += NT) acc += read_partial(partials + (x + 8 * i) * K + k);
        s[k] = block_sum<NT>(acc, sh);
    }
    if (threadIdx.x == 0) {
        __hip_atomic_store(&t->xcd[x][0], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
#pragma unroll
        for (int k = 0; k < K; ++k) publish_partial(group + x * K + k, s[k]);
        stores_complete();
        const unsigned int a = __hip_atomic_fetch_add(&t->top[0], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        flag = (a == 7u) ? 1 : 0;
    }
    __syncthreads();
    if (!flag) return false;
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < K; ++k) {
            double total = 0.0;
            for (int g = 0; g < 8; ++g) total += read_partial(group + g * K + k);
            out[k] = total;
        }
        __hip_atomic_store(&t->top[0], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    return true;
}

// Blocks are dispatched round-robin over the 8 XCDs (block b runs on XCD b % 8).
// Remapping gives each XCD one contiguous range of logical blocks, so the rows a
// XCD works on (and the x entries they gather) stay in that XCD's L2.
__device__ __forceinline__ int xcd_block(int b, int nblocks) {
    const int per = nblocks >> 3;  // nblocks is a multiple of 8
    return (b & 7) * per + (b >> 3);
}

__device__ __forceinline__ double max0(double v) { return (0.0 < v) ? v : 0.0; }

// Value pairs of one 2x2 block row (16-byte FP64 or 8-byte FP32 loads).
typedef double dpair __attribute__((ext_vector_type(2)));

// A 2x2 block (rows u, l) times the pair x, added to the row sums a0, a1: one product
// at a time, as the CSR row loop (sparse.cpp:113-117), or with pair = true the block's
// two products of a row summed first, as the block storage product (sparse.cpp:188-193).
__device__ __forceinline__ void add_block(double& a0, double& a1, double ux, double uy, double lx, double ly, dpair x,
                                          bool pair) {
    if (pair) {
        a0 += ux * x.x + uy * x.y;
        a1 += lx * x.x + ly * x.y;
    } else {
        a0 += ux * x.x;
        a0 += uy * x.y;
        a1 += lx * x.x;
        a1 += ly * x.y;
    }
}
typedef float fpair __attribute__((ext_vector_type(2)));

template <typename VT>
struct PairOf;
template <>
struct PairOf<double> {
    using type = dpair;
};
template <>
struct PairOf<float> {
    using type = fpair;
};

}  // namespace lfg
