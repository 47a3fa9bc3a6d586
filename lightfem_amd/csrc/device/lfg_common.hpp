// Shared device helpers and the device-resident solver state of the
// LightFEM gfx950 backend. Wave size is 64 (CDNA4); blocks are multiples of it.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace lfg {

constexpr int kWave = 64;

// Device-resident Krylov state. One instance per solve lives in device memory;
// the last block of each reducing kernel (ticket pattern) updates it, and the
// host reads it back once per captured graph launch.
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
    long long iter;
    long long max_iter;
    long long replacements;
    int done;          // 1 once converged, failed or out of iterations
    int status;        // lfg_status
    int converged;
    int need_replace;  // mixed: the replacement kernels run this iteration
    int phase32;       // mixed: SpMVs read the FP32 values
    unsigned int ticket[4];
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

// Ticket pattern: thread 0 of every block publishes its partial, then takes a
// ticket; the block that takes the last ticket sees every partial (release /
// acquire at agent scope) and finishes the reduction. Returns true in every
// thread of that last block.
__device__ __forceinline__ bool ticket_last(unsigned int* ticket, unsigned int nblocks) {
    __shared__ int last;
    if (threadIdx.x == 0) {
        const unsigned int t = __hip_atomic_fetch_add(ticket, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
        last = (t == nblocks - 1) ? 1 : 0;
    }
    __syncthreads();
    const bool is_last = last != 0;
    if (is_last) __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    return is_last;
}

// Deterministic final reduction of `count` partials (fixed assignment + fixed tree).
template <int NT>
__device__ __forceinline__ double reduce_partials(const double* partials, int count, int stride, double* sh) {
    double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
    int i = threadIdx.x;
    for (; i + 3 * NT < count; i += 4 * NT) {  // four independent loads in flight
        s0 += partials[i * stride];
        s1 += partials[(i + NT) * stride];
        s2 += partials[(i + 2 * NT) * stride];
        s3 += partials[(i + 3 * NT) * stride];
    }
    for (; i < count; i += NT) s0 += partials[i * stride];
    return block_sum<NT>((s0 + s1) + (s2 + s3), sh);
}

// Blocks are dispatched round-robin over the 8 XCDs (block b runs on XCD b % 8).
// Remapping gives each XCD one contiguous range of logical blocks, so the rows a
// XCD works on (and the x entries they gather) stay in that XCD's L2.
__device__ __forceinline__ int xcd_block(int b, int nblocks) {
    const int per = nblocks >> 3;  // nblocks is a multiple of 8
    return (b & 7) * per + (b >> 3);
}

__device__ __forceinline__ double max0(double v) { return (0.0 < v) ? v : 0.0; }

}  // namespace lfg
