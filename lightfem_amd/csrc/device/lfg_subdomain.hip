// Subdomain block-Jacobi (SubdomainBlockJacobi, precond.cpp:121-281) on gfx950.
//
// Each subdomain's dofs (matrix order) have a packed lower Cholesky factor with the
// reciprocal pivots on its diagonal, formed on the host in the reference's order.
// One wavefront solves one subdomain: lane l owns the rows l, l + 64, ... of the
// subdomain (kR rows, registers), the factor is staged in LDS when it fits (by all
// four wavefronts of the block, eight loads in flight per thread: the copy is bound
// by memory latency, not by the sweeps), and the two triangular sweeps run column
// by column. At step k the lane owning row k has
// the complete sum of that row; v_readlane broadcasts it, every lane forms
// w_k = sum_k * (1 / l_kk) itself, and the rows still open subtract l_ik * w_k.
// Each row therefore subtracts its terms in the reference's order (ascending k
// forward, descending k backward), one product at a time, and the result is
// bit-identical to the serial solve.
#include "lfg_kernels.hpp"

#include "lightfem_gpu.h"

namespace lfg {
namespace {

constexpr int kSubThreads = 256;  // four wavefronts stage the factor, the first one solves
constexpr int kFillUnroll = 8;

__device__ __forceinline__ double readlane_f64(double v, int lane) {
    const long long bits = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane(static_cast<int>(bits), lane);
    const int hi = __builtin_amdgcn_readlane(static_cast<int>(bits >> 32), lane);
    return __longlong_as_double((static_cast<long long>(hi) << 32) | static_cast<unsigned int>(lo));
}

// Forward substitution with L, then back substitution with L^T (precond.cpp:258-276)
// on the rows of this lane; F is the packed factor (row i starts at i(i+1)/2). The
// step loops run slot by slot (k = 64 m + kk) so that every register index is a
// compile-time constant. Each lane keeps the reciprocal pivots of its own rows in
// registers and the row owner forms w_k = sum_k * (1 / l_kk) (the same product the
// serial solve forms), so a step's dependency chain is one multiply, one broadcast,
// one multiply and one subtraction; the row entries of the next step are loaded
// while a step computes. Rows outside a step's update keep their value (select, not
// branch); loads past the factor are clamped to its last entry.
template <int kR>
__device__ __forceinline__ void triangular_sweeps(const double* __restrict__ F, int size, double (&acc)[kR]) {
    const int lane = static_cast<int>(threadIdx.x) & (kWave - 1);
    const int last = size * (size + 1) / 2 - 1;
    auto at = [&](int idx) { return F[min(idx, last)]; };
    int rows[kR];
    double piv[kR];
#pragma unroll
    for (int t = 0; t < kR; ++t) {
        const int i = lane + kWave * t;
        rows[t] = i * (i + 1) / 2;
        piv[t] = at(rows[t] + i);
    }
    // Steps go in pairs with two register sets for the row entries (a: even steps,
    // b: odd steps), so a loaded set is consumed one step later without copies.
#pragma unroll
    for (int m = 0; m < kR; ++m) {
        const int count = min(size - kWave * m, kWave);
        if (count <= 0) continue;
        auto load = [&](double (&f)[kR], int k) {
#pragma unroll
            for (int t = m; t < kR; ++t) f[t] = at(rows[t] + k);
        };
        auto step = [&](const double (&f)[kR], int kk) {
            const int k = kWave * m + kk;
            const double wk = readlane_f64(acc[m] * piv[m], kk);
#pragma unroll
            for (int t = m; t < kR; ++t) {
                const int i = lane + kWave * t;
                const double upd = acc[t] - f[t] * wk;
                acc[t] = i == k ? wk : ((i > k && i < size) ? upd : acc[t]);
            }
        };
        double fa[kR], fb[kR];
        load(fa, kWave * m);
        int kk = 0;
        for (; kk + 1 < count; kk += 2) {
            load(fb, kWave * m + kk + 1);
            __builtin_amdgcn_sched_barrier(0);  // issue the loads before the step that hides them
            step(fa, kk);
            load(fa, kWave * m + kk + 2);
            __builtin_amdgcn_sched_barrier(0);
            step(fb, kk + 1);
        }
        if (kk < count) step(fa, kk);
    }
#pragma unroll
    for (int m = kR - 1; m >= 0; --m) {
        const int count = min(size - kWave * m, kWave);
        if (count <= 0) continue;
        auto load = [&](double (&f)[kR], int k) {
            const int row_k = k > 0 ? k * (k + 1) / 2 : 0;
#pragma unroll
            for (int t = 0; t <= m; ++t) f[t] = at(row_k + lane + kWave * t);
        };
        auto step = [&](const double (&f)[kR], int kk) {
            const int k = kWave * m + kk;
            const double wk = readlane_f64(acc[m] * piv[m], kk);
#pragma unroll
            for (int t = 0; t <= m; ++t) {
                const int i = lane + kWave * t;
                const double upd = acc[t] - f[t] * wk;
                acc[t] = i == k ? wk : (i < k ? upd : acc[t]);
            }
        };
        double fa[kR], fb[kR];
        int kk = count - 1;
        load(fa, kWave * m + kk);
        for (; kk >= 1; kk -= 2) {
            load(fb, kWave * m + kk - 1);
            __builtin_amdgcn_sched_barrier(0);
            step(fa, kk);
            load(fa, kWave * m + kk - 2);
            __builtin_amdgcn_sched_barrier(0);
            step(fb, kk - 1);
        }
        if (kk == 0) step(fa, 0);
    }
}

// out = M^-1 in over the subdomains of this block (blockIdx.x, + gridDim.x, ...); the
// lane's share of in.out (r.z) is returned. kLds: the factor is copied to LDS first
// (subdomains of at most 64 kR dofs).
template <int kR, bool kLds>
__device__ __forceinline__ double subdomain_solves(const SubdomainView& v, const double* __restrict__ in,
                                                   double* __restrict__ out) {
    constexpr int kCap = kLds ? (kWave * kR) * (kWave * kR + 1) / 2 : 1;
    __shared__ double fl[kCap];
    const int lane = static_cast<int>(threadIdx.x) & (kWave - 1);
    const bool solver = threadIdx.x < kWave;
    double dot = 0.0;
    for (int s = blockIdx.x; s < v.n_sub; s += gridDim.x) {
        const int begin = v.dof_begin[s];
        const int size = v.dof_begin[s + 1] - begin;
        // an empty subdomain (an id no node carries) has no factor: its pivot loads
        // would read the entry before it; size is the same in every thread of the block
        if (size == 0) continue;
        const double* __restrict__ fg = v.factor + v.factor_begin[s];
        const double* F = fg;
        if (kLds) {
            const int fsz = size * (size + 1) / 2;
            __syncthreads();  // the previous subdomain's sweeps are done with fl
            for (int base = 0; base < fsz; base += kSubThreads * kFillUnroll) {
                double vbuf[kFillUnroll];
#pragma unroll
                for (int u = 0; u < kFillUnroll; ++u) {
                    const int t = base + u * kSubThreads + static_cast<int>(threadIdx.x);
                    vbuf[u] = t < fsz ? __builtin_nontemporal_load(fg + t) : 0.0;
                }
#pragma unroll
                for (int u = 0; u < kFillUnroll; ++u) {
                    const int t = base + u * kSubThreads + static_cast<int>(threadIdx.x);
                    if (t < fsz) fl[t] = vbuf[u];
                }
            }
            __syncthreads();
            F = fl;
        }
        if (!solver) continue;
        double acc[kR];
        int dof[kR];
#pragma unroll
        for (int t = 0; t < kR; ++t) {
            const int i = lane + kWave * t;
            dof[t] = i < size ? v.dofs[begin + i] : 0;
            acc[t] = i < size ? in[dof[t]] : 0.0;
        }
        triangular_sweeps<kR>(F, size, acc);
#pragma unroll
        for (int t = 0; t < kR; ++t) {
            if (lane + kWave * t < size) {
                out[dof[t]] = acc[t];
                dot += in[dof[t]] * acc[t];
            }
        }
    }
    return dot;
}

template <int kR, bool kLds>
__global__ __launch_bounds__(kSubThreads) void subdomain_apply_kernel(SubdomainView v, const double* __restrict__ in,
                                                                      double* __restrict__ out) {
    (void)subdomain_solves<kR, kLds>(v, in, out);
}

// z = M^-1 r, rho = r.z and beta (solver.cpp:575-583 on the first call, :611-620
// after), as the IC(0) step does for its sweeps.
template <int kR, bool kLds>
__global__ __launch_bounds__(kSubThreads) void pcg_subdomain_step_kernel(PcgBuffers b, SubdomainView v) {
    if (b.st->done) return;
    __shared__ double sh[kSubThreads / kWave];
    const double rz = subdomain_solves<kR, kLds>(v, b.r, b.z);  // 0 outside the first wavefront
    const double vals[1] = {block_sum<kSubThreads>(rz, sh)};
    double tot[1];
    if (!finish_reduction<kSubThreads, 1>(b.tickets + 5, b.partials, vals, tot, sh)) return;
    if (threadIdx.x == 0) {
        PcgState* st = b.st;
        const double trz = tot[0];
        if (!(trz > 0.0) || !isfinite(trz)) {  // :579-581, :616-618
            st->status = LFG_ERR_NONFINITE;
            st->done = 1;
        } else if (st->iter == 0) {
            st->rho = trz;
            st->beta = 0.0;
        } else {
            st->beta = trz / st->rho;
            st->rho = trz;
        }
    }
}

unsigned int subdomain_grid(const SubdomainView& v) {
    const unsigned int g = static_cast<unsigned int>(v.n_sub < 8192 ? (v.n_sub > 0 ? v.n_sub : 1) : 8192);
    return (g + 7u) & ~7u;
}

// Rows per lane kR and LDS staging by the largest subdomain: up to 192 dofs the factor
// fits in LDS (144.75 KB at 192), beyond that it is read from global memory.
#define LFG_SUBDOMAIN_DISPATCH(KERNEL, ...)                                                   \
    switch ((v.max_size + kWave - 1) / kWave) {                                             \
        case 0:                                                                             \
        case 1: KERNEL<1, true><<<grid, kSubThreads, 0, s>>>(__VA_ARGS__); break;           \
        case 2: KERNEL<2, true><<<grid, kSubThreads, 0, s>>>(__VA_ARGS__); break;           \
        case 3: KERNEL<3, true><<<grid, kSubThreads, 0, s>>>(__VA_ARGS__); break;           \
        case 4: KERNEL<4, false><<<grid, kSubThreads, 0, s>>>(__VA_ARGS__); break;          \
        case 5:                                                                             \
        case 6: KERNEL<6, false><<<grid, kSubThreads, 0, s>>>(__VA_ARGS__); break;          \
        default: KERNEL<8, false><<<grid, kSubThreads, 0, s>>>(__VA_ARGS__); break;         \
    }

}  // namespace

void launch_subdomain_apply(hipStream_t s, const SubdomainView& v, const double* in, double* out) {
    const unsigned int grid = subdomain_grid(v);
    LFG_SUBDOMAIN_DISPATCH(subdomain_apply_kernel, v, in, out)
}

void launch_pcg_subdomain_step(hipStream_t s, const PcgBuffers& b, const SubdomainView& v) {
    const unsigned int grid = subdomain_grid(v);
    LFG_SUBDOMAIN_DISPATCH(pcg_subdomain_step_kernel, b, v)
}

}  // namespace lfg
