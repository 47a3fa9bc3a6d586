// Solver kernels for gfx950: symmetric Dirichlet elimination, reactions, the
// fused preconditioned-CG iteration (FP64 and FP32-value mixed precision) and
// the level-scheduled IC(0) sweeps.
//
// Layout: CSR with int32 row offsets and columns, FP64 (or FP32) values; node i
// owns rows 2i and 2i+1, which share one column set. The SpMV therefore works
// on node groups: 2W lanes per node, W lanes per row, each group loading its
// row segments with consecutive lanes (coalesced) and reducing with wave
// shuffles. Every reduction is finished in-kernel by the last block to arrive
// (ticket pattern), so one CG iteration is three launches (SpMV + p.Ap, update
// + r.z, direction) that a hipGraph replays without host round trips; each
// kernel returns immediately once the device state says the solve is done.
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

// ------------------------------------------------------------ zero pivots

// max |f| into *out as the bit pattern of a non-negative double (ordered like its value).
__global__ __launch_bounds__(256) void abs_max_kernel(int n, const double* __restrict__ f,
                                                      unsigned long long* __restrict__ out) {
    double m = 0.0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) m = fmax(m, fabs(f[i]));
#pragma unroll
    for (int o = kWave / 2; o > 0; o >>= 1) m = fmax(m, __shfl_xor(m, o, kWave));
    if ((threadIdx.x & (kWave - 1)) == 0 && m > 0.0)
        atomicMax(out, static_cast<unsigned long long>(__double_as_longlong(m)));
}

// solver.cpp:916-940, one row per thread: a diagonal of magnitude at most 1e-30 becomes
// 1 with its right-hand side 0, unless that right-hand side exceeds 1e-12 of the
// largest (1 when all are zero): then the row is reported (the smallest such row in
// counters[1]). counters[0] counts the pinned rows.
__global__ __launch_bounds__(256) void pin_pivots_kernel(int n, const int* __restrict__ rp,
                                                         const int* __restrict__ col, double* __restrict__ vals,
                                                         double* __restrict__ f,
                                                         const unsigned long long* __restrict__ scale_bits,
                                                         int* __restrict__ counters) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int lo = rp[i], hi = rp[i + 1];
    while (lo < hi) {  // the diagonal's position (columns sorted; the pattern always holds it)
        const int mid = (lo + hi) >> 1;
        if (col[mid] < i) lo = mid + 1;
        else hi = mid;
    }
    const bool present = lo < rp[i + 1] && col[lo] == i;
    const double diag = present ? vals[lo] : 0.0;
    if (fabs(diag) > 1.0e-30) return;
    double scale = __longlong_as_double(static_cast<long long>(*scale_bits));
    if (scale <= 0.0) scale = 1.0;
    if (fabs(f[i]) > 1.0e-12 * scale || !present) {
        atomicMin(counters + 1, i);
        return;
    }
    vals[lo] = 1.0;
    f[i] = 0.0;
    atomicAdd(counters, 1);
}

// ------------------------------------------------------------ Dirichlet

// solver.cpp:355-413. One node group of 16 lanes per row pair would scatter the
// RHS sum; the sum must run in column order to stay bit-exact, so lane 0 of each
// 8-lane row group does it while all 8 lanes copy/zero the row coalesced. The 8
// lanes flag the row's constrained nonzero entries as they copy it; a row with none
// (all but the rows next to the supports) keeps f unchanged, so lane 0 runs the
// serial sum only where it subtracts something.
__global__ __launch_bounds__(256) void dirichlet_kernel(int n, const int* __restrict__ rp, const int* __restrict__ col,
                                                        const double* __restrict__ values, const double* __restrict__ f,
                                                        const int* __restrict__ constrained,
                                                        const double* __restrict__ prescribed, double* __restrict__ kc,
                                                        double* __restrict__ fc) {
    const int row = (blockIdx.x * blockDim.x + threadIdx.x) >> 3;
    const int lane = threadIdx.x & 7;
    if (row >= n) return;
    const int begin = rp[row], end = rp[row + 1];
    if (constrained[row]) {
        for (int k = begin + lane; k < end; k += 8) kc[k] = (col[k] == row) ? 1.0 : 0.0;
        if (lane == 0) fc[row] = prescribed[row];
        return;
    }
    bool hit = false;
    for (int k = begin + lane; k < end; k += 8) {
        const double v = values[k];
        const bool eliminated = constrained[col[k]] != 0 && v != 0.0;
        kc[k] = eliminated ? 0.0 : v;
        hit = hit || eliminated;
    }
    // the 8-lane group of this row within the wavefront's ballot
    const unsigned long long group = (__ballot(hit) >> (threadIdx.x & (kWave - 8))) & 0xffull;
    if (lane == 0 && group == 0) fc[row] = f[row];
    if (lane == 0 && group != 0) {
        double rhs = f[row];
        for (int k = begin; k < end; ++k) {
            const int c = col[k];
            const double v = values[k];
            if (constrained[c] == 0 || v == 0.0) continue;
            rhs -= v * prescribed[c];
        }
        fc[row] = rhs;
    }
}

// y = A x - sub with each row summed serially in column order (sparse.cpp:104-120
// followed by solver.cpp:953-954), so reactions are bit-identical to the CPU.
__global__ __launch_bounds__(256) void spmv_serial_rows_kernel(int n, const int* __restrict__ rp,
                                                               const int* __restrict__ col,
                                                               const double* __restrict__ values,
                                                               const double* __restrict__ x,
                                                               const double* __restrict__ sub, double* __restrict__ y) {
    const int row = blockIdx.x * blockDim.x + threadIdx.x;
    if (row >= n) return;
    double sum = 0.0;
    for (int k = rp[row]; k < rp[row + 1]; ++k) sum += values[k] * x[col[k]];
    y[row] = sub ? sum - sub[row] : sum;
}

__device__ __forceinline__ int find_in_row(const int* col, int begin, int end, int c) {
    int lo = begin, hi = end;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (col[mid] < c) lo = mid + 1; else hi = mid;
    }
    return (lo < end && col[lo] == c) ? lo : -1;
}

__global__ void extract_diag_kernel(int n, const int* __restrict__ rp, const int* __restrict__ col,
                                    const double* __restrict__ values, double* __restrict__ diag, int* status) {
    const int row = blockIdx.x * blockDim.x + threadIdx.x;
    if (row >= n) return;
    const int k = find_in_row(col, rp[row], rp[row + 1], row);
    if (k < 0) {
        atomicOr(status, 1);
        diag[row] = 0.0;
        return;
    }
    diag[row] = values[k];
}

// inverse_diagonal_of (solver.cpp:656-669).
__global__ void jacobi_inv_kernel(int n, const double* __restrict__ diag, double* __restrict__ inv, int* status) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double d = diag[i];
    if (!isfinite(d)) atomicOr(status, 2);
    inv[i] = jacobi_inverse(d);
}

// SubdomainBlockJacobi factor for one-node subdomains (precond.cpp:178-232):
// packed [1/l00, l10, 1/l11] per node, identity when the block is not SPD.
__global__ void block2_factor_kernel(int n_nodes, const int* __restrict__ rp, const int* __restrict__ col,
                                     const double* __restrict__ values, double* __restrict__ factor, int* degenerate) {
    const int node = blockIdx.x * blockDim.x + threadIdx.x;
    if (node >= n_nodes) return;
    double dense[4] = {0.0, 0.0, 0.0, 0.0};
    for (int i = 0; i < 2; ++i) {
        const int dof = 2 * node + i;
        for (int k = rp[dof]; k < rp[dof + 1]; ++k) {
            const int c = col[k];
            if ((c >> 1) != node) continue;
            dense[i * 2 + (c - 2 * node)] = values[k];
        }
    }
    bool factorized = true;
    for (int i = 0; i < 2 && factorized; ++i)
        for (int j = 0; j <= i; ++j) {
            double sum = dense[i * 2 + j];
            for (int k = 0; k < j; ++k) sum -= dense[i * 2 + k] * dense[j * 2 + k];
            if (i == j) {
                if (!(sum > 0.0) || !isfinite(sum)) {
                    factorized = false;
                    break;
                }
                dense[i * 2 + i] = sqrt(sum);
            } else {
                dense[i * 2 + j] = sum / dense[j * 2 + j];
            }
        }
    if (!factorized) {
        atomicAdd(degenerate, 1);
        dense[0] = 1.0;
        dense[2] = 0.0;
        dense[3] = 1.0;
    }
    double* t = factor + 3LL * node;
    t[0] = 1.0 / dense[0];
    t[1] = dense[2];
    t[2] = 1.0 / dense[3];
}

__global__ void precond_apply_kernel(int kind, int n_nodes, const double* __restrict__ m, const double* __restrict__ in,
                                     double* __restrict__ out) {
    const int node = blockIdx.x * blockDim.x + threadIdx.x;
    if (node >= n_nodes) return;
    const double2 r = reinterpret_cast<const double2*>(in)[node];
    double z0, z1;
    if (kind == LFG_PRECOND_JACOBI) precond_node<LFG_PRECOND_JACOBI>(m, node, r.x, r.y, z0, z1);
    else precond_node<LFG_PRECOND_BLOCK2>(m, node, r.x, r.y, z0, z1);
    reinterpret_cast<double2*>(out)[node] = make_double2(z0, z1);
}

__global__ void to_f32_kernel(const double* __restrict__ in, float* __restrict__ out, long long n) {
    const long long i = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) out[i] = static_cast<float>(in[i]);
}

// ------------------------------------------------------------ PCG kernels

// Row sum of a W-lane row group: lanes stride the row segment (coalesced loads of
// values and columns), gather x, accumulate in FP64, reduce with shuffles. The
// result is valid in every lane of the group.
template <int W, typename VT>
__device__ __forceinline__ double row_dot(const int* __restrict__ rp, const int* __restrict__ col,
                                          const VT* __restrict__ val, const double* __restrict__ x, int row,
                                          int lane) {
    const int begin = rp[row], end = rp[row + 1];
    double sum = 0.0;
    for (int k = begin + lane; k < end; k += W) sum += static_cast<double>(val[k]) * x[col[k]];
#pragma unroll
    for (int o = W / 2; o > 0; o >>= 1) sum += __shfl_xor(sum, o, W);
    return sum;
}

// r0 = b, x0 = 0, z0 = M^-1 r0, p0 = z0; ||r0||^2 and r0.z0; threshold.
// switch_tau: 0 keeps FP64 values throughout, > 0 switches to the FP32 values once
// ||r|| <= switch_tau ||r0||, < 0 uses the FP32 values from the first iteration.
template <int Prec>
__global__ __launch_bounds__(kThreads) void pcg_init_kernel(PcgBuffers b, int stop_rule, double tol,
                                                            long long max_iter, double drop, double switch_tau,
                                                            int fused) {
    __shared__ double sh[kThreads / kWave];
    const int n_nodes = b.n >> 1;
    double rr = 0.0, rz = 0.0;
    for (int node = blockIdx.x * kThreads + threadIdx.x; node < n_nodes; node += gridDim.x * kThreads) {
        const double2 r = reinterpret_cast<const double2*>(b.b)[node];
        reinterpret_cast<double2*>(b.r)[node] = r;
        reinterpret_cast<double2*>(b.x)[node] = make_double2(0.0, 0.0);
        rr += r.x * r.x + r.y * r.y;
        if (Prec == LFG_PRECOND_IC0) {
            reinterpret_cast<double2*>(b.p)[node] = make_double2(0.0, 0.0);
        } else {
            double z0, z1;
            precond_node<Prec>(b.m, node, r.x, r.y, z0, z1);
            reinterpret_cast<double2*>(b.p)[node] = fused ? make_double2(0.0, 0.0) : make_double2(z0, z1);
            if (fused) reinterpret_cast<double2*>(b.q)[node] = make_double2(0.0, 0.0);
            rz += r.x * z0 + r.y * z1;
        }
    }
    const double vals[2] = {block_sum<kThreads>(rr, sh), block_sum<kThreads>(rz, sh)};
    double tot[2];
    if (!finish_reduction<kThreads, 2>(b.tickets + 0, b.partials, vals, tot, sh)) return;
    if (threadIdx.x == 0) {
        const double trr = tot[0], trz = tot[1];
        PcgState* st = b.st;
        const double rnorm0 = sqrt(max0(trr));
        const double base = stop_rule == LFG_STOP_REFERENCE ? ((1.0 < rnorm0) ? rnorm0 : 1.0) : rnorm0;
        st->rnorm0 = rnorm0;
        st->target = tol * base;
        st->rr = trr;
        st->rho = trz;
        st->beta = 0.0;
        st->alpha = 0.0;
        st->alpha_prev = 0.0;
        st->iter = fused ? -1 : 0;  // the first fused launch only forms p0 and A p0
        st->max_iter = max_iter;
        st->replacements = 0;
        st->maxr = rnorm0;
        st->drop = drop;
        st->need_replace = 0;
        st->replaced = 0;
        st->pending = 0;
        st->phase32 = switch_tau < 0.0 ? 1 : 0;
        st->true_rr = 0.0;
        st->switch_rr = switch_tau > 0.0 ? (switch_tau * rnorm0) * (switch_tau * rnorm0) : -1.0;
        st->converged = 0;
        st->status = LFG_OK;
        st->done = 0;
        if (rnorm0 <= st->target) {  // solver.cpp:573-578
            st->converged = 1;
            st->done = 1;
        } else if (Prec != LFG_PRECOND_IC0 && (!(trz > 0.0) || !isfinite(trz))) {  // :579-581
            st->status = LFG_ERR_NONFINITE;
            st->done = 1;
        }
    }
}

// q = A p (node groups of 2W lanes), fused p.q, alpha = rho / p.q. Each block owns
// one contiguous node range (XCD-remapped), so partials stay few and the rows a
// XCD touches, with the p entries they gather, stay in that XCD's L2.
template <int W, typename VT>
__device__ __forceinline__ double spmv_range(const PcgBuffers& b, const int* __restrict__ rp,
                                             const int* __restrict__ col, const VT* __restrict__ val, int start,
                                             int stop) {
    constexpr int kNodesPerStep = kThreads / (2 * W);
    const int group = threadIdx.x / (2 * W);
    const int half = (threadIdx.x / W) & 1;
    const int lane = threadIdx.x & (W - 1);
    double pq = 0.0;
    for (int base = start; base < stop; base += kNodesPerStep) {
        const int node = base + group;
        if (node < stop) {
            const int row = 2 * node + half;
            const double s = row_dot<W, VT>(rp, col, val, b.p, row, lane);
            if (lane == 0) {
                b.q[row] = s;
                pq += b.p[row] * s;
            }
        }
    }
    return pq;
}

// q = A p (node groups of 2W lanes), fused p.q, alpha = rho / p.q. Each block owns
// one contiguous node range (XCD-remapped), so partials stay few and the rows a
// XCD touches, with the p entries they gather, stay in that XCD's L2. With FP32
// values present, the state's phase selects which value array is streamed.
template <int W>
__global__ __launch_bounds__(kThreads) void pcg_spmv_kernel(PcgBuffers b, const int* __restrict__ rp,
                                                            const int* __restrict__ col,
                                                            const double* __restrict__ val64,
                                                            const float* __restrict__ val32) {
    if (b.st->done) return;
    __shared__ double sh[kThreads / kWave];
    const bool use32 = val32 != nullptr && (val64 == nullptr || b.st->phase32 != 0);
    const int n_nodes = b.n >> 1;
    const int lb = xcd_block(blockIdx.x, gridDim.x);
    const int per_block = (n_nodes + gridDim.x - 1) / gridDim.x;
    const int start = lb * per_block;
    const int stop = min(start + per_block, n_nodes);
    const double pq = use32 ? spmv_range<W, float>(b, rp, col, val32, start, stop)
                            : spmv_range<W, double>(b, rp, col, val64, start, stop);
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

// Row-segment streaming SpMV (variant 4). A wave owns 64 consecutive rows, whose
// nonzeros form one contiguous segment: lane l loads entries seg + l, seg + 64 + l,
// ... (every load instruction moves 64 consecutive values / columns), all value
// and column loads are issued before the gathers of p, the products are staged in
// LDS, and lane l then sums row l of the wave. Rows longer than the staging room
// (kSlots per lane on average) fall back to 8 lanes per row.
// Rows per wave chunk: 64 in variant 4, 32 in variant 5.
constexpr int slots_for(int rows) { return (rows * 18 + kWave - 1) / kWave; }

// kMaxRow > 0: every row has at most kMaxRow entries (checked on the host), so
// the row sums are unrolled with clamped, unconditional LDS reads; 0: generic loop.
template <typename VT, int kMaxRow, int kRows>
__device__ __forceinline__ double stream_rows(const PcgBuffers& b, const int* __restrict__ rp,
                                              const int* __restrict__ col, const VT* __restrict__ val, int row0,
                                              int row_end, double* __restrict__ stage) {
    constexpr int kSlots = slots_for(kRows);
    const int lane = threadIdx.x & (kWave - 1);
    const int seg = rp[row0];
    const int len = rp[row_end] - seg;
    double pq = 0.0;
    if (len > kWave * kSlots) {  // wave-uniform fallback for long rows
        for (int r = row0 + (lane >> 3); r < row_end + 7; r += 8) {
            const int row = min(r, row_end - 1);
            const double s = row_dot<8, VT>(rp, col, val, b.p, row, lane & 7);
            if ((lane & 7) == 0 && r < row_end) {
                b.q[row] = s;
                pq += b.p[row] * s;
            }
        }
        return pq;
    }
    const int row = row0 + lane;
    const int my_row = min(row, row_end - 1);
    const int a = rp[my_row] - seg, e = rp[my_row + 1] - seg;
    // Loads are unconditional (indices clamped into the segment) so the compiler
    // issues all of them back to back; only the LDS stores are predicated.
    VT v[kSlots];
    int c[kSlots];
#pragma unroll
    for (int j = 0; j < kSlots; ++j) {
        const int k = min(j * kWave + lane, len - 1);
        v[j] = __builtin_nontemporal_load(val + seg + k);
        c[j] = __builtin_nontemporal_load(col + seg + k);
    }
    double x[kSlots];
#pragma unroll
    for (int j = 0; j < kSlots; ++j) x[j] = b.p[c[j]];
#pragma unroll
    for (int j = 0; j < kSlots; ++j) {
        const int k = j * kWave + lane;
        if (k < len) stage[k] = static_cast<double>(v[j]) * x[j];
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    double sum = 0.0;
    if (kMaxRow > 0) {
#pragma unroll
        for (int t = 0; t < kMaxRow; ++t) {
            const double term = stage[min(a + t, e - 1)];
            sum += (a + t < e) ? term : 0.0;
        }
    } else {
        for (int k = a; k < e; ++k) sum += stage[k];
    }
    if (lane < kRows && row < row_end) {
        b.q[row] = sum;
        pq = b.p[row] * sum;
    }
    __builtin_amdgcn_wave_barrier();  // the stage is reused by the wave's next rows
    return pq;
}

template <typename VT, int kMaxRow, int kRows>
__device__ __forceinline__ double stream_range(const PcgBuffers& b, const int* __restrict__ rp,
                                               const int* __restrict__ col, const VT* __restrict__ val,
                                               int row_start, int row_stop, double* stage) {
    const int wave = threadIdx.x / kWave;
    constexpr int kWaves = kThreads / kWave;
    double pq = 0.0;
    for (int r0 = row_start + wave * kRows; r0 < row_stop; r0 += kWaves * kRows)
        pq += stream_rows<VT, kMaxRow, kRows>(b, rp, col, val, r0, min(r0 + kRows, row_stop), stage);
    return pq;
}

template <int kMaxRow, int kRows>
__global__ __launch_bounds__(kThreads) void pcg_spmv_stream_kernel(PcgBuffers b, const int* __restrict__ rp,
                                                                   const int* __restrict__ col,
                                                                   const double* __restrict__ val64,
                                                                   const float* __restrict__ val32) {
    if (b.st->done) return;
    __shared__ double sh[kThreads / kWave];
    __shared__ double stage_all[kThreads / kWave][kWave * slots_for(kRows)];
    double* stage = stage_all[threadIdx.x / kWave];
    const bool use32 = val32 != nullptr && (val64 == nullptr || b.st->phase32 != 0);
    const int lb = xcd_block(blockIdx.x, gridDim.x);
    // contiguous row range per block, in whole waves' worth of rows (rows stay node-aligned)
    const int chunks = (b.n + kRows - 1) / kRows;
    const int per_block = (chunks + gridDim.x - 1) / gridDim.x;
    const int start = min(lb * per_block * kRows, b.n);
    const int stop = min(start + per_block * kRows, b.n);
    const double pq = use32 ? stream_range<float, kMaxRow, kRows>(b, rp, col, val32, start, stop, stage)
                            : stream_range<double, kMaxRow, kRows>(b, rp, col, val64, start, stop, stage);
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

// ------------------------------------------------------------ node-slice SpMV

// One thread per (slice, lane): copies node 64s + l's blocks from its two CSR rows
// (entries 2k, 2k+1 of each row form block k) into the planes; padding is zero.
__global__ __launch_bounds__(256) void pack_slices_kernel(int n_nodes, int n_slices, long long blocks,
                                                          const int* __restrict__ rp, const int* __restrict__ col,
                                                          const int* __restrict__ ptr,
                                                          const double* __restrict__ values, double* __restrict__ out64,
                                                          float* __restrict__ out32, int* __restrict__ col_out,
                                                          double* __restrict__ diag_out) {
    const long long t = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
    const int s = static_cast<int>(t >> 6);
    const int lane = static_cast<int>(t & (kWave - 1));
    if (s >= n_slices) return;
    const int node = s * kWave + lane;
    const int begin = ptr[s];
    const int width = (ptr[s + 1] - begin) / kWave;
    int r0 = 0, r1 = 0, deg = 0;
    bool lone = false;  // a node in no element: rows hold only their diagonals
    if (node < n_nodes) {
        r0 = rp[2 * node];
        r1 = rp[2 * node + 1];
        lone = r1 - r0 == 1;
        deg = lone ? 1 : (r1 - r0) >> 1;
    }
    const int self = node < n_nodes ? node : n_nodes - 1;
    double d0 = 0.0, d1 = 0.0, d2 = 0.0, d3 = 0.0;  // diagonal block, then its compensation
    for (int k = 0; k < width; ++k) {
        const long long at = begin + static_cast<long long>(k) * kWave + lane;
        double u0 = 0.0, u1 = 0.0, l0 = 0.0, l1 = 0.0;
        int c = self;
        if (lone && k == 0) {
            if (values) {
                u0 = values[r0];
                l1 = values[r1];
            }
        } else if (k < deg) {
            c = col[r0 + 2 * k] >> 1;
            if (values) {
                u0 = values[r0 + 2 * k];
                u1 = values[r0 + 2 * k + 1];
                l0 = values[r1 + 2 * k];
                l1 = values[r1 + 2 * k + 1];
            }
        }
        if (out64) {
            out64[2 * at] = u0;
            out64[2 * at + 1] = u1;
            out64[2 * (blocks + at)] = l0;
            out64[2 * (blocks + at) + 1] = l1;
        }
        if (out32) {
            const bool diag_block = diag_out != nullptr && c == self && k < deg && node < n_nodes;
            const float f0 = diag_block ? 0.0f : static_cast<float>(u0);
            const float f1 = diag_block ? 0.0f : static_cast<float>(u1);
            const float f2 = diag_block ? 0.0f : static_cast<float>(l0);
            const float f3 = diag_block ? 0.0f : static_cast<float>(l1);
            out32[2 * at] = f0;
            out32[2 * at + 1] = f1;
            out32[2 * (blocks + at)] = f2;
            out32[2 * (blocks + at) + 1] = f3;
            if (diag_block) {
                d0 += u0;
                d1 += u1;
                d2 += l0;
                d3 += l1;
            } else if (k < deg) {
                // the row keeps its FP64 sum: the diagonal absorbs each rounding error
                d0 -= static_cast<double>(f0) - u0;
                d1 -= static_cast<double>(f1) - u1;
                d2 -= static_cast<double>(f2) - l0;
                d3 -= static_cast<double>(f3) - l1;
            }
        }
        if (col_out) col_out[at] = c;
    }
    if (diag_out && node < n_nodes) {
        reinterpret_cast<dpair*>(diag_out)[2 * node] = dpair{d0, d1};
        reinterpret_cast<dpair*>(diag_out)[2 * node + 1] = dpair{d2, d3};
    }
}


// Rows 2i, 2i+1 of node i = 64s + lane: every load instruction of the wave moves 64
// consecutive blocks of a plane (or columns), p is gathered as one pair per block,
// and each row accumulates its products in column order from 0.0, exactly as the
// serial CSR row loop (sparse.cpp:104-120), or block by block as the block storage
// product (sparse.cpp:176-198) when v.pair_sum is set. kW > 0: slices up to kW wide take the
// unrolled path, whose loads are unconditional (indices clamped to the last
// block) and whose accumulation skips the clamped entries.
template <typename VT, int kW>
__device__ __forceinline__ double slice_rows(const PcgBuffers& b, const SliceView& v, const VT* __restrict__ planes,
                                             int s) {
    using P = typename PairOf<VT>::type;
    const int lane = threadIdx.x & (kWave - 1);
    const int node = s * kWave + lane;
    const int begin = v.ptr[s];
    const int width = (v.ptr[s + 1] - begin) / kWave;
    const P* __restrict__ upper = reinterpret_cast<const P*>(planes);
    const P* __restrict__ lower = upper + v.blocks;
    const dpair* __restrict__ p2 = reinterpret_cast<const dpair*>(b.p);
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
    double pq = 0.0;
    if (node < (b.n >> 1)) {
        const dpair pv = p2[node];
        if (sizeof(VT) == sizeof(float)) add_diag32(v, node, pv, a0, a1);
        const dpair qv = {a0, a1};
        reinterpret_cast<dpair*>(b.q)[node] = qv;
        pq = pv.x * a0 + pv.y * a1;
    }
    return pq;
}

// q = A p over the node-slice layout, fused p.q and alpha = rho / p.q. Each block
// owns a contiguous slice range (XCD-remapped) and its waves take one slice each.
template <int kW>
__global__ __launch_bounds__(kThreads) void pcg_spmv_slices_kernel(PcgBuffers b, SliceView v,
                                                                   const double* __restrict__ planes64,
                                                                   const float* __restrict__ planes32) {
    if (b.st->done) return;
    __shared__ double sh[kThreads / kWave];
    const bool use32 = planes32 != nullptr && (planes64 == nullptr || b.st->phase32 != 0);
    const int lb = xcd_block(blockIdx.x, gridDim.x);
    const int per_block = (v.n_slices + gridDim.x - 1) / gridDim.x;
    const int start = lb * per_block;
    const int stop = min(start + per_block, v.n_slices);
    double pq = 0.0;
    for (int s = start + static_cast<int>(threadIdx.x) / kWave; s < stop; s += kThreads / kWave)
        pq += use32 ? slice_rows<float, kW>(b, v, planes32, s) : slice_rows<double, kW>(b, v, planes64, s);
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

// The same multiply with the search direction of this iteration formed on the fly:
// p = z + beta p_old with z = M^-1 r (solver.cpp:620), computed for every column node a
// lane gathers (from r, the preconditioner and p_old, all read-only here) and written
// for the lane's own node into p_new. The values equal those of the separate
// direction kernel, so the iteration is bit-identical with one launch less.
template <typename VT, int kW, int Prec>
__device__ __forceinline__ double slice_rows_dir(const PcgBuffers& b, const SliceView& v, const VT* __restrict__ planes,
                                                 int s, const double* __restrict__ p_old, double* __restrict__ p_new,
                                                 double beta) {
    using P = typename PairOf<VT>::type;
    const int lane = threadIdx.x & (kWave - 1);
    const int node = s * kWave + lane;
    const int begin = v.ptr[s];
    const int width = (v.ptr[s + 1] - begin) / kWave;
    const P* __restrict__ upper = reinterpret_cast<const P*>(planes);
    const P* __restrict__ lower = upper + v.blocks;
    const dpair* __restrict__ r2 = reinterpret_cast<const dpair*>(b.st->replaced ? b.z : b.r);
    const dpair* __restrict__ po2 = reinterpret_cast<const dpair*>(p_old);
    auto direction = [&](int c) {
        const dpair r = r2[c];
        double z0, z1;
        precond_node<Prec>(b.m, c, r.x, r.y, z0, z1);
        const dpair po = po2[c];
        const dpair pn = {z0 + beta * po.x, z1 + beta * po.y};
        return pn;
    };
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
#pragma unroll
        for (int k = 0; k < kW; ++k) {
            if (k < width) {
                const dpair x = direction(c[k]);
                add_block(a0, a1, static_cast<double>(u[k].x), static_cast<double>(u[k].y),
                          static_cast<double>(l[k].x), static_cast<double>(l[k].y), x, v.pair_sum != 0);
            }
        }
    } else {
        for (int k = 0; k < width; ++k) {
            const int at = begin + k * kWave + lane;
            const P uk = __builtin_nontemporal_load(upper + at);
            const P lk = __builtin_nontemporal_load(lower + at);
            const dpair x = direction(__builtin_nontemporal_load(v.col + at));
            add_block(a0, a1, static_cast<double>(uk.x), static_cast<double>(uk.y), static_cast<double>(lk.x),
                      static_cast<double>(lk.y), x, v.pair_sum != 0);
        }
    }
    double pq = 0.0;
    if (node < (b.n >> 1)) {
        const dpair pv = direction(node);
        if (sizeof(VT) == sizeof(float)) add_diag32(v, node, pv, a0, a1);
        const dpair qv = {a0, a1};
        reinterpret_cast<dpair*>(b.q)[node] = qv;
        reinterpret_cast<dpair*>(p_new)[node] = pv;
        pq = pv.x * a0 + pv.y * a1;
    }
    return pq;
}

template <int kW, int Prec>
__global__ __launch_bounds__(kThreads) void pcg_spmv_slices_dir_kernel(PcgBuffers b, SliceView v,
                                                                       const double* __restrict__ planes64,
                                                                       const float* __restrict__ planes32,
                                                                       const double* __restrict__ p_old,
                                                                       double* __restrict__ p_new) {
    if (b.st->done) return;
    __shared__ double sh[kThreads / kWave];
    const bool use32 = planes32 != nullptr && (planes64 == nullptr || b.st->phase32 != 0);
    const double beta = b.st->beta;
    const int lb = xcd_block(blockIdx.x, gridDim.x);
    const int per_block = (v.n_slices + gridDim.x - 1) / gridDim.x;
    const int start = lb * per_block;
    const int stop = min(start + per_block, v.n_slices);
    double pq = 0.0;
    for (int s = start + static_cast<int>(threadIdx.x) / kWave; s < stop; s += kThreads / kWave)
        pq += use32 ? slice_rows_dir<float, kW, Prec>(b, v, planes32, s, p_old, p_new, beta)
                    : slice_rows_dir<double, kW, Prec>(b, v, planes64, s, p_old, p_new, beta);
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

// Finishes an iteration after r.z is known (solver.cpp:614-621 checks).
__device__ __forceinline__ void accept_rz(PcgState* st, double rz) {
    if (!(rz > 0.0) || !isfinite(rz)) {
        st->status = LFG_ERR_NONFINITE;
        st->done = 1;
        return;
    }
    st->beta = rz / st->rho;
    st->rho = rz;
    if (st->iter >= st->max_iter) st->done = 1;
}

// x += alpha p, r -= alpha q (solver.cpp:599-604), ||r||^2, r.z with z = M^-1 r.
template <int Prec, bool Mixed>
__global__ __launch_bounds__(kThreads) void pcg_update_kernel(PcgBuffers b) {
    if (b.st->done) return;
    __shared__ double sh[kThreads / kWave];
    const double alpha = b.st->alpha;
    // after an accepted replacement the current residual is the FP64 one in z
    const double* __restrict__ rsrc = (Mixed && b.st->replaced) ? b.z : b.r;
    const int n_nodes = b.n >> 1;
    double rr = 0.0, rz = 0.0;
    for (int node = blockIdx.x * kThreads + threadIdx.x; node < n_nodes; node += gridDim.x * kThreads) {
        const double2 p = reinterpret_cast<const double2*>(b.p)[node];
        const double2 q = reinterpret_cast<const double2*>(b.q)[node];
        double2 x = reinterpret_cast<double2*>(b.x)[node];
        double2 r = reinterpret_cast<const double2*>(rsrc)[node];
        x.x += alpha * p.x;
        x.y += alpha * p.y;
        r.x = r.x - alpha * q.x;
        r.y = r.y - alpha * q.y;
        reinterpret_cast<double2*>(b.x)[node] = x;
        reinterpret_cast<double2*>(b.r)[node] = r;
        rr += r.x * r.x + r.y * r.y;
        if (Prec != LFG_PRECOND_IC0) {
            double z0, z1;
            precond_node<Prec>(b.m, node, r.x, r.y, z0, z1);
            rz += r.x * z0 + r.y * z1;
        }
    }
    const double vals[2] = {block_sum<kThreads>(rr, sh), block_sum<kThreads>(rz, sh)};
    double tot[2];
    if (!finish_reduction<kThreads, 2>(b.tickets + 2, b.partials, vals, tot, sh)) return;
    if (threadIdx.x == 0) {
        const double trr = tot[0], trz = tot[1];
        PcgState* st = b.st;
        st->replaced = 0;  // every block has read it; r now holds the updated residual
        st->iter += 1;
        st->rr = trr;
        st->rz_new = trz;
        const double rnorm = sqrt(max0(trr));
        if (rnorm <= st->target) {  // solver.cpp:606-611, on the recurrence residual in every form
            st->converged = 1;
            st->done = 1;
            return;
        }
        if (Mixed && st->phase32 && rnorm < st->drop * st->maxr) {
            // FP32 phase: replace the recurrence residual by the FP64 residual once
            // it has dropped by `drop` since the last replacement.
            st->need_replace = 1;
            return;
        }
        if (Mixed && !st->phase32 && trr <= st->switch_rr) {
            st->phase32 = 1;  // the next SpMVs stream the FP32 values
            st->maxr = rnorm;
        }
        if (Prec == LFG_PRECOND_IC0) {
            if (st->iter >= st->max_iter) st->done = 1;  // z and beta come from the IC(0) step
        } else {
            accept_rz(st, trz);
        }
    }
}

// Mixed precision, when the update kernel asked for it: the FP64 residual b - K x
// into z (tentatively), with its ||.||^2 and r.z. It is accepted as the residual of
// this iteration (state `replaced`: the direction and the next update read it from
// z) while replacements keep halving the FP64 residual; the first one that does not
// has reached the residual floor of FP64 arithmetic on this system, where a replaced
// residual is rounding noise that would restart the recurrence, so it is dropped,
// replacements stop and the recurrence continues as the FP64 form's does.
template <int W, int Prec>
__global__ __launch_bounds__(kThreads) void pcg_replace_kernel(PcgBuffers b, const int* __restrict__ rp,
                                                               const int* __restrict__ col,
                                                               const double* __restrict__ val) {
    if (b.st->done || !b.st->need_replace) return;
    __shared__ double sh[kThreads / kWave];
    constexpr int kNodesPerBlock = kThreads / (2 * W);
    const int group = threadIdx.x / (2 * W);
    const int half = (threadIdx.x / W) & 1;
    const int lane = threadIdx.x & (W - 1);
    const int n_nodes = b.n >> 1;
    double rr = 0.0, rz = 0.0;
    for (int base = blockIdx.x * kNodesPerBlock; base < n_nodes; base += gridDim.x * kNodesPerBlock) {
        const int node = base + group;
        const bool active = node < n_nodes;
        const int row = 2 * min(node, n_nodes - 1) + half;
        const double s = row_dot<W, double>(rp, col, val, b.x, row, lane);
        const double ri = b.b[row] - s;
        const double r_other = __shfl_xor(ri, W, 2 * W);
        if (active && lane == 0) {
            b.z[row] = ri;
            if (half == 0) {
                double z0, z1;
                precond_node<Prec>(b.m, node, ri, r_other, z0, z1);
                rr += ri * ri + r_other * r_other;
                rz += ri * z0 + r_other * z1;
            }
        }
    }
    const double vals[2] = {block_sum<kThreads>(rr, sh), block_sum<kThreads>(rz, sh)};
    double tot[2];
    if (!finish_reduction<kThreads, 2>(b.tickets + 3, b.partials, vals, tot, sh)) return;
    if (threadIdx.x == 0) {
        const double trr = tot[0], trz = tot[1];
        PcgState* st = b.st;
        st->need_replace = 0;
        if (st->true_rr > 0.0 && !(trr <= 0.25 * st->true_rr)) {
            st->drop = 0.0;  // keep the recurrence residual and its r.z
            accept_rz(st, st->rz_new);
            return;
        }
        st->replacements += 1;
        st->replaced = 1;
        st->rr = trr;
        st->true_rr = trr;
        const double rnorm = sqrt(max0(trr));
        st->maxr = rnorm;
        if (rnorm <= st->target) {
            st->converged = 1;
            st->done = 1;
            return;
        }
        accept_rz(st, trz);
    }
}

// p = z + beta p (solver.cpp:620).
template <int Prec>
__global__ __launch_bounds__(kThreads) void pcg_direction_kernel(PcgBuffers b) {
    if (b.st->done) return;
    const double beta = b.st->beta;
    const double* __restrict__ rsrc = b.st->replaced ? b.z : b.r;  // mixed: accepted FP64 residual
    const int n_nodes = b.n >> 1;
    for (int node = blockIdx.x * kThreads + threadIdx.x; node < n_nodes; node += gridDim.x * kThreads) {
        double z0, z1;
        if (Prec == LFG_PRECOND_IC0) {
            const double2 z = reinterpret_cast<const double2*>(b.z)[node];
            z0 = z.x;
            z1 = z.y;
        } else {
            const double2 r = reinterpret_cast<const double2*>(rsrc)[node];
            precond_node<Prec>(b.m, node, r.x, r.y, z0, z1);
        }
        double2 p = reinterpret_cast<double2*>(b.p)[node];
        p.x = z0 + beta * p.x;
        p.y = z1 + beta * p.y;
        reinterpret_cast<double2*>(b.p)[node] = p;
    }
}

// ||b - A x||^2 into out2 and, when r_out is given, r = b - A x into r_out; blocks own
// contiguous row ranges, the last block folds the partials.
template <int W>
__global__ __launch_bounds__(kThreads) void residual_norm_kernel(int n, const int* __restrict__ rp,
                                                                 const int* __restrict__ col,
                                                                 const double* __restrict__ val,
                                                                 const double* __restrict__ x,
                                                                 const double* __restrict__ bvec,
                                                                 double* __restrict__ partials,
                                                                 TicketSet* tickets, double* out2,
                                                                 double* __restrict__ r_out) {
    __shared__ double sh[kThreads / kWave];
    constexpr int kRows = kThreads / W;
    const int per_block = (n + gridDim.x - 1) / gridDim.x;
    const int start = blockIdx.x * per_block;
    const int stop = min(start + per_block, n);
    const int lane = threadIdx.x & (W - 1);
    double acc = 0.0;
    for (int base = start; base < stop; base += kRows) {
        const int row = base + threadIdx.x / W;
        if (row < stop) {
            const double s = row_dot<W, double>(rp, col, val, x, row, lane);
            if (lane == 0) {
                const double ri = bvec[row] - s;
                acc += ri * ri;
                if (r_out) r_out[row] = ri;
            }
        }
    }
    const double vals[1] = {block_sum<kThreads>(acc, sh)};
    double tot[1];
    if (!finish_reduction<kThreads, 1>(tickets, partials, vals, tot, sh)) return;
    if (threadIdx.x == 0) *out2 = tot[0];
}

// y += a x (the iterative-refinement correction of the solution).
__global__ __launch_bounds__(kThreads) void axpy_kernel(long long n, double a, const double* __restrict__ x,
                                                        double* __restrict__ y) {
    for (long long i = blockIdx.x * static_cast<long long>(kThreads) + threadIdx.x; i < n;
         i += static_cast<long long>(gridDim.x) * kThreads)
        y[i] += a * x[i];
}

// Streaming calibration: e = a + b + c + d over n pairs, one 16-B pair per thread (four
// streams read, one written, the read / write mix of the fused PCG iterations).
__global__ __launch_bounds__(kThreads) void stream_probe_kernel(long long n, const dpair* __restrict__ a,
                                                                const dpair* __restrict__ b,
                                                                const dpair* __restrict__ c,
                                                                const dpair* __restrict__ d, dpair* __restrict__ e) {
    const long long i = blockIdx.x * static_cast<long long>(kThreads) + threadIdx.x;
    if (i < n) {
        const dpair u = a[i], v = b[i], w = c[i], z = d[i];
        e[i] = dpair{u.x + v.x + w.x + z.x, u.y + v.y + w.y + z.y};
    }
}

__global__ void mask_constrained_kernel(int n, const int* __restrict__ constrained, const double* __restrict__ in,
                                        double* __restrict__ out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = constrained[i] ? 0.0 : in[i];
}

// SpMV variants: 1, 2, 3 -> node groups with 4, 8, 16 lanes per row; 4 (the
// default) and 5 -> row-segment streaming with 64 / 32 rows per wave chunk.
constexpr int kDefaultVariant = 4;

inline int variant_width(int variant) {
    switch (variant) {
        case 1: return 4;
        case 3: return 16;
        default: return 8;
    }
}

inline unsigned int update_blocks(int n) {
    const unsigned int b = blocks_for(n >> 1, kThreads);
    return round8(b < 4096u ? b : 4096u);
}

}  // namespace

// ------------------------------------------------------------ launchers

void launch_dirichlet(hipStream_t s, const CsrView& a, const double* values, const double* f, const int* constrained,
                      const double* prescribed, double* kc, double* fc) {
    if (a.n <= 0) return;
    dirichlet_kernel<<<blocks_for(8LL * a.n, 256), 256, 0, s>>>(a.n, a.row_ptr, a.col, values, f, constrained,
                                                                 prescribed, kc, fc);
}

void launch_pin_zero_pivots(hipStream_t s, const CsrView& a, double* values, double* f,
                            unsigned long long* scale_bits, int* counters) {
    if (a.n <= 0) return;
    const unsigned int grid = blocks_for(a.n, 256);
    abs_max_kernel<<<grid < 1024u ? grid : 1024u, 256, 0, s>>>(a.n, f, scale_bits);
    pin_pivots_kernel<<<grid, 256, 0, s>>>(a.n, a.row_ptr, a.col, values, f, scale_bits, counters);
}

void launch_spmv_serial_rows(hipStream_t s, const CsrView& a, const double* values, const double* x,
                             const double* sub, double* y) {
    if (a.n <= 0) return;
    spmv_serial_rows_kernel<<<blocks_for(a.n, 256), 256, 0, s>>>(a.n, a.row_ptr, a.col, values, x, sub, y);
}

void launch_extract_diag(hipStream_t s, const CsrView& a, const double* values, double* diag, int* status) {
    extract_diag_kernel<<<blocks_for(a.n, 256), 256, 0, s>>>(a.n, a.row_ptr, a.col, values, diag, status);
}

void launch_block2_factor(hipStream_t s, const CsrView& a, const double* values, double* factor, int* degenerate) {
    block2_factor_kernel<<<blocks_for(a.n / 2, 256), 256, 0, s>>>(a.n / 2, a.row_ptr, a.col, values, factor,
                                                                   degenerate);
}

void launch_jacobi_inv(hipStream_t s, int n, const double* diag, double* inv, int* status) {
    jacobi_inv_kernel<<<blocks_for(n, 256), 256, 0, s>>>(n, diag, inv, status);
}

void launch_precond_apply(hipStream_t s, int kind, int n, const double* m, const double* in, double* out) {
    precond_apply_kernel<<<blocks_for(n / 2, 256), 256, 0, s>>>(kind, n / 2, m, in, out);
}

void launch_to_f32(hipStream_t s, const double* in, float* out, long long n) {
    if (n <= 0) return;
    to_f32_kernel<<<blocks_for(n, 256), 256, 0, s>>>(in, out, n);
}

void launch_mask_constrained(hipStream_t s, int n, const int* constrained, const double* in, double* out) {
    if (n <= 0) return;
    mask_constrained_kernel<<<blocks_for(n, 256), 256, 0, s>>>(n, constrained, in, out);
}

// Fixed SpMV grid: at most kSpmvBlocks blocks, each looping over its node range.
constexpr unsigned int kSpmvBlocks = 2048;

int spmv_blocks(int n, int variant) {
    if (variant == 0) variant = kDefaultVariant;
    if (variant == 4 || variant == 5) {
        const int rows = variant == 4 ? 64 : 32;
        const unsigned int need = blocks_for((n + rows - 1) / rows, kThreads / kWave);
        return static_cast<int>(round8(need < kSpmvBlocks ? need : kSpmvBlocks));
    }
    const int w = variant_width(variant);
    const int nodes_per_step = kThreads / (2 * w);
    const unsigned int need = blocks_for(n >> 1, nodes_per_step);
    return static_cast<int>(round8(need < kSpmvBlocks ? need : kSpmvBlocks));
}

void launch_pcg_init(hipStream_t s, const PcgBuffers& b, int stop_rule, double tol, long long max_iter, double drop,
                     double switch_tau, int fused) {
    const unsigned int grid = update_blocks(b.n);
    switch (b.precond) {
        case LFG_PRECOND_JACOBI:
            pcg_init_kernel<LFG_PRECOND_JACOBI>
                <<<grid, kThreads, 0, s>>>(b, stop_rule, tol, max_iter, drop, switch_tau, fused);
            break;
        case LFG_PRECOND_BLOCK2:
            pcg_init_kernel<LFG_PRECOND_BLOCK2>
                <<<grid, kThreads, 0, s>>>(b, stop_rule, tol, max_iter, drop, switch_tau, fused);
            break;
        default:
            pcg_init_kernel<LFG_PRECOND_IC0>
                <<<grid, kThreads, 0, s>>>(b, stop_rule, tol, max_iter, drop, switch_tau, 0);
            break;
    }
}

template <int W>
static void spmv_w(hipStream_t s, const PcgBuffers& b, const CsrView& a, const double* v64, const float* v32) {
    const unsigned int grid = static_cast<unsigned int>(spmv_blocks(b.n, W == 4 ? 1 : (W == 16 ? 3 : 2)));
    pcg_spmv_kernel<W><<<grid, kThreads, 0, s>>>(b, a.row_ptr, a.col, v64, v32);
}

void launch_pcg_spmv(hipStream_t s, const PcgBuffers& b, const CsrView& a, const double* vals64, const float* vals32,
                     int variant) {
    if (variant == 0) variant = kDefaultVariant;
    if (variant == 4 || variant == 5) {
        const unsigned int grid = static_cast<unsigned int>(spmv_blocks(b.n, variant));
        if (variant == 4) {
            if (a.max_row <= 24)
                pcg_spmv_stream_kernel<24, 64><<<grid, kThreads, 0, s>>>(b, a.row_ptr, a.col, vals64, vals32);
            else
                pcg_spmv_stream_kernel<0, 64><<<grid, kThreads, 0, s>>>(b, a.row_ptr, a.col, vals64, vals32);
        } else {
            if (a.max_row <= 24)
                pcg_spmv_stream_kernel<24, 32><<<grid, kThreads, 0, s>>>(b, a.row_ptr, a.col, vals64, vals32);
            else
                pcg_spmv_stream_kernel<0, 32><<<grid, kThreads, 0, s>>>(b, a.row_ptr, a.col, vals64, vals32);
        }
        return;
    }
    switch (variant_width(variant)) {
        case 4: spmv_w<4>(s, b, a, vals64, vals32); break;
        case 16: spmv_w<16>(s, b, a, vals64, vals32); break;
        default: spmv_w<8>(s, b, a, vals64, vals32); break;
    }
}

void launch_pack_slices(hipStream_t s, const CsrView& a, const SliceView& v, const double* values, double* out64,
                        float* out32, int* col_out, double* diag_out) {
    if (v.n_slices <= 0) return;
    pack_slices_kernel<<<blocks_for(static_cast<long long>(v.n_slices) * kWave, 256), 256, 0, s>>>(
        a.n >> 1, v.n_slices, v.blocks, a.row_ptr, a.col, v.ptr, values, out64, out32, col_out, diag_out);
}

int slice_spmv_blocks(int n_slices) {
    const unsigned int need = blocks_for(n_slices, kThreads / kWave);
    return static_cast<int>(round8(need < 4096u ? need : 4096u));
}

void launch_pcg_spmv_slices(hipStream_t s, const PcgBuffers& b, const SliceView& v, const double* planes64,
                            const float* planes32) {
    const unsigned int grid = static_cast<unsigned int>(slice_spmv_blocks(v.n_slices));
    if (v.max_width <= 9)
        pcg_spmv_slices_kernel<9><<<grid, kThreads, 0, s>>>(b, v, planes64, planes32);
    else if (v.max_width <= 16)
        pcg_spmv_slices_kernel<16><<<grid, kThreads, 0, s>>>(b, v, planes64, planes32);
    else
        pcg_spmv_slices_kernel<0><<<grid, kThreads, 0, s>>>(b, v, planes64, planes32);
}

void launch_pcg_spmv_slices_dir(hipStream_t s, const PcgBuffers& b, const SliceView& v, const double* planes64,
                                const float* planes32, const double* p_old, double* p_new) {
    const unsigned int grid = static_cast<unsigned int>(slice_spmv_blocks(v.n_slices));
    const bool narrow = v.max_width <= 9;
    if (b.precond == LFG_PRECOND_BLOCK2) {
        if (narrow)
            pcg_spmv_slices_dir_kernel<9, LFG_PRECOND_BLOCK2><<<grid, kThreads, 0, s>>>(b, v, planes64, planes32, p_old,
                                                                                      p_new);
        else
            pcg_spmv_slices_dir_kernel<0, LFG_PRECOND_BLOCK2><<<grid, kThreads, 0, s>>>(b, v, planes64, planes32, p_old,
                                                                                      p_new);
    } else {
        if (narrow)
            pcg_spmv_slices_dir_kernel<9, LFG_PRECOND_JACOBI><<<grid, kThreads, 0, s>>>(b, v, planes64, planes32, p_old,
                                                                                      p_new);
        else
            pcg_spmv_slices_dir_kernel<0, LFG_PRECOND_JACOBI><<<grid, kThreads, 0, s>>>(b, v, planes64, planes32, p_old,
                                                                                      p_new);
    }
}

void launch_pcg_update(hipStream_t s, const PcgBuffers& b, int mixed) {
    const unsigned int grid = update_blocks(b.n);
    if (mixed) {
        if (b.precond == LFG_PRECOND_BLOCK2) pcg_update_kernel<LFG_PRECOND_BLOCK2, true><<<grid, kThreads, 0, s>>>(b);
        else pcg_update_kernel<LFG_PRECOND_JACOBI, true><<<grid, kThreads, 0, s>>>(b);
        return;
    }
    switch (b.precond) {
        case LFG_PRECOND_JACOBI: pcg_update_kernel<LFG_PRECOND_JACOBI, false><<<grid, kThreads, 0, s>>>(b); break;
        case LFG_PRECOND_BLOCK2: pcg_update_kernel<LFG_PRECOND_BLOCK2, false><<<grid, kThreads, 0, s>>>(b); break;
        default: pcg_update_kernel<LFG_PRECOND_IC0, false><<<grid, kThreads, 0, s>>>(b); break;
    }
}

void launch_pcg_replace(hipStream_t s, const PcgBuffers& b, const CsrView& a, const double* vals64) {
    constexpr int W = 8;
    const unsigned int grid = 512;
    if (b.precond == LFG_PRECOND_BLOCK2)
        pcg_replace_kernel<W, LFG_PRECOND_BLOCK2><<<grid, kThreads, 0, s>>>(b, a.row_ptr, a.col, vals64);
    else
        pcg_replace_kernel<W, LFG_PRECOND_JACOBI><<<grid, kThreads, 0, s>>>(b, a.row_ptr, a.col, vals64);
}

void launch_pcg_direction(hipStream_t s, const PcgBuffers& b) {
    const unsigned int grid = update_blocks(b.n);
    switch (b.precond) {
        case LFG_PRECOND_JACOBI: pcg_direction_kernel<LFG_PRECOND_JACOBI><<<grid, kThreads, 0, s>>>(b); break;
        case LFG_PRECOND_BLOCK2: pcg_direction_kernel<LFG_PRECOND_BLOCK2><<<grid, kThreads, 0, s>>>(b); break;
        default: pcg_direction_kernel<LFG_PRECOND_IC0><<<grid, kThreads, 0, s>>>(b); break;
    }
}

void launch_residual_norm(hipStream_t s, const CsrView& a, const double* vals64, const double* x, const double* bvec,
                          double* partials, TicketSet* tickets, double* out2, double* r_out) {
    constexpr int W = 8;
    const unsigned int need = round8(blocks_for(a.n, kThreads / W));
    const unsigned int grid = need < 2048u ? need : 2048u;
    residual_norm_kernel<W>
        <<<grid, kThreads, 0, s>>>(a.n, a.row_ptr, a.col, vals64, x, bvec, partials, tickets, out2, r_out);
}

void launch_stream_probe(hipStream_t s, long long n, const double* a, const double* b, const double* c,
                         const double* d, double* e) {
    const long long blocks = (n + kThreads - 1) / kThreads;
    stream_probe_kernel<<<static_cast<unsigned>(blocks > 0 ? blocks : 1), kThreads, 0, s>>>(
        n, reinterpret_cast<const dpair*>(a), reinterpret_cast<const dpair*>(b), reinterpret_cast<const dpair*>(c),
        reinterpret_cast<const dpair*>(d), reinterpret_cast<dpair*>(e));
}

void launch_axpy(hipStream_t s, long long n, double a, const double* x, double* y) {
    const unsigned int need = blocks_for(n, kThreads);
    axpy_kernel<<<need < 4096u ? (need > 0 ? need : 1u) : 4096u, kThreads, 0, s>>>(n, a, x, y);
}

}  // namespace lfg
