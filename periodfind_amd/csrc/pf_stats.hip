// On-device 'stats' output: per curve, the summary numpy computes in
// Statistics.statistics_from_data (periodfind/__init__.py:337-427) — mean,
// std, median, MAD over the whole [period][pdt] grid — and the top-n trials.
//
// Bit-compatibility with numpy float32:
//   * np.sum / np.mean: pairwise summation (8 accumulators over leaves of at
//     most 128 elements, halving splits rounded down to a multiple of 8)
//     within blocks of 8192 elements, blocks accumulated in order; mean is
//     f32(f64(sum) / n) (numpy divides by an intp count, which promotes the
//     division to float64 before the float32 cast).
//   * np.std: x = a - mean, x*x, the same summation, f32(f64 / n), sqrt.
//   * np.median: middle element, or fl(fl(a + b) / 2) of the two middle ones;
//     MAD = median(|a - median|).  Order statistics by radix selection on
//     order-preserving uint32 keys (no sort).
//   * top-n: the n best values; among equal values the lower flat index wins
//     (numpy's choice among exact ties is unspecified).
// One workgroup per curve, T <= kStatsMaxT values staged in LDS; a curve
// containing NaN is flagged so the caller can defer to numpy.
#include "pf_device.hpp"
#include "pf_kernels.hpp"
#include "pf_numpy.hpp"

namespace pf {

namespace {

constexpr int kThreads = 256;

__device__ __forceinline__ uint32_t order_key(float f) {
    const uint32_t u = __float_as_uint(f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}

__device__ __forceinline__ float key_value(uint32_t k) {
    const uint32_t u = (k & 0x80000000u) ? (k & 0x7fffffffu) : ~k;
    return __uint_as_float(u);
}

using npsum::LeafPlan;
using npsum::plan_leaves;

// np.add.reduce of x[0..n), n <= 8192 (one numpy block), by the whole block;
// result in *out.  L must hold the plan of n (plan_leaves, then a barrier).
__device__ void block_numpy_sum(const float* x, int n, LeafPlan& L, float* out) {
    const float s = npsum::block_pairwise(x, n, L);
    if (threadIdx.x == 0) *out = 0.0f + s;
    __syncthreads();
}

// Exclusive prefix of a per-thread flag over the block in thread order;
// returns this thread's prefix and writes the block total to *total.
__device__ int block_scan(bool flag, int* wave_tot, int* total) {
    const int lane = threadIdx.x & 63;
    const int wv = threadIdx.x >> 6;
    const uint64_t m = __ballot(flag);
    const int in_wave = __popcll(m & ((lane == 0) ? 0ull : (~0ull >> (64 - lane))));
    if (lane == 0) wave_tot[wv] = __popcll(m);
    __syncthreads();
    int base = 0, tot = 0;
    for (int w = 0; w < kThreads / 64; ++w) {
        if (w < wv) base += wave_tot[w];
        tot += wave_tot[w];
    }
    __syncthreads();
    *total = tot;
    return base + in_wave;
}

static_assert(kThreads == 256, "radix_select maps one histogram bin to each thread");

// Block reductions (red: kThreads / 64 slots of LDS).
__device__ int block_sum_int(int v, int* red) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    int tot = 0;
    for (int w = 0; w < kThreads / 64; ++w) tot += red[w];
    __syncthreads();
    return tot;
}

__device__ uint32_t block_min_u32(uint32_t v, uint32_t* red) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        const uint32_t u = __shfl_xor(v, o);
        v = u < v ? u : v;
    }
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    uint32_t m = red[0];
    for (int w = 1; w < kThreads / 64; ++w) m = red[w] < m ? red[w] : m;
    __syncthreads();
    return m;
}

// The k-th smallest (0-based) of keys[0..n): four 8-bit radix passes, each
// histogramming the keys that match the digits fixed so far (per-wave LDS
// histograms, hist: 4 x 256 ints) and fixing the digit whose bin holds rank
// k (one bin per thread, block inclusive scan).  sel: 2 + kThreads/64 ints.
__device__ uint32_t radix_select(const uint32_t* keys, int n, int k, int* hist, int* sel) {
    const int tid = threadIdx.x, wv = tid >> 6, lane = tid & 63;
    uint32_t prefix = 0, mask = 0;
    for (int shift = 24; shift >= 0; shift -= 8) {
        for (int i = tid; i < 4 * 256; i += kThreads) hist[i] = 0;
        __syncthreads();
        int* h = hist + wv * 256;
        for (int i = tid; i < n; i += kThreads) {
            const uint32_t key = keys[i];
            if ((key & mask) == prefix) atomicAdd(&h[(key >> shift) & 255u], 1);
        }
        __syncthreads();
        const int cnt = hist[tid] + hist[256 + tid] + hist[512 + tid] + hist[768 + tid];
        int incl = cnt;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const int v = __shfl_up(incl, o);
            if (lane >= o) incl += v;
        }
        if (lane == 63) sel[2 + wv] = incl;
        __syncthreads();
        for (int w = 0; w < wv; ++w) incl += sel[2 + w];
        const int excl = incl - cnt;
        if (k >= excl && k < incl) {
            sel[0] = tid;
            sel[1] = excl;
        }
        __syncthreads();
        prefix |= static_cast<uint32_t>(sel[0]) << shift;
        mask |= 255u << shift;
        k -= sel[1];
        __syncthreads();
    }
    return prefix;
}

// numpy median of the n keys: the middle one, or fl(fl(a + b) / 2) of the
// two middle ones (the upper is the lower again if it repeats, else the
// smallest key above it).
__device__ float block_median(const uint32_t* keys, int n, int* hist, int* sel, int* red) {
    const uint32_t lo = radix_select(keys, n, (n - 1) / 2, hist, sel);
    if (n & 1) return key_value(lo);
    int le = 0;
    uint32_t above = 0xffffffffu;
    for (int i = threadIdx.x; i < n; i += kThreads) {
        const uint32_t key = keys[i];
        le += key <= lo ? 1 : 0;
        if (key > lo && key < above) above = key;
    }
    le = block_sum_int(le, red);
    above = block_min_u32(above, reinterpret_cast<uint32_t*>(red));
    const uint32_t hi = le >= n / 2 + 1 ? lo : above;
    float s = -0.0f;
    s += key_value(lo);
    s += key_value(hi);
    return s / 2.0f;
}

__global__ __launch_bounds__(kThreads) void stats_kernel(const float* __restrict__ data,
                                                         int64_t T, int n_top, int use_max,
                                                         float* __restrict__ summary,
                                                         int64_t* __restrict__ top_idx,
                                                         float* __restrict__ top_val,
                                                         int* __restrict__ flags) {
    extern __shared__ uint32_t sm[];
    const int n = static_cast<int>(T);
    float* a = reinterpret_cast<float*>(sm);  // original order [n]
    uint32_t* k = sm + n;                     // order keys [n]
    float* d = reinterpret_cast<float*>(k);   // scratch aliasing keys
    __shared__ float s_mean, s_sum;
    __shared__ int s_nan;
    __shared__ int wave_tot[kThreads / 64];
    __shared__ int hist[4 * 256];
    __shared__ int sel[2 + kThreads / 64];
    __shared__ LeafPlan plan;
    const int64_t c = blockIdx.x;
    const float* row = data + c * T;
    if (threadIdx.x == 0) {
        s_nan = 0;
        plan_leaves(n, plan);
    }
    __syncthreads();
    for (int i = threadIdx.x; i < n; i += kThreads) {
        const float v = row[i];
        a[i] = v;
        if (v != v) s_nan = 1;
    }
    __syncthreads();
    if (s_nan) {
        if (threadIdx.x == 0) flags[c] = 1;
        return;
    }
    // mean
    block_numpy_sum(a, n, plan, &s_sum);
    if (threadIdx.x == 0) {
        s_mean = static_cast<float>(static_cast<double>(s_sum) / static_cast<double>(n));
        flags[c] = 0;
    }
    __syncthreads();
    const float mean = s_mean;
    // std: x = a - mean; x = x * x; sum; / n; sqrt.
    for (int i = threadIdx.x; i < n; i += kThreads) {
        const float x = a[i] - mean;
        d[i] = x * x;
    }
    __syncthreads();
    block_numpy_sum(d, n, plan, &s_sum);
    if (threadIdx.x == 0) {
        const float var =
            static_cast<float>(static_cast<double>(s_sum) / static_cast<double>(n));
        summary[c * 4 + 0] = mean;
        // f64 sqrt then f32: correctly rounded (__fsqrt_rn maps to the native
        // approximate sqrt unless OCML's rounded operations are enabled).
        summary[c * 4 + 1] = static_cast<float>(__builtin_sqrt(static_cast<double>(var)));
    }
    __syncthreads();
    // median, and the top-n threshold: the n_top-th best key and the number
    // of keys strictly better than it.
    for (int i = threadIdx.x; i < n; i += kThreads) k[i] = order_key(a[i]);
    __syncthreads();
    const float med = block_median(k, n, hist, sel, wave_tot);
    uint32_t thr_key = 0;
    int n_better = 0;
    if (n_top == 1) {
        // The best key by a block reduction; nothing is strictly better.
        uint32_t best = use_max ? 0u : 0xffffffffu;
        for (int i = threadIdx.x; i < n; i += kThreads)
            best = use_max ? (k[i] > best ? k[i] : best) : (k[i] < best ? k[i] : best);
        thr_key = use_max ? ~block_min_u32(~best, reinterpret_cast<uint32_t*>(wave_tot))
                          : block_min_u32(best, reinterpret_cast<uint32_t*>(wave_tot));
    } else if (n_top > 1) {
        thr_key = radix_select(k, n, use_max ? n - n_top : n_top - 1, hist, sel);
        int b = 0;
        for (int i = threadIdx.x; i < n; i += kThreads)
            b += (use_max ? k[i] > thr_key : k[i] < thr_key) ? 1 : 0;
        n_better = block_sum_int(b, wave_tot);
    }
    // MAD
    for (int i = threadIdx.x; i < n; i += kThreads) k[i] = order_key(__builtin_fabsf(a[i] - med));
    __syncthreads();
    const float mad = block_median(k, n, hist, sel, wave_tot);
    if (threadIdx.x == 0) {
        summary[c * 4 + 2] = med;
        summary[c * 4 + 3] = mad;
    }
    // Top-n: values strictly better than the threshold, then equal values in
    // index order, compacted in index order by block scans; then a best-first
    // insertion sort of the n_top entries (equal values keep index order).
    if (n_top > 0) {
        const float thr = key_value(thr_key);
        const int nb = n_better;
        int64_t* oi = top_idx + c * n_top;
        float* ov = top_val + c * n_top;
        int base_b = 0, base_t = 0;
        for (int r = 0; r < n; r += kThreads) {
            const int i = r + threadIdx.x;
            const float v = i < n ? a[i] : 0.0f;
            const bool bt = i < n && (use_max ? (v > thr) : (v < thr));
            const bool tie = i < n && v == thr;
            int tot_b, tot_t;
            const int pb = block_scan(bt, wave_tot, &tot_b);
            const int pt = block_scan(tie, wave_tot, &tot_t);
            if (bt) {
                oi[base_b + pb] = i;
                ov[base_b + pb] = v;
            }
            if (tie && nb + base_t + pt < n_top) {
                oi[nb + base_t + pt] = i;
                ov[nb + base_t + pt] = v;
            }
            base_b += tot_b;
            base_t += tot_t;
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            for (int x = 1; x < n_top; ++x) {
                const float v = ov[x];
                const int64_t id = oi[x];
                int y = x - 1;
                while (y >= 0 && (use_max ? (ov[y] < v || (ov[y] == v && oi[y] > id))
                                          : (ov[y] > v || (ov[y] == v && oi[y] > id)))) {
                    ov[y + 1] = ov[y];
                    oi[y + 1] = oi[y];
                    --y;
                }
                ov[y + 1] = v;
                oi[y + 1] = id;
            }
        }
    }
}

}  // namespace

// Dynamic LDS: values and keys, T each.
size_t stats_lds_bytes(int64_t T) { return static_cast<size_t>(2 * T) * 4; }

size_t stats_lds_static_bytes() { return sizeof(LeafPlan) + (4 * 256 + 16) * sizeof(int) + 64; }

hipError_t launch_stats(const float* data, int64_t n_curves, int64_t T, int n_top, int use_max,
                        float* summary, int64_t* top_idx, float* top_val, int* flags,
                        hipStream_t s) {
    if (n_curves <= 0 || T <= 0) return hipSuccess;
    if (T > kStatsMaxT || n_top > T) return hipErrorInvalidValue;
    hipLaunchKernelGGL(stats_kernel, dim3(static_cast<uint32_t>(n_curves)), dim3(kThreads),
                       stats_lds_bytes(T), s, data, T, n_top, use_max, summary, top_idx, top_val,
                       flags);
    return hipGetLastError();
}

}  // namespace pf
