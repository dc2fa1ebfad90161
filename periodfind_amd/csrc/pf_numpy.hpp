// numpy's float32 summation, bit for bit, by one workgroup.
//
// np.add.reduce over a contiguous float32 array (np.sum, np.mean) splits the
// array into blocks of 8192 elements (the ufunc buffer size) and adds the
// blocks' pairwise sums in order to an accumulator that starts at 0.  The
// pairwise sum of a block of n elements: n < 8 -> serial; n <= 128 -> eight
// accumulators over strided elements, combined ((r0+r1)+(r2+r3)) +
// ((r4+r5)+(r6+r7)), then the remainder serially; else the halves [0, n2) and
// [n2, n), n2 = n/2 rounded down to a multiple of 8, summed recursively.
//
// Here the split tree of one block is enumerated once into LDS (its leaves
// in left-to-right order); the leaf sums run in parallel, one thread per
// leaf; thread 0 then combines them in the tree's order.
#pragma once

#include <hip/hip_runtime.h>

namespace pf {
namespace npsum {

constexpr int kBlock = 8192;    // numpy's ufunc buffer size
constexpr int kMaxLeaves = 72;  // the tree of n <= 8192 has at most 65 leaves

struct LeafPlan {
    int start[kMaxLeaves];
    int len[kMaxLeaves];
    float sum[kMaxLeaves];
    int count;
};

// Left-to-right leaves of the split tree of [0, n), n <= kBlock (thread 0).
// The explicit stack lives in LDS: a dynamically indexed private array would
// be placed in scratch memory, one global-memory round trip per push and pop.
__device__ inline void plan_leaves(int n, LeafPlan& L) {
    __shared__ int st_s[16], st_n[16];
    int sp = 0;
    st_s[0] = 0;
    st_n[0] = n;
    sp = 1;
    int count = 0;
    while (sp > 0) {
        --sp;
        const int s0 = st_s[sp], m = st_n[sp];
        if (m <= 128) {
            L.start[count] = s0;
            L.len[count] = m;
            ++count;
            continue;
        }
        int n2 = m / 2;
        n2 -= n2 % 8;
        // push right first so the left half is visited first
        st_s[sp] = s0 + n2;
        st_n[sp] = m - n2;
        ++sp;
        st_s[sp] = s0;
        st_n[sp] = n2;
        ++sp;
    }
    L.count = count;
}

// One leaf (n <= 128).
__device__ inline float leaf_sum(const float* x, int n) {
    if (n < 8) {
        float res = -0.0f;
        for (int i = 0; i < n; ++i) res += x[i];
        return res;
    }
    float r[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) r[j] = x[j];
    int i = 8;
    for (; i < n - (n % 8); i += 8) {
#pragma unroll
        for (int j = 0; j < 8; ++j) r[j] += x[i + j];
    }
    float res = ((r[0] + r[1]) + (r[2] + r[3])) + ((r[4] + r[5]) + (r[6] + r[7]));
    for (; i < n; ++i) res += x[i];
    return res;
}

// Post-order combination of the leaf sums (thread 0; stack in LDS, as above).
__device__ inline float combine_leaves(int n, const LeafPlan& L) {
    __shared__ int st_n[16], st_state[16];
    __shared__ float st_left[16];
    int sp = 0;
    st_n[0] = n;
    st_state[0] = 0;
    int next = 0;
    float ret = 0.0f;
    while (true) {
        const int fn = st_n[sp];
        if (fn <= 128) {
            ret = L.sum[next++];
            if (sp == 0) return ret;
            --sp;
            continue;
        }
        int n2 = fn / 2;
        n2 -= n2 % 8;
        const int state = st_state[sp];
        if (state == 0) {
            st_state[sp] = 1;
            st_n[sp + 1] = n2;
            st_state[sp + 1] = 0;
            ++sp;
        } else if (state == 1) {
            st_left[sp] = ret;
            st_state[sp] = 2;
            st_n[sp + 1] = fn - n2;
            st_state[sp + 1] = 0;
            ++sp;
        } else {
            ret = st_left[sp] + ret;
            if (sp == 0) return ret;
            --sp;
        }
    }
}

// Pairwise sum of one block x[0..n), n <= kBlock, by the whole workgroup;
// L must hold the plan of n (plan_leaves, then a barrier).  The result is
// returned to thread 0 only.
__device__ inline float block_pairwise(const float* x, int n, LeafPlan& L) {
    for (int i = threadIdx.x; i < L.count; i += blockDim.x)
        L.sum[i] = leaf_sum(x + L.start[i], L.len[i]);
    __syncthreads();
    float r = 0.0f;
    if (threadIdx.x == 0) r = combine_leaves(n, L);
    __syncthreads();
    return r;
}

// np.add.reduce of x[0..n) for any n >= 0, by the whole workgroup; the
// result is written to *out (LDS) and visible to every thread on return.
// Full blocks share one plan; the last, shorter block gets its own.
__device__ inline void block_numpy_sum(const float* x, int64_t n, LeafPlan& L, float* out) {
    float acc = 0.0f;  // thread 0's running sum over the blocks
    int planned = -1;
    for (int64_t b = 0; b < n; b += kBlock) {
        const int m = static_cast<int>(n - b < kBlock ? n - b : kBlock);
        if (m != planned) {
            if (threadIdx.x == 0) plan_leaves(m, L);
            __syncthreads();
            planned = m;
        }
        const float s = block_pairwise(x + b, m, L);
        if (threadIdx.x == 0) acc = acc + s;
    }
    if (threadIdx.x == 0) *out = acc;
    __syncthreads();
}

}  // namespace npsum
}  // namespace pf
