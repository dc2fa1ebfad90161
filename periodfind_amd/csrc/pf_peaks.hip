// On-device top-N peak selection with the semantics of rust/src/peaks.rs.
//
// The reference scans the flattened [period][pdt] periodogram once, keeping
// a candidate strict local extremum that better extrema may replace within
// `min_distance` samples, and a best-first list of at most n_peaks accepted
// peaks whose worst entry is the acceptance threshold (peaks.rs:18-238).
//
// One wave64 per curve.  The scan advances 64 samples at a time: each lane
// tests one sample for being a strict local extremum better than the
// threshold at the start of the chunk (the threshold only tightens, so this
// filter never drops a sample the serial scan would accept), and the wave
// then replays the serial state machine over the surviving samples in index
// order.  Between two surviving samples the serial scan only advances the
// candidate's distance counter, so a pending candidate whose promotion index
// (candidate + min_distance - 1) precedes the next survivor is promoted
// first.  The accepted list is held one entry per lane (n_peaks <= 64):
// insertion is a ballot for the position plus a one-lane shift.
//
// Few long curves (a single curve per call, say) leave most of the device
// idle under one wave per curve.  With min_distance 1 the serial scan keeps
// exactly the n_peaks best strict local extrema under the order "better
// value first, then lower index" (every extremum that beats the current
// worst is inserted at once, ties keep the earlier index), so a curve's
// samples can be cut into segments: one wave keeps each segment's best
// n_peaks (peaks_kernel over a sample range), and a merge wave per curve
// selects the best n_peaks of the union by the same order (peaks_merge_kernel).
// Every extremum of the global selection is among its segment's selection,
// so the result is the serial scan's.
#include "pf_device.hpp"
#include "pf_kernels.hpp"

namespace pf {

namespace {

struct PeakList {
    int64_t li;   // this lane's entry index
    float lv;     // this lane's entry value
    int count;    // entries in use (uniform)
};

__device__ __forceinline__ bool better(float a, float b, bool use_max) {
    return use_max ? (a > b) : (a < b);
}

__device__ __forceinline__ float lane_value(float v, int lane) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), lane));
}

__device__ __forceinline__ int64_t lane_value64(int64_t v, int lane) {
    const int lo = __builtin_amdgcn_readlane(static_cast<int>(v & 0xffffffff), lane);
    const int hi = __builtin_amdgcn_readlane(static_cast<int>(v >> 32), lane);
    return (static_cast<int64_t>(hi) << 32) | static_cast<uint32_t>(lo);
}

// peaks.rs:104-130: append or replace the worst, then bubble toward the
// front while strictly better than the predecessor.
__device__ __forceinline__ void list_insert(PeakList& L, int64_t idx, float v, int n_peaks,
                                            bool use_max, float worst) {
    const int lane = threadIdx.x;
    int last;
    if (L.count < n_peaks) {
        last = L.count;           // new slot
    } else {
        const float wv = lane_value(L.lv, L.count - 1);
        if (!better(v, wv, use_max)) return;
        last = L.count - 1;       // replace the worst
    }
    // Final position: first slot whose entry v is strictly better than.
    const bool beats = (lane < last) && better(v, L.lv, use_max);
    const uint64_t m = __ballot(beats);
    const int pos = m ? __builtin_ctzll(m) : last;
    // Shift [pos, last) right by one.
    const float up_v = __shfl_up(L.lv, 1);
    const int64_t up_i = __shfl_up(L.li, 1);
    if (lane > pos && lane <= last) {
        L.lv = up_v;
        L.li = up_i;
    }
    if (lane == pos) {
        L.lv = v;
        L.li = idx;
    }
    if (L.count < n_peaks) L.count += 1;
    (void)worst;
}

__device__ __forceinline__ float list_threshold(const PeakList& L, int n_peaks, float worst) {
    return L.count < n_peaks ? worst : lane_value(L.lv, L.count - 1);
}

// Scans interior samples [lo, hi) of curve c, where segment `seg` of `n_seg`
// covers lo = 1 + seg*seg_len, hi = min(lo + seg_len, T - 1); one segment
// (n_seg = 1, seg_len >= T) is the whole curve.  Writes n_peaks entries at
// (c*n_seg + seg)*n_peaks.
__global__ __launch_bounds__(kWave) void peaks_kernel(const float* __restrict__ data,
                                                      int64_t n_curves, int64_t T, int n_peaks,
                                                      int64_t min_distance, int use_max_i,
                                                      int64_t* __restrict__ out_idx,
                                                      float* __restrict__ out_val, int n_seg,
                                                      int64_t seg_len) {
    const int64_t c = blockIdx.x / n_seg;
    const int seg = static_cast<int>(blockIdx.x % n_seg);
    if (c >= n_curves) return;
    const int64_t lo = 1 + static_cast<int64_t>(seg) * seg_len;
    const int64_t hi = (T - 1 - lo < seg_len) ? T - 1 : lo + seg_len;  // exclusive
    const int lane = threadIdx.x;
    const bool use_max = use_max_i != 0;
    const float worst = use_max ? -__builtin_inff() : __builtin_inff();
    const int64_t md = min_distance < 1 ? 1 : min_distance;
    const float* __restrict__ row = data + c * T;

    PeakList L;
    L.li = 0;
    L.lv = worst;
    L.count = 0;
    bool has_cand = false;
    int64_t cand_i = -1;
    float cand_v = worst;

    auto promote = [&]() {
        if (has_cand && better(cand_v, list_threshold(L, n_peaks, worst), use_max))
            list_insert(L, cand_i, cand_v, n_peaks, use_max, worst);
        has_cand = false;
    };

    // Interior samples 1..T-2 (endpoints are never peaks, peaks.rs:186-188).
    for (int64_t base = lo; base < hi; base += kWave) {
        const int64_t i = base + lane;
        bool ext = false;
        float v = 0.0f;
        if (i < hi) {
            v = row[i];
            const float l = row[i - 1];
            const float r = row[i + 1];
            ext = use_max ? (v > l && v > r) : (v < l && v < r);
            ext = ext && better(v, list_threshold(L, n_peaks, worst), use_max);
        }
        uint64_t m = __ballot(ext);
        while (m) {
            const int src = __builtin_ctzll(m);
            m &= m - 1;
            const int64_t j = base + src;
            const float vj = lane_value(v, src);
            // Promotion index of a pending candidate precedes j.
            if (has_cand && j - cand_i >= md) promote();
            if (better(vj, list_threshold(L, n_peaks, worst), use_max) &&
                (!has_cand || better(vj, cand_v, use_max))) {
                has_cand = true;
                cand_i = j;
                cand_v = vj;
            }
            // Distance after feeding j (peaks.rs:150-156).
            if (has_cand && j - cand_i + 1 >= md) promote();
        }
    }
    promote();  // finalize (peaks.rs:203-205)

    if (lane < n_peaks) {
        const bool used = lane < L.count;
        const int64_t o = (c * n_seg + seg) * n_peaks + lane;
        out_idx[o] = used ? L.li : -1;
        out_val[o] = used ? L.lv : 0.0f;
    }
}

// a before b in the selection order: better value, then lower index.
__device__ __forceinline__ bool precedes(float av, int64_t ai, float bv, int64_t bi,
                                         bool use_max) {
    return better(av, bv, use_max) || (av == bv && ai < bi);
}

// Best n_peaks of the n_seg segment selections of curve c (entries with
// index -1 are unused), in the selection order, padded with -1 / 0.
__global__ __launch_bounds__(kWave) void peaks_merge_kernel(const int64_t* __restrict__ seg_idx,
                                                            const float* __restrict__ seg_val,
                                                            int64_t n_curves, int n_peaks,
                                                            int n_seg, int use_max_i,
                                                            int64_t* __restrict__ out_idx,
                                                            float* __restrict__ out_val) {
    const int64_t c = blockIdx.x;
    if (c >= n_curves) return;
    const int lane = threadIdx.x;
    const bool use_max = use_max_i != 0;
    const int64_t total = static_cast<int64_t>(n_seg) * n_peaks;
    const int64_t* __restrict__ ci_all = seg_idx + c * total;
    const float* __restrict__ cv_all = seg_val + c * total;
    int64_t li = -1;  // this lane's list entry
    float lv = 0.0f;
    int count = 0;
    for (int64_t base = 0; base < total; base += kWave) {
        const int64_t k = base + lane;
        const int64_t ci = k < total ? ci_all[k] : -1;
        const float cv = k < total ? cv_all[k] : 0.0f;
        // Filter against the worst entry at the chunk start (it only improves).
        bool pass = ci >= 0;
        if (count == n_peaks) {
            const float wv = lane_value(lv, n_peaks - 1);
            const int64_t wi = lane_value64(li, n_peaks - 1);
            pass = pass && precedes(cv, ci, wv, wi, use_max);
        }
        uint64_t m = __ballot(pass);
        while (m) {
            const int src = __builtin_ctzll(m);
            m &= m - 1;
            const float v = lane_value(cv, src);
            const int64_t i = lane_value64(ci, src);
            int last;
            if (count < n_peaks) {
                last = count;
            } else {
                if (!precedes(v, i, lane_value(lv, n_peaks - 1), lane_value64(li, n_peaks - 1),
                              use_max))
                    continue;
                last = n_peaks - 1;
            }
            const bool beats = lane < last && precedes(v, i, lv, li, use_max);
            const uint64_t bm = __ballot(beats);
            const int pos = bm ? __builtin_ctzll(bm) : last;
            const float up_v = __shfl_up(lv, 1);
            const int64_t up_i = __shfl_up(li, 1);
            if (lane > pos && lane <= last) {
                lv = up_v;
                li = up_i;
            }
            if (lane == pos) {
                lv = v;
                li = i;
            }
            if (count < n_peaks) count += 1;
        }
    }
    if (lane < n_peaks) {
        const bool used = lane < count;
        out_idx[c * n_peaks + lane] = used ? li : -1;
        out_val[c * n_peaks + lane] = used ? lv : 0.0f;
    }
}

}  // namespace

hipError_t launch_peaks(const float* data, int64_t n_curves, int64_t T, int64_t n_peaks,
                        int64_t min_distance, int use_max, int64_t* idx, float* val,
                        hipStream_t s, const PeaksWorkspace* ws) {
    if (n_curves <= 0 || n_peaks <= 0) return hipSuccess;
    if (n_peaks > kWave) return hipErrorInvalidValue;
    const int n_seg = ws ? peaks_segments(n_curves, T, min_distance) : 1;
    if (n_seg > 1) {
        const int64_t interior = T - 2;
        const int64_t seg_len = (interior + n_seg - 1) / n_seg;
        hipLaunchKernelGGL(peaks_kernel, dim3(static_cast<uint32_t>(n_curves * n_seg)),
                           dim3(kWave), 0, s, data, n_curves, T, static_cast<int>(n_peaks),
                           min_distance, use_max, ws->idx, ws->val, n_seg, seg_len);
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
        hipLaunchKernelGGL(peaks_merge_kernel, dim3(static_cast<uint32_t>(n_curves)),
                           dim3(kWave), 0, s, ws->idx, ws->val, n_curves,
                           static_cast<int>(n_peaks), n_seg, use_max, idx, val);
        return hipGetLastError();
    }
    hipLaunchKernelGGL(peaks_kernel, dim3(static_cast<uint32_t>(n_curves)), dim3(kWave), 0, s,
                       data, n_curves, T, static_cast<int>(n_peaks), min_distance, use_max, idx,
                       val, 1, T);
    return hipGetLastError();
}

int peaks_segments(int64_t n_curves, int64_t T, int64_t min_distance) {
    if (min_distance > 1 || n_curves <= 0 || T < 3) return 1;
    const int64_t by_waves = (kPeaksSplitWaves + n_curves - 1) / n_curves;
    const int64_t by_len = (T - 2) / kPeaksMinSegment;
    int64_t n = by_waves < by_len ? by_waves : by_len;
    if (n > kPeaksMaxSegments) n = kPeaksMaxSegments;
    return n < 2 ? 1 : static_cast<int>(n);
}

}  // namespace pf
