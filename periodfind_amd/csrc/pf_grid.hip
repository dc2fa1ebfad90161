// Periodogram grid kernels for gfx950: CE, AOV, LS, FPW, BLS.
//
// Launch shape (all kernels): one wave64 workgroup per (curve, block of 64
// consecutive trials); lane = trial.  Trials are the flattened
// [period][pdt] index, the reference's output order (rust/src/ce.rs:37-39).
// The workgroup id is remapped so blocks of one curve share an XCD.
//
// Every kernel reads, per point, a wave-uniform record (t as f64, optional
// t*t, and one or two f32 aux values prepared once per curve set) and
// accumulates into lane-private state: LDS bins for the binned statistics
// (updated by lds_rmw, no atomics), registers for LS.  Accumulation is serial
// in point order per trial, as in the Rust reference, so f32 bin sums are
// bit-identical to it.
//
// Template parameters: HAS_PDT (any nonzero period_dt: the fold subtracts
// k*t^2), EXT1 (CE/AOV bin extents of 1), G (points per lds_rmw group, a
// tuning knob: 1, 2 or 4; for CE, G = 0 selects non-returning LDS u32
// atomics, which the hardware executes fast, unlike its f32 LDS atomics).
#include <type_traits>

#include "pf_device.hpp"
#include "pf_kernels.hpp"

namespace pf {

// Resolve (curve, trial) of this lane.  Lanes past the last trial compute
// the last trial again and do not store.
struct LaneTrial {
    int64_t curve;
    int64_t trial;
    bool active;
};

__device__ __forceinline__ LaneTrial lane_trial(uint64_t n_blocks, int64_t tblocks, int64_t T,
                                                int lane = static_cast<int>(threadIdx.x)) {
    uint64_t L = xcd_remap(blockIdx.x, n_blocks);
    LaneTrial lt;
    lt.curve = static_cast<int64_t>(L / static_cast<uint64_t>(tblocks));
    int64_t tb = static_cast<int64_t>(L % static_cast<uint64_t>(tblocks));
    int64_t tr = tb * kWave + lane;
    lt.active = tr < T;
    lt.trial = lt.active ? tr : T - 1;
    return lt;
}

// Persistent-launch work queue, one part per XCD.  Ranks order curves
// longest first (so the longest items start first and the tail is short);
// rank r belongs to part r % 8, so each part keeps that order and gets an
// even share.  A wave claims from the part of its own XCD (gfx950 deals
// workgroup i to XCD i % 8) and, once that is empty, from the others in
// turn.  The waves of one XCD thus work on a few curves at a time, whose
// points stay in that XCD's L2: with one queue for all, consecutive claims
// went to every XCD and the resident waves spread over hundreds of curves
// (CE on 10,000 ragged curves fetched 152 GiB from HBM, against 0.7 GiB for
// one plain launch).  Within a part, items are numbered virtually: the
// parts' items are laid end to end, part p's rank j being virtual rank
// base(p) + j, item = virtual rank * tblocks + trial block, so a claim is a
// contiguous range (item_rank maps it back to the rank).  Claims are guided:
// a wave takes remaining / (2 * waves per part) items (at least 1, at most
// the claim bound of the bucket it is in), estimated from the counter value
// it saw last, so claims are large while the part is full and single items
// at the end, with few atomics on each counter.  The state is wave-uniform.
constexpr int kQueueParts = kQueuePartsHost;

struct Claim {
    int64_t begin, end;
};

// A wave-uniform 64-bit value, read from the wave's first active lane into
// scalar registers.  What is derived from a claim (curve, point offsets) is
// then visibly uniform, and the point stream takes scalar loads: a value
// passed through a lane shuffle or LDS stays a vector register, and every
// load indexed by it becomes a 64-lane vector load of one address.
__device__ __forceinline__ int64_t wave_uniform(int64_t v) {
    const uint64_t u = static_cast<uint64_t>(v);
    const uint32_t lo = __builtin_amdgcn_readfirstlane(static_cast<uint32_t>(u));
    const uint32_t hi = __builtin_amdgcn_readfirstlane(static_cast<uint32_t>(u >> 32));
    return static_cast<int64_t>((static_cast<uint64_t>(hi) << 32) | lo);
}

struct QueueState {
    int part;      // the part being drained
    int64_t seen;  // its items known to be claimed
    int tried;     // parts found empty in a row
};

__device__ __forceinline__ QueueState queue_start() {
    return QueueState{static_cast<int>(blockIdx.x % kQueueParts), 0, 0};
}

// Ranks of the parts: the first nr % 8 parts hold one more.
struct QueueParts {
    int64_t q, r;  // nr = 8 q + r
    __device__ int64_t ranks(int p) const { return q + (p < r ? 1 : 0); }
    __device__ int64_t base(int p) const { return p * q + (p < r ? p : r); }
    // virtual rank -> rank
    __device__ int64_t rank(int64_t v) const {
        const int64_t big = r * (q + 1);
        const int64_t p = v < big ? v / (q + 1) : r + (v - big) / (q > 0 ? q : 1);
        return (v - base(static_cast<int>(p))) * kQueueParts + p;
    }
};

__device__ __forceinline__ QueueParts queue_parts(const WorkQueue& q, const GridArgs& g) {
    const int64_t nr = q.total / g.tblocks;
    return QueueParts{nr / kQueueParts, nr % kQueueParts};
}

// The rank of a (virtual) queue item.
__device__ __forceinline__ int64_t item_rank(const WorkQueue& q, const GridArgs& g,
                                             int64_t item) {
    return queue_parts(q, g).rank(item / g.tblocks);
}

// Binning variant and claim bound of a rank's bucket (buckets are contiguous
// rank ranges; wave-uniform).  Constant indices only, so the queue stays in
// kernel-argument registers.
__device__ __forceinline__ int rank_group(const WorkQueue& q, int64_t rank) {
    int v = q.bucket_group[0];
#pragma unroll
    for (int i = 1; i < kQueueBuckets; ++i)
        if (i < q.n_buckets && rank >= q.bucket_end[i - 1]) v = q.bucket_group[i];
    return v;
}

__device__ __forceinline__ int rank_claim(const WorkQueue& q, int64_t rank) {
    int v = q.bucket_claim[0];
#pragma unroll
    for (int i = 1; i < kQueueBuckets; ++i)
        if (i < q.n_buckets && rank >= q.bucket_end[i - 1]) v = q.bucket_claim[i];
    return v;
}

// The next claim, or {total, total} once every part is empty (each wave
// then has found all 8 parts empty: every item was claimed exactly once, by
// the atomics).
__device__ __forceinline__ Claim next_claim(const WorkQueue& q, const GridArgs& g,
                                            QueueState& st) {
    const QueueParts qp = queue_parts(q, g);
    const int64_t waves = gridDim.x / kQueueParts > 0 ? gridDim.x / kQueueParts : 1;
    while (st.tried < kQueueParts) {
        const int p = st.part;
        const int64_t total = qp.ranks(p) * g.tblocks;
        if (st.seen < total) {
            const int64_t vbase = qp.base(p) * g.tblocks;
            const int64_t bound = rank_claim(q, qp.rank((vbase + st.seen) / g.tblocks));
            int64_t c = (total - st.seen) / (2 * waves);
            c = c < 1 ? 1 : (c > bound ? bound : c);
            // Lane 0 claims (every caller reaches here with the whole wave
            // active, so the first active lane is lane 0).
            unsigned long long v = 0;
            if (threadIdx.x == 0) v = atomicAdd(q.counter + p, static_cast<unsigned long long>(c));
            const int64_t b = wave_uniform(static_cast<int64_t>(v));
            st.seen = b + c;
            if (b < total) {
                st.tried = 0;
                return Claim{vbase + b, vbase + (b + c < total ? b + c : total)};
            }
        }
        st.part = (p + 1) % kQueueParts;
        st.seen = 0;
        st.tried += 1;
    }
    return Claim{q.total, q.total};
}

__device__ __forceinline__ LaneTrial queue_trial(const WorkQueue& q, const GridArgs& g,
                                                 int64_t item,
                                                 int lane = static_cast<int>(threadIdx.x)) {
    LaneTrial lt;
    const int64_t vr = item / g.tblocks;
    const int64_t tb = item - vr * g.tblocks;
    lt.curve = q.order[queue_parts(q, g).rank(vr)];
    const int64_t tr = tb * kWave + lane;
    lt.active = tr < g.T;
    lt.trial = lt.active ? tr : g.T - 1;
    return lt;
}

// Wrap `k + o` into [0, n) for k <= n and small o (rust `% n`, ce.rs:57-58).
__device__ __forceinline__ uint32_t wrap_add(uint32_t k, uint32_t o, uint32_t n) {
    uint32_t v = k + o;
    while (v >= n) v -= n;
    return v;
}

// s_waitcnt lgkmcnt(0) (vmcnt/expcnt left at their maxima), as a real
// instruction the compiler's wait-count tracking sees, fenced so the
// scheduler does not move loads across it.
__device__ __forceinline__ void wait_scalar_loads() {
    __builtin_amdgcn_sched_barrier(0);
    __builtin_amdgcn_s_waitcnt(0xC07F);
    __builtin_amdgcn_sched_barrier(0);
}

// Number of points in a body group (the extent of its phase array).
template <class A>
constexpr int group_size_v = static_cast<int>(std::extent<std::remove_reference_t<A>>::value);

// A value the register allocator must keep rather than recompute: an LDS
// offset used by a read and, a group later, by the write back (MF: the
// allocator otherwise re-derives it from the bin, 200 instead of 186 VALU
// per 16 points; elsewhere keeping it measured no better).  An empty asm
// statement with the value as an in/out operand; no instruction.
__device__ __forceinline__ uint32_t keep_reg(uint32_t v) {
    asm("" : "+v"(v));
    return v;
}

// The scans hand their bodies phases (NoBins) or, given FastBins, the
// points' final bins (fold_bins: fast where certain, else the exact rule).
struct NoBins {};
template <class B>
constexpr bool is_fast_bins_v = !std::is_same<B, NoBins>::value;

// Walk the curve's points in order.  Points are loaded in chunks of C: the
// wave-uniform loads of chunk i+1 (times and the statistic's aux record,
// scalar loads) are issued before chunk i is folded and accumulated, so the
// single lgkmcnt wait per chunk (scalar loads and LDS traffic share that
// counter) falls after a whole chunk of VALU work.  The prefetch index is
// clamped to the last full chunk, so it never reads past the curve.
// `body(ph[G], rec[G])` receives the folded phases of G consecutive points
// (G inside full chunks, 1 for the remainder); with FastBins, `body(k[G],
// rec[G])` receives their bins instead, a whole chunk's computed together.
// SCALAR = false reads the times with vector loads (the caller's `load`
// must too): they count on vmcnt, so no wait on the point stream also waits
// for the chunk's LDS operations.
template <bool HAS_PDT, int G, bool EXACT = true, bool SCALAR = true, class Load, class Body,
          class Bins = NoBins>
__device__ __forceinline__ void scan_points(const double* __restrict__ t64,
                                            const double* __restrict__ tt, int64_t n,
                                            const FoldTrial& f, Load load, Body body,
                                            const Bins& bins = Bins{}) {
    constexpr bool FASTB = is_fast_bins_v<Bins>;
    using Rec = decltype(load(int64_t(0)));
    constexpr int C = HAS_PDT ? kChunk / 2 : kChunk;
    static_assert(C % G == 0, "chunk must hold whole groups");
    struct Chunk {
        double t[C];
        double t2[C];
        Rec r[C];
    };
    auto fetch = [&](int64_t base, Chunk& ch) {
#pragma unroll
        for (int j = 0; j < C; ++j) {
            ch.t[j] = SCALAR ? uniform_load(t64 + base + j) : t64[base + j];
            ch.t2[j] = HAS_PDT ? (SCALAR ? uniform_load(tt + base + j) : tt[base + j]) : 0.0;
            ch.r[j] = load(base + j);
        }
    };
    using Ph = std::conditional_t<FASTB, uint32_t, float>;
    auto run = [&](const Chunk& ch) {
        Ph phc[C];
        if constexpr (FASTB) fold_bins<HAS_PDT, C>(ch.t, ch.t2, f, bins.rule, bins.exact, phc);
        else fold_phases<HAS_PDT, C, EXACT>(ch.t, ch.t2, f, phc);
#pragma unroll
        for (int gi = 0; gi < C; gi += G) {
            Ph ph[G];
            Rec rc[G];
#pragma unroll
            for (int j = 0; j < G; ++j) {
                ph[j] = phc[gi + j];
                rc[j] = ch.r[gi + j];
            }
            body(ph, rc);
        }
    };
    // Ping-pong over two buffers.  Scalar loads return out of order, so any
    // wait on them is lgkmcnt(0); the explicit wait sits before each
    // prefetch, where the only outstanding scalar load is the buffer issued
    // one whole chunk earlier.  The scheduling barriers keep that order (left
    // alone, the scheduler sinks loads below the consuming chunk to save
    // SGPRs, which exposes their latency).
    int64_t i = 0;
    if (n >= 2 * C) {
        const int64_t last = n - C;
        Chunk a, b;
        fetch(0, a);
        for (; i + 2 * C <= n; i += 2 * C) {
            if (SCALAR) wait_scalar_loads();
            fetch(i + C, b);
            __builtin_amdgcn_sched_barrier(0);
            run(a);
            if (SCALAR) wait_scalar_loads();
            const int64_t na = (i + 2 * C <= last) ? i + 2 * C : last;
            fetch(na, a);
            __builtin_amdgcn_sched_barrier(0);
            run(b);
        }
    }
    for (; i < n; ++i) {
        const float p1 = fold_phase<HAS_PDT, EXACT>(t64[i], HAS_PDT ? tt[i] : 0.0, f);
        Ph ph[1];
        if constexpr (FASTB) ph[0] = bins.exact(p1);
        else ph[0] = p1;
        Rec rc[1] = {load(i)};
        body(ph, rc);
    }
}

// scan_points for the statistics that update lane-private LDS bins by
// read-modify-write (AOV, FPW, MF, BLS), software-pipelined: the reads of group
// g are issued, then group g+1 is folded and binned (independent VALU work,
// mostly f64), then group g is finished (added, written).  The scheduling
// barriers keep that order, so the fold covers the LDS read latency instead
// of each group waiting on its own reads.  `prep(ph[K], rec[K], off[K],
// val[A][K])` maps a group's phases and records to bin offsets and addends;
// results are identical to updating one group after another.  CH points are
// prefetched per chunk (a tuning knob: results do not depend on it).  SCALAR
// selects how the times are read (the caller's `load` reads the records the
// same way): scalar loads deliver a point once per wave, into SGPRs, but
// count on lgkmcnt like the LDS reads, so a wait on an LDS read also waits
// for the next chunk's prefetch; vector loads count on vmcnt instead, at the
// cost of one copy of the point per lane.  Measured at 8192 points: scalar
// is 3-5% faster for AOV and FPW, vector 28% faster for BLS (16-point
// chunks, one LDS wait per point).  With FastBins, `prep` receives the
// group's bins (k[K]) in place of its phases.
template <bool HAS_PDT, int G, int A, int CH = kChunk, bool SCALAR = true, class Load, class Prep,
          class Bins = NoBins>
__device__ __forceinline__ void scan_points_rmw(const double* __restrict__ t64,
                                                const double* __restrict__ tt, int64_t n,
                                                const FoldTrial& f, Load load, float* col,
                                                Prep prep, const Bins& bins = Bins{}) {
    constexpr bool FASTB = is_fast_bins_v<Bins>;
    using Ph = std::conditional_t<FASTB, uint32_t, float>;
    using Rec = decltype(load(int64_t(0)));
    constexpr int C = HAS_PDT ? CH / 2 : CH;
    constexpr int NG = C / G;
    static_assert(C % G == 0, "chunk must hold whole groups");
    struct Chunk {
        double t[C];
        double t2[C];
        Rec r[C];
    };
    struct Group {
        uint32_t off[G];
        float val[A][G];
    };
    auto fetch = [&](int64_t base, Chunk& ch) {
#pragma unroll
        for (int j = 0; j < C; ++j) {
            ch.t[j] = SCALAR ? uniform_load(t64 + base + j) : t64[base + j];
            ch.t2[j] = HAS_PDT ? (SCALAR ? uniform_load(tt + base + j) : tt[base + j]) : 0.0;
            ch.r[j] = load(base + j);
        }
    };
    // A group's points (times and records) out of the chunk.
    struct Pts {
        double t[G], t2[G];
        Rec rc[G];
    };
    auto pts_of = [&](const Chunk& ch, int gi) {
        Pts p;
#pragma unroll
        for (int j = 0; j < G; ++j) {
            p.t[j] = ch.t[gi + j];
            p.t2[j] = ch.t2[gi + j];
            p.rc[j] = ch.r[gi + j];
        }
        return p;
    };
    auto prep_group = [&](const Chunk& ch, int gi, Group& gr) {
        const Pts p = pts_of(ch, gi);
        Ph ph[G];
        if constexpr (FASTB) fold_bins<HAS_PDT, G>(p.t, p.t2, f, bins.rule, bins.exact, ph);
        else fold_phases<HAS_PDT, G>(p.t, p.t2, f, ph);
        prep(ph, p.rc, gr.off, gr.val);
    };
    // Fast bins: group g+1's bins are computed while group g's reads are in
    // flight, and its exact-path branch (which waits on the compares) comes
    // after group g is written back.
    struct Pending {
        uint32_t k[G];
        bool slow[G];
    };
    auto run = [&](const Chunk& ch) {
        Group cur, nxt;
        prep_group(ch, 0, cur);
#pragma unroll
        for (int gi = 0; gi < NG; ++gi) {
            float r[A][G];
            lds_read_group(col, cur.off, r);
            __builtin_amdgcn_sched_barrier(0);
            if constexpr (FASTB) {
                Pending pn;
                Pts p;
                if (gi + 1 < NG) {
                    p = pts_of(ch, (gi + 1) * G);
                    fold_bins_start<HAS_PDT, G>(p.t, p.t2, f, bins.rule, pn.k, pn.slow);
                }
                __builtin_amdgcn_sched_barrier(0);
                lds_finish_group(col, cur.off, cur.val, r);
                if (gi + 1 < NG) {
                    fold_bins_fix<HAS_PDT, G>(p.t, p.t2, f, bins.exact, pn.slow, pn.k);
                    prep(pn.k, p.rc, nxt.off, nxt.val);
                    cur = nxt;
                }
            } else {
                if (gi + 1 < NG) prep_group(ch, (gi + 1) * G, nxt);
                __builtin_amdgcn_sched_barrier(0);
                lds_finish_group(col, cur.off, cur.val, r);
                if (gi + 1 < NG) cur = nxt;
            }
        }
    };
    // Ping-pong chunk prefetch, as in scan_points (vector loads need no
    // explicit wait: the compiler waits on vmcnt for the chunk in use).
    int64_t i = 0;
    if (n >= 2 * C) {
        const int64_t last = n - C;
        Chunk a, b;
        fetch(0, a);
        for (; i + 2 * C <= n; i += 2 * C) {
            if (SCALAR) wait_scalar_loads();
            fetch(i + C, b);
            __builtin_amdgcn_sched_barrier(0);
            run(a);
            if (SCALAR) wait_scalar_loads();
            const int64_t na = (i + 2 * C <= last) ? i + 2 * C : last;
            fetch(na, a);
            __builtin_amdgcn_sched_barrier(0);
            run(b);
        }
    }
    for (; i < n; ++i) {
        const float p1 = fold_phase<HAS_PDT>(t64[i], HAS_PDT ? tt[i] : 0.0, f);
        Ph ph[1];
        if constexpr (FASTB) ph[0] = bins.exact(p1);
        else ph[0] = p1;
        Rec rc[1] = {load(i)};
        uint32_t off[1];
        float val[A][1];
        prep(ph, rc, off, val);
        lds_rmw(col, off, val);
    }
}

struct F2 {
    float a, b;
};

// How AOV, FPW and MF read their point stream in scan_points_rmw: by vector
// loads.  With a scalar prefetch in flight, every wait for a group's LDS
// reads is lgkmcnt(0), which also waits for that prefetch (scalar loads
// return out of order and share the counter); rocprofv3 showed these waves
// waiting 52-55% of their time.  Vector loads count on vmcnt instead.
constexpr bool kRmwScalarStream = false;

template <typename T>
__device__ __forceinline__ T rmw_load(const T* p) {
    if constexpr (kRmwScalarStream) return uniform_load(p);
    else return *p;
}

// ---------------------------------------------------------------------------
// Conditional Entropy — rust/src/ce.rs:13-91.
// LDS: u32 hist[n_phase+1][n_mag][64].  Row n_phase catches phase == 1.0
// (bin n) and is folded into row 0 after the scan (counts commute), so the
// hot loop needs no wrap.  mrow[i] is the wave-uniform byte offset
// (mag bin % n_mag) * 256 of the point's magnitude column.
// ---------------------------------------------------------------------------
// The (phase, magnitude) count histogram of one trial per lane, as CE builds
// it (ce.rs:44-63); `epi(count, cell_count, ncell, n, col)` then turns it into
// the lane's statistic.  count(j, i) is the count of phase row j, magnitude
// column i (the spare row folded into row 0).
template <bool HAS_PDT, bool EXT1, int G, class Epi>
__device__ __forceinline__ void hist_block(const GridArgs& g, const CeArgs& a, const LaneTrial lt,
                                           Epi epi) {
    extern __shared__ uint32_t hist[];
    // G < 0: packed atomics, cells c and c^1 share a word (16-bit counts).
    constexpr bool PACK = EXT1 && G < 0;
    const int lane = threadIdx.x;
    const int nrows = EXT1 ? a.n_phase + 1 : a.n_phase;
    const int ncell = nrows * a.n_mag;
    const int nword = PACK ? (ncell + 1) / 2 : ncell;
    for (int b = 0; b < nword; ++b) hist[b * kWave + lane] = 0u;

    const FoldTrial f = make_fold_trial(g.periods[lt.trial / g.nd], g.pdts[lt.trial % g.nd]);
    const int64_t p0 = g.off[lt.curve];
    const int64_t n = g.off[lt.curve + 1] - p0;
    const uint32_t* __restrict__ mrow = a.mrow + p0;
    const DivBin db = make_div_bin(a.n_phase);
    const uint32_t np_ = static_cast<uint32_t>(a.n_phase);
    const uint32_t nm = static_cast<uint32_t>(a.n_mag);
    uint32_t* __restrict__ col = hist + lane;
    const uint32_t row_bytes = nm * kPlaneBytes;
    const bool nm_even = (nm & 1u) == 0;
    const uint32_t half_row_bytes = (nm / 2) * kPlaneBytes;

    // The atomic variant takes whole chunks (the histogram is order-free), so
    // the bins of a chunk are computed stage by stage across its points.
    // The point stream comes by scalar loads.  (Vector loads, whose waits do
    // not also wait for the chunk's LDS atomics, measured 6% slower at 8192
    // points: 102 instead of 54 VGPRs and a scalar move of each magnitude
    // row.)
    constexpr int kCeChunk = HAS_PDT ? kChunk / 2 : kChunk;
    // The point body for a compile-time choice of the packed-even layout
    // (kept out of the loop: one straight-line loop per variant).
    auto body_for = [&](auto even_tag) {
        constexpr bool EVEN = decltype(even_tag)::value;
        // ph: the points' phase bins with EXT1 (scan with fast bins), else
        // their phases.
        return [&](const auto& ph, const auto& mb) {
            constexpr int K = group_size_v<decltype(ph)>;
            if constexpr (PACK) {
                const auto& kb = ph;
                if constexpr (EVEN) {
                    // Cell kb*nm + m: word kb*(nm/2) + m/2, half m & 1.  The
                    // magnitude part (word offset, increment) comes with the
                    // point: mb[j] is its CeArgs::mpk record.
#pragma unroll
                    for (int j = 0; j < K; ++j) {
                        const uint32_t word = static_cast<uint32_t>(mb[j]);
                        const uint32_t inc = static_cast<uint32_t>(mb[j] >> 32);
                        char* cell = reinterpret_cast<char*>(col) +
                                     __umul24(kb[j], half_row_bytes) + word;
                        __hip_atomic_fetch_add(reinterpret_cast<uint32_t*>(cell), inc,
                                               __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
                    }
                } else {
#pragma unroll
                    for (int j = 0; j < K; ++j) {
                        const uint32_t c = __umul24(kb[j], nm) + mb[j] / kPlaneBytes;
                        uint32_t* cell = col + (c >> 1) * kWave;
                        __hip_atomic_fetch_add(cell, 1u << ((c & 1u) << 4), __ATOMIC_RELAXED,
                                               __HIP_MEMORY_SCOPE_WAVEFRONT);
                    }
                }
            } else if constexpr (EXT1 && G == 0) {
                const auto& kb = ph;
#pragma unroll
                for (int j = 0; j < K; ++j) {
                    uint32_t* cell = reinterpret_cast<uint32_t*>(
                        reinterpret_cast<char*>(col) + __umul24(kb[j], row_bytes) + mb[j]);
                    __hip_atomic_fetch_add(cell, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
                }
            } else if constexpr (EXT1) {
                uint32_t off[K];
                uint32_t one[1][K];
#pragma unroll
                for (int j = 0; j < K; ++j) {
                    off[j] = __umul24(ph[j], row_bytes) + mb[j];
                    one[0][j] = 1u;
                }
                lds_rmw(col, off, one);
            } else {
#pragma unroll
                for (int j = 0; j < K; ++j) {
                    const uint32_t k = div_bin(ph[j], db);
                    const uint32_t mbin = mb[j] / kPlaneBytes;
                    for (int po = 0; po < a.phase_ext; ++po) {
                        const uint32_t pb = wrap_add(k, po, np_);
                        for (int mo = 0; mo < a.mag_ext; ++mo) {
                            const uint32_t m2 = wrap_add(mbin, mo, nm);
                            const uint32_t off1[1] = {(pb * nm + m2) * kPlaneBytes};
                            const uint32_t one1[1][1] = {{1u}};
                            lds_rmw(col, off1, one1);
                        }
                    }
                }
            }
        };
    };
    // With extents of 1, bins: bin n (phase 1.0) is kept, for the spare row.
    auto scan = [&](auto body, auto load) {
        if constexpr (EXT1)
            scan_points<HAS_PDT, (G > 0 ? G : kCeChunk)>(
                g.t64 + p0, g.tt + p0, n, f, load, body,
                fast_bins(lane_bin_rule(g.tmax, lt.curve, f, a.n_phase, true),
                          [&](float ph) { return div_bin(ph, db); }));
        else
            scan_points<HAS_PDT, 1>(g.t64 + p0, g.tt + p0, n, f, load, body);
    };
    auto rows = [&](int64_t i) { return uniform_load(mrow + i); };
    if constexpr (PACK) {
        // (read as one 64-bit word: x in the low half)
        const uint64_t* __restrict__ mpk =
            a.mpk ? reinterpret_cast<const uint64_t*>(a.mpk + p0) : nullptr;
        if (nm_even && mpk)
            scan(body_for(std::true_type{}), [&](int64_t i) { return uniform_load(mpk + i); });
        else
            scan(body_for(std::false_type{}), rows);
    } else {
        scan(body_for(std::false_type{}), rows);
    }
    // Count of cell (row j, mag i); with EXT1 the spare row n_phase (phase
    // 1.0) belongs to row 0 (ce.rs:57 wraps it there).
    auto cell_count = [&](int c) -> uint32_t {
        if constexpr (PACK) return (col[(c >> 1) * kWave] >> ((c & 1) << 4)) & 0xffffu;
        else return col[c * kWave];
    };
    auto count = [&](int j, int i) -> uint32_t {
        uint32_t v = cell_count(j * a.n_mag + i);
        if (EXT1 && j == 0) v += cell_count(a.n_phase * a.n_mag + i);
        return v;
    };
    epi(count, cell_count, ncell, n, col);
}

// Largest CE normaliser (points x extents) whose probabilities take the
// division-free quotient.
constexpr float kCeFastDivMax = 131072.0f;
// Cells of a phase row taken together in the CE epilogue, so that their
// count and table reads are in flight at once.
constexpr int kCeCells = 5;

template <bool HAS_PDT, bool EXT1, int G>
__device__ __forceinline__ void ce_block(const GridArgs& g, const CeArgs& a, const LaneTrial lt) {
    constexpr bool PACK = EXT1 && G < 0;
    // The epilogue's logf table (1/c, log c) in LDS: the subinterval varies
    // from lane to lane, and an LDS read is the shortest wait for it.
    __shared__ double2 logtab[16];
    if (threadIdx.x < 16)
        logtab[threadIdx.x] = make_double2(logf_glibc_data::kTab[threadIdx.x][0],
                                           logf_glibc_data::kTab[threadIdx.x][1]);
    __syncthreads();
    hist_block<HAS_PDT, EXT1, G>(g, a, lt, [&](auto count, auto cell_count, int ncell, int64_t n,
                                                 uint32_t*) {
    // ce.rs:66-86: p = count / (N*ext_p*ext_m); CE = sum_j sum_i p_ij ln(p_j/p_ij).
    const float norm = static_cast<float>(static_cast<uint64_t>(n) *
                                          static_cast<uint64_t>(a.phase_ext) *
                                          static_cast<uint64_t>(a.mag_ext));
    // Packed counts of a 65536-point curve overflow only if all points share
    // one cell (see kCePackMaxLen); then p = 1 there and CE = 1 * ln(1/1) = 0.
    bool one_cell = false;
    if constexpr (PACK) {
        if (n > 65535) {
            uint32_t total = 0;
            for (int c = 0; c < ncell; ++c) total += cell_count(c);
            one_cell = total != static_cast<uint32_t>(n);
        }
    }
    // p = fl(count / norm): for norm <= kCeFastDivMax by one reciprocal
    // multiply and one FMA correction, which gives the IEEE quotient for every
    // integer count <= norm (tests/exactness/check_count_div.c checks all of
    // them); above that, the division itself.
    const float rn = 1.0f / norm;
    const int nm = a.n_mag;
    auto epilogue = [&](auto fast_tag) {
        constexpr bool FAST = decltype(fast_tag)::value;
        auto prob = [&](uint32_t c) -> float {
            const float cf = static_cast<float>(c);
            if constexpr (FAST) {
                const float q0 = cf * rn;
                return __builtin_fmaf(__builtin_fmaf(-q0, norm, cf), rn, q0);
            } else {
                return cf / norm;
            }
        };
        // Cells i0 .. i0 + kCeCells - 1 of row j; past the row's end the
        // last cell is read again and `valid` is false.
        auto cells = [&](int j, int i0, float (&p)[kCeCells], bool (&valid)[kCeCells]) {
            uint32_t c[kCeCells];
#pragma unroll
            for (int u = 0; u < kCeCells; ++u) c[u] = count(j, min(i0 + u, nm - 1));
#pragma unroll
            for (int u = 0; u < kCeCells; ++u) {
                p[u] = prob(c[u]);
                valid[u] = i0 + u < nm;
            }
        };
        // Every p is >= +0, so p_j and the sum never hold -0 and adding +0
        // in place of a skipped cell leaves them unchanged.
        float ce = 0.0f;
        for (int j = 0; j < a.n_phase && !one_cell; ++j) {
            float pj = 0.0f;
            for (int i0 = 0; i0 < nm; i0 += kCeCells) {
                float p[kCeCells];
                bool valid[kCeCells];
                cells(j, i0, p, valid);
#pragma unroll
                for (int u = 0; u < kCeCells; ++u) pj += valid[u] ? p[u] : 0.0f;
            }
            for (int i0 = 0; i0 < nm; i0 += kCeCells) {
                float pij[kCeCells];
                bool valid[kCeCells];
                cells(j, i0, pij, valid);
                // p_j >= p_ij (a rounded sum of non-negative terms never falls
                // below one of them), so p_j / p_ij is a positive normal float
                // wherever p_ij != 0: logf's special cases cannot arise.
                LogfSplit sp[kCeCells];
                double2 tab[kCeCells];
#pragma unroll
                for (int u = 0; u < kCeCells; ++u) {
                    const float x = pij[u] != 0.0f ? pj / pij[u] : 1.0f;
                    sp[u] = logf_glibc_split(__builtin_bit_cast(uint32_t, x));
                    tab[u] = logtab[sp[u].i];
                }
#pragma unroll
                for (int u = 0; u < kCeCells; ++u) {
                    const float term = pij[u] * logf_glibc_eval(sp[u], tab[u].x, tab[u].y);
                    ce += valid[u] && pij[u] != 0.0f ? term : 0.0f;
                }
            }
        }
        return ce;
    };
    const float ce = norm <= kCeFastDivMax ? epilogue(std::true_type{})
                                           : epilogue(std::false_type{});
    if (lt.active) g.out[lt.curve * g.T + lt.trial] = ce;
    });
}

// One workgroup per (curve, 64-trial block), hardware-scheduled.
template <bool HAS_PDT, bool EXT1, int G>
__global__ __launch_bounds__(kWave) void ce_kernel(GridArgs g, CeArgs a) {
    ce_block<HAS_PDT, EXT1, G>(g, a, lane_trial(g.n_blocks, g.tblocks, g.T));
}

// Persistent: each wave pulls (curve, 64-trial block) items from the queue.
// PACKED: every bucket takes the packed atomic variant, the only code path
// compiled in (78 instead of 83 VGPRs with period derivatives, and about a
// third of the code).
template <bool HAS_PDT, bool EXT1, bool PACKED = false>
__global__ __launch_bounds__(kWave) void ce_persistent(GridArgs g, CeArgs a, WorkQueue q) {
    QueueState seen = queue_start();
    for (Claim cl = next_claim(q, g, seen); cl.begin < q.total; cl = next_claim(q, g, seen))
        for (int64_t it = cl.begin; it < cl.end; ++it) {
            const LaneTrial lt = queue_trial(q, g, it);
            if (!EXT1) {
                ce_block<HAS_PDT, false, 1>(g, a, lt);
                continue;
            }
            if constexpr (PACKED) {
                ce_block<HAS_PDT, EXT1, -1>(g, a, lt);
                continue;
            }
            switch (rank_group(q, item_rank(q, g, it))) {
                case -1: ce_block<HAS_PDT, EXT1, -1>(g, a, lt); break;
                case 0: ce_block<HAS_PDT, EXT1, 0>(g, a, lt); break;
                case 2: ce_block<HAS_PDT, EXT1, 2>(g, a, lt); break;
                case 4: ce_block<HAS_PDT, EXT1, 4>(g, a, lt); break;
                default: ce_block<HAS_PDT, EXT1, 1>(g, a, lt); break;
            }
        }
}

// ---------------------------------------------------------------------------
// Viterbi Narrowband — rust/src/vn.rs:17-219.
// The histogram is CE's (hist_block).  Per lane, the circular Viterbi of
// vn.rs:36-95 runs in registers: the state vector over the magnitude bins is
// padded to MAXM with -inf (a padded cell never wins a max, so the clamped
// transition window of vn.rs:55-61 equals the unclamped one), the window max
// of radius `bandwidth` is `bandwidth` rounds of a radius-1 max (max is exact
// and associative), and start states advance vn_start_group<MAXM>() at a
// time so each row of emissions is gathered once per group: 8 up to 24
// magnitude bins (measured 3.5% faster than 4 at 20 bins, no spills), 4 at 32
// (8 would spill).  Emissions come from the
// per-curve table lut[count].  The best start is the first strict maximum of
// the circular totals (vn.rs:81).  Its path is traced back without stored
// backpointers: for each phase, last to first, the winning pass is recomputed
// up to the previous phase, and the previous state is the first strict
// maximum over the transition window (vn.rs:63-68).  Each row's path-mass
// fraction (vn.rs:120-145) is computed during the trace and stored over that
// row's counts (the recomputations only read the rows before it), then the
// fractions are summed in phase order.
// ---------------------------------------------------------------------------
template <int MAXM>
constexpr int vn_start_group() {
    return MAXM <= 24 ? 8 : 4;
}

template <int MAXM>
__device__ __forceinline__ void vn_window(float (&v)[MAXM], int bw) {
    for (int r = 0; r < bw; ++r) {
        float t[MAXM];
#pragma unroll
        for (int m = 0; m < MAXM; ++m) {
            float x = v[m];
            if (m > 0) x = fmaxf(x, v[m - 1]);
            if (m + 1 < MAXM) x = fmaxf(x, v[m + 1]);
            t[m] = x;
        }
#pragma unroll
        for (int m = 0; m < MAXM; ++m) v[m] = t[m];
    }
}

template <int MAXM, bool PACK, class Count>
__device__ __forceinline__ float vn_score(const VnArgs& a, Count count, uint32_t* col, int64_t n,
                                          const float* __restrict__ lut) {
    constexpr float kNegInf = -__builtin_inff();
    const int np_ = a.h.n_phase, nm = a.h.n_mag, bw = a.bandwidth, mg = a.margin;
    // Emissions of a row, gathered from the curve's table.  (An LDS copy of
    // the table's head measured slower: the extra LDS costs the third wave
    // per CU that the histogram leaves room for.)
    auto load_row = [&](int phi, float (&e)[MAXM]) {
#pragma unroll
        for (int m = 0; m < MAXM; ++m) {
            e[m] = kNegInf;
            if (m < nm) e[m] = lut[count(phi, m)];
        }
    };
    // One transition (vn.rs:52-75): window max, plus this row's emission.
    auto step = [&](float (&v)[MAXM], const float (&e)[MAXM]) {
        vn_window<MAXM>(v, bw);
#pragma unroll
        for (int m = 0; m < MAXM; ++m) v[m] = m < nm ? e[m] + v[m] : kNegInf;
    };
    float e0[MAXM];
    load_row(0, e0);
    float best_total = kNegInf;
    int best_s = 0;
    constexpr int kVnStartGroup = vn_start_group<MAXM>();
    for (int s0 = 0; s0 < nm; s0 += kVnStartGroup) {
        float v[kVnStartGroup][MAXM];
#pragma unroll
        for (int j = 0; j < kVnStartGroup; ++j)
#pragma unroll
            for (int m = 0; m < MAXM; ++m) v[j][m] = (m == s0 + j) ? e0[m] : kNegInf;
        for (int phi = 1; phi < np_; ++phi) {
            float e[MAXM];
            load_row(phi, e);
#pragma unroll
            for (int j = 0; j < kVnStartGroup; ++j) step(v[j], e);
        }
#pragma unroll
        for (int j = 0; j < kVnStartGroup; ++j) {
            const int st = s0 + j;
            float total = kNegInf;
#pragma unroll
            for (int m = 0; m < MAXM; ++m)
                if (m == st) total = v[j][m];
            if (st < nm && total > best_total) {
                best_total = total;
                best_s = st;
            }
        }
    }
    if (best_total == kNegInf) return 0.0f;  // vn.rs:109-111

    // Trace back, the concentration fraction of each row on the way.
    auto frac_word = [&](int phi) -> float* {
        const int c0 = phi * nm;  // packed rows own the word holding cells c0+1.. (n_mag >= 3)
        return reinterpret_cast<float*>(col + (PACK ? (c0 + 1) / 2 : c0) * kWave);
    };
    constexpr float kSkip = -1.0f;  // an empty column (vn.rs:127-129)
    const float norm = static_cast<float>(static_cast<uint64_t>(n) *
                                          static_cast<uint64_t>(a.h.phase_ext) *
                                          static_cast<uint64_t>(a.h.mag_ext));
    int pm = best_s;
    for (int phi = np_ - 1; phi >= 0; --phi) {
        float col_mass = 0.0f;
        for (int m = 0; m < nm; ++m) col_mass += static_cast<float>(count(phi, m)) / norm;
        float frac = kSkip;
        if (!(col_mass <= 0.0f)) {
            const int lo = pm >= mg ? pm - mg : 0;
            const int hi = pm + mg < nm ? pm + mg : nm - 1;
            float path_mass = 0.0f;
            for (int m = lo; m <= hi; ++m) path_mass += static_cast<float>(count(phi, m)) / norm;
            frac = path_mass / col_mass;
        }
        *frac_word(phi) = frac;
        if (phi == 0) break;
        float v[MAXM];
#pragma unroll
        for (int m = 0; m < MAXM; ++m) v[m] = (m == best_s) ? e0[m] : kNegInf;
        for (int q = 1; q < phi; ++q) {
            float e[MAXM];
            load_row(q, e);
            step(v, e);
        }
        const int lo = pm >= bw ? pm - bw : 0;
        const int hi = pm + bw < nm ? pm + bw : nm - 1;
        float best_prev = kNegInf;
        int arg = 0;
#pragma unroll
        for (int m = 0; m < MAXM; ++m)
            if (m >= lo && m <= hi && v[m] > best_prev) {
                best_prev = v[m];
                arg = m;
            }
        pm = arg;
    }
    float sum_frac = 0.0f;
    uint32_t n_cols = 0;
    for (int phi = 0; phi < np_; ++phi) {
        const float f = *frac_word(phi);
        if (f != kSkip) {
            sum_frac += f;
            n_cols += 1;
        }
    }
    if (n_cols == 0) return 0.0f;
    const float avg = sum_frac / static_cast<float>(n_cols);
    const int band = (2 * mg + 1) < nm ? (2 * mg + 1) : nm;
    const float baseline = static_cast<float>(band) / static_cast<float>(nm);
    float r = (avg - baseline) / (1.0f - baseline);
    r = (r != r) ? 0.0f : (r > 0.0f ? r : 0.0f);  // f32::max(0.0), NaN -> 0.0
    return r < 1.0f ? r : 1.0f;                    // f32::min(1.0)
}

// LDS: the histogram, as CE's.
template <bool HAS_PDT, bool EXT1, int G, int MAXM>
__device__ __forceinline__ void vn_block(const GridArgs& g, const VnArgs& a, const LaneTrial lt) {
    constexpr bool PACK = EXT1 && G < 0;
    hist_block<HAS_PDT, EXT1, G>(g, a.h, lt, [&](auto count, auto, int, int64_t n, uint32_t* col) {
        const float sc = vn_score<MAXM, PACK>(a, count, col, n, a.lut + a.lut_off[lt.curve]);
        if (lt.active) g.out[lt.curve * g.T + lt.trial] = sc;
    });
}

template <bool HAS_PDT, bool EXT1, int G, int MAXM>
__global__ __launch_bounds__(kWave) void vn_kernel(GridArgs g, VnArgs a) {
    vn_block<HAS_PDT, EXT1, G, MAXM>(g, a, lane_trial(g.n_blocks, g.tblocks, g.T));
}

template <bool HAS_PDT, bool EXT1, int MAXM>
__global__ __launch_bounds__(kWave) void vn_persistent(GridArgs g, VnArgs a, WorkQueue q) {
    QueueState seen = queue_start();
    for (Claim cl = next_claim(q, g, seen); cl.begin < q.total; cl = next_claim(q, g, seen))
        for (int64_t it = cl.begin; it < cl.end; ++it) {
            const LaneTrial lt = queue_trial(q, g, it);
            if (!EXT1) {
                vn_block<HAS_PDT, false, 1, MAXM>(g, a, lt);
                continue;
            }
            if (rank_group(q, item_rank(q, g, it)) < 0) vn_block<HAS_PDT, EXT1, -1, MAXM>(g, a, lt);
            else vn_block<HAS_PDT, EXT1, 0, MAXM>(g, a, lt);
        }
}

// ---------------------------------------------------------------------------
// AOV — rust/src/aov.rs:12-89.
// LDS: per bin three planes (count, sum y, sum y^2), [bin][3][64]; the count
// is a u32, as Rust's (converted to f32 in the epilogue, aov.rs:68), added by
// a non-returning LDS atomic (counts do not depend on order); the two f32
// sums are lane-private read-modify-writes in point order.  Per-curve mean
// is the serial f32 mean (aov.rs:27).
// ---------------------------------------------------------------------------
template <bool HAS_PDT, bool EXT1, int G>
__device__ __forceinline__ void aov_block(const GridArgs& g, const AovArgs& a, const LaneTrial lt) {
    extern __shared__ float lds[];
    const int lane = threadIdx.x;
    const int nb = a.n_bins;
    float* col = lds + lane;
    for (int b = 0; b < 3 * nb; ++b) col[b * kWave] = 0.0f;

    const FoldTrial f = make_fold_trial(g.periods[lt.trial / g.nd], g.pdts[lt.trial % g.nd]);
    const int64_t p0 = g.off[lt.curve];
    const int64_t n = g.off[lt.curve + 1] - p0;
    const float* __restrict__ y = a.y + p0;
    const float* __restrict__ yy = a.yy + p0;
    const float nf = static_cast<float>(nb);
    const DivBin db = make_div_bin(nb);
    const uint32_t nbu = static_cast<uint32_t>(nb);
    constexpr uint32_t kBinBytes = 3 * kPlaneBytes;

    if constexpr (EXT1) {
        // Offsets from the LDS base, this lane's column included (one
        // multiply-add; bin nb, phase 1.0, wraps to bin 0, aov.rs:57-58):
        // the count at +0, the sums y and y^2 at +256 and +512 (planes 0 and
        // 1 of the read-modify-write, from base + 256).
        const uint32_t lane4 = static_cast<uint32_t>(lane) * 4u;
        char* base = reinterpret_cast<char*>(lds);
        // Bins already wrapped (the fast bins never reach n).
        scan_points_rmw<HAS_PDT, G, 2, kChunk, kRmwScalarStream>(
            g.t64 + p0, g.tt + p0, n, f,
            [&](int64_t i) { return F2{rmw_load(y + i), rmw_load(yy + i)}; },
            reinterpret_cast<float*>(base + kPlaneBytes),
            [&](const auto& kb, const auto& v, auto& off, auto& val) {
                constexpr int K = group_size_v<decltype(kb)>;
#pragma unroll
                for (int j = 0; j < K; ++j) {
                    off[j] = __umul24(kb[j], kBinBytes) + lane4;
                    __hip_atomic_fetch_add(reinterpret_cast<uint32_t*>(base + off[j]), 1u,
                                           __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
                    val[0][j] = v[j].a;
                    val[1][j] = v[j].b;
                }
            },
            fast_bins(lane_bin_rule(g.tmax, lt.curve, f, nb, true), [&](float ph) {
                const uint32_t k = div_bin(ph, db);
                return (k == nbu) ? 0u : k;
            }));
    } else {
        scan_points<HAS_PDT, 1>(
            g.t64 + p0, g.tt + p0, n, f,
            [&](int64_t i) { return F2{uniform_load(y + i), uniform_load(yy + i)}; },
            [&](const auto& ph, const auto& v) {
                const uint32_t k = div_bin(ph[0], db);
                for (int o = 0; o < a.ext; ++o) {
                    const uint32_t off1[1] = {
                        static_cast<uint32_t>(__umul24(wrap_add(k, o, nbu), kBinBytes))};
                    __hip_atomic_fetch_add(reinterpret_cast<uint32_t*>(col) + off1[0] / 4, 1u,
                                           __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
                    const float val1[2][1] = {{v[0].a}, {v[0].b}};
                    lds_rmw(col + kWave, off1, val1);
                }
            });
    }

    // aov.rs:67-84.
    const float avg = a.avg[lt.curve];
    float s1 = 0.0f, s2 = 0.0f;
    const uint32_t* ccount = reinterpret_cast<const uint32_t*>(col);
    for (int b = 0; b < nb; ++b) {
        const float c = static_cast<float>(ccount[(3 * b) * kWave]);
        if (c != 0.0f) {
            const float aux = col[(3 * b + 1) * kWave] / c;
            const float res = aux - avg;
            float t1 = c * res;
            t1 = t1 * res;
            s1 += t1;
            float t2 = c * aux;
            t2 = t2 * aux;
            s2 += col[(3 * b + 2) * kWave] - t2;
        }
    }
    float F = 0.0f;
    if (s2 != 0.0f) {
        const float lhs = (static_cast<float>(n) - nf) / (nf - 1.0f);
        F = lhs * (s1 / s2);
    }
    if (lt.active) g.out[lt.curve * g.T + lt.trial] = F;
}

// One workgroup per (curve, 64-trial block), hardware-scheduled.
template <bool HAS_PDT, bool EXT1, int G>
__global__ __launch_bounds__(kWave) void aov_kernel(GridArgs g, AovArgs a) {
    aov_block<HAS_PDT, EXT1, G>(g, a, lane_trial(g.n_blocks, g.tblocks, g.T));
}

// Persistent: each wave pulls (curve, 64-trial block) items from the queue.
template <bool HAS_PDT, bool EXT1>
__global__ __launch_bounds__(kWave) void aov_persistent(GridArgs g, AovArgs a, WorkQueue q) {
    QueueState seen = queue_start();
    for (Claim cl = next_claim(q, g, seen); cl.begin < q.total; cl = next_claim(q, g, seen))
        for (int64_t it = cl.begin; it < cl.end; ++it) {
            const LaneTrial lt = queue_trial(q, g, it);
            if (!EXT1) {
                aov_block<HAS_PDT, false, 1>(g, a, lt);
                continue;
            }
            switch (rank_group(q, item_rank(q, g, it))) {
                case 2: aov_block<HAS_PDT, EXT1, 2>(g, a, lt); break;
                case 4: aov_block<HAS_PDT, EXT1, 4>(g, a, lt); break;
                default: aov_block<HAS_PDT, EXT1, 1>(g, a, lt); break;
            }
        }
}

// ---------------------------------------------------------------------------
// Lomb-Scargle — rust/src/ls.rs:15-79.
// Registers only: C = sum y cos, S = sum y sin, CC = sum cos^2, CS = sum cos sin
// accumulated serially per trial.  FAST uses the hardware v_sin/v_cos, which
// take the angle in revolutions (the fold phase itself); otherwise sincosf of
// fl(2pi_f32 * phase) as the reference computes it.
// ---------------------------------------------------------------------------
template <bool HAS_PDT, bool FAST>
__device__ __forceinline__ void ls_block(const GridArgs& g, const LsArgs& a, const LaneTrial lt) {
    const FoldTrial f = make_fold_trial(g.periods[lt.trial / g.nd], g.pdts[lt.trial % g.nd]);
    const int64_t p0 = g.off[lt.curve];
    const int64_t n = g.off[lt.curve + 1] - p0;
    const float* __restrict__ y = a.y + p0;
    const float two_pi = 2.0f * 3.14159274101257324f;  // 2 * (f32)pi, ls.rs:9

    float mc = 0.0f, ms = 0.0f, cc = 0.0f, cs = 0.0f;
    // The hardware sin/cos path takes the uncorrected fold (its bar is a
    // tolerance); the precise path keeps the exact one.
    scan_points<HAS_PDT, 4, !FAST>(
        g.t64 + p0, g.tt + p0, n, f, [&](int64_t i) { return uniform_load(y + i); },
        [&](const auto& ph, const auto& yv) {
            constexpr int K = group_size_v<decltype(ph)>;
#pragma unroll
            for (int j = 0; j < K; ++j) {
                float s, c;
                if (FAST) {
                    s = __builtin_amdgcn_sinf(ph[j]);
                    c = __builtin_amdgcn_cosf(ph[j]);
                } else {
                    sincosf(two_pi * ph[j], &s, &c);
                }
                const float yi = yv[j];
                if (FAST) {
                    // Fused: each term differs from the separate multiply
                    // and add by at most half an ulp of the product, far
                    // below the hardware sin/cos error this path accepts.
                    mc = __builtin_fmaf(yi, c, mc);
                    ms = __builtin_fmaf(yi, s, ms);
                    cc = __builtin_fmaf(c, c, cc);
                    cs = __builtin_fmaf(c, s, cs);
                } else {
                    float p;
                    p = yi * c; mc += p;
                    p = yi * s; ms += p;
                    p = c * c;  cc += p;
                    p = c * s;  cs += p;
                }
            }
        });
    // ls.rs:47-77.
    const float ss = static_cast<float>(n) - cc;
    const float tau = 0.5f * atan2f(2.0f * cs, cc - ss);
    float st, ct;
    sincosf(tau, &st, &ct);
    float v = ct * mc + st * ms;
    const float num_l = v * v;
    v = ct * ms - st * mc;
    const float num_r = v * v;
    const float den_l = ct * ct * cc + 2.0f * ct * st * cs + st * st * ss;
    const float den_r = ct * ct * ss - 2.0f * ct * st * cs + st * st * cc;
    float P = 0.0f;
    if (!(den_l == 0.0f || den_r == 0.0f)) P = 0.5f * (num_l / den_l + num_r / den_r);
    if (lt.active) g.out[lt.curve * g.T + lt.trial] = P;
}

// One workgroup per (curve, 64-trial block), hardware-scheduled.
template <bool HAS_PDT, bool FAST>
__global__ __launch_bounds__(kWave) void ls_kernel(GridArgs g, LsArgs a) {
    ls_block<HAS_PDT, FAST>(g, a, lane_trial(g.n_blocks, g.tblocks, g.T));
}

// Persistent: each wave pulls (curve, 64-trial block) items from the queue.
template <bool HAS_PDT, bool FAST>
__global__ __launch_bounds__(kWave) void ls_persistent(GridArgs g, LsArgs a, WorkQueue q) {
    QueueState seen = queue_start();
    for (Claim cl = next_claim(q, g, seen); cl.begin < q.total; cl = next_claim(q, g, seen))
        for (int64_t it = cl.begin; it < cl.end; ++it)
            ls_block<HAS_PDT, FAST>(g, a, queue_trial(q, g, it));
}

// ---------------------------------------------------------------------------
// FPW — rust/src/fpw.rs:13-67.
// LDS: [bin][2][64] (sum w, sum w*y); w = 1/(e*e), wy = w*y prepared per
// curve set (fpw.rs:26-27, uncentred).  With equal weights (FpwArgs::wsum,
// every curve's w bitwise equal, *unequal == 0) the first plane counts the
// bin's points by non-returning atomics instead (u32), and a bin of c points
// weighs wsum[c - 1], the curve's running serial sum of c weights: the same
// value as adding them; only the w*y plane stays a read-modify-write.
// ---------------------------------------------------------------------------
template <bool HAS_PDT, int G>
__device__ __forceinline__ void fpw_block(const GridArgs& g, const FpwArgs& a, const LaneTrial lt) {
    extern __shared__ float fl[];
    const int lane = threadIdx.x;
    const int nb = a.n_bins;
    float* col = fl + lane;
    for (int b = 0; b < 2 * nb; ++b) col[b * kWave] = 0.0f;
    const FoldTrial f = make_fold_trial(g.periods[lt.trial / g.nd], g.pdts[lt.trial % g.nd]);
    const int64_t p0 = g.off[lt.curve];
    const int64_t n = g.off[lt.curve + 1] - p0;
    const float* __restrict__ w = a.w + p0;
    const float* __restrict__ wy = a.wy + p0;
    const float nf = static_cast<float>(nb);
    const uint32_t nm1 = static_cast<uint32_t>(nb - 1);
    // (wave-uniform: one scalar load of the set's flag)
    const bool ew = a.wsum != nullptr && uniform_load(a.unequal) == 0;
    // Bins already clamped to n - 1 (fpw.rs:48; the fast bins never reach n).
    const auto bins = fast_bins(lane_bin_rule(g.tmax, lt.curve, f, nb, false),
                                [&](float ph) { return mul_bin(ph, nf, nm1); });

    if (ew) {
        uint32_t* ccol = reinterpret_cast<uint32_t*>(col);
        scan_points_rmw<HAS_PDT, G, 1, kChunk, kRmwScalarStream>(
            g.t64 + p0, g.tt + p0, n, f, [&](int64_t i) { return rmw_load(wy + i); },
            col + kWave,
            [&](const auto& kb, const auto& v, auto& off, auto& val) {
                constexpr int K = group_size_v<decltype(kb)>;
#pragma unroll
                for (int j = 0; j < K; ++j) {
                    off[j] = kb[j] * (2 * kPlaneBytes);
                    val[0][j] = v[j];
                    __hip_atomic_fetch_add(
                        reinterpret_cast<uint32_t*>(reinterpret_cast<char*>(ccol) + off[j]), 1u,
                        __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
                }
            },
            bins);
    } else {
        scan_points_rmw<HAS_PDT, G, 2, kChunk, kRmwScalarStream>(
            g.t64 + p0, g.tt + p0, n, f,
            [&](int64_t i) { return F2{rmw_load(w + i), rmw_load(wy + i)}; }, col,
            [&](const auto& kb, const auto& v, auto& off, auto& val) {
                constexpr int K = group_size_v<decltype(kb)>;
#pragma unroll
                for (int j = 0; j < K; ++j) {
                    off[j] = kb[j] * (2 * kPlaneBytes);
                    val[0][j] = v[j].a;
                    val[1][j] = v[j].b;
                }
            },
            bins);
    }
    // fpw.rs:55-60.
    const float* __restrict__ ws = ew ? a.wsum + p0 : nullptr;
    float s = 0.0f;
    for (int b = 0; b < nb; ++b) {
        float W;
        if (ew) {
            const uint32_t c = reinterpret_cast<const uint32_t*>(col)[(2 * b) * kWave];
            W = c > 0 ? ws[c - 1] : 0.0f;
        } else {
            W = col[(2 * b) * kWave];
        }
        if (W > 0.0f) {
            const float Y = col[(2 * b + 1) * kWave];
            const float num = Y * Y;
            const float den = 2.0f * W;
            s += num / den;
        }
    }
    if (lt.active) g.out[lt.curve * g.T + lt.trial] = s;
}

// One workgroup per (curve, 64-trial block), hardware-scheduled.
template <bool HAS_PDT, int G>
__global__ __launch_bounds__(kWave) void fpw_kernel(GridArgs g, FpwArgs a) {
    fpw_block<HAS_PDT, G>(g, a, lane_trial(g.n_blocks, g.tblocks, g.T));
}

// Persistent: each wave pulls (curve, 64-trial block) items from the queue.
template <bool HAS_PDT>
__global__ __launch_bounds__(kWave) void fpw_persistent(GridArgs g, FpwArgs a, WorkQueue q) {
    QueueState seen = queue_start();
    for (Claim cl = next_claim(q, g, seen); cl.begin < q.total; cl = next_claim(q, g, seen))
        for (int64_t it = cl.begin; it < cl.end; ++it) {
            const LaneTrial lt = queue_trial(q, g, it);
            switch (rank_group(q, item_rank(q, g, it))) {
                case 2: fpw_block<HAS_PDT, 2>(g, a, lt); break;
                case 4: fpw_block<HAS_PDT, 4>(g, a, lt); break;
                default: fpw_block<HAS_PDT, 1>(g, a, lt); break;
            }
        }
}

// ---------------------------------------------------------------------------
// Matched Filter — rust/src/mf.rs:172-330 (eval_mf_at_period).
// LDS: [bin][3][64] f32: sum w, sum w*y, sum w*y*y, in point order
// (mf.rs:181-187); w = 1/(e*e), w*y and (w*y)*y are prepared per curve set.
// Epilogue per lane, every f32 operation as in Rust:
//   * one pass over the bins (k order) for the profile Y/W, filled bins,
//     total weights, within-bin sums of squares, extremes and RMS (mf.rs:
//     196-291; separate accumulators, so one pass gives the same sums);
//   * R^2 needs every point again: the weighted sum of squares about this
//     trial's grand mean, in point order (mf.rs:230-234);
//   * circular Pearson maxima against the 14 templates (mf.rs:124-153).  The
//     profile mean, variance and denominator are the same for every template,
//     so they are computed once; the centred profile c[k] is kept in the lane's
//     LDS column.  Shifts are evaluated kMfShiftBlock at a time with one
//     accumulator each, each summing c[k]*t[k+s] over k in ascending order
//     (the reference's order).  As the denominator is positive and rounding
//     is monotone, the best fl(dot/denom) over shifts is fl(max dot / denom):
//     one division per template instead of one per shift.
// ---------------------------------------------------------------------------
struct F3 {
    float a, b, c;
};

template <bool HAS_PDT, int G>
__device__ __forceinline__ void mf_block(const GridArgs& g, const MfArgs& a, const LaneTrial lt) {
    extern __shared__ float fl[];
    const int lane = threadIdx.x;
    const int nb = a.n_bins;
    float* col = fl + lane;  // bin k: W at plane 3k, Y at 3k+1, V at 3k+2
    for (int b = 0; b < 3 * nb; ++b) col[b * kWave] = 0.0f;
    const FoldTrial f =
        a.curve_period ? make_fold_trial(a.curve_period[lt.curve], 0.0f)
                       : make_fold_trial(g.periods[lt.trial / g.nd], g.pdts[lt.trial % g.nd]);
    const int64_t p0 = g.off[lt.curve];
    const int64_t n = g.off[lt.curve + 1] - p0;
    const float* __restrict__ w = a.w + p0;
    const float* __restrict__ wy = a.wy + p0;
    const float* __restrict__ wyy = a.wyy + p0;
    const float nf = static_cast<float>(nb);
    const uint32_t nm1 = static_cast<uint32_t>(nb - 1);
    constexpr uint32_t kBinBytes = 3 * kPlaneBytes;

    scan_points_rmw<HAS_PDT, G, 3, kChunk, kRmwScalarStream>(
        g.t64 + p0, g.tt + p0, n, f, [&](int64_t i) {
            return F3{rmw_load(w + i), rmw_load(wy + i), rmw_load(wyy + i)};
        }, col,
        [&](const auto& kb, const auto& v, auto& off, auto& val) {
            constexpr int K = group_size_v<decltype(kb)>;
#pragma unroll
            for (int j = 0; j < K; ++j) {
                off[j] = keep_reg(__umul24(kb[j], kBinBytes));
                val[0][j] = v[j].a;
                val[1][j] = v[j].b;
                val[2][j] = v[j].c;
            }
        },
        // Bins clamped to n - 1 (mf.rs:193).
        fast_bins(lane_bin_rule(g.tmax, lt.curve, f, nb, false),
                  [&](float ph) { return mul_bin(ph, nf, nm1); }));

    // Bins (mf.rs:196-291).  The profile replaces Y in LDS (0 for an empty bin).
    uint32_t n_filled = 0, n_rms = 0;
    float total_wt = 0.0f, total_wy = 0.0f, ss_within = 0.0f, sum_rms = 0.0f;
    float min_p = __builtin_inff(), max_p = -__builtin_inff();
    for (int k = 0; k < nb; ++k) {
        const float W = col[(3 * k) * kWave];
        float prof = 0.0f;
        if (W > 0.0f) {
            const float Y = col[(3 * k + 1) * kWave];
            const float V = col[(3 * k + 2) * kWave];
            prof = Y / W;
            n_filled += 1;
            total_wt += W;
            total_wy += Y;
            const float yyw = Y * Y / W;
            ss_within += V - yyw;
            if (prof < min_p) min_p = prof;
            if (prof > max_p) max_p = prof;
            const float wv = (V - yyw) / W;
            if (wv > 0.0f) {
                sum_rms += __builtin_sqrtf(wv);
                n_rms += 1;
            }
        }
        col[(3 * k + 1) * kWave] = prof;
    }
    const float coverage = static_cast<float>(n_filled) / nf;
    const bool scored = n_filled >= 3;  // mf.rs:208-219

    // R^2 (mf.rs:225-245): a second pass over the points, wave-uniform loads.
    float r2 = 0.0f;
    if (scored) {
        const float gm = total_wy / total_wt;
        const float* __restrict__ yv = a.y + p0;
        float ss_total = 0.0f;
        int64_t i = 0;
        for (; i + 8 <= n; i += 8) {
            float yy[8], ww[8];
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                yy[j] = uniform_load(yv + i + j);
                ww[j] = uniform_load(w + i + j);
            }
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                const float d = yy[j] - gm;
                ss_total += ww[j] * d * d;
            }
        }
        for (; i < n; ++i) {
            const float d = uniform_load(yv + i) - gm;
            ss_total += uniform_load(w + i) * d * d;
        }
        if (ss_total > 0.0f) {
            const float q = 1.0f - ss_within / ss_total;
            r2 = q > 0.0f ? q : 0.0f;  // f32::max(q, 0.0), NaN -> 0.0
        }
    }

    // Template correlations (mf.rs:124-153, 293-316).
    float best[3] = {0.0f, 0.0f, 0.0f};  // sawtooth, sinusoidal, eclipsing
    if (scored) {
        float p_mean = -0.0f;
        for (int k = 0; k < nb; ++k) p_mean = p_mean + col[(3 * k + 1) * kWave];
        p_mean = p_mean / nf;
        float p_var = -0.0f;
        for (int k = 0; k < nb; ++k) {
            const float c = col[(3 * k + 1) * kWave] - p_mean;
            p_var = p_var + c * c;
            col[(3 * k + 2) * kWave] = c;  // centred profile
        }
        const float denom = __builtin_sqrtf(p_var) * __builtin_sqrtf(nf);
        const bool flat = p_var < 1e-20f || denom < 1e-20f;  // correlation 0
        const int stride = mf_template_stride(nb);  // indices below reach 2*nb + 6
        for (int tp = 0; tp < kMfShapes; ++tp) {
            const float* __restrict__ tt = a.tmpl + tp * stride;
            float maxdot = -__builtin_inff();
            for (int s0 = 0; s0 < nb; s0 += kMfShiftBlock) {
                float dot[kMfShiftBlock];
#pragma unroll
                for (int j = 0; j < kMfShiftBlock; ++j) dot[j] = 0.0f;
                for (int k = 0; k < nb; ++k) {
                    const float c = col[(3 * k + 2) * kWave];
#pragma unroll
                    for (int j = 0; j < kMfShiftBlock; ++j)
                        dot[j] = dot[j] + c * uniform_load(tt + s0 + k + j);
                }
#pragma unroll
                for (int j = 0; j < kMfShiftBlock; ++j)
                    if (s0 + j < nb && dot[j] > maxdot) maxdot = dot[j];
            }
            float corr = 0.0f;
            if (!flat) {
                const float r = maxdot / denom;
                corr = r > -1.0f ? r : -1.0f;  // best starts at -1 (mf.rs:140)
            }
            const int kind = tp < 5 ? 0 : (tp == 5 ? 1 : 2);
            if (corr > best[kind]) best[kind] = corr;
        }
    }
    const float max_corr = fmaxf(fmaxf(best[0], best[1]), best[2]);
    const float combined = max_corr * r2 * coverage;

    if (!lt.active) return;
    if (a.feat) {
        float amp_snr = 0.0f;
        if (scored) {
            const float amplitude = max_p - min_p;
            const float mean_rms = n_rms > 0 ? sum_rms / static_cast<float>(n_rms) : 1.0f;
            amp_snr = mean_rms > 1e-12f ? amplitude / mean_rms : 0.0f;
        }
        float* o = a.feat + lt.curve * kMfFeatures;
        o[0] = best[0];
        o[1] = best[1];
        o[2] = best[2];
        o[3] = r2;
        o[4] = amp_snr;
        o[5] = static_cast<float>(n_filled);
        o[6] = combined;
    } else {
        g.out[lt.curve * g.T + lt.trial] = combined;
    }
}

template <bool HAS_PDT, int G>
__global__ __launch_bounds__(kWave) void mf_kernel(GridArgs g, MfArgs a) {
    mf_block<HAS_PDT, G>(g, a, lane_trial(g.n_blocks, g.tblocks, g.T));
}

template <bool HAS_PDT>
__global__ __launch_bounds__(kWave) void mf_persistent(GridArgs g, MfArgs a, WorkQueue q) {
    QueueState seen = queue_start();
    for (Claim cl = next_claim(q, g, seen); cl.begin < q.total; cl = next_claim(q, g, seen))
        for (int64_t it = cl.begin; it < cl.end; ++it) {
            const LaneTrial lt = queue_trial(q, g, it);
            switch (rank_group(q, item_rank(q, g, it))) {
                case 2: mf_block<HAS_PDT, 2>(g, a, lt); break;
                case 4: mf_block<HAS_PDT, 4>(g, a, lt); break;
                default: mf_block<HAS_PDT, 1>(g, a, lt); break;
            }
        }
}

// ---------------------------------------------------------------------------
// Multi-Harmonic Fourier — rust/src/mhf.rs:75-226 (eval_mhf_single_period).
// Weighted least squares of y on [1, x, cos h phi, sin h phi (h = 1..K)],
// phi = 2 pi phase, for the models k = 0..K, scored by BIC; the trial's value
// is the best BIC gain over the flat model k = 0.  FP64 throughout, in
// registers (no LDS).
//
// Normal equations.  The reference sums every product of two design columns
// over the points (mhf.rs:129-145).  Products of harmonics reduce to single
// harmonics (cos a cos b = (cos(a-b) + cos(a+b))/2, and likewise), so the
// point loop accumulates only
//   C[m], S[m] = sum w cos(m phi), sum w sin(m phi)       m = 1..2K
//   PC[h], PS[h] = sum (w x) cos(h phi), sum (w x) sin     h = 1..K
//   YC[h], YS[h] = sum (w y) cos(h phi), sum (w y) sin
// (8K sums instead of the (2K+2)(2K+3)/2 + 2K+2 entries), the harmonics by
// the reference's Chebyshev recurrence from one sincos per point; the
// offset/slope sums do not depend on the trial and come per curve.
//
// Models.  The leading (2k+2)-block of the normal matrix is model k's, and
// Cholesky factors and forward substitutions are prefix-consistent: model k's
// factor and z = L^-1 b are the leading parts of the full ones, computed by the
// same operations.  So one factorization serves every model (the reference
// factors each model's block anew with the same operations, mhf.rs:159-177),
// and chi2_k = y'Wy - sum_{j < 2k+2} z_j^2, so that
//   BIC_flat - BIC_k = sum_{2 <= j < 2k+2} z_j^2 - 2k ln n
// without the residual pass of mhf.rs:181-189 (and without cancelling two
// large chi2 values).  A pivot that is not positive fails its model and every
// larger one (cholesky_solve's failure, fourier.rs:37-39); a NaN pivot, which
// the reference carries into a NaN score clamped to 0, gives the same 0 here.
//
// sin/cos of phi per point: a table of (cos, sin)(2 pi j/256), j = 0..256, in
// the wave's LDS (filled once per wave with the device sincospi), and the
// angle-addition formula with Taylor polynomials for the remainder
// |d| <= pi/256 (sin to d^7, cos to d^6: truncation below 1e-20 relative).
// About 20 VALU operations and one LDS read per point instead of a full
// double-precision sincos with its range reduction.
// ---------------------------------------------------------------------------
constexpr int kMhfTab = 257;
constexpr size_t kMhfLds = kMhfTab * sizeof(double2);

__device__ __forceinline__ void mhf_fill_table(double2* tab) {
    for (int j = threadIdx.x; j < kMhfTab; j += kWave) {
        double s, c;
        sincospi(static_cast<double>(j) * (1.0 / 128.0), &s, &c);
        tab[j] = double2{c, s};
    }
    __syncthreads();
}

// sin and cos of 2 pi ph for K phases ph in [0, 1], stage by stage: the K
// table reads are issued together and the polynomials computed while they
// are in flight.
template <int K>
__device__ __forceinline__ void mhf_sincos_n(const float (&ph)[K], const double2* tab,
                                             double (&s)[K], double (&c)[K]) {
    double d[K];
    double2 t0[K];
#pragma unroll
    for (int k = 0; k < K; ++k) {
        const double u = static_cast<double>(ph[k]) * 256.0;
        const double j = __builtin_rint(u);
        d[k] = (u - j) * (6.283185307179586476925286766559 / 256.0);
        int ji = static_cast<int>(j);
        ji = ji < 0 ? 0 : (ji > kMhfTab - 1 ? kMhfTab - 1 : ji);
        t0[k] = tab[ji];
    }
    __builtin_amdgcn_sched_barrier(0);
    double sd[K], cd[K];
#pragma unroll
    for (int k = 0; k < K; ++k) {
        const double d2 = d[k] * d[k];
        double p = __builtin_fma(d2, -1.0 / 5040.0, 1.0 / 120.0);
        p = __builtin_fma(d2, p, -1.0 / 6.0);
        sd[k] = __builtin_fma(d[k], p * d2, d[k]);
        double q = __builtin_fma(d2, -1.0 / 720.0, 1.0 / 24.0);
        q = __builtin_fma(d2, q, -0.5);
        cd[k] = __builtin_fma(d2, q, 1.0);
    }
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int k = 0; k < K; ++k) {
        s[k] = __builtin_fma(t0[k].y, cd[k], t0[k].x * sd[k]);
        c[k] = __builtin_fma(t0[k].x, cd[k], -(t0[k].y * sd[k]));
    }
}

// sin and cos of 2 pi r for an f64 revolution count r in [0, 1] (the table
// and polynomials of mhf_sincos_n).
__device__ __forceinline__ void mhf_sincos_rev(double r, const double2* tab, double& s, double& c) {
    const double u = r * 256.0;  // exact
    const double j = __builtin_rint(u);
    const double d = (u - j) * (6.283185307179586476925286766559 / 256.0);
    int ji = static_cast<int>(j);
    ji = ji < 0 ? 0 : (ji > kMhfTab - 1 ? kMhfTab - 1 : ji);
    const double2 t0 = tab[ji];
    const double d2 = d * d;
    double p = __builtin_fma(d2, -1.0 / 5040.0, 1.0 / 120.0);
    p = __builtin_fma(d2, p, -1.0 / 6.0);
    const double sd = __builtin_fma(d, p * d2, d);
    double q = __builtin_fma(d2, -1.0 / 720.0, 1.0 / 24.0);
    q = __builtin_fma(d2, q, -0.5);
    const double cd = __builtin_fma(d2, q, 1.0);
    s = __builtin_fma(t0.y, cd, t0.x * sd);
    c = __builtin_fma(t0.x, cd, -(t0.y * sd));
}

// Walk a curve's MHF records in order.  Records carry their own time (t*t
// for the period derivative is formed in the kernel, exactly).  Chunks of C
// records are prefetched with scalar loads one chunk ahead, as in
// scan_points, but in two steps per chunk: first the fold and the sin/cos
// (whose table reads are LDS traffic, waited for on the same counter as
// scalar loads, so they run while no scalar load is outstanding), then the
// next chunk's prefetch, then the chunk's accumulation, which covers it.
// `body(s1, c1, rec)` receives sin and cos of phi for one point.
__device__ __forceinline__ MhfPoint load_record(const MhfPoint* p) {
    return MhfPoint{uniform_load(&p->t), uniform_load(&p->w), uniform_load(&p->wx),
                    uniform_load(&p->wy)};
}

template <bool HAS_PDT, int C, class Body>
__device__ __forceinline__ void scan_records(const MhfPoint* __restrict__ pt, int64_t n,
                                             const FoldTrial& f, const double2* tab, Body body) {
    struct Chunk {
        MhfPoint r[C];
    };
    auto fetch = [&](int64_t base, Chunk& ch) {
#pragma unroll
        for (int j = 0; j < C; ++j) ch.r[j] = load_record(pt + base + j);
    };
    auto trig = [&](const Chunk& ch, double (&s)[C], double (&c)[C]) {
        double t[C], tt[C];
        float ph[C];
#pragma unroll
        for (int j = 0; j < C; ++j) {
            t[j] = ch.r[j].t;
            tt[j] = HAS_PDT ? t[j] * t[j] : 0.0;
        }
        fold_phases<HAS_PDT, C>(t, tt, f, ph);
        mhf_sincos_n<C>(ph, tab, s, c);
    };
    auto accumulate = [&](const Chunk& ch, const double (&s)[C], const double (&c)[C]) {
#pragma unroll
        for (int j = 0; j < C; ++j) body(s[j], c[j], ch.r[j]);
    };
    int64_t i = 0;
    if (n >= 2 * C) {
        const int64_t last = n - C;
        Chunk a, b;
        double s[C], c[C];
        fetch(0, a);
        wait_scalar_loads();
        for (; i + 2 * C <= n; i += 2 * C) {
            trig(a, s, c);
            __builtin_amdgcn_sched_barrier(0);
            fetch(i + C, b);
            __builtin_amdgcn_sched_barrier(0);
            accumulate(a, s, c);
            wait_scalar_loads();
            trig(b, s, c);
            __builtin_amdgcn_sched_barrier(0);
            const int64_t na = (i + 2 * C <= last) ? i + 2 * C : last;
            fetch(na, a);
            __builtin_amdgcn_sched_barrier(0);
            accumulate(b, s, c);
            wait_scalar_loads();
        }
    }
    for (; i < n; ++i) {
        const MhfPoint r = load_record(pt + i);
        float ph[1] = {fold_phase<HAS_PDT>(r.t, HAS_PDT ? r.t * r.t : 0.0, f)};
        double s1[1], c1[1];
        mhf_sincos_n<1>(ph, tab, s1, c1);
        body(s1[0], c1[0], r);
    }
}

// The 8K harmonic sums of one trial (see above), harmonics 1-based.
template <int K>
struct MhfAcc {
    static constexpr int M = 2 * K;
    static constexpr int N = 8 * K;
    double v[N > 0 ? N : 1];
    __device__ __forceinline__ double& C(int m) { return v[m - 1]; }            // m = 1..2K
    __device__ __forceinline__ double& S(int m) { return v[M + m - 1]; }
    __device__ __forceinline__ double& PC(int h) { return v[2 * M + h - 1]; }   // h = 1..K
    __device__ __forceinline__ double& PS(int h) { return v[2 * M + K + h - 1]; }
    __device__ __forceinline__ double& YC(int h) { return v[2 * M + 2 * K + h - 1]; }
    __device__ __forceinline__ double& YS(int h) { return v[2 * M + 3 * K + h - 1]; }
};

// The trial's period and period derivative (per-model mode: the curve's).
struct MhfTrial {
    float period, pdt;
};

__device__ __forceinline__ MhfTrial mhf_trial(const GridArgs& g, const MhfArgs& a,
                                              const LaneTrial lt) {
    return a.feat ? MhfTrial{g.periods[lt.curve], g.pdts[lt.curve]}
                  : MhfTrial{g.periods[lt.trial / g.nd], g.pdts[lt.trial % g.nd]};
}

// Add the harmonic sums of points [0, n) of `pt` to acc, in point order.
template <bool HAS_PDT, int K>
__device__ __forceinline__ void mhf_accumulate(const MhfPoint* pt, int64_t n, const MhfTrial tr,
                                               const double2* tab, MhfAcc<K>& acc) {
    constexpr int M = 2 * K;
    if constexpr (K > 0) {
        const FoldTrial f = make_fold_trial(tr.period, tr.pdt);
        scan_records<HAS_PDT, 4>(pt, n, f, tab, [&](double s1, double c1, const MhfPoint& r) {
            // phi = 2 pi phase (mhf.rs:100).
            double c[M + 1], s[M + 1];
            c[1] = c1;
            s[1] = s1;
            c[0] = 1.0;
            s[0] = 0.0;
            const double two_c1 = 2.0 * c[1];
#pragma unroll
            for (int m = 2; m <= M; ++m) {  // mhf.rs:118-128
                c[m] = __builtin_fma(two_c1, c[m - 1], -c[m - 2]);
                s[m] = __builtin_fma(two_c1, s[m - 1], -s[m - 2]);
            }
#pragma unroll
            for (int m = 1; m <= M; ++m) {
                acc.C(m) = __builtin_fma(r.w, c[m], acc.C(m));
                acc.S(m) = __builtin_fma(r.w, s[m], acc.S(m));
            }
#pragma unroll
            for (int h = 1; h <= K; ++h) {
                acc.PC(h) = __builtin_fma(r.wx, c[h], acc.PC(h));
                acc.PS(h) = __builtin_fma(r.wx, s[h], acc.PS(h));
                acc.YC(h) = __builtin_fma(r.wy, c[h], acc.YC(h));
                acc.YS(h) = __builtin_fma(r.wy, s[h], acc.YS(h));
            }
        });
    }
}

// Normal equations from the sums, one Cholesky, every model's BIC gain
// (mhf.rs:147-225); writes the lane's output.
template <int K>
__device__ __forceinline__ void mhf_finish(const GridArgs& g, const MhfArgs& a, const LaneTrial lt,
                                           const MhfTrial tr, MhfAcc<K>& acc) {
    constexpr int NP = 2 + 2 * K;  // parameters of the largest model
    const int64_t n = g.off[lt.curve + 1] - g.off[lt.curve];
    const double* u = a.sums + lt.curve * kMhfSums;
    const double sw = u[0], swx = u[1], swxx = u[2], swy = u[3], swxy = u[4], ln_n = u[5];
    // Normal matrix entry (i >= j) and right-hand side; columns 0 = offset,
    // 1 = slope, 2h = cos h phi, 2h+1 = sin h phi (mhf.rs:103-116).
    auto A = [&](int i, int j) -> double {
        if (i < 2) return i == 0 ? sw : (j == 0 ? swx : swxx);
        const int hi = i >> 1;
        const bool si = i & 1;
        if (j == 0) return si ? acc.S(hi) : acc.C(hi);
        if (j == 1) return si ? acc.PS(hi) : acc.PC(hi);
        const int hj = j >> 1;
        const bool sj = j & 1;
        const int d = hi - hj;  // >= 0
        const double cd = d == 0 ? sw : acc.C(d);
        const double sd = d == 0 ? 0.0 : acc.S(d);
        if (!si && !sj) return 0.5 * (cd + acc.C(hi + hj));
        if (si && sj) return 0.5 * (cd - acc.C(hi + hj));
        if (si) return 0.5 * (acc.S(hi + hj) + sd);  // sin hi cos hj
        return 0.5 * (acc.S(hi + hj) - sd);          // cos hi sin hj
    };
    auto B = [&](int i) -> double {
        if (i < 2) return i == 0 ? swy : swxy;
        return (i & 1) ? acc.YS(i >> 1) : acc.YC(i >> 1);
    };

    // Cholesky, row by row as fourier.rs:30-47 (each entry's sum in ascending
    // k), with the forward substitution of fourier.rs:49-53 per row.
    double L[NP * (NP + 1) / 2];
    double rd[NP], z[NP];
    int bad = NP;  // first pivot that is not positive
#pragma unroll
    for (int i = 0; i < NP; ++i) {
#pragma unroll
        for (int j = 0; j <= i; ++j) {
            double v = A(i, j);
#pragma unroll
            for (int k = 0; k < j; ++k)
                v = __builtin_fma(-L[i * (i + 1) / 2 + k], L[j * (j + 1) / 2 + k], v);
            if (j < i) {
                L[i * (i + 1) / 2 + j] = v * rd[j];
            } else {
                bad = (!(v > 0.0) && bad == NP) ? i : bad;
                const double d = sqrt(v);
                L[i * (i + 1) / 2 + i] = d;
                rd[i] = 1.0 / d;
            }
        }
        double v = B(i);
#pragma unroll
        for (int k = 0; k < i; ++k) v = __builtin_fma(-L[i * (i + 1) / 2 + k], z[k], v);
        z[i] = v * rd[i];
    }

    // mhf.rs:192-225.  Model k needs n > 2k + 2 points and a factor through
    // pivot 2k+1; without model 0 (or for an invalid period) every score is 0.
    const bool valid = tr.period > 0.0f && isfinite(tr.period) && n >= 3 && bad >= 2;
    float delta[K + 1];
    delta[0] = 0.0f;
    double best = 0.0, q = 0.0;
    int best_k = 0;
#pragma unroll
    for (int k = 1; k <= K; ++k) {
        q = __builtin_fma(z[2 * k], z[2 * k], q);
        q = __builtin_fma(z[2 * k + 1], z[2 * k + 1], q);
        const bool ok = valid && bad >= 2 * k + 2 && n > 2 * k + 2;
        double d = q - static_cast<double>(2 * k) * ln_n;
        d = (ok && d > 0.0) ? d : 0.0;  // f64::max(0.0) maps NaN to 0
        delta[k] = static_cast<float>(d);
        if (d > best) {
            best = d;
            best_k = k;
        }
    }

    if (!lt.active) return;
    if (a.feat) {
        float* o = a.feat + lt.curve * (K + 2);
#pragma unroll
        for (int k = 0; k <= K; ++k) o[k] = delta[k];
        o[K + 1] = static_cast<float>(best_k);
    } else {
        g.out[lt.curve * g.T + lt.trial] = static_cast<float>(best);
    }
}

template <bool HAS_PDT, int K>
__device__ __forceinline__ void mhf_block(const GridArgs& g, const MhfArgs& a, const LaneTrial lt,
                                          const double2* tab) {
    const MhfTrial tr = mhf_trial(g, a, lt);
    const int64_t p0 = g.off[lt.curve];
    MhfAcc<K> acc;
#pragma unroll
    for (int k = 0; k < MhfAcc<K>::N; ++k) acc.v[k] = 0.0;
    mhf_accumulate<HAS_PDT, K>(a.pt + p0, g.off[lt.curve + 1] - p0, tr, tab, acc);
    mhf_finish<K>(g, a, lt, tr, acc);
}

// One workgroup per (curve, 64-trial block), hardware-scheduled.
template <bool HAS_PDT, int K>
__global__ __launch_bounds__(kWave) void mhf_kernel(GridArgs g, MhfArgs a) {
    extern __shared__ double2 tab[];
    if constexpr (K > 0) mhf_fill_table(tab);
    mhf_block<HAS_PDT, K>(g, a, lane_trial(g.n_blocks, g.tblocks, g.T), tab);
}

// Split launch, part 1: one workgroup per (curve, 64-trial block, point
// slice); slice s of a curve of n points covers [s*len, min(n, (s+1)*len))
// with len = ceil(n / split).  Writes the slice's sums to
// part[((s * n_curves + curve) * 8K + k) * tblocks*64 + trial slot].
// g.n_blocks counts (curve, trial block) pairs in both kernels; the split
// kernel's grid is g.n_blocks * split workgroups.
template <bool HAS_PDT, int K>
__global__ __launch_bounds__(kWave) void mhf_split_kernel(GridArgs g, MhfArgs a) {
    extern __shared__ double2 tab[];
    mhf_fill_table(tab);
    const uint64_t L = xcd_remap(blockIdx.x, g.n_blocks * static_cast<uint64_t>(a.split));
    const int64_t sp = static_cast<int64_t>(L % static_cast<uint64_t>(a.split));
    const int64_t pair = static_cast<int64_t>(L / static_cast<uint64_t>(a.split));
    LaneTrial lt;
    lt.curve = pair / g.tblocks;
    const int64_t slot = (pair % g.tblocks) * kWave + threadIdx.x;  // raw trial index
    lt.active = slot < g.T;
    lt.trial = lt.active ? slot : g.T - 1;
    const MhfTrial tr = mhf_trial(g, a, lt);
    const int64_t p0 = g.off[lt.curve];
    const int64_t n = g.off[lt.curve + 1] - p0;
    const int64_t len = (n + a.split - 1) / a.split;
    const int64_t b = sp * len < n ? sp * len : n;
    const int64_t e = b + len < n ? b + len : n;
    MhfAcc<K> acc;
#pragma unroll
    for (int k = 0; k < MhfAcc<K>::N; ++k) acc.v[k] = 0.0;
    mhf_accumulate<HAS_PDT, K>(a.pt + p0 + b, e - b, tr, tab, acc);
    const int64_t ts = g.tblocks * kWave;
    const int64_t n_curves = static_cast<int64_t>(g.n_blocks) / g.tblocks;
    double* o = a.part + (sp * n_curves + lt.curve) * MhfAcc<K>::N * ts + slot;
#pragma unroll
    for (int k = 0; k < MhfAcc<K>::N; ++k) o[k * ts] = acc.v[k];
}

// Split launch, part 2: one workgroup per (curve, 64-trial block); adds the
// slices' sums in slice order and finishes as mhf_block.
template <int K>
__global__ __launch_bounds__(kWave) void mhf_finish_kernel(GridArgs g, MhfArgs a) {
    const uint64_t L = xcd_remap(blockIdx.x, g.n_blocks);
    LaneTrial lt;
    lt.curve = static_cast<int64_t>(L / static_cast<uint64_t>(g.tblocks));
    const int64_t slot = static_cast<int64_t>(L % static_cast<uint64_t>(g.tblocks)) * kWave +
                         threadIdx.x;
    lt.active = slot < g.T;
    lt.trial = lt.active ? slot : g.T - 1;
    const int64_t ts = g.tblocks * kWave;
    const int64_t n_curves = static_cast<int64_t>(g.n_blocks) / g.tblocks;
    MhfAcc<K> acc;
#pragma unroll
    for (int k = 0; k < MhfAcc<K>::N; ++k) acc.v[k] = 0.0;
    for (int sp = 0; sp < a.split; ++sp) {
        const double* p = a.part + (sp * n_curves + lt.curve) * MhfAcc<K>::N * ts + slot;
#pragma unroll
        for (int k = 0; k < MhfAcc<K>::N; ++k) acc.v[k] += p[k * ts];
    }
    mhf_finish<K>(g, a, lt, mhf_trial(g, a, lt), acc);
}

// Persistent: each wave pulls (curve, 64-trial block) items from the queue.
template <bool HAS_PDT, int K>
__global__ __launch_bounds__(kWave) void mhf_persistent(GridArgs g, MhfArgs a, WorkQueue q) {
    extern __shared__ double2 tab[];
    if constexpr (K > 0) mhf_fill_table(tab);
    QueueState seen = queue_start();
    for (Claim cl = next_claim(q, g, seen); cl.begin < q.total; cl = next_claim(q, g, seen))
        for (int64_t it = cl.begin; it < cl.end; ++it)
            mhf_block<HAS_PDT, K>(g, a, queue_trial(q, g, it), tab);
}

// ---------------------------------------------------------------------------
// Fourier decomposition features — rust/src/fourier.rs:68-230
// (calc_fourier_single): at one period per curve, the models k = 0..5 of
// MHF's design with phi = 2 pi t / period (not folded: the revolution count
// t/period is formed as an exact f64 quotient and its fraction taken exactly),
// the BIC choice among the models that factor, and 14 outputs [power, BIC,
// offset, slope, A1, B1, ..., A5, B5] (NaN for fewer than 3 points, an
// invalid period or a failed flat model).
//
// One wave per curve: lanes take points lane, lane+64, ... and accumulate the
// 40 harmonic sums of the MHF kernel, which a butterfly then adds across the
// wave (lane 0's order is the one written).  Unlike the reference, every
// model comes from one Cholesky (see MHF), chi2_k = chi2_0 - sum_{2<=j<2k+2}
// z_j^2 with chi2_0 from the curve's residual pass (clamped at 0 against
// rounding), and the chosen model's coefficients by back substitution in its
// leading block.  Every model is attempted whatever n is (as the reference
// does): with fewer points than parameters its matrix is singular and whether
// its Cholesky fails depends on rounding, in the reference too.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kWave) void fourier_kernel(const int64_t* __restrict__ off,
                                                        const MhfPoint* __restrict__ pt,
                                                        const double* __restrict__ sums,
                                                        const float* __restrict__ periods,
                                                        float* __restrict__ out) {
    constexpr int K = kMhfMaxK, M = 2 * K, NP = 2 + M;
    extern __shared__ double2 tab[];
    mhf_fill_table(tab);
    const int64_t c = blockIdx.x;
    const int64_t p0 = off[c];
    const int64_t n = off[c + 1] - p0;
    const float period = periods[c];
    const double P = static_cast<double>(period);
    const double inv_p = 1.0 / P;
    MhfAcc<K> acc;
#pragma unroll
    for (int k = 0; k < MhfAcc<K>::N; ++k) acc.v[k] = 0.0;
    for (int64_t i = threadIdx.x; i < n; i += kWave) {
        const MhfPoint r = pt[p0 + i];
        // RN(t / P) by one correction step of the reciprocal (as fold_phase).
        const double q0 = r.t * inv_p;
        const double q = __builtin_fma(__builtin_fma(-q0, P, r.t), inv_p, q0);
        double cs[M + 1], sn[M + 1];
        mhf_sincos_rev(__builtin_amdgcn_fract(q), tab, sn[1], cs[1]);
        cs[0] = 1.0;
        sn[0] = 0.0;
        const double two_c1 = 2.0 * cs[1];
#pragma unroll
        for (int m = 2; m <= M; ++m) {
            cs[m] = __builtin_fma(two_c1, cs[m - 1], -cs[m - 2]);
            sn[m] = __builtin_fma(two_c1, sn[m - 1], -sn[m - 2]);
        }
#pragma unroll
        for (int m = 1; m <= M; ++m) {
            acc.C(m) = __builtin_fma(r.w, cs[m], acc.C(m));
            acc.S(m) = __builtin_fma(r.w, sn[m], acc.S(m));
        }
#pragma unroll
        for (int h = 1; h <= K; ++h) {
            acc.PC(h) = __builtin_fma(r.wx, cs[h], acc.PC(h));
            acc.PS(h) = __builtin_fma(r.wx, sn[h], acc.PS(h));
            acc.YC(h) = __builtin_fma(r.wy, cs[h], acc.YC(h));
            acc.YS(h) = __builtin_fma(r.wy, sn[h], acc.YS(h));
        }
    }
#pragma unroll
    for (int k = 0; k < MhfAcc<K>::N; ++k)
#pragma unroll
        for (int o2 = 32; o2 >= 1; o2 >>= 1) acc.v[k] += __shfl_xor(acc.v[k], o2);

    const double* u = sums + c * kMhfSums;
    const double sw = u[0], swx = u[1], swxx = u[2], swy = u[3], swxy = u[4], ln_n = u[5];
    const double chi2_0 = u[6];
    auto A = [&](int i, int j) -> double {
        if (i < 2) return i == 0 ? sw : (j == 0 ? swx : swxx);
        const int hi = i >> 1;
        const bool si = i & 1;
        if (j == 0) return si ? acc.S(hi) : acc.C(hi);
        if (j == 1) return si ? acc.PS(hi) : acc.PC(hi);
        const int hj = j >> 1;
        const bool sj = j & 1;
        const int d = hi - hj;
        const double cd = d == 0 ? sw : acc.C(d);
        const double sd = d == 0 ? 0.0 : acc.S(d);
        if (!si && !sj) return 0.5 * (cd + acc.C(hi + hj));
        if (si && sj) return 0.5 * (cd - acc.C(hi + hj));
        if (si) return 0.5 * (acc.S(hi + hj) + sd);
        return 0.5 * (acc.S(hi + hj) - sd);
    };
    auto B = [&](int i) -> double {
        if (i < 2) return i == 0 ? swy : swxy;
        return (i & 1) ? acc.YS(i >> 1) : acc.YC(i >> 1);
    };
    double L[NP * (NP + 1) / 2];
    double rd[NP], z[NP];
    int bad = NP;
#pragma unroll
    for (int i = 0; i < NP; ++i) {
#pragma unroll
        for (int j = 0; j <= i; ++j) {
            double v = A(i, j);
#pragma unroll
            for (int k = 0; k < j; ++k)
                v = __builtin_fma(-L[i * (i + 1) / 2 + k], L[j * (j + 1) / 2 + k], v);
            if (j < i) {
                L[i * (i + 1) / 2 + j] = v * rd[j];
            } else {
                bad = (v <= 0.0 && bad == NP) ? i : bad;  // a NaN pivot carries on
                const double d = sqrt(v);
                L[i * (i + 1) / 2 + i] = d;
                rd[i] = 1.0 / d;
            }
        }
        double v = B(i);
#pragma unroll
        for (int k = 0; k < i; ++k) v = __builtin_fma(-L[i * (i + 1) / 2 + k], z[k], v);
        z[i] = v * rd[i];
    }
    if (threadIdx.x != 0) return;
    float* o = out + c * kFourierFeatures;
    if (n < 3 || !(period > 0.0f) || !isfinite(period) || bad < 2) {
        for (int j = 0; j < kFourierFeatures; ++j) o[j] = __builtin_nanf("");
        return;
    }
    // BIC choice (fourier.rs:184-196): the first minimum, from +inf, over the
    // models that factor (a NaN BIC never wins).
    int best = 0;
    double best_bic = INFINITY, best_gain = 0.0, q = 0.0;
    if (chi2_0 + ln_n * 2.0 < best_bic) best_bic = chi2_0 + ln_n * 2.0;
#pragma unroll
    for (int k = 1; k <= K; ++k) {
        q = __builtin_fma(z[2 * k], z[2 * k], q);
        q = __builtin_fma(z[2 * k + 1], z[2 * k + 1], q);
        const double gain = q < chi2_0 ? q : chi2_0;  // chi2_0 - chi2_k, chi2_k >= 0
        const double bic = (chi2_0 - gain) + ln_n * static_cast<double>(2 + 2 * k);
        if (bad >= 2 * k + 2 && bic < best_bic) {
            best_bic = bic;
            best = k;
            best_gain = gain;
        }
    }
    // Back substitution L^T beta = z in the chosen model's leading block
    // (fourier.rs:55-60).
    const int np = 2 + 2 * best;
    double beta[NP];
#pragma unroll
    for (int i = NP - 1; i >= 0; --i) {
        double v = z[i];
#pragma unroll
        for (int k = i + 1; k < NP; ++k)
            if (k < np) v = __builtin_fma(-L[k * (k + 1) / 2 + i], beta[k], v);
        beta[i] = i < np ? v * rd[i] : 0.0;
    }
    o[0] = static_cast<float>(chi2_0 > 0.0 ? best_gain / chi2_0 : 0.0);
    o[1] = static_cast<float>(best_bic);
#pragma unroll
    for (int j = 0; j < NP; ++j) o[2 + j] = static_cast<float>(beta[j]);
}

// ---------------------------------------------------------------------------
// BLS — rust/src/bls.rs:20-134.
// LDS: [slot 0..nb][2][64] f32 (W, Y).  Binning adds bin k into slot k+1; the
// prefix pass then turns slot k into prefix[k] in place (bls.rs:85-91).  The
// (duration, phase) search (bls.rs:100-127) is serial per lane, with uniform
// loop bounds, so lanes never diverge on them.
//
// Exact max without a division per window: windows compare by the exact
// cross products of their (num, den) pairs, and the best pair is divided once.
// ---------------------------------------------------------------------------
// Waves per BLS workgroup.  Measured at 100 curves x 1000 periods (ms for
// 256 / 1024 / 4096 / 8192 points): 1 wave 0.18 / 0.26 / 0.58 / 0.99, 2 waves
// 0.12 / 0.20 / 0.53 / 0.93, 4 waves 0.12 / 0.23 / 0.66 / 1.26 (the kernel's
// 144 VGPRs hold 12 waves per CU, so 4 waves cost workgroups per CU).
constexpr int kBlsWaves = 2;

// The largest float <= wt * 2^-150 for wt > 0 (the product is exact in f64).
__device__ __forceinline__ float bls_window_floor(float wt) {
    const double x = static_cast<double>(wt) * 0x1p-150;
    float f = static_cast<float>(x);
    if (static_cast<double>(f) > x) f = __builtin_bit_cast(float, __builtin_bit_cast(uint32_t, f) - 1u);
    return f;
}

// W waves share one (curve, 64-trial block): wave 0 bins the points and forms
// the prefix sums (serial per trial, as Rust), then every wave searches the
// phases p = wave, wave + W, ... and the waves' best pairs are merged.  The
// LDS is the same for any W, so W > 1 adds search issue slots at the same
// occupancy.
template <bool HAS_PDT, int G, int W = 1>
__device__ __forceinline__ void bls_block(const GridArgs& g, const BlsArgs& a, const LaneTrial lt,
                                          const int lane = static_cast<int>(threadIdx.x),
                                          const int wave = 0) {
    extern __shared__ float fl[];
    const int nb = a.n_bins;
    float* col = fl + lane;  // slot s: W at plane 0, Y at plane 1
    if (wave == 0) {
    for (int b = 0; b < 2 * (nb + 1); ++b) col[b * kWave] = 0.0f;
    const FoldTrial f = make_fold_trial(g.periods[lt.trial / g.nd], g.pdts[lt.trial % g.nd]);
    const int64_t p0 = g.off[lt.curve];
    const int64_t n = g.off[lt.curve + 1] - p0;
    const float* __restrict__ w = a.w + p0;
    const float* __restrict__ yw = a.yw + p0;
    const float nf = static_cast<float>(nb);
    const uint32_t nm1 = static_cast<uint32_t>(nb - 1);
    constexpr uint32_t kSlotBytes = 2 * kPlaneBytes;
    float* col1 = col + 2 * kWave;  // slot 1

    // Chunks of 16 points: measured 6% faster than 8 for BLS at 8192 points
    // (slower for AOV and FPW, which keep the default).
    scan_points_rmw<HAS_PDT, G, 2, 16, false>(
        g.t64 + p0, g.tt + p0, n, f, [&](int64_t i) { return F2{w[i], yw[i]}; }, col1,
        [&](const auto& kb, const auto& v, auto& off, auto& val) {
            constexpr int K = group_size_v<decltype(kb)>;
#pragma unroll
            for (int j = 0; j < K; ++j) {
                off[j] = kb[j] * kSlotBytes;
                val[0][j] = v[j].a;
                val[1][j] = v[j].b;
            }
        },
        // Bins clamped to n - 1 (bls.rs:79).
        fast_bins(lane_bin_rule(g.tmax, lt.curve, f, nb, false),
                  [&](float ph) { return mul_bin(ph, nf, nm1); }));
    // Prefix sums in place (bls.rs:85-91): slot k+1 <- slot k + bin k.
    // Eight slots are read before any is added, so the LDS latency is paid
    // once per eight slots rather than once per slot.
    {
        float pw = 0.0f, py = 0.0f;
        int k = 1;
        for (; k + 7 <= nb; k += 8) {
            float bw[8], by[8];
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                bw[j] = col[(2 * (k + j)) * kWave];
                by[j] = col[(2 * (k + j) + 1) * kWave];
            }
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                pw = pw + bw[j];
                py = py + by[j];
                col[(2 * (k + j)) * kWave] = pw;
                col[(2 * (k + j) + 1) * kWave] = py;
            }
        }
        for (; k <= nb; ++k) {
            pw = pw + col[(2 * k) * kWave];
            py = py + col[(2 * k + 1) * kWave];
            col[(2 * k) * kWave] = pw;
            col[(2 * k + 1) * kWave] = py;
        }
    }
    }  // wave 0
    if constexpr (W > 1) __syncthreads();
    const float wt = col[(2 * nb) * kWave];
    const float ytot = col[(2 * nb + 1) * kWave];
    double bn = 0.0, bd = 1.0;
    if (wt > 0.0f) {
        // Window test (bls.rs:118-119), 0 < fl(r/wt) < 1: for positive wt,
        // fl(r/wt) < 1 iff r < wt, i.e. t = fl(wt - r) > 0, and fl(r/wt) > 0
        // iff r > wt * 2^-150 (the quotient must round up to the smallest
        // subnormal), i.e. r > rlo, the largest float <= wt * 2^-150.  For
        // wt = inf, rlo = inf and no window passes, as in Rust.
        const float rlo = bls_window_floor(wt);
        // Rust keeps the largest fl(num/den) over the windows that pass.
        // Rounding is monotone, so that is fl(num/den) of the window with the
        // largest exact quotient, and quotients compare exactly by cross
        // products in f64 (a product of two floats has at most 48 significant
        // bits).  The running best is the pair (bn, bd), from 0/1; the one
        // division is at the end.  Special values keep the quotient order:
        // den = 0 with num > 0 (+inf) beats every finite quotient and ties
        // another +inf; 0/0, NaN operands and x/inf (0) never compare greater,
        // as their products are NaN or not larger.  No branch depends on the
        // data, so all lanes stay converged.
        auto cand = [&](float r, float s) {
            const float t = wt - r;
            const float num = s * s;
            const float den = r * t;
            const double dn = num, dd = den;
            const bool take = (r > rlo) & (t > 0.0f) & (dn * bd > bn * dd);
            bn = take ? dn : bn;
            bd = take ? dd : bd;
        };
        constexpr int kG = 8;  // windows whose prefix sums are read together
        const float* pw = col;          // W plane of slot k at pw[2k*64]
        const float* py = col + kWave;  // Y plane
        for (int p = wave; p < nb; p += W) {
            const float w0 = pw[(2 * p) * kWave];
            const float y0 = py[(2 * p) * kWave];
            const float wtail = wt - w0;    // wp[nb] - wp[phi]
            const float ytail = ytot - y0;  // yp[nb] - yp[phi]
            // Durations up to nb - p stay inside the bins: r = wp[p+d] - wp[p].
            // Longer ones wrap: r = (wp[nb] - wp[p]) + wp[p+d-nb].
            const int d_in = min(a.d_hi, nb - p);
            int d = a.d_lo;
            for (; d + kG - 1 <= d_in; d += kG) {
                float r[kG], s[kG];
#pragma unroll
                for (int j = 0; j < kG; ++j) {
                    r[j] = pw[(2 * (p + d + j)) * kWave] - w0;
                    s[j] = py[(2 * (p + d + j)) * kWave] - y0;
                }
#pragma unroll
                for (int j = 0; j < kG; ++j) cand(r[j], s[j]);
            }
            for (; d <= d_in; ++d) cand(pw[(2 * (p + d)) * kWave] - w0, py[(2 * (p + d)) * kWave] - y0);
            for (; d + kG - 1 <= a.d_hi; d += kG) {
                float r[kG], s[kG];
#pragma unroll
                for (int j = 0; j < kG; ++j) {
                    r[j] = wtail + pw[(2 * (p + d + j - nb)) * kWave];
                    s[j] = ytail + py[(2 * (p + d + j - nb)) * kWave];
                }
#pragma unroll
                for (int j = 0; j < kG; ++j) cand(r[j], s[j]);
            }
            for (; d <= a.d_hi; ++d)
                cand(wtail + pw[(2 * (p + d - nb)) * kWave], ytail + py[(2 * (p + d - nb)) * kWave]);
        }
    }
    if constexpr (W > 1) {
        // Merge: waves 1.. leave their pairs in slots 1.. (the prefix sums
        // are no longer read), wave 0 keeps the largest quotient.  The pairs
        // are float values, so they survive the round trip exactly.
        __syncthreads();
        if (wave > 0) {
            col[(2 * wave) * kWave] = static_cast<float>(bn);
            col[(2 * wave + 1) * kWave] = static_cast<float>(bd);
        }
        __syncthreads();
        if (wave != 0) return;
#pragma unroll
        for (int v = 1; v < W; ++v) {
            const double vn = col[(2 * v) * kWave];
            const double vd = col[(2 * v + 1) * kWave];
            const bool take = vn * bd > bn * vd;
            bn = take ? vn : bn;
            bd = take ? vd : bd;
        }
    }
    const float best = static_cast<float>(bn) / static_cast<float>(bd);
    if (lt.active) g.out[lt.curve * g.T + lt.trial] = best;
}

// One workgroup of kBlsWaves waves per (curve, 64-trial block),
// hardware-scheduled.
template <bool HAS_PDT, int G>
__global__ __launch_bounds__(kWave * kBlsWaves) void bls_kernel(GridArgs g, BlsArgs a) {
    const int lane = static_cast<int>(threadIdx.x) & (kWave - 1);
    const int wave = static_cast<int>(threadIdx.x) / kWave;
    bls_block<HAS_PDT, G, kBlsWaves>(g, a, lane_trial(g.n_blocks, g.tblocks, g.T, lane), lane,
                                     wave);
}

// Persistent: each workgroup of kBlsWaves waves pulls (curve, 64-trial
// block) items from the queue; wave 0 claims and hands the claim to the
// other waves through LDS.
template <bool HAS_PDT>
__global__ __launch_bounds__(kWave * kBlsWaves) void bls_persistent(GridArgs g, BlsArgs a,
                                                                    WorkQueue q) {
    __shared__ int64_t claim[2];
    const int lane = static_cast<int>(threadIdx.x) & (kWave - 1);
    const int wave = static_cast<int>(threadIdx.x) / kWave;
    QueueState seen = queue_start();  // used by wave 0 only
    for (;;) {
        if (wave == 0) {
            const Claim cl = next_claim(q, g, seen);
            if (lane == 0) {
                claim[0] = cl.begin;
                claim[1] = cl.end;
            }
        }
        __syncthreads();
        const int64_t begin = wave_uniform(claim[0]), end = wave_uniform(claim[1]);
        __syncthreads();  // every wave holds the claim before wave 0 takes the next
        if (begin >= q.total) break;
        for (int64_t it = begin; it < end; ++it) {
            const LaneTrial lt = queue_trial(q, g, it, lane);
            switch (rank_group(q, item_rank(q, g, it))) {
                case 2: bls_block<HAS_PDT, 2, kBlsWaves>(g, a, lt, lane, wave); break;
                case 4: bls_block<HAS_PDT, 4, kBlsWaves>(g, a, lt, lane, wave); break;
                default: bls_block<HAS_PDT, 1, kBlsWaves>(g, a, lt, lane, wave); break;
            }
        }
    }
}

// ---------------------------------------------------------------------------
// BLS with equal weights (BlsArgs.wsum set): every curve's weights w are
// bitwise equal, so the serial f32 sum of the w of a bin's c points is the
// curve's running sum wsum[c - 1], the same additions in the same order.  The
// scan then counts points per bin (16-bit counts, two lanes' per LDS word,
// curves of up to 65536 points; non-returning atomics: no dependency between
// points)
// and keeps only the Y
// plane as a lane-private read-modify-write.  LDS per 64-trial block drops
// from 2 (nb + 1) to nb + 1 + (nb + 1) / 2 words per lane (26.1 -> 19.5 KiB at
// nb = 50), so 8 blocks fit per CU instead of 6.  The weight prefix sums are
// formed from the counts after the scan and held in registers for the search
// (two 32-element vectors read with the wave-uniform window index), the Y
// prefix sums stay in LDS.  One wave per block.  Same value for every trial as
// bls_block: the bins, the prefix sums and the search are the reference's
// f32 operations in the reference's order.
// ---------------------------------------------------------------------------
typedef float f32x32 __attribute__((ext_vector_type(32)));

// Weight prefix sums wp[0..nb] of one lane, nb <= 63.
struct BlsWeightPrefix {
    f32x32 lo, hi;
    __device__ __forceinline__ float operator()(int q) const {
        return q < 32 ? lo[q] : hi[q - 32];
    }
};

template <bool HAS_PDT, int G, int W>
__device__ __forceinline__ void bls_block_ew(const GridArgs& g, const BlsArgs& a,
                                             const LaneTrial lt, const int lane,
                                             const int wave) {
    extern __shared__ float fl[];
    const int nb = a.n_bins;
    float* ycol = fl + lane;  // Y slot s at ycol[s * 64]
    // Counts: lanes l and l + 32 share a word, l < 32 in the low half, so the
    // count of bin k is at ccol[k * 32] (one multiply-add, as Y's slot) and
    // the addend is a per-lane constant.
    uint32_t* ccol = reinterpret_cast<uint32_t*>(fl + (nb + 1) * kWave) + (lane & 31);
    const uint32_t cshift = static_cast<uint32_t>(lane >> 5) << 4;
    const uint32_t cadd = 1u << cshift;
    const int64_t p0 = g.off[lt.curve];
    const int64_t n = g.off[lt.curve + 1] - p0;
    if (wave == 0) {
    for (int b = 0; b <= nb; ++b) ycol[b * kWave] = 0.0f;
    for (int k = 0; k < nb; ++k) ccol[k * 32] = 0u;
    const FoldTrial f = make_fold_trial(g.periods[lt.trial / g.nd], g.pdts[lt.trial % g.nd]);
    const float* __restrict__ yw = a.yw + p0;
    const float nf = static_cast<float>(nb);
    const uint32_t nm1 = static_cast<uint32_t>(nb - 1);

    // Bin k: Y into slot k + 1, one count into the lane's half of count k.
    // (8-point chunks: fewer registers than BLS's 16.)
    scan_points_rmw<HAS_PDT, G, 1, 8, false>(
        g.t64 + p0, g.tt + p0, n, f, [&](int64_t i) { return yw[i]; }, ycol + kWave,
        [&](const auto& kb, const auto& v, auto& off, auto& val) {
            constexpr int K = group_size_v<decltype(kb)>;
#pragma unroll
            for (int j = 0; j < K; ++j) {
                const uint32_t k = kb[j];
                off[j] = k * kPlaneBytes;
                val[0][j] = v[j];
                __hip_atomic_fetch_add(ccol + k * 32, cadd, __ATOMIC_RELAXED,
                                       __HIP_MEMORY_SCOPE_WAVEFRONT);
            }
        },
        // Bins clamped to n - 1 (bls.rs:79).
        fast_bins(lane_bin_rule(g.tmax, lt.curve, f, nb, false),
                  [&](float ph) { return mul_bin(ph, nf, nm1); }));
    // Y prefix sums in place (bls.rs:85-91): slot k+1 <- slot k + bin k.
    {
        float py = 0.0f;
        int k = 1;
        for (; k + 7 <= nb; k += 8) {
            float by[8];
#pragma unroll
            for (int j = 0; j < 8; ++j) by[j] = ycol[(k + j) * kWave];
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                py = py + by[j];
                ycol[(k + j) * kWave] = py;
            }
        }
        for (; k <= nb; ++k) {
            py = py + ycol[k * kWave];
            ycol[k * kWave] = py;
        }
    }
    }  // wave 0
    if constexpr (W > 1) __syncthreads();
    // Weight of bin k: c_k points of equal weight, wsum[c_k - 1] (0 for an
    // empty bin); prefix sums wp[k + 1] = wp[k] + W_k in bin order.  Indices
    // are compile-time (registers); bins k >= nb read word 0 and count 0.
    // (An empty curve has no running sums to read: every count is 0.)
    const float* __restrict__ ws = a.wsum + p0;
    BlsWeightPrefix wp;
    float pw = 0.0f;
    // Sixteen bins at a time: their count reads and table reads are in flight
    // together.  `total` sums the decoded counts.
    uint32_t total = 0;
    auto bin_weights = [&](auto q_tag, f32x32& dst) {
        constexpr int q = decltype(q_tag)::value;  // quarter of the 64 bins
        constexpr int k0 = q * 16, j0 = (q & 1) * 16;
        float wk[16];
#pragma unroll
        for (int j = 0; j < 16; ++j) {
            const int k = k0 + j;
            const bool in = k < nb;
            const uint32_t word = ccol[(in ? k : 0) * 32];
            const uint32_t c = in ? ((word >> cshift) & 0xffffu) : 0u;
            const float x = n > 0 ? ws[c > 0 ? c - 1 : 0] : 0.0f;
            wk[j] = c > 0 ? x : 0.0f;
            total += c;
        }
#pragma unroll
        for (int j = 0; j < 16; ++j) {
            dst[j0 + j] = pw;  // wp[k0 + j]
            pw = pw + wk[j];
        }
    };
    bin_weights(std::integral_constant<int, 0>{}, wp.lo);
    bin_weights(std::integral_constant<int, 1>{}, wp.lo);
    bin_weights(std::integral_constant<int, 2>{}, wp.hi);
    bin_weights(std::integral_constant<int, 3>{}, wp.hi);
    // For k > nb the stored values are wp[nb] (the empty bins add 0), so
    // wp(nb) is the total weight.  A 16-bit count overflows only for a curve
    // of exactly 65536 points all in one bin; the decoded counts then no
    // longer sum to n, and the statistic is 0 (every window holds all the
    // weight or none, so none passes 0 < r/W < 1).
    const bool overflow = total != static_cast<uint32_t>(n);
    const float wt = overflow ? 0.0f : wp(nb);
    const float ytot = ycol[nb * kWave];
    double bn = 0.0, bd = 1.0;
    if (wt > 0.0f) {
        // As bls_block: exact maximum of fl(s^2 / (r (W - r))) over the
        // windows with 0 < fl(r/W) < 1 (bls.rs:100-127).
        const float rlo = bls_window_floor(wt);
        auto cand = [&](float r, float sv) {
            const float t = wt - r;
            const float num = sv * sv;
            const float den = r * t;
            const double dn = num, dd = den;
            const bool take = (r > rlo) & (t > 0.0f) & (dn * bd > bn * dd);
            bn = take ? dn : bn;
            bd = take ? dd : bd;
        };
        constexpr int kG = 8;
        for (int p = wave; p < nb; p += W) {
            const float w0 = wp(p);
            const float y0 = ycol[p * kWave];
            const float wtail = wt - w0;
            const float ytail = ytot - y0;
            const int d_in = min(a.d_hi, nb - p);
            int d = a.d_lo;
            for (; d + kG - 1 <= d_in; d += kG) {
                float r[kG], sv[kG];
#pragma unroll
                for (int j = 0; j < kG; ++j) {
                    r[j] = wp(p + d + j) - w0;
                    sv[j] = ycol[(p + d + j) * kWave] - y0;
                }
#pragma unroll
                for (int j = 0; j < kG; ++j) cand(r[j], sv[j]);
            }
            for (; d <= d_in; ++d) cand(wp(p + d) - w0, ycol[(p + d) * kWave] - y0);
            for (; d + kG - 1 <= a.d_hi; d += kG) {
                float r[kG], sv[kG];
#pragma unroll
                for (int j = 0; j < kG; ++j) {
                    r[j] = wtail + wp(p + d + j - nb);
                    sv[j] = ytail + ycol[(p + d + j - nb) * kWave];
                }
#pragma unroll
                for (int j = 0; j < kG; ++j) cand(r[j], sv[j]);
            }
            for (; d <= a.d_hi; ++d)
                cand(wtail + wp(p + d - nb), ytail + ycol[(p + d - nb) * kWave]);
        }
    }
    if constexpr (W > 1) {
        // Merge as bls_block: waves 1.. leave their pairs in Y slots 2v-2,
        // 2v-1 (no longer read once every wave has searched), wave 0 keeps
        // the largest quotient.
        __syncthreads();
        if (wave > 0) {
            ycol[(2 * wave - 2) * kWave] = static_cast<float>(bn);
            ycol[(2 * wave - 1) * kWave] = static_cast<float>(bd);
        }
        __syncthreads();
        if (wave != 0) return;
#pragma unroll
        for (int v = 1; v < W; ++v) {
            const double vn = ycol[(2 * v - 2) * kWave];
            const double vd = ycol[(2 * v - 1) * kWave];
            const bool take = vn * bd > bn * vd;
            bn = take ? vn : bn;
            bd = take ? vd : bd;
        }
    }
    const float best = static_cast<float>(bn) / static_cast<float>(bd);
    if (lt.active) g.out[lt.curve * g.T + lt.trial] = best;
}

// Waves per equal-weight BLS workgroup: wave 0 scans, all search.
constexpr int kBlsEwWaves = 2;

// One workgroup per (curve, 64-trial block), hardware-scheduled.
template <bool HAS_PDT, int G>
__global__ __launch_bounds__(kWave * kBlsEwWaves)
__attribute__((amdgpu_waves_per_eu(2 * kBlsEwWaves))) void bls_ew_kernel(GridArgs g, BlsArgs a) {
    const int lane = static_cast<int>(threadIdx.x) & (kWave - 1);
    // (wave-uniform: the search indexes registers by p = wave + 2k)
    const int wave = __builtin_amdgcn_readfirstlane(static_cast<int>(threadIdx.x) / kWave);
    bls_block_ew<HAS_PDT, G, kBlsEwWaves>(g, a, lane_trial(g.n_blocks, g.tblocks, g.T, lane),
                                          lane, wave);
}

// Persistent: each workgroup pulls (curve, 64-trial block) items from the
// queue; wave 0 claims and hands the claim to the other waves through LDS.
template <bool HAS_PDT>
__global__ __launch_bounds__(kWave * kBlsEwWaves)
__attribute__((amdgpu_waves_per_eu(2 * kBlsEwWaves))) void bls_ew_persistent(GridArgs g, BlsArgs a,
                                                                WorkQueue q) {
    __shared__ int64_t claim[2];
    const int lane = static_cast<int>(threadIdx.x) & (kWave - 1);
    // (wave-uniform: the search indexes registers by p = wave + 2k)
    const int wave = __builtin_amdgcn_readfirstlane(static_cast<int>(threadIdx.x) / kWave);
    QueueState seen = queue_start();  // used by wave 0 only
    for (;;) {
        if (wave == 0) {
            const Claim cl = next_claim(q, g, seen);
            if (lane == 0) {
                claim[0] = cl.begin;
                claim[1] = cl.end;
            }
        }
        __syncthreads();
        const int64_t begin = wave_uniform(claim[0]), end = wave_uniform(claim[1]);
        __syncthreads();  // every wave holds the claim before wave 0 takes the next
        if (begin >= q.total) break;
        for (int64_t it = begin; it < end; ++it) {
            const LaneTrial lt = queue_trial(q, g, it, lane);
            switch (rank_group(q, item_rank(q, g, it))) {
                case 2: bls_block_ew<HAS_PDT, 2, kBlsEwWaves>(g, a, lt, lane, wave); break;
                case 4: bls_block_ew<HAS_PDT, 4, kBlsEwWaves>(g, a, lt, lane, wave); break;
                default: bls_block_ew<HAS_PDT, 1, kBlsEwWaves>(g, a, lt, lane, wave); break;
            }
        }
    }
}

// ---------------------------------------------------------------------------
// Host-side launchers.  `grp` selects the lds_rmw group size (1, 2 or 4; CE
// also 0 = LDS atomics).  With q == nullptr the grid kernel runs one
// workgroup per (curve, trial block); otherwise the persistent kernel runs
// as many single-wave workgroups as fit on the device at once.
// ---------------------------------------------------------------------------
static inline dim3 grid_of(const GridArgs& g) { return dim3(static_cast<uint32_t>(g.n_blocks)); }

// Resident single-wave workgroups of `kernel` with `lds` bytes, device-wide.
template <class K>
static uint32_t persistent_grid(K kernel, size_t lds, int waves) {
    int dev = 0, cus = 0, per_cu = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 256;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
        cus = 256;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, kWave * waves, lds) !=
            hipSuccess ||
        per_cu < 1)
        per_cu = 1;
    return static_cast<uint32_t>(cus) * static_cast<uint32_t>(per_cu);
}

template <class K, class Args>
void launch_grid_kernel(K kernel, const GridArgs& g, const Args& a, size_t lds, hipStream_t s,
                        int waves = 1) {
    hipLaunchKernelGGL(kernel, grid_of(g), dim3(kWave * waves), lds, s, g, a);
}

template <class K, class Args>
void launch_persistent(K kernel, const GridArgs& g, const Args& a, size_t lds,
                       const WorkQueue& q, hipStream_t s, int waves = 1) {
    hipLaunchKernelGGL(kernel, dim3(persistent_grid(kernel, lds, waves)), dim3(kWave * waves),
                       lds, s, g, a, q);
}

// Plain launches take the binning variant `grp`; persistent launches carry
// one per length bucket in the queue (CE: grp < 0 when every bucket's is
// the packed one).
template <bool P, bool E>
void ce_g(const GridArgs& g, const CeArgs& a, size_t lds, int grp, const WorkQueue* q,
          hipStream_t s) {
    if (q && E && grp < 0) launch_persistent(ce_persistent<P, E, true>, g, a, lds, *q, s);
    else if (q) launch_persistent(ce_persistent<P, E>, g, a, lds, *q, s);
    else if (!E) launch_grid_kernel(ce_kernel<P, false, 1>, g, a, lds, s);
    else if (grp < 0) launch_grid_kernel(ce_kernel<P, E, -1>, g, a, lds, s);
    else if (grp == 0) launch_grid_kernel(ce_kernel<P, E, 0>, g, a, lds, s);
    else if (grp >= 4) launch_grid_kernel(ce_kernel<P, E, 4>, g, a, lds, s);
    else if (grp == 2) launch_grid_kernel(ce_kernel<P, E, 2>, g, a, lds, s);
    else launch_grid_kernel(ce_kernel<P, E, 1>, g, a, lds, s);
}

template <bool P, bool E>
void aov_g(const GridArgs& g, const AovArgs& a, size_t lds, int grp, const WorkQueue* q,
           hipStream_t s) {
    if (q) launch_persistent(aov_persistent<P, E>, g, a, lds, *q, s);
    else if (!E) launch_grid_kernel(aov_kernel<P, false, 1>, g, a, lds, s);
    else if (grp >= 4) launch_grid_kernel(aov_kernel<P, E, 4>, g, a, lds, s);
    else if (grp == 2) launch_grid_kernel(aov_kernel<P, E, 2>, g, a, lds, s);
    else launch_grid_kernel(aov_kernel<P, E, 1>, g, a, lds, s);
}

template <bool P>
void fpw_g(const GridArgs& g, const FpwArgs& a, size_t lds, int grp, const WorkQueue* q,
           hipStream_t s) {
    if (q) launch_persistent(fpw_persistent<P>, g, a, lds, *q, s);
    else if (grp >= 4) launch_grid_kernel(fpw_kernel<P, 4>, g, a, lds, s);
    else if (grp == 2) launch_grid_kernel(fpw_kernel<P, 2>, g, a, lds, s);
    else launch_grid_kernel(fpw_kernel<P, 1>, g, a, lds, s);
}

template <bool P, bool E, int MAXM>
void vn_g2(const GridArgs& g, const VnArgs& a, size_t lds, int grp, const WorkQueue* q,
           hipStream_t s) {
    if (q) launch_persistent(vn_persistent<P, E, MAXM>, g, a, lds, *q, s);
    else if (!E) launch_grid_kernel(vn_kernel<P, false, 1, MAXM>, g, a, lds, s);
    else if (grp < 0) launch_grid_kernel(vn_kernel<P, E, -1, MAXM>, g, a, lds, s);
    else launch_grid_kernel(vn_kernel<P, E, 0, MAXM>, g, a, lds, s);
}

template <bool P, bool E>
void vn_g(const GridArgs& g, const VnArgs& a, size_t lds, int grp, const WorkQueue* q,
          hipStream_t s) {
    const int nm = a.h.n_mag;
    if (nm <= 8) vn_g2<P, E, 8>(g, a, lds, grp, q, s);
    else if (nm <= 16) vn_g2<P, E, 16>(g, a, lds, grp, q, s);
    else if (nm <= 24) vn_g2<P, E, 24>(g, a, lds, grp, q, s);
    else vn_g2<P, E, kVnMaxMag>(g, a, lds, grp, q, s);
}

template <bool P>
void mf_g(const GridArgs& g, const MfArgs& a, size_t lds, int grp, const WorkQueue* q,
          hipStream_t s) {
    if (q) launch_persistent(mf_persistent<P>, g, a, lds, *q, s);
    else if (grp >= 4) launch_grid_kernel(mf_kernel<P, 4>, g, a, lds, s);
    else if (grp == 2) launch_grid_kernel(mf_kernel<P, 2>, g, a, lds, s);
    else launch_grid_kernel(mf_kernel<P, 1>, g, a, lds, s);
}

template <bool P>
void bls_g(const GridArgs& g, const BlsArgs& a, size_t lds, int grp, const WorkQueue* q,
           hipStream_t s) {
    if (a.wsum) {
        if (q) launch_persistent(bls_ew_persistent<P>, g, a, lds, *q, s, kBlsEwWaves);
        else if (grp >= 4) launch_grid_kernel(bls_ew_kernel<P, 4>, g, a, lds, s, kBlsEwWaves);
        else if (grp == 2) launch_grid_kernel(bls_ew_kernel<P, 2>, g, a, lds, s, kBlsEwWaves);
        else launch_grid_kernel(bls_ew_kernel<P, 1>, g, a, lds, s, kBlsEwWaves);
        return;
    }
    if (q) launch_persistent(bls_persistent<P>, g, a, lds, *q, s, kBlsWaves);
    else if (grp >= 4) launch_grid_kernel(bls_kernel<P, 4>, g, a, lds, s, kBlsWaves);
    else if (grp == 2) launch_grid_kernel(bls_kernel<P, 2>, g, a, lds, s, kBlsWaves);
    else launch_grid_kernel(bls_kernel<P, 1>, g, a, lds, s, kBlsWaves);
}

hipError_t launch_ce(const GridArgs& g, const CeArgs& a, size_t lds, int grp, const WorkQueue* q,
                     hipStream_t s) {
    const bool ext1 = a.phase_ext == 1 && a.mag_ext == 1;
    if (g.has_pdt) ext1 ? ce_g<true, true>(g, a, lds, grp, q, s) : ce_g<true, false>(g, a, lds, grp, q, s);
    else ext1 ? ce_g<false, true>(g, a, lds, grp, q, s) : ce_g<false, false>(g, a, lds, grp, q, s);
    return hipGetLastError();
}

hipError_t launch_aov(const GridArgs& g, const AovArgs& a, size_t lds, int grp,
                      const WorkQueue* q, hipStream_t s) {
    const bool ext1 = a.ext == 1;
    if (g.has_pdt) ext1 ? aov_g<true, true>(g, a, lds, grp, q, s) : aov_g<true, false>(g, a, lds, grp, q, s);
    else ext1 ? aov_g<false, true>(g, a, lds, grp, q, s) : aov_g<false, false>(g, a, lds, grp, q, s);
    return hipGetLastError();
}

hipError_t launch_ls(const GridArgs& g, const LsArgs& a, const WorkQueue* q, hipStream_t s) {
    if (g.has_pdt) {
        if (q) a.fast ? launch_persistent(ls_persistent<true, true>, g, a, 0, *q, s)
                      : launch_persistent(ls_persistent<true, false>, g, a, 0, *q, s);
        else a.fast ? launch_grid_kernel(ls_kernel<true, true>, g, a, 0, s)
                    : launch_grid_kernel(ls_kernel<true, false>, g, a, 0, s);
    } else {
        if (q) a.fast ? launch_persistent(ls_persistent<false, true>, g, a, 0, *q, s)
                      : launch_persistent(ls_persistent<false, false>, g, a, 0, *q, s);
        else a.fast ? launch_grid_kernel(ls_kernel<false, true>, g, a, 0, s)
                    : launch_grid_kernel(ls_kernel<false, false>, g, a, 0, s);
    }
    return hipGetLastError();
}

template <bool P, int K>
void mhf_g2(const GridArgs& g, const MhfArgs& a, const WorkQueue* q, hipStream_t s) {
    const size_t lds = K > 0 ? kMhfLds : 0;
    if (q) {
        launch_persistent(mhf_persistent<P, K>, g, a, lds, *q, s);
    } else if (K > 0 && a.split > 1) {
        const uint32_t nb = static_cast<uint32_t>(g.n_blocks * static_cast<uint64_t>(a.split));
        hipLaunchKernelGGL((mhf_split_kernel<P, K>), dim3(nb), dim3(kWave), lds, s, g, a);
        hipLaunchKernelGGL((mhf_finish_kernel<K>), grid_of(g), dim3(kWave), 0, s, g, a);
    } else {
        launch_grid_kernel(mhf_kernel<P, K>, g, a, lds, s);
    }
}

template <bool P>
void mhf_g(const GridArgs& g, const MhfArgs& a, const WorkQueue* q, hipStream_t s) {
    switch (a.max_k) {
        case 0: mhf_g2<P, 0>(g, a, q, s); break;
        case 1: mhf_g2<P, 1>(g, a, q, s); break;
        case 2: mhf_g2<P, 2>(g, a, q, s); break;
        case 3: mhf_g2<P, 3>(g, a, q, s); break;
        case 4: mhf_g2<P, 4>(g, a, q, s); break;
        default: mhf_g2<P, kMhfMaxK>(g, a, q, s); break;
    }
}

hipError_t launch_mhf(const GridArgs& g, const MhfArgs& a, const WorkQueue* q, hipStream_t s) {
    if (a.max_k < 0 || a.max_k > kMhfMaxK) return hipErrorInvalidValue;
    if (a.split > 1 && (q || !a.part || g.n_blocks * static_cast<uint64_t>(a.split) > 0xffffffffull))
        return hipErrorInvalidValue;
    g.has_pdt ? mhf_g<true>(g, a, q, s) : mhf_g<false>(g, a, q, s);
    return hipGetLastError();
}

hipError_t launch_fourier(const int64_t* off, const MhfPoint* pt, const double* sums,
                          const float* periods, float* out, int64_t n_curves, hipStream_t s) {
    if (n_curves <= 0) return hipSuccess;
    if (n_curves > 0xffffffffll) return hipErrorInvalidValue;
    hipLaunchKernelGGL(fourier_kernel, dim3(static_cast<uint32_t>(n_curves)), dim3(kWave), kMhfLds,
                       s, off, pt, sums, periods, out);
    return hipGetLastError();
}

hipError_t launch_vn(const GridArgs& g, const VnArgs& a, size_t lds, int grp, const WorkQueue* q,
                     hipStream_t s) {
    if (a.h.n_mag < 1 || a.h.n_mag > kVnMaxMag) return hipErrorInvalidValue;
    const bool ext1 = a.h.phase_ext == 1 && a.h.mag_ext == 1;
    if (g.has_pdt) ext1 ? vn_g<true, true>(g, a, lds, grp, q, s) : vn_g<true, false>(g, a, lds, grp, q, s);
    else ext1 ? vn_g<false, true>(g, a, lds, grp, q, s) : vn_g<false, false>(g, a, lds, grp, q, s);
    return hipGetLastError();
}

hipError_t launch_mf(const GridArgs& g, const MfArgs& a, size_t lds, int grp, const WorkQueue* q,
                     hipStream_t s) {
    g.has_pdt ? mf_g<true>(g, a, lds, grp, q, s) : mf_g<false>(g, a, lds, grp, q, s);
    return hipGetLastError();
}

hipError_t launch_fpw(const GridArgs& g, const FpwArgs& a, size_t lds, int grp,
                      const WorkQueue* q, hipStream_t s) {
    g.has_pdt ? fpw_g<true>(g, a, lds, grp, q, s) : fpw_g<false>(g, a, lds, grp, q, s);
    return hipGetLastError();
}

hipError_t launch_bls(const GridArgs& g, const BlsArgs& a, size_t lds, int grp,
                      const WorkQueue* q, hipStream_t s) {
    g.has_pdt ? bls_g<true>(g, a, lds, grp, q, s) : bls_g<false>(g, a, lds, grp, q, s);
    return hipGetLastError();
}

}  // namespace pf
