// Uneven batches: many independent trajectories (different atom counts, boxes
// and frame counts) through one RDF pair launch and one RMSD launch, replayed
// as a captured HIP graph.  The reference processes one trajectory per run()
// with a fresh context (crates/traj-py/src/pack.rs:1320-1323) and tunes only
// its chunk size (pack.rs:1366-1405); per trajectory the results here are the
// RdfPlan counts (rdf.rs:882-941, bit-exact through the thresholds of
// wmd_rdf_core.hpp) and the RmsdPlan(frame0, align=true) series
// (rmsd/plan.rs:336-392).
//
// Layout.  The batch store holds trajectory t's frames as SoA planes
// [F_t][3][stride_t] (stride_t = N_t rounded up to 4, padding zero) starting at
// float offset off_t; boxes are [sum F_t][9] row-major.
//
// Work.  Every trajectory is enumerated either by the cell list of the single
// plan (frame x home cell units, shift / rint image modes) or, when its box is
// too small for a stencil of at least 2k+1 cells per side (r_max near or above
// half the box), by brute force: units (frame, block of 64 atoms i) pair the
// block with every atom j > i under the reference's roundf(dx / L) wrap
// (pair_s_checked).  A cell-sorted AoS copy of every frame (store order for
// brute-force trajectories) is built by batched assign / scan / scatter
// kernels over all trajectories at once.
//
// The pair kernel is persistent with a work queue.  A trajectory's units are
// grouped into tasks of `grab` consecutive units (a task never spans two
// trajectories); the task range is split into 8 slices of equal estimated
// cost (slice q is served by workgroups g % 8 == q, i.e. one XCD under
// round-robin dispatch), and each wave takes one task at a time from its
// slice's atomic counter.  A wave whose slice is drained moves on to the next
// slices in turn and helps there, so a slice whose cost was underestimated
// does not leave the other XCDs idle; every wave exits after it has found all
// 8 counters past their slice ends.  Each wave keeps a private u64 LDS
// histogram for the trajectory it is on and merges it into that trajectory's
// global counts when it moves to another trajectory and at the end.
//
// Tuning per trajectory: cells per r_max k (or brute force), the brute-force
// home block (home atoms per unit, a multiple of 64) and the units per task;
// wmd_batch_autotune picks them per class of similar trajectories by timing.
//
// The whole step (clears by fill kernels, cell build, pair kernel, RMSD
// moments and finalize) is captured once into a HIP graph and replayed;
// eager launches give the same results.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <mutex>
#include <sstream>
#include <tuple>

#include "wmd_rdf_core.hpp"

using namespace wmd;

namespace {

constexpr int kBatchBlock = 768;  // 12 waves: two workgroups per CU with u32 wave histograms
constexpr int kBatchWaves = kBatchBlock / 64;
constexpr int kBruteHome = 64;    // default home atoms per brute-force unit
constexpr int kMaxBruteHome = 512;
constexpr int kGrab = 2;          // default units per task
constexpr int kMaxGrab = 64;
constexpr int kSlices = 8;        // queue slices (one per XCD)
constexpr int kClock = 4;         // per slice: first start, drained, last end, tasks taken by helpers

enum { KIND_CELLS = 0, KIND_BRUTE = 1 };

struct BatchTraj {
    long long store_off;  // float offset of frame 0 in the batch store
    long long af0;        // first atom-frame; also the float4 offset of its sorted atoms
    long long start_off;  // u32 offset of frame 0's cell starts [F][ncell + 1] (cells)
    long long count_off;  // u32 offset of frame 0's cell counts [F][ncell] (cells)
    long long unit0;      // first pair unit
    long long task0;      // first task (grab consecutive units of this trajectory)
    long long scan0;      // first scan workgroup (cells)
    int stride, n, n_frames, box0;  // box0: index of frame 0's box and frame flag
    int kind, k, ncx, ncy, ncz, shift_ok, upf;  // upf: units per frame
    int bh, grab;         // brute-force home atoms per unit; units per task
    float cull2;
};

template <class K>
__device__ __forceinline__ int find_last_le(int n, long long x, K key) {
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (key(mid) <= x) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// ---- batched cell build ---------------------------------------------------
template <int PBC>
__global__ __launch_bounds__(kBlock) void batch_assign_kernel(
    const BatchTraj *__restrict__ traj, int n_traj, long long total_af,
    const float *__restrict__ store, const FrameBox *__restrict__ boxes,
    unsigned *__restrict__ cell_of, unsigned *__restrict__ cell_slot,
    unsigned *__restrict__ cell_count, unsigned *__restrict__ image_of,
    unsigned *__restrict__ frame_flag) {
    for (long long g = (long long)blockIdx.x * kBlock + threadIdx.x; g < total_af;
         g += (long long)gridDim.x * kBlock) {
        const int t = find_last_le(n_traj, g, [&](int i) { return traj[i].af0; });
        const BatchTraj &T = traj[t];
        if (T.kind != KIND_CELLS) continue;
        const long long local = g - T.af0;
        const int f = (int)(local / T.n);
        const int i = (int)(local - (long long)f * T.n);
        const float *X = store + T.store_off + (long long)f * 3 * T.stride;
        unsigned img;
        bool far_out;
        const int cid = assign_atom<PBC>(X[i], X[T.stride + i], X[2 * T.stride + i], boxes[T.box0 + f],
                                         T.ncx, T.ncy, T.ncz, img, far_out);
        if (far_out) atomicOr(&frame_flag[T.box0 + f], 1u);
        const int ncell = T.ncx * T.ncy * T.ncz;
        cell_slot[g] = atomicAdd(&cell_count[T.count_off + (long long)f * ncell + cid], 1u);
        cell_of[g] = (unsigned)cid;
        image_of[g] = img;
    }
}

// One workgroup per (cell-list trajectory, frame).
__global__ __launch_bounds__(kBlock) void batch_scan_kernel(const BatchTraj *__restrict__ traj,
                                                            int n_traj,
                                                            const unsigned *__restrict__ count,
                                                            unsigned *__restrict__ start) {
    const long long s = blockIdx.x;
    const int t = find_last_le(n_traj, s, [&](int i) { return traj[i].scan0; });
    const BatchTraj &T = traj[t];
    const int f = (int)(s - T.scan0);
    const int ncell = T.ncx * T.ncy * T.ncz;
    scan_cells(count + T.count_off + (long long)f * ncell,
               start + T.start_off + (long long)f * (ncell + 1), ncell);
}

// sorted[af0 + f n + pos] = (x, y, z, meta); pos = cell position (cells) or
// the atom index (brute force).  Unit selection multiplicities (mA = mB = 1).
__global__ __launch_bounds__(kBlock) void batch_scatter_kernel(
    const BatchTraj *__restrict__ traj, int n_traj, long long total_af,
    const float *__restrict__ store, const unsigned *__restrict__ cell_of,
    const unsigned *__restrict__ cell_slot, const unsigned *__restrict__ image_of,
    const unsigned *__restrict__ start, float4 *__restrict__ sorted) {
    for (long long g = (long long)blockIdx.x * kBlock + threadIdx.x; g < total_af;
         g += (long long)gridDim.x * kBlock) {
        const int t = find_last_le(n_traj, g, [&](int i) { return traj[i].af0; });
        const BatchTraj &T = traj[t];
        const long long local = g - T.af0;
        const int f = (int)(local / T.n);
        const int i = (int)(local - (long long)f * T.n);
        const float *X = store + T.store_off + (long long)f * 3 * T.stride;
        unsigned pos, meta;
        if (T.kind == KIND_CELLS) {
            const int ncell = T.ncx * T.ncy * T.ncz;
            pos = start[T.start_off + (long long)f * (ncell + 1) + cell_of[g]] + cell_slot[g];
            meta = (image_of[g] & 0xFFFFFFu) | 0x11000000u;
        } else {
            pos = (unsigned)i;
            meta = 0x11202020u;  // image bytes m + 32 with m = 0 (unused by brute force)
        }
        sorted[T.af0 + (long long)f * T.n + pos] =
            make_float4(X[i], X[T.stride + i], X[2 * T.stride + i], __uint_as_float(meta));
    }
}

// Fused build, one workgroup per frame of the batch (global frame index s =
// the box index): cell-list trajectories are counting-sorted in LDS
// (build_frame), brute-force trajectories copied in store order.  Used when
// every cell-list trajectory has at most kBuildMaxCells cells.
template <int PBC>
__global__ __launch_bounds__(kBuildBlock) void batch_build_kernel(
    const BatchTraj *__restrict__ traj, int n_traj, const float *__restrict__ store,
    const FrameBox *__restrict__ boxes, unsigned *__restrict__ start, float4 *__restrict__ sorted,
    unsigned *__restrict__ frame_flag) {
    extern __shared__ __attribute__((aligned(16))) char lds[];
    unsigned *part = (unsigned *)lds;
    unsigned *far = part + kBuildBlock;
    unsigned *cnt = far + 4;
    const long long s = blockIdx.x;
    const int t = find_last_le(n_traj, s, [&](int i) { return (long long)traj[i].box0; });
    const BatchTraj T = uload(traj, t);
    const int f = (int)(s - T.box0);
    const float *X = store + T.store_off + (long long)f * 3 * T.stride;
    float4 *dst = sorted + T.af0 + (long long)f * T.n;
    if (T.kind == KIND_CELLS) {
        const int ncell = T.ncx * T.ncy * T.ncz;
        build_frame<PBC>(X, T.stride, T.n, boxes[s], T.ncx, T.ncy, T.ncz, nullptr,
                         start + T.start_off + (long long)f * (ncell + 1), dst, nullptr, frame_flag + s,
                         cnt, part, far);
    } else {
        for (int i = threadIdx.x; i < T.n; i += kBuildBlock)
            dst[i] = make_float4(X[i], X[T.stride + i], X[2 * T.stride + i], __uint_as_float(0x11202020u));
    }
}

// ---- persistent pair kernel -------------------------------------------------
struct BatchPairArgs {
    const BatchTraj *traj;
    int n_traj;
    const float4 *sorted;
    const unsigned *start;
    const unsigned *frame_flag;
    const FrameBox *boxes;
    unsigned long long *counts;  // [n_traj][row_bins] (+ the pass's first bin)
    unsigned long long *queue;   // [kSlices] task counters, zero at launch
    unsigned long long *clock;   // [kSlices][kClock] wall-clock marks, zero at the step's start
    long long slice_lo[kSlices + 1];  // task ranges of the slices
    int row_bins;                // bins of a trajectory's histogram row
    HistParams p;
};

__host__ __device__ inline size_t batch_lds_bytes(int bins, int tab_n) {
    return align16((size_t)kBatchWaves * bins * 4) + align16((size_t)tab_n * 8) +
           align16(((size_t)bins + 1) * 4) + (size_t)kBatchWaves * kRingStride * 16;
}

// Brute-force unit: home block [I0, I0 + bh) of frame f against every j > i.
// Chunks of j inside the home block pair only with the home atoms below them.
template <class C>
__device__ __forceinline__ void brute_unit(const CellGrid &G, int f, int blk, int bh, C &cons) {
    const int lane = threadIdx.x & 63;
    const float nanv = __builtin_nanf("");
    const float4 *fp = G.sorted + (long long)f * G.n;
    const int I0 = blk * bh;
    const int n_home = min(bh, G.n - I0);
    const float4 *hp = fp + I0;
    const FrameBox b = uload(G.boxes, f);
    for (int j0 = I0; j0 < G.n; j0 += 64) {
        const int j = j0 + lane;
        float4 q = make_float4(nanv, nanv, nanv, __uint_as_float(0x11202020u));
        if (j < G.n) q = fp[j];
        const unsigned mj = __float_as_uint(q.w);
        if (j0 < I0 + n_home)  // pairs (i, j) with i < j inside the home block
            cons.template chunk<3, true>(hp, min(n_home, j0 - I0 + 64), q.x, q.y, q.z, mj, j - I0, 0u, f, b);
        else
            cons.template chunk<3, false>(hp, n_home, q.x, q.y, q.z, mj, 0x7FFFFFFF, 0u, f, b);
    }
}

__device__ __forceinline__ unsigned long long clock_now() { return (unsigned long long)wall_clock64(); }

template <int PBC, bool TABLE>
__global__ __launch_bounds__(kBatchBlock) void batch_pair_kernel(BatchPairArgs a) {
    extern __shared__ __attribute__((aligned(16))) char lds[];
    const HistParams &p = a.p;
    const int bins = p.bins;
    unsigned *hist_all = (unsigned *)lds;  // [kBatchWaves][bins] u32, one histogram per wave
    size_t off = align16((size_t)kBatchWaves * bins * 4);
    uint2 *tab2 = (uint2 *)(lds + off);
    off += align16((size_t)(TABLE ? p.tab_n : 0) * 8);
    float *thr = (float *)(lds + off);
    off += align16(((size_t)bins + 1) * 4);
    float4 *rings = (float4 *)(lds + off);
    for (int i = threadIdx.x; i < kBatchWaves * bins; i += kBatchBlock) hist_all[i] = 0u;
    for (int i = threadIdx.x; i <= bins; i += kBatchBlock) thr[i] = p.thr[i];
    if constexpr (TABLE)
        for (int i = threadIdx.x; i < p.tab_n; i += kBatchBlock) tab2[i] = p.tab2[i];
    __syncthreads();

    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    CellLds L;
    L.hist = (unsigned long long *)(hist_all + wave * bins);
    L.tab2 = tab2;
    L.thr = thr;
    L.elo = thr;
    L.ehi = thr;
    L.ring = rings + wave * kRingStride;
    // u32 counters (table offsets 4 x bin), drained into the trajectory's row
    PlainCons<PBC, false, false, TABLE, 4> cons{p, L, 0ull, 0ull};

    const int home_q = blockIdx.x % kSlices;
    if (a.clock && lane == 0) atomicMax(&a.clock[home_q * kClock], ~clock_now());
    int q = home_q, visited = 0;
    int cur = -1;
    BatchTraj T{};
    CellGrid G{};
    StencilLane sl{0, 0, 0, -1};
    while (visited < kSlices) {
        unsigned long long g0 = 0;
        if (lane == 0) g0 = atomicAdd(&a.queue[q], 1ull);
        const unsigned g_lo = __builtin_amdgcn_readfirstlane((unsigned)g0);
        const unsigned g_hi = __builtin_amdgcn_readfirstlane((unsigned)(g0 >> 32));
        const long long task = a.slice_lo[q] + (long long)(((unsigned long long)g_hi << 32) | g_lo);
        if (task >= a.slice_lo[q + 1]) {  // slice drained: help the next one
            if (a.clock && lane == 0 && visited == 0) atomicMax(&a.clock[q * kClock + 1], ~clock_now());
            ++visited;
            q = q + 1 == kSlices ? 0 : q + 1;
            continue;
        }
        if (a.clock && lane == 0 && visited > 0) atomicAdd(&a.clock[q * kClock + 3], 1ull);
        if (cur < 0 || task < T.task0 || task >= T.task0 + ((long long)T.n_frames * T.upf + T.grab - 1) / T.grab) {
            if (cur >= 0) cons.drain();
            cur = __builtin_amdgcn_readfirstlane(
                find_last_le(a.n_traj, task, [&](int i) { return a.traj[i].task0; }));
            cons.row = a.counts + (long long)cur * a.row_bins;
            T = uload(a.traj, cur);
            G = CellGrid{a.sorted + T.af0, a.start + T.start_off, a.frame_flag + T.box0,
                         a.boxes + T.box0, T.n, T.n_frames, T.ncx, T.ncy, T.ncz, T.k,
                         T.shift_ok, kSlices, T.cull2, nullptr, 0};
            sl = stencil_lane(lane, T.kind == KIND_CELLS ? T.k : 1);
        }
        const long long u0 = (task - T.task0) * T.grab;
        const long long u1 = min(u0 + (long long)T.grab, (long long)T.n_frames * T.upf);
        for (long long u = u0; u < u1; ++u) {
            const int f = (int)(u / T.upf);
            const int h = (int)(u - (long long)f * T.upf);
            if constexpr (PBC != WMD_PBC_NONE) {
                if (T.kind == KIND_CELLS) {
                    cell_unit<PBC, false>(G, f, h, sl, L.ring, nullptr, cons);
                    continue;
                }
            }
            brute_unit(G, f, h, T.bh, cons);
        }
    }
    if (cur >= 0) cons.drain();
    if (a.clock && lane == 0) atomicMax(&a.clock[home_q * kClock + 2], clock_now());
}

// ---- autotuner cache ---------------------------------------------------------
// key: (arch, pbc, bins, r_max bits, atom-count bucket, density bucket) ->
// (k: -1 brute force or 1..3 cells per r_max, brute-force home atoms per unit
// (0 default), units per task (0 default)).  Buckets: quarter steps of log2.
using TuneKey = std::tuple<std::string, int, int, uint32_t, int, int>;
struct TuneVal {
    int k, bh, grab;
};
std::mutex g_tune_mu;
std::map<TuneKey, TuneVal> g_tune;
constexpr double kTuneAtomFrames = 4e6;  // work of one timed tuning run (atom-frames)

}  // namespace

// ---------------------------------------------------------------------------
// Batch handle
// ---------------------------------------------------------------------------
struct wmd_batch {
    wmd_ctx *ctx = nullptr;
    int n_traj = 0;
    std::vector<uint32_t> n_atoms, n_frames;
    std::vector<long long> store_off;
    std::vector<uint32_t> stride;
    long long store_floats = 0, total_frames = 0;
    int bins = 0;
    float r_max = 0.f;
    int pbc = 0;
    uint32_t flags = 0;
    BinTables tables;
    bool table = false;
    std::vector<int> tune_k;     // per trajectory: 0 automatic, -1 brute, 1..3
    std::vector<int> tune_bh;    // per trajectory: brute-force home atoms per unit (0 default)
    std::vector<int> tune_grab;  // per trajectory: units per task (0 default)
    bool use_graph = true;
    // plan (set_boxes)
    bool planned = false;
    std::vector<float> boxes_host;
    std::vector<BatchTraj> traj;
    std::vector<BatchRmsdTraj> rtraj;
    long long total_af = 0, n_units = 0, n_tasks = 0, scan_blocks = 0, cell_counts = 0, cell_starts = 0;
    long long rmsd_units = 0;
    long long slice_lo[kSlices + 1] = {};
    bool any_cells = false;
    int max_ncell = 0;  // largest cell grid of a cell-list trajectory
    // device
    // d_queue: [kSlices] task counters, then [kSlices][kClock] wall-clock marks
    DevBuf d_counts, d_thr, d_tab2, d_traj, d_rtraj, d_boxes, d_flag, d_count, d_start, d_cell_of,
        d_slot, d_image, d_sorted, d_queue, d_partials;
    // bins beyond the per-wave LDS histograms: slice_bins per pass (0 = one
    // pass) and per slice [slice_bins + 2] local thresholds
    int slice_bins = 0;
    DevBuf d_sthr;
    HostBuf h_boxes;
    // graph
    hipStream_t cap = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    const void *g_store = nullptr;
    const void *g_out32 = nullptr, *g_out64 = nullptr, *g_err = nullptr;
    ~wmd_batch() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        if (cap) (void)hipStreamDestroy(cap);
    }
};

namespace {

void drop_graph(wmd_batch *b) {
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
    b->g_store = nullptr;
}

int grid_1d(long long total) {
    long long g = (total + kBlock - 1) / kBlock;
    return (int)std::max<long long>(1, std::min<long long>(g, 65536));
}

// Enumeration of one trajectory: k > 0 cells per r_max, or 0 (brute force).
int traj_cells(const wmd_batch *b, int t, const BoxSet &B, int *nc) {
    const int n = (int)b->n_atoms[t];
    if (b->pbc == WMD_PBC_NONE || n < 64) return 0;
    const double mw = std::min(B.min_width[0], std::min(B.min_width[1], B.min_width[2]));
    if (!((double)b->r_max <= 0.5 * mw * (1.0 - 1e-5))) return 0;
    const int forced = b->tune_k[t];
    if (forced < 0) return 0;
    return choose_cells((double)b->r_max, B.min_width, n, forced, nc);
}

void plan_batch(wmd_batch *b) {
    const int T = b->n_traj;
    hipStream_t st = b->ctx->stream;
    b->traj.assign(T, BatchTraj{});
    b->rtraj.assign(T, BatchRmsdTraj{});
    std::vector<FrameBox> fb;
    fb.reserve((size_t)b->total_frames);
    std::vector<double> cost(T, 0.0);
    long long af = 0, units = 0, tasks = 0, scan = 0, cnt = 0, sta = 0, box0 = 0, runits = 0, out0 = 0;
    b->any_cells = false;
    b->max_ncell = 0;
    for (int t = 0; t < T; ++t) {
        const int n = (int)b->n_atoms[t], F = (int)b->n_frames[t];
        BatchTraj &d = b->traj[t];
        d.store_off = b->store_off[t];
        d.stride = (int)b->stride[t];
        d.n = n;
        d.n_frames = F;
        d.af0 = af;
        d.box0 = (int)box0;
        d.unit0 = units;
        d.task0 = tasks;
        d.scan0 = scan;
        d.count_off = cnt;
        d.start_off = sta;
        BoxSet B;
        if (b->pbc != WMD_PBC_NONE && F > 0)
            B = make_boxes(b->pbc, b->boxes_host.data() + 9 * box0, (size_t)F);
        else
            B = make_boxes(WMD_PBC_NONE, nullptr, (size_t)F);
        for (const FrameBox &x : B.fb) fb.push_back(x);
        int nc[3] = {0, 0, 0};
        const int k = F > 0 && n > 1 ? traj_cells(b, t, B, nc) : 0;
        if (k > 0) {
            const int ncell = nc[0] * nc[1] * nc[2];
            d.kind = KIND_CELLS;
            d.k = k;
            d.ncx = nc[0]; d.ncy = nc[1]; d.ncz = nc[2];
            d.shift_ok = (nc[0] >= 2 * k + 1 && nc[1] >= 2 * k + 1 && nc[2] >= 2 * k + 1) ? 1 : 0;
            d.cull2 = cull2_for((double)b->r_max, B);
            d.upf = ncell;
            scan += F;
            cnt += (long long)F * ncell;
            sta += (long long)F * (ncell + 1);
            b->any_cells = true;
            b->max_ncell = std::max(b->max_ncell, ncell);
            // per-unit cost: home atoms x half-stencil atoms, plus a fixed part
            const double nh = (double)n / ncell;
            const double st_cells = (std::pow(2.0 * k + 1.0, 3) - 1.0) / 2.0 + 1.0;
            cost[t] = (double)F * ncell * (nh * nh * st_cells + 2000.0);
        } else {
            d.kind = KIND_BRUTE;
            d.k = 0;
            d.ncx = d.ncy = d.ncz = 1;
            d.bh = b->tune_bh[t] > 0 ? b->tune_bh[t] : kBruteHome;
            d.upf = n > 1 ? (n + d.bh - 1) / d.bh : 0;
            cost[t] = (double)F * d.upf * (d.bh * 0.5 * n * 1.5 + 2000.0);
        }
        if (F == 0) d.upf = 0;
        d.grab = b->tune_grab[t] > 0 ? b->tune_grab[t] : kGrab;
        units += (long long)F * d.upf;
        tasks += ((long long)F * d.upf + d.grab - 1) / d.grab;
        af += (long long)F * n;
        box0 += F;
        // RMSD descriptor: tiles of up to 2048 float4 groups
        BatchRmsdTraj &r = b->rtraj[t];
        const int groups = (int)(b->stride[t] / 4);
        r.store_off = b->store_off[t];
        r.unit0 = runits;
        r.out0 = out0;
        r.stride = (int)b->stride[t];
        r.n = n;
        r.n_frames = F;
        r.tiles = std::max(1, (groups + 2047) / 2048);
        r.gpt = std::max(1, (groups + r.tiles - 1) / r.tiles);
        runits += (long long)F * r.tiles;
        out0 += F;
    }
    b->total_af = af;
    b->n_units = units;
    b->n_tasks = tasks;
    b->scan_blocks = scan;
    b->cell_counts = cnt;
    b->cell_starts = sta;
    b->rmsd_units = runits;
    // slices of equal estimated cost, cut at task granularity
    double total = 0.0;
    for (double c : cost) total += c;
    b->slice_lo[0] = 0;
    {
        int t = 0;
        double acc = 0.0;
        for (int q = 1; q < kSlices; ++q) {
            const double target = total * q / kSlices;
            while (t < T && acc + cost[t] < target) acc += cost[t++];
            long long cut = tasks;
            if (t < T) {
                const BatchTraj &d = b->traj[t];
                const long long tt = ((long long)d.n_frames * d.upf + d.grab - 1) / d.grab;
                const double frac = cost[t] > 0 ? (target - acc) / cost[t] : 0.0;
                cut = d.task0 + (long long)std::llround(frac * (double)tt);
            }
            b->slice_lo[q] = std::max(b->slice_lo[q - 1], std::min(cut, tasks));
        }
        b->slice_lo[kSlices] = tasks;
    }
    // device buffers
    const size_t nT = std::max(T, 1);
    b->d_traj.reserve(sizeof(BatchTraj) * nT);
    b->d_rtraj.reserve(sizeof(BatchRmsdTraj) * nT);
    b->d_boxes.reserve(sizeof(FrameBox) * std::max<size_t>(fb.size(), 1));
    b->d_flag.reserve(4 * std::max<long long>(b->total_frames, 1));
    b->d_count.reserve(4 * std::max<long long>(cnt, 1));
    b->d_start.reserve(4 * std::max<long long>(sta, 1));
    const long long afc = std::max<long long>(af, 1);
    if (b->flags & WMD_BATCH_RDF) {
        if (b->any_cells && b->max_ncell > kBuildMaxCells) {
            b->d_cell_of.reserve(4 * afc);
            b->d_slot.reserve(4 * afc);
            b->d_image.reserve(4 * afc);
        }
        b->d_sorted.reserve(16 * afc);
    }
    b->d_queue.reserve(8 * kSlices * (1 + kClock));
    b->d_partials.reserve(8 * 13 * std::max<long long>(runits, 1));
    const size_t tb = sizeof(BatchTraj) * T, rb = sizeof(BatchRmsdTraj) * T, bb = sizeof(FrameBox) * fb.size();
    char *hs = (char *)b->h_boxes.reserve(std::max<size_t>(tb + rb + bb, 1));
    WMD_HIP_CHECK(hipStreamSynchronize(st));  // staging free
    std::memcpy(hs, b->traj.data(), tb);
    std::memcpy(hs + tb, b->rtraj.data(), rb);
    std::memcpy(hs + tb + rb, fb.data(), bb);
    if (tb) WMD_HIP_CHECK(hipMemcpyAsync(b->d_traj.ptr, hs, tb, hipMemcpyHostToDevice, st));
    if (rb) WMD_HIP_CHECK(hipMemcpyAsync(b->d_rtraj.ptr, hs + tb, rb, hipMemcpyHostToDevice, st));
    if (bb) WMD_HIP_CHECK(hipMemcpyAsync(b->d_boxes.ptr, hs + tb + rb, bb, hipMemcpyHostToDevice, st));
    WMD_HIP_CHECK(hipStreamSynchronize(st));
    b->planned = true;
    drop_graph(b);
}

template <int PBC>
void enqueue_pairs(wmd_batch *b, hipStream_t st, const float *d_store) {
    const int T = b->n_traj;
    const BatchTraj *dt = b->d_traj.as<BatchTraj>();
    const bool fused = b->max_ncell <= kBuildMaxCells;
    if (fused) {
        hipLaunchKernelGGL(batch_build_kernel<PBC == WMD_PBC_NONE ? WMD_PBC_ORTHO : PBC>,
                           dim3((unsigned)b->total_frames), dim3(kBuildBlock), build_lds_bytes(b->max_ncell), st, dt, T,
                           d_store, (const FrameBox *)b->d_boxes.ptr, b->d_start.as<unsigned>(),
                           b->d_sorted.as<float4>(), b->d_flag.as<unsigned>());
        WMD_HIP_CHECK(hipGetLastError());
    } else if (b->any_cells && PBC != WMD_PBC_NONE) {
        hipLaunchKernelGGL(batch_assign_kernel<PBC == WMD_PBC_NONE ? WMD_PBC_ORTHO : PBC>,
                           dim3(grid_1d(b->total_af)), dim3(kBlock), 0, st, dt, T, b->total_af,
                           d_store, (const FrameBox *)b->d_boxes.ptr, b->d_cell_of.as<unsigned>(),
                           b->d_slot.as<unsigned>(), b->d_count.as<unsigned>(),
                           b->d_image.as<unsigned>(), b->d_flag.as<unsigned>());
        WMD_HIP_CHECK(hipGetLastError());
        hipLaunchKernelGGL(batch_scan_kernel, dim3((unsigned)b->scan_blocks), dim3(kBlock), 0, st, dt, T,
                           (const unsigned *)b->d_count.ptr, b->d_start.as<unsigned>());
        WMD_HIP_CHECK(hipGetLastError());
    }
    if (!fused) {
        hipLaunchKernelGGL(batch_scatter_kernel, dim3(grid_1d(b->total_af)), dim3(kBlock), 0, st, dt, T,
                           b->total_af, d_store, (const unsigned *)b->d_cell_of.ptr,
                           (const unsigned *)b->d_slot.ptr, (const unsigned *)b->d_image.ptr,
                           (const unsigned *)b->d_start.ptr, b->d_sorted.as<float4>());
        WMD_HIP_CHECK(hipGetLastError());
    }

    BatchPairArgs a;
    a.traj = dt;
    a.n_traj = T;
    a.sorted = b->d_sorted.as<float4>();
    a.start = b->d_start.as<unsigned>();
    a.frame_flag = b->d_flag.as<unsigned>();
    a.boxes = b->d_boxes.as<FrameBox>();
    a.counts = b->d_counts.as<unsigned long long>();
    a.queue = b->d_queue.as<unsigned long long>();
    a.clock = a.queue + kSlices;
    for (int q = 0; q <= kSlices; ++q) a.slice_lo[q] = b->slice_lo[q];
    HistParams &p = a.p;
    std::memset(&p, 0, sizeof(p));
    p.bins = b->bins;
    p.hist_copies = 1;
    p.cut = b->tables.cut;
    p.inv_w = 1.0f / b->tables.w;
    p.thr = b->d_thr.as<float>();
    p.tab2 = b->table ? b->d_tab2.as<uint2>() : nullptr;
    p.tab_n = b->table ? (int)b->tables.tab2.size() : 0;
    p.tab_sh = b->tables.tab_sh;
    p.tab_base = b->tables.tab_base;
    p.counts = a.counts;
    p.count_evals = 1;
    a.row_bins = b->bins;
    auto launch = [&](auto kern, size_t lds) {
        int occ = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, kern, kBatchBlock, lds) != hipSuccess || occ < 1)
            occ = 1;
        int grid = std::max(kSlices, b->ctx->cu_count * occ / kSlices * kSlices);
        hipLaunchKernelGGL(kern, dim3(grid), dim3(kBatchBlock), lds, st, a);
        WMD_HIP_CHECK(hipGetLastError());
    };
    if (b->slice_bins == 0) {
        const size_t lds = batch_lds_bytes(b->bins, p.tab_n);
        if (b->table) launch(batch_pair_kernel<PBC, true>, lds);
        else launch(batch_pair_kernel<PBC, false>, lds);
        return;
    }
    // bins beyond the per-wave LDS histograms: one pass (a full enumeration)
    // per slice of bins, with a local bin for the pairs below the slice
    const int S = b->slice_bins;
    const int n_pass = (b->bins + S - 1) / S;
    for (int k = 0; k < n_pass; ++k) {
        const int b0 = k * S, nb = std::min(S, b->bins - b0);
        if (k > 0) {
            hipLaunchKernelGGL(fill_u64_kernel, dim3(1), dim3(kBlock), 0, st, b->d_queue.as<unsigned long long>(),
                               (long long)kSlices);
            WMD_HIP_CHECK(hipGetLastError());
        }
        p.bins = nb + 1;
        p.cut = b->tables.thr[b0 + nb];
        p.thr = b->d_sthr.as<float>() + (size_t)k * (S + 2);
        p.tab2 = nullptr;
        p.tab_n = 0;
        p.shift = b0 - 1;
        p.skip = 1;
        a.counts = b->d_counts.as<unsigned long long>() + b0;
        p.counts = a.counts;
        launch(batch_pair_kernel<PBC, false>, batch_lds_bytes(nb + 1, 0));
    }
}

// The whole step on `st`: clears, cell build, pair kernel, RMSD.
void enqueue_step(wmd_batch *b, hipStream_t st, const float *d_store, float *d_out32, double *d_out64,
                  double *d_err) {
    const bool rdf = (b->flags & WMD_BATCH_RDF) && b->bins > 0 && b->n_units > 0;
    if (rdf) {
        if (b->max_ncell > kBuildMaxCells)
            hipLaunchKernelGGL(fill_u32_kernel, dim3(grid_1d(std::max<long long>(b->cell_counts, 1))), dim3(kBlock), 0, st,
                               b->d_count.as<unsigned>(), b->cell_counts, 0u);
        hipLaunchKernelGGL(fill_u32_kernel, dim3(grid_1d(std::max<long long>(b->total_frames, 1))), dim3(kBlock), 0, st,
                           b->d_flag.as<unsigned>(), b->total_frames, 0u);
        hipLaunchKernelGGL(fill_u64_kernel, dim3(1), dim3(kBlock), 0, st, b->d_queue.as<unsigned long long>(),
                           (long long)kSlices * (1 + kClock));
        WMD_HIP_CHECK(hipGetLastError());
        if (b->pbc == WMD_PBC_NONE) enqueue_pairs<WMD_PBC_NONE>(b, st, d_store);
        else if (b->pbc == WMD_PBC_ORTHO) enqueue_pairs<WMD_PBC_ORTHO>(b, st, d_store);
        else enqueue_pairs<WMD_PBC_TRICLINIC>(b, st, d_store);
    }
    if ((b->flags & WMD_BATCH_RMSD) && d_out32)
        batch_rmsd_enqueue(st, b->d_rtraj.as<BatchRmsdTraj>(), b->n_traj, b->rmsd_units,
                           b->total_frames, d_store, b->d_partials.as<double>(), d_out32, d_out64, d_err);
}

void run_batch(wmd_batch *b, const float *d_store, float *d_out32, double *d_out64, double *d_err,
               hipStream_t st, bool graph_ok) {
    WMD_REQUIRE(b->planned, WMD_ERR_INVALID, "wmd_batch_set_boxes must be called before a run");
    WMD_REQUIRE(d_store != nullptr || b->store_floats == 0, WMD_ERR_INVALID, "null store");
    WMD_REQUIRE(!(b->flags & WMD_BATCH_RMSD) || d_out32 || b->total_frames == 0, WMD_ERR_INVALID,
                "RMSD output required (the batch was created with WMD_BATCH_RMSD)");
    if (!(b->use_graph && graph_ok)) {
        ProfScope prof(b->ctx, st, "batch_step");
        enqueue_step(b, st, d_store, d_out32, d_out64, d_err);
        return;
    }
    if (!b->exec || b->g_store != d_store || b->g_out32 != d_out32 || b->g_out64 != d_out64 ||
        b->g_err != d_err) {
        drop_graph(b);
        if (!b->cap) WMD_HIP_CHECK(hipStreamCreateWithFlags(&b->cap, hipStreamNonBlocking));
        WMD_HIP_CHECK(hipStreamBeginCapture(b->cap, hipStreamCaptureModeThreadLocal));
        try {
            enqueue_step(b, b->cap, d_store, d_out32, d_out64, d_err);
        } catch (...) {
            hipGraph_t g = nullptr;
            (void)hipStreamEndCapture(b->cap, &g);
            if (g) (void)hipGraphDestroy(g);
            throw;
        }
        WMD_HIP_CHECK(hipStreamEndCapture(b->cap, &b->graph));
        WMD_HIP_CHECK(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
        b->g_store = d_store;
        b->g_out32 = d_out32;
        b->g_out64 = d_out64;
        b->g_err = d_err;
    }
    ProfScope prof(b->ctx, st, "batch_step");
    WMD_HIP_CHECK(hipGraphLaunch(b->exec, st));
}

TuneKey tune_key(const wmd_batch *b, int t) {
    const int n = (int)b->n_atoms[t];
    double vol = 0.0;
    long long box0 = 0;
    for (int s = 0; s < t; ++s) box0 += b->n_frames[s];
    if (b->pbc != WMD_PBC_NONE && b->n_frames[t] > 0) {
        const float *m = b->boxes_host.data() + 9 * box0;
        const double a[9] = {m[0], m[1], m[2], m[3], m[4], m[5], m[6], m[7], m[8]};
        vol = std::fabs(a[0] * (a[4] * a[8] - a[5] * a[7]) - a[1] * (a[3] * a[8] - a[5] * a[6]) +
                        a[2] * (a[3] * a[7] - a[4] * a[6]));
    }
    const int nb = n > 0 ? (int)std::floor(std::log2((double)n) * 4.0) : -1;
    const int db = vol > 0 ? (int)std::floor(std::log2((double)n / vol) * 4.0) : -999;
    uint32_t rb;
    std::memcpy(&rb, &b->r_max, 4);
    return TuneKey(b->ctx->arch, b->pbc, b->bins, rb, nb, db);
}

// offsets / strides (optional, both or neither): an explicit store layout, as
// the autotuner's sub-batches use to read members of the caller's store.
wmd_batch *create_batch(wmd_ctx *ctx, size_t n_traj, const uint32_t *n_atoms, const uint32_t *n_frames,
                        size_t bins, float r_max, int pbc, uint32_t flags,
                        const long long *offsets = nullptr, const uint32_t *strides = nullptr) {
    WMD_REQUIRE(ctx, WMD_ERR_INVALID, "null context");
    WMD_REQUIRE(n_traj == 0 || (n_atoms && n_frames), WMD_ERR_INVALID, "null size arrays");
    WMD_REQUIRE(n_traj < (1u << 20), WMD_ERR_UNSUPPORTED, "too many trajectories");
    WMD_REQUIRE(pbc >= WMD_PBC_NONE && pbc <= WMD_PBC_TRICLINIC, WMD_ERR_PARSE,
                "pbc must be none, orthorhombic or triclinic");
    WMD_REQUIRE(bins <= (size_t(1) << 22), WMD_ERR_UNSUPPORTED, "a batch supports at most 4194304 bins");
    WMD_REQUIRE(std::isfinite(r_max) && r_max > 0.0f, WMD_ERR_PARSE, "r_max must be finite and positive");
    WMD_REQUIRE((flags & ~(uint32_t)(WMD_BATCH_RDF | WMD_BATCH_RMSD | WMD_BATCH_NO_GRAPH)) == 0,
                WMD_ERR_INVALID, "unknown batch flags");
    bind(ctx);
    auto *b = new wmd_batch();
    try {
        b->ctx = ctx;
        b->n_traj = (int)n_traj;
        b->n_atoms.assign(n_atoms, n_atoms + n_traj);
        b->n_frames.assign(n_frames, n_frames + n_traj);
        b->bins = (int)bins;
        b->r_max = r_max;
        b->pbc = pbc;
        b->flags = flags;
        b->use_graph = !(flags & WMD_BATCH_NO_GRAPH);
        b->tune_k.assign(n_traj, 0);
        b->tune_bh.assign(n_traj, 0);
        b->tune_grab.assign(n_traj, 0);
        long long off = 0, frames = 0;
        for (size_t t = 0; t < n_traj; ++t) {
            WMD_REQUIRE(n_atoms[t] < (1u << 27), WMD_ERR_UNSUPPORTED, "trajectory too large");
            const uint32_t s = (uint32_t)round_up(std::max<uint32_t>(n_atoms[t], 1), 4);
            b->store_off.push_back(offsets ? offsets[t] : off);
            b->stride.push_back(strides ? strides[t] : s);
            off = std::max(off, b->store_off.back() + 3LL * b->stride.back() * n_frames[t]);
            frames += n_frames[t];
        }
        b->store_floats = off;
        b->total_frames = frames;
        WMD_REQUIRE(frames < (1ll << 31), WMD_ERR_UNSUPPORTED, "too many frames");
        hipStream_t st = ctx->stream;
        b->d_counts.reserve(8 * std::max<size_t>(n_traj * bins, 1));
        hipLaunchKernelGGL(fill_u64_kernel, dim3((unsigned)((std::max<size_t>(n_traj * bins, 1) + kBlock - 1) / kBlock)),
                           dim3(kBlock), 0, st, b->d_counts.as<unsigned long long>(),
                           (long long)(n_traj * bins));
        WMD_HIP_CHECK(hipGetLastError());
        std::vector<uint2> t4;  // host source of the table copy, alive until the sync below
        if (bins > 0) {
            b->tables = make_bin_tables((int)bins, r_max);
            b->table = !b->tables.tab2.empty() && (int)b->tables.tab2.size() <= kMaxTab2;
            b->d_thr.reserve((bins + 1) * 4);
            WMD_HIP_CHECK(hipMemcpyAsync(b->d_thr.ptr, b->tables.thr.data(), (bins + 1) * 4,
                                         hipMemcpyHostToDevice, st));
            if (b->table) {
                // the pair kernel's wave histograms hold u32 counters: byte
                // offsets 4 x bin in place of the table's 8 x bin
                t4 = b->tables.tab2;
                for (uint2 &e : t4) e.y >>= 1;
                b->d_tab2.reserve(t4.size() * 8);
                WMD_HIP_CHECK(hipMemcpyAsync(b->d_tab2.ptr, t4.data(), t4.size() * 8, hipMemcpyHostToDevice, st));
            }
        }
        WMD_HIP_CHECK(hipStreamSynchronize(st));
        if (bins > 0 &&
            batch_lds_bytes((int)bins, b->table ? (int)b->tables.tab2.size() : 0) > ctx->lds_per_block) {
            // slices of bins, one pass each: the largest multiple of 64 whose
            // per-wave histograms (plus a bin below the slice) fit LDS
            int S = (int)std::min<size_t>(bins, 1 << 16) / 64 * 64;
            while (S > 64 && batch_lds_bytes(S + 1, 0) > ctx->lds_per_block) S -= 64;
            WMD_REQUIRE(S >= 64 && batch_lds_bytes(S + 1, 0) <= ctx->lds_per_block, WMD_ERR_UNSUPPORTED,
                        "too many bins for the batch kernel's LDS");
            const int n_pass = ((int)bins + S - 1) / S;
            const size_t per = (size_t)S + 2;
            std::vector<float> sthr(per * n_pass, 0.0f);
            for (int k = 0; k < n_pass; ++k) {
                const int b0 = k * S, nb = std::min(S, (int)bins - b0);
                for (int j = 0; j <= nb; ++j) sthr[k * per + j + 1] = b->tables.thr[b0 + j];
            }
            b->d_sthr.reserve(sthr.size() * 4);
            WMD_HIP_CHECK(hipMemcpy(b->d_sthr.ptr, sthr.data(), sthr.size() * 4, hipMemcpyHostToDevice));
            b->slice_bins = S;
            b->table = false;
        }
    } catch (...) {
        delete b;
        throw;
    }
    return b;
}

void set_boxes(wmd_batch *b, const float *boxes) {
    WMD_REQUIRE(b->pbc == WMD_PBC_NONE || boxes || b->total_frames == 0, WMD_ERR_INVALID,
                "boxes required for PBC");
    if (b->pbc != WMD_PBC_NONE && b->total_frames > 0)
        b->boxes_host.assign(boxes, boxes + 9 * b->total_frames);
    else
        b->boxes_host.clear();
    plan_batch(b);
}

}  // namespace

extern "C" {

wmd_status wmd_batch_create(wmd_ctx *ctx, size_t n_traj, const uint32_t *n_atoms,
                            const uint32_t *n_frames, size_t bins, float r_max, int pbc,
                            uint32_t flags, wmd_batch **out) {
    WMD_API_BEGIN
    WMD_REQUIRE(out, WMD_ERR_INVALID, "null output handle");
    *out = nullptr;
    *out = create_batch(ctx, n_traj, n_atoms, n_frames, bins, r_max, pbc, flags);
    WMD_API_END
}

wmd_status wmd_batch_store_layout(const wmd_batch *b, uint64_t *offsets, uint32_t *strides,
                                  uint64_t *total_floats) {
    WMD_API_BEGIN
    WMD_REQUIRE(b, WMD_ERR_INVALID, "null batch");
    for (int t = 0; t < b->n_traj; ++t) {
        if (offsets) offsets[t] = (uint64_t)b->store_off[t];
        if (strides) strides[t] = b->stride[t];
    }
    if (total_floats) *total_floats = (uint64_t)b->store_floats;
    WMD_API_END
}

wmd_status wmd_batch_set_boxes(wmd_batch *b, const float *boxes) {
    WMD_API_BEGIN
    WMD_REQUIRE(b, WMD_ERR_INVALID, "null batch");
    bind(b->ctx);
    set_boxes(b, boxes);
    WMD_API_END
}

wmd_status wmd_batch_set_tuning(wmd_batch *b, const int *cell_k, const int *home_block, const int *grab) {
    WMD_API_BEGIN
    WMD_REQUIRE(b, WMD_ERR_INVALID, "null batch");
    for (int t = 0; t < b->n_traj; ++t) {
        const int k = cell_k ? cell_k[t] : 0;
        const int bh = home_block ? home_block[t] : 0;
        const int g = grab ? grab[t] : 0;
        WMD_REQUIRE(k >= -1 && k <= 3, WMD_ERR_INVALID, "cell_k must be -1, 0 or 1..3");
        WMD_REQUIRE(bh == 0 || (bh % 64 == 0 && bh >= 64 && bh <= kMaxBruteHome), WMD_ERR_INVALID,
                    "home_block must be 0 or a multiple of 64 in [64, 512]");
        WMD_REQUIRE(g >= 0 && g <= kMaxGrab, WMD_ERR_INVALID, "grab must be 0..64");
    }
    for (int t = 0; t < b->n_traj; ++t) {
        b->tune_k[t] = cell_k ? cell_k[t] : 0;
        b->tune_bh[t] = home_block ? home_block[t] : 0;
        b->tune_grab[t] = grab ? grab[t] : 0;
    }
    if (b->planned) {
        bind(b->ctx);
        plan_batch(b);
    }
    WMD_API_END
}

wmd_status wmd_batch_run_device(wmd_batch *b, const float *d_store, float *d_rmsd, double *d_rmsd64,
                                double *d_err, void *stream) {
    WMD_API_BEGIN
    WMD_REQUIRE(b, WMD_ERR_INVALID, "null batch");
    bind(b->ctx);
    run_batch(b, d_store, d_rmsd, d_rmsd64, d_err, pick_stream(b->ctx, stream), true);
    WMD_API_END
}

wmd_status wmd_batch_read_counts(wmd_batch *b, uint64_t *counts) {
    WMD_API_BEGIN
    WMD_REQUIRE(b && (counts || b->n_traj * b->bins == 0), WMD_ERR_INVALID, "null argument");
    bind(b->ctx);
    WMD_HIP_CHECK(hipDeviceSynchronize());
    const size_t bytes = (size_t)b->n_traj * b->bins * 8;
    if (bytes) WMD_HIP_CHECK(hipMemcpy(counts, b->d_counts.ptr, bytes, hipMemcpyDeviceToHost));
    WMD_API_END
}

wmd_status wmd_batch_reset(wmd_batch *b) {
    WMD_API_BEGIN
    WMD_REQUIRE(b, WMD_ERR_INVALID, "null batch");
    bind(b->ctx);
    WMD_HIP_CHECK(hipDeviceSynchronize());
    const long long n = (long long)b->n_traj * b->bins;
    if (n) {
        hipLaunchKernelGGL(fill_u64_kernel, dim3((unsigned)((n + kBlock - 1) / kBlock)), dim3(kBlock), 0,
                           b->ctx->stream, b->d_counts.as<unsigned long long>(), n);
        WMD_HIP_CHECK(hipGetLastError());
        WMD_HIP_CHECK(hipStreamSynchronize(b->ctx->stream));
    }
    WMD_API_END
}

wmd_status wmd_batch_plan_info(const wmd_batch *b, int *cell_k, int *home_block, int *grab,
                               long long *n_units) {
    WMD_API_BEGIN
    WMD_REQUIRE(b, WMD_ERR_INVALID, "null batch");
    WMD_REQUIRE(b->planned, WMD_ERR_INVALID, "wmd_batch_set_boxes must be called first");
    for (int t = 0; t < b->n_traj; ++t) {
        const BatchTraj &d = b->traj[t];
        if (cell_k) cell_k[t] = d.kind == KIND_CELLS ? d.k : 0;
        if (home_block) home_block[t] = d.kind == KIND_CELLS ? 0 : d.bh;
        if (grab) grab[t] = d.grab;
    }
    if (n_units) *n_units = b->n_units;
    WMD_API_END
}

wmd_status wmd_batch_slice_times(wmd_batch *b, double *out) {
    WMD_API_BEGIN
    WMD_REQUIRE(b && out, WMD_ERR_INVALID, "null argument");
    WMD_REQUIRE(b->planned, WMD_ERR_INVALID, "wmd_batch_set_boxes must be called first");
    bind(b->ctx);
    WMD_HIP_CHECK(hipDeviceSynchronize());
    unsigned long long c[kSlices * kClock];
    WMD_HIP_CHECK(hipMemcpy(c, b->d_queue.as<unsigned long long>() + kSlices, sizeof(c), hipMemcpyDeviceToHost));
    int khz = 0;
    WMD_HIP_CHECK(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, b->ctx->device));
    WMD_REQUIRE(khz > 0, WMD_ERR_DEVICE, "device reports no wall clock rate");
    unsigned long long t0 = ~0ull;
    for (int q = 0; q < kSlices; ++q)
        if (c[q * kClock]) t0 = std::min(t0, ~c[q * kClock]);
    const double per_ms = (double)khz;
    for (int q = 0; q < kSlices; ++q) {
        const unsigned long long *r = c + q * kClock;
        auto rel = [&](unsigned long long t) { return t && t0 != ~0ull ? ((double)t - (double)t0) / per_ms : -1.0; };
        out[q * 4 + 0] = r[0] ? rel(~r[0]) : -1.0;
        out[q * 4 + 1] = r[1] ? rel(~r[1]) : -1.0;
        out[q * 4 + 2] = rel(r[2]);
        out[q * 4 + 3] = (double)r[3];
    }
    WMD_API_END
}

wmd_status wmd_batch_autotune(wmd_batch *b, const float *d_store, int reps, void *stream) {
    WMD_API_BEGIN
    WMD_REQUIRE(b, WMD_ERR_INVALID, "null batch");
    WMD_REQUIRE(b->planned, WMD_ERR_INVALID, "wmd_batch_set_boxes must be called first");
    WMD_REQUIRE(d_store || b->store_floats == 0, WMD_ERR_INVALID, "null store");
    bind(b->ctx);
    hipStream_t st = pick_stream(b->ctx, stream);
    reps = std::max(1, std::min(reps, 20));
    std::map<TuneKey, std::vector<int>> classes;
    for (int t = 0; t < b->n_traj; ++t)
        if (b->n_atoms[t] > 1 && b->n_frames[t] > 0) classes[tune_key(b, t)].push_back(t);
    std::vector<long long> box0(b->n_traj + 1, 0);
    for (int t = 0; t < b->n_traj; ++t) box0[t + 1] = box0[t] + b->n_frames[t];
    hipEvent_t e0, e1;
    WMD_HIP_CHECK(hipEventCreate(&e0));
    WMD_HIP_CHECK(hipEventCreate(&e1));
    wmd_batch *s = nullptr;
    try {
        for (auto &kv : classes) {
            TuneVal choice{0, 0, 0};
            bool cached;
            {
                std::lock_guard<std::mutex> lk(g_tune_mu);
                auto it = g_tune.find(kv.first);
                cached = it != g_tune.end();
                if (cached) choice = it->second;
            }
            if (!cached) {
                // A sub-batch of the class's members (a frame prefix of each,
                // replicated until one run holds about kTuneAtomFrames
                // atom-frames, so that it fills the device) reading the
                // caller's store in place.
                const std::vector<int> &ts = kv.second;
                double atoms = 0.0;
                for (int t : ts) atoms += b->n_atoms[t];
                const int fmax = (int)std::max(1.0, std::ceil(kTuneAtomFrames / atoms));
                double af = 0.0;
                for (int t : ts) af += (double)b->n_atoms[t] * std::min<int>(b->n_frames[t], fmax);
                const int rep = (int)std::max(1.0, std::min(64.0, std::ceil(kTuneAtomFrames / af)));
                std::vector<uint32_t> na, nf, sd;
                std::vector<long long> off;
                std::vector<float> bx;
                for (int r = 0; r < rep; ++r)
                    for (int t : ts) {
                        const uint32_t f = (uint32_t)std::min<int>(b->n_frames[t], fmax);
                        na.push_back(b->n_atoms[t]);
                        nf.push_back(f);
                        off.push_back(b->store_off[t]);
                        sd.push_back(b->stride[t]);
                        if (!b->boxes_host.empty())
                            bx.insert(bx.end(), b->boxes_host.begin() + 9 * box0[t],
                                      b->boxes_host.begin() + 9 * (box0[t] + f));
                    }
                s = create_batch(b->ctx, na.size(), na.data(), nf.data(), b->bins, b->r_max, b->pbc,
                                 WMD_BATCH_RDF | WMD_BATCH_NO_GRAPH, off.data(), sd.data());
                s->boxes_host = bx;
                auto time_cfg = [&](const TuneVal &c) -> double {
                    std::fill(s->tune_k.begin(), s->tune_k.end(), c.k);
                    std::fill(s->tune_bh.begin(), s->tune_bh.end(), c.bh);
                    std::fill(s->tune_grab.begin(), s->tune_grab.end(), c.grab);
                    plan_batch(s);
                    const BatchTraj &d = s->traj[0];
                    const bool taken = c.k < 0 ? d.kind == KIND_BRUTE : (d.kind == KIND_CELLS && d.k == c.k);
                    if (!taken) return 1e300;
                    run_batch(s, d_store, nullptr, nullptr, nullptr, st, false);
                    WMD_HIP_CHECK(hipEventRecord(e0, st));
                    for (int r = 0; r < reps; ++r) run_batch(s, d_store, nullptr, nullptr, nullptr, st, false);
                    WMD_HIP_CHECK(hipEventRecord(e1, st));
                    WMD_HIP_CHECK(hipEventSynchronize(e1));
                    float ms = 0.f;
                    WMD_HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
                    return ms;
                };
                // coordinate search: enumeration, then home block (brute
                // force), then units per task
                double best = 1e300;
                choice = TuneVal{-1, 0, 0};
                for (int cand : {-1, 1, 2, 3}) {
                    const double ms = time_cfg(TuneVal{cand, 0, 0});
                    if (ms < best) { best = ms; choice = TuneVal{cand, 0, 0}; }
                }
                if (choice.k < 0)
                    for (int bh : {128, 256}) {
                        const double ms = time_cfg(TuneVal{-1, bh, 0});
                        if (ms < best) { best = ms; choice.bh = bh; }
                    }
                for (int g : {1, 4, 8}) {
                    const double ms = time_cfg(TuneVal{choice.k, choice.bh, g});
                    if (ms < best) { best = ms; choice.grab = g; }
                }
                delete s;
                s = nullptr;
                std::lock_guard<std::mutex> lk(g_tune_mu);
                g_tune[kv.first] = choice;
            }
            for (int t : kv.second) {
                b->tune_k[t] = choice.k;
                b->tune_bh[t] = choice.bh;
                b->tune_grab[t] = choice.grab;
            }
        }
    } catch (...) {
        delete s;
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        throw;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    plan_batch(b);
    WMD_API_END
}

wmd_status wmd_autotune_export(char *buf, size_t cap, size_t *len) {
    WMD_API_BEGIN
    std::ostringstream os;
    {
        std::lock_guard<std::mutex> lk(g_tune_mu);
        for (const auto &kv : g_tune)
            os << std::get<0>(kv.first) << ' ' << std::get<1>(kv.first) << ' ' << std::get<2>(kv.first)
               << ' ' << std::get<3>(kv.first) << ' ' << std::get<4>(kv.first) << ' '
               << std::get<5>(kv.first) << ' ' << kv.second.k << ' ' << kv.second.bh << ' '
               << kv.second.grab << '\n';
    }
    const std::string s = os.str();
    if (len) *len = s.size();
    if (buf && cap) {
        const size_t n = std::min(cap - 1, s.size());
        std::memcpy(buf, s.data(), n);
        buf[n] = '\0';
    }
    WMD_API_END
}

wmd_status wmd_autotune_import(const char *text) {
    WMD_API_BEGIN
    WMD_REQUIRE(text, WMD_ERR_INVALID, "null text");
    std::istringstream is(text);
    std::string line;
    std::map<TuneKey, TuneVal> add;
    while (std::getline(is, line)) {
        if (line.find_first_not_of(" \t\r") == std::string::npos) continue;
        std::istringstream ls(line);
        std::string arch;
        int pbc, bins, nb, db, k;
        uint32_t rb;
        WMD_REQUIRE(bool(ls >> arch >> pbc >> bins >> rb >> nb >> db >> k), WMD_ERR_PARSE,
                    "malformed autotune line: " + line);
        int bh = 0, g = 0;
        if (ls >> bh) {  // the seven-field form (k only) keeps the default home block and task size
            WMD_REQUIRE(bool(ls >> g), WMD_ERR_PARSE, "malformed autotune line: " + line);
        }
        std::string rest;
        WMD_REQUIRE(!(ls >> rest), WMD_ERR_PARSE, "malformed autotune line: " + line);
        WMD_REQUIRE(k == -1 || (k >= 1 && k <= 3), WMD_ERR_PARSE, "autotune choice out of range: " + line);
        WMD_REQUIRE(bh == 0 || (bh % 64 == 0 && bh >= 64 && bh <= kMaxBruteHome), WMD_ERR_PARSE,
                    "autotune home block out of range: " + line);
        WMD_REQUIRE(g >= 0 && g <= kMaxGrab, WMD_ERR_PARSE, "autotune task size out of range: " + line);
        add[TuneKey(arch, pbc, bins, rb, nb, db)] = TuneVal{k, bh, g};
    }
    std::lock_guard<std::mutex> lk(g_tune_mu);
    for (auto &kv : add) g_tune[kv.first] = kv.second;
    WMD_API_END
}

void wmd_batch_destroy(wmd_batch *b) {
    if (!b) return;
    bind(b->ctx);
    (void)hipDeviceSynchronize();
    delete b;
}

}  // extern "C"
