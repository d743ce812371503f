// RDF pair-distance histogram for gfx950: the engine behind RdfPlan's device
// branch (crates/traj-engine/src/plans/rdf.rs:743-791, 882-941) replacing
// GpuContext::rdf_accum (crates/traj-gpu/src/cuda/structure_polymer.rs:246-313)
// and the rdf_hist kernel (crates/traj-kernels/src/kernels/part4.cu:29-70).
//
// Exactness.  Each pair's squared distance s is formed with the reference's f32
// operations in the reference's order (rdf.rs:921-929, rdf_distance :473-478;
// built with -ffp-contract=off).  The reference then takes r = sqrt(s) (IEEE),
// tests r < r_max and bins (r / bin_width) as usize.  That map s -> bin is a
// monotone step function of the f32 value s, so the host precomputes its exact
// breakpoints T[k] = min{ s : bin(s) >= k } by bisection over f32 bit patterns
// with IEEE sqrtf / division.  A pair is counted iff s < T[bins]; its bin comes
// from a fast v_sqrt estimate corrected by one comparison against T on each
// side.  Counts are therefore bit-identical to the reference with neither the
// IEEE divide nor the IEEE square root in the pair loop.
//
// Minimum image.  When r_max < half the smallest plane spacing (less 1e-5
// relative), every in-range pair's image index round(dx/L) is far from a .5
// tie, so rint(dx * (1/L)) gives the same integer as the reference's
// roundf(dx / L) and the wrapped dx (dx - n*L, not fused) is bit-identical; a
// pair where the two could differ lies near L/2 > r_max and is not counted by
// either.  Otherwise the EXACT variant evaluates roundf(dx / L) with IEEE
// division.  The triclinic expression is the reference's f32 device formula
// wrap_triclinic (part1.cu:12-27), without FMA contraction.
//
// Enumeration.  Counts are integer sums, so any enumeration that visits every
// ordered pair's multiset once reproduces the reference loop.  Two pair kernels:
//  * brute force over 256x256 atom tiles (upper triangle of tile pairs; diagonal
//    tiles as ordered pairs, off-diagonal as unordered pairs of weight 2);
//  * a periodic cell list, cells >= (r_max + 1e-2 A) / k wide along each plane
//    spacing, home cell x half stencil of (2k+1)^3 neighbour cells (unordered
//    pairs, weight 2) plus the home cell itself (ordered pairs, weight 1).
// Both are persistent: each workgroup keeps wave-private u32 sub-histograms in
// LDS for all the units it processes and merges them into the u64 global
// counts once at the end.  Units are assigned in contiguous ranges per XCD so
// that an XCD's L2 holds the frames it works on.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <unordered_map>

#include "wmd_internal.hpp"

using namespace wmd;

namespace {

constexpr int kBlock = 256;
constexpr int kWaves = kBlock / 64;
constexpr int kTile = 256;        // brute-force tile
constexpr int kMaxJ = 1024;       // cell-list j atoms staged per pass
constexpr int kMaxStencil = 172;  // (7^3 - 1) / 2 + 1 for k = 3
constexpr uint32_t kPad = 0xFFFFFFFFu;

struct FrameBox {
    float lx, ly, lz, ilx, ily, ilz;
    float r[9];  // inverse rows: f_d = r[0*3+d]*dx + r[1*3+d]*dy + r[2*3+d]*dz
    float c[9];  // box rows: d_d = f0*c[0*3+d] + f1*c[1*3+d] + f2*c[2*3+d]
};

struct HistParams {
    int bins;
    int hist_copies;  // 4 = one sub-histogram per wave, 1 = shared
    float cut;        // counted iff s < cut
    float inv_w;      // approximate 1 / bin_width for the bin estimate
    const float *thr; // [bins + 1]
    const float *elo; // [bins + 1] edge diagnostics (EDGE variants)
    const float *ehi;
    unsigned long long *counts;  // [bins]
    unsigned long long *edge;    // [1]
};

// -------------------------------------------------------------------------
// The per-pair squared distance, reference operation order.
// -------------------------------------------------------------------------
template <int PBC, bool EXACT>
__device__ __forceinline__ float pair_s(float xi, float yi, float zi, float xj, float yj, float zj,
                                        const FrameBox &b) {
    float dx = xj - xi;
    float dy = yj - yi;
    float dz = zj - zi;
    if constexpr (PBC == WMD_PBC_ORTHO) {
        if constexpr (EXACT) {
            dx -= roundf(dx / b.lx) * b.lx;
            dy -= roundf(dy / b.ly) * b.ly;
            dz -= roundf(dz / b.lz) * b.lz;
        } else {
            dx -= rintf(dx * b.ilx) * b.lx;
            dy -= rintf(dy * b.ily) * b.ly;
            dz -= rintf(dz * b.ilz) * b.lz;
        }
    } else if constexpr (PBC == WMD_PBC_TRICLINIC) {
        float f0 = b.r[0] * dx + b.r[3] * dy + b.r[6] * dz;
        float f1 = b.r[1] * dx + b.r[4] * dy + b.r[7] * dz;
        float f2 = b.r[2] * dx + b.r[5] * dy + b.r[8] * dz;
        if constexpr (EXACT) {
            f0 -= roundf(f0);
            f1 -= roundf(f1);
            f2 -= roundf(f2);
        } else {
            f0 -= rintf(f0);
            f1 -= rintf(f1);
            f2 -= rintf(f2);
        }
        dx = f0 * b.c[0] + f1 * b.c[3] + f2 * b.c[6];
        dy = f0 * b.c[1] + f1 * b.c[4] + f2 * b.c[7];
        dz = f0 * b.c[2] + f1 * b.c[5] + f2 * b.c[8];
    }
    return dx * dx + dy * dy + dz * dz;
}

// Exact bin of a counted pair (s < cut): estimate then one-step correction.
__device__ __forceinline__ int exact_bin(float s, const float *thr_lds, float inv_w, int bins) {
    float r_est = __builtin_amdgcn_sqrtf(s);
    int b = (int)(r_est * inv_w);
    b = b < bins - 1 ? b : bins - 1;
    float lo = thr_lds[b];
    float hi = thr_lds[b + 1];
    b += (int)(s >= hi) - (int)(s < lo);
    return b;
}

template <bool EDGE>
__device__ __forceinline__ void count_pair(float s, unsigned w, const HistParams &p,
                                           const float *thr_lds, const float *elo_lds,
                                           const float *ehi_lds, unsigned *hist,
                                           unsigned &edge_local) {
    if (s < p.cut) {
        int b = exact_bin(s, thr_lds, p.inv_w, p.bins);
        atomicAdd(&hist[b], w);
        if constexpr (EDGE) {
            bool near = (s >= elo_lds[b] && s < ehi_lds[b]) || (s >= elo_lds[b + 1] && s < ehi_lds[b + 1]);
            edge_local += near ? w : 0u;
        }
    }
}

// LDS layout shared by the pair kernels:
//   thr[bins+1] | elo[bins+1] | ehi[bins+1] (EDGE) | hist[copies][bins] | kernel-specific
struct LdsHist {
    float *thr, *elo, *ehi;
    unsigned *hist;
    char *rest;
};

__device__ __forceinline__ LdsHist lds_hist_layout(char *base, const HistParams &p, bool edge) {
    LdsHist L;
    size_t nb1 = (size_t)p.bins + 1;
    L.thr = (float *)base;
    L.elo = L.thr + nb1;
    L.ehi = L.elo + nb1;
    size_t nf = edge ? 3 * nb1 : nb1;
    size_t off = ((nf * 4 + 15) / 16) * 16;
    L.hist = (unsigned *)(base + off);
    size_t hb = (size_t)p.hist_copies * p.bins * 4;
    L.rest = base + off + ((hb + 15) / 16) * 16;
    return L;
}

template <bool EDGE>
__device__ void lds_hist_init(const LdsHist &L, const HistParams &p) {
    for (int i = threadIdx.x; i <= p.bins; i += kBlock) {
        L.thr[i] = p.thr[i];
        if constexpr (EDGE) {
            L.elo[i] = p.elo[i];
            L.ehi[i] = p.ehi[i];
        }
    }
    for (int i = threadIdx.x; i < p.hist_copies * p.bins; i += kBlock) L.hist[i] = 0u;
}

// Merges the LDS sub-histograms into the global u64 counts (and clears them).
template <bool EDGE>
__device__ void lds_hist_flush(const LdsHist &L, const HistParams &p, unsigned &edge_local) {
    __syncthreads();
    for (int i = threadIdx.x; i < p.bins; i += kBlock) {
        unsigned long long sum = 0;
        for (int c = 0; c < p.hist_copies; ++c) {
            sum += L.hist[c * p.bins + i];
            L.hist[c * p.bins + i] = 0u;
        }
        if (sum) atomicAdd(&p.counts[i], sum);
    }
    if constexpr (EDGE) {
        if (edge_local) atomicAdd(p.edge, (unsigned long long)edge_local);
        edge_local = 0;
    }
    __syncthreads();
}

// XCD-aware persistent schedule: WG g runs on XCD g % 8 (round-robin dispatch);
// each XCD owns a contiguous slice of the unit range.
struct UnitRange {
    long long begin, end, step;
};

__device__ __forceinline__ UnitRange xcd_units(long long n_units, int n_xcd) {
    int g = blockIdx.x;
    int G = gridDim.x;
    int xcd = g % n_xcd;
    int local = g / n_xcd;
    int per_xcd = G / n_xcd;  // host launches a multiple of n_xcd workgroups
    long long lo = n_units * xcd / n_xcd;
    long long hi = n_units * (xcd + 1) / n_xcd;
    return {lo + local, hi, (long long)per_xcd};
}

// -------------------------------------------------------------------------
// Brute force over 256-atom tiles of the IO selection.
// soa: [F][3][stride]; weights (WEIGHTED): packed mA | mB << 8 per IO atom.
// tile_pairs: (it, jt) with it <= jt, upper-triangle order.
// -------------------------------------------------------------------------
template <int PBC, bool EXACT, bool WEIGHTED, bool EDGE>
__global__ __launch_bounds__(kBlock) void rdf_brute_kernel(
    const float *__restrict__ soa, long long frame_stride, int stride, int n, int n_frames,
    const FrameBox *__restrict__ boxes, const unsigned *__restrict__ wts,
    const int2 *__restrict__ tile_pairs, int n_tile_pairs, long long frame0, HistParams p,
    int n_xcd, unsigned w_max) {
    extern __shared__ __attribute__((aligned(16))) char lds[];
    LdsHist L = lds_hist_layout(lds, p, EDGE);
    float *jx = (float *)L.rest;
    float *jy = jx + kTile;
    float *jz = jy + kTile;
    unsigned *jw = (unsigned *)(jz + kTile);
    lds_hist_init<EDGE>(L, p);
    unsigned edge_local = 0;
    const int wave = threadIdx.x >> 6;
    unsigned *hist = L.hist + (p.hist_copies == 1 ? 0 : wave) * p.bins;
    const long long n_units = (long long)n_frames * n_tile_pairs;
    UnitRange ur = xcd_units(n_units, n_xcd);
    unsigned long long budget = 0;
    __syncthreads();
    for (long long u = ur.begin; u < ur.end; u += ur.step) {
        const int f = (int)(u / n_tile_pairs);
        const int2 tp = tile_pairs[u - (long long)f * n_tile_pairs];
        const float *fx = soa + (frame0 + f) * frame_stride;
        const float *fy = fx + stride;
        const float *fz = fy + stride;
        const FrameBox b = boxes[f];
        const int i = tp.x * kTile + threadIdx.x;
        const float nanv = __builtin_nanf("");
        float xi = nanv, yi = nanv, zi = nanv;
        unsigned wi = 0;
        if (i < n) {
            xi = fx[i];
            yi = fy[i];
            zi = fz[i];
            if constexpr (WEIGHTED) wi = wts[i];
        }
        const int j0 = tp.y * kTile;
        const int nj = min(kTile, n - j0);
        __syncthreads();  // previous unit done with the j tile
        if ((int)threadIdx.x < nj) {
            jx[threadIdx.x] = fx[j0 + threadIdx.x];
            jy[threadIdx.x] = fy[j0 + threadIdx.x];
            jz[threadIdx.x] = fz[j0 + threadIdx.x];
            if constexpr (WEIGHTED) jw[threadIdx.x] = wts[j0 + threadIdx.x];
        }
        __syncthreads();
        if (tp.x == tp.y) {
            // ordered pairs (i, j), i != j
            for (int jj = 0; jj < nj; ++jj) {
                float s = pair_s<PBC, EXACT>(xi, yi, zi, jx[jj], jy[jj], jz[jj], b);
                if (jj == (int)threadIdx.x) s = nanv;
                unsigned w = 1u;
                if constexpr (WEIGHTED) w = (wi & 0xFFu) * (jw[jj] >> 8);
                count_pair<EDGE>(s, w, p, L.thr, L.elo, L.ehi, hist, edge_local);
            }
        } else {
            // unordered pairs: both orders at once
            for (int jj = 0; jj < nj; ++jj) {
                float s = pair_s<PBC, EXACT>(xi, yi, zi, jx[jj], jy[jj], jz[jj], b);
                unsigned w = 2u;
                if constexpr (WEIGHTED) {
                    unsigned wj = jw[jj];
                    w = (wi & 0xFFu) * (wj >> 8) + (wj & 0xFFu) * (wi >> 8);
                }
                count_pair<EDGE>(s, w, p, L.thr, L.elo, L.ehi, hist, edge_local);
            }
        }
        // at most kTile pairs per thread column and sub-histogram copy per unit
        budget += (unsigned long long)kTile * kTile * w_max;
        if (budget > 0x7FFFFFFFull) {
            budget = 0;
            lds_hist_flush<EDGE>(L, p, edge_local);
        }
    }
    lds_hist_flush<EDGE>(L, p, edge_local);
}

// -------------------------------------------------------------------------
// Cell list construction.
// cell key = (cz * ncy + cy) * ncx + cx from the wrapped fractional position.
// -------------------------------------------------------------------------
template <int PBC>
__device__ __forceinline__ int cell_of_pos(float x, float y, float z, const FrameBox &b, int ncx,
                                           int ncy, int ncz) {
    float f0, f1, f2;
    if constexpr (PBC == WMD_PBC_ORTHO) {
        f0 = x * b.ilx;
        f1 = y * b.ily;
        f2 = z * b.ilz;
    } else {
        f0 = b.r[0] * x + b.r[3] * y + b.r[6] * z;
        f1 = b.r[1] * x + b.r[4] * y + b.r[7] * z;
        f2 = b.r[2] * x + b.r[5] * y + b.r[8] * z;
    }
    f0 -= floorf(f0);
    f1 -= floorf(f1);
    f2 -= floorf(f2);
    int cx = min(max((int)(f0 * (float)ncx), 0), ncx - 1);
    int cy = min(max((int)(f1 * (float)ncy), 0), ncy - 1);
    int cz = min(max((int)(f2 * (float)ncz), 0), ncz - 1);
    return (cz * ncy + cy) * ncx + cx;
}

__global__ void fill_u32_kernel(unsigned *p, long long n, unsigned v) {
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    long long step = (long long)gridDim.x * blockDim.x;
    for (; i < n; i += step) p[i] = v;
}

__global__ void fill_u64_kernel(unsigned long long *p, long long n) {
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0ull;
}

template <int PBC>
__global__ __launch_bounds__(kBlock) void cell_assign_kernel(
    const float *__restrict__ soa, long long frame_stride, int stride, int n, int n_frames,
    long long frame0, const FrameBox *__restrict__ boxes, int ncx, int ncy, int ncz,
    unsigned *__restrict__ cell_of, unsigned *__restrict__ cell_slot, unsigned *__restrict__ cell_count) {
    const int ncell = ncx * ncy * ncz;
    long long total = (long long)n * n_frames;
    for (long long t = (long long)blockIdx.x * kBlock + threadIdx.x; t < total;
         t += (long long)gridDim.x * kBlock) {
        int f = (int)(t / n);
        int i = (int)(t - (long long)f * n);
        const float *fx = soa + (frame0 + f) * frame_stride;
        float x = fx[i], y = fx[stride + i], z = fx[2 * stride + i];
        int c = cell_of_pos<PBC>(x, y, z, boxes[f], ncx, ncy, ncz);
        unsigned slot = atomicAdd(&cell_count[(long long)f * ncell + c], 1u);
        cell_of[t] = (unsigned)c;
        cell_slot[t] = slot;
    }
}

// Exclusive scan of each frame's cell counts (one workgroup per frame);
// start has ncell + 1 entries per frame.
__global__ __launch_bounds__(kBlock) void cell_scan_kernel(const unsigned *__restrict__ count,
                                                           unsigned *__restrict__ start, int ncell) {
    __shared__ unsigned part[kBlock];
    const long long f = blockIdx.x;
    const unsigned *c = count + f * ncell;
    unsigned *s = start + f * (ncell + 1);
    int per = (ncell + kBlock - 1) / kBlock;
    int lo = threadIdx.x * per;
    int hi = min(lo + per, ncell);
    unsigned sum = 0;
    for (int k = lo; k < hi; ++k) sum += c[k];
    part[threadIdx.x] = sum;
    __syncthreads();
    // Hillis-Steele inclusive scan over 256 partial sums
    for (int off = 1; off < kBlock; off <<= 1) {
        unsigned v = threadIdx.x >= (unsigned)off ? part[threadIdx.x - off] : 0u;
        __syncthreads();
        part[threadIdx.x] += v;
        __syncthreads();
    }
    unsigned run = threadIdx.x ? part[threadIdx.x - 1] : 0u;
    for (int k = lo; k < hi; ++k) {
        s[k] = run;
        run += c[k];
    }
    if (threadIdx.x == kBlock - 1) s[ncell] = part[kBlock - 1];
}

// sorted[f][start[c] + slot] = (x, y, z, packed weights)
__global__ __launch_bounds__(kBlock) void cell_scatter_kernel(
    const float *__restrict__ soa, long long frame_stride, int stride, int n, int n_frames,
    long long frame0, const unsigned *__restrict__ cell_of, const unsigned *__restrict__ cell_slot,
    const unsigned *__restrict__ start, int ncell, const unsigned *__restrict__ wts,
    float4 *__restrict__ sorted) {
    long long total = (long long)n * n_frames;
    for (long long t = (long long)blockIdx.x * kBlock + threadIdx.x; t < total;
         t += (long long)gridDim.x * kBlock) {
        int f = (int)(t / n);
        int i = (int)(t - (long long)f * n);
        const float *fx = soa + (frame0 + f) * frame_stride;
        unsigned pos = start[(long long)f * (ncell + 1) + cell_of[t]] + cell_slot[t];
        float4 v;
        v.x = fx[i];
        v.y = fx[stride + i];
        v.z = fx[2 * stride + i];
        v.w = __uint_as_float(wts ? wts[i] : 0x0101u);
        sorted[(long long)f * n + pos] = v;
    }
}

// -------------------------------------------------------------------------
// Cell-list pair kernel.  Unit = (frame, home cell).  The j list is the home
// cell followed by the half-stencil neighbour cells, staged into LDS in
// passes of kMaxJ atoms.  Lanes: i = t % I over the home atoms (I = min(n_home,
// 256)), j phase = t / I.
// -------------------------------------------------------------------------
template <int PBC, bool EXACT, bool WEIGHTED, bool EDGE>
__global__ __launch_bounds__(kBlock) void rdf_cell_kernel(
    const float4 *__restrict__ sorted, const unsigned *__restrict__ start, int n, int n_frames,
    const FrameBox *__restrict__ boxes, int ncx, int ncy, int ncz,
    const int4 *__restrict__ offsets, int n_off, HistParams p, int n_xcd, unsigned w_max) {
    extern __shared__ __attribute__((aligned(16))) char lds[];
    __shared__ int wave_tot[kWaves];
    LdsHist L = lds_hist_layout(lds, p, EDGE);
    float4 *jpos = (float4 *)L.rest;
    int *c_start = (int *)(jpos + kMaxJ);
    int *c_pref = c_start + kMaxStencil;  // kMaxStencil + 1 entries
    lds_hist_init<EDGE>(L, p);
    unsigned edge_local = 0;
    const int wave = threadIdx.x >> 6;
    const int lane = threadIdx.x & 63;
    unsigned *hist = L.hist + (p.hist_copies == 1 ? 0 : wave) * p.bins;
    const int ncell = ncx * ncy * ncz;
    const int n_sten = n_off + 1;
    const long long n_units = (long long)n_frames * ncell;
    UnitRange ur = xcd_units(n_units, n_xcd);
    const float nanv = __builtin_nanf("");
    // Upper bound of what any u32 sub-histogram slot has received since the
    // last flush (workgroup-uniform); flush before it could wrap.
    unsigned long long budget = 0;
    const unsigned long long budget_cap = 0x7FFFFFFFull;
    __syncthreads();
    for (long long u = ur.begin; u < ur.end; u += ur.step) {
        const int f = (int)(u / ncell);
        const int h = (int)(u - (long long)f * ncell);
        const int hx = h % ncx;
        const int hy = (h / ncx) % ncy;
        const int hz = h / (ncx * ncy);
        const unsigned *fs = start + (long long)f * (ncell + 1);
        const float4 *fp = sorted + (long long)f * n;
        const FrameBox b = boxes[f];
        __syncthreads();  // previous unit done with LDS
        int cnt = 0;
        if ((int)threadIdx.x < n_sten) {
            int cid = h;
            if (threadIdx.x > 0) {
                int4 o = offsets[threadIdx.x - 1];
                int cx = hx + o.x, cy = hy + o.y, cz = hz + o.z;
                cx += cx < 0 ? ncx : 0; cx -= cx >= ncx ? ncx : 0;
                cy += cy < 0 ? ncy : 0; cy -= cy >= ncy ? ncy : 0;
                cz += cz < 0 ? ncz : 0; cz -= cz >= ncz ? ncz : 0;
                cid = (cz * ncy + cy) * ncx + cx;
            }
            int s0 = (int)fs[cid];
            cnt = (int)fs[cid + 1] - s0;
            c_start[threadIdx.x] = s0;
        }
        // block-wide exclusive prefix of the stencil cell counts
        int v = cnt;
        for (int off = 1; off < 64; off <<= 1) {
            int t = __shfl_up(v, off, 64);
            if (lane >= off) v += t;
        }
        if (lane == 63) wave_tot[wave] = v;
        __syncthreads();
        int add = 0;
        for (int w = 0; w < wave; ++w) add += wave_tot[w];
        if ((int)threadIdx.x < n_sten) c_pref[threadIdx.x + 1] = v + add;
        if (threadIdx.x == 0) c_pref[0] = 0;
        __syncthreads();
        const int n_home = c_pref[1];
        const int total = c_pref[n_sten];
        if (n_home == 0) continue;
        const int I = min(n_home, kBlock);
        const int J = kBlock / I;
        const int islot = (int)threadIdx.x % I;
        const int jph = (int)threadIdx.x / I;
        const bool lane_on = jph < J;
        for (int ib = 0; ib < n_home; ib += I) {
            const int i = ib + islot;
            float xi = nanv, yi = nanv, zi = nanv;
            unsigned wi = 0;
            if (lane_on && i < n_home) {
                float4 pi = fp[c_start[0] + i];
                xi = pi.x; yi = pi.y; zi = pi.z;
                wi = __float_as_uint(pi.w);
            }
            for (int c0 = 0; c0 < total; c0 += kMaxJ) {
                const int c1 = min(total, c0 + kMaxJ);
                __syncthreads();
                for (int idx = c0 + (int)threadIdx.x; idx < c1; idx += kBlock) {
                    // stencil cell holding list position idx: c_pref[lo] <= idx < c_pref[lo+1]
                    int lo = 0, hi = n_sten;
                    while (hi - lo > 1) {
                        int mid = (lo + hi) >> 1;
                        if (c_pref[mid] <= idx) lo = mid; else hi = mid;
                    }
                    jpos[idx - c0] = fp[c_start[lo] + (idx - c_pref[lo])];
                }
                __syncthreads();
                if (lane_on) {
                    // home cell: ordered pairs, skipping the atom itself
                    const int own_end = min(c1, n_home);
                    for (int jj = c0 + jph; jj < own_end; jj += J) {
                        float4 pj = jpos[jj - c0];
                        float s = pair_s<PBC, EXACT>(xi, yi, zi, pj.x, pj.y, pj.z, b);
                        if (jj == i) s = nanv;
                        unsigned w = 1u;
                        if constexpr (WEIGHTED) w = (wi & 0xFFu) * (__float_as_uint(pj.w) >> 8);
                        count_pair<EDGE>(s, w, p, L.thr, L.elo, L.ehi, hist, edge_local);
                    }
                    // neighbour cells: unordered pairs, both orders at once
                    for (int jj = max(c0, n_home) + jph; jj < c1; jj += J) {
                        float4 pj = jpos[jj - c0];
                        float s = pair_s<PBC, EXACT>(xi, yi, zi, pj.x, pj.y, pj.z, b);
                        unsigned w = 2u;
                        if constexpr (WEIGHTED) {
                            unsigned wj = __float_as_uint(pj.w);
                            w = (wi & 0xFFu) * (wj >> 8) + (wj & 0xFFu) * (wi >> 8);
                        }
                        count_pair<EDGE>(s, w, p, L.thr, L.elo, L.ehi, hist, edge_local);
                    }
                }
                // at most I * (c1 - c0) pairs per sub-histogram copy in this pass
                budget += (unsigned long long)I * (unsigned long long)(c1 - c0) * w_max;
                if (budget > budget_cap) {
                    budget = 0;
                    lds_hist_flush<EDGE>(L, p, edge_local);
                }
            }
        }
    }
    lds_hist_flush<EDGE>(L, p, edge_local);
}

// AoS [F][n_src][4] -> SoA [F][3][stride] through a gather list (kPad -> 0).
__global__ __launch_bounds__(kBlock) void gather_soa_kernel(const float4 *__restrict__ aos,
                                                            long long n_src, const unsigned *__restrict__ idx,
                                                            int n_store, int stride, int n_frames,
                                                            float *__restrict__ soa) {
    long long total = (long long)stride * n_frames;
    for (long long t = (long long)blockIdx.x * kBlock + threadIdx.x; t < total;
         t += (long long)gridDim.x * kBlock) {
        int f = (int)(t / stride);
        int k = (int)(t - (long long)f * stride);
        float x = 0.f, y = 0.f, z = 0.f;
        if (k < n_store) {
            unsigned a = idx ? idx[k] : (unsigned)k;
            if (a != kPad) {
                float4 v = aos[(long long)f * n_src + a];
                x = v.x; y = v.y; z = v.z;
            }
        }
        float *o = soa + (long long)f * 3 * stride;
        o[k] = x;
        o[stride + k] = y;
        o[2 * stride + k] = z;
    }
}

// ---------------------------------------------------------------------------
// Host-side exact thresholds.
// ---------------------------------------------------------------------------
static inline float bits_f(uint32_t u) {
    float f;
    std::memcpy(&f, &u, 4);
    return f;
}

// Reference key of a squared distance: its bin if counted, else bins.
// IEEE single sqrt and division (x86 SSE), each rounding once (rdf.rs:929-935).
static int ref_key(float s, float r_max, float w, int bins) {
    volatile float sv = s;
    volatile float r = std::sqrt((float)sv);
    if (!(r < r_max)) return bins;
    volatile float q = r / w;
    double qd = (double)q;
    if (qd >= (double)bins) return bins;
    return (int)qd;
}

// min{ f32 s >= 0 : pred(s) } by bisection over bit patterns (pred monotone).
template <class P> static float min_s(P pred) {
    uint32_t lo = 0, hi = 0x7F800000u;  // pred(+inf) assumed true
    if (pred(bits_f(lo))) return 0.0f;
    while (hi - lo > 1) {
        uint32_t mid = lo + (hi - lo) / 2;
        if (pred(bits_f(mid))) hi = mid; else lo = mid;
    }
    return bits_f(hi);
}

}  // namespace

// ---------------------------------------------------------------------------
// Plan handle
// ---------------------------------------------------------------------------
struct wmd_rdf {
    wmd_ctx *ctx = nullptr;
    std::vector<uint32_t> io;
    std::vector<unsigned> wts;  // packed mA | mB << 8 per IO atom
    bool weighted = false;
    unsigned w_max = 2;  // largest weight one pair can add (2 * max mA * max mB)
    int bins = 0;
    float r_max = 0.f, bin_width = 0.f, cut = 0.f;
    int pbc = 0;
    uint32_t flags = 0;
    std::vector<float> thr, elo, ehi;
    int cell_k = 0;        // 0 = automatic
    int last_path = 0;
    int last_nc[3] = {0, 0, 0};
    DevBuf d_counts, d_thr, d_elo, d_ehi, d_wts, d_io, d_boxes, d_aos, d_soa;
    DevBuf d_cell_of, d_cell_slot, d_cell_count, d_cell_start, d_sorted, d_offsets, d_tiles;
    HostBuf h_stage, h_boxes;
};

namespace {

int hist_copies_for(int bins) { return bins * 4 * kWaves <= 32768 ? kWaves : 1; }

size_t lds_hist_bytes(int bins, bool edge) {
    size_t nb1 = (size_t)bins + 1;
    size_t nf = edge ? 3 * nb1 : nb1;
    size_t off = ((nf * 4 + 15) / 16) * 16;
    size_t hb = (size_t)hist_copies_for(bins) * bins * 4;
    return off + ((hb + 15) / 16) * 16;
}

// Half stencil: offsets o in [-k, k]^3 with (oz, oy, ox) > (0, 0, 0) lexicographically.
std::vector<int4> half_stencil(int k) {
    std::vector<int4> out;
    for (int oz = -k; oz <= k; ++oz)
        for (int oy = -k; oy <= k; ++oy)
            for (int ox = -k; ox <= k; ++ox) {
                bool pos = oz > 0 || (oz == 0 && (oy > 0 || (oy == 0 && ox > 0)));
                if (pos) out.push_back(make_int4(ox, oy, oz, 0));
            }
    return out;
}

template <class K> int occupancy(K kernel, size_t lds) {
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, kernel, kBlock, lds) != hipSuccess ||
        nb < 1)
        nb = 1;
    return nb;
}

// Dispatch helpers over the template flags.
template <int PBC, bool EXACT, bool WEIGHTED, bool EDGE>
void launch_brute_t(wmd_rdf *h, hipStream_t st, const float *soa, long long frame_stride,
                    int stride, int n, int nf, long long frame0, const FrameBox *boxes,
                    const int2 *tiles, int n_tp, const HistParams &p) {
    auto kern = rdf_brute_kernel<PBC, EXACT, WEIGHTED, EDGE>;
    ProfScope prof(h->ctx, st, "rdf_pair");
    size_t lds = lds_hist_bytes(p.bins, EDGE) + kTile * 16;
    int occ = occupancy(kern, lds);
    long long n_units = (long long)nf * n_tp;
    int grid = h->ctx->cu_count * occ;
    grid = std::max(8, grid / 8 * 8);
    long long max_grid = (n_units + 7) / 8 * 8;
    if (grid > max_grid) grid = (int)std::max<long long>(8, max_grid);
    WMD_REQUIRE(lds <= h->ctx->lds_per_block, WMD_ERR_UNSUPPORTED, "too many bins for LDS");
    hipLaunchKernelGGL(kern, dim3(grid), dim3(kBlock), lds, st, soa, frame_stride, stride, n, nf,
                       boxes, (const unsigned *)h->d_wts.ptr, tiles, n_tp, frame0, p, 8,
                       h->w_max);
    WMD_HIP_CHECK(hipGetLastError());
}

template <int PBC, bool EXACT, bool WEIGHTED, bool EDGE>
void launch_cells_t(wmd_rdf *h, hipStream_t st, const float4 *sorted, const unsigned *start, int n,
                    int nf, const FrameBox *boxes, const int *nc, const int4 *offs, int n_off,
                    const HistParams &p) {
    auto kern = rdf_cell_kernel<PBC, EXACT, WEIGHTED, EDGE>;
    ProfScope prof(h->ctx, st, "rdf_pair");
    size_t lds = lds_hist_bytes(p.bins, EDGE) + kMaxJ * 16 + (2 * kMaxStencil + 4) * 4;
    int occ = occupancy(kern, lds);
    long long n_units = (long long)nf * nc[0] * nc[1] * nc[2];
    int grid = h->ctx->cu_count * occ;
    grid = std::max(8, grid / 8 * 8);
    long long max_grid = (n_units + 7) / 8 * 8;
    if (grid > max_grid) grid = (int)std::max<long long>(8, max_grid);
    WMD_REQUIRE(lds <= h->ctx->lds_per_block, WMD_ERR_UNSUPPORTED, "too many bins for LDS");
    hipLaunchKernelGGL(kern, dim3(grid), dim3(kBlock), lds, st, sorted, start, n, nf, boxes, nc[0],
                       nc[1], nc[2], offs, n_off, p, 8, h->w_max);
    WMD_HIP_CHECK(hipGetLastError());
}

template <bool WEIGHTED, bool EDGE, class... A>
void brute_by_pbc(int pbc, bool exact, A... a) {
    if (pbc == WMD_PBC_NONE) launch_brute_t<WMD_PBC_NONE, false, WEIGHTED, EDGE>(a...);
    else if (pbc == WMD_PBC_ORTHO) {
        if (exact) launch_brute_t<WMD_PBC_ORTHO, true, WEIGHTED, EDGE>(a...);
        else launch_brute_t<WMD_PBC_ORTHO, false, WEIGHTED, EDGE>(a...);
    } else {
        if (exact) launch_brute_t<WMD_PBC_TRICLINIC, true, WEIGHTED, EDGE>(a...);
        else launch_brute_t<WMD_PBC_TRICLINIC, false, WEIGHTED, EDGE>(a...);
    }
}

template <bool WEIGHTED, bool EDGE, class... A>
void cells_by_pbc(int pbc, A... a) {
    // the cell list is only used when r_max < half the plane spacing: fast wrap
    if (pbc == WMD_PBC_ORTHO) launch_cells_t<WMD_PBC_ORTHO, false, WEIGHTED, EDGE>(a...);
    else launch_cells_t<WMD_PBC_TRICLINIC, false, WEIGHTED, EDGE>(a...);
}

HistParams hist_params(wmd_rdf *h) {
    HistParams p;
    p.bins = h->bins;
    p.hist_copies = hist_copies_for(h->bins);
    p.cut = h->cut;
    p.inv_w = 1.0f / h->bin_width;
    p.thr = h->d_thr.as<float>();
    p.elo = h->d_elo.as<float>();
    p.ehi = h->d_ehi.as<float>();
    p.counts = h->d_counts.as<unsigned long long>();
    p.edge = h->d_counts.as<unsigned long long>() + h->bins;
    return p;
}

void grid1d(long long total, int *grid) {
    long long g = (total + kBlock - 1) / kBlock;
    if (g > 65536) g = 65536;
    if (g < 1) g = 1;
    *grid = (int)g;
}

// Core: accumulate frames of an IO-selection SoA store already on the device.
void accumulate_device(wmd_rdf *h, const float *d_soa, size_t stride, size_t n_frames,
                       const float *boxes_host, hipStream_t st) {
    const int n = (int)h->io.size();
    if (n_frames == 0 || h->bins == 0) return;
    WMD_REQUIRE(stride >= (size_t)n && stride % 4 == 0, WMD_ERR_INVALID,
                "stride must be a multiple of 4 and >= the IO selection size");
    WMD_REQUIRE(h->pbc == WMD_PBC_NONE || boxes_host, WMD_ERR_INVALID, "boxes required for PBC");
    // per-frame box parameters
    std::vector<FrameBox> fb(n_frames);
    double min_width[3] = {1e300, 1e300, 1e300};
    for (size_t f = 0; f < n_frames; ++f) {
        BoxInfo bi;
        if (h->pbc != WMD_PBC_NONE) {
            if (!make_box(h->pbc, boxes_host + 9 * f, &bi))
                fail(WMD_ERR_MISMATCH, h->pbc == WMD_PBC_ORTHO
                                           ? "RDF requires a positive orthorhombic box"
                                           : "box matrix not invertible");
            for (int d = 0; d < 3; ++d) min_width[d] = std::min(min_width[d], bi.width[d]);
        } else {
            std::memset(&bi, 0, sizeof(bi));
        }
        FrameBox &b = fb[f];
        b.lx = bi.lx; b.ly = bi.ly; b.lz = bi.lz;
        b.ilx = bi.lx > 0 ? 1.0f / bi.lx : 0.f;
        b.ily = bi.ly > 0 ? 1.0f / bi.ly : 0.f;
        b.ilz = bi.lz > 0 ? 1.0f / bi.lz : 0.f;
        if (h->pbc == WMD_PBC_ORTHO) {
            // cell assignment uses the inverse rows for both kinds
            for (int i = 0; i < 9; ++i) b.r[i] = 0.f, b.c[i] = 0.f;
            b.r[0] = b.ilx; b.r[4] = b.ily; b.r[8] = b.ilz;
            b.c[0] = b.lx; b.c[4] = b.ly; b.c[8] = b.lz;
        } else {
            for (int i = 0; i < 9; ++i) b.r[i] = bi.r[i], b.c[i] = bi.c[i];
        }
    }
    size_t box_bytes = sizeof(FrameBox) * n_frames;
    void *hb = h->h_boxes.reserve(box_bytes);
    // the pinned buffer may still be read by an earlier async copy
    WMD_HIP_CHECK(hipStreamSynchronize(st));
    std::memcpy(hb, fb.data(), box_bytes);
    FrameBox *d_boxes = (FrameBox *)h->d_boxes.reserve(box_bytes);
    WMD_HIP_CHECK(hipMemcpyAsync(d_boxes, hb, box_bytes, hipMemcpyHostToDevice, st));

    const bool edge = h->flags & WMD_RDF_EDGE_DIAGNOSTICS;
    HistParams p = hist_params(h);
    double mw = std::min(min_width[0], std::min(min_width[1], min_width[2]));
    bool fast_wrap = h->pbc == WMD_PBC_NONE || (double)h->r_max <= 0.5 * mw * (1.0 - 1e-5);

    // ---- choose the enumeration -------------------------------------------
    int k_use = 0, nc[3] = {0, 0, 0};
    if (h->pbc != WMD_PBC_NONE && fast_wrap && !(h->flags & WMD_RDF_FORCE_BRUTE) && n >= 64) {
        int ks[3] = {2, 1, 3};
        if (h->cell_k >= 1 && h->cell_k <= 3) { ks[0] = h->cell_k; ks[1] = 0; ks[2] = 0; }
        for (int kk : ks) {
            if (kk == 0) continue;
            double reach = ((double)h->r_max + 1e-2) / kk;
            int c[3];
            bool ok = true;
            for (int d = 0; d < 3; ++d) {
                double m = std::floor(min_width[d] / reach);
                c[d] = (int)std::min(m, 96.0);
                if (c[d] < 2 * kk + 1) ok = false;
            }
            if (!ok) continue;
            k_use = kk;
            nc[0] = c[0]; nc[1] = c[1]; nc[2] = c[2];
            break;
        }
        // a cell list only pays off with several atoms per stencil
        if (k_use && (double)n / ((double)nc[0] * nc[1] * nc[2]) * std::pow(2 * k_use + 1, 3) < 48.0)
            k_use = 0;
    }

    if (k_use == 0) {
        // ---- brute force ----------------------------------------------------
        int tiles = (n + kTile - 1) / kTile;
        int n_tp = tiles * (tiles + 1) / 2;
        std::vector<int2> tp;
        tp.reserve(n_tp);
        for (int i = 0; i < tiles; ++i)
            for (int j = i; j < tiles; ++j) tp.push_back(make_int2(i, j));
        int2 *d_tp = (int2 *)h->d_tiles.reserve(sizeof(int2) * n_tp);
        WMD_HIP_CHECK(hipMemcpyAsync(d_tp, tp.data(), sizeof(int2) * n_tp, hipMemcpyHostToDevice, st));
        // the vector dies at scope exit: make the copy complete first
        WMD_HIP_CHECK(hipStreamSynchronize(st));
        long long fs = 3 * (long long)stride;
        if (h->weighted) {
            if (edge) brute_by_pbc<true, true>(h->pbc, !fast_wrap, h, st, d_soa, fs, (int)stride, n, (int)n_frames, 0LL, (const FrameBox *)d_boxes, (const int2 *)d_tp, n_tp, p);
            else brute_by_pbc<true, false>(h->pbc, !fast_wrap, h, st, d_soa, fs, (int)stride, n, (int)n_frames, 0LL, (const FrameBox *)d_boxes, (const int2 *)d_tp, n_tp, p);
        } else {
            if (edge) brute_by_pbc<false, true>(h->pbc, !fast_wrap, h, st, d_soa, fs, (int)stride, n, (int)n_frames, 0LL, (const FrameBox *)d_boxes, (const int2 *)d_tp, n_tp, p);
            else brute_by_pbc<false, false>(h->pbc, !fast_wrap, h, st, d_soa, fs, (int)stride, n, (int)n_frames, 0LL, (const FrameBox *)d_boxes, (const int2 *)d_tp, n_tp, p);
        }
        h->last_path = 1;
        h->last_nc[0] = h->last_nc[1] = h->last_nc[2] = 0;
        return;
    }

    // ---- cell list ----------------------------------------------------------
    std::vector<int4> offs = half_stencil(k_use);
    int n_off = (int)offs.size();
    int4 *d_offs = (int4 *)h->d_offsets.reserve(sizeof(int4) * n_off);
    WMD_HIP_CHECK(hipMemcpyAsync(d_offs, offs.data(), sizeof(int4) * n_off, hipMemcpyHostToDevice, st));
    WMD_HIP_CHECK(hipStreamSynchronize(st));
    const int ncell = nc[0] * nc[1] * nc[2];
    // frames per batch: bound the scratch (sorted float4 + 2 u32 per atom)
    size_t per_frame = (size_t)n * (16 + 8) + (size_t)(2 * ncell + 1) * 4;
    size_t batch = std::max<size_t>(1, std::min<size_t>(n_frames, (size_t)(768ull << 20) / per_frame));
    unsigned *d_cell_of = (unsigned *)h->d_cell_of.reserve((size_t)n * batch * 4);
    unsigned *d_cell_slot = (unsigned *)h->d_cell_slot.reserve((size_t)n * batch * 4);
    unsigned *d_count = (unsigned *)h->d_cell_count.reserve((size_t)ncell * batch * 4);
    unsigned *d_start = (unsigned *)h->d_cell_start.reserve((size_t)(ncell + 1) * batch * 4);
    float4 *d_sorted = (float4 *)h->d_sorted.reserve((size_t)n * batch * 16);
    const unsigned *d_w = h->weighted ? h->d_wts.as<unsigned>() : nullptr;
    long long fs = 3 * (long long)stride;
    for (size_t f0 = 0; f0 < n_frames; f0 += batch) {
        int nf = (int)std::min(batch, n_frames - f0);
        int g;
        const FrameBox *bx = (const FrameBox *)d_boxes + f0;
        {
        ProfScope build_prof(h->ctx, st, "cell_build");
        grid1d((long long)ncell * nf, &g);
        hipLaunchKernelGGL(fill_u32_kernel, dim3(g), dim3(kBlock), 0, st, d_count, (long long)ncell * nf, 0u);
        WMD_HIP_CHECK(hipGetLastError());
        grid1d((long long)n * nf, &g);
        if (h->pbc == WMD_PBC_ORTHO)
            hipLaunchKernelGGL(cell_assign_kernel<WMD_PBC_ORTHO>, dim3(g), dim3(kBlock), 0, st, d_soa, fs, (int)stride, n, nf, (long long)f0, bx, nc[0], nc[1], nc[2], d_cell_of, d_cell_slot, d_count);
        else
            hipLaunchKernelGGL(cell_assign_kernel<WMD_PBC_TRICLINIC>, dim3(g), dim3(kBlock), 0, st, d_soa, fs, (int)stride, n, nf, (long long)f0, bx, nc[0], nc[1], nc[2], d_cell_of, d_cell_slot, d_count);
        WMD_HIP_CHECK(hipGetLastError());
        hipLaunchKernelGGL(cell_scan_kernel, dim3(nf), dim3(kBlock), 0, st, d_count, d_start, ncell);
        WMD_HIP_CHECK(hipGetLastError());
        hipLaunchKernelGGL(cell_scatter_kernel, dim3(g), dim3(kBlock), 0, st, d_soa, fs, (int)stride, n, nf, (long long)f0, d_cell_of, d_cell_slot, d_start, ncell, d_w, d_sorted);
        WMD_HIP_CHECK(hipGetLastError());
        }
        if (h->weighted) {
            if (edge) cells_by_pbc<true, true>(h->pbc, h, st, (const float4 *)d_sorted, (const unsigned *)d_start, n, nf, bx, (const int *)nc, (const int4 *)d_offs, n_off, p);
            else cells_by_pbc<true, false>(h->pbc, h, st, (const float4 *)d_sorted, (const unsigned *)d_start, n, nf, bx, (const int *)nc, (const int4 *)d_offs, n_off, p);
        } else {
            if (edge) cells_by_pbc<false, true>(h->pbc, h, st, (const float4 *)d_sorted, (const unsigned *)d_start, n, nf, bx, (const int *)nc, (const int4 *)d_offs, n_off, p);
            else cells_by_pbc<false, false>(h->pbc, h, st, (const float4 *)d_sorted, (const unsigned *)d_start, n, nf, bx, (const int *)nc, (const int4 *)d_offs, n_off, p);
        }
    }
    h->last_path = 2;
    h->last_nc[0] = nc[0]; h->last_nc[1] = nc[1]; h->last_nc[2] = nc[2];
}

// Host AoS chunk -> device SoA store of the IO atoms -> accumulate.
void accumulate_host(wmd_rdf *h, const float *coords, size_t n_src, size_t n_frames,
                     const float *boxes, bool selected) {
    if (n_frames == 0 || h->bins == 0) return;
    WMD_REQUIRE(coords != nullptr, WMD_ERR_INVALID, "null coords");
    const int n = (int)h->io.size();
    if (!selected) {
        for (uint32_t a : h->io)
            WMD_REQUIRE(a < n_src, WMD_ERR_MISMATCH,
                        "selection index " + std::to_string(a) + " out of range for " +
                            std::to_string(n_src) + " atoms");
    }
    hipStream_t st = h->ctx->stream;
    size_t stride = round_up(std::max(n, 1), 4);
    // frames per staging batch (~256 MB of AoS)
    size_t fbytes = n_src * 16;
    size_t batch = std::max<size_t>(1, std::min<size_t>(n_frames, (256ull << 20) / std::max<size_t>(fbytes, 1)));
    float4 *d_aos = (float4 *)h->d_aos.reserve(batch * fbytes);
    float *d_soa = (float *)h->d_soa.reserve(batch * 3 * stride * 4);
    void *stage = h->h_stage.reserve(batch * fbytes);
    for (size_t f0 = 0; f0 < n_frames; f0 += batch) {
        size_t nf = std::min(batch, n_frames - f0);
        WMD_HIP_CHECK(hipStreamSynchronize(st));  // staging and store free again
        std::memcpy(stage, coords + f0 * n_src * 4, nf * fbytes);
        WMD_HIP_CHECK(hipMemcpyAsync(d_aos, stage, nf * fbytes, hipMemcpyHostToDevice, st));
        int g;
        grid1d((long long)stride * nf, &g);
        hipLaunchKernelGGL(gather_soa_kernel, dim3(g), dim3(kBlock), 0, st, (const float4 *)d_aos,
                           (long long)n_src, selected ? nullptr : h->d_io.as<unsigned>(), n,
                           (int)stride, (int)nf, d_soa);
        WMD_HIP_CHECK(hipGetLastError());
        accumulate_device(h, d_soa, stride, nf, boxes ? boxes + 9 * f0 : nullptr, st);
    }
}

}  // namespace

extern "C" {

wmd_status wmd_rdf_create(wmd_ctx *ctx, const uint32_t *sel_a, size_t n_a, const uint32_t *sel_b,
                          size_t n_b, size_t bins, float r_max, int pbc, uint32_t flags,
                          wmd_rdf **out) {
    WMD_API_BEGIN
    WMD_REQUIRE(ctx && out, WMD_ERR_INVALID, "null context or output handle");
    *out = nullptr;
    WMD_REQUIRE((n_a == 0 || sel_a) && (n_b == 0 || sel_b), WMD_ERR_INVALID, "null selection");
    WMD_REQUIRE(pbc >= WMD_PBC_NONE && pbc <= WMD_PBC_TRICLINIC, WMD_ERR_PARSE,
                "pbc must be none, orthorhombic or triclinic");
    WMD_REQUIRE(bins <= 12288, WMD_ERR_UNSUPPORTED, "at most 12288 bins are supported");
    WMD_REQUIRE(std::isfinite(r_max) && r_max > 0.0f, WMD_ERR_PARSE, "r_max must be finite and positive");
    bind(ctx);
    auto *h = new wmd_rdf();
    try {
        h->ctx = ctx;
        h->bins = (int)bins;
        h->r_max = r_max;
        h->pbc = pbc;
        h->flags = flags;
        // rdf.rs:498-526: order-preserving union, sel_a first
        std::vector<uint16_t> ma, mb;
        {
            std::unordered_map<uint32_t, size_t> pos;
            pos.reserve((n_a + n_b) * 2 + 1);
            for (size_t i = 0; i < n_a + n_b; ++i) {
                uint32_t a = i < n_a ? sel_a[i] : sel_b[i - n_a];
                if (pos.emplace(a, h->io.size()).second) h->io.push_back(a);
            }
            ma.assign(h->io.size(), 0);
            mb.assign(h->io.size(), 0);
            for (size_t i = 0; i < n_a; ++i) ma[pos[sel_a[i]]] += 1;
            for (size_t i = 0; i < n_b; ++i) mb[pos[sel_b[i]]] += 1;
        }
        bool unit = true;
        unsigned max_a = 0, max_b = 0;
        h->wts.resize(h->io.size());
        for (size_t i = 0; i < h->io.size(); ++i) {
            WMD_REQUIRE(ma[i] <= 64 && mb[i] <= 64, WMD_ERR_UNSUPPORTED,
                        "an atom repeated more than 64 times in a selection");
            h->wts[i] = (unsigned)ma[i] | ((unsigned)mb[i] << 8);
            if (ma[i] != 1 || mb[i] != 1) unit = false;
            max_a = std::max<unsigned>(max_a, ma[i]);
            max_b = std::max<unsigned>(max_b, mb[i]);
        }
        h->weighted = !unit;
        h->w_max = std::max(2u, 2u * max_a * max_b);
        WMD_REQUIRE(h->io.size() < (1u << 31), WMD_ERR_UNSUPPORTED, "selection too large");
        if (bins > 0) {
            // exact thresholds (rdf.rs:896, 929-935)
            h->bin_width = r_max / (float)bins;
            const float w = h->bin_width;
            const int B = (int)bins;
            h->thr.resize(bins + 1);
            h->thr[0] = 0.0f;
            for (int k = 1; k <= B; ++k)
                h->thr[k] = min_s([&](float s) { return ref_key(s, r_max, w, B) >= k; });
            h->cut = h->thr[B];
            h->elo.assign(bins + 1, 0.0f);
            h->ehi.assign(bins + 1, 0.0f);
            if (flags & WMD_RDF_EDGE_DIAGNOSTICS) {
                WMD_REQUIRE(w >= 1e-4f, WMD_ERR_UNSUPPORTED, "edge diagnostics need bin width >= 1e-4 A");
                for (int k = 1; k <= B; ++k) {
                    double e = (double)k * (double)w;
                    h->elo[k] = min_s([&](float s) { volatile float r = std::sqrt(s); return (double)r > e - 1e-5; });
                    h->ehi[k] = min_s([&](float s) { volatile float r = std::sqrt(s); return (double)r >= e + 1e-5; });
                }
            }
            h->d_counts.reserve((bins + 1) * 8);
            WMD_HIP_CHECK(hipMemsetAsync(h->d_counts.ptr, 0, (bins + 1) * 8, ctx->stream));
            h->d_thr.reserve((bins + 1) * 4);
            h->d_elo.reserve((bins + 1) * 4);
            h->d_ehi.reserve((bins + 1) * 4);
            WMD_HIP_CHECK(hipMemcpyAsync(h->d_thr.ptr, h->thr.data(), (bins + 1) * 4, hipMemcpyHostToDevice, ctx->stream));
            WMD_HIP_CHECK(hipMemcpyAsync(h->d_elo.ptr, h->elo.data(), (bins + 1) * 4, hipMemcpyHostToDevice, ctx->stream));
            WMD_HIP_CHECK(hipMemcpyAsync(h->d_ehi.ptr, h->ehi.data(), (bins + 1) * 4, hipMemcpyHostToDevice, ctx->stream));
        }
        size_t nio = std::max<size_t>(h->io.size(), 1);
        h->d_wts.reserve(nio * 4);
        h->d_io.reserve(nio * 4);
        if (!h->io.empty()) {
            WMD_HIP_CHECK(hipMemcpyAsync(h->d_wts.ptr, h->wts.data(), h->io.size() * 4, hipMemcpyHostToDevice, ctx->stream));
            WMD_HIP_CHECK(hipMemcpyAsync(h->d_io.ptr, h->io.data(), h->io.size() * 4, hipMemcpyHostToDevice, ctx->stream));
        }
        WMD_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    } catch (...) {
        delete h;
        throw;
    }
    *out = h;
    WMD_API_END
}

wmd_status wmd_rdf_io_selection(const wmd_rdf *h, uint32_t *out, size_t cap, size_t *n) {
    WMD_API_BEGIN
    WMD_REQUIRE(h && n, WMD_ERR_INVALID, "null handle");
    *n = h->io.size();
    if (out) {
        WMD_REQUIRE(cap >= h->io.size(), WMD_ERR_INVALID, "output buffer too small");
        std::memcpy(out, h->io.data(), h->io.size() * 4);
    }
    WMD_API_END
}

wmd_status wmd_rdf_accumulate(wmd_rdf *h, const float *coords, size_t n_atoms, size_t n_frames,
                              const float *boxes) {
    WMD_API_BEGIN
    WMD_REQUIRE(h != nullptr, WMD_ERR_INVALID, "null handle");
    bind(h->ctx);
    accumulate_host(h, coords, n_atoms, n_frames, boxes, false);
    WMD_API_END
}

wmd_status wmd_rdf_accumulate_selected(wmd_rdf *h, const float *coords, size_t n_frames,
                                       const float *boxes) {
    WMD_API_BEGIN
    WMD_REQUIRE(h != nullptr, WMD_ERR_INVALID, "null handle");
    bind(h->ctx);
    accumulate_host(h, coords, h->io.size(), n_frames, boxes, true);
    WMD_API_END
}

wmd_status wmd_rdf_accumulate_device(wmd_rdf *h, const float *d_soa, size_t stride,
                                     size_t n_frames, const float *boxes, void *stream) {
    WMD_API_BEGIN
    WMD_REQUIRE(h != nullptr, WMD_ERR_INVALID, "null handle");
    WMD_REQUIRE(d_soa != nullptr || n_frames == 0, WMD_ERR_INVALID, "null device store");
    bind(h->ctx);
    hipStream_t st = pick_stream(h->ctx, stream);
    accumulate_device(h, d_soa, stride, n_frames, boxes, st);
    WMD_API_END
}

wmd_status wmd_rdf_read(wmd_rdf *h, uint64_t *counts, uint64_t *edge) {
    WMD_API_BEGIN
    WMD_REQUIRE(h != nullptr, WMD_ERR_INVALID, "null handle");
    bind(h->ctx);
    if (h->bins == 0) {
        if (edge) *edge = 0;
        return WMD_OK;
    }
    WMD_REQUIRE(counts != nullptr, WMD_ERR_INVALID, "null counts buffer");
    std::vector<uint64_t> tmp(h->bins + 1);
    WMD_HIP_CHECK(hipDeviceSynchronize());
    WMD_HIP_CHECK(hipMemcpy(tmp.data(), h->d_counts.ptr, (h->bins + 1) * 8, hipMemcpyDeviceToHost));
    std::memcpy(counts, tmp.data(), h->bins * 8);
    if (edge) *edge = tmp[h->bins];
    WMD_API_END
}

wmd_status wmd_rdf_reset(wmd_rdf *h) {
    WMD_API_BEGIN
    WMD_REQUIRE(h != nullptr, WMD_ERR_INVALID, "null handle");
    bind(h->ctx);
    if (h->bins) {
        int g = (h->bins + 1 + kBlock - 1) / kBlock;
        hipLaunchKernelGGL(fill_u64_kernel, dim3(g), dim3(kBlock), 0, h->ctx->stream,
                           h->d_counts.as<unsigned long long>(), (long long)h->bins + 1);
        WMD_HIP_CHECK(hipGetLastError());
        WMD_HIP_CHECK(hipStreamSynchronize(h->ctx->stream));
    }
    WMD_API_END
}

void wmd_rdf_destroy(wmd_rdf *h) {
    if (!h) return;
    (void)hipSetDevice(h->ctx->device);
    // work may still be queued on a caller's stream: wait for the whole device
    (void)hipDeviceSynchronize();
    delete h;
}

wmd_status wmd_rdf_last_path(const wmd_rdf *h, int *path, int *nc) {
    WMD_API_BEGIN
    WMD_REQUIRE(h && path, WMD_ERR_INVALID, "null handle");
    *path = h->last_path;
    if (nc) for (int d = 0; d < 3; ++d) nc[d] = h->last_nc[d];
    WMD_API_END
}

wmd_status wmd_rdf_set_tuning(wmd_rdf *h, int cell_k, int block_threads) {
    WMD_API_BEGIN
    WMD_REQUIRE(h != nullptr, WMD_ERR_INVALID, "null handle");
    WMD_REQUIRE(cell_k >= 0 && cell_k <= 3, WMD_ERR_INVALID, "cell_k must be 0..3");
    WMD_REQUIRE(block_threads == 0 || block_threads == kBlock, WMD_ERR_UNSUPPORTED,
                "only 256-thread workgroups are built");
    h->cell_k = cell_k;
    WMD_API_END
}

}  // extern "C"
