// Shared core of the RDF engine (included by wmd_rdf.hip and wmd_batch.hip):
// the per-frame box parameters, the exact-threshold histogram parameters and
// their LDS layout, the reference-order pair expressions, the brute-force and
// cell-list pair kernels with their unit loop, the cell-list construction
// kernels, and the host helpers that choose and build a cell grid.  The
// exactness and enumeration arguments are stated at the top of wmd_rdf.hip.
#pragma once

#include <algorithm>
#include <cmath>
#include <cstring>
#include <unordered_map>

#include "wmd_internal.hpp"

using namespace wmd;

namespace {

constexpr int kBlock = 256;
constexpr int kTile = 256;        // brute-force tile
constexpr int kMaxTab = 8192;     // s-bits bin table entries (u16)
constexpr uint32_t kPad = 0xFFFFFFFFu;
constexpr int kImageLimit = 8;  // |image index| beyond which a frame uses the rint path
constexpr int kImageBias = 32;  // image indices are stored as m + 32 in one byte

struct FrameBox {
    float lx, ly, lz, ilx, ily, ilz;
    float r[9];  // inverse rows: f_d = r[0*3+d]*dx + r[1*3+d]*dy + r[2*3+d]*dz
    float c[9];  // box rows: d_d = f0*c[0*3+d] + f1*c[1*3+d] + f2*c[2*3+d]
};

struct HistParams {
    int bins;
    int hist_copies;  // 4 = one sub-histogram per wave, 1 = shared
    float cut;        // counted iff s < cut
    float inv_w;      // approximate 1 / bin_width for the sqrt estimate
    const float *thr; // [bins + 1]
    const float *elo; // [bins + 1] edge diagnostics (EDGE variants)
    const float *ehi;
    const unsigned short *tab;  // [tab_n] s-bits bin table (TABLE variants)
    const uint2 *tab2;          // [tab_n] per bucket (split threshold bits, 8 x bin below it)
    int tab_n, tab_sh, tab_base;
    unsigned long long *counts;  // [bins - skip]: counts[b - skip] gets local bin b
    unsigned long long *edge;    // [2]: edge pairs, pair evaluations
    // Bin slices (a histogram larger than LDS, counted one slice of bins per
    // pass): local bin b is global bin b + shift of the plain tables, local
    // bins below `skip` collect the pairs under the slice and are dropped.
    // shift = skip = 0 for an unsliced histogram.
    int shift = 0, skip = 0;
    int count_evals = 1;  // EDGE variants: add the pair evaluations (first pass only)
};

// LDS layout shared by the pair kernels:
//   thr[bins+1] | elo, ehi (EDGE) | tab[tab_n] (TABLE) | hist[copies][bins] | kernel-specific
struct LdsHist {
    float *thr, *elo, *ehi;
    unsigned short *tab;
    unsigned *hist;
    char *rest;
};

__host__ __device__ inline size_t align16(size_t x) { return (x + 15) / 16 * 16; }

__host__ __device__ inline size_t lds_hist_size(int bins, int copies, bool edge, int tab_n) {
    size_t nb1 = (size_t)bins + 1;
    return align16(nb1 * 4 * (edge ? 3 : 1)) + align16((size_t)tab_n * 2) +
           align16((size_t)copies * bins * 4);
}

__device__ __forceinline__ LdsHist lds_hist_layout(char *base, const HistParams &p, bool edge,
                                                   bool table) {
    LdsHist L;
    size_t nb1 = (size_t)p.bins + 1;
    L.thr = (float *)base;
    L.elo = L.thr + nb1;
    L.ehi = L.elo + nb1;
    size_t off = align16(nb1 * 4 * (edge ? 3 : 1));
    L.tab = (unsigned short *)(base + off);
    off += align16((size_t)(table ? p.tab_n : 0) * 2);
    L.hist = (unsigned *)(base + off);
    off += align16((size_t)p.hist_copies * p.bins * 4);
    L.rest = base + off;
    return L;
}

template <bool EDGE, bool TABLE>
__device__ void lds_hist_init(const LdsHist &L, const HistParams &p) {
    for (int i = threadIdx.x; i <= p.bins; i += kBlock) {
        L.thr[i] = p.thr[i];
        if constexpr (EDGE) {
            L.elo[i] = p.elo[i];
            L.ehi[i] = p.ehi[i];
        }
    }
    if constexpr (TABLE)
        for (int i = threadIdx.x; i < p.tab_n; i += kBlock) L.tab[i] = p.tab[i];
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
        if (sum && i >= p.skip) atomicAdd(&p.counts[i - p.skip], sum);
    }
    if constexpr (EDGE) {
        if (edge_local) atomicAdd(p.edge, (unsigned long long)edge_local);
        edge_local = 0;
    }
    __syncthreads();
}

// Exact bin of a counted pair (s < cut).
template <bool TABLE>
__device__ __forceinline__ int exact_bin(float s, const LdsHist &L, const HistParams &p) {
    if constexpr (TABLE) {
        // bucket of s: saturating (bits - base) >> sh; s < cut keeps it below tab_n
        const unsigned idx = __builtin_elementwise_sub_sat(__float_as_uint(s), (unsigned)p.tab_base) >> p.tab_sh;
        int b = L.tab[idx];
        b += (int)(s >= L.thr[b + 1]);
        return b;
    } else {
        float r_est = __builtin_amdgcn_sqrtf(s);
        int b = (int)(r_est * p.inv_w) - p.shift;
        b = b < p.bins - 1 ? b : p.bins - 1;
        b = b > 0 ? b : 0;
        float lo = L.thr[b];
        float hi = L.thr[b + 1];
        b += (int)(s >= hi) - (int)(s < lo);
        return b;
    }
}

template <bool EDGE, bool TABLE>
__device__ __forceinline__ void count_pair(float s, unsigned w, const HistParams &p,
                                           const LdsHist &L, unsigned *hist,
                                           unsigned &edge_local) {
    if (s < p.cut) {
        int b = exact_bin<TABLE>(s, L, p);
        atomicAdd(&hist[b], w);
        if constexpr (EDGE) {
            bool near = (s >= L.elo[b] && s < L.ehi[b]) || (s >= L.elo[b + 1] && s < L.ehi[b + 1]);
            edge_local += (near && b >= p.skip) ? w : 0u;
        }
    }
}

// -------------------------------------------------------------------------
// The per-pair squared distance (rint / exact wrap).  Orthorhombic: the
// reference's operations in its order, dx -= round(dx / L) * L.  Triclinic
// (which the reference RdfPlan refuses, rdf.rs:797-802): the image n is the
// reference's f32 fractional rounding n = round(f), f = inv . d (wrap_triclinic,
// part1.cu:12-27), and the lattice shift is taken off in Cartesian space,
// d -= (n0 a + n1 b) + n2 c, each component one f32 sum of f32 products.  The
// shift path below subtracts the same per-j vector, so a pair costs the same
// as an orthorhombic one; the oracle states this order.
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
        const float f0 = b.r[0] * dx + b.r[3] * dy + b.r[6] * dz;
        const float f1 = b.r[1] * dx + b.r[4] * dy + b.r[7] * dz;
        const float f2 = b.r[2] * dx + b.r[5] * dy + b.r[8] * dz;
        const float n0 = EXACT ? roundf(f0) : rintf(f0);
        const float n1 = EXACT ? roundf(f1) : rintf(f1);
        const float n2 = EXACT ? roundf(f2) : rintf(f2);
        dx -= n0 * b.c[0] + n1 * b.c[3] + n2 * b.c[6];
        dy -= n0 * b.c[1] + n1 * b.c[4] + n2 * b.c[7];
        dz -= n0 * b.c[2] + n1 * b.c[5] + n2 * b.c[8];
    }
    return dx * dx + dy * dy + dz * dz;
}

// The lattice shift of image n as pair_s subtracts it: n * L (orthorhombic,
// the reference's f32 product round(dx / L) * L) or (n0 a + n1 b) + n2 c.
template <int PBC>
__device__ __forceinline__ void image_shift(float n0, float n1, float n2, const FrameBox &b,
                                            float &sx, float &sy, float &sz) {
    if constexpr (PBC == WMD_PBC_ORTHO) {
        sx = n0 * b.lx;
        sy = n1 * b.ly;
        sz = n2 * b.lz;
    } else {
        sx = n0 * b.c[0] + n1 * b.c[3] + n2 * b.c[6];
        sy = n0 * b.c[1] + n1 * b.c[4] + n2 * b.c[7];
        sz = n0 * b.c[2] + n1 * b.c[5] + n2 * b.c[8];
    }
}

// Shift path: the image is known and (sx, sy, sz) = image_shift(n).
__device__ __forceinline__ float pair_s_shift(float xi, float yi, float zi, float xj, float yj,
                                              float zj, float sx, float sy, float sz) {
    const float dx = (xj - xi) - sx;
    const float dy = (yj - yi) - sy;
    const float dz = (zj - zi) - sz;
    return dx * dx + dy * dy + dz * dz;
}

// The reference wrap roundf(dx / L) for any pair, without the IEEE division
// in the common case.  q = dx * (1 / L) differs from the correctly rounded
// quotient by less than 5e-5 for |q| < 256, so when q lies more than 1e-4
// from a half-integer, rint(q) equals roundf(dx / L); otherwise (or for
// |q| >= 256, or NaN) the division is done.  Triclinic boxes use the
// reference device formula (roundf of products, no division).
__device__ __forceinline__ float wrap_checked(float d, float l, float il) {
    const float q = d * il;
    float n = rintf(q);
    if (!(fabsf(q - n) <= 0.4999f) || !(fabsf(q) < 256.f)) n = roundf(d / l);
    return d - n * l;
}

template <int PBC>
__device__ __forceinline__ float pair_s_checked(float xi, float yi, float zi, float xj, float yj,
                                                float zj, const FrameBox &b) {
    if constexpr (PBC == WMD_PBC_ORTHO) {
        const float dx = wrap_checked(xj - xi, b.lx, b.ilx);
        const float dy = wrap_checked(yj - yi, b.ly, b.ily);
        const float dz = wrap_checked(zj - zi, b.lz, b.ilz);
        return dx * dx + dy * dy + dz * dz;
    } else {
        return pair_s<PBC, true>(xi, yi, zi, xj, yj, zj, b);
    }
}

// Load through the constant address space: a wave-uniform load this way
// becomes a scalar (SMEM) load even in kernels that also write global memory
// (where the compiler cannot otherwise prove the data unchanged).  Only for
// data that no kernel writes while the loading kernel runs.  (The host pass
// of the compiler parses device functions too; it takes the plain load.)
template <class T>
__device__ __forceinline__ T uload(const T *p, long long i) {
#if defined(__HIP_DEVICE_COMPILE__)
    return ((const __attribute__((address_space(4))) T *)p)[i];
#else
    return p[i];
#endif
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
template <int PBC, bool EXACT, bool WEIGHTED, bool EDGE, bool TABLE>
__global__ __launch_bounds__(kBlock) void rdf_brute_kernel(
    const float *__restrict__ soa, long long frame_stride, int stride, int n, int n_frames,
    const FrameBox *__restrict__ boxes, const unsigned *__restrict__ wts,
    const int2 *__restrict__ tile_pairs, int n_tile_pairs, long long frame0, HistParams p,
    int n_xcd, unsigned w_max) {
    extern __shared__ __attribute__((aligned(16))) char lds[];
    LdsHist L = lds_hist_layout(lds, p, EDGE, TABLE);
    float *jx = (float *)L.rest;
    float *jy = jx + kTile;
    float *jz = jy + kTile;
    unsigned *jw = (unsigned *)(jz + kTile);
    lds_hist_init<EDGE, TABLE>(L, p);
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
                count_pair<EDGE, TABLE>(s, w, p, L, hist, edge_local);
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
                count_pair<EDGE, TABLE>(s, w, p, L, hist, edge_local);
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
// cell key = (cz * ncy + cy) * ncx + cx from the wrapped fractional position;
// meta word of a sorted atom: image index (mx, my, mz) + 32 as 3 x u8 in bits
// 0..23, selection multiplicities mA in bits 24..27 and mB in bits 28..31.
// -------------------------------------------------------------------------
template <int PBC>
__device__ __forceinline__ void frac_of(float x, float y, float z, const FrameBox &b, float *f) {
    if constexpr (PBC == WMD_PBC_ORTHO) {
        f[0] = x * b.ilx;
        f[1] = y * b.ily;
        f[2] = z * b.ilz;
    } else {
        f[0] = b.r[0] * x + b.r[3] * y + b.r[6] * z;
        f[1] = b.r[1] * x + b.r[4] * y + b.r[7] * z;
        f[2] = b.r[2] * x + b.r[5] * y + b.r[8] * z;
    }
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

// Cell of one atom and its image word ((m + 32) per byte, m = floor of the
// fractional coordinate); far_out: an image index beyond kImageLimit or a
// non-finite coordinate (the frame then takes the rint path).
template <int PBC>
__device__ __forceinline__ int assign_atom(float x, float y, float z, const FrameBox &b, int ncx,
                                           int ncy, int ncz, unsigned &img, bool &far_out) {
    const int nc[3] = {ncx, ncy, ncz};
    float fr[3];
    frac_of<PBC>(x, y, z, b, fr);
    int c[3];
    img = 0;
    far_out = false;
    for (int d = 0; d < 3; ++d) {
        float fl = floorf(fr[d]);
        float w = fr[d] - fl;  // in [0, 1) (NaN for non-finite input)
        int cc = (int)(w * (float)nc[d]);
        c[d] = min(max(cc, 0), nc[d] - 1);
        bool ok = fl >= -(float)kImageLimit && fl <= (float)kImageLimit;
        int m = ok ? (int)fl : 0;
        far_out |= !ok;
        img |= ((unsigned)(m + kImageBias) & 0xFFu) << (8 * d);
    }
    return (c[2] * ncy + c[1]) * ncx + c[0];
}

template <int PBC>
__global__ __launch_bounds__(kBlock) void cell_assign_kernel(
    const float *__restrict__ soa, long long frame_stride, int stride, int n, int n_frames,
    long long frame0, const FrameBox *__restrict__ boxes, int ncx, int ncy, int ncz,
    unsigned *__restrict__ cell_of, unsigned *__restrict__ cell_slot,
    unsigned *__restrict__ cell_count, unsigned *__restrict__ image_of,
    unsigned *__restrict__ frame_flag) {
    const int ncell = ncx * ncy * ncz;
    long long total = (long long)n * n_frames;
    for (long long t = (long long)blockIdx.x * kBlock + threadIdx.x; t < total;
         t += (long long)gridDim.x * kBlock) {
        int f = (int)(t / n);
        int i = (int)(t - (long long)f * n);
        const float *fx = soa + (frame0 + f) * frame_stride;
        unsigned img;
        bool far_out;
        const int cid = assign_atom<PBC>(fx[i], fx[stride + i], fx[2 * stride + i], boxes[f], ncx,
                                         ncy, ncz, img, far_out);
        if (far_out) atomicOr(&frame_flag[f], 1u);
        unsigned slot = atomicAdd(&cell_count[(long long)f * ncell + cid], 1u);
        cell_of[t] = (unsigned)cid;
        cell_slot[t] = slot;
        image_of[t] = img;
    }
}

// Exclusive scan of one frame's cell counts by one workgroup of kBlock
// threads; s has ncell + 1 entries.
__device__ __forceinline__ void scan_cells(const unsigned *__restrict__ c, unsigned *__restrict__ s,
                                           int ncell) {
    __shared__ unsigned part[kBlock];
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

// Exclusive scan of each frame's cell counts (one workgroup per frame);
// start has ncell + 1 entries per frame.
__global__ __launch_bounds__(kBlock) void cell_scan_kernel(const unsigned *__restrict__ count,
                                                           unsigned *__restrict__ start, int ncell) {
    const long long f = blockIdx.x;
    scan_cells(count + f * ncell, start + f * (ncell + 1), ncell);
}

// Fused cell build of one frame by one workgroup of kBuildBlock threads: a
// counting sort in LDS.  Pass 1 counts the atoms of every cell with LDS
// atomics; a block scan gives the cell starts (written to start[0..ncell]);
// pass 2 recomputes each atom's cell and takes its slot from an LDS cursor.
// The frame is read twice from HBM and the sorted copy written once, with no
// global atomics.  cnt: LDS [ncell] (ncell <= kBuildMaxCells).
constexpr int kBuildBlock = 1024;
constexpr int kBuildMaxCells = 16384;

constexpr int kBuildUnroll = 16;  // atoms per thread in flight in the per-frame build

__device__ __forceinline__ unsigned block_exclusive_scan_1024(unsigned v, unsigned *part,
                                                              unsigned *total) {
    part[threadIdx.x] = v;
    __syncthreads();
    for (int off = 1; off < kBuildBlock; off <<= 1) {
        const unsigned u = threadIdx.x >= (unsigned)off ? part[threadIdx.x - off] : 0u;
        __syncthreads();
        part[threadIdx.x] += u;
        __syncthreads();
    }
    const unsigned incl = part[threadIdx.x];
    *total = part[kBuildBlock - 1];
    __syncthreads();
    return incl - v;
}

template <int PBC>
__device__ __forceinline__ void build_frame(const float *__restrict__ X, int stride, int n,
                                            const FrameBox &b, int ncx, int ncy, int ncz,
                                            const unsigned *__restrict__ wts,
                                            unsigned *__restrict__ start, float4 *__restrict__ sorted,
                                            unsigned *__restrict__ order, unsigned *frame_flag,
                                            unsigned *cnt, unsigned *part, unsigned *far,
                                            float4 *__restrict__ bbox = nullptr) {
    const int ncell = ncx * ncy * ncz;
    for (int c = threadIdx.x; c < ncell; c += kBuildBlock) cnt[c] = 0u;
    if (threadIdx.x == 0) *far = 0u;
    __syncthreads();
    // both passes take kBuildUnroll atoms per thread at a time, loads first,
    // so that several loads of a thread are in flight together
    for (int i0 = threadIdx.x; i0 < n; i0 += kBuildUnroll * kBuildBlock) {
        float xs[kBuildUnroll], ys[kBuildUnroll], zs[kBuildUnroll];
#pragma unroll
        for (int u = 0; u < kBuildUnroll; ++u) {
            const int i = min(i0 + u * kBuildBlock, n - 1);
            xs[u] = X[i];
            ys[u] = X[stride + i];
            zs[u] = X[2 * stride + i];
        }
#pragma unroll
        for (int u = 0; u < kBuildUnroll; ++u) {
            if (i0 + u * kBuildBlock >= n) break;
            unsigned img;
            bool far_out;
            const int cid = assign_atom<PBC>(xs[u], ys[u], zs[u], b, ncx, ncy, ncz, img, far_out);
            atomicAdd(&cnt[cid], 1u);
            if (far_out) *far = 1u;
        }
    }
    __syncthreads();
    // exclusive scan of cnt: thread t owns cells [t * per, (t + 1) * per)
    const int per = (ncell + kBuildBlock - 1) / kBuildBlock;
    const int lo = min(ncell, (int)threadIdx.x * per), hi = min(ncell, lo + per);
    unsigned sum = 0;
    for (int c = lo; c < hi; ++c) sum += cnt[c];
    unsigned total;
    unsigned run = block_exclusive_scan_1024(sum, part, &total);
    for (int c = lo; c < hi; ++c) {
        const unsigned k = cnt[c];
        start[c] = run;
        cnt[c] = run;  // the cursor of pass 2
        run += k;
    }
    if (threadIdx.x == 0) {
        start[ncell] = total;
        *frame_flag = *far;
    }
    __syncthreads();
    for (int i0 = threadIdx.x; i0 < n; i0 += kBuildUnroll * kBuildBlock) {
        float xs[kBuildUnroll], ys[kBuildUnroll], zs[kBuildUnroll];
#pragma unroll
        for (int u = 0; u < kBuildUnroll; ++u) {
            const int i = min(i0 + u * kBuildBlock, n - 1);
            xs[u] = X[i];
            ys[u] = X[stride + i];
            zs[u] = X[2 * stride + i];
        }
#pragma unroll
        for (int u = 0; u < kBuildUnroll; ++u) {
            const int i = i0 + u * kBuildBlock;
            if (i >= n) break;
            unsigned img;
            bool far_out;
            const int cid = assign_atom<PBC>(xs[u], ys[u], zs[u], b, ncx, ncy, ncz, img, far_out);
            const unsigned pos = atomicAdd(&cnt[cid], 1u);
            const unsigned w = wts ? wts[i] : 0x0101u;  // mA | mB << 8
            const unsigned meta = (img & 0xFFFFFFu) | ((w & 0xFu) << 24) | (((w >> 8) & 0xFu) << 28);
            sorted[pos] = make_float4(xs[u], ys[u], zs[u], __uint_as_float(meta));
            if (order) order[i] = pos;  // store slot -> sorted position
        }
    }
    if (bbox) {
        // per cell: the bounding box of its atoms' raw coordinates and whether
        // they share one image index (bbox[2c] = min x, y, z and the flag
        // 1.0 when they do, bbox[2c + 1] = max x, y, z); the cell kernels read
        // it instead of reducing over the home atoms in every unit.  After
        // pass 2, cnt[c] is the end of cell c; this thread wrote start[c].
        // The first kBuildUnroll atoms of a cell are loaded together (the
        // last one repeated past the cell's end, harmless for a min or max).
        __syncthreads();
        for (int c = lo; c < hi; ++c) {
            const unsigned beg = start[c], end = cnt[c];
            if (beg == end) continue;
            float4 q[kBuildUnroll];
#pragma unroll
            for (int u = 0; u < kBuildUnroll; ++u) q[u] = sorted[min(beg + (unsigned)u, end - 1u)];
            const unsigned m0 = __float_as_uint(q[0].w) & 0xFFFFFFu;
            float mnx = q[0].x, mny = q[0].y, mnz = q[0].z, mxx = q[0].x, mxy = q[0].y, mxz = q[0].z;
            bool same = true;
#pragma unroll
            for (int u = 1; u < kBuildUnroll; ++u) {
                mnx = fminf(mnx, q[u].x); mny = fminf(mny, q[u].y); mnz = fminf(mnz, q[u].z);
                mxx = fmaxf(mxx, q[u].x); mxy = fmaxf(mxy, q[u].y); mxz = fmaxf(mxz, q[u].z);
                same = same && (__float_as_uint(q[u].w) & 0xFFFFFFu) == m0;
            }
            for (unsigned a = beg + kBuildUnroll; a < end; ++a) {
                const float4 r = sorted[a];
                mnx = fminf(mnx, r.x); mny = fminf(mny, r.y); mnz = fminf(mnz, r.z);
                mxx = fmaxf(mxx, r.x); mxy = fmaxf(mxy, r.y); mxz = fmaxf(mxz, r.z);
                same = same && (__float_as_uint(r.w) & 0xFFFFFFu) == m0;
            }
            bbox[2 * c] = make_float4(mnx, mny, mnz, same ? 1.0f : 0.0f);
            bbox[2 * c + 1] = make_float4(mxx, mxy, mxz, 0.0f);
        }
    }
}

template <int PBC>
__global__ __launch_bounds__(kBuildBlock) void cell_build_frame_kernel(
    const float *__restrict__ soa, long long frame_stride, int stride, int n, long long frame0,
    const FrameBox *__restrict__ boxes, int ncx, int ncy, int ncz, const unsigned *__restrict__ wts,
    unsigned *__restrict__ start, float4 *__restrict__ sorted, unsigned *__restrict__ order,
    unsigned *__restrict__ frame_flag, float4 *__restrict__ bbox) {
    extern __shared__ __attribute__((aligned(16))) char lds[];
    unsigned *part = (unsigned *)lds;
    unsigned *far = part + kBuildBlock;
    unsigned *cnt = far + 4;
    const long long f = blockIdx.x;
    const int ncell = ncx * ncy * ncz;
    build_frame<PBC>(soa + (frame0 + f) * frame_stride, stride, n, boxes[f], ncx, ncy, ncz, wts,
                     start + f * (ncell + 1), sorted + f * n, order ? order + f * n : nullptr,
                     frame_flag + f, cnt, part, far, bbox ? bbox + f * 2 * ncell : nullptr);
}

__host__ __device__ inline size_t build_lds_bytes(int ncell) {
    return (size_t)(kBuildBlock + 4 + ncell) * 4;
}

// sorted[f][start[c] + slot] = (x, y, z, meta)
__global__ __launch_bounds__(kBlock) void cell_scatter_kernel(
    const float *__restrict__ soa, long long frame_stride, int stride, int n, int n_frames,
    long long frame0, const unsigned *__restrict__ cell_of, const unsigned *__restrict__ cell_slot,
    const unsigned *__restrict__ image_of, const unsigned *__restrict__ start, int ncell,
    const unsigned *__restrict__ wts, float4 *__restrict__ sorted, unsigned *__restrict__ order) {
    long long total = (long long)n * n_frames;
    for (long long t = (long long)blockIdx.x * kBlock + threadIdx.x; t < total;
         t += (long long)gridDim.x * kBlock) {
        int f = (int)(t / n);
        int i = (int)(t - (long long)f * n);
        const float *fx = soa + (frame0 + f) * frame_stride;
        unsigned pos = start[(long long)f * (ncell + 1) + cell_of[t]] + cell_slot[t];
        unsigned w = wts ? wts[i] : 0x0101u;  // mA | mB << 8
        unsigned meta = (image_of[t] & 0xFFFFFFu) | ((w & 0xFu) << 24) | (((w >> 8) & 0xFu) << 28);
        float4 v;
        v.x = fx[i];
        v.y = fx[stride + i];
        v.z = fx[2 * stride + i];
        v.w = __uint_as_float(meta);
        sorted[(long long)f * n + pos] = v;
        if (order) order[t] = pos;  // store slot -> sorted position
    }
}


// -------------------------------------------------------------------------
// Cell-list pair kernel.  One wave per unit (frame, home cell); no block-level
// synchronisation between units.  The unit's j stream is the home atoms, then
// the neighbour cells of the half stencil taken as rows along x: for each
// (dy, dz) with dz > 0, or dz = 0 and dy > 0, the cells hx-k..hx+k, and in the
// home row the cells hx+1..hx+k.  Such a row is one contiguous run of the
// cell-sorted atoms (two runs where it wraps in x); runs that wrap the box are
// streamed after those that do not.  Lanes load 64 stream atoms at a time,
// drop neighbours farther than the cull radius from the home atoms' bounding
// box (shift path), and append the rest to a per-wave LDS ring; every full
// ring chunk of 64 j atoms is paired with all home atoms i, which are uniform
// across the wave and read with scalar loads.  Home atoms in the stream pair
// only with home atoms of lower index (unordered pairs, weight 2).  The
// histogram is a u64 LDS array shared by the workgroup's waves.
// -------------------------------------------------------------------------
constexpr int kCellBlock = 1024;  // 16 waves: two workgroups per CU, 8 waves per SIMD (registers capped for 8 waves)
constexpr int kCellWaves = kCellBlock / 64;
constexpr int kRing = 128;       // per-wave ring entries (float4 each)
constexpr int kRingStride = kRing + 2;  // per wave: the ring, then the unit's cull bounds
constexpr int kMaxTab2 = 5056;   // split-table entries the cell kernel keeps in LDS

struct CellLds {
    unsigned long long *hist;
    uint2 *tab2;
    float *thr, *elo, *ehi;
    float4 *ring;  // [kCellWaves][kRingStride]
};

__host__ __device__ inline size_t cell_lds_bytes(int bins, bool edge, int tab_n) {
    return align16((size_t)bins * 8) + align16((size_t)tab_n * 8) +
           align16(((size_t)bins + 1) * 4 * (edge ? 3 : 1)) + (size_t)kCellWaves * kRingStride * 16;
}

__device__ __forceinline__ CellLds cell_lds_layout(char *base, const HistParams &p, bool edge,
                                                   bool table) {
    CellLds L;
    const size_t nb1 = (size_t)p.bins + 1;
    size_t off = 0;
    L.hist = (unsigned long long *)base;
    off += align16((size_t)p.bins * 8);
    L.tab2 = (uint2 *)(base + off);
    off += align16((size_t)(table ? p.tab_n : 0) * 8);
    L.thr = (float *)(base + off);
    L.elo = L.thr + nb1;
    L.ehi = L.elo + nb1;
    off += align16(nb1 * 4 * (edge ? 3 : 1));
    L.ring = (float4 *)(base + off);
    return L;
}

// Byte offset (HB x bin) of the exact bin of a counted pair (s < cut) in a
// histogram of HB-byte counters; the table's offsets are scaled by HB.
template <bool TABLE, int HB = 8>
__device__ __forceinline__ unsigned cell_bin8(float s, const CellLds &L, const HistParams &p) {
    if constexpr (TABLE) {
        const unsigned t = __builtin_elementwise_sub_sat(__float_as_uint(s), (unsigned)p.tab_base);
        const uint2 e = L.tab2[t >> p.tab_sh];
        return e.y + (s >= __uint_as_float(e.x) ? (unsigned)HB : 0u);
    } else {
        float r_est = __builtin_amdgcn_sqrtf(s);
        int b = (int)(r_est * p.inv_w) - p.shift;
        b = b < p.bins - 1 ? b : p.bins - 1;
        b = b > 0 ? b : 0;
        b += (int)(s >= L.thr[b + 1]) - (int)(s < L.thr[b]);
        return (unsigned)b * (unsigned)HB;
    }
}

// Wave-wide min of (a0, a1, a2) and max of (b0, b1, b2); every lane must be
// active.  Six DPP row steps per value, each one v_min_f32 / v_max_f32 that
// reads the permuted operand (quad swaps, half-row and row mirrors, then
// row_bcast:15 into rows 1 and 3 and row_bcast:31 into rows 2 and 3, whose
// other rows keep their value), after which lane 63 holds the result.  The six
// chains are interleaved, so each DPP read comes five instructions after the
// write of its operand (two wait states are required).
#define WMD_BBOX_STEP(ctrl)                      \
    "v_min_f32_dpp %0, %0, %0 " ctrl "\n\t"      \
    "v_min_f32_dpp %1, %1, %1 " ctrl "\n\t"      \
    "v_min_f32_dpp %2, %2, %2 " ctrl "\n\t"      \
    "v_max_f32_dpp %3, %3, %3 " ctrl "\n\t"      \
    "v_max_f32_dpp %4, %4, %4 " ctrl "\n\t"      \
    "v_max_f32_dpp %5, %5, %5 " ctrl "\n\t"
__device__ __forceinline__ void wave_bbox(float &a0, float &a1, float &a2, float &b0, float &b1,
                                          float &b2) {
    asm volatile("s_nop 1\n\t"
                 WMD_BBOX_STEP("quad_perm:[1,0,3,2] row_mask:0xf bank_mask:0xf")
                 WMD_BBOX_STEP("quad_perm:[2,3,0,1] row_mask:0xf bank_mask:0xf")
                 WMD_BBOX_STEP("row_half_mirror row_mask:0xf bank_mask:0xf")
                 WMD_BBOX_STEP("row_mirror row_mask:0xf bank_mask:0xf")
                 WMD_BBOX_STEP("row_bcast:15 row_mask:0xa bank_mask:0xf")
                 WMD_BBOX_STEP("row_bcast:31 row_mask:0xc bank_mask:0xf")
                 "s_nop 1"
                 : "+v"(a0), "+v"(a1), "+v"(a2), "+v"(b0), "+v"(b1), "+v"(b2));
    a0 = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(a0), 63));
    a1 = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(a1), 63));
    a2 = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(a2), 63));
    b0 = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(b0), 63));
    b1 = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(b1), 63));
    b2 = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(b2), 63));
}
#undef WMD_BBOX_STEP

// Number of set bits of m below this lane.
__device__ __forceinline__ int lanes_below(unsigned long long m) {
    return (int)__builtin_amdgcn_mbcnt_hi((unsigned)(m >> 32),
                                          __builtin_amdgcn_mbcnt_lo((unsigned)m, 0u));
}

// Counts one pair (s, i's meta word) of a chunk into HB-byte counters.
template <bool WEIGHTED, bool EDGE, bool TABLE, int HB = 8>
__device__ __forceinline__ void cell_count(float s, unsigned mi, unsigned mAj, unsigned mBj,
                                           const HistParams &p, const CellLds &L,
                                           unsigned long long &edge_local) {
    if (s < p.cut) {
        unsigned w = 2u;
        if constexpr (WEIGHTED) w = ((mi >> 24) & 0xFu) * mBj + (mi >> 28) * mAj;
        const unsigned off = cell_bin8<TABLE, HB>(s, L, p);
        if constexpr (HB == 8)
            atomicAdd((unsigned long long *)((char *)L.hist + off), (unsigned long long)w);
        else
            atomicAdd((unsigned *)((char *)L.hist + off), w);
        if constexpr (EDGE) {
            const unsigned bin = off / (unsigned)HB;
            bool near = (s >= L.elo[bin] && s < L.ehi[bin]) ||
                        (s >= L.elo[bin + 1] && s < L.ehi[bin + 1]);
            edge_local += (near && bin >= (unsigned)p.skip) ? w : 0u;
        }
    }
}

// Pairs every j of one ring chunk with every home atom i (i unrolled by 4:
// one 64-byte scalar load per group).
// MODE 0: zero image shift (orthorhombic: dx = xj - xi); 1: j carries its
// image shift n (n + 128 as 3 x u8 in the meta word); 2: the rint wrap;
// 3: the reference's roundf(dx / L) wrap (pair_s_checked).
// HOME: the chunk holds home atoms (stream positions below n_home), which
// pair only with home atoms of lower index.
template <int PBC, bool WEIGHTED, bool EDGE, bool TABLE, int MODE, bool HOME, int HB = 8>
__device__ __forceinline__ void cell_chunk(const float4 *__restrict__ hp, int n_home, float xj,
                                           float yj, float zj, unsigned mj, int rank,
                                           const FrameBox &b, const HistParams &p,
                                           const CellLds &L, unsigned long long &edge_local) {
    float sx = 0.f, sy = 0.f, sz = 0.f;
    if constexpr (MODE == 1) {
        // image shift n + 128 per byte
        image_shift<PBC>((float)((int)(mj & 0xFFu) - 128), (float)((int)((mj >> 8) & 0xFFu) - 128),
                         (float)((int)((mj >> 16) & 0xFFu) - 128), b, sx, sy, sz);
    }
    const unsigned mAj = (mj >> 24) & 0xFu, mBj = mj >> 28;
    const float nanv = __builtin_nanf("");
    auto dist = [&](const float4 &pi) {
        if constexpr (MODE == 0) {  // no image shift in the chunk
            const float dx = xj - pi.x, dy = yj - pi.y, dz = zj - pi.z;
            return dx * dx + dy * dy + dz * dz;
        } else if constexpr (MODE == 1) {
            return pair_s_shift(pi.x, pi.y, pi.z, xj, yj, zj, sx, sy, sz);
        } else if constexpr (MODE == 2) {
            return pair_s<PBC, false>(pi.x, pi.y, pi.z, xj, yj, zj, b);
        } else {
            return pair_s_checked<PBC>(pi.x, pi.y, pi.z, xj, yj, zj, b);
        }
    };
    int i = 0;
    for (; i + 4 <= n_home; i += 4) {
        const float4 a0 = uload(hp, i), a1 = uload(hp, i + 1), a2 = uload(hp, i + 2),
                     a3 = uload(hp, i + 3);
        float s0 = dist(a0), s1 = dist(a1), s2 = dist(a2), s3 = dist(a3);
        if constexpr (HOME) {
            s0 = i >= rank ? nanv : s0;
            s1 = i + 1 >= rank ? nanv : s1;
            s2 = i + 2 >= rank ? nanv : s2;
            s3 = i + 3 >= rank ? nanv : s3;
        }
        cell_count<WEIGHTED, EDGE, TABLE, HB>(s0, __float_as_uint(a0.w), mAj, mBj, p, L, edge_local);
        cell_count<WEIGHTED, EDGE, TABLE, HB>(s1, __float_as_uint(a1.w), mAj, mBj, p, L, edge_local);
        cell_count<WEIGHTED, EDGE, TABLE, HB>(s2, __float_as_uint(a2.w), mAj, mBj, p, L, edge_local);
        cell_count<WEIGHTED, EDGE, TABLE, HB>(s3, __float_as_uint(a3.w), mAj, mBj, p, L, edge_local);
    }
    for (; i < n_home; ++i) {
        const float4 a0 = uload(hp, i);
        float s0 = dist(a0);
        if constexpr (HOME) s0 = i >= rank ? nanv : s0;
        cell_count<WEIGHTED, EDGE, TABLE, HB>(s0, __float_as_uint(a0.w), mAj, mBj, p, L, edge_local);
    }
}

// Inputs of the cell-list pair kernels: a batch of cell-sorted frames.
struct CellGrid {
    const float4 *sorted;       // [F][n] (x, y, z, meta) in cell order
    const unsigned *start;      // [F][ncell + 1]
    const unsigned *frame_flag; // [F] nonzero: the frame takes the rint path
    const FrameBox *boxes;      // [F]
    int n, n_frames, ncx, ncy, ncz, k, shift_ok, n_xcd;
    float cull2;                // squared cull radius (pair reach plus rounding margin)
    unsigned long long *queue;  // [n_xcd] zeroed grab counters (dynamic schedule), or null
    int grab;                   // units per grab of the dynamic schedule (0: kCellGrab)
    const float4 *bbox = nullptr;  // [F][ncell][2] per-cell bounding boxes from the build, or null
};

// A consumer declaring kPreShift = true moves the chunk's j atoms to their
// images next to the home cell itself (cons.preshift) and takes image-shift
// chunks (mode 1) in its mode-0 pair loop.
template <class C, class = void>
struct cons_preshift {
    static constexpr bool value = false;
};
template <class C>
struct cons_preshift<C, decltype((void)C::kPreShift)> {
    static constexpr bool value = C::kPreShift;
};

// Dispatches one chunk to the consumer's pair loop for its image mode.
template <class C>
__device__ __forceinline__ void cell_dispatch(C &cons, int mode, bool home, const float4 *hp,
                                              int n_home, float xj, float yj, float zj,
                                              unsigned mj, int rank, unsigned jpos, int f,
                                              const FrameBox &b) {
    if constexpr (cons_preshift<C>::value) {
        if (mode == 1) {
            cons.preshift(xj, yj, zj, mj, b);
            mode = 0;
        }
    }
    if (home) {
        if (mode == 0) cons.template chunk<0, true>(hp, n_home, xj, yj, zj, mj, rank, jpos, f, b);
        else if (mode == 1) cons.template chunk<1, true>(hp, n_home, xj, yj, zj, mj, rank, jpos, f, b);
        else if (mode == 2) cons.template chunk<2, true>(hp, n_home, xj, yj, zj, mj, rank, jpos, f, b);
        else cons.template chunk<3, true>(hp, n_home, xj, yj, zj, mj, rank, jpos, f, b);
    } else {
        if (mode == 0) cons.template chunk<0, false>(hp, n_home, xj, yj, zj, mj, rank, jpos, f, b);
        else if (mode == 1) cons.template chunk<1, false>(hp, n_home, xj, yj, zj, mj, rank, jpos, f, b);
        else if (mode == 2) cons.template chunk<2, false>(hp, n_home, xj, yj, zj, mj, rank, jpos, f, b);
        else cons.template chunk<3, false>(hp, n_home, xj, yj, zj, mj, rank, jpos, f, b);
    }
}

// Geometry of the half stencil's pieces for reach k: lanes 1..n_pieces each
// own one piece (odd lane: the in-range part of a row of cells along x, even
// lane: its part wrapped across the box): row offsets (dy, dz) and the x range
// [xa, xb] relative to the home cell.  Rows: k rows (dz = 0, dy > 0),
// k (2k + 1) rows with dz > 0, and the home row's cells hx + 1 .. hx + k.
struct StencilLane {
    int dy, dz, xa, xb;
};

__device__ __forceinline__ int stencil_pieces(int k) { return 2 * (2 * k * k + 2 * k + 1); }

__device__ __forceinline__ StencilLane stencil_lane(int lane, int k) {
    StencilLane s{0, 0, 0, -1};
    const int n_rows = 2 * k * k + 2 * k + 1;
    if (lane >= 1 && lane <= 2 * n_rows) {
        const int r = (lane - 1) >> 1;
        if (r < k) {
            s.dz = 0; s.dy = r + 1; s.xa = -k; s.xb = k;
        } else if (r < n_rows - 1) {
            const int q = r - k;
            s.dz = 1 + q / (2 * k + 1); s.dy = q % (2 * k + 1) - k; s.xa = -k; s.xb = k;
        } else {
            s.dz = 0; s.dy = 0; s.xa = 1; s.xb = k;
        }
    }
    return s;
}

// One unit (frame f, home cell h) of one wave.  It streams the half-stencil j
// list through the wave's ring (POS: also each j's position in the sorted
// frame) and hands every chunk of up to 64 j atoms to cons.chunk<MODE, HOME>;
// cons.unit_begin(f, hs, n_home) / cons.unit_end(f, hs, n_home) bracket a
// non-empty unit.  sl: this lane's stencil_lane(lane, G.k).
template <int PBC, bool POS, class C>
__device__ __forceinline__ void cell_unit(const CellGrid &G, int f, int h, const StencilLane &sl,
                                          float4 *ring, unsigned *ringp, C &cons) {
    const int lane = threadIdx.x & 63;
    const float nanv = __builtin_nanf("");
    const int ncx = G.ncx, ncy = G.ncy, ncz = G.ncz;
    const int ncell = ncx * ncy * ncz;
    const int n_pieces = stencil_pieces(G.k);
    const float cull2 = G.cull2;
    {
        const int hx = h % ncx;
        const int hy = (h / ncx) % ncy;
        const int hz = h / (ncx * ncy);
        const unsigned *fs = G.start + (long long)f * (ncell + 1);
        const float4 *fp = G.sorted + (long long)f * G.n;
        const int hs = (int)uload(fs, h);
        const int n_home = (int)uload(fs, h + 1) - hs;
        if (n_home == 0) return;
        const float4 *hp = fp + hs;
        const FrameBox b = uload(G.boxes, f);
        cons.unit_begin(f, hs, n_home);

        // home atoms: bounding box and image uniformity (the box is only used
        // on the shift path, whose frames hold finite coordinates), from the
        // build's per-cell boxes when it wrote them
        const unsigned m0 = __float_as_uint(uload(hp, 0).w) & 0xFFFFFFu;
        // far: the frame has atoms beyond kImageLimit box lengths (or non-finite
        // ones); its pairs take the checked wrap, exact for any quotient
        const bool far = uload(G.frame_flag, f) != 0u;
        float mn0, mn1, mn2, mx0, mx1, mx2;
        bool use_shift;
        if (G.bbox) {
            const float4 lo4 = uload(G.bbox, ((long long)f * ncell + h) * 2);
            const float4 hi4 = uload(G.bbox, ((long long)f * ncell + h) * 2 + 1);
            mn0 = lo4.x; mn1 = lo4.y; mn2 = lo4.z;
            mx0 = hi4.x; mx1 = hi4.y; mx2 = hi4.z;
            use_shift = G.shift_ok && !far && lo4.w != 0.0f;
        } else {
            // (only used when every coordinate of the frame is finite)
            const float inf = __builtin_inff();
            mn0 = inf; mn1 = inf; mn2 = inf;
            mx0 = -inf; mx1 = -inf; mx2 = -inf;
            bool diff = false;
            for (int hb = 0; hb < n_home; hb += 64) {
                if (hb + lane < n_home) {
                    const float4 q = hp[hb + lane];
                    mn0 = fminf(mn0, q.x); mx0 = fmaxf(mx0, q.x);
                    mn1 = fminf(mn1, q.y); mx1 = fmaxf(mx1, q.y);
                    mn2 = fminf(mn2, q.z); mx2 = fmaxf(mx2, q.z);
                    diff |= (__float_as_uint(q.w) & 0xFFFFFFu) != m0;
                }
            }
            use_shift = G.shift_ok && !far && !__any(diff);
            wave_bbox(mn0, mn1, mn2, mx0, mx1, mx2);
        }
        // the cull's bounds go to the two LDS entries past the wave's ring and
        // are read there in the stream loop (held in SGPRs across it, they are
        // spilled to VGPR lanes and read back with one v_readlane each)
        // The bounds are stored shifted by -128 lattice vectors (S), so that
        // the cull can take the biased image index n + 128 of each byte as it
        // is (its image positions are then off by the same -S; the rounding
        // this adds is covered by cull2_for's margin).
        float4 *const ubox = ring + kRing;
        if (lane == 0) {
            float sx, sy, sz;
            if constexpr (PBC == WMD_PBC_ORTHO) {
                sx = 128.f * b.lx; sy = 128.f * b.ly; sz = 128.f * b.lz;
            } else {
                sx = 128.f * (b.c[0] + b.c[3] + b.c[6]);
                sy = 128.f * (b.c[1] + b.c[4] + b.c[7]);
                sz = 128.f * (b.c[2] + b.c[5] + b.c[8]);
            }
            ubox[0] = make_float4(mn0 - sx, mn1 - sy, mn2 - sz, cull2);
            ubox[1] = make_float4(mx0 - sx, mx1 - sy, mx2 - sz, 0.f);
        }
        // per-piece byte constants K: n + 128 = (m_j + 32) + K per byte, with
        // K = 128 - w - (m_home + 32) (no carries: every byte stays below 256)
        const unsigned k_home = ((128u - (m0 & 0xFFu)) & 0xFFu) | (((128u - ((m0 >> 8) & 0xFFu)) & 0xFFu) << 8) |
                                (((128u - ((m0 >> 16) & 0xFFu)) & 0xFFu) << 16);

        // the stream's pieces: lane 0 the home cell, lanes 1..n_pieces the runs
        int pbeg = hs, plen = n_home;
        unsigned pk = k_home;
        if (lane >= 1 && lane <= n_pieces) {
            int cy = hy + sl.dy, cz = hz + sl.dz;
            const int wy = cy < 0 ? -1 : (cy >= ncy ? 1 : 0);
            const int wz = cz < 0 ? -1 : (cz >= ncz ? 1 : 0);
            cy -= wy * ncy;
            cz -= wz * ncz;
            const int xa = hx + sl.xa, xb = hx + sl.xb;
            // odd lane: [max(xa, 0), min(xb, ncx - 1)]; even lane: the part
            // below 0 (wrapped up, wx = -1) or above ncx - 1 (wrapped down,
            // wx = +1), else empty (grids have at least 2k + 1 cells per side,
            // so not both).  Selects, not branches: lo and hi + 1 stay within
            // [0, ncx] for every lane, empty runs included.
            const bool odd = (lane & 1) != 0, wl = xa < 0, wr = xb >= ncx;
            const int lo = odd ? max(xa, 0) : wl ? xa + ncx : wr ? 0 : 1;
            const int hi = odd ? min(xb, ncx - 1) : wl ? ncx - 1 : wr ? xb - ncx : 0;
            const int wx = odd ? 0 : wl ? -1 : wr ? 1 : 0;
            const int rowbase = (cz * ncy + cy) * ncx;
            pbeg = (int)fs[rowbase + lo];
            const int pend = (int)fs[rowbase + hi + 1];
            plen = lo <= hi ? pend - pbeg : 0;
            // K per byte: 128 - w - (m_home + 32); w = +1 subtracts 1, w = -1 adds 1
            pk = k_home + (unsigned)(-wx) + ((unsigned)(-wy) << 8) + ((unsigned)(-wz) << 16);
        }
        // stream order: home, runs without a box wrap, runs with one (so that
        // chunks of unshifted atoms come first), empty runs dropped (most of
        // the wrapped halves are empty); permuted through the empty ring
        int n_runs;  // non-empty runs after the home cell
        {
            const bool act = lane <= n_pieces && plen > 0;
            const bool wrapped = pk != k_home;
            const unsigned long long b0m = __ballot(act && !wrapped);
            const unsigned long long b1m = __ballot(act && wrapped);
            const int n0 = (int)__popcll(b0m);
            n_runs = n0 + (int)__popcll(b1m) - 1;
            const int dst = wrapped ? n0 + lanes_below(b1m) : lanes_below(b0m);
            if (act) ring[dst] = make_float4(__int_as_float(pbeg), __int_as_float(plen), __uint_as_float(pk), 0.f);
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            if (lane <= n_runs) {
                const float4 t = ring[lane];
                pbeg = __float_as_int(t.x);
                plen = __float_as_int(t.y);
                pk = __float_as_uint(t.z);
            }
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        }

        // Stream positions P0 .. P0 + 63 per step, across piece boundaries.
        // Uniform cursor: piece c covers stream positions [c_off, c_end).
        int head = 0, tail = 0;  // ring positions
        int c = 0, c_beg = hs, c_off = 0, c_end = n_home;
        unsigned c_k = k_home;
        int P0 = 0;
        while (true) {
            while (c_end <= P0 && c < n_runs) {
                ++c;
                c_beg = __builtin_amdgcn_readlane(pbeg, c);
                c_off = c_end;
                c_end += __builtin_amdgcn_readlane(plen, c);
                c_k = __builtin_amdgcn_readlane(pk, c);
            }
            const bool more = c_end > P0;
            if (more) {
                const int pos = P0 + lane;
                int src = c_beg + (pos - c_off);
                unsigned kk = c_k;
                int e = c, e_end = c_end;
                while (e_end < P0 + 64 && e < n_runs) {
                    ++e;
                    const int eb = __builtin_amdgcn_readlane(pbeg, e);
                    const int el = __builtin_amdgcn_readlane(plen, e);
                    const unsigned ek = __builtin_amdgcn_readlane(pk, e);
                    if (pos >= e_end) {
                        src = eb + (pos - e_end);
                        kk = ek;
                    }
                    e_end += el;
                }
                const bool valid = pos < e_end;
                const float4 q = fp[valid ? src : hs];  // lanes past the stream end load the home start
                const float4 lo = ubox[0];
                unsigned nb = 0x808080u;  // without the shift path: multiplicities only
                float d2 = 0.f;           // squared distance to the home atoms' box (shift path)
                if (use_shift) {
                    // n + 128 per byte; home atoms get n = 0 and lie inside the box
                    nb = (__float_as_uint(q.w) & 0xFFFFFFu) + kk;
                    const float bx = (float)(nb & 0xFFu);
                    const float by = (float)((nb >> 8) & 0xFFu);
                    const float bz = (float)((nb >> 16) & 0xFFu);
                    float px, py, pz;  // image of j next to the home atoms, less S (approximate)
                    if constexpr (PBC == WMD_PBC_ORTHO) {
                        px = __builtin_fmaf(-bx, b.lx, q.x);
                        py = __builtin_fmaf(-by, b.ly, q.y);
                        pz = __builtin_fmaf(-bz, b.lz, q.z);
                    } else {
                        px = __builtin_fmaf(-bx, b.c[0], __builtin_fmaf(-by, b.c[3], __builtin_fmaf(-bz, b.c[6], q.x)));
                        py = __builtin_fmaf(-bx, b.c[1], __builtin_fmaf(-by, b.c[4], __builtin_fmaf(-bz, b.c[7], q.y)));
                        pz = __builtin_fmaf(-bx, b.c[2], __builtin_fmaf(-by, b.c[5], __builtin_fmaf(-bz, b.c[8], q.z)));
                    }
                    const float4 hi = ubox[1];
                    const float ex = fmaxf(fmaxf(lo.x - px, px - hi.x), 0.f);
                    const float ey = fmaxf(fmaxf(lo.y - py, py - hi.y), 0.f);
                    const float ez = fmaxf(fmaxf(lo.z - pz, pz - hi.z), 0.f);
                    d2 = ex * ex + ey * ey + ez * ez;
                }
                const unsigned meta = (__float_as_uint(q.w) & 0xFF000000u) | nb;
                const bool keep = valid && d2 <= lo.w;
                const unsigned long long bal = __ballot(keep);
                if (keep) {
                    const int slot = (tail + lanes_below(bal)) & (kRing - 1);
                    ring[slot] = make_float4(q.x, q.y, q.z, __uint_as_float(meta));
                    if constexpr (POS) ringp[slot] = (unsigned)src;
                }
                tail += (int)__popcll(bal);
                P0 += 64;
            }
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            while (tail - head >= 64 || (!more && tail > head)) {
                const int cnt = min(64, tail - head);
                const int slot = (head + lane) & (kRing - 1);
                const float4 ej = ring[slot];
                const bool v = lane < cnt;
                const float xj = v ? ej.x : nanv;
                const float yj = v ? ej.y : nanv;
                const float zj = v ? ej.z : nanv;
                const unsigned mj = v ? __float_as_uint(ej.w) : 0x808080u;
                unsigned jpos = 0;
                if constexpr (POS) jpos = v ? ringp[slot] : 0xFFFFFFFFu;
                // 0: no lane has an image shift; 1: some lane has one; 2: rint wrap;
                // 3: checked wrap
                const int mode = far ? 3 : !use_shift ? 2 : (__any((mj & 0xFFFFFFu) != 0x808080u) ? 1 : 0);
                const bool home = head < n_home;
                const int rank = home && head + lane < n_home ? head + lane : 0x7FFFFFFF;
                cell_dispatch(cons, mode, home, hp, n_home, xj, yj, zj, mj, rank, jpos, f, b);
                head += cnt;
            }
            if (!more) break;
        }
        cons.unit_end(f, hs, n_home);
    }
}

// The unit loop of one wave over a batch of frames.  XCD-aware: workgroup g
// runs on XCD g % n_xcd and works on that XCD's contiguous slice of the
// units.  Static schedule (G.queue null): the slice's units are dealt to its
// waves round robin.  Dynamic schedule: each wave takes G.grab (default
// kCellGrab) consecutive units at a time from the slice's atomic counter until
// the counter passes the slice end (every wave reaches that exit), which evens
// out units of uneven cost.
constexpr int kCellGrab = 2;

template <int PBC, bool POS, class C>
__device__ __forceinline__ void cell_waves(const CellGrid &G, float4 *ring, unsigned *ringp,
                                           int wave, C &cons) {
    const StencilLane sl = stencil_lane(threadIdx.x & 63, G.k);
    const int ncell = G.ncx * G.ncy * G.ncz;
    const int n_xcd = G.n_xcd;
    const long long n_units = (long long)G.n_frames * ncell;
    const int xcd = blockIdx.x % n_xcd;
    const long long u_lo = n_units * xcd / n_xcd;
    const long long u_hi = n_units * (xcd + 1) / n_xcd;
    if (G.queue) {
        const int lane = threadIdx.x & 63;
        const int grab = G.grab > 0 ? G.grab : kCellGrab;
        while (true) {
            unsigned long long g0 = 0;
            if (lane == 0) g0 = atomicAdd(&G.queue[xcd], (unsigned long long)grab);
            const unsigned glo = __builtin_amdgcn_readfirstlane((unsigned)g0);
            const unsigned ghi = __builtin_amdgcn_readfirstlane((unsigned)(g0 >> 32));
            const long long u0 = u_lo + (long long)(((unsigned long long)ghi << 32) | glo);
            if (u0 >= u_hi) break;
            const long long u1 = min(u0 + (long long)grab, u_hi);
            for (long long u = u0; u < u1; ++u) {
                const int f = (int)(u / ncell);
                const int h = (int)(u - (long long)f * ncell);
                cell_unit<PBC, POS>(G, f, h, sl, ring, ringp, cons);
            }
        }
        return;
    }
    const int waves_per_wg = blockDim.x >> 6;
    const long long wave_local = (long long)(blockIdx.x / n_xcd) * waves_per_wg + wave;
    const long long waves_per_xcd = (long long)(gridDim.x / n_xcd) * waves_per_wg;
    for (long long u = u_lo + wave_local; u < u_hi; u += waves_per_xcd) {
        const int f = (int)(u / ncell);
        const int h = (int)(u - (long long)f * ncell);
        cell_unit<PBC, POS>(G, f, h, sl, ring, ringp, cons);
    }
}

// Plain histogram consumer: exact bins, LDS counts of HB bytes (u64, or
// u32 for a wave-private histogram: HB = 4).
// With EDGE (diagnostics) it also counts the pair evaluations: a valid j lane
// meets min(rank, n_home) home atoms in a home chunk and all n_home otherwise.
// With HB = 4 it bounds what any counter can have gained since the last drain
// (a chunk adds at most 2 x 64 x n_home to one bin) and drains the wave's
// histogram into `row` (u64 global counts of its bins from p.skip on) before
// a counter could wrap.
template <int PBC, bool WEIGHTED, bool EDGE, bool TABLE, int HB = 8>
struct PlainCons {
    static_assert(HB == 8 || (HB == 4 && !WEIGHTED), "u32 counters take unit weights only");
    HistParams p;
    CellLds L;
    unsigned long long edge_local;
    unsigned long long evals_local;
    unsigned long long *row = nullptr;
    unsigned pend = 0;
    template <int MODE, bool HOME>
    __device__ __forceinline__ void chunk(const float4 *hp, int n_home, float xj, float yj, float zj,
                                          unsigned mj, int rank, unsigned, int, const FrameBox &b) {
        cell_chunk<PBC, WEIGHTED, EDGE, TABLE, MODE, HOME, HB>(hp, n_home, xj, yj, zj, mj, rank, b, p,
                                                                L, edge_local);
        if constexpr (EDGE) {
            const int m = HOME ? min(rank, n_home) : n_home;
            evals_local += xj == xj ? (unsigned long long)m : 0ull;
        }
        if constexpr (HB == 4) {
            pend += 128u * (unsigned)n_home;
            if (pend >= 0x80000000u) drain();
        }
    }
    // Adds the wave's u32 histogram to row and clears it (every lane of the
    // wave, uniform control flow).
    __device__ __forceinline__ void drain() {
        unsigned *h = (unsigned *)L.hist;
        const int lane = threadIdx.x & 63;
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        for (int i = lane; i < p.bins; i += 64) {
            const unsigned v = h[i];
            if (v) {
                if (i >= p.skip) atomicAdd(row + (i - p.skip), (unsigned long long)v);
                h[i] = 0u;
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        pend = 0;
    }
    __device__ __forceinline__ void unit_begin(int, int, int) {}
    __device__ __forceinline__ void unit_end(int, int, int) {}
};

template <int PBC, bool WEIGHTED, bool EDGE, bool TABLE>
__global__ __launch_bounds__(kCellBlock) __attribute__((amdgpu_waves_per_eu(8, 8))) void rdf_cell_kernel(CellGrid G, HistParams p) {
    extern __shared__ __attribute__((aligned(16))) char lds[];
    const CellLds L = cell_lds_layout(lds, p, EDGE, TABLE);
    for (int i = threadIdx.x; i < p.bins; i += kCellBlock) L.hist[i] = 0ull;
    for (int i = threadIdx.x; i <= p.bins; i += kCellBlock) {
        L.thr[i] = p.thr[i];
        if constexpr (EDGE) {
            L.elo[i] = p.elo[i];
            L.ehi[i] = p.ehi[i];
        }
    }
    if constexpr (TABLE)
        for (int i = threadIdx.x; i < p.tab_n; i += kCellBlock) L.tab2[i] = p.tab2[i];
    __syncthreads();

    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);  // uniform: unit indices in SGPRs
    const int lane = threadIdx.x & 63;
    PlainCons<PBC, WEIGHTED, EDGE, TABLE> cons{p, L, 0ull, 0ull};
    cell_waves<PBC, false>(G, L.ring + wave * kRingStride, nullptr, wave, cons);

    __syncthreads();
    for (int i = threadIdx.x; i < p.bins; i += kCellBlock) {
        const unsigned long long v = L.hist[i];
        if (v && i >= p.skip) atomicAdd(&p.counts[i - p.skip], v);
    }
    if constexpr (EDGE) {
        unsigned long long e = cons.edge_local, v = cons.evals_local;
        for (int off = 32; off > 0; off >>= 1) {
            e += __shfl_xor(e, off, 64);
            v += __shfl_xor(v, off, 64);
        }
        if (lane == 0 && e) atomicAdd(p.edge, e);
        if (lane == 0 && v && p.count_evals) atomicAdd(p.edge + 1, v);
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

static inline uint32_t f_bits(float f) {
    uint32_t u;
    std::memcpy(&u, &f, 4);
    return u;
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

// Exact bin breakpoints of the reference's map s -> bin (rdf.rs:896, 929-935):
// thr[k] = min{ f32 s : bin(s) >= k }, cut = thr[bins]; and the s-bits bin
// table: the largest shift such that every bucket of 2^sh bit patterns in
// [thr[1], cut) holds at most one breakpoint thr[1..bins-1] (tab: the bin of
// the bucket's first value; tab2: the breakpoint inside the bucket, +inf if
// none, and 8 x the bin of the values below it).  Empty tables when no shift
// keeps the table within kMaxTab entries.
struct BinTables {
    float w = 0.f, cut = 0.f;
    std::vector<float> thr;
    std::vector<unsigned short> tab;
    std::vector<uint2> tab2;
    int tab_sh = 0, tab_base = 0;
};

inline BinTables make_bin_tables(int bins, float r_max) {
    BinTables T;
    T.w = r_max / (float)bins;
    const float w = T.w;
    T.thr.resize((size_t)bins + 1);
    T.thr[0] = 0.0f;
    for (int k = 1; k <= bins; ++k)
        T.thr[k] = min_s([&](float s) { return ref_key(s, r_max, w, bins) >= k; });
    T.cut = T.thr[bins];
    if (bins < 2) return T;
    const uint32_t lo_bits = f_bits(T.thr[1]);
    const uint32_t hi_bits = f_bits(T.cut) - 1;  // largest s below the cut
    for (int sh = 23; sh >= 8; --sh) {
        const int base = (int)(lo_bits >> sh);
        const int n = (int)(hi_bits >> sh) - base + 1;
        if (n > kMaxTab) break;
        bool ok = true;
        for (int k = 1; k + 1 < bins && ok; ++k)
            if ((f_bits(T.thr[k]) >> sh) == (f_bits(T.thr[k + 1]) >> sh)) ok = false;
        if (!ok) continue;
        std::vector<unsigned short> tab(n);
        int b = 0;
        for (int k = 0; k < n; ++k) {
            const uint32_t start_bits = (uint32_t)(base + k) << sh;
            const float s0 = k == 0 ? 0.0f : bits_f(start_bits);
            while (b + 1 < bins && s0 >= T.thr[b + 1]) ++b;
            tab[k] = (unsigned short)b;
        }
        std::vector<uint2> tab2(n);
        for (int k = 0; k < n; ++k) {
            const int bk = tab[k];
            uint32_t split = 0x7F800000u;
            if (bk + 1 <= bins && (f_bits(T.thr[bk + 1]) >> sh) == (uint32_t)(base + k))
                split = f_bits(T.thr[bk + 1]);
            tab2[k] = make_uint2(split, (unsigned)bk * 8u);  // byte offset of the bin
        }
        T.tab.swap(tab);
        T.tab2.swap(tab2);
        T.tab_sh = sh;
        T.tab_base = base << sh;  // bit pattern where bucket 0 starts
        return T;
    }
    return T;
}

// Scratch of the cell build.
struct CellScratch {
    DevBuf cell_of, cell_slot, image, count, start, flag, sorted, bbox;
    DevBuf order;  // [F][n] store slot -> sorted position (want_order)
    size_t frames_for(int n, int ncell, size_t n_frames, bool want_order, bool want_bbox = false) const {
        const size_t per_frame = (size_t)n * (16 + 12 + (want_order ? 4 : 0)) + (size_t)(2 * ncell + 2) * 4 +
                                 (want_bbox ? (size_t)ncell * 32 : 0);
        return std::max<size_t>(1, std::min<size_t>(n_frames, (size_t)(1536ull << 20) / per_frame));
    }
    void reserve(int n, int ncell, size_t batch, bool want_order, bool want_bbox = false) {
        cell_of.reserve((size_t)n * batch * 4);
        cell_slot.reserve((size_t)n * batch * 4);
        image.reserve((size_t)n * batch * 4);
        count.reserve((size_t)ncell * batch * 4);
        start.reserve((size_t)(ncell + 1) * batch * 4);
        flag.reserve(batch * 4);
        sorted.reserve((size_t)n * batch * 16);
        if (want_bbox) bbox.reserve((size_t)ncell * batch * 32);
        if (want_order) order.reserve((size_t)n * batch * 4);
    }
};

void grid1d(long long total, int *grid) {
    long long g = (total + kBlock - 1) / kBlock;
    if (g > 65536) g = 65536;
    if (g < 1) g = 1;
    *grid = (int)g;
}

// Per-frame box parameters of a batch (throws on an invalid box).
struct BoxSet {
    std::vector<FrameBox> fb;
    double min_width[3] = {1e300, 1e300, 1e300};
    bool lower_tri = true;
};

BoxSet make_boxes(int pbc, const float *boxes_host, size_t n_frames) {
    BoxSet B;
    B.fb.resize(n_frames);
    for (size_t f = 0; f < n_frames; ++f) {
        BoxInfo bi;
        if (pbc != WMD_PBC_NONE) {
            const float *m = boxes_host + 9 * f;
            if (!make_box(pbc, m, &bi))
                fail(WMD_ERR_MISMATCH, pbc == WMD_PBC_ORTHO ? "RDF requires a positive orthorhombic box"
                                                            : "box matrix not invertible");
            for (int d = 0; d < 3; ++d) B.min_width[d] = std::min(B.min_width[d], bi.width[d]);
            if (m[1] != 0.0f || m[2] != 0.0f || m[5] != 0.0f) B.lower_tri = false;
        } else {
            std::memset(&bi, 0, sizeof(bi));
        }
        FrameBox &b = B.fb[f];
        b.lx = bi.lx; b.ly = bi.ly; b.lz = bi.lz;
        b.ilx = bi.lx > 0 ? 1.0f / bi.lx : 0.f;
        b.ily = bi.ly > 0 ? 1.0f / bi.ly : 0.f;
        b.ilz = bi.lz > 0 ? 1.0f / bi.lz : 0.f;
        if (pbc == WMD_PBC_ORTHO) {
            for (int i = 0; i < 9; ++i) b.r[i] = 0.f, b.c[i] = 0.f;
            b.r[0] = b.ilx; b.r[4] = b.ily; b.r[8] = b.ilz;
            b.c[0] = b.lx; b.c[4] = b.ly; b.c[8] = b.lz;
        } else {
            for (int i = 0; i < 9; ++i) b.r[i] = bi.r[i], b.c[i] = bi.c[i];
        }
    }
    return B;
}

// Copies the batch's boxes to the device through a pinned staging buffer.
// guard: NULL, or an event recorded after the previous copy out of hbuf; the
// host then waits for that copy only, not for the whole stream.
FrameBox *upload_boxes(const BoxSet &B, HostBuf &hbuf, DevBuf &dbuf, hipStream_t st,
                       Event *guard = nullptr) {
    const size_t bytes = sizeof(FrameBox) * std::max<size_t>(B.fb.size(), 1);
    if (guard) guard->wait();
    else WMD_HIP_CHECK(hipStreamSynchronize(st));  // the staging buffer may still be in use
    if (guard && (bytes > hbuf.bytes || bytes > dbuf.bytes))
        WMD_HIP_CHECK(hipStreamSynchronize(st));  // growing frees buffers queued work may use
    void *hb = hbuf.reserve(bytes);
    std::memcpy(hb, B.fb.data(), sizeof(FrameBox) * B.fb.size());
    FrameBox *d = (FrameBox *)dbuf.reserve(bytes);
    WMD_HIP_CHECK(hipMemcpyAsync(d, hb, sizeof(FrameBox) * B.fb.size(), hipMemcpyHostToDevice, st));
    if (guard) WMD_HIP_CHECK(hipEventRecord(guard->get(), st));
    return d;
}

// Cell grid for pair reach `reach`: cells at least (reach + 1e-2) / k wide per
// plane spacing, >= 2k + 1 per side, several atoms per stencil.  Returns k
// (0 = use brute force); forced_k in 1..3 restricts the choice.
int choose_cells(double reach, const double *min_width, int n, int forced_k, int *nc) {
    int ks[3] = {2, 1, 3};
    if (forced_k >= 1 && forced_k <= 3) { ks[0] = forced_k; ks[1] = 0; ks[2] = 0; }
    for (int kk : ks) {
        if (kk == 0) continue;
        const double cw = (reach + 1e-2) / kk;
        int c[3];
        bool ok = true;
        for (int d = 0; d < 3; ++d) {
            c[d] = (int)std::min(std::floor(min_width[d] / cw), 96.0);
            if (c[d] < 2 * kk + 1) ok = false;
        }
        if (!ok) continue;
        if ((double)n / ((double)c[0] * c[1] * c[2]) * std::pow(2 * kk + 1, 3) < 48.0) return 0;
        nc[0] = c[0]; nc[1] = c[1]; nc[2] = c[2];
        return kk;
    }
    return 0;
}

// Squared cull radius: reach plus a margin covering the f32 rounding of image
// positions and pair differences for coordinates within (kImageLimit + 1) box
// lengths (frames beyond that take the rint path, which does not cull).
float cull2_for(double reach, const BoxSet &B) {
    double coord_bound = 0.0;
    for (const FrameBox &b : B.fb) {
        double row = 0.0;
        for (int i = 0; i < 9; ++i) row += std::fabs((double)b.c[i]);
        coord_bound = std::max(coord_bound, row);
    }
    // the cull's biased image positions q - (n + 128) . box (n + 128 in
    // [111, 145] for frames within kImageLimit box lengths, |q| below 9 box
    // lengths) and its bounds shifted by -128 lattice vectors carry at most
    // five roundings of magnitude below 160 x row per axis
    const double biased = std::sqrt(3.0) * 5.0 * 160.0 * coord_bound * std::ldexp(1.0, -24);
    coord_bound *= (double)(kImageLimit + 1);
    const double margin = 0.05 + 32.0 * coord_bound * std::ldexp(1.0, -24) + biased;
    const float cull = (float)(reach + margin);
    return cull * cull;
}

// Sorts frames [f0, f0 + nf) of an SoA store into cells (bx: their boxes).
// With want_bbox the one-workgroup-per-frame build also writes per-cell
// bounding boxes (S.bbox), returned; null otherwise and on the three-kernel
// path for grids beyond kBuildMaxCells.  The extra pass re-reads the sorted
// frame, which costs the plain histogram more than its units' own box
// reductions; the soft-RDF kernels, whose forward and backward share one
// build and whose units are heavier, take the boxes.
const float4 *build_cells(wmd_ctx *ctx, hipStream_t st, int pbc, const float *d_soa, int stride, int n,
                          int nf, long long f0, const FrameBox *bx, const int *nc, const unsigned *d_w,
                          CellScratch &S, bool want_order, bool want_bbox = false) {
    ProfScope build_prof(ctx, st, "cell_build");
    const int ncell = nc[0] * nc[1] * nc[2];
    const long long fs = 3 * (long long)stride;
    if (ncell <= kBuildMaxCells) {
        const size_t lds = build_lds_bytes(ncell);
        unsigned *ord = want_order ? S.order.as<unsigned>() : (unsigned *)nullptr;
        float4 *bb = want_bbox ? S.bbox.as<float4>() : (float4 *)nullptr;
        if (pbc == WMD_PBC_ORTHO)
            hipLaunchKernelGGL(cell_build_frame_kernel<WMD_PBC_ORTHO>, dim3(nf), dim3(kBuildBlock), lds, st, d_soa, fs, stride, n, f0, bx, nc[0], nc[1], nc[2], d_w, S.start.as<unsigned>(), S.sorted.as<float4>(), ord, S.flag.as<unsigned>(), bb);
        else
            hipLaunchKernelGGL(cell_build_frame_kernel<WMD_PBC_TRICLINIC>, dim3(nf), dim3(kBuildBlock), lds, st, d_soa, fs, stride, n, f0, bx, nc[0], nc[1], nc[2], d_w, S.start.as<unsigned>(), S.sorted.as<float4>(), ord, S.flag.as<unsigned>(), bb);
        WMD_HIP_CHECK(hipGetLastError());
        return bb;
    }
    unsigned *d_count = S.count.as<unsigned>(), *d_fflag = S.flag.as<unsigned>();
    int g;
    grid1d((long long)ncell * nf, &g);
    hipLaunchKernelGGL(fill_u32_kernel, dim3(g), dim3(kBlock), 0, st, d_count, (long long)ncell * nf, 0u);
    hipLaunchKernelGGL(fill_u32_kernel, dim3(1), dim3(kBlock), 0, st, d_fflag, (long long)nf, 0u);
    WMD_HIP_CHECK(hipGetLastError());
    grid1d((long long)n * nf, &g);
    if (pbc == WMD_PBC_ORTHO)
        hipLaunchKernelGGL(cell_assign_kernel<WMD_PBC_ORTHO>, dim3(g), dim3(kBlock), 0, st, d_soa, fs, stride, n, nf, f0, bx, nc[0], nc[1], nc[2], S.cell_of.as<unsigned>(), S.cell_slot.as<unsigned>(), d_count, S.image.as<unsigned>(), d_fflag);
    else
        hipLaunchKernelGGL(cell_assign_kernel<WMD_PBC_TRICLINIC>, dim3(g), dim3(kBlock), 0, st, d_soa, fs, stride, n, nf, f0, bx, nc[0], nc[1], nc[2], S.cell_of.as<unsigned>(), S.cell_slot.as<unsigned>(), d_count, S.image.as<unsigned>(), d_fflag);
    WMD_HIP_CHECK(hipGetLastError());
    hipLaunchKernelGGL(cell_scan_kernel, dim3(nf), dim3(kBlock), 0, st, d_count, S.start.as<unsigned>(), ncell);
    WMD_HIP_CHECK(hipGetLastError());
    hipLaunchKernelGGL(cell_scatter_kernel, dim3(g), dim3(kBlock), 0, st, d_soa, fs, stride, n, nf, f0,
                       (const unsigned *)S.cell_of.ptr, (const unsigned *)S.cell_slot.ptr,
                       (const unsigned *)S.image.ptr, (const unsigned *)S.start.ptr, ncell, d_w,
                       S.sorted.as<float4>(), want_order ? S.order.as<unsigned>() : (unsigned *)nullptr);
    WMD_HIP_CHECK(hipGetLastError());
    return nullptr;
}

}  // namespace
