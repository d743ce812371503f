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
// with IEEE sqrtf / division.  A pair is counted iff s < T[bins].  Its bin is
// read from a table indexed by the high bits of s (built so that every bucket
// holds at most one breakpoint) and corrected by one comparison with T; when
// the bins are too fine for such a table, a v_sqrt estimate corrected on both
// sides is used instead.  Counts are therefore bit-identical to the reference
// with neither the IEEE divide nor the IEEE square root in the pair loop.
//
// Minimum image.  The reference wraps each component with round(dx / L)
// (roundf).  Three exact ways are used:
//  * shift: the cell path runs only when r_max is below half the smallest
//    plane spacing (less 1e-5 relative) on a periodic grid of >= 2k+1 cells per
//    side, cells >= (r_max + 1e-2 A) / k wide.  A pair the reference counts has
//    its minimum image within k cells of the home cell, where the stencil
//    visits it exactly once (offsets -k..k are distinct cells mod nc), with
//    wrap w such that round(dx / L) = m_j - w - m_i (m: the atom's image index
//    floor(x / L)).  A visited pair whose cell image is not round(dx / L) lies
//    at least L/2 > r_max away in that image and is counted by neither.  The
//    per-j value n*L (the reference's own f32 product round(dx/L)*L) is staged
//    in LDS and the pair costs dx = (xj - xi) - S: the reference's two f32
//    roundings, nothing else.
//    Triclinic boxes (refused by the reference RdfPlan, rdf.rs:797-802) take
//    n = round(f) of the reference's f32 fractional displacement f = inv . d
//    (wrap_triclinic, part1.cu:12-27) and subtract the lattice vector in
//    Cartesian space, S = (n0 a + n1 b) + n2 c per component in f32; the shift
//    path stages that S per j, so a pair costs what an orthorhombic one does.
//    For a pair in range, |f - n| < 1/2 on every axis (r_max is below half the
//    smallest plane spacing), so the cells' image is round(f); any other image
//    lies at least half a plane spacing away and is counted by neither.
//  * rint: when r_max < half the smallest plane spacing (less 1e-5 relative),
//    rint(dx * (1/L)) equals roundf(dx / L) for every pair in range; a pair
//    where they could differ lies near L/2 > r_max and is counted by neither.
//  * exact: IEEE division and roundf, for small boxes.
//
// Enumeration.  Counts are integer sums, so any enumeration that visits every
// ordered pair's multiset once reproduces the reference loop.  Two kernels:
//  * brute force over 256x256 atom tiles (upper triangle of tile pairs; diagonal
//    tiles as ordered pairs, off-diagonal as unordered pairs of weight 2);
//  * a periodic cell list, cells >= (r_max + 1e-2 A) / k wide along each plane
//    spacing, home cell x half stencil of (2k+1)^3 neighbour cells (unordered
//    pairs, weight 2) plus the home cell itself (ordered pairs, weight 1).
//    Neighbour atoms farther than r_max + 0.05 A from the home atoms' bounding
//    box are dropped while staging (their pairs are provably out of range).
// Both are persistent.  The brute-force kernel keeps u32 sub-histograms in LDS
// and merges them into the u64 global counts when a uniform bound says a slot
// could wrap, and at the end; the cell kernel keeps one u64 LDS histogram per
// workgroup.  Units are split into contiguous ranges per XCD so that an XCD's
// L2 holds the frames it works on; inside its range, each wave of the cell
// kernel takes units from an atomic counter (units differ in cost).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <unordered_map>

#include "wmd_internal.hpp"

#include "wmd_rdf_core.hpp"

using namespace wmd;

namespace {

// -------------------------------------------------------------------------
// Gaussian-smoothed RDF (no counterpart in the reference; the differentiable
// companion of the histogram above, over the same unordered-pair set and the
// same minimum-image rule), in particle-mesh form:
//
//   H_m = sum over ordered pairs (i, j), i != j, of B(r_ij / h - m)
//   c_k = sum over nodes m with |m h - r_k| <= 6 sigma of H_m w phi(r_k - m h),
//
// where B is the quadratic B-spline (support 3 nodes, a partition of unity),
// h = w / 4 the node spacing, r_k = (k + 1/2) w, w = r_max / bins and phi the
// N(0, sigma^2) density.  c is a fixed linear map of H, so it is C1 in the
// positions; the effective kernel is the Gaussian convolved with B
// (variance sigma^2 + h^2 / 4).  Each pair deposits onto 3 nodes (f64 LDS
// atomics); a small kernel applies the map.  The backward pass tabulates
// D_m = sum_k dL/dc_k w phi(r_k - m h) on the nodes, so that one unordered
// pair's dL/dr = 2 sum_m D_m B'(r / h - m) / h, and enumerates the pairs again
// rather than storing them.  B' is linear on each interval |u - n0| <= 1/2, so
// dL/dr there is A_n0 + t S_n0 (t = u - n0) with
//   A_n = (D_{n+1} - D_{n-1}) / h,  S_n = 2 (D_{n-1} - 2 D_n + D_{n+1}) / h,
// one (A, S) table entry per node.
// -------------------------------------------------------------------------
constexpr int kSoftNodesPerBin = 4;   // h = w / 4
constexpr int kSoftPadL = 1;          // node padding below node 0
constexpr int kSoftPadR = 3;          // and above the last node

struct SoftParams {
    int n_nodes;        // nodes 0 .. n_nodes - 1 carry weight
    float inv_h;        // 1 / node spacing
    float cut2;         // squared pair reach
    double *nodes;      // [n_nodes] H (forward output, accumulated)
    const float2 *table; // padded (A, S) per node (backward)
    float *grad;        // brute force: [F][3][stride] output; cell list: [F][3][n] in cell order
    long long grad_frame0;  // frame offset of this batch in grad (brute force)
    int stride;
    int n;              // atoms per frame
};

// Quadratic B-spline weights of position u on nodes n0 - 1, n0, n0 + 1
// (n0 = rint(u), t = u - n0 in [-1/2, 1/2]) and their u-derivatives.
__device__ __forceinline__ int bspline2(float u, int n_max, float &t) {
    const int n0 = min((int)rintf(u), n_max);
    t = u - (float)n0;
    return n0;
}

// Deposits one unordered pair (both orders) at squared distance s onto the
// padded node array (nodes[m + kSoftPadL] holds node m) as unsigned fixed
// point, 2^30 per unit: the three weights (at most 1.5) are formed in f32 and
// converted once, so integer LDS atomics add them exactly and the node sums
// of a workgroup do not depend on the order of its deposits.  A node sum
// overflows only past 1.1e10 deposits in one workgroup.
constexpr double kSoftUnfix = 1.0 / 1073741824.0;
__device__ __forceinline__ void soft_deposit(float s, const SoftParams &sp, unsigned long long *nodes) {
    const float u = __builtin_amdgcn_sqrtf(s) * sp.inv_h;
    float t;
    const int n0 = bspline2(u, sp.n_nodes, t);
    // The scale 2^30 is folded into the operands: scaling by a power of two
    // commutes with rounding, so each product is bit-identical to the f32
    // weight times 2^30 (2^15 (1/2 -+ t) is one rounding of 1/2 -+ t, scaled).
    const float a = __builtin_fmaf(-t, 32768.0f, 16384.0f);                 // 2^15 (1/2 - t)
    const float c = __builtin_fmaf(t, 32768.0f, 16384.0f);                  // 2^15 (1/2 + t)
    const float b = __builtin_fmaf(t * t, -2147483648.0f, 1610612736.0f);   // 2^30 (3/2 - 2 t^2)
    unsigned long long *p = nodes + kSoftPadL + n0;
    atomicAdd(p - 1, (unsigned long long)(unsigned)(a * a));  // 2 x (1/2 - t)^2 / 2
    atomicAdd(p, (unsigned long long)(unsigned)b);            // 2 x (3/4 - t^2)
    atomicAdd(p + 1, (unsigned long long)(unsigned)(c * c));  // 2 x (1/2 + t)^2 / 2
}

// One unordered pair's dL/dr divided by r, from the padded (A, S) table.
// Pairs within reach have u <= n_nodes, so t = u - rint(u) needs no clamp;
// the index is clamped for the lanes a caller masks out.
__device__ __forceinline__ float soft_force_over_r(float s, const SoftParams &sp, const float2 *tab) {
    const float rinv = __builtin_amdgcn_rsqf(s);
    const float u = s * rinv * sp.inv_h;
    const float n0f = rintf(u);
    const float t = u - n0f;
    const int n0 = min((int)n0f, sp.n_nodes);
    const float2 as = tab[kSoftPadL + n0];
    return __builtin_fmaf(t, as.y, as.x) * rinv;
}

// Minimum-image displacement x_j - x_i of the cell stream's pair, returned
// with its squared length.  MODE 0: the j atoms sit next to the home cell
// (image shifts of mode-1 chunks already applied by soft_preshift); 2 and 3:
// the rint wrap (the smoothed histogram is this engine's own quantity, so the
// checked wrap of mode 3 is not needed).
template <int PBC, bool LT, int MODE>
__device__ __forceinline__ float soft_disp(const float4 &pi, float xj, float yj, float zj,
                                           const FrameBox &b, float &dx, float &dy, float &dz) {
    dx = xj - pi.x;
    dy = yj - pi.y;
    dz = zj - pi.z;
    if constexpr (MODE >= 2) {
        if constexpr (PBC == WMD_PBC_ORTHO) {
            dx -= rintf(dx * b.ilx) * b.lx;
            dy -= rintf(dy * b.ily) * b.ly;
            dz -= rintf(dz * b.ilz) * b.lz;
        } else {
            float f0 = b.r[0] * dx + b.r[3] * dy + b.r[6] * dz;
            float f1 = b.r[1] * dx + b.r[4] * dy + b.r[7] * dz;
            float f2 = b.r[2] * dx + b.r[5] * dy + b.r[8] * dz;
            f0 -= rintf(f0); f1 -= rintf(f1); f2 -= rintf(f2);
            dx = f0 * b.c[0] + f1 * b.c[3] + f2 * b.c[6];
            dy = f0 * b.c[1] + f1 * b.c[4] + f2 * b.c[7];
            dz = f0 * b.c[2] + f1 * b.c[5] + f2 * b.c[8];
        }
    }
    return __builtin_fmaf(dz, dz, __builtin_fmaf(dy, dy, dx * dx));
}

// Moves a stream atom to its image next to the home cell: x - n . box, with
// the image shift n + 128 per byte of the meta word (cell_unit).
template <int PBC>
__device__ __forceinline__ void soft_preshift(float &xj, float &yj, float &zj, unsigned mj,
                                              const FrameBox &b) {
    const float n0 = (float)((int)(mj & 0xFFu) - 128);
    const float n1 = (float)((int)((mj >> 8) & 0xFFu) - 128);
    const float n2 = (float)((int)((mj >> 16) & 0xFFu) - 128);
    if constexpr (PBC == WMD_PBC_ORTHO) {
        xj -= n0 * b.lx;
        yj -= n1 * b.ly;
        zj -= n2 * b.lz;
    } else {
        xj -= n0 * b.c[0] + n1 * b.c[3] + n2 * b.c[6];
        yj -= n0 * b.c[1] + n1 * b.c[4] + n2 * b.c[7];
        zj -= n0 * b.c[2] + n1 * b.c[5] + n2 * b.c[8];
    }
}

// Forward consumer: in-range pairs are compacted into a per-wave buffer of
// squared distances and spread onto the histogram 64 at a time.
template <int PBC, bool LT>
struct SoftFwdCons {
    SoftParams sp;
    unsigned long long *hist;   // LDS, padded, fixed point
    float *pbuf;    // this wave's 128-entry pair buffer, filled from its start
    int pt;         // entries in the buffer (uniform, below 64 between pairs)
    // Deposits the first cnt entries (cnt = 64, or the rest at the end).
    __device__ __forceinline__ void drain(int cnt) {
        const int lane = threadIdx.x & 63;
        if (lane < cnt) soft_deposit(pbuf[lane], sp, hist);
    }
    static constexpr bool kPreShift = true;
    __device__ __forceinline__ void preshift(float &xj, float &yj, float &zj, unsigned mj,
                                             const FrameBox &b) {
        soft_preshift<PBC>(xj, yj, zj, mj, b);
    }
    template <int MODE, bool HOME>
    __device__ __forceinline__ void chunk(const float4 *hp, int n_home, float xj, float yj, float zj,
                                          unsigned, int rank, unsigned, int, const FrameBox &b) {
        auto pair = [&](int i, const float4 &a) {
            float dx, dy, dz;
            const float s = soft_disp<PBC, LT, MODE>(a, xj, yj, zj, b, dx, dy, dz);
            const bool in = s < sp.cut2 && (!HOME || i < rank);
            const unsigned long long bal = __ballot(in);
            // slot pt + (in-range lanes below): pt stays below 64, so the
            // slot stays below 128 without a wrap (pt goes into the uniform
            // base address, not into the lane count)
            float *const dst = pbuf + pt;
            if (in) dst[__builtin_amdgcn_mbcnt_hi((unsigned)(bal >> 32),
                                                  __builtin_amdgcn_mbcnt_lo((unsigned)bal, 0u))] = s;
            pt += (int)__popcll(bal);
            if (pt >= 64) {
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                __builtin_amdgcn_wave_barrier();
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
                // both halves read together (one LDS wait), the first 64
                // deposited, the entries past them moved to the front
                const int lane = threadIdx.x & 63;
                const float first = pbuf[lane];
                const float rest = pbuf[64 + lane];
                soft_deposit(first, sp, hist);
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                __builtin_amdgcn_wave_barrier();
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
                pt -= 64;
                if (lane < pt) pbuf[lane] = rest;
            }
        };
        // home atoms four at a time (scalar loads)
        int i = 0;
        for (; i + 4 <= n_home; i += 4) {
            const float4 a0 = uload(hp, i), a1 = uload(hp, i + 1), a2 = uload(hp, i + 2),
                         a3 = uload(hp, i + 3);
            pair(i, a0);
            pair(i + 1, a1);
            pair(i + 2, a2);
            pair(i + 3, a3);
        }
        for (; i < n_home; ++i) pair(i, uload(hp, i));
    }
    __device__ __forceinline__ void unit_begin(int, int, int) {}
    __device__ __forceinline__ void unit_end(int, int, int) {}
};

// Backward consumer.  Lanes hold the chunk's j atoms and the home atoms i
// come four at a time (uniform, scalar loads); every lane evaluates its pair
// with each of them without compaction (masked lanes give a zero force).  A
// pair's gradient g = (dL/dr / r) (x_j - x_i) goes to j (+g) and to i (-g):
//  * j's share stays in the lane's registers for the whole chunk and is
//    added to the output once;
//  * i's share is summed over lanes: 5 home atoms' 3 components (plus a
//    zero pad, 16 values) are reduced over each 16-lane row by recursive
//    halving (4 DPP steps; afterwards lane l holds the row sum of value
//    v(l) = 8 b3 + 4 b2 + 2 b0 + b1 of its lane bits), and each lane adds its
//    value to its own running sum (one register per 5 home atoms for the
//    first 35, a per-row LDS slot of the home atom beyond: 64 distinct
//    addresses per instruction); the 4 row copies are summed once at the end
//    of the unit.
// Home atoms past the consumer's MAXH LDS slots add their row sums to the
// output directly (MAXH = 32 for the usual cells of about 20 atoms, 128 for
// dense cells, whose LDS footprint is larger).
// The output is indexed by cell-sorted position, where a chunk's lanes mostly
// hold consecutive atoms, so the atomics of one instruction fall on a few
// cache lines.  (Keeping each lane's home-atom sums in registers across the
// unit instead, reduced once per unit, needs 3 VGPRs per home atom; at 24
// home atoms the kernel spilled and ran slower at three waves per SIMD.)
// floats per row copy of the home sums (padded: rows on distinct banks)
__host__ __device__ constexpr int soft_gi_row(int maxh) { return maxh * 4 + 16; }

template <int CTRL>
__device__ __forceinline__ float dpp_f(float v) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, 0xF, 0xF, false));
}

// The value of lane l ^ (1 << B) within a 16-lane row.
template <int B>
__device__ __forceinline__ float row_xor(float v) {
    if constexpr (B == 0) return dpp_f<0xB1>(v);                 // quad_perm [1,0,3,2]
    else if constexpr (B == 1) return dpp_f<0x4E>(v);            // quad_perm [2,3,0,1]
    else if constexpr (B == 2) return dpp_f<0x141>(dpp_f<0x1B>(v));   // half-row mirror of quad_perm [3,2,1,0]
    else return dpp_f<0x141>(dpp_f<0x140>(v));                   // half-row mirror of the row mirror
}

// One recursive-halving step over lane bit B: M values -> M / 2, a lane with
// the bit set keeping the upper half.
template <int M, int B>
__device__ __forceinline__ void row_halve(float (&v)[16], bool hi) {
#pragma unroll
    for (int k = 0; k < M / 2; ++k) {
        const float keep = hi ? v[k + M / 2] : v[k];
        const float send = hi ? v[k] : v[k + M / 2];
        v[k] = keep + row_xor<B>(send);
    }
}

// Halving steps over lane bits 3 and 2, where the lanes that keep the lower
// and the upper half fill whole 4-lane banks: one DPP add per half, each
// writing only its own banks (bank_mask), instead of two selects and an add.
// d = lo + lo of lane l ^ 8 in lanes 0-7 of a row, hi + hi of lane l ^ 8 in
// lanes 8-15 (row_ror:8 reads lane (l - 8) mod 16).  The s_nop covers the
// VALU-write-to-DPP-read wait of the inputs.
__device__ __forceinline__ float bank_halve_b3(float lo, float hi) {
    float d;
    asm volatile("s_nop 1\n\t"
                 "v_add_f32_dpp %0, %1, %1 row_ror:8 row_mask:0xf bank_mask:0x3\n\t"
                 "v_add_f32_dpp %0, %2, %2 row_ror:8 row_mask:0xf bank_mask:0xc"
                 : "=&v"(d)
                 : "v"(lo), "v"(hi));
    return d;
}
// d = lo + lo of lane l + 4 in banks 0 and 2 (row_ror:12), hi + hi of lane
// l - 4 in banks 1 and 3 (row_ror:4).
__device__ __forceinline__ float bank_halve_b2(float lo, float hi) {
    float d;
    asm volatile("s_nop 1\n\t"
                 "v_add_f32_dpp %0, %1, %1 row_ror:12 row_mask:0xf bank_mask:0x5\n\t"
                 "v_add_f32_dpp %0, %2, %2 row_ror:4 row_mask:0xf bank_mask:0xa"
                 : "=&v"(d)
                 : "v"(lo), "v"(hi));
    return d;
}

// Row sums of 16 values per lane: afterwards v[0] of lane l holds the sum
// over its 16-lane row of value v(l) = 8 b3 + 4 b2 + 2 b0 + b1 of its lane
// bits (bits 3 and 2 by bank-masked DPP adds, then bits 0 and 1).
__device__ __forceinline__ void row_reduce16(float (&g)[16]) {
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int k = 0; k < 8; ++k) g[k] = bank_halve_b3(g[k], g[k + 8]);
#pragma unroll
    for (int k = 0; k < 4; ++k) g[k] = bank_halve_b2(g[k], g[k + 4]);
    row_halve<4, 0>(g, lane & 1);
    row_halve<2, 1>(g, lane & 2);
}

template <int PBC, bool LT, int MAXH>
struct SoftBwdCons {
    static constexpr int kGiRow = soft_gi_row(MAXH);
    static constexpr bool kPreShift = true;
    static constexpr int kG = 5;          // home atoms per row reduction (15 values + a pad)
    static constexpr int kRegGroups = 7;  // home atoms 0..34: row sums kept in registers
    // home atoms with LDS row slots (the register groups' sums land there at
    // the end of the unit); soft_gi_row(32) = 144 floats hold 35 atoms
    static constexpr int kSlots = MAXH > kG * kRegGroups ? MAXH : kG * kRegGroups;
    static_assert(kGiRow >= 4 * kSlots, "row slots");
    SoftParams sp;
    const float2 *tab;  // LDS (A, S) table (padded)
    float *gi;          // this wave's [4][kGiRow] row copies of the home-atom sums ([i][4] per row)
    int hs;             // the current unit's first home position
    float racc[kRegGroups];  // this lane's row sum of value v for home atoms kG q .. kG q + kG - 1
    // a full chunk of unshifted, non-home j atoms waiting for a second one
    // (two j atoms per lane halve the row reductions per pair)
    bool held;
    float hx, hy, hz;
    unsigned hpos;
    const float4 *hhp;     // the unit's home atoms
    const FrameBox *hbox;  // the unit's box (alive until unit_end)
    __device__ __forceinline__ void add_grad(int f, unsigned pos, float gx, float gy, float gz) {
        float *g = sp.grad + (long long)f * 3 * sp.n + pos;
        atomicAdd(g, gx);
        atomicAdd(g + sp.n, gy);
        atomicAdd(g + 2 * sp.n, gz);
    }
    __device__ __forceinline__ static void wave_sync() {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
    // value v = 3 ii + component of a lane after row_reduce16 (v = 15: the pad)
    __device__ __forceinline__ static int lane_value() {
        const int lane = threadIdx.x & 63;
        return (lane & 12) | ((lane & 1) << 1) | ((lane >> 1) & 1);
    }
    __device__ __forceinline__ void preshift(float &xj, float &yj, float &zj, unsigned mj,
                                             const FrameBox &b) {
        soft_preshift<PBC>(xj, yj, zj, mj, b);
    }
    __device__ __forceinline__ void unit_begin(int, int h0, int) {
        hs = h0;
        held = false;
#pragma unroll
        for (int q = 0; q < kRegGroups; ++q) racc[q] = 0.f;
    }
    template <int MODE, bool HOME>
    __device__ __forceinline__ void chunk(const float4 *hp, int n_home, float xj, float yj, float zj,
                                          unsigned, int rank, unsigned jpos, int f, const FrameBox &b) {
        // a full chunk has finite positions in every lane, so a masked force
        // of zero gives zero products; lanes past a partial chunk hold NaN
        // positions and their products are masked instead
        const bool full = !__builtin_amdgcn_readfirstlane((int)__any(xj != xj));
        if constexpr (MODE == 0 && !HOME) {
            if (full) {
                if (held) {
                    pairs2(hp, n_home, hx, hy, hz, hpos, xj, yj, zj, jpos, f, b);
                    held = false;
                } else {
                    hx = xj; hy = yj; hz = zj; hpos = jpos;
                    hhp = hp;
                    hbox = &b;
                    held = true;
                }
                return;
            }
        }
        flush(hp, n_home, f, b);
        if (full) pairs<MODE, HOME, true>(hp, n_home, xj, yj, zj, rank, jpos, f, b);
        else pairs<MODE, HOME, false>(hp, n_home, xj, yj, zj, rank, jpos, f, b);
    }
    // the held chunk alone
    __device__ __forceinline__ void flush(const float4 *hp, int n_home, int f, const FrameBox &b) {
        if (held) {
            pairs<0, false, true>(hp, n_home, hx, hy, hz, 0x7FFFFFFF, hpos, f, b);
            held = false;
        }
    }
    // Two full chunks of unshifted, non-home j atoms (one each per lane): the
    // row reductions take both lanes' shares at once.
    __device__ __forceinline__ void pairs2(const float4 *hp, int n_home, float xa, float ya, float za,
                                           unsigned pa, float xb, float yb, float zb, unsigned pb,
                                           int f, const FrameBox &b) {
        float gax = 0.f, gay = 0.f, gaz = 0.f, gbx = 0.f, gby = 0.f, gbz = 0.f;
#pragma unroll
        for (int q = 0; q < kRegGroups; ++q) {
            if (kG * q >= n_home) break;
            racc[q] += group2(hp, kG * q, n_home, xa, ya, za, xb, yb, zb, b, gax, gay, gaz, gbx, gby, gbz);
        }
        const int row = (threadIdx.x & 63) >> 4;
        const int v = lane_value();
        const int ii = v / 3, c = v - 3 * ii;
        for (int i0 = kG * kRegGroups; i0 < n_home; i0 += kG) {
            const float r = group2(hp, i0, n_home, xa, ya, za, xb, yb, zb, b, gax, gay, gaz, gbx, gby, gbz);
            const int i = i0 + ii;
            if (v < 15 && i < n_home) {
                if (i < MAXH) atomicAdd(&gi[row * kGiRow + i * 4 + c], r);
                else atomicAdd(sp.grad + (long long)f * 3 * sp.n + c * sp.n + hs + i, -r);
            }
        }
        if (gax != 0.f || gay != 0.f || gaz != 0.f) add_grad(f, pa, gax, gay, gaz);
        if (gbx != 0.f || gby != 0.f || gbz != 0.f) add_grad(f, pb, gbx, gby, gbz);
    }
    __device__ __forceinline__ float group2(const float4 *hp, int i0, int n_home, float xa, float ya,
                                            float za, float xb, float yb, float zb, const FrameBox &b,
                                            float &gax, float &gay, float &gaz, float &gbx, float &gby,
                                            float &gbz) {
        float g[16];
#pragma unroll
        for (int ii = 0; ii < kG; ++ii) {
            const int i = i0 + ii;
            const float4 a = uload(hp, min(i, n_home - 1));
            float dx, dy, dz, ex, ey, ez;
            const float s = soft_disp<PBC, LT, 0>(a, xa, ya, za, b, dx, dy, dz);
            const float t = soft_disp<PBC, LT, 0>(a, xb, yb, zb, b, ex, ey, ez);
            const bool in_s = i < n_home && s < sp.cut2 && s > 0.f;
            const bool in_t = i < n_home && t < sp.cut2 && t > 0.f;
            const float fs0 = soft_force_over_r(s, sp, tab);  // both evaluated, then masked
            const float ft0 = soft_force_over_r(t, sp, tab);
            const float fs = in_s ? fs0 : 0.f;
            const float ft = in_t ? ft0 : 0.f;
            const float g0x = fs * dx, g0y = fs * dy, g0z = fs * dz;
            gax += g0x; gay += g0y; gaz += g0z;
            gbx = __builtin_fmaf(ft, ex, gbx);
            gby = __builtin_fmaf(ft, ey, gby);
            gbz = __builtin_fmaf(ft, ez, gbz);
            g[3 * ii] = __builtin_fmaf(ft, ex, g0x);
            g[3 * ii + 1] = __builtin_fmaf(ft, ey, g0y);
            g[3 * ii + 2] = __builtin_fmaf(ft, ez, g0z);
        }
        g[15] = 0.f;
        row_reduce16(g);
        return g[0];
    }
    // Home atoms i0 .. i0 + kG - 1 against the chunk: j's share added to gj,
    // the returned value is this lane's row sum of value v (the home share).
    template <int MODE, bool HOME, bool FULL>
    __device__ __forceinline__ float group(const float4 *hp, int i0, int n_home, float xj, float yj,
                                           float zj, int rank, const FrameBox &b, float &gjx,
                                           float &gjy, float &gjz) {
        float g[16];
#pragma unroll
        for (int ii = 0; ii < kG; ++ii) {
            const int i = i0 + ii;
            const float4 a = uload(hp, min(i, n_home - 1));
            float dx, dy, dz;
            const float s = soft_disp<PBC, LT, MODE>(a, xj, yj, zj, b, dx, dy, dz);
            const bool in = i < n_home && s < sp.cut2 && s > 0.f && (!HOME || i < rank);
            const float fr = soft_force_over_r(s, sp, tab);  // (dL/dr) / r
            float gx, gy, gz;
            if constexpr (FULL) {
                const float frm = in ? fr : 0.f;
                gx = frm * dx;
                gy = frm * dy;
                gz = frm * dz;
            } else {
                gx = in ? fr * dx : 0.f;
                gy = in ? fr * dy : 0.f;
                gz = in ? fr * dz : 0.f;
            }
            gjx += gx;
            gjy += gy;
            gjz += gz;
            g[3 * ii] = gx;
            g[3 * ii + 1] = gy;
            g[3 * ii + 2] = gz;
        }
        g[15] = 0.f;
        row_reduce16(g);
        return g[0];
    }
    template <int MODE, bool HOME, bool FULL>
    __device__ __forceinline__ void pairs(const float4 *hp, int n_home, float xj, float yj, float zj,
                                          int rank, unsigned jpos, int f, const FrameBox &b) {
        const int row = (threadIdx.x & 63) >> 4;
        const int v = lane_value();
        const int ii = v / 3, c = v - 3 * ii;
        float gjx = 0.f, gjy = 0.f, gjz = 0.f;
        // the first kG kRegGroups home atoms: this lane's row sums stay in registers
#pragma unroll
        for (int q = 0; q < kRegGroups; ++q) {
            if (kG * q >= n_home) break;
            racc[q] += group<MODE, HOME, FULL>(hp, kG * q, n_home, xj, yj, zj, rank, b, gjx, gjy, gjz);
        }
        // dense cells: LDS row slots up to MAXH, the output beyond
        for (int i0 = kG * kRegGroups; i0 < n_home; i0 += kG) {
            const float r = group<MODE, HOME, FULL>(hp, i0, n_home, xj, yj, zj, rank, b, gjx, gjy, gjz);
            const int i = i0 + ii;
            if (v < 15 && i < n_home) {
                if (i < MAXH) atomicAdd(&gi[row * kGiRow + i * 4 + c], r);
                else atomicAdd(sp.grad + (long long)f * 3 * sp.n + c * sp.n + hs + i, -r);
            }
        }
        if (jpos != 0xFFFFFFFFu && (gjx != 0.f || gjy != 0.f || gjz != 0.f)) add_grad(f, jpos, gjx, gjy, gjz);
    }
    __device__ __forceinline__ void unit_end(int f, int, int n_home) {
        if (held) flush(hhp, n_home, f, *hbox);
        const int lane = threadIdx.x & 63;
        const int row = lane >> 4;
        const int v = lane_value();
        const int ii = v / 3, c = v - 3 * ii;
        // register row sums into this lane's own slots (no other lane writes them)
#pragma unroll
        for (int q = 0; q < kRegGroups; ++q) {
            if (kG * q >= n_home) break;
            if (v < 15) gi[row * kGiRow + (kG * q + ii) * 4 + c] = racc[q];
        }
        wave_sync();
        for (int i = lane; i < min(n_home, kSlots); i += 64) {
            float t[3];
#pragma unroll
            for (int cc = 0; cc < 3; ++cc) {
                float acc = 0.f;
#pragma unroll
                for (int r = 0; r < 4; ++r) acc += gi[r * kGiRow + i * 4 + cc];
                t[cc] = acc;
            }
            add_grad(f, (unsigned)(hs + i), -t[0], -t[1], -t[2]);
        }
        wave_sync();
        // clear the slots the atomics use (past the register groups)
        const int used = min(n_home, MAXH) * 4;
        for (int r = 0; r < 4; ++r)
            for (int e = 4 * kG * kRegGroups + lane; e < used; e += 64) gi[r * kGiRow + e] = 0.f;
        wave_sync();
    }
};

// grad[f][d][i] += gs[f][d][pos[f][i]]: cell order back to store slots
// (pos: store slot -> sorted position; writes in store order, reads gathered
// from the frame's cell-order gradient, which stays in L2).
__global__ __launch_bounds__(kBlock) void soft_unsort_kernel(const float *__restrict__ gs,
                                                             const unsigned *__restrict__ pos,
                                                             int n, int n_frames, int stride,
                                                             float *__restrict__ grad) {
    const long long total = (long long)n * n_frames;
    for (long long t = (long long)blockIdx.x * kBlock + threadIdx.x; t < total;
         t += (long long)gridDim.x * kBlock) {
        const int f = (int)(t / n);
        const int i = (int)(t - (long long)f * n);
        const float *src = gs + (long long)f * 3 * n + pos[t];
        float *dst = grad + (long long)f * 3 * stride + i;
        dst[0] += src[0];
        dst[stride] += src[n];
        dst[2 * stride] += src[2 * n];
    }
}

constexpr int kSoftBlock = 1024;  // 16 waves: two workgroups per CU, 8 waves per SIMD (registers capped for 8)
constexpr int kSoftWaves = kSoftBlock / 64;

template <int PBC, bool LT>
__global__ __launch_bounds__(kSoftBlock) __attribute__((amdgpu_waves_per_eu(8, 8))) void soft_cell_fwd_kernel(CellGrid G, SoftParams sp) {
    extern __shared__ __attribute__((aligned(16))) char lds[];
    const int hn = sp.n_nodes + kSoftPadL + kSoftPadR;
    unsigned long long *hist = (unsigned long long *)lds;
    float *pbufs = (float *)(lds + align16((size_t)hn * 8));
    float4 *rings = (float4 *)((char *)pbufs + kSoftWaves * 128 * 4);
    for (int i = threadIdx.x; i < hn; i += kSoftBlock) hist[i] = 0ull;
    __syncthreads();
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    SoftFwdCons<PBC, LT> cons{sp, hist, pbufs + wave * 128, 0};
    cell_waves<PBC, false>(G, rings + wave * kRingStride, nullptr, wave, cons);
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    if (cons.pt > 0) cons.drain(cons.pt);
    __syncthreads();
    for (int i = threadIdx.x; i < sp.n_nodes; i += kSoftBlock) {
        const unsigned long long v = hist[kSoftPadL + i];
        if (v != 0ull) atomicAdd(&sp.nodes[i], (double)v * kSoftUnfix);
    }
}

constexpr int kSoftBwdBlock = 512;  // 8 waves: two workgroups per CU, 126 VGPRs without spills
constexpr int kSoftBwdWaves = kSoftBwdBlock / 64;

// LDS of the backward cell kernel: (A, S) table | per wave: ring[kRingStride]
// float4, ring positions [kRing], the home-atom row sums [4][soft_gi_row(maxh)].
__host__ __device__ inline size_t soft_bwd_lds_bytes(int tn, int maxh) {
    return align16((size_t)tn * 8) + (size_t)kSoftBwdWaves * (kRingStride * 16 + kRing * 4 + 4 * soft_gi_row(maxh) * 4);
}

template <int PBC, bool LT, int MAXH>
__global__ __launch_bounds__(kSoftBwdBlock) void soft_cell_bwd_kernel(CellGrid G, SoftParams sp) {
    constexpr int kGiRow = soft_gi_row(MAXH);
    extern __shared__ __attribute__((aligned(16))) char lds[];
    const int tn = sp.n_nodes + kSoftPadL + kSoftPadR;
    float2 *tab = (float2 *)lds;
    float4 *rings = (float4 *)(lds + align16((size_t)tn * 8));
    unsigned *ringps = (unsigned *)(rings + kSoftBwdWaves * kRingStride);
    float *gis = (float *)(ringps + kSoftBwdWaves * kRing);
    for (int i = threadIdx.x; i < tn; i += kSoftBwdBlock) tab[i] = sp.table[i];
    for (int i = threadIdx.x; i < kSoftBwdWaves * 4 * kGiRow; i += kSoftBwdBlock) gis[i] = 0.f;
    __syncthreads();
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    SoftBwdCons<PBC, LT, MAXH> cons{sp, tab, gis + wave * 4 * kGiRow, 0};
    cell_waves<PBC, true>(G, rings + wave * kRingStride, ringps + wave * kRing, wave, cons);
}

// Brute-force soft kernels (boxes too small for the cell list): one thread per
// atom i of a frame; the forward takes j > i, the backward every j != i and
// writes dL/dx_i itself.  Minimum image by roundf(d / L), the reference rule.
template <int PBC>
__device__ __forceinline__ float soft_pair_vec(float xi, float yi, float zi, float xj, float yj,
                                               float zj, const FrameBox &b, float &dx, float &dy,
                                               float &dz) {
    dx = xj - xi;
    dy = yj - yi;
    dz = zj - zi;
    if constexpr (PBC == WMD_PBC_ORTHO) {
        dx -= roundf(dx / b.lx) * b.lx;
        dy -= roundf(dy / b.ly) * b.ly;
        dz -= roundf(dz / b.lz) * b.lz;
    } else if constexpr (PBC == WMD_PBC_TRICLINIC) {
        float f0 = b.r[0] * dx + b.r[3] * dy + b.r[6] * dz;
        float f1 = b.r[1] * dx + b.r[4] * dy + b.r[7] * dz;
        float f2 = b.r[2] * dx + b.r[5] * dy + b.r[8] * dz;
        f0 -= roundf(f0);
        f1 -= roundf(f1);
        f2 -= roundf(f2);
        dx = f0 * b.c[0] + f1 * b.c[3] + f2 * b.c[6];
        dy = f0 * b.c[1] + f1 * b.c[4] + f2 * b.c[7];
        dz = f0 * b.c[2] + f1 * b.c[5] + f2 * b.c[8];
    }
    return dx * dx + dy * dy + dz * dz;
}

template <int PBC>
__global__ __launch_bounds__(kBlock) void soft_brute_fwd_kernel(const float *__restrict__ soa,
                                                                int stride, int n,
                                                                const FrameBox *__restrict__ boxes,
                                                                SoftParams sp) {
    extern __shared__ __attribute__((aligned(16))) char lds[];
    const int hn = sp.n_nodes + kSoftPadL + kSoftPadR;
    unsigned long long *hist = (unsigned long long *)lds;
    for (int t = threadIdx.x; t < hn; t += kBlock) hist[t] = 0ull;
    __syncthreads();
    const int f = blockIdx.y;
    const float *X = soa + (long long)f * 3 * stride;
    const FrameBox b = boxes[f];
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i < n) {
        const float xi = X[i], yi = X[stride + i], zi = X[2 * stride + i];
        for (int j = i + 1; j < n; ++j) {
            float dx, dy, dz;
            const float s = soft_pair_vec<PBC>(xi, yi, zi, X[j], X[stride + j], X[2 * stride + j], b, dx, dy, dz);
            if (s < sp.cut2) soft_deposit(s, sp, hist);
        }
    }
    __syncthreads();
    for (int t = threadIdx.x; t < sp.n_nodes; t += kBlock) {
        const unsigned long long v = hist[kSoftPadL + t];
        if (v != 0ull) atomicAdd(&sp.nodes[t], (double)v * kSoftUnfix);
    }
}

template <int PBC>
__global__ __launch_bounds__(kBlock) void soft_brute_bwd_kernel(const float *__restrict__ soa,
                                                                int stride, int n,
                                                                const FrameBox *__restrict__ boxes,
                                                                SoftParams sp) {
    extern __shared__ __attribute__((aligned(16))) char lds[];
    const int tn = sp.n_nodes + kSoftPadL + kSoftPadR;
    float2 *tab = (float2 *)lds;
    for (int t = threadIdx.x; t < tn; t += kBlock) tab[t] = sp.table[t];
    __syncthreads();
    const int f = blockIdx.y;
    const float *X = soa + (long long)f * 3 * stride;
    const FrameBox b = boxes[f];
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    const float xi = X[i], yi = X[stride + i], zi = X[2 * stride + i];
    float gx = 0.f, gy = 0.f, gz = 0.f;
    for (int j = 0; j < n; ++j) {
        if (j == i) continue;
        float dx, dy, dz;
        const float s = soft_pair_vec<PBC>(xi, yi, zi, X[j], X[stride + j], X[2 * stride + j], b, dx, dy, dz);
        if (s < sp.cut2 && s > 0.f) {
            const float fr = soft_force_over_r(s, sp, tab);
            gx -= fr * dx;
            gy -= fr * dy;
            gz -= fr * dz;
        }
    }
    float *g = sp.grad + (sp.grad_frame0 + f) * 3 * (long long)stride + i;
    g[0] += gx;
    g[stride] += gy;
    g[2 * stride] += gz;
}

// c_k += sum over |m h - r_k| <= 6 sigma of H_m w phi(r_k - m h), one thread per bin.
// The bin-node weights depend only on j = kSoftNodesPerBin k + kSoftNodesPerBin / 2 - m
// (r_k - m h = j h): wt[J + j] = w phi_sigma(j h) for |j| <= J (6 sigma / h),
// tabulated once per plan.
constexpr int kSoftMid = kSoftNodesPerBin / 2;

// c_k += sum_m H_m wt[J + j(k, m)].
__global__ void soft_conv_kernel(const double *__restrict__ nodes, int n_nodes, int bins,
                                 const double *__restrict__ wt, int J, double *__restrict__ counts) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= bins) return;
    const int c0 = kSoftNodesPerBin * k + kSoftMid;
    const int m0 = max(0, c0 - J), m1 = min(n_nodes - 1, c0 + J);
    double c = 0.0;
    for (int m = m0; m <= m1; ++m) c += nodes[m] * wt[J + c0 - m];
    counts[k] += c;
}

// D_m = sum_k dL/dc_k wt[J + j(k, m)] for a padded node e (node e - kSoftPadL;
// zero outside 0 .. n_nodes - 1).
__device__ __forceinline__ double soft_node_d(const double *__restrict__ dldc, int bins,
                                              const double *__restrict__ wt, int J, int n_nodes,
                                              int e) {
    const int m = e - kSoftPadL;
    double v = 0.0;
    if (m >= 0 && m < n_nodes) {
        // k with |kSoftNodesPerBin k + kSoftMid - m| <= J (floor divisions)
        auto fdiv = [](int a) { return a >= 0 ? a / kSoftNodesPerBin : -((-a + kSoftNodesPerBin - 1) / kSoftNodesPerBin); };
        const int k0 = max(0, -fdiv(-(m - kSoftMid - J)));
        const int k1 = min(bins - 1, fdiv(m - kSoftMid + J));
        for (int k = k0; k <= k1; ++k) v += dldc[k] * wt[J + kSoftNodesPerBin * k + kSoftMid - m];
    }
    return v;
}

// The padded (A, S) table of the backward: A_e = (D_{e+1} - D_{e-1}) / h,
// S_e = 2 (D_{e-1} - 2 D_e + D_{e+1}) / h, each thread forming the three
// node values it needs.
__global__ void soft_table_kernel(const double *__restrict__ dldc, int bins,
                                  const double *__restrict__ wt, int J, double inv_h, int n_nodes,
                                  float2 *__restrict__ table) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    const int tn = n_nodes + kSoftPadL + kSoftPadR;
    if (e >= tn) return;
    const double dm = e > 0 ? soft_node_d(dldc, bins, wt, J, n_nodes, e - 1) : 0.0;
    const double d0 = soft_node_d(dldc, bins, wt, J, n_nodes, e);
    const double dp = e + 1 < tn ? soft_node_d(dldc, bins, wt, J, n_nodes, e + 1) : 0.0;
    table[e] = make_float2((float)((dp - dm) * inv_h), (float)(2.0 * (dm - 2.0 * d0 + dp) * inv_h));
}

__global__ __launch_bounds__(kBlock) void gather_soa_kernel(const float4 *__restrict__ aos,
                                                            long long n_src, const unsigned *__restrict__ idx,
                                                            int n_store, int stride, int n_frames,
                                                            float *__restrict__ soa, int planar = 0) {
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
                x = v.x; y = v.y; z = planar ? 0.f : v.z;
            }
        }
        float *o = soa + (long long)f * 3 * stride;
        o[k] = x;
        o[stride + k] = y;
        o[2 * stride + k] = z;
    }
}

// Group centres of the grouped RDF modes (group_position, rdf.rs:353-366, over
// center_of_selection, traj-core/src/centers.rs:3-29): per frame and IO group,
// f64 sums of x * m over the group's atoms in order divided by the summed m
// (m = 1 unless mass weighted) and cast to f32; a one-atom group is the atom
// itself, an empty group or a zero mass sum the origin.  planar: z = 0.
__global__ __launch_bounds__(kBlock) void group_center_kernel(
    const float4 *__restrict__ aos, long long n_src, const unsigned long long *__restrict__ goff,
    const unsigned *__restrict__ gatoms, const float *__restrict__ gmass, int n_groups, int stride,
    int n_frames, int planar, float *__restrict__ soa) {
    const long long total = (long long)stride * n_frames;
    for (long long t = (long long)blockIdx.x * kBlock + threadIdx.x; t < total;
         t += (long long)gridDim.x * kBlock) {
        const int f = (int)(t / stride);
        const int k = (int)(t - (long long)f * stride);
        float x = 0.f, y = 0.f, z = 0.f;
        if (k < n_groups) {
            const float4 *fr = aos + (long long)f * n_src;
            const unsigned long long b = goff[k], e = goff[k + 1];
            if (e - b == 1) {
                const float4 p = fr[gatoms[b]];
                x = p.x; y = p.y; z = p.z;
            } else if (e > b) {
                double s0 = 0.0, s1 = 0.0, s2 = 0.0, ms = 0.0;
                for (unsigned long long i = b; i < e; ++i) {
                    const float4 p = fr[gatoms[i]];
                    const double m = gmass ? (double)gmass[i] : 1.0;
                    s0 += (double)p.x * m;
                    s1 += (double)p.y * m;
                    s2 += (double)p.z * m;
                    ms += m;
                }
                if (ms != 0.0) {
                    x = (float)(s0 / ms);
                    y = (float)(s1 / ms);
                    z = (float)(s2 / ms);
                }
            }
        }
        if (planar) z = 0.f;
        float *o = soa + (long long)f * 3 * stride;
        o[k] = x;
        o[stride + k] = y;
        o[2 * stride + k] = z;
    }
}

// Copy of a SoA store with z = 0 (PlanarXY on a caller's device store).
__global__ __launch_bounds__(kBlock) void planar_copy_kernel(const float *__restrict__ src,
                                                             float *__restrict__ dst, long long stride,
                                                             long long n_frames) {
    for (long long t = (long long)blockIdx.x * kBlock + threadIdx.x; t < stride * n_frames;
         t += (long long)gridDim.x * kBlock) {
        const long long f = t / stride, k = t - f * stride;
        const long long o = 3 * f * stride + k;
        dst[o] = src[o];
        dst[o + stride] = src[o + stride];
        dst[o + 2 * stride] = 0.f;
    }
}

// Removes the intramolecular pairs of no_intramol (rdf.rs:822-824,
// same_intramolecular_group :368-373) from the counts the pair kernels made:
// each excluded ordered pair (u from A, v from B, weight mA(u) mB(v)) is
// evaluated with the reference's own expression (f32 difference, dx -=
// round(dx / l) l, s = dx^2 + dy^2 + dz^2, no contraction) and its weight
// subtracted from the bin the exact breakpoints thr give.
template <int PBC>
__global__ __launch_bounds__(kBlock) void rdf_excl_kernel(const float *__restrict__ soa, int stride,
                                                          int n_frames, const FrameBox *__restrict__ boxes,
                                                          const uint4 *__restrict__ excl, long long n_excl,
                                                          const float *__restrict__ thr, int bins,
                                                          unsigned long long *counts) {
    const long long total = n_excl * n_frames;
    for (long long t = (long long)blockIdx.x * kBlock + threadIdx.x; t < total;
         t += (long long)gridDim.x * kBlock) {
        const int f = (int)(t / n_excl);
        const uint4 e = excl[t - (long long)f * n_excl];
        const float *x = soa + (long long)f * 3 * stride;
        float dx = x[e.y] - x[e.x];
        float dy = x[stride + e.y] - x[stride + e.x];
        float dz = x[2 * stride + e.y] - x[2 * stride + e.x];
        if constexpr (PBC == WMD_PBC_ORTHO) {
            const FrameBox b = boxes[f];
            dx -= roundf(dx / b.lx) * b.lx;
            dy -= roundf(dy / b.ly) * b.ly;
            dz -= roundf(dz / b.lz) * b.lz;
        }
        const float s = dx * dx + dy * dy + dz * dz;
        if (s < thr[bins]) {
            int lo = 0, hi = bins;  // largest k < bins with thr[k] <= s
            while (hi - lo > 1) {
                const int mid = (lo + hi) >> 1;
                if (thr[mid] <= s) lo = mid;
                else hi = mid;
            }
            atomicAdd(&counts[lo], (unsigned long long)0 - (unsigned long long)e.z);
        }
    }
}

// -------------------------------------------------------------------------
// f64 distances (StructureFactorPlan's CPU path, structure_factor.rs:160-198):
// the f32 difference cast to f64 and scaled, the orthorhombic wrap in f64 on
// the scaled box, s = dx^2 + dy^2 + dz^2 in f64 (left-to-right, unfused).  A
// pair counts iff s < thr64[bins]; its bin comes from an estimate corrected
// against the f64 breakpoints thr64 (exact, from host bisection over f64 bit
// patterns with IEEE sqrt and division).  Brute force over 256-atom tiles.
// -------------------------------------------------------------------------
// round(d / l) * l subtracted from d, with the division only near a tie:
// q = d * (1 / l) is within 5e-7 of the correctly rounded quotient for
// |q| < 1e9, so rint(q) is exact when q lies more than 1e-4 from a half-integer.
__device__ __forceinline__ double wrap64(double d, double l, double il) {
    const double q = d * il;
    double n = rint(q);
    if (!(fabs(q - n) <= 0.4999) || !(fabs(q) < 1e9)) n = round(d / l);
    return d - n * l;
}

template <int PBC>
__device__ __forceinline__ double pair_s64(float xi, float yi, float zi, float xj, float yj, float zj,
                                           double scale, double lx, double ly, double lz, double ilx,
                                           double ily, double ilz) {
    double dx = (double)(xj - xi) * scale;
    double dy = (double)(yj - yi) * scale;
    double dz = (double)(zj - zi) * scale;
    if constexpr (PBC == WMD_PBC_ORTHO) {
        dx = wrap64(dx, lx, ilx);
        dy = wrap64(dy, ly, ily);
        dz = wrap64(dz, lz, ilz);
    }
    return dx * dx + dy * dy + dz * dz;
}

template <int PBC, bool WEIGHTED>
__global__ __launch_bounds__(kBlock) void rdf_brute64_kernel(
    const float *__restrict__ soa, long long frame_stride, int stride, int n, int n_frames,
    const FrameBox *__restrict__ boxes, const unsigned *__restrict__ wts,
    const int2 *__restrict__ tile_pairs, int n_tile_pairs, long long frame0, int bins,
    const double *__restrict__ thr64, double inv_w, double scale, unsigned long long *counts,
    int n_xcd, unsigned w_max) {
    extern __shared__ __attribute__((aligned(16))) char lds[];
    double *thr = (double *)lds;                                  // [bins + 1]
    unsigned *hist = (unsigned *)(thr + bins + 1);                // [bins]
    float *jx = (float *)(hist + ((bins + 3) & ~3));              // [kTile] x 3
    float *jy = jx + kTile;
    float *jz = jy + kTile;
    unsigned *jw = (unsigned *)(jz + kTile);
    for (int i = threadIdx.x; i <= bins; i += kBlock) thr[i] = thr64[i];
    for (int i = threadIdx.x; i < bins; i += kBlock) hist[i] = 0u;
    const double cut = thr64[bins];
    const long long n_units = (long long)n_frames * n_tile_pairs;
    UnitRange ur = xcd_units(n_units, n_xcd);
    unsigned long long budget = 0;
    auto flush = [&]() {
        __syncthreads();
        for (int i = threadIdx.x; i < bins; i += kBlock) {
            const unsigned v = hist[i];
            if (v) {
                atomicAdd(&counts[i], (unsigned long long)v);
                hist[i] = 0u;
            }
        }
        __syncthreads();
    };
    auto count = [&](double s, unsigned w) {
        if (s < cut) {
            int b = (int)(sqrt(s) * inv_w);
            b = min(max(b, 0), bins - 1);
            b += (int)(s >= thr[b + 1]) - (int)(s < thr[b]);
            atomicAdd(&hist[b], w);
        }
    };
    __syncthreads();
    for (long long u = ur.begin; u < ur.end; u += ur.step) {
        const int f = (int)(u / n_tile_pairs);
        const int2 tp = tile_pairs[u - (long long)f * n_tile_pairs];
        const float *fx = soa + (frame0 + f) * frame_stride;
        const float *fy = fx + stride;
        const float *fz = fy + stride;
        const FrameBox b = boxes[f];
        const double lx = (double)b.lx * scale, ly = (double)b.ly * scale, lz = (double)b.lz * scale;
        const double ilx = PBC == WMD_PBC_ORTHO ? 1.0 / lx : 0.0;
        const double ily = PBC == WMD_PBC_ORTHO ? 1.0 / ly : 0.0;
        const double ilz = PBC == WMD_PBC_ORTHO ? 1.0 / lz : 0.0;
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
                double s = pair_s64<PBC>(xi, yi, zi, jx[jj], jy[jj], jz[jj], scale, lx, ly, lz, ilx, ily, ilz);
                if (jj == (int)threadIdx.x) s = __builtin_nan("");
                unsigned w = 1u;
                if constexpr (WEIGHTED) w = (wi & 0xFFu) * (jw[jj] >> 8);
                count(s, w);
            }
        } else {
            // unordered pairs: both orders give the same f64 s (the f32
            // difference, the scaling and the wrap are all odd functions)
            for (int jj = 0; jj < nj; ++jj) {
                const double s = pair_s64<PBC>(xi, yi, zi, jx[jj], jy[jj], jz[jj], scale, lx, ly, lz, ilx, ily, ilz);
                unsigned w = 2u;
                if constexpr (WEIGHTED) {
                    const unsigned wj = jw[jj];
                    w = (wi & 0xFFu) * (wj >> 8) + (wj & 0xFFu) * (wi >> 8);
                }
                count(s, w);
            }
        }
        budget += (unsigned long long)kTile * kTile * w_max;
        if (budget > 0x7FFFFFFFull) {
            budget = 0;
            flush();
        }
    }
    flush();
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
    std::vector<unsigned short> tab;
    std::vector<uint2> tab2;
    int tab_sh = 0, tab_base = 0;
    int cell_k = 0;        // 0 = automatic
    int grab = 0;          // cell units per queue grab, 0 = automatic
    int last_path = 0;
    int last_nc[3] = {0, 0, 0};
    DevBuf d_counts, d_thr, d_elo, d_ehi, d_tab, d_tab2, d_wts, d_io, d_boxes;
    DevBuf d_tiles, d_queue;
    int tiles_n = -1;  // atom count the tile pairs in d_tiles were built for
    CellScratch cells;
    HostBuf h_boxes;
    // host-chunk pipeline: two staging / device buffers, events guarding reuse
    HostBuf p_stage[2];
    DevBuf p_aos[2], p_soa[2];
    Event p_copied[2], p_used[2], ev_boxes;
    // grouped modes (centre / byres / bymol / no_intramol) and PlanarXY: the
    // IO "atoms" are group centres computed from full chunks; exclusions are
    // the no_intramol ordered pairs (u, v, weight) in IO positions
    bool grouped = false, planar = false;
    DevBuf d_goff, d_gatoms, d_gmass, d_excl, d_pstore;
    long long n_excl = 0;
    uint64_t g_max_atom = 0;  // chunks must hold at least this many atoms
    // f64 distance mode (StructureFactorPlan): thresholds on the f64 squared
    // distance, r_max and the length scale in f64
    bool f64 = false;
    double r_max64 = 0.0, scale64 = 1.0;
    std::vector<double> thr64;
    DevBuf d_thr64;
    // bin slices of a histogram larger than the pair kernels' LDS: slice_bins
    // global bins per pass; per slice [slice_bins + 2] local thresholds (and
    // edge bounds): a bin below the slice, then the slice's own
    int slice_bins = 0;
    DevBuf d_sthr, d_selo, d_sehi;
};

namespace {

// One shared LDS histogram: LDS atomics of different waves are separate
// instructions, so per-wave copies would not reduce conflicts.
int hist_copies_for(int) { return 1; }

template <class K> int occupancy(K kernel, size_t lds) {
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, kernel, kBlock, lds) != hipSuccess ||
        nb < 1)
        nb = 1;
    return nb;
}

int persistent_grid(wmd_rdf *h, int occ, long long n_units) {
    int grid = h->ctx->cu_count * occ;
    grid = std::max(8, grid / 8 * 8);
    long long max_grid = (n_units + 7) / 8 * 8;
    if (grid > max_grid) grid = (int)std::max<long long>(8, max_grid);
    return grid;
}

HistParams hist_params(wmd_rdf *h, bool table) {
    HistParams p;
    p.bins = h->bins;
    p.hist_copies = hist_copies_for(h->bins);
    p.cut = h->cut;
    p.inv_w = 1.0f / h->bin_width;
    p.thr = h->d_thr.as<float>();
    p.elo = h->d_elo.as<float>();
    p.ehi = h->d_ehi.as<float>();
    p.tab = table ? h->d_tab.as<unsigned short>() : nullptr;
    p.tab2 = table ? h->d_tab2.as<uint2>() : nullptr;
    p.tab_n = table ? (int)h->tab.size() : 0;
    p.tab_sh = h->tab_sh;
    p.tab_base = h->tab_base;
    p.counts = h->d_counts.as<unsigned long long>();
    p.edge = h->d_counts.as<unsigned long long>() + h->bins;
    return p;
}

size_t lds_hist_bytes(const HistParams &p, bool edge) {
    return lds_hist_size(p.bins, p.hist_copies, edge, p.tab_n);
}

// The histogram passes of one pair enumeration.  When the kernel's LDS
// (lds_of(local bins, with table)) holds every bin: one pass with p.  Else
// the bins are counted slice by slice, each pass a full enumeration whose
// histogram is one local bin for the pairs below the slice (dropped) plus
// the slice's bins; pairs above it fail the pass's cut.  *table is cleared
// for sliced passes (the bin estimate from sqrt(s) needs no table).
template <class LdsOf>
std::vector<HistParams> hist_passes(wmd_rdf *h, const HistParams &p, bool edge, bool *table,
                                    LdsOf lds_of, hipStream_t st) {
    const size_t limit = h->ctx->lds_per_block;
    if (lds_of(h->bins, *table) <= limit) return {p};
    int S = std::min(h->bins, 1 << 16) / 64 * 64;
    while (S > 64 && lds_of(S + 1, false) > limit) S -= 64;
    WMD_REQUIRE(S >= 64 && lds_of(S + 1, false) <= limit, WMD_ERR_UNSUPPORTED, "too many bins for LDS");
    const int n_slices = (h->bins + S - 1) / S;
    if (h->slice_bins != S) {
        const size_t per = (size_t)S + 2;
        std::vector<float> thr(per * n_slices, 0.0f), elo(per * n_slices, 0.0f), ehi(per * n_slices, 0.0f);
        for (int k = 0; k < n_slices; ++k) {
            const int b0 = k * S, nb = std::min(S, h->bins - b0);
            for (int j = 0; j <= nb; ++j) {
                thr[k * per + j + 1] = h->thr[b0 + j];
                elo[k * per + j + 1] = h->elo[b0 + j];
                ehi[k * per + j + 1] = h->ehi[b0 + j];
            }
        }
        WMD_HIP_CHECK(hipStreamSynchronize(st));  // growing frees buffers queued work may use
        h->d_sthr.reserve(thr.size() * 4);
        h->d_selo.reserve(elo.size() * 4);
        h->d_sehi.reserve(ehi.size() * 4);
        WMD_HIP_CHECK(hipMemcpyAsync(h->d_sthr.ptr, thr.data(), thr.size() * 4, hipMemcpyHostToDevice, st));
        WMD_HIP_CHECK(hipMemcpyAsync(h->d_selo.ptr, elo.data(), elo.size() * 4, hipMemcpyHostToDevice, st));
        WMD_HIP_CHECK(hipMemcpyAsync(h->d_sehi.ptr, ehi.data(), ehi.size() * 4, hipMemcpyHostToDevice, st));
        WMD_HIP_CHECK(hipStreamSynchronize(st));  // the host vectors die at scope exit
        h->slice_bins = S;
    }
    std::vector<HistParams> out;
    for (int k = 0; k < n_slices; ++k) {
        const int b0 = k * S, nb = std::min(S, h->bins - b0);
        HistParams q = p;
        q.bins = nb + 1;
        q.cut = h->thr[b0 + nb];
        q.thr = h->d_sthr.as<float>() + (size_t)k * (S + 2);
        q.elo = h->d_selo.as<float>() + (size_t)k * (S + 2);
        q.ehi = h->d_sehi.as<float>() + (size_t)k * (S + 2);
        q.tab = nullptr;
        q.tab2 = nullptr;
        q.tab_n = 0;
        q.shift = b0 - 1;
        q.skip = 1;
        q.counts = p.counts + b0;
        q.count_evals = k == 0;
        out.push_back(q);
    }
    *table = false;
    return out;
}

// Dispatch helpers over the template flags.
template <int PBC, bool EXACT, bool WEIGHTED, bool EDGE, bool TABLE>
void launch_brute_t(wmd_rdf *h, hipStream_t st, const float *soa, long long frame_stride,
                    int stride, int n, int nf, long long frame0, const FrameBox *boxes,
                    const int2 *tiles, int n_tp, const HistParams &p) {
    auto kern = rdf_brute_kernel<PBC, EXACT, WEIGHTED, EDGE, TABLE>;
    ProfScope prof(h->ctx, st, "rdf_pair");
    size_t lds = lds_hist_bytes(p, EDGE) + kTile * 16;
    WMD_REQUIRE(lds <= h->ctx->lds_per_block, WMD_ERR_UNSUPPORTED, "too many bins for LDS");
    int grid = persistent_grid(h, occupancy(kern, lds), (long long)nf * n_tp);
    hipLaunchKernelGGL(kern, dim3(grid), dim3(kBlock), lds, st, soa, frame_stride, stride, n, nf,
                       boxes, (const unsigned *)h->d_wts.ptr, tiles, n_tp, frame0, p, 8, h->w_max);
    WMD_HIP_CHECK(hipGetLastError());
}

template <int PBC, bool WEIGHTED, bool EDGE, bool TABLE>
void launch_cells_t(wmd_rdf *h, hipStream_t st, const float4 *sorted, const unsigned *start,
                    const unsigned *frame_flag, const float4 *bbox, int n, int nf, const FrameBox *boxes,
                    const int *nc, int k, int shift_ok, float cull2, const HistParams &p) {
    auto kern = rdf_cell_kernel<PBC, WEIGHTED, EDGE, TABLE>;
    ProfScope prof(h->ctx, st, "rdf_pair");
    size_t lds = cell_lds_bytes(p.bins, EDGE, TABLE ? p.tab_n : 0);
    WMD_REQUIRE(lds <= h->ctx->lds_per_block, WMD_ERR_UNSUPPORTED, "too many bins for LDS");
    int occ = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, kern, kCellBlock, lds) != hipSuccess || occ < 1)
        occ = 1;
    const long long n_units = (long long)nf * nc[0] * nc[1] * nc[2];
    int grid = std::max(8, h->ctx->cu_count * occ / 8 * 8);
    const long long max_grid = (n_units + kCellWaves * 8 - 1) / (kCellWaves * 8) * 8;
    WMD_REQUIRE(k >= 1 && 2 * (2 * k * k + 2 * k + 1) < 64, WMD_ERR_INVALID, "cell stencil reach out of range");
    if (grid > max_grid) grid = (int)std::max<long long>(8, max_grid);
    // dynamic unit schedule: the per-XCD grab counters start at zero
    unsigned long long *queue = (unsigned long long *)h->d_queue.reserve(8 * 8);
    hipLaunchKernelGGL(fill_u64_kernel, dim3(1), dim3(kBlock), 0, st, queue, 8LL);
    WMD_HIP_CHECK(hipGetLastError());
    // a larger grab halves the traffic on each XCD's queue counter (which
    // otherwise stalls the waves of a long launch); short launches keep
    // grabs of 2 so that the last grabs stay balanced
    const int grab = h->grab > 0 ? h->grab : n_units >= 64LL * grid * kCellWaves ? 4 : 2;
    CellGrid G{sorted, start, frame_flag, boxes, n, nf, nc[0], nc[1], nc[2], k, shift_ok, 8, cull2, queue, grab, bbox};
    hipLaunchKernelGGL(kern, dim3(grid), dim3(kCellBlock), lds, st, G, p);
    WMD_HIP_CHECK(hipGetLastError());
}

template <bool EXACT, bool WEIGHTED, bool EDGE, bool TABLE, class... A>
void brute_pbc(int pbc, A... a) {
    if (pbc == WMD_PBC_NONE) launch_brute_t<WMD_PBC_NONE, false, WEIGHTED, EDGE, TABLE>(a...);
    else if (pbc == WMD_PBC_ORTHO) launch_brute_t<WMD_PBC_ORTHO, EXACT, WEIGHTED, EDGE, TABLE>(a...);
    else launch_brute_t<WMD_PBC_TRICLINIC, EXACT, WEIGHTED, EDGE, TABLE>(a...);
}

template <class... A>
void brute_dispatch(bool exact, bool weighted, bool edge, bool table, int pbc, A... a) {
#define WMD_BRUTE(E, W, D, T) \
    if (exact == E && weighted == W && edge == D && table == T) return brute_pbc<E, W, D, T>(pbc, a...);
    WMD_BRUTE(false, false, false, false) WMD_BRUTE(false, false, false, true)
    WMD_BRUTE(false, false, true, false) WMD_BRUTE(false, false, true, true)
    WMD_BRUTE(false, true, false, false) WMD_BRUTE(false, true, false, true)
    WMD_BRUTE(false, true, true, false) WMD_BRUTE(false, true, true, true)
    WMD_BRUTE(true, false, false, false) WMD_BRUTE(true, false, false, true)
    WMD_BRUTE(true, false, true, false) WMD_BRUTE(true, false, true, true)
    WMD_BRUTE(true, true, false, false) WMD_BRUTE(true, true, false, true)
    WMD_BRUTE(true, true, true, false) WMD_BRUTE(true, true, true, true)
#undef WMD_BRUTE
}

template <class... A>
void cells_dispatch(int pbc, bool weighted, bool edge, bool table, A... a) {
#define WMD_CELLS(P, W, D, T)                                       \
    if (pbc == P && weighted == W && edge == D && table == T)       \
        return launch_cells_t<P, W, D, T>(a...);
#define WMD_CELLS_P(P) \
    WMD_CELLS(P, false, false, false) WMD_CELLS(P, false, false, true) \
    WMD_CELLS(P, false, true, false) WMD_CELLS(P, false, true, true)   \
    WMD_CELLS(P, true, false, false) WMD_CELLS(P, true, false, true)   \
    WMD_CELLS(P, true, true, false) WMD_CELLS(P, true, true, true)
    WMD_CELLS_P(WMD_PBC_ORTHO)
    WMD_CELLS_P(WMD_PBC_TRICLINIC)
#undef WMD_CELLS_P
#undef WMD_CELLS
    fail(WMD_ERR_INVALID, "no cell-list kernel for this combination");
}


// Tile pairs (it <= jt, upper triangle) of the brute-force kernels, uploaded
// to h->d_tiles; returns their number.
int upload_tile_pairs(wmd_rdf *h, int n, hipStream_t st) {
    const int tiles = (n + kTile - 1) / kTile;
    const int n_tp = tiles * (tiles + 1) / 2;
    if (h->tiles_n == n) return n_tp;  // d_tiles already holds them
    std::vector<int2> tp;
    tp.reserve(n_tp);
    for (int i = 0; i < tiles; ++i)
        for (int j = i; j < tiles; ++j) tp.push_back(make_int2(i, j));
    int2 *d_tp = (int2 *)h->d_tiles.reserve(sizeof(int2) * std::max(n_tp, 1));
    WMD_HIP_CHECK(hipMemcpyAsync(d_tp, tp.data(), sizeof(int2) * n_tp, hipMemcpyHostToDevice, st));
    WMD_HIP_CHECK(hipStreamSynchronize(st));  // the vector dies at scope exit
    h->tiles_n = n;
    return n_tp;
}

// f64-distance counts (StructureFactorPlan) on the device store.
void accumulate_f64(wmd_rdf *h, const float *d_soa, size_t stride, size_t n_frames,
                    const FrameBox *d_boxes, hipStream_t st) {
    const int n = (int)h->io.size();
    const int n_tp = upload_tile_pairs(h, n, st);
    const size_t lds = (size_t)(h->bins + 1) * 8 + (size_t)((h->bins + 3) & ~3) * 4 + (size_t)kTile * 16;
    WMD_REQUIRE(lds <= h->ctx->lds_per_block, WMD_ERR_UNSUPPORTED, "too many bins for LDS");
    auto launch = [&](auto kern) {
        ProfScope prof(h->ctx, st, "rdf_pair");
        int grid = persistent_grid(h, occupancy(kern, lds), (long long)n_frames * n_tp);
        hipLaunchKernelGGL(kern, dim3(grid), dim3(kBlock), lds, st, d_soa, 3 * (long long)stride,
                           (int)stride, n, (int)n_frames, d_boxes, (const unsigned *)h->d_wts.ptr,
                           (const int2 *)h->d_tiles.ptr, n_tp, 0LL, h->bins,
                           (const double *)h->d_thr64.ptr, (double)h->bins / h->r_max64, h->scale64,
                           h->d_counts.as<unsigned long long>(), 8, h->w_max);
        WMD_HIP_CHECK(hipGetLastError());
    };
    if (h->pbc == WMD_PBC_ORTHO) {
        if (h->weighted) launch(rdf_brute64_kernel<WMD_PBC_ORTHO, true>);
        else launch(rdf_brute64_kernel<WMD_PBC_ORTHO, false>);
    } else {
        if (h->weighted) launch(rdf_brute64_kernel<WMD_PBC_NONE, true>);
        else launch(rdf_brute64_kernel<WMD_PBC_NONE, false>);
    }
    h->last_path = 1;
    h->last_nc[0] = h->last_nc[1] = h->last_nc[2] = 0;
}

// Core: accumulate frames of an IO-selection SoA store already on the device.
void accumulate_device(wmd_rdf *h, const float *d_soa, size_t stride, size_t n_frames,
                       const float *boxes_host, hipStream_t st) {
    const int n = (int)h->io.size();
    if (n_frames == 0 || h->bins == 0) return;
    WMD_REQUIRE(stride >= (size_t)n && stride % 4 == 0, WMD_ERR_INVALID,
                "stride must be a multiple of 4 and >= the IO selection size");
    WMD_REQUIRE(h->pbc == WMD_PBC_NONE || boxes_host, WMD_ERR_INVALID, "boxes required for PBC");
    const BoxSet B = make_boxes(h->pbc, boxes_host, n_frames);
    const FrameBox *d_boxes = upload_boxes(B, h->h_boxes, h->d_boxes, st, &h->ev_boxes);

    if (h->f64) {
        accumulate_f64(h, d_soa, stride, n_frames, d_boxes, st);
        return;
    }
    const bool edge = h->flags & WMD_RDF_EDGE_DIAGNOSTICS;
    const bool table = !h->tab.empty();
    HistParams p = hist_params(h, table);
    const double mw = std::min(B.min_width[0], std::min(B.min_width[1], B.min_width[2]));
    const bool fast_wrap = h->pbc == WMD_PBC_NONE || (double)h->r_max <= 0.5 * mw * (1.0 - 1e-5);

    // ---- choose the enumeration -------------------------------------------
    int k_use = 0, nc[3] = {0, 0, 0};
    if (h->pbc != WMD_PBC_NONE && fast_wrap && !(h->flags & WMD_RDF_FORCE_BRUTE) && n >= 64)
        k_use = choose_cells((double)h->r_max, B.min_width, n, h->cell_k, nc);

    if (k_use == 0) {
        // ---- brute force ----------------------------------------------------
        const int n_tp = upload_tile_pairs(h, n, st);
        const int2 *d_tp = (const int2 *)h->d_tiles.ptr;
        long long fs = 3 * (long long)stride;
        bool tb = table;
        const auto passes = hist_passes(h, p, edge, &tb, [&](int bins, bool t) {
            return lds_hist_size(bins, p.hist_copies, edge, t ? p.tab_n : 0) + (size_t)kTile * 16;
        }, st);
        for (const HistParams &q : passes)
            brute_dispatch(!fast_wrap, h->weighted, edge, tb, h->pbc, h, st, d_soa, fs, (int)stride,
                           n, (int)n_frames, 0LL, d_boxes, (const int2 *)d_tp, n_tp, q);
        h->last_path = 1;
        h->last_nc[0] = h->last_nc[1] = h->last_nc[2] = 0;
        return;
    }

    // ---- cell list ----------------------------------------------------------
    const int ncell = nc[0] * nc[1] * nc[2];
    // every counted pair takes its image from the cells on a grid of >= 2k+1
    // cells per side (r_max is below half the smallest plane spacing here)
    const int shift_ok = (nc[0] >= 2 * k_use + 1 && nc[1] >= 2 * k_use + 1 && nc[2] >= 2 * k_use + 1) ? 1 : 0;
    const float cull2 = cull2_for((double)h->r_max, B);
    // frames per build: a whole number of CUs' worth when the scratch holds
    // more than one (the build runs one workgroup per frame and per CU, so
    // 276 frames take two rounds where 256 take one)
    size_t batch = h->cells.frames_for(n, ncell, n_frames, false);
    const size_t cus = (size_t)std::max(1, h->ctx->cu_count);
    if (batch < n_frames && batch > cus) batch = batch / cus * cus;
    h->cells.reserve(n, ncell, batch, false);
    const unsigned *d_w = h->weighted ? h->d_wts.as<unsigned>() : nullptr;
    bool cell_table = table && (int)h->tab.size() <= kMaxTab2;
    const auto passes = hist_passes(h, p, edge, &cell_table, [&](int bins, bool t) {
        return cell_lds_bytes(bins, edge, t ? p.tab_n : 0);
    }, st);
    for (size_t f0 = 0; f0 < n_frames; f0 += batch) {
        const int nf = (int)std::min(batch, n_frames - f0);
        const FrameBox *bx = d_boxes + f0;
        const float4 *bbox = build_cells(h->ctx, st, h->pbc, d_soa, (int)stride, n, nf, (long long)f0, bx, nc, d_w,
                                         h->cells, false);
        for (const HistParams &q : passes)
            cells_dispatch(h->pbc, h->weighted, edge, cell_table, h, st, h->cells.sorted.as<float4>(),
                           (const unsigned *)h->cells.start.ptr, (const unsigned *)h->cells.flag.ptr, bbox, n, nf,
                           bx, (const int *)nc, k_use, shift_ok, cull2, q);
    }
    h->last_path = 2;
    h->last_nc[0] = nc[0]; h->last_nc[1] = nc[1]; h->last_nc[2] = nc[2];
}

// Host AoS chunk -> device SoA store of the IO atoms -> accumulate, as a
// two-stage pipeline: batch k + 1 is copied into the other pinned staging
// buffer (several host threads) while the GPU runs batch k's H2D copy, gather
// and pair kernels; all device work stays in order on the context's stream,
// and an event per staging buffer tells the host when it may be refilled.
void accumulate_host(wmd_rdf *h, const float *coords, size_t n_src, size_t n_frames,
                     const float *boxes, bool selected) {
    if (n_frames == 0 || h->bins == 0) return;
    WMD_REQUIRE(coords != nullptr, WMD_ERR_INVALID, "null coords");
    WMD_REQUIRE(!(selected && h->grouped), WMD_ERR_UNSUPPORTED,
                "grouped RDF modes take full chunks (wmd_rdf_accumulate)");
    const int n = (int)h->io.size();
    if (h->grouped)
        WMD_REQUIRE(h->g_max_atom <= n_src, WMD_ERR_MISMATCH,
                    "group atom index out of range for " + std::to_string(n_src) + " atoms");
    if (!selected && !h->grouped) {
        for (uint32_t a : h->io)
            WMD_REQUIRE(a < n_src, WMD_ERR_MISMATCH,
                        "selection index " + std::to_string(a) + " out of range for " +
                            std::to_string(n_src) + " atoms");
    }
    // every frame's box is checked before anything is enqueued, so that a bad
    // box in a later batch leaves the counts untouched (the reference checks
    // the whole chunk's boxes before its pair loop, rdf.rs:746-788)
    if (h->pbc != WMD_PBC_NONE) {
        WMD_REQUIRE(boxes != nullptr, WMD_ERR_INVALID, "boxes required for PBC");
        (void)make_boxes(h->pbc, boxes, n_frames);
    }
    hipStream_t st = h->ctx->stream;
    hipStream_t cs = copy_stream(h->ctx);
    const size_t stride = round_up(std::max(n, 1), 4);
    // frames per batch: one staging buffer of AoS (48 MB unless the context
    // sets another size)
    const size_t fbytes = n_src * 16;
    const size_t batch = staging_frames(h->ctx, fbytes, n_frames);
    for (int b = 0; b < 2; ++b) {
        if (h->p_stage[b].bytes < batch * fbytes || h->p_soa[b].bytes < batch * 3 * stride * 4) {
            WMD_HIP_CHECK(hipStreamSynchronize(st));  // growing frees buffers queued work may use
            WMD_HIP_CHECK(hipStreamSynchronize(cs));
            h->p_stage[b].reserve(batch * fbytes);
            h->p_aos[b].reserve(batch * fbytes);
            h->p_soa[b].reserve(batch * 3 * stride * 4);
        }
    }
    // Three overlapping stages: host copies into pinned staging buffer b, the
    // H2D copy of b on the copy stream, and the gather and kernels on the
    // context's stream.  p_copied[b]: the H2D of b is done (the host may refill
    // the staging buffer, the kernels may read the device copy); p_used[b]: the
    // gather has read the device copy (the next H2D into it may start).
    WMD_HIP_CHECK(hipEventRecord(h->p_used[0].get(), st));
    WMD_HIP_CHECK(hipEventRecord(h->p_used[1].get(), st));
    size_t k = 0;
    for (size_t f0 = 0; f0 < n_frames; f0 += batch, ++k) {
        const int b = (int)(k & 1);
        const size_t nf = std::min(batch, n_frames - f0);
        h->p_copied[b].wait();  // the copy out of this staging buffer two batches ago is done
        host_copy(h->p_stage[b].ptr, coords + f0 * n_src * 4, nf * fbytes);
        WMD_HIP_CHECK(hipStreamWaitEvent(cs, h->p_used[b].get(), 0));
        WMD_HIP_CHECK(hipMemcpyAsync(h->p_aos[b].ptr, h->p_stage[b].ptr, nf * fbytes,
                                     hipMemcpyHostToDevice, cs));
        WMD_HIP_CHECK(hipEventRecord(h->p_copied[b].get(), cs));
        WMD_HIP_CHECK(hipStreamWaitEvent(st, h->p_copied[b].get(), 0));
        float *d_soa = (float *)h->p_soa[b].ptr;
        int g;
        grid1d((long long)stride * nf, &g);
        {
        ProfScope prof(h->ctx, st, "rdf_gather");
        if (h->grouped) {
            hipLaunchKernelGGL(group_center_kernel, dim3(g), dim3(kBlock), 0, st,
                               (const float4 *)h->p_aos[b].ptr, (long long)n_src,
                               h->d_goff.as<unsigned long long>(), h->d_gatoms.as<unsigned>(),
                               h->d_gmass.as<float>(), n, (int)stride, (int)nf, (int)h->planar, d_soa);
        } else {
            hipLaunchKernelGGL(gather_soa_kernel, dim3(g), dim3(kBlock), 0, st, (const float4 *)h->p_aos[b].ptr,
                               (long long)n_src, selected ? nullptr : h->d_io.as<unsigned>(), n,
                               (int)stride, (int)nf, d_soa, (int)h->planar);
        }
        WMD_HIP_CHECK(hipGetLastError());
        }
        WMD_HIP_CHECK(hipEventRecord(h->p_used[b].get(), st));
        accumulate_device(h, d_soa, stride, nf, boxes ? boxes + 9 * f0 : nullptr, st);
        if (h->n_excl > 0) {
            grid1d(h->n_excl * (long long)nf, &g);
            auto kern = h->pbc == WMD_PBC_ORTHO ? rdf_excl_kernel<WMD_PBC_ORTHO> : rdf_excl_kernel<WMD_PBC_NONE>;
            hipLaunchKernelGGL(kern, dim3(g), dim3(kBlock), 0, st, (const float *)d_soa, (int)stride, (int)nf,
                               h->d_boxes.as<FrameBox>(), h->d_excl.as<uint4>(), h->n_excl,
                               h->d_thr.as<float>(), h->bins, h->d_counts.as<unsigned long long>());
            WMD_HIP_CHECK(hipGetLastError());
        }
    }
    WMD_HIP_CHECK(hipStreamSynchronize(st));
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
    WMD_REQUIRE(bins <= (size_t(1) << 22), WMD_ERR_UNSUPPORTED, "at most 4194304 bins are supported");
    WMD_REQUIRE(std::isfinite(r_max) && r_max > 0.0f, WMD_ERR_PARSE, "r_max must be finite and positive");
    bind(ctx);
    auto *h = new wmd_rdf();
    try {
        h->ctx = ctx;
        h->bins = (int)bins;
        h->r_max = r_max;
        h->pbc = pbc;
        h->flags = flags;
        h->planar = (flags & WMD_RDF_PLANAR_XY) != 0;
        WMD_REQUIRE(!h->planar || pbc != WMD_PBC_TRICLINIC, WMD_ERR_UNSUPPORTED,
                    "the xy dimension supports pbc none or orthorhombic");
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
            WMD_REQUIRE(ma[i] <= 15 && mb[i] <= 15, WMD_ERR_UNSUPPORTED,
                        "an atom repeated more than 15 times in a selection");
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
            BinTables T = make_bin_tables((int)bins, r_max);
            h->bin_width = T.w;
            const float w = h->bin_width;
            const int B = (int)bins;
            h->thr.swap(T.thr);
            h->cut = T.cut;
            h->tab.swap(T.tab);
            h->tab2.swap(T.tab2);
            h->tab_sh = T.tab_sh;
            h->tab_base = T.tab_base;
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
            hipStream_t st = ctx->stream;
            // counts [bins], edge pairs, pair evaluations
            h->d_counts.reserve((bins + 2) * 8);
            hipLaunchKernelGGL(fill_u64_kernel, dim3((bins + 2 + kBlock - 1) / kBlock), dim3(kBlock), 0, st,
                               h->d_counts.as<unsigned long long>(), (long long)bins + 2);
            WMD_HIP_CHECK(hipGetLastError());
            h->d_thr.reserve((bins + 1) * 4);
            h->d_elo.reserve((bins + 1) * 4);
            h->d_ehi.reserve((bins + 1) * 4);
            WMD_HIP_CHECK(hipMemcpyAsync(h->d_thr.ptr, h->thr.data(), (bins + 1) * 4, hipMemcpyHostToDevice, st));
            WMD_HIP_CHECK(hipMemcpyAsync(h->d_elo.ptr, h->elo.data(), (bins + 1) * 4, hipMemcpyHostToDevice, st));
            WMD_HIP_CHECK(hipMemcpyAsync(h->d_ehi.ptr, h->ehi.data(), (bins + 1) * 4, hipMemcpyHostToDevice, st));
            if (!h->tab.empty()) {
                h->d_tab.reserve(h->tab.size() * 2);
                WMD_HIP_CHECK(hipMemcpyAsync(h->d_tab.ptr, h->tab.data(), h->tab.size() * 2, hipMemcpyHostToDevice, st));
                h->d_tab2.reserve(h->tab2.size() * 8);
                WMD_HIP_CHECK(hipMemcpyAsync(h->d_tab2.ptr, h->tab2.data(), h->tab2.size() * 8, hipMemcpyHostToDevice, st));
            }
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

wmd_status wmd_rdf_create_f64(wmd_ctx *ctx, const uint32_t *sel, size_t n_sel, size_t bins,
                              double r_max, int pbc, double length_scale, wmd_rdf **out) {
    WMD_API_BEGIN
    WMD_REQUIRE(ctx && out, WMD_ERR_INVALID, "null context or output handle");
    *out = nullptr;
    WMD_REQUIRE(pbc == WMD_PBC_NONE || pbc == WMD_PBC_ORTHO, WMD_ERR_MISMATCH,
                "structure factor requires orthorhombic box for PBC");
    WMD_REQUIRE(bins <= 4096, WMD_ERR_UNSUPPORTED, "at most 4096 bins in f64 mode");
    WMD_REQUIRE(std::isfinite(r_max) && r_max > 0.0, WMD_ERR_PARSE, "r_max must be finite and positive");
    WMD_REQUIRE(std::isfinite(length_scale) && length_scale > 0.0, WMD_ERR_PARSE,
                "length_scale must be finite and positive");
    wmd_rdf *h = nullptr;
    const wmd_status st = wmd_rdf_create(ctx, sel, n_sel, sel, n_sel, bins, (float)r_max, pbc, 0, &h);
    if (st != WMD_OK) return st;
    try {
        h->f64 = true;
        h->r_max64 = r_max;
        h->scale64 = length_scale;
        if (bins > 0) {
            // thr64[k] = min{ f64 s >= 0 : bin(s) >= k }, bin(s) = the reference's
            // r = sqrt(s) (f64), r < r_max, (r / (r_max / bins)) as usize
            const double w = r_max / (double)bins;
            auto key = [&](double s) -> size_t {
                volatile double r = std::sqrt(s);
                if (!(r < r_max)) return bins;
                volatile double q = r / w;
                const size_t b = (size_t)q;
                return b < bins ? b : bins;
            };
            h->thr64.assign(bins + 1, 0.0);
            for (size_t k = 1; k <= bins; ++k) {
                uint64_t lo = 0, hi = 0x7FF0000000000000ull;  // key(+inf) = bins >= k
                double lo_s;
                std::memcpy(&lo_s, &lo, 8);
                if (key(lo_s) >= k) { h->thr64[k] = 0.0; continue; }
                while (hi - lo > 1) {
                    const uint64_t mid = lo + (hi - lo) / 2;
                    double ms;
                    std::memcpy(&ms, &mid, 8);
                    if (key(ms) >= k) hi = mid; else lo = mid;
                }
                std::memcpy(&h->thr64[k], &hi, 8);
            }
            h->d_thr64.reserve((bins + 1) * 8);
            WMD_HIP_CHECK(hipMemcpy(h->d_thr64.ptr, h->thr64.data(), (bins + 1) * 8, hipMemcpyHostToDevice));
        }
    } catch (...) {
        delete h;
        throw;
    }
    *out = h;
    WMD_API_END
}

wmd_status wmd_rdf_create_groups(wmd_ctx *ctx, size_t n_groups, const uint64_t *group_offsets,
                                 const uint32_t *group_atoms, const float *group_masses,
                                 const uint32_t *groups_a, size_t n_a, const uint32_t *groups_b,
                                 size_t n_b, const int64_t *mol_key, const int64_t *res_key,
                                 size_t bins, float r_max, int pbc, uint32_t flags, wmd_rdf **out) {
    WMD_API_BEGIN
    WMD_REQUIRE(ctx && out, WMD_ERR_INVALID, "null context or output handle");
    *out = nullptr;
    WMD_REQUIRE(group_offsets && (n_groups == 0 || group_offsets[n_groups] == 0 || group_atoms),
                WMD_ERR_INVALID, "null group layout");
    WMD_REQUIRE(pbc == WMD_PBC_NONE || pbc == WMD_PBC_ORTHO, WMD_ERR_UNSUPPORTED,
                "grouped RDF modes support pbc none or orthorhombic");
    WMD_REQUIRE(!(flags & WMD_RDF_EDGE_DIAGNOSTICS), WMD_ERR_UNSUPPORTED,
                "edge diagnostics are not available in grouped modes");
    for (size_t g = 0; g < n_groups; ++g)
        WMD_REQUIRE(group_offsets[g] <= group_offsets[g + 1], WMD_ERR_INVALID, "group offsets must not decrease");
    for (size_t i = 0; i < n_a; ++i) WMD_REQUIRE(groups_a[i] < n_groups, WMD_ERR_INVALID, "group id out of range");
    for (size_t i = 0; i < n_b; ++i) WMD_REQUIRE(groups_b[i] < n_groups, WMD_ERR_INVALID, "group id out of range");
    const bool no_intramol = (flags & WMD_RDF_NO_INTRAMOL) != 0;
    WMD_REQUIRE(!no_intramol || (mol_key && res_key), WMD_ERR_INVALID, "no_intramol needs group keys");
    wmd_rdf *h = nullptr;
    const wmd_status st0 = wmd_rdf_create(ctx, groups_a, n_a, groups_b, n_b, bins, r_max, pbc,
                                          flags & ~WMD_RDF_NO_INTRAMOL, &h);
    if (st0 != WMD_OK) return st0;
    try {
        h->grouped = true;
        // the IO "atoms" are the groups in IO order: their atom lists as CSR
        const size_t nio = h->io.size();
        std::vector<unsigned long long> off(nio + 1, 0);
        std::vector<uint32_t> atoms;
        std::vector<float> masses;
        for (size_t i = 0; i < nio; ++i) {
            const uint32_t g = h->io[i];
            for (uint64_t k = group_offsets[g]; k < group_offsets[g + 1]; ++k) {
                atoms.push_back(group_atoms[k]);
                if (group_masses) masses.push_back(group_masses[k]);
            }
            off[i + 1] = atoms.size();
        }
        h->d_goff.reserve(off.size() * 8);
        WMD_HIP_CHECK(hipMemcpy(h->d_goff.ptr, off.data(), off.size() * 8, hipMemcpyHostToDevice));
        h->d_gatoms.reserve(std::max<size_t>(atoms.size(), 1) * 4);
        if (!atoms.empty())
            WMD_HIP_CHECK(hipMemcpy(h->d_gatoms.ptr, atoms.data(), atoms.size() * 4, hipMemcpyHostToDevice));
        if (group_masses) {
            h->d_gmass.reserve(std::max<size_t>(masses.size(), 1) * 4);
            if (!masses.empty())
                WMD_HIP_CHECK(hipMemcpy(h->d_gmass.ptr, masses.data(), masses.size() * 4, hipMemcpyHostToDevice));
        }
        h->g_max_atom = 0;
        for (uint32_t a : atoms) h->g_max_atom = std::max<uint64_t>(h->g_max_atom, (uint64_t)a + 1);
        if (no_intramol) {
            // same_intramolecular_group (rdf.rs:368-373) for u in its sel_a role and
            // v in its sel_b role (keys [g] and [n_groups + g]): molecule keys equal
            // when both have one, else residue keys equal when both have one
            const int64_t kNone = INT64_MIN;
            auto key = [&](const int64_t *k, uint32_t io_pos, int role) { return k[(size_t)role * n_groups + h->io[io_pos]]; };
            std::unordered_map<int64_t, std::vector<uint32_t>> mol_a, mol_b, res_a, res_b;
            for (uint32_t i = 0; i < (uint32_t)nio; ++i) {
                if (h->wts[i] & 0xFFu) {
                    if (key(mol_key, i, 0) != kNone) mol_a[key(mol_key, i, 0)].push_back(i);
                    if (key(res_key, i, 0) != kNone) res_a[key(res_key, i, 0)].push_back(i);
                }
                if (h->wts[i] >> 8) {
                    if (key(mol_key, i, 1) != kNone) mol_b[key(mol_key, i, 1)].push_back(i);
                    if (key(res_key, i, 1) != kNone) res_b[key(res_key, i, 1)].push_back(i);
                }
            }
            std::vector<uint4> ex;
            auto emit = [&](uint32_t u, uint32_t v) {
                const unsigned w = (h->wts[u] & 0xFFu) * (h->wts[v] >> 8);
                if (u != v && w) ex.push_back(make_uint4(u, v, w, 0u));
                WMD_REQUIRE(ex.size() <= (1ull << 27), WMD_ERR_UNSUPPORTED,
                            "too many intramolecular pairs to exclude");
            };
            for (auto &kv : mol_a) {
                auto it = mol_b.find(kv.first);
                if (it == mol_b.end()) continue;
                for (uint32_t u : kv.second)
                    for (uint32_t v : it->second) emit(u, v);
            }
            for (auto &kv : res_a) {
                auto it = res_b.find(kv.first);
                if (it == res_b.end()) continue;
                for (uint32_t u : kv.second)
                    for (uint32_t v : it->second)
                        if (!(key(mol_key, u, 0) != kNone && key(mol_key, v, 1) != kNone)) emit(u, v);
            }
            h->n_excl = (long long)ex.size();
            if (!ex.empty()) {
                h->d_excl.reserve(ex.size() * sizeof(uint4));
                WMD_HIP_CHECK(hipMemcpy(h->d_excl.ptr, ex.data(), ex.size() * sizeof(uint4), hipMemcpyHostToDevice));
            }
        }
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
    WMD_REQUIRE(!h->grouped, WMD_ERR_UNSUPPORTED, "grouped RDF modes take full chunks (wmd_rdf_accumulate)");
    bind(h->ctx);
    hipStream_t st = pick_stream(h->ctx, stream);
    if (h->planar && n_frames > 0) {
        // PlanarXY: the store's x and y planes with z = 0
        WMD_REQUIRE(stride % 4 == 0, WMD_ERR_INVALID, "stride must be a multiple of 4");
        const size_t fl = 3 * stride;
        if (h->d_pstore.bytes < n_frames * fl * 4) WMD_HIP_CHECK(hipStreamSynchronize(st));
        float *dst = (float *)h->d_pstore.reserve(n_frames * fl * 4);
        int g;
        grid1d((long long)n_frames * stride, &g);
        hipLaunchKernelGGL(planar_copy_kernel, dim3(g), dim3(kBlock), 0, st, d_soa, dst, (long long)stride,
                           (long long)n_frames);
        WMD_HIP_CHECK(hipGetLastError());
        d_soa = dst;
    }
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

wmd_status wmd_rdf_pair_evaluations(wmd_rdf *h, uint64_t *evaluations) {
    WMD_API_BEGIN
    WMD_REQUIRE(h && evaluations, WMD_ERR_INVALID, "null handle or output");
    bind(h->ctx);
    *evaluations = 0;
    if (h->bins == 0) return WMD_OK;
    WMD_HIP_CHECK(hipDeviceSynchronize());
    WMD_HIP_CHECK(hipMemcpy(evaluations, h->d_counts.as<unsigned long long>() + h->bins + 1, 8,
                            hipMemcpyDeviceToHost));
    WMD_API_END
}

wmd_status wmd_rdf_reset(wmd_rdf *h) {
    WMD_API_BEGIN
    WMD_REQUIRE(h != nullptr, WMD_ERR_INVALID, "null handle");
    bind(h->ctx);
    if (h->bins) {
        int g = (h->bins + 2 + kBlock - 1) / kBlock;
        hipLaunchKernelGGL(fill_u64_kernel, dim3(g), dim3(kBlock), 0, h->ctx->stream,
                           h->d_counts.as<unsigned long long>(), (long long)h->bins + 2);
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

wmd_status wmd_rdf_set_tuning(wmd_rdf *h, int cell_k, int grab) {
    WMD_API_BEGIN
    WMD_REQUIRE(h != nullptr, WMD_ERR_INVALID, "null handle");
    WMD_REQUIRE(cell_k >= 0 && cell_k <= 3, WMD_ERR_INVALID, "cell_k must be 0..3");
    WMD_REQUIRE(grab >= 0 && grab <= 64, WMD_ERR_INVALID, "grab must be 0..64");
    h->cell_k = cell_k;
    h->grab = grab;
    WMD_API_END
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Gaussian-smoothed RDF handle
// ---------------------------------------------------------------------------
struct wmd_softrdf {
    wmd_ctx *ctx = nullptr;
    int bins = 0;
    float r_max = 0.f, sigma = 0.f, w = 0.f;
    int pbc = 0;
    int cell_k = 0;
    int last_path = 0;
    int last_nc[3] = {0, 0, 0};
    DevBuf d_boxes, d_table, d_nodes, d_gsorted, d_queue;
    DevBuf d_wt;                       // bin-node weights wt[2 wt_j + 1] (f64)
    int wt_j = 0;
    HostBuf h_boxes;
    Event ev_boxes;                    // the last box upload (guards h_boxes)
    std::vector<FrameBox> boxes_up;    // what d_boxes holds
    CellScratch cells;
    // cell lists of the last forward, kept for the backward (keep_cells):
    // valid when they cover every frame of that call
    bool keep_cells = false, cells_valid = false;
    const float *cells_soa = nullptr;
    size_t cells_stride = 0, cells_n = 0, cells_frames = 0;
    int cells_nc[3] = {0, 0, 0};
};

namespace {

double soft_h(const wmd_softrdf *h) { return (double)h->w / kSoftNodesPerBin; }

// Pair reach: a pair deposits onto nodes within 1.5 h of r; the last bin uses
// nodes up to r_last + 6 sigma.
double soft_reach(const wmd_softrdf *h) {
    return ((double)h->bins - 0.5) * (double)h->w + 6.0 * (double)h->sigma + 2.0 * soft_h(h);
}

SoftParams soft_params(const wmd_softrdf *h) {
    SoftParams sp{};
    const double hn = soft_h(h);
    sp.n_nodes = (int)std::ceil((((double)h->bins - 0.5) * (double)h->w + 6.0 * (double)h->sigma) / hn) + 2;
    sp.inv_h = (float)(1.0 / hn);
    const float reach = (float)soft_reach(h);
    sp.cut2 = reach * reach;
    return sp;
}

struct SoftPlanOut {
    int k = 0;
    int nc[3] = {0, 0, 0};
    bool lt = false;
};

// Cell list when the pair reach is below half the smallest plane spacing.
SoftPlanOut soft_plan(wmd_softrdf *h, const BoxSet &B, int n) {
    SoftPlanOut o;
    if (h->pbc == WMD_PBC_NONE || h->cell_k < 0) return o;
    const double mw = std::min(B.min_width[0], std::min(B.min_width[1], B.min_width[2]));
    if (soft_reach(h) > 0.5 * mw * (1.0 - 1e-5) || n < 64) return o;
    o.k = choose_cells(soft_reach(h), B.min_width, n, h->cell_k, o.nc);
    o.lt = h->pbc == WMD_PBC_TRICLINIC && B.lower_tri;
    return o;
}

template <class K>
void launch_soft_cells(wmd_softrdf *h, hipStream_t st, K kern, int block, size_t lds, const CellGrid &G,
                       const SoftParams &sp, const char *name) {
    ProfScope prof(h->ctx, st, name);
    WMD_REQUIRE(lds <= h->ctx->lds_per_block, WMD_ERR_UNSUPPORTED, "too many bins for LDS");
    int occ = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, kern, block, lds) != hipSuccess || occ < 1)
        occ = 1;
    const long long n_units = (long long)G.n_frames * G.ncx * G.ncy * G.ncz;
    const int waves = block / 64;
    int grid = std::max(8, h->ctx->cu_count * occ / 8 * 8);
    const long long max_grid = (n_units + waves * 8 - 1) / (waves * 8) * 8;
    if (grid > max_grid) grid = (int)std::max<long long>(8, max_grid);
    hipLaunchKernelGGL(kern, dim3(grid), dim3(block), lds, st, G, sp);
    WMD_HIP_CHECK(hipGetLastError());
}

void soft_run(wmd_softrdf *h, bool backward, const float *d_soa, size_t stride, size_t n_atoms,
              size_t n_frames, const float *boxes_host, double *d_counts, const double *d_dc,
              float *d_grad, hipStream_t st) {
    if (n_frames == 0 || n_atoms < 2) return;
    WMD_REQUIRE(d_soa != nullptr, WMD_ERR_INVALID, "null device store");
    WMD_REQUIRE(stride >= n_atoms && stride % 4 == 0, WMD_ERR_INVALID,
                "stride must be a multiple of 4 and >= the atom count");
    WMD_REQUIRE(n_atoms < (1u << 20), WMD_ERR_UNSUPPORTED, "at most 2^20 atoms per frame");
    WMD_REQUIRE(h->pbc == WMD_PBC_NONE || boxes_host, WMD_ERR_INVALID, "boxes required for PBC");
    const int n = (int)n_atoms;
    const BoxSet B = make_boxes(h->pbc, boxes_host, n_frames);
    // boxes: uploaded when they differ from what the device copy holds (the
    // backward usually repeats the forward's), without a stream sync
    const bool same_boxes = h->d_boxes.ptr && h->boxes_up.size() == B.fb.size() &&
                            std::memcmp(h->boxes_up.data(), B.fb.data(), B.fb.size() * sizeof(FrameBox)) == 0;
    const FrameBox *d_boxes = same_boxes ? (const FrameBox *)h->d_boxes.ptr
                                         : upload_boxes(B, h->h_boxes, h->d_boxes, st, &h->ev_boxes);
    if (!same_boxes) {
        h->boxes_up = B.fb;
        h->cells_valid = false;
    }
    SoftParams sp = soft_params(h);
    sp.grad = d_grad;
    sp.stride = (int)stride;
    sp.n = n;
    const int tn = sp.n_nodes + kSoftPadL + kSoftPadR;
    if (backward) {
        float2 *tab = (float2 *)h->d_table.reserve((size_t)tn * 8);
        hipLaunchKernelGGL(soft_table_kernel, dim3((tn + 255) / 256), dim3(256), 0, st, d_dc,
                           h->bins, (const double *)h->d_wt.ptr, h->wt_j, 1.0 / soft_h(h), sp.n_nodes, tab);
        WMD_HIP_CHECK(hipGetLastError());
        sp.table = tab;
    } else {
        sp.nodes = (double *)h->d_nodes.reserve((size_t)sp.n_nodes * 8);
        hipLaunchKernelGGL(fill_u64_kernel, dim3((sp.n_nodes + kBlock - 1) / kBlock), dim3(kBlock), 0, st,
                           (unsigned long long *)sp.nodes, (long long)sp.n_nodes);
        WMD_HIP_CHECK(hipGetLastError());
    }
    // node sums -> bins (forward), after the pair kernels
    auto conv = [&]() {
        if (backward) return;
        hipLaunchKernelGGL(soft_conv_kernel, dim3((h->bins + 255) / 256), dim3(256), 0, st,
                           (const double *)sp.nodes, sp.n_nodes, h->bins, (const double *)h->d_wt.ptr,
                           h->wt_j, d_counts);
        WMD_HIP_CHECK(hipGetLastError());
    };
    const SoftPlanOut pl = soft_plan(h, B, n);
    if (pl.k == 0) {
        // brute force: one thread per atom, frames along y
        for (size_t f0 = 0; f0 < n_frames; f0 += 65535) {
            const int nf = (int)std::min<size_t>(65535, n_frames - f0);
            dim3 grid((n + kBlock - 1) / kBlock, nf);
            const float *soa = d_soa + f0 * 3 * stride;
            const FrameBox *bx = d_boxes + f0;
            sp.grad_frame0 = (long long)f0;
            ProfScope prof(h->ctx, st, backward ? "softrdf_bwd" : "softrdf_fwd");
            if (!backward) {
                const size_t lds = (size_t)tn * 8;
                WMD_REQUIRE(lds <= h->ctx->lds_per_block, WMD_ERR_UNSUPPORTED, "too many bins for LDS");
                if (h->pbc == WMD_PBC_NONE)
                    hipLaunchKernelGGL(soft_brute_fwd_kernel<WMD_PBC_NONE>, grid, dim3(kBlock), lds, st, soa, (int)stride, n, bx, sp);
                else if (h->pbc == WMD_PBC_ORTHO)
                    hipLaunchKernelGGL(soft_brute_fwd_kernel<WMD_PBC_ORTHO>, grid, dim3(kBlock), lds, st, soa, (int)stride, n, bx, sp);
                else
                    hipLaunchKernelGGL(soft_brute_fwd_kernel<WMD_PBC_TRICLINIC>, grid, dim3(kBlock), lds, st, soa, (int)stride, n, bx, sp);
            } else {
                sp.grad_frame0 = 0;
                SoftParams spb = sp;
                spb.grad = d_grad + f0 * 3 * stride;
                const size_t lds = (size_t)tn * 8;
                WMD_REQUIRE(lds <= h->ctx->lds_per_block, WMD_ERR_UNSUPPORTED, "too many bins for LDS");
                if (h->pbc == WMD_PBC_NONE)
                    hipLaunchKernelGGL(soft_brute_bwd_kernel<WMD_PBC_NONE>, grid, dim3(kBlock), lds, st, soa, (int)stride, n, bx, spb);
                else if (h->pbc == WMD_PBC_ORTHO)
                    hipLaunchKernelGGL(soft_brute_bwd_kernel<WMD_PBC_ORTHO>, grid, dim3(kBlock), lds, st, soa, (int)stride, n, bx, spb);
                else
                    hipLaunchKernelGGL(soft_brute_bwd_kernel<WMD_PBC_TRICLINIC>, grid, dim3(kBlock), lds, st, soa, (int)stride, n, bx, spb);
            }
            WMD_HIP_CHECK(hipGetLastError());
        }
        conv();
        h->cells_valid = false;
        h->last_path = 1;
        h->last_nc[0] = h->last_nc[1] = h->last_nc[2] = 0;
        return;
    }
    const int ncell = pl.nc[0] * pl.nc[1] * pl.nc[2];
    const int shift_ok = (pl.nc[0] >= 2 * pl.k + 1 && pl.nc[1] >= 2 * pl.k + 1 && pl.nc[2] >= 2 * pl.k + 1) ? 1 : 0;
    const float cull2 = cull2_for(soft_reach(h), B);
    // with keep_cells the forward also writes the slot map, and a backward
    // on the same store, sizes and boxes takes the forward's cell lists
    const bool want_order = backward || h->keep_cells;
    const size_t batch = h->cells.frames_for(n, ncell, n_frames, want_order, true);
    const bool reuse = backward && h->keep_cells && h->cells_valid && h->cells_soa == d_soa &&
                       h->cells_stride == stride && h->cells_n == n_atoms && h->cells_frames == n_frames &&
                       h->cells_nc[0] == pl.nc[0] && h->cells_nc[1] == pl.nc[1] && h->cells_nc[2] == pl.nc[2];
    if (!reuse) {
        h->cells_valid = false;
        h->cells.reserve(n, ncell, batch, want_order, true);
    }
    for (size_t f0 = 0; f0 < n_frames; f0 += batch) {
        const int nf = (int)std::min(batch, n_frames - f0);
        const FrameBox *bx = d_boxes + f0;
        if (!reuse)
            build_cells(h->ctx, st, h->pbc, d_soa, (int)stride, n, nf, (long long)f0, bx, pl.nc, nullptr,
                        h->cells, want_order, true);
        // the per-cell boxes of this batch's (or the reused) build, when its path writes them
        const float4 *bbox = ncell <= kBuildMaxCells ? h->cells.bbox.as<float4>() : nullptr;
        // dynamic unit schedule: the per-XCD grab counters start at zero; the
        // forward grabs 2 units at a time (20,000 beads x 64 frames: 0.88 ms
        // against 0.96 with 1 and 0.98 with 4), the heavier backward units 1
        unsigned long long *queue = (unsigned long long *)h->d_queue.reserve(8 * 8);
        hipLaunchKernelGGL(fill_u64_kernel, dim3(1), dim3(kBlock), 0, st, queue, 8LL);
        WMD_HIP_CHECK(hipGetLastError());
        CellGrid G{h->cells.sorted.as<float4>(), (const unsigned *)h->cells.start.ptr,
                   (const unsigned *)h->cells.flag.ptr, bx, n, nf, pl.nc[0], pl.nc[1], pl.nc[2], pl.k,
                   shift_ok, 8, cull2, queue, backward ? 1 : 2, bbox};
        if (backward) {
            // per-batch gradient in cell order, scattered to the store afterwards
            sp.grad = (float *)h->d_gsorted.reserve((size_t)3 * n * nf * 4);
            hipLaunchKernelGGL(fill_u32_kernel, dim3(std::min<long long>(65536, ((long long)3 * n * nf + kBlock - 1) / kBlock)),
                               dim3(kBlock), 0, st, (unsigned *)sp.grad, (long long)3 * n * nf, 0u);
            WMD_HIP_CHECK(hipGetLastError());
        }
        if (!backward) {
            const size_t lds = align16((size_t)tn * 8) + (size_t)kSoftWaves * 128 * 4 +
                               (size_t)kSoftWaves * kRingStride * 16;
            if (h->pbc == WMD_PBC_ORTHO)
                launch_soft_cells(h, st, soft_cell_fwd_kernel<WMD_PBC_ORTHO, false>, kSoftBlock, lds, G, sp, "softrdf_fwd");
            else if (pl.lt)
                launch_soft_cells(h, st, soft_cell_fwd_kernel<WMD_PBC_TRICLINIC, true>, kSoftBlock, lds, G, sp, "softrdf_fwd");
            else
                launch_soft_cells(h, st, soft_cell_fwd_kernel<WMD_PBC_TRICLINIC, false>, kSoftBlock, lds, G, sp, "softrdf_fwd");
        } else {
            // home slots in LDS: 32 for cells of up to ~24 atoms on average, else 128
            const double per_cell = (double)n / ((double)pl.nc[0] * pl.nc[1] * pl.nc[2]);
            auto bwd = [&](auto k32, auto k128) {
                if (per_cell <= 24.0)
                    launch_soft_cells(h, st, k32, kSoftBwdBlock, soft_bwd_lds_bytes(tn, 32), G, sp, "softrdf_bwd");
                else
                    launch_soft_cells(h, st, k128, kSoftBwdBlock, soft_bwd_lds_bytes(tn, 128), G, sp, "softrdf_bwd");
            };
            if (h->pbc == WMD_PBC_ORTHO)
                bwd(soft_cell_bwd_kernel<WMD_PBC_ORTHO, false, 32>, soft_cell_bwd_kernel<WMD_PBC_ORTHO, false, 128>);
            else if (pl.lt)
                bwd(soft_cell_bwd_kernel<WMD_PBC_TRICLINIC, true, 32>, soft_cell_bwd_kernel<WMD_PBC_TRICLINIC, true, 128>);
            else
                bwd(soft_cell_bwd_kernel<WMD_PBC_TRICLINIC, false, 32>, soft_cell_bwd_kernel<WMD_PBC_TRICLINIC, false, 128>);
            int g;
            grid1d((long long)n * nf, &g);
            hipLaunchKernelGGL(soft_unsort_kernel, dim3(g), dim3(kBlock), 0, st, (const float *)sp.grad,
                               (const unsigned *)h->cells.order.ptr, n, nf, (int)stride,
                               d_grad + f0 * 3 * stride);
            WMD_HIP_CHECK(hipGetLastError());
        }
    }
    conv();
    if (h->keep_cells && batch >= n_frames) {
        // the cell lists now hold every frame of this call
        h->cells_valid = true;
        h->cells_soa = d_soa;
        h->cells_stride = stride;
        h->cells_n = n_atoms;
        h->cells_frames = n_frames;
        for (int d = 0; d < 3; ++d) h->cells_nc[d] = pl.nc[d];
    }
    h->last_path = 2;
    h->last_nc[0] = pl.nc[0]; h->last_nc[1] = pl.nc[1]; h->last_nc[2] = pl.nc[2];
}

}  // namespace

extern "C" {

wmd_status wmd_softrdf_create(wmd_ctx *ctx, size_t bins, float r_max, float sigma, int pbc,
                              wmd_softrdf **out) {
    WMD_API_BEGIN
    WMD_REQUIRE(ctx && out, WMD_ERR_INVALID, "null argument");
    *out = nullptr;
    WMD_REQUIRE(bins >= 1 && bins <= 8192, WMD_ERR_INVALID, "bins must be in 1..8192");
    WMD_REQUIRE(std::isfinite(r_max) && r_max > 0.f, WMD_ERR_INVALID, "r_max must be positive");
    WMD_REQUIRE(pbc >= WMD_PBC_NONE && pbc <= WMD_PBC_TRICLINIC, WMD_ERR_INVALID, "unknown pbc kind");
    const float w = r_max / (float)bins;
    WMD_REQUIRE(std::isfinite(sigma) && sigma >= 0.25f * w && sigma <= 4.0f * w, WMD_ERR_UNSUPPORTED,
                "sigma must lie in [1/4, 4] bin widths");
    bind(ctx);
    wmd_softrdf *h = new wmd_softrdf;
    h->ctx = ctx;
    h->bins = (int)bins;
    h->r_max = r_max;
    h->sigma = sigma;
    h->w = w;
    h->pbc = pbc;
    {
        // wt[J + j] = w phi_sigma(j h), |j h| <= 6 sigma (J rounded so that an
        // exact multiple of h at 6 sigma is kept)
        const double hh = soft_h(h), sg = (double)sigma;
        const int J = (int)std::floor(6.0 * sg / hh * (1.0 + 1e-12));
        std::vector<double> wt(2 * J + 1);
        const double norm = (double)w / (sg * 2.5066282746310002);
        for (int j = -J; j <= J; ++j) {
            const double d = j * hh;
            wt[J + j] = norm * std::exp(-0.5 * d * d / (sg * sg));
        }
        h->wt_j = J;
        WMD_HIP_CHECK(hipMemcpy(h->d_wt.reserve(wt.size() * 8), wt.data(), wt.size() * 8,
                                hipMemcpyHostToDevice));
    }
    *out = h;
    WMD_API_END
}

wmd_status wmd_softrdf_forward(wmd_softrdf *h, const float *d_soa, size_t stride, size_t n_atoms,
                               size_t n_frames, const float *boxes, double *d_counts, void *stream) {
    WMD_API_BEGIN
    WMD_REQUIRE(h != nullptr, WMD_ERR_INVALID, "null handle");
    WMD_REQUIRE(d_counts != nullptr || n_frames == 0, WMD_ERR_INVALID, "null counts buffer");
    bind(h->ctx);
    soft_run(h, false, d_soa, stride, n_atoms, n_frames, boxes, d_counts, nullptr, nullptr,
             pick_stream(h->ctx, stream));
    WMD_API_END
}

wmd_status wmd_softrdf_backward(wmd_softrdf *h, const float *d_soa, size_t stride, size_t n_atoms,
                                size_t n_frames, const float *boxes, const double *d_dcounts,
                                float *d_grad, void *stream) {
    WMD_API_BEGIN
    WMD_REQUIRE(h != nullptr, WMD_ERR_INVALID, "null handle");
    WMD_REQUIRE((d_dcounts && d_grad) || n_frames == 0, WMD_ERR_INVALID, "null device buffer");
    bind(h->ctx);
    soft_run(h, true, d_soa, stride, n_atoms, n_frames, boxes, nullptr, d_dcounts, d_grad,
             pick_stream(h->ctx, stream));
    WMD_API_END
}

wmd_status wmd_softrdf_last_path(const wmd_softrdf *h, int *path, int *nc) {
    WMD_API_BEGIN
    WMD_REQUIRE(h && path, WMD_ERR_INVALID, "null handle");
    *path = h->last_path;
    if (nc) for (int d = 0; d < 3; ++d) nc[d] = h->last_nc[d];
    WMD_API_END
}

wmd_status wmd_softrdf_set_tuning(wmd_softrdf *h, int cell_k) {
    WMD_API_BEGIN
    WMD_REQUIRE(h != nullptr, WMD_ERR_INVALID, "null handle");
    WMD_REQUIRE(cell_k >= -1 && cell_k <= 3, WMD_ERR_INVALID, "cell_k must be -1..3");
    h->cell_k = cell_k;
    h->cells_valid = false;
    WMD_API_END
}

wmd_status wmd_softrdf_keep_cells(wmd_softrdf *h, int keep) {
    WMD_API_BEGIN
    WMD_REQUIRE(h != nullptr, WMD_ERR_INVALID, "null handle");
    h->keep_cells = keep != 0;
    h->cells_valid = false;
    WMD_API_END
}

void wmd_softrdf_destroy(wmd_softrdf *h) {
    if (!h) return;
    (void)hipSetDevice(h->ctx->device);
    (void)hipDeviceSynchronize();
    delete h;
}

}  // extern "C"
