// PairwiseRmsdPlan on the device: the F x F matrix of RMSDs between every two
// frames of one selection, after optimal superposition ("rms",
// crates/traj-engine/src/plans/rmsd/plan.rs:654-668 with rmsd_aligned
// :1112-1159) or without it ("nofit", rmsd_raw :1081-1091).
//
// The reference fits every pair separately (O(F^2 n) scalar work on the CPU).
// Here the O(F^2 n) part is one symmetric Gram product on the FP64 matrix
// cores: with X the (3F x n) matrix of centred coordinates (row 3f + d holds
// frame f's d coordinate of every atom, minus the frame's f64 centroid),
// G = X X^T holds in its 3 x 3 block (i, j) exactly the cross-covariance
// H_ij = sum_k (x_ik - c_i)(x_jk - c_j)^T the reference accumulates.  Each
// workgroup computes a 96 x 96 tile of G (32 x 32 frame pairs) with
// v_mfma_f64_16x16x4 and finishes its frame pairs in the epilogue: the FP64
// QCP eigenvalue (wmd_qcp.hpp) for "rms", and
//   sum |x_i - x_j|^2 = G_i + G_j - 2 tr H_ij + n |c_i - c_j|^2
// for "nofit" (G_f = sum_k |x_fk - c_f|^2 from the centring pass).  Centring
// before the product keeps every sum free of cancellation; all arithmetic is
// f64 like the reference's, and the result is cast to f32 as it does.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>

#include "wmd_internal.hpp"
#include "wmd_qcp.hpp"

namespace {
using namespace wmd;

typedef double d4 __attribute__((ext_vector_type(4)));

constexpr int kPwBlock = 256;                      // 4 waves
constexpr int kPwFrames = 32;                      // frames per tile side
constexpr int kPwRows = 3 * kPwFrames;             // Gram rows per tile side
constexpr int kPwKc = 32;                          // atoms per staged chunk
constexpr int kPwQuads = kPwKc / 4;                // float4 groups per row and chunk
constexpr int kPwS4 = kPwRows + 1;                 // quad-row stride (odd: conflict-free staging)
constexpr int kPwPanel = kPwQuads * kPwS4 * 4;     // doubles per staged panel
constexpr int kPwGs = kPwRows + 1;                 // row stride of the tile Gram in LDS
constexpr int kPwLdsDoubles = (2 * kPwPanel > (kPwRows / 2) * kPwGs) ? 2 * kPwPanel : (kPwRows / 2) * kPwGs;

// [F][n_src] xyzw chunk -> SoA [F][3][stride] of the selected atoms (pads 0).
__global__ __launch_bounds__(kPwBlock) void pw_gather_kernel(const float4 *__restrict__ aos,
                                                             long long n_src,
                                                             const unsigned *__restrict__ idx,
                                                             int n, int stride, int n_frames,
                                                             float *__restrict__ soa) {
    const long long total = (long long)stride * n_frames;
    for (long long t = (long long)blockIdx.x * kPwBlock + threadIdx.x; t < total;
         t += (long long)gridDim.x * kPwBlock) {
        const int f = (int)(t / stride);
        const int k = (int)(t - (long long)f * stride);
        float x = 0.f, y = 0.f, z = 0.f;
        if (k < n) {
            const float4 v = aos[(long long)f * n_src + idx[k]];
            x = v.x;
            y = v.y;
            z = v.z;
        }
        float *o = soa + (long long)f * 3 * stride;
        o[k] = x;
        o[stride + k] = y;
        o[2 * stride + k] = z;
    }
}

__device__ __forceinline__ double block_sum(double v, double *red) {
    // 256 threads: wave sums through LDS, then one wave adds the four
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    __syncthreads();
    if (lane == 0) red[w] = v;
    __syncthreads();
    return red[0] + red[1] + red[2] + red[3];
}

// Per frame: centroid c = (sum_k x_k) / n in f64 (rmsd/math.rs:40-46) and
// G = sum_k |x_k - c|^2.  One workgroup per frame (grid-stride).
__global__ __launch_bounds__(kPwBlock) void pw_center_kernel(const float *__restrict__ soa,
                                                             int stride, int n, int n_frames,
                                                             double *__restrict__ cen,
                                                             double *__restrict__ gsum) {
    __shared__ double red[4];
    for (int f = blockIdx.x; f < n_frames; f += gridDim.x) {
        const float *x = soa + (long long)f * 3 * stride;
        double s[3] = {0.0, 0.0, 0.0};
        for (int k = threadIdx.x; k < n; k += kPwBlock) {
            s[0] += (double)x[k];
            s[1] += (double)x[stride + k];
            s[2] += (double)x[2 * stride + k];
        }
        double c[3];
        for (int d = 0; d < 3; ++d) c[d] = block_sum(s[d], red) / (double)n;
        double g = 0.0;
        for (int k = threadIdx.x; k < n; k += kPwBlock) {
            const double dx = (double)x[k] - c[0];
            const double dy = (double)x[stride + k] - c[1];
            const double dz = (double)x[2 * stride + k] - c[2];
            g += dx * dx + dy * dy + dz * dz;
        }
        g = block_sum(g, red);
        if (threadIdx.x == 0) {
            cen[3 * f] = c[0];
            cen[3 * f + 1] = c[1];
            cen[3 * f + 2] = c[2];
            gsum[f] = g;
        }
    }
}

// Staging of rows [row0, row0 + 96) x atoms [k0, k0 + 32) of X (centred,
// f64) in two halves so the global loads of the next chunk overlap the MFMAs
// of the current one: pw_load fetches a thread's three float4 groups,
// pw_store centres them and writes the panel, laid out [quad][row][4]
// (element (k, r) at ((k / 4) * S4 + r) * 4 + k % 4).
constexpr int kPwPer = (kPwRows * kPwQuads) / kPwBlock;  // float4 groups per thread and panel

__device__ __forceinline__ void pw_load(float4 *v, const float *__restrict__ soa, int stride, int n,
                                        int rows, int row0, int k0) {
#pragma unroll
    for (int e = 0; e < kPwPer; ++e) {
        const int idx = e * kPwBlock + threadIdx.x;
        const int r = idx / kPwQuads;
        const int k = k0 + 4 * (idx - r * kPwQuads);
        v[e] = make_float4(0.f, 0.f, 0.f, 0.f);
        if (row0 + r < rows && k < n) v[e] = *(const float4 *)(soa + (long long)(row0 + r) * stride + k);
    }
}

__device__ __forceinline__ void pw_store(double *panel, const float4 *v, int n, int rows, int row0,
                                         int k0, const double *__restrict__ cen) {
#pragma unroll
    for (int e = 0; e < kPwPer; ++e) {
        const int idx = e * kPwBlock + threadIdx.x;
        const int r = idx / kPwQuads;
        const int q = idx - r * kPwQuads;
        const int k = k0 + 4 * q;
        double v0 = 0.0, v1 = 0.0, v2 = 0.0, v3 = 0.0;
        if (row0 + r < rows && k < n) {
            const double c = cen[row0 + r];
            v0 = (double)v[e].x - c;
            if (k + 1 < n) v1 = (double)v[e].y - c;
            if (k + 2 < n) v2 = (double)v[e].z - c;
            if (k + 3 < n) v3 = (double)v[e].w - c;
        }
        // slots swizzled by row: rows with bit 3 set keep k % 4 = 0, 1 in slots
        // 2, 3 (and 2, 3 in 0, 1), so that the 32 lanes of an operand read
        // (16 rows x 2 k) fall on 64 distinct banks
        double *p = panel + (q * kPwS4 + r) * 4;
        const int sw = ((r >> 3) & 1) << 1;
        p[0 ^ sw] = v0;
        p[1 ^ sw] = v1;
        p[2 ^ sw] = v2;
        p[3 ^ sw] = v3;
    }
}

// One 96 x 96 tile of G per workgroup (tile list: upper triangle, ti <= tj),
// then the RMSD of its frame pairs.  Wave w owns rows 48 (w / 2) .. + 48 and
// columns 48 (w % 2) .. + 48 as 3 x 3 blocks of 16 x 16.
template <bool FIT>
__global__ __launch_bounds__(kPwBlock) void pw_gram_kernel(
    const float *__restrict__ soa, int stride, int n, int n_frames,
    const double *__restrict__ cen, const double *__restrict__ gsum,
    const int2 *__restrict__ tiles, int n_tiles, int per_xcd, int n_xcd, float *__restrict__ out,
    double *__restrict__ out64) {
    extern __shared__ __attribute__((aligned(16))) double lds[];
    // XCD-aware order: workgroups on one XCD take consecutive tiles (same row panel)
    const int b = blockIdx.x;
    const int t = (b % n_xcd) * per_xcd + b / n_xcd;
    if (t >= n_tiles) return;
    const int2 tp = tiles[t];
    const int rows = 3 * n_frames;
    const int ra = kPwRows * tp.x, rb = kPwRows * tp.y;
    double *pa = lds;
    double *pb = lds + kPwPanel;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int mb0 = 3 * (w >> 1), nb0 = 3 * (w & 1);
    d4 acc[3][3];
#pragma unroll
    for (int m = 0; m < 3; ++m)
#pragma unroll
        for (int j = 0; j < 3; ++j) acc[m][j] = (d4){0.0, 0.0, 0.0, 0.0};
    const int lr = lane & 15, lk = lane >> 4;
    const int lsw = ((lr >> 3) & 1) << 1;  // the staging's slot swizzle (rows 16 b + lr)
    float4 va[kPwPer], vb[kPwPer];
    pw_load(va, soa, stride, n, rows, ra, 0);
    pw_load(vb, soa, stride, n, rows, rb, 0);
    for (int k0 = 0; k0 < n; k0 += kPwKc) {
        __syncthreads();  // the previous chunk is consumed
        pw_store(pa, va, n, rows, ra, k0, cen);
        pw_store(pb, vb, n, rows, rb, k0, cen);
        __syncthreads();
        if (k0 + kPwKc < n) {  // next chunk in flight during this one's MFMAs
            pw_load(va, soa, stride, n, rows, ra, k0 + kPwKc);
            pw_load(vb, soa, stride, n, rows, rb, k0 + kPwKc);
        }
#pragma unroll
        for (int q = 0; q < kPwQuads; ++q) {
            double av[3], bv[3];
#pragma unroll
            for (int m = 0; m < 3; ++m) av[m] = pa[(q * kPwS4 + 16 * (mb0 + m) + lr) * 4 + (lk ^ lsw)];
#pragma unroll
            for (int j = 0; j < 3; ++j) bv[j] = pb[(q * kPwS4 + 16 * (nb0 + j) + lr) * 4 + (lk ^ lsw)];
#pragma unroll
            for (int m = 0; m < 3; ++m)
#pragma unroll
                for (int j = 0; j < 3; ++j)
                    acc[m][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[m], bv[j], acc[m][j], 0, 0, 0);
        }
    }
    // Epilogue in two halves of 16 frames x 32 frames: the waves owning rows
    // 48 h .. 48 h + 47 put their blocks in LDS (lane l, register r of a block
    // holds row 4 r + l / 16, column l % 16), then every thread finishes pairs.
    double *gt = lds;
    const double nd = (double)n;
    const int jj = threadIdx.x & 31;
    for (int half = 0; half < 2; ++half) {
        __syncthreads();
        if ((w >> 1) == half) {
#pragma unroll
            for (int m = 0; m < 3; ++m)
#pragma unroll
                for (int j = 0; j < 3; ++j)
#pragma unroll
                    for (int r = 0; r < 4; ++r)
                        gt[(16 * m + 4 * r + lk) * kPwGs + 16 * (nb0 + j) + lr] = acc[m][j][r];
        }
        __syncthreads();
        for (int il = threadIdx.x >> 5; il < kPwFrames / 2; il += kPwBlock / 32) {
            const int ii = (kPwFrames / 2) * half + il;
            const int i = kPwFrames * tp.x + ii, j = kPwFrames * tp.y + jj;
            if (i >= n_frames || j >= n_frames) continue;
            if (tp.x == tp.y && ii >= jj) {
                if (ii == jj) {
                    out[(long long)i * n_frames + i] = 0.f;
                    if (out64) out64[(long long)i * n_frames + i] = 0.0;
                }
                continue;
            }
            double M[9];
#pragma unroll
            for (int a = 0; a < 3; ++a)
#pragma unroll
                for (int c = 0; c < 3; ++c) M[3 * a + c] = gt[(3 * il + a) * kPwGs + 3 * jj + c];
            const double gi = gsum[i], gj = gsum[j];
            double msd;
            if (FIT) {
                QcpResult res;
                qcp_solve(M, gi, gj, false, &res);
                msd = 2.0 * (res.e0 - res.lambda) / nd;
            } else {
                const double dx = cen[3 * i] - cen[3 * j], dy = cen[3 * i + 1] - cen[3 * j + 1],
                             dz = cen[3 * i + 2] - cen[3 * j + 2];
                msd = (gi + gj - 2.0 * (M[0] + M[4] + M[8]) + nd * (dx * dx + dy * dy + dz * dz)) / nd;
            }
            const double v = sqrt(msd > 0.0 ? msd : 0.0);
            out[(long long)i * n_frames + j] = (float)v;
            out[(long long)j * n_frames + i] = (float)v;
            if (out64) {
                out64[(long long)i * n_frames + j] = v;
                out64[(long long)j * n_frames + i] = v;
            }
        }
    }
}

__global__ __launch_bounds__(kPwBlock) void pw_zero_kernel(float *out, double *out64, long long n) {
    for (long long t = (long long)blockIdx.x * kPwBlock + threadIdx.x; t < n;
         t += (long long)gridDim.x * kPwBlock) {
        out[t] = 0.f;
        if (out64) out64[t] = 0.0;
    }
}

int pw_grid(long long total) {
    return (int)std::max<long long>(1, std::min<long long>((total + kPwBlock - 1) / kPwBlock, 65536));
}

// ---------------------------------------------------------------------------
// TrajectoryClusterPlan's k-means (crates/traj-engine/src/plans/analysis/
// clustering.rs:384-519) on the plan's store: features are every frame's
// selected atoms superposed onto frame 0 (align_to_reference :394-418,
// R (x - cx) + cy in f64 with R from the FP64 QCP fit), then Lloyd iterations
// whose distance and sum loops run in the reference's order (sequential over
// features, resp. over samples), so they are bit-exact given the features.
// ---------------------------------------------------------------------------

// features [F][3 n] f64; one workgroup per frame (grid-stride).
__global__ __launch_bounds__(kPwBlock) void km_features_kernel(const float *__restrict__ soa, int stride,
                                                               int n, int n_frames,
                                                               double *__restrict__ feat) {
    __shared__ double red[4];
    __shared__ double rot[12];
    const float *y = soa;  // frame 0 is the reference
    double cy[3];
    for (int d = 0; d < 3; ++d) {
        double t = 0.0;
        for (int k = threadIdx.x; k < n; k += kPwBlock) t += (double)y[d * stride + k];
        cy[d] = block_sum(t, red) / (double)n;
    }
    double gy = 0.0;
    for (int k = threadIdx.x; k < n; k += kPwBlock)
        for (int d = 0; d < 3; ++d) {
            const double e = (double)y[d * stride + k] - cy[d];
            gy += e * e;
        }
    gy = block_sum(gy, red);
    for (int f = blockIdx.x; f < n_frames; f += gridDim.x) {
        const float *x = soa + (long long)f * 3 * stride;
        double cx[3];
        for (int d = 0; d < 3; ++d) {
            double t = 0.0;
            for (int k = threadIdx.x; k < n; k += kPwBlock) t += (double)x[d * stride + k];
            cx[d] = block_sum(t, red) / (double)n;
        }
        double m[10];
        for (int c = 0; c < 10; ++c) m[c] = 0.0;
        for (int k = threadIdx.x; k < n; k += kPwBlock) {
            const double xr[3] = {(double)x[k] - cx[0], (double)x[stride + k] - cx[1], (double)x[2 * stride + k] - cx[2]};
            const double yr[3] = {(double)y[k] - cy[0], (double)y[stride + k] - cy[1], (double)y[2 * stride + k] - cy[2]};
            for (int a = 0; a < 3; ++a)
                for (int b = 0; b < 3; ++b) m[3 * a + b] += xr[a] * yr[b];
            m[9] += xr[0] * xr[0] + xr[1] * xr[1] + xr[2] * xr[2];
        }
        for (int c = 0; c < 10; ++c) m[c] = block_sum(m[c], red);
        if (threadIdx.x == 0) {
            QcpResult res;
            qcp_solve(m, m[9], gy, true, &res);
            for (int i = 0; i < 9; ++i) rot[i] = res.rot[i];
        }
        __syncthreads();
        double R[9];
        for (int i = 0; i < 9; ++i) R[i] = rot[i];
        double *o = feat + (long long)f * 3 * n;
        for (int k = threadIdx.x; k < n; k += kPwBlock) {
            const double c0 = (double)x[k] - cx[0], c1 = (double)x[stride + k] - cx[1],
                         c2 = (double)x[2 * stride + k] - cx[2];
            for (int a = 0; a < 3; ++a) o[3 * k + a] = (R[3 * a] * c0 + R[3 * a + 1] * c1 + R[3 * a + 2] * c2) + cy[a];
        }
        __syncthreads();  // rot is reused by the next frame
    }
}

// dist [F][K]: sq_dist(sample i, centre c) summed over features in order (clustering.rs:521-528).
__global__ __launch_bounds__(kPwBlock) void km_dist_kernel(const double *__restrict__ feat,
                                                           const double *__restrict__ centers,
                                                           long long n_feat, int n_samples, int k,
                                                           double *__restrict__ dist) {
    const long long total = (long long)n_samples * k;
    for (long long t = (long long)blockIdx.x * kPwBlock + threadIdx.x; t < total;
         t += (long long)gridDim.x * kPwBlock) {
        const long long i = t / k, c = t - i * k;
        const double *a = feat + i * n_feat, *b = centers + c * n_feat;
        double s = 0.0;
        for (long long q = 0; q < n_feat; ++q) {
            const double d = a[q] - b[q];
            s += d * d;
        }
        dist[t] = s;
    }
}

// sums [K][n_feat]: per cluster and feature, the samples' values added in sample order.
__global__ __launch_bounds__(kPwBlock) void km_sums_kernel(const double *__restrict__ feat,
                                                           const int *__restrict__ labels,
                                                           long long n_feat, int n_samples, int k,
                                                           double *__restrict__ sums) {
    const long long total = (long long)k * n_feat;
    for (long long t = (long long)blockIdx.x * kPwBlock + threadIdx.x; t < total;
         t += (long long)gridDim.x * kPwBlock) {
        const int c = (int)(t / n_feat);
        const long long q = t - (long long)c * n_feat;
        double s = 0.0;
        for (int i = 0; i < n_samples; ++i)
            if (labels[i] == c) s += feat[(long long)i * n_feat + q];
        sums[t] = s;
    }
}

// ---------------------------------------------------------------------------
// Metric "dme" (plan.rs:612-650): the RMS difference of two frames' compact
// pair-distance vectors (pair_distances_compact, plan.rs:1279-1302).  The
// store is transposed to atom-major planes so that consecutive threads read
// consecutive frames; blocks of K pair distances D[k][f] (f64) are computed
// for every frame, and the frame-pair sums S[i][j] = sum_k (D[k][i] -
// D[k][j])^2 advance block by block in k order with the reference's own
// roundings (a separate multiply and add, no contraction), so every S[i][j]
// is the reference's sequential f64 sum.
// ---------------------------------------------------------------------------
constexpr int kDmeTile = 64;  // frames per tile side (4 x 4 per thread)
constexpr int kDmeKc = 16;    // pair distances per LDS chunk

// store [F][3][stride] -> T [3][n][fp], fp = F rounded up to kDmeTile (0 pads).
__global__ __launch_bounds__(kPwBlock) void dme_transpose_kernel(const float *__restrict__ store,
                                                                 int stride, int n, int F, int fp,
                                                                 float *__restrict__ T) {
    const long long total = (long long)n * fp;
    for (long long t = (long long)blockIdx.x * kPwBlock + threadIdx.x; t < total;
         t += (long long)gridDim.x * kPwBlock) {
        const int a = (int)(t / fp);
        const int f = (int)(t - (long long)a * fp);
        float x = 0.f, y = 0.f, z = 0.f;
        if (f < F) {
            const float *s = store + (long long)f * 3 * stride;
            x = s[a];
            y = s[stride + a];
            z = s[2 * stride + a];
        }
        T[(long long)a * fp + f] = x;
        T[((long long)n + a) * fp + f] = y;
        T[(2LL * n + a) * fp + f] = z;
    }
}

// Pair k of the compact order (i < j, i major): row i starts at i (2n - i - 1) / 2.
__device__ __forceinline__ void dme_pair(long long k, int n, int *pi, int *pj) {
    const double b = 2.0 * n - 1.0;
    int i = (int)floor((b - sqrt(fmax(b * b - 8.0 * (double)k, 0.0))) * 0.5);
    i = max(0, min(i, n - 2));
    while (i > 0 && (long long)i * (2LL * n - i - 1) / 2 > k) --i;
    while (i < n - 2 && (long long)(i + 1) * (2LL * n - i - 2) / 2 <= k) ++i;
    *pi = i;
    *pj = (int)(k - (long long)i * (2LL * n - i - 1) / 2) + i + 1;
}

// D[kk][f] = distance of pair k0 + kk in frame f (0 for padding frames):
// the f32 difference pb - pa cast to f64, apply_pbc (pbc_math.rs:17-27) on
// the frame's f64 box lengths when box3 is given, sqrt in f64.
__global__ __launch_bounds__(kPwBlock) void dme_dist_kernel(const float *__restrict__ T, int n,
                                                            int F, int fp, long long k0, int K,
                                                            const double *__restrict__ box3,
                                                            double *__restrict__ D) {
    for (int kk = blockIdx.x; kk < K; kk += gridDim.x) {
        int i, j;
        dme_pair(k0 + kk, n, &i, &j);
        const float *xi = T + (long long)i * fp, *xj = T + (long long)j * fp;
        const float *yi = T + ((long long)n + i) * fp, *yj = T + ((long long)n + j) * fp;
        const float *zi = T + (2LL * n + i) * fp, *zj = T + (2LL * n + j) * fp;
        double *out = D + (long long)kk * fp;
        for (int f = threadIdx.x; f < fp; f += kPwBlock) {
            double d = 0.0;
            if (f < F) {
                double dx = (double)(xj[f] - xi[f]);
                double dy = (double)(yj[f] - yi[f]);
                double dz = (double)(zj[f] - zi[f]);
                if (box3) {
                    const double lx = box3[3 * f], ly = box3[3 * f + 1], lz = box3[3 * f + 2];
                    if (lx > 0.0) dx -= round(dx / lx) * lx;
                    if (ly > 0.0) dy -= round(dy / ly) * ly;
                    if (lz > 0.0) dz -= round(dz / lz) * lz;
                }
                d = sqrt(dx * dx + dy * dy + dz * dz);
            }
            out[f] = d;
        }
    }
}

// For tiles (ti <= tj) of kDmeTile x kDmeTile frame pairs: S[i][j] (row
// stride fp) continues its sum over the K pair distances of the block, k in
// order (first block: from 0).  Thread (ty, tx) owns rows ty + 16 p and
// columns tx + 16 q of the tile.
__global__ __launch_bounds__(kPwBlock) void dme_sum_kernel(const double *__restrict__ D, int fp, int K,
                                                           const int2 *__restrict__ tiles, int n_tiles,
                                                           int first, double *__restrict__ S) {
    __shared__ double A[kDmeKc][kDmeTile], B[kDmeKc][kDmeTile];
    const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
    for (int tt = blockIdx.x; tt < n_tiles; tt += gridDim.x) {
        const int2 tl = tiles[tt];
        const int i0 = tl.x * kDmeTile, j0 = tl.y * kDmeTile;
        double acc[4][4];
#pragma unroll
        for (int p = 0; p < 4; ++p)
#pragma unroll
            for (int q = 0; q < 4; ++q)
                acc[p][q] = first ? 0.0 : S[(long long)(i0 + ty + 16 * p) * fp + j0 + tx + 16 * q];
        for (int c0 = 0; c0 < K; c0 += kDmeKc) {
            // rows past K stage zeros on both sides: (0 - 0)^2 adds an exact 0
            for (int e = threadIdx.x; e < kDmeKc * kDmeTile; e += kPwBlock) {
                const int r = e / kDmeTile, c = e - r * kDmeTile;
                const bool in = c0 + r < K;
                A[r][c] = in ? D[(long long)(c0 + r) * fp + i0 + c] : 0.0;
                B[r][c] = in ? D[(long long)(c0 + r) * fp + j0 + c] : 0.0;
            }
            __syncthreads();
#pragma unroll 4
            for (int r = 0; r < kDmeKc; ++r) {
                double a[4], b[4];
#pragma unroll
                for (int p = 0; p < 4; ++p) {
                    a[p] = A[r][ty + 16 * p];
                    b[p] = B[r][tx + 16 * p];
                }
#pragma unroll
                for (int p = 0; p < 4; ++p)
#pragma unroll
                    for (int q = 0; q < 4; ++q) {
                        const double d = a[p] - b[q];
                        acc[p][q] = acc[p][q] + d * d;  // two roundings, as the reference
                    }
            }
            __syncthreads();
        }
#pragma unroll
        for (int p = 0; p < 4; ++p)
#pragma unroll
            for (int q = 0; q < 4; ++q)
                S[(long long)(i0 + ty + 16 * p) * fp + j0 + tx + 16 * q] = acc[p][q];
    }
}

// out[i][j] = out[j][i] = (f32) sqrt(S / n_pairs) (plan.rs:637-644); 0 on the diagonal.
__global__ __launch_bounds__(kPwBlock) void dme_finalize_kernel(const double *__restrict__ S, int fp,
                                                                int F, double n_pairs,
                                                                float *__restrict__ out,
                                                                double *__restrict__ out64) {
    const long long total = (long long)F * F;
    for (long long t = (long long)blockIdx.x * kPwBlock + threadIdx.x; t < total;
         t += (long long)gridDim.x * kPwBlock) {
        const int i = (int)(t / F), j = (int)(t - (long long)i * F);
        double v = 0.0;
        if (i != j) {
            const double sum = i < j ? S[(long long)i * fp + j] : S[(long long)j * fp + i];
            v = sqrt(sum / n_pairs);
        }
        out[t] = (float)v;
        if (out64) out64[t] = v;
    }
}

}  // namespace

struct wmd_pairwise {
    wmd_ctx *ctx = nullptr;
    std::vector<uint32_t> sel;
    int metric = WMD_PAIRWISE_RMS;
    int stride = 0;
    size_t frames = 0, cap = 0;
    DevBuf d_sel, d_cen, d_g, d_out, d_tiles;
    DevBuf p_aos[2];  // host-chunk pipeline of wmd_pairwise_append
    DevBuf d_feat, d_centers, d_dist, d_labels, d_sums;
    std::unique_ptr<DevBuf> d_store;
    HostBuf p_stage[2];
    Event p_copied[2], p_used[2];
    std::vector<int2> tiles;
    // metric "dme": pbc (none / orthorhombic) and, for orthorhombic, the box
    // lengths of every appended frame ([F][3], f64 of the f32 lengths)
    int pbc = WMD_PBC_NONE;
    std::vector<double> box3;
    DevBuf d_dme_t, d_dme_d, d_dme_s, d_box3, d_dme_tiles;
    size_t dme_tiles_n = 0;
};

namespace {

// Metric "dme" on a store [F][3][stride]; d_box3: NULL or [F][3] f64 box lengths.
void dme_run(wmd_pairwise *h, const float *d_soa, size_t stride, size_t F, float *d_out,
             double *d_out64, const double *d_box3, hipStream_t st) {
    const int n = (int)h->sel.size();
    const long long n_pairs = (long long)n * (n - 1) / 2;
    const int T = (int)((F + kDmeTile - 1) / kDmeTile);
    const int fp = T * kDmeTile;
    WMD_REQUIRE((size_t)fp * fp * 8 <= (size_t(32) << 30), WMD_ERR_UNSUPPORTED,
                "dme: the frame-pair sums exceed 32 GB");
    float *d_t = (float *)h->d_dme_t.reserve((size_t)3 * n * fp * 4);
    hipLaunchKernelGGL(dme_transpose_kernel, dim3(pw_grid((long long)n * fp)), dim3(kPwBlock), 0, st,
                       d_soa, (int)stride, n, (int)F, fp, d_t);
    WMD_HIP_CHECK(hipGetLastError());
    const size_t nt = (size_t)T * (T + 1) / 2;
    if (h->dme_tiles_n != nt) {
        std::vector<int2> tl;
        tl.reserve(nt);
        for (int i = 0; i < T; ++i)
            for (int j = i; j < T; ++j) tl.push_back(make_int2(i, j));
        int2 *d = (int2 *)h->d_dme_tiles.reserve(nt * sizeof(int2));
        WMD_HIP_CHECK(hipMemcpyAsync(d, tl.data(), nt * sizeof(int2), hipMemcpyHostToDevice, st));
        WMD_HIP_CHECK(hipStreamSynchronize(st));  // tl dies at scope exit
        h->dme_tiles_n = nt;
    }
    // blocks of pair distances: about 256 MB of D per block
    const long long per = std::max<long long>(kDmeKc, (long long)((size_t(256) << 20) / ((size_t)fp * 8)) / kDmeKc * kDmeKc);
    const long long K = std::min<long long>(per, n_pairs);
    double *d_d = (double *)h->d_dme_d.reserve((size_t)K * fp * 8);
    double *d_s = (double *)h->d_dme_s.reserve((size_t)fp * fp * 8);
    ProfScope prof(h->ctx, st, "pairwise_dme");
    for (long long k0 = 0; k0 < n_pairs; k0 += K) {
        const int kb = (int)std::min<long long>(K, n_pairs - k0);
        hipLaunchKernelGGL(dme_dist_kernel, dim3((unsigned)std::min(kb, 65536)), dim3(kPwBlock), 0, st,
                           (const float *)d_t, n, (int)F, fp, k0, kb, d_box3, d_d);
        WMD_HIP_CHECK(hipGetLastError());
        hipLaunchKernelGGL(dme_sum_kernel, dim3((unsigned)std::min<size_t>(nt, 65536)), dim3(kPwBlock), 0, st,
                           (const double *)d_d, fp, kb, (const int2 *)h->d_dme_tiles.ptr, (int)nt,
                           k0 == 0 ? 1 : 0, d_s);
        WMD_HIP_CHECK(hipGetLastError());
    }
    hipLaunchKernelGGL(dme_finalize_kernel, dim3(pw_grid((long long)(F * F))), dim3(kPwBlock), 0, st,
                       (const double *)d_s, fp, (int)F, (double)n_pairs, d_out, d_out64);
    WMD_HIP_CHECK(hipGetLastError());
}

// The F x F matrix of a store [F][3][stride] on stream st (d_box3: the dme
// metric's box lengths, NULL without PBC).
void pw_run(wmd_pairwise *h, const float *d_soa, size_t stride, size_t F, float *d_out,
            double *d_out64, hipStream_t st, const double *d_box3 = nullptr) {
    const int n = (int)h->sel.size();
    if (F == 0) return;
    WMD_REQUIRE(stride >= (size_t)n && stride % 4 == 0, WMD_ERR_INVALID,
                "stride must be a multiple of 4 and >= the selection size");
    WMD_REQUIRE(F <= 46000, WMD_ERR_UNSUPPORTED, "at most 46000 frames (F x F output)");
    WMD_REQUIRE(3 * F * stride < (1ull << 40), WMD_ERR_UNSUPPORTED, "store too large");
    if (n < 2) {  // plan.rs:597-603: a zero matrix
        hipLaunchKernelGGL(pw_zero_kernel, dim3(pw_grid((long long)(F * F))), dim3(kPwBlock), 0, st,
                           d_out, d_out64, (long long)(F * F));
        WMD_HIP_CHECK(hipGetLastError());
        return;
    }
    if (h->metric == WMD_PAIRWISE_DME) {
        dme_run(h, d_soa, stride, F, d_out, d_out64, d_box3, st);
        return;
    }
    double *cen = (double *)h->d_cen.reserve(F * 3 * sizeof(double));
    double *g = (double *)h->d_g.reserve(F * sizeof(double));
    {
        ProfScope prof(h->ctx, st, "pairwise_center");
        hipLaunchKernelGGL(pw_center_kernel, dim3((unsigned)std::min<size_t>(F, 65536)), dim3(kPwBlock),
                           0, st, d_soa, (int)stride, n, (int)F, cen, g);
        WMD_HIP_CHECK(hipGetLastError());
    }
    const int T = (int)((F + kPwFrames - 1) / kPwFrames);
    const size_t nt = (size_t)T * (T + 1) / 2;
    if (h->tiles.size() != nt) {
        h->tiles.clear();
        h->tiles.reserve(nt);
        for (int i = 0; i < T; ++i)
            for (int j = i; j < T; ++j) h->tiles.push_back(make_int2(i, j));
        int2 *d = (int2 *)h->d_tiles.reserve(nt * sizeof(int2));
        WMD_HIP_CHECK(hipMemcpyAsync(d, h->tiles.data(), nt * sizeof(int2), hipMemcpyHostToDevice, st));
        WMD_HIP_CHECK(hipStreamSynchronize(st));
    }
    const int n_xcd = std::max(1, h->ctx->xcd_count);
    const int per_xcd = (int)((nt + n_xcd - 1) / n_xcd);
    const size_t lds = (size_t)kPwLdsDoubles * sizeof(double);
    ProfScope prof(h->ctx, st, "pairwise_gram");
    auto kern = h->metric == WMD_PAIRWISE_RMS ? pw_gram_kernel<true> : pw_gram_kernel<false>;
    hipLaunchKernelGGL(kern, dim3((unsigned)(per_xcd * n_xcd)), dim3(kPwBlock), lds, st, d_soa,
                       (int)stride, n, (int)F, (const double *)cen, (const double *)g,
                       (const int2 *)h->d_tiles.ptr, (int)nt, per_xcd, n_xcd, d_out, d_out64);
    WMD_HIP_CHECK(hipGetLastError());
}

}  // namespace

extern "C" {

wmd_status wmd_pairwise_create(wmd_ctx *ctx, const uint32_t *sel, size_t n_sel, int metric,
                               wmd_pairwise **out) {
    WMD_API_BEGIN
    WMD_REQUIRE(ctx && out, WMD_ERR_INVALID, "null context or output handle");
    *out = nullptr;
    WMD_REQUIRE(n_sel == 0 || sel, WMD_ERR_INVALID, "null selection");
    WMD_REQUIRE(metric == WMD_PAIRWISE_RMS || metric == WMD_PAIRWISE_NOFIT ||
                    metric == WMD_PAIRWISE_DME,
                WMD_ERR_PARSE, "metric must be rms, nofit or dme");
    WMD_REQUIRE(n_sel < (1u << 30), WMD_ERR_UNSUPPORTED, "selection too large");
    bind(ctx);
    auto h = std::make_unique<wmd_pairwise>();
    h->ctx = ctx;
    h->metric = metric;
    h->sel.assign(sel, sel + n_sel);
    h->stride = (int)round_up(std::max<size_t>(n_sel, 1), 4);
    h->d_sel.reserve(std::max<size_t>(n_sel, 1) * 4);
    if (n_sel)
        WMD_HIP_CHECK(hipMemcpy(h->d_sel.ptr, sel, n_sel * 4, hipMemcpyHostToDevice));
    *out = h.release();
    WMD_API_END
}

wmd_status wmd_pairwise_create_dme(wmd_ctx *ctx, const uint32_t *sel, size_t n_sel, int pbc,
                                   wmd_pairwise **out) {
    if (pbc != WMD_PBC_NONE && pbc != WMD_PBC_ORTHO) {  // parse_pbc (pack.rs:2665-2673)
        if (out) *out = nullptr;
        set_error("pbc must be 'orthorhombic' or 'none'");
        return WMD_ERR_PARSE;
    }
    wmd_status st = wmd_pairwise_create(ctx, sel, n_sel, WMD_PAIRWISE_DME, out);
    if (st == WMD_OK) (*out)->pbc = pbc;
    return st;
}

static wmd_status pairwise_append_impl(wmd_pairwise *h, const float *coords_xyzw, size_t n_atoms,
                                       size_t n_frames, const float *boxes, bool with_boxes);

wmd_status wmd_pairwise_append(wmd_pairwise *h, const float *coords_xyzw, size_t n_atoms,
                               size_t n_frames) {
    return pairwise_append_impl(h, coords_xyzw, n_atoms, n_frames, nullptr, false);
}

wmd_status wmd_pairwise_append_boxes(wmd_pairwise *h, const float *coords_xyzw, size_t n_atoms,
                                     size_t n_frames, const float *boxes) {
    return pairwise_append_impl(h, coords_xyzw, n_atoms, n_frames, boxes, true);
}

static wmd_status pairwise_append_impl(wmd_pairwise *h, const float *coords_xyzw, size_t n_atoms,
                                       size_t n_frames, const float *boxes, bool with_boxes) {
    WMD_API_BEGIN
    WMD_REQUIRE(h != nullptr, WMD_ERR_INVALID, "null handle");
    if (n_frames == 0) return WMD_OK;
    // dme with PBC keeps every frame's box lengths (plan.rs:549-556, box_lengths
    // pbc_math.rs:11-15); other metrics ignore boxes, as the reference does
    std::vector<double> new_box3;
    if (h->metric == WMD_PAIRWISE_DME && h->pbc == WMD_PBC_ORTHO) {
        WMD_REQUIRE(with_boxes && boxes != nullptr, WMD_ERR_MISMATCH, "pairwise_rmsd requires orthorhombic box");
        new_box3.resize(3 * n_frames);
        for (size_t f = 0; f < n_frames; ++f) {
            const float *m = boxes + 9 * f;
            WMD_REQUIRE(m[1] == 0.f && m[2] == 0.f && m[3] == 0.f && m[5] == 0.f && m[6] == 0.f && m[7] == 0.f,
                        WMD_ERR_MISMATCH, "orthorhombic box required for PBC");
            new_box3[3 * f] = m[0];
            new_box3[3 * f + 1] = m[4];
            new_box3[3 * f + 2] = m[8];
        }
    }
    WMD_REQUIRE(coords_xyzw != nullptr, WMD_ERR_INVALID, "null coords");
    for (uint32_t a : h->sel)
        WMD_REQUIRE(a < n_atoms, WMD_ERR_MISMATCH,
                    "selection index " + std::to_string(a) + " out of range for " +
                        std::to_string(n_atoms) + " atoms");
    bind(h->ctx);
    hipStream_t st = h->ctx->stream;
    const size_t fl = (size_t)3 * h->stride;  // floats per stored frame
    if (h->frames + n_frames > h->cap) {
        const size_t cap = std::max({h->frames + n_frames, 2 * h->cap, (size_t)64});
        auto nb = std::make_unique<DevBuf>();
        nb->reserve(cap * fl * 4);
        if (h->frames)
            WMD_HIP_CHECK(hipMemcpyAsync(nb->ptr, h->d_store->ptr, h->frames * fl * 4,
                                         hipMemcpyDeviceToDevice, st));
        WMD_HIP_CHECK(hipStreamSynchronize(st));
        h->d_store = std::move(nb);
        h->cap = cap;
    }
    // pipeline as in the RDF host path: host copies into pinned staging b,
    // H2D on the copy stream (p_copied[b]), gather into the store on st
    // (p_used[b]: the gather has read the device copy of b)
    hipStream_t cs = copy_stream(h->ctx);
    const size_t fbytes = n_atoms * 16;
    const size_t batch = staging_frames(h->ctx, fbytes, n_frames);
    if (h->p_stage[0].bytes < batch * fbytes || h->p_stage[1].bytes < batch * fbytes) {
        WMD_HIP_CHECK(hipStreamSynchronize(st));  // growing frees buffers queued work may use
        WMD_HIP_CHECK(hipStreamSynchronize(cs));
        for (int b = 0; b < 2; ++b) {
            h->p_stage[b].reserve(batch * fbytes);
            h->p_aos[b].reserve(batch * fbytes);
        }
    }
    WMD_HIP_CHECK(hipEventRecord(h->p_used[0].get(), st));
    WMD_HIP_CHECK(hipEventRecord(h->p_used[1].get(), st));
    size_t k = 0;
    for (size_t f0 = 0; f0 < n_frames; f0 += batch, ++k) {
        const int b = (int)(k & 1);
        const size_t nf = std::min(batch, n_frames - f0);
        h->p_copied[b].wait();  // the H2D out of this staging buffer two batches ago is done
        host_copy(h->p_stage[b].ptr, coords_xyzw + f0 * n_atoms * 4, nf * fbytes);
        WMD_HIP_CHECK(hipStreamWaitEvent(cs, h->p_used[b].get(), 0));
        WMD_HIP_CHECK(hipMemcpyAsync(h->p_aos[b].ptr, h->p_stage[b].ptr, nf * fbytes, hipMemcpyHostToDevice, cs));
        WMD_HIP_CHECK(hipEventRecord(h->p_copied[b].get(), cs));
        WMD_HIP_CHECK(hipStreamWaitEvent(st, h->p_copied[b].get(), 0));
        float *dst = (float *)h->d_store->ptr + (h->frames + f0) * fl;
        hipLaunchKernelGGL(pw_gather_kernel, dim3(pw_grid((long long)h->stride * nf)), dim3(kPwBlock), 0,
                           st, (const float4 *)h->p_aos[b].ptr, (long long)n_atoms, h->d_sel.as<unsigned>(),
                           (int)h->sel.size(), h->stride, (int)nf, dst);
        WMD_HIP_CHECK(hipGetLastError());
        WMD_HIP_CHECK(hipEventRecord(h->p_used[b].get(), st));
    }
    WMD_HIP_CHECK(hipStreamSynchronize(st));
    h->frames += n_frames;
    h->box3.insert(h->box3.end(), new_box3.begin(), new_box3.end());
    WMD_API_END
}

wmd_status wmd_pairwise_frames(const wmd_pairwise *h, size_t *n_frames) {
    WMD_API_BEGIN
    WMD_REQUIRE(h && n_frames, WMD_ERR_INVALID, "null handle");
    *n_frames = h->frames;
    WMD_API_END
}

wmd_status wmd_pairwise_finalize(wmd_pairwise *h, float *out) {
    WMD_API_BEGIN
    WMD_REQUIRE(h != nullptr, WMD_ERR_INVALID, "null handle");
    const size_t F = h->frames;
    if (F == 0) return WMD_OK;
    WMD_REQUIRE(out != nullptr, WMD_ERR_INVALID, "null output");
    bind(h->ctx);
    hipStream_t st = h->ctx->stream;
    float *d_out = (float *)h->d_out.reserve(F * F * 4);
    const double *d_box3 = nullptr;
    if (h->metric == WMD_PAIRWISE_DME && h->pbc == WMD_PBC_ORTHO) {
        WMD_REQUIRE(h->box3.size() == 3 * F, WMD_ERR_MISMATCH, "pairwise_rmsd requires orthorhombic box");
        double *d = (double *)h->d_box3.reserve(h->box3.size() * 8);
        WMD_HIP_CHECK(hipMemcpyAsync(d, h->box3.data(), h->box3.size() * 8, hipMemcpyHostToDevice, st));
        d_box3 = d;
    }
    pw_run(h, (const float *)h->d_store->ptr, (size_t)h->stride, F, d_out, nullptr, st, d_box3);
    WMD_HIP_CHECK(hipMemcpyAsync(out, d_out, F * F * 4, hipMemcpyDeviceToHost, st));
    WMD_HIP_CHECK(hipStreamSynchronize(st));
    WMD_API_END
}

wmd_status wmd_pairwise_store_stride(const wmd_pairwise *h, size_t *stride) {
    WMD_API_BEGIN
    WMD_REQUIRE(h && stride, WMD_ERR_INVALID, "null handle");
    *stride = (size_t)h->stride;
    WMD_API_END
}

wmd_status wmd_pairwise_run_device(wmd_pairwise *h, const float *d_store, size_t stride,
                                   size_t n_frames, float *d_out, double *d_out64, void *stream) {
    WMD_API_BEGIN
    WMD_REQUIRE(h != nullptr, WMD_ERR_INVALID, "null handle");
    WMD_REQUIRE(n_frames == 0 || (d_store && d_out), WMD_ERR_INVALID, "null device buffer");
    WMD_REQUIRE(((uintptr_t)d_store & 15) == 0, WMD_ERR_INVALID, "device store must be 16-byte aligned");
    WMD_REQUIRE(!(h->metric == WMD_PAIRWISE_DME && h->pbc == WMD_PBC_ORTHO), WMD_ERR_UNSUPPORTED,
                "dme with PBC takes per-frame boxes: use wmd_pairwise_append_boxes and wmd_pairwise_finalize");
    bind(h->ctx);
    pw_run(h, d_store, stride, n_frames, d_out, d_out64, pick_stream(h->ctx, stream));
    WMD_API_END
}

// clustering.rs:530-550: xorshift64 and the Fisher-Yates shuffle it drives.
static uint64_t km_xorshift64(uint64_t *state) {
    uint64_t x = *state;
    if (x == 0) x = 0x9E3779B97F4A7C15ull;
    x ^= x << 13;
    x ^= x >> 7;
    x ^= x << 17;
    *state = x;
    return x;
}

static double km_sq_dist(const double *a, const double *b, size_t n) {
    double s = 0.0;
    for (size_t q = 0; q < n; ++q) {
        const double d = a[q] - b[q];
        s += d * d;
    }
    return s;
}

wmd_status wmd_pairwise_kmeans(wmd_pairwise *h, size_t n_clusters, size_t max_iter, double tol,
                               uint64_t seed, int32_t *labels, uint32_t *centroids, uint32_t *sizes) {
    WMD_API_BEGIN
    WMD_REQUIRE(h != nullptr, WMD_ERR_INVALID, "null handle");
    const size_t F = h->frames;
    if (F == 0) return WMD_OK;
    WMD_REQUIRE(labels && centroids && sizes, WMD_ERR_INVALID, "null output");
    bind(h->ctx);
    hipStream_t st = h->ctx->stream;
    const int n = (int)h->sel.size();
    const size_t K = std::min(std::max<size_t>(n_clusters, 1), F);  // clustering.rs:195
    max_iter = std::max<size_t>(max_iter, 1);
    tol = std::max(tol, 0.0);
    const size_t nf = 3 * (size_t)n;
    WMD_REQUIRE(F * std::max<size_t>(nf, 1) * 8 < (64ull << 30), WMD_ERR_UNSUPPORTED, "k-means features exceed 64 GB");
    double *feat = (double *)h->d_feat.reserve(std::max<size_t>(F * nf, 1) * 8);
    if (n > 0) {
        hipLaunchKernelGGL(km_features_kernel, dim3((unsigned)std::min<size_t>(F, 65536)), dim3(kPwBlock), 0, st,
                           (const float *)h->d_store->ptr, h->stride, n, (int)F, feat);
        WMD_HIP_CHECK(hipGetLastError());
    }
    // kmeans_cluster (clustering.rs:420-519)
    uint64_t state = seed == 0 ? 0x9E3779B97F4A7C15ull : seed;
    std::vector<size_t> order(F);
    for (size_t i = 0; i < F; ++i) order[i] = i;
    if (F >= 2)
        for (size_t i = F - 1; i >= 1; --i) {
            const size_t j = (size_t)(km_xorshift64(&state) % (uint64_t)(i + 1));
            std::swap(order[i], order[j]);
        }
    std::vector<double> centers(K * nf), sums(K * nf), dist(F * K);
    std::vector<int32_t> lab(F, 0);
    std::vector<size_t> counts(K, 0);
    auto fetch_sample = [&](size_t i, double *dst) {
        if (nf) WMD_HIP_CHECK(hipMemcpy(dst, feat + i * nf, nf * 8, hipMemcpyDeviceToHost));
    };
    for (size_t c = 0; c < K; ++c) fetch_sample(order[c], centers.data() + c * nf);
    double *d_cent = (double *)h->d_centers.reserve(std::max<size_t>(K * nf, 1) * 8);
    double *d_dist = (double *)h->d_dist.reserve(F * K * 8);
    int *d_lab = (int *)h->d_labels.reserve(F * 4);
    double *d_sums = (double *)h->d_sums.reserve(std::max<size_t>(K * nf, 1) * 8);
    auto distances = [&]() {
        if (nf) WMD_HIP_CHECK(hipMemcpyAsync(d_cent, centers.data(), K * nf * 8, hipMemcpyHostToDevice, st));
        hipLaunchKernelGGL(km_dist_kernel, dim3(pw_grid((long long)(F * K))), dim3(kPwBlock), 0, st,
                           (const double *)feat, (const double *)d_cent, (long long)nf, (int)F, (int)K, d_dist);
        WMD_HIP_CHECK(hipGetLastError());
        WMD_HIP_CHECK(hipMemcpyAsync(dist.data(), d_dist, F * K * 8, hipMemcpyDeviceToHost, st));
        WMD_HIP_CHECK(hipStreamSynchronize(st));
    };
    for (size_t it = 0; it < max_iter; ++it) {
        distances();
        std::fill(counts.begin(), counts.end(), 0);
        for (size_t i = 0; i < F; ++i) {
            size_t best = 0;
            double best_d = INFINITY;
            for (size_t c = 0; c < K; ++c)
                if (dist[i * K + c] < best_d) {
                    best_d = dist[i * K + c];
                    best = c;
                }
            lab[i] = (int32_t)best;
            counts[best] += 1;
        }
        WMD_HIP_CHECK(hipMemcpyAsync(d_lab, lab.data(), F * 4, hipMemcpyHostToDevice, st));
        hipLaunchKernelGGL(km_sums_kernel, dim3(pw_grid((long long)(K * std::max<size_t>(nf, 1)))), dim3(kPwBlock), 0, st,
                           (const double *)feat, (const int *)d_lab, (long long)nf, (int)F, (int)K, d_sums);
        WMD_HIP_CHECK(hipGetLastError());
        if (nf) WMD_HIP_CHECK(hipMemcpyAsync(sums.data(), d_sums, K * nf * 8, hipMemcpyDeviceToHost, st));
        WMD_HIP_CHECK(hipStreamSynchronize(st));
        double max_shift = 0.0;
        std::vector<double> nc(nf);
        for (size_t c = 0; c < K; ++c) {
            if (counts[c] == 0) {
                const size_t idx = (size_t)(km_xorshift64(&state) % (uint64_t)F);
                fetch_sample(idx, nc.data());
            } else {
                const double inv = 1.0 / (double)counts[c];
                for (size_t q = 0; q < nf; ++q) nc[q] = sums[c * nf + q] * inv;
            }
            max_shift = std::max(max_shift, std::sqrt(km_sq_dist(centers.data() + c * nf, nc.data(), nf)));
            std::copy(nc.begin(), nc.end(), centers.begin() + c * nf);
        }
        if (max_shift <= tol) break;
    }
    std::fill(counts.begin(), counts.end(), 0);
    for (size_t i = 0; i < F; ++i) counts[(size_t)lab[i]] += 1;
    distances();  // to the final centres
    for (size_t c = 0; c < K; ++c) {
        size_t best_i = 0;
        double best_d = INFINITY;
        for (size_t i = 0; i < F; ++i) {
            if (lab[i] != (int32_t)c) continue;
            if (dist[i * K + c] < best_d) {
                best_d = dist[i * K + c];
                best_i = i;
            }
        }
        centroids[c] = (uint32_t)best_i;
        sizes[c] = (uint32_t)counts[c];
    }
    for (size_t i = 0; i < F; ++i) labels[i] = lab[i];
    WMD_API_END
}

// clustering.rs:290-382 dbscan_cluster + neighbors_within on a full [n][n]
// distance matrix (the triangle's tri_get: the diagonal is 0).
wmd_status wmd_dbscan(const float *dist, size_t n, double eps, size_t min_samples, int32_t *labels,
                      uint32_t *centroids, uint32_t *sizes, size_t *n_clusters) {
    WMD_API_BEGIN
    WMD_REQUIRE(n_clusters != nullptr, WMD_ERR_INVALID, "null output");
    *n_clusters = 0;
    if (n == 0) return WMD_OK;
    WMD_REQUIRE(dist && labels && centroids && sizes, WMD_ERR_INVALID, "null buffer");
    const int32_t kUnclassified = -2, kNoise = -1;
    auto d = [&](size_t i, size_t j) -> double { return i == j ? 0.0 : (double)dist[i * n + j]; };
    std::vector<int32_t> lab(n, kUnclassified);
    std::vector<char> visited(n, 0), queued(n, 0);
    std::vector<size_t> neigh, queue;
    auto neighbors = [&](size_t p) {
        neigh.clear();
        for (size_t j = 0; j < n; ++j)
            if (d(p, j) <= eps) neigh.push_back(j);
    };
    int32_t cluster = 0;
    for (size_t i = 0; i < n; ++i) {
        if (visited[i]) continue;
        visited[i] = 1;
        neighbors(i);
        if (neigh.size() < min_samples) {
            lab[i] = kNoise;
            continue;
        }
        lab[i] = cluster;
        std::fill(queued.begin(), queued.end(), 0);
        queue.clear();
        size_t head = 0;
        for (size_t p : neigh)
            if (p != i && !queued[p]) {
                queue.push_back(p);
                queued[p] = 1;
            }
        while (head < queue.size()) {
            const size_t p = queue[head++];
            if (!visited[p]) {
                visited[p] = 1;
                neighbors(p);
                if (neigh.size() >= min_samples)
                    for (size_t q : neigh)
                        if (!queued[q]) {
                            queue.push_back(q);
                            queued[q] = 1;
                        }
            }
            if (lab[p] == kUnclassified || lab[p] == kNoise) lab[p] = cluster;
        }
        ++cluster;
    }
    // medoids: the member with the smallest summed distance to the others
    for (int32_t c = 0; c < cluster; ++c) {
        std::vector<size_t> members;
        for (size_t i = 0; i < n; ++i)
            if (lab[i] == c) members.push_back(i);
        sizes[c] = (uint32_t)members.size();
        if (members.empty()) {
            centroids[c] = 0;
            continue;
        }
        size_t best = members[0];
        double best_sum = INFINITY;
        for (size_t a : members) {
            double sum = 0.0;
            for (size_t b : members) sum += d(a, b);
            if (sum < best_sum) {
                best_sum = sum;
                best = a;
            }
        }
        centroids[c] = (uint32_t)best;
    }
    for (size_t i = 0; i < n; ++i) labels[i] = lab[i];
    *n_clusters = (size_t)cluster;
    WMD_API_END
}

void wmd_pairwise_destroy(wmd_pairwise *h) {
    if (!h) return;
    int cur = 0;
    if (hipGetDevice(&cur) == hipSuccess && h->ctx) (void)hipSetDevice(h->ctx->device);
    delete h;
    (void)hipSetDevice(cur);
}

}  // extern "C"
