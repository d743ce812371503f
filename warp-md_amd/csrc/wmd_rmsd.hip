// RMSD superposition for gfx950: the engine behind RmsdPlan's device branch
// (crates/traj-engine/src/plans/rmsd/plan.rs:361-392) replacing
// GpuContext::rmsd_covariance / rmsd_raw (crates/traj-gpu/src/cuda/runtime_core.rs:375-535),
// the rmsd_centroid / rmsd_cov / rmsd_raw_accum / rmsd_finalize kernels
// (crates/traj-kernels/src/kernels/part1.cu:299-393) and the host-side
// rmsd_from_cov (plans/rmsd/math.rs:7-38).
//
// One pass over HBM.  A frame's store is SoA planes x[stride], y[stride],
// z[stride] of the plan's atoms, permuted into three 4-aligned segments:
// atoms in both the fit and the RMSD selection, fit-only atoms, RMSD-only
// atoms.  Each wave owns one (frame, tile) and streams float4 groups of the
// three planes; the reference planes stay L2-resident.  Per atom it adds the
// raw moments  Sx += x,  Sxx += |x|^2,  Sxy += x y^T  in FP64: an f32 x f32
// product is exact in FP64 and v_fma_f64 issues at the same rate as a
// non-packed FP32 FMA on gfx950, so this is both cheaper and more accurate
// than a compensated FP32 chain.  Lanes are combined by a wave tree reduction
// and tiles by a fixed-order sum, so results are deterministic.
//
// The finalize kernel (one thread per frame) centres the moments, solves the
// QCP polynomial in FP64 (wmd_qcp.hpp) and evaluates
//   RMSD^2 = (Gx_B + Gy_B - 2 tr(R M_B)) / n_B
// with the fit set's centroids and rotation, i.e. the composition of
// kabsch_rotation (geometry_math.rs:169-297) and the residual of rms.rs:656-672;
// for fit == RMSD selection it is 2 (E0 - lambda) / n, the RmsdPlan value.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <mutex>

#include "wmd_internal.hpp"
#include "wmd_qcp.hpp"

using namespace wmd;

// Scratch of the allocation-free entry point (wmd_rmsd_run_device_noalloc):
// the moments partials, at most kScratchBytes, in a fixed array that every
// GPU the library runs on holds until the process ends.  In host code the
// name `scratch` is only a host-side stand-in: kernels get the address that
// hipGetSymbolAddress returns for the current device (scratch_address).
static constexpr size_t kScratchBytes = size_t(16) << 20;
static __device__ __align__(256) double scratch[kScratchBytes / sizeof(double)];

namespace {

constexpr int kBlock = 256;
constexpr uint32_t kPad = 0xFFFFFFFFu;
constexpr int kMom = 13;  // Sx[3], Sxy[9], Sxx

enum { MODE_SAME = 0, MODE_SPLIT = 1, MODE_RAW = 2 };

std::mutex g_scratch_mu;
std::map<int, double *> g_scratch_addr;  // device -> its own copy of `scratch`

// The device address of `scratch` on `device` (which must be current),
// resolved once per device.
double *scratch_address(int device) {
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    auto it = g_scratch_addr.find(device);
    if (it != g_scratch_addr.end()) return it->second;
    void *p = nullptr;
    WMD_HIP_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    WMD_REQUIRE(p != nullptr, WMD_ERR_DEVICE, "no device address for the RMSD scratch array");
    g_scratch_addr[device] = static_cast<double *>(p);
    return static_cast<double *>(p);
}

struct RefMoments {
    double n_a, n_b;
    double sy_a[3], syy_a;  // reference moments over the fit set
    double sy_b[3], syy_b;  // over the RMSD set
};

__device__ __forceinline__ double wave_sum(double v) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

__device__ __forceinline__ void acc_atom(double *a, float xf, float yf, float zf, float rxf,
                                         float ryf, float rzf) {
    const double x = xf, y = yf, z = zf;
    const double rx = rxf, ry = ryf, rz = rzf;
    a[0] += x;
    a[1] += y;
    a[2] += z;
    a[3] = fma(x, rx, a[3]);
    a[4] = fma(x, ry, a[4]);
    a[5] = fma(x, rz, a[5]);
    a[6] = fma(y, rx, a[6]);
    a[7] = fma(y, ry, a[7]);
    a[8] = fma(y, rz, a[8]);
    a[9] = fma(z, rx, a[9]);
    a[10] = fma(z, ry, a[10]);
    a[11] = fma(z, rz, a[11]);
    a[12] = fma(x, x, fma(y, y, fma(z, z, a[12])));
}

__device__ __forceinline__ void acc_raw(double *a, float xf, float yf, float zf, float rxf,
                                        float ryf, float rzf) {
    // (f32 - f32) is exact in FP64, like the reference's `as f64` differences
    const double dx = (double)xf - (double)rxf;
    const double dy = (double)yf - (double)ryf;
    const double dz = (double)zf - (double)rzf;
    a[0] = fma(dx, dx, fma(dy, dy, fma(dz, dz, a[0])));
}

// The six second moments x x^T (xx, xy, xz, yy, yz, zz) of the RMSD set, for
// the adjoint of a split fit.
__device__ __forceinline__ void acc_xx(double *C, float xf, float yf, float zf) {
    const double x = xf, y = yf, z = zf;
    C[0] = fma(x, x, C[0]);
    C[1] = fma(x, y, C[1]);
    C[2] = fma(x, z, C[2]);
    C[3] = fma(y, y, C[3]);
    C[4] = fma(y, z, C[4]);
    C[5] = fma(z, z, C[5]);
}

template <int MODE>
__device__ __forceinline__ void acc_group(double *A, double *B, int seg, const float4 &x4,
                                          const float4 &y4, const float4 &z4, const float4 &rx,
                                          const float4 &ry, const float4 &rz) {
    if constexpr (MODE == MODE_RAW) {
        acc_raw(A, x4.x, y4.x, z4.x, rx.x, ry.x, rz.x);
        acc_raw(A, x4.y, y4.y, z4.y, rx.y, ry.y, rz.y);
        acc_raw(A, x4.z, y4.z, z4.z, rx.z, ry.z, rz.z);
        acc_raw(A, x4.w, y4.w, z4.w, rx.w, ry.w, rz.w);
    } else if constexpr (MODE == MODE_SAME) {
        acc_atom(A, x4.x, y4.x, z4.x, rx.x, ry.x, rz.x);
        acc_atom(A, x4.y, y4.y, z4.y, rx.y, ry.y, rz.y);
        acc_atom(A, x4.z, y4.z, z4.z, rx.z, ry.z, rz.z);
        acc_atom(A, x4.w, y4.w, z4.w, rx.w, ry.w, rz.w);
    } else {
        if (seg != 2) {  // fit set: both or fit-only
            acc_atom(A, x4.x, y4.x, z4.x, rx.x, ry.x, rz.x);
            acc_atom(A, x4.y, y4.y, z4.y, rx.y, ry.y, rz.y);
            acc_atom(A, x4.z, y4.z, z4.z, rx.z, ry.z, rz.z);
            acc_atom(A, x4.w, y4.w, z4.w, rx.w, ry.w, rz.w);
        }
        if (seg != 1) {  // RMSD set: both or RMSD-only
            acc_atom(B, x4.x, y4.x, z4.x, rx.x, ry.x, rz.x);
            acc_atom(B, x4.y, y4.y, z4.y, rx.y, ry.y, rz.y);
            acc_atom(B, x4.z, y4.z, z4.z, rx.z, ry.z, rz.z);
            acc_atom(B, x4.w, y4.w, z4.w, rx.w, ry.w, rz.w);
        }
    }
}

// partials: [F][n_tiles][NV] doubles, NV = 1 (raw), 13 (same), 26 (split;
// 32 with XX: the RMSD set's second moments for the split-fit adjoint)
template <int MODE, bool XX = false>
__global__ __launch_bounds__(kBlock) void rmsd_moments_kernel(
    const float *__restrict__ store, long long frame_stride, int stride, int n_frames,
    const float *__restrict__ ref, int n_groups, int seg1, int seg2, int n_tiles, int gpt,
    double *__restrict__ partials) {
    static_assert(!XX || MODE == MODE_SPLIT, "second moments only for a split fit");
    constexpr int NV = MODE == MODE_RAW ? 1 : (MODE == MODE_SAME ? kMom : 2 * kMom + (XX ? 6 : 0));
    const int wave = threadIdx.x >> 6;
    const int lane = threadIdx.x & 63;
    const long long unit = (long long)blockIdx.x * (kBlock / 64) + wave;
    if (unit >= (long long)n_frames * n_tiles) return;
    const int f = (int)(unit / n_tiles);
    const int t = (int)(unit - (long long)f * n_tiles);
    const int g0 = t * gpt;
    const int g1 = min(n_groups, g0 + gpt);
    const float4 *X = (const float4 *)(store + (long long)f * frame_stride);
    const float4 *Y = (const float4 *)(store + (long long)f * frame_stride + stride);
    const float4 *Z = (const float4 *)(store + (long long)f * frame_stride + 2 * stride);
    const float4 *RX = (const float4 *)ref;
    const float4 *RY = (const float4 *)(ref + stride);
    const float4 *RZ = (const float4 *)(ref + 2 * stride);
    double A[MODE == MODE_RAW ? 1 : kMom];
    double B[MODE == MODE_SPLIT ? kMom : 1];
    double C[XX ? 6 : 1];
#pragma unroll
    for (int k = 0; k < (MODE == MODE_RAW ? 1 : kMom); ++k) A[k] = 0.0;
#pragma unroll
    for (int k = 0; k < (MODE == MODE_SPLIT ? kMom : 1); ++k) B[k] = 0.0;
#pragma unroll
    for (int k = 0; k < (XX ? 6 : 1); ++k) C[k] = 0.0;
    auto second = [&](int seg, const float4 &x4, const float4 &y4, const float4 &z4) {
        if constexpr (XX) {
            if (seg != 1) {  // RMSD set: both or RMSD-only
                acc_xx(C, x4.x, y4.x, z4.x);
                acc_xx(C, x4.y, y4.y, z4.y);
                acc_xx(C, x4.z, y4.z, z4.z);
                acc_xx(C, x4.w, y4.w, z4.w);
            }
        }
    };
    int g = g0 + lane;
    // two groups in flight per lane
    for (; g + 64 < g1; g += 128) {
        float4 x0 = X[g], y0 = Y[g], z0 = Z[g];
        float4 x1 = X[g + 64], y1 = Y[g + 64], z1 = Z[g + 64];
        float4 a0 = RX[g], b0 = RY[g], c0 = RZ[g];
        float4 a1 = RX[g + 64], b1 = RY[g + 64], c1 = RZ[g + 64];
        const int s0 = (g >= seg1) + (g >= seg2);
        const int s1 = (g + 64 >= seg1) + (g + 64 >= seg2);
        acc_group<MODE>(A, B, s0, x0, y0, z0, a0, b0, c0);
        acc_group<MODE>(A, B, s1, x1, y1, z1, a1, b1, c1);
        second(s0, x0, y0, z0);
        second(s1, x1, y1, z1);
    }
    if (g < g1) {
        float4 x0 = X[g], y0 = Y[g], z0 = Z[g];
        float4 a0 = RX[g], b0 = RY[g], c0 = RZ[g];
        const int s0 = (g >= seg1) + (g >= seg2);
        acc_group<MODE>(A, B, s0, x0, y0, z0, a0, b0, c0);
        second(s0, x0, y0, z0);
    }
    double *out = partials + ((long long)f * n_tiles + t) * NV;
#pragma unroll
    for (int k = 0; k < (MODE == MODE_RAW ? 1 : kMom); ++k) {
        double v = wave_sum(A[k]);
        if (lane == 0) out[k] = v;
    }
    if constexpr (MODE == MODE_SPLIT) {
#pragma unroll
        for (int k = 0; k < kMom; ++k) {
            double v = wave_sum(B[k]);
            if (lane == 0) out[kMom + k] = v;
        }
    }
    if constexpr (XX) {
#pragma unroll
        for (int k = 0; k < 6; ++k) {
            double v = wave_sum(C[k]);
            if (lane == 0) out[2 * kMom + k] = v;
        }
    }
}

// Centred moments of a set relative to centroids (cx of the frame set, cy of
// the reference set): M = Sxy - cx Sy^T - Sx cy^T + n cx cy^T, G likewise.
__device__ __host__ inline void centred(const double *S, const double *sy, double syy, double n,
                                        const double *cx, const double *cy, double *M, double *gx,
                                        double *gy) {
    for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b)
            M[a * 3 + b] = S[3 + a * 3 + b] - cx[a] * sy[b] - S[a] * cy[b] + n * cx[a] * cy[b];
    *gx = S[12] - 2.0 * (cx[0] * S[0] + cx[1] * S[1] + cx[2] * S[2]) +
          n * (cx[0] * cx[0] + cx[1] * cx[1] + cx[2] * cx[2]);
    *gy = syy - 2.0 * (cy[0] * sy[0] + cy[1] * sy[1] + cy[2] * sy[2]) +
          n * (cy[0] * cy[0] + cy[1] * cy[1] + cy[2] * cy[2]);
}

template <int MODE>
__global__ void rmsd_finalize_kernel(const double *__restrict__ partials, int n_frames, int n_tiles,
                                     RefMoments rm, double lane_terms, float *__restrict__ out,
                                     double *__restrict__ out64, double *__restrict__ err,
                                     double *__restrict__ rot_out, double *__restrict__ cx_out) {
    constexpr int NV = MODE == MODE_RAW ? 1 : (MODE == MODE_SAME ? kMom : 2 * kMom);
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= n_frames) return;
    double S[NV];
    for (int k = 0; k < NV; ++k) S[k] = 0.0;
    for (int t = 0; t < n_tiles; ++t) {
        const double *p = partials + ((long long)f * n_tiles + t) * NV;
        for (int k = 0; k < NV; ++k) S[k] += p[k];
    }
    const double eps = 1.1102230246251565e-16;
    const double growth = (lane_terms + 8.0 + n_tiles) * eps;
    double rmsd = 0.0, e = 0.0;
    double R[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
    double cx[3] = {0, 0, 0};
    if constexpr (MODE == MODE_RAW) {
        if (rm.n_b > 0) {
            rmsd = sqrt(S[0] / rm.n_b);
            e = growth * S[0] / (2.0 * rm.n_b * fmax(rmsd, 1e-300));
        }
    } else {
        const double *SA = S;
        const double *SB = MODE == MODE_SPLIT ? S + kMom : S;
        double cy[3] = {0, 0, 0};
        if (rm.n_a > 0) {
            for (int d = 0; d < 3; ++d) {
                cx[d] = SA[d] / rm.n_a;
                cy[d] = rm.sy_a[d] / rm.n_a;
            }
        }
        double MA[9], gxa, gya;
        centred(SA, rm.sy_a, rm.syy_a, rm.n_a, cx, cy, MA, &gxa, &gya);
        QcpResult qr;
        double resid, mag;
        if constexpr (MODE == MODE_SAME) {
            qcp_solve(MA, gxa, gya, rot_out != nullptr, &qr);
            resid = gxa + gya - 2.0 * qr.lambda;
            mag = SA[12] + rm.syy_a + 2.0 * fabs(qr.lambda);
            if (rot_out) for (int k = 0; k < 9; ++k) R[k] = qr.rot[k];
        } else {
            if (rm.n_a > 0) {
                qcp_solve(MA, gxa, gya, true, &qr);
                for (int k = 0; k < 9; ++k) R[k] = qr.rot[k];
            }
            double MB[9], gxb, gyb;
            centred(SB, rm.sy_b, rm.syy_b, rm.n_b, cx, cy, MB, &gxb, &gyb);
            double tr = 0.0;
            for (int a = 0; a < 3; ++a)
                for (int b = 0; b < 3; ++b) tr += R[a * 3 + b] * MB[b * 3 + a];
            resid = gxb + gyb - 2.0 * tr;
            mag = SB[12] + rm.syy_b + 2.0 * fabs(tr);
        }
        const double n = MODE == MODE_SAME ? rm.n_a : rm.n_b;
        if (n > 0) {
            double r2 = resid / n;
            rmsd = r2 > 0.0 ? sqrt(r2) : 0.0;
            double de = growth * mag / n;  // error of r2
            e = de / (rmsd + sqrt(de));
        }
    }
    out[f] = (float)rmsd;
    if (out64) out64[f] = rmsd;
    if (err) err[f] = e;
    if (rot_out) for (int k = 0; k < 9; ++k) rot_out[(long long)f * 9 + k] = R[k];
    if (cx_out) for (int d = 0; d < 3; ++d) cx_out[(long long)f * 3 + d] = cx[d];
}

// AoS [F][n_src][4] -> SoA [F][3][stride] through the store's atom list.
__global__ __launch_bounds__(kBlock) void rmsd_gather_kernel(const float4 *__restrict__ aos,
                                                             long long n_src,
                                                             const unsigned *__restrict__ idx,
                                                             int stride, int n_frames,
                                                             float *__restrict__ soa) {
    long long total = (long long)stride * n_frames;
    for (long long t = (long long)blockIdx.x * kBlock + threadIdx.x; t < total;
         t += (long long)gridDim.x * kBlock) {
        int f = (int)(t / stride);
        int k = (int)(t - (long long)f * stride);
        unsigned a = idx[k];
        float x = 0.f, y = 0.f, z = 0.f;
        if (a != kPad) {
            float4 v = aos[(long long)f * n_src + a];
            x = v.x; y = v.y; z = v.z;
        }
        float *o = soa + (long long)f * 3 * stride;
        o[k] = x;
        o[stride + k] = y;
        o[2 * stride + k] = z;
    }
}

// Backward of the per-frame RMSD (no counterpart in the reference; the
// adjoint for structure-matching losses).  For the fitted RMSD over the fit
// set itself, the envelope theorem gives dRMSD/dx_i = ((x_i - cx) -
// R^T (y_i - cy)) / (n RMSD) with the optimal R and centroids (the centroid
// and rotation terms vanish at the optimum); for the unfitted RMSD,
// (x_i - y_i) / (n RMSD).  grad[f][d][slot] = gout[f] * dRMSD_f/dx (0 for
// padding slots and frames with RMSD = 0).
template <bool FIT>
__global__ __launch_bounds__(kBlock) void rmsd_backward_kernel(
    const float *__restrict__ store, int stride, int n_frames, const float *__restrict__ ref,
    const unsigned *__restrict__ atoms, double cy0, double cy1, double cy2, double n,
    const double *__restrict__ rmsd64, const double *__restrict__ rot,
    const double *__restrict__ cx, const float *__restrict__ gout, float *__restrict__ grad) {
    const long long total = (long long)n_frames * stride;
    for (long long t = (long long)blockIdx.x * kBlock + threadIdx.x; t < total;
         t += (long long)gridDim.x * kBlock) {
        const int f = (int)(t / stride);
        const int k = (int)(t - (long long)f * stride);
        const float *X = store + (long long)f * 3 * stride;
        float *G = grad + (long long)f * 3 * stride;
        const double r = rmsd64[f];
        double g0 = 0.0, g1 = 0.0, g2 = 0.0;
        if (atoms[k] != kPad && r > 0.0) {
            const double sc = (double)gout[f] / (n * r);
            const double x0 = X[k], x1 = X[stride + k], x2 = X[2 * stride + k];
            const double y0 = ref[k], y1 = ref[stride + k], y2 = ref[2 * stride + k];
            if constexpr (FIT) {
                const double *R = rot + (long long)f * 9;
                const double *c = cx + (long long)f * 3;
                const double v0 = y0 - cy0, v1 = y1 - cy1, v2 = y2 - cy2;
                g0 = sc * ((x0 - c[0]) - (R[0] * v0 + R[3] * v1 + R[6] * v2));
                g1 = sc * ((x1 - c[1]) - (R[1] * v0 + R[4] * v1 + R[7] * v2));
                g2 = sc * ((x2 - c[2]) - (R[2] * v0 + R[5] * v1 + R[8] * v2));
            } else {
                g0 = sc * (x0 - y0);
                g1 = sc * (x1 - y1);
                g2 = sc * (x2 - y2);
            }
        }
        G[k] = (float)g0;
        G[stride + k] = (float)g1;
        G[2 * stride + k] = (float)g2;
    }
}

// Adjoint of the split fit (fit on set A, RMSD over set B; the composition of
// rms.rs:643-676), per frame from the moments with second moments (NV = 32)
// and the forward's rotation.  With u = x - cx, v = y - cy, r = R u - v and
// E = sum_B |r|^2:
//   dE/dx_i = 2 R^T r_i [i in B] - (2 / n_A) R^T sum_B r [i in A] + 2 G^T v_i [i in A],
// where G = [z]x R, z = (tr(P) I - P)^-1 q, P = R H (H the fit set's centred
// cross-covariance, kept symmetric by the optimum), q the axial vector
// (Q23 - Q32, Q31 - Q13, Q12 - Q21) of Q = R W, W = sum_B u r^T
// = Suu R^T - sum_B u v^T.  coef [F][16]: cx (3), G (9), the centroid term
// (2 / n_A) R^T sum_B r (3), and sc = dL/dRMSD / (2 n_B RMSD) (0 when RMSD = 0).
__global__ void rmsd_split_adjoint_kernel(const double *__restrict__ partials, int n_frames,
                                          int n_tiles, RefMoments rm,
                                          const double *__restrict__ rot,
                                          const double *__restrict__ rmsd64,
                                          const float *__restrict__ gout,
                                          double *__restrict__ coef) {
    constexpr int NV = 2 * kMom + 6;
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= n_frames) return;
    double S[NV];
    for (int k = 0; k < NV; ++k) S[k] = 0.0;
    for (int t = 0; t < n_tiles; ++t) {
        const double *p = partials + ((long long)f * n_tiles + t) * NV;
        for (int k = 0; k < NV; ++k) S[k] += p[k];
    }
    const double *SA = S, *SB = S + kMom, *XXB = S + 2 * kMom;
    double R[9];
    for (int k = 0; k < 9; ++k) R[k] = rot[(long long)f * 9 + k];
    double cx[3] = {0, 0, 0}, cy[3] = {0, 0, 0};
    if (rm.n_a > 0)
        for (int d = 0; d < 3; ++d) {
            cx[d] = SA[d] / rm.n_a;
            cy[d] = rm.sy_a[d] / rm.n_a;
        }
    double H[9], MB[9], gx, gy;
    centred(SA, rm.sy_a, rm.syy_a, rm.n_a, cx, cy, H, &gx, &gy);
    centred(SB, rm.sy_b, rm.syy_b, rm.n_b, cx, cy, MB, &gx, &gy);
    // Suu = sum_B (x - cx)(x - cx)^T from the raw second moments
    const int ix[3][3] = {{0, 1, 2}, {1, 3, 4}, {2, 4, 5}};
    double Suu[9];
    for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b)
            Suu[a * 3 + b] = XXB[ix[a][b]] - cx[a] * SB[b] - SB[a] * cx[b] + rm.n_b * cx[a] * cx[b];
    double W[9], Q[9], P[9];
    for (int a = 0; a < 3; ++a)
        for (int c = 0; c < 3; ++c) {
            double w = 0.0;
            for (int b = 0; b < 3; ++b) w += Suu[a * 3 + b] * R[c * 3 + b];
            W[a * 3 + c] = w - MB[a * 3 + c];
        }
    for (int a = 0; a < 3; ++a)
        for (int c = 0; c < 3; ++c) {
            double q = 0.0, pp = 0.0;
            for (int b = 0; b < 3; ++b) {
                q += R[a * 3 + b] * W[b * 3 + c];
                pp += R[a * 3 + b] * H[b * 3 + c];
            }
            Q[a * 3 + c] = q;
            P[a * 3 + c] = pp;
        }
    const double qv[3] = {Q[5] - Q[7], Q[6] - Q[2], Q[1] - Q[3]};
    const double tr = P[0] + P[4] + P[8];
    // M = tr(P) I - sym(P), solved by its adjugate
    double M[9];
    for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b)
            M[a * 3 + b] = (a == b ? tr : 0.0) - 0.5 * (P[a * 3 + b] + P[b * 3 + a]);
    double adj[9];
    adj[0] = M[4] * M[8] - M[5] * M[7];
    adj[1] = M[2] * M[7] - M[1] * M[8];
    adj[2] = M[1] * M[5] - M[2] * M[4];
    adj[3] = M[5] * M[6] - M[3] * M[8];
    adj[4] = M[0] * M[8] - M[2] * M[6];
    adj[5] = M[2] * M[3] - M[0] * M[5];
    adj[6] = M[3] * M[7] - M[4] * M[6];
    adj[7] = M[1] * M[6] - M[0] * M[7];
    adj[8] = M[0] * M[4] - M[1] * M[3];
    const double det = M[0] * adj[0] + M[1] * adj[3] + M[2] * adj[6];
    double z[3] = {0, 0, 0};
    const double scale = fabs(tr) + 1e-300;
    if (fabs(det) > 1e-14 * scale * scale * scale)  // else a degenerate fit: no rotation term
        for (int a = 0; a < 3; ++a) z[a] = (adj[a * 3] * qv[0] + adj[a * 3 + 1] * qv[1] + adj[a * 3 + 2] * qv[2]) / det;
    const double Z[9] = {0.0, -z[2], z[1], z[2], 0.0, -z[0], -z[1], z[0], 0.0};
    double *o = coef + (long long)f * 16;
    for (int d = 0; d < 3; ++d) o[d] = cx[d];
    for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) {
            double g = 0.0;
            for (int c = 0; c < 3; ++c) g += Z[a * 3 + c] * R[c * 3 + b];
            o[3 + a * 3 + b] = g;
        }
    // sum_B r = R sum_B u - sum_B v
    double su[3], sr[3];
    for (int d = 0; d < 3; ++d) su[d] = SB[d] - rm.n_b * cx[d];
    for (int a = 0; a < 3; ++a)
        sr[a] = R[a * 3] * su[0] + R[a * 3 + 1] * su[1] + R[a * 3 + 2] * su[2] - (rm.sy_b[a] - rm.n_b * cy[a]);
    const double ca = rm.n_a > 0 ? 2.0 / rm.n_a : 0.0;
    for (int b = 0; b < 3; ++b) o[12 + b] = ca * (R[b] * sr[0] + R[3 + b] * sr[1] + R[6 + b] * sr[2]);
    const double r = rmsd64[f];
    o[15] = (r > 0.0 && rm.n_b > 0) ? (double)gout[f] / (2.0 * rm.n_b * r) : 0.0;
}

// grad[f][d][slot] of the split fit from the per-frame coefficients (0 for
// padding slots); a slot is in the fit set unless it is RMSD-only, in the
// RMSD set unless it is fit-only.
__global__ __launch_bounds__(kBlock) void rmsd_split_backward_kernel(
    const float *__restrict__ store, int stride, int n_frames, const float *__restrict__ ref,
    const unsigned *__restrict__ atoms, int seg1, int seg2, double cy0, double cy1, double cy2,
    const double *__restrict__ rot, const double *__restrict__ coef, float *__restrict__ grad) {
    const long long total = (long long)n_frames * stride;
    for (long long t = (long long)blockIdx.x * kBlock + threadIdx.x; t < total;
         t += (long long)gridDim.x * kBlock) {
        const int f = (int)(t / stride);
        const int k = (int)(t - (long long)f * stride);
        const double *o = coef + (long long)f * 16;
        const float *X = store + (long long)f * 3 * stride;
        float *Gr = grad + (long long)f * 3 * stride;
        double g[3] = {0.0, 0.0, 0.0};
        const double sc = o[15];
        if (atoms[k] != kPad && sc != 0.0) {
            const double *R = rot + (long long)f * 9;
            const int seg = (k / 4 >= seg1) + (k / 4 >= seg2);
            const double u[3] = {X[k] - o[0], X[stride + k] - o[1], X[2 * stride + k] - o[2]};
            const double v[3] = {ref[k] - cy0, ref[stride + k] - cy1, ref[2 * stride + k] - cy2};
            if (seg != 1) {  // RMSD set: 2 R^T (R u - v)
                double rr[3];
                for (int a = 0; a < 3; ++a) rr[a] = R[a * 3] * u[0] + R[a * 3 + 1] * u[1] + R[a * 3 + 2] * u[2] - v[a];
                for (int b = 0; b < 3; ++b) g[b] += 2.0 * (R[b] * rr[0] + R[3 + b] * rr[1] + R[6 + b] * rr[2]);
            }
            if (seg != 2) {  // fit set: 2 G^T v - (2 / n_A) R^T sum_B r
                const double *G = o + 3;
                for (int b = 0; b < 3; ++b)
                    g[b] += 2.0 * (G[b] * v[0] + G[3 + b] * v[1] + G[6 + b] * v[2]) - o[12 + b];
            }
            for (int b = 0; b < 3; ++b) g[b] *= sc;
        }
        Gr[k] = (float)g[0];
        Gr[stride + k] = (float)g[1];
        Gr[2 * stride + k] = (float)g[2];
    }
}

// ---------------------------------------------------------------------------
// Batched RMSD to frame 0 (the RMSD half of an uneven batch; RmsdPlan with
// reference = frame0 and align = true, rmsd/plan.rs:336-354, per trajectory).
// The reference moments of a trajectory are its frame 0's own moments, so
// no separate reference pass is needed.
// ---------------------------------------------------------------------------
// Index t in [0, n) with key(t) <= x < key(t + 1) (keys non-decreasing,
// key(0) <= x); trajectories without units share their successor's key and
// are never returned.
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

__global__ __launch_bounds__(kBlock) void batch_rmsd_moments_kernel(
    const BatchRmsdTraj *__restrict__ traj, int n_traj, long long n_units,
    const float *__restrict__ store, double *__restrict__ partials) {
    const int lane = threadIdx.x & 63;
    const long long unit = (long long)blockIdx.x * (kBlock / 64) +
                           __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    if (unit >= n_units) return;
    const int t = find_last_le(n_traj, unit, [&](int i) { return traj[i].unit0; });
    const BatchRmsdTraj T = traj[t];
    const long long local = unit - T.unit0;
    const int f = (int)(local / T.tiles);
    const int tile = (int)(local - (long long)f * T.tiles);
    const int n_groups = T.stride / 4;
    const int g0 = tile * T.gpt;
    const int g1 = min(n_groups, g0 + T.gpt);
    const float *F0 = store + T.store_off;
    const float *Fx = F0 + (long long)f * 3 * T.stride;
    const float4 *X = (const float4 *)Fx;
    const float4 *Y = (const float4 *)(Fx + T.stride);
    const float4 *Z = (const float4 *)(Fx + 2 * T.stride);
    const float4 *RX = (const float4 *)F0;
    const float4 *RY = (const float4 *)(F0 + T.stride);
    const float4 *RZ = (const float4 *)(F0 + 2 * T.stride);
    double A[kMom];
#pragma unroll
    for (int k = 0; k < kMom; ++k) A[k] = 0.0;
    int g = g0 + lane;
    for (; g + 64 < g1; g += 128) {
        float4 x0 = X[g], y0 = Y[g], z0 = Z[g];
        float4 x1 = X[g + 64], y1 = Y[g + 64], z1 = Z[g + 64];
        float4 a0 = RX[g], b0 = RY[g], c0 = RZ[g];
        float4 a1 = RX[g + 64], b1 = RY[g + 64], c1 = RZ[g + 64];
        acc_group<MODE_SAME>(A, nullptr, 0, x0, y0, z0, a0, b0, c0);
        acc_group<MODE_SAME>(A, nullptr, 0, x1, y1, z1, a1, b1, c1);
    }
    if (g < g1) {
        float4 x0 = X[g], y0 = Y[g], z0 = Z[g];
        float4 a0 = RX[g], b0 = RY[g], c0 = RZ[g];
        acc_group<MODE_SAME>(A, nullptr, 0, x0, y0, z0, a0, b0, c0);
    }
    double *out = partials + unit * kMom;
#pragma unroll
    for (int k = 0; k < kMom; ++k) {
        const double v = wave_sum(A[k]);
        if (lane == 0) out[k] = v;
    }
}

__global__ void batch_rmsd_finalize_kernel(const BatchRmsdTraj *__restrict__ traj, int n_traj,
                                           long long n_out, const double *__restrict__ partials,
                                           float *__restrict__ out32, double *__restrict__ out64,
                                           double *__restrict__ err) {
    const long long o = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (o >= n_out) return;
    const int t = find_last_le(n_traj, o, [&](int i) { return traj[i].out0; });
    const BatchRmsdTraj T = traj[t];
    const int f = (int)(o - T.out0);
    double S[kMom], R0[kMom];
    for (int k = 0; k < kMom; ++k) S[k] = R0[k] = 0.0;
    for (int tile = 0; tile < T.tiles; ++tile) {
        const double *p = partials + (T.unit0 + (long long)f * T.tiles + tile) * kMom;
        const double *q = partials + (T.unit0 + tile) * kMom;  // frame 0: x = y
        for (int k = 0; k < kMom; ++k) {
            S[k] += p[k];
            R0[k] += q[k];
        }
    }
    const double n = (double)T.n;
    double rmsd = 0.0, e = 0.0;
    if (T.n > 0) {
        double cx[3], cy[3];
        for (int d = 0; d < 3; ++d) {
            cx[d] = S[d] / n;
            cy[d] = R0[d] / n;
        }
        double M[9], gx, gy;
        centred(S, R0, R0[12], n, cx, cy, M, &gx, &gy);
        QcpResult qr;
        qcp_solve(M, gx, gy, false, &qr);
        const double resid = gx + gy - 2.0 * qr.lambda;
        const double mag = S[12] + R0[12] + 2.0 * fabs(qr.lambda);
        const double lane_terms = 4.0 * ceil((double)T.gpt / 64.0) + 6.0;
        const double growth = (lane_terms + 8.0 + T.tiles) * 1.1102230246251565e-16;
        const double r2 = resid / n;
        rmsd = r2 > 0.0 ? sqrt(r2) : 0.0;
        const double de = growth * mag / n;
        e = de / (rmsd + sqrt(de));
    }
    out32[o] = (float)rmsd;
    if (out64) out64[o] = rmsd;
    if (err) err[o] = e;
}

}  // namespace

struct wmd_rmsd {
    wmd_ctx *ctx = nullptr;
    int mode = MODE_SAME;
    std::vector<uint32_t> atoms;  // store slot -> atom (kPad = zero padding)
    size_t n_store = 0, stride = 0;
    int seg1 = 0, seg2 = 0;  // group indices where fit-only / RMSD-only segments start
    size_t n_ref_atoms = 0;
    RefMoments rm{};
    DevBuf d_ref, d_atoms, d_partials, d_out32, d_out64, d_err;
    double *scratch = nullptr;  // this device's address of the fixed scratch array
    DevBuf d_bwd_partials, d_bwd_coef;  // split-fit adjoint: moments with second moments, per-frame terms
    // host-chunk pipeline: two staging / device buffers, events guarding reuse
    HostBuf p_stage[2];
    DevBuf p_aos[2], p_soa[2];
    Event p_copied[2], p_used[2];
};

namespace {

int moment_count(int mode) { return mode == MODE_RAW ? 1 : (mode == MODE_SAME ? kMom : 2 * kMom); }

// Tiles per frame of the moments kernel: aim for ~12k waves in flight, each
// lane streaming >= 32 float4 groups.  gpt: groups per tile.
void rmsd_tiling(const wmd_rmsd *h, size_t n_frames, int *n_tiles_out, int *gpt_out) {
    const int n_groups = (int)(h->stride / 4);
    long long want = std::max<long long>(1, (12288 + (long long)n_frames - 1) / (long long)n_frames);
    long long cap = std::max<long long>(1, n_groups / (64 * 32));
    int n_tiles = (int)std::min(want, cap);
    int gpt = (n_groups + n_tiles - 1) / n_tiles;
    n_tiles = (n_groups + gpt - 1) / std::max(gpt, 1);
    *n_tiles_out = std::max(n_tiles, 1);
    *gpt_out = gpt;
}

// partials: [n_frames][n_tiles][moments] doubles (every entry is written by
// the moments kernel before the finalize reads it, so nothing is cleared).
// prof: record HIP events around the kernels when the context profiles.
template <int MODE>
void launch_rmsd(wmd_rmsd *h, const float *d_store, size_t stride, size_t n_frames, float *d_rmsd,
                 double *d_r64, double *d_err, double *d_rot, double *d_cx, hipStream_t st,
                 double *partials, int n_tiles, int gpt, bool prof) {
    const int n_groups = (int)(h->stride / 4);
    long long waves = (long long)n_frames * n_tiles;
    long long blocks = (waves + (kBlock / 64) - 1) / (kBlock / 64);
    WMD_REQUIRE(blocks < (1ll << 31), WMD_ERR_UNSUPPORTED, "too many frames for one launch");
    {
    ProfScope scope(prof ? h->ctx : nullptr, st, "rmsd_moments");
    hipLaunchKernelGGL(rmsd_moments_kernel<MODE>, dim3((unsigned)blocks), dim3(kBlock), 0, st,
                       d_store, 3 * (long long)stride, (int)stride, (int)n_frames,
                       h->d_ref.as<float>(), n_groups, h->seg1, h->seg2, n_tiles, gpt, partials);
    }
    WMD_HIP_CHECK(hipGetLastError());
    double lane_terms = 4.0 * std::ceil((double)gpt / 64.0) + 6.0;
    int fb = (int)((n_frames + 63) / 64);
    {
    ProfScope scope(prof ? h->ctx : nullptr, st, "rmsd_finalize");
    hipLaunchKernelGGL(rmsd_finalize_kernel<MODE>, dim3(fb), dim3(64), 0, st, partials,
                       (int)n_frames, n_tiles, h->rm, lane_terms, d_rmsd, d_r64, d_err, d_rot, d_cx);
    }
    WMD_HIP_CHECK(hipGetLastError());
}

// scratch_partials: NULL (the handle's own partials buffer, grown as needed)
// or the address of the library's fixed scratch array on this device, which
// must hold n_frames x tiles x moments doubles (checked by the caller).
void run_device(wmd_rmsd *h, const float *d_store, size_t stride, size_t n_frames, float *d_rmsd,
                double *d_r64, double *d_err, double *d_rot, double *d_cx, hipStream_t st,
                double *scratch_partials = nullptr) {
    if (n_frames == 0) return;
    WMD_REQUIRE(stride == h->stride, WMD_ERR_INVALID,
                "store stride must equal the plan's stride (" + std::to_string(h->stride) + ")");
    WMD_REQUIRE(d_store && d_rmsd, WMD_ERR_INVALID, "null device buffer");
    int n_tiles = 1, gpt = 1;
    rmsd_tiling(h, n_frames, &n_tiles, &gpt);
    double *partials = scratch_partials;
    if (!partials)
        partials = (double *)h->d_partials.reserve((size_t)n_frames * n_tiles * moment_count(h->mode) * 8);
    const bool prof = scratch_partials == nullptr;  // event creation is left out of the allocation-free path
    if (h->mode == MODE_SAME) launch_rmsd<MODE_SAME>(h, d_store, stride, n_frames, d_rmsd, d_r64, d_err, d_rot, d_cx, st, partials, n_tiles, gpt, prof);
    else if (h->mode == MODE_SPLIT) launch_rmsd<MODE_SPLIT>(h, d_store, stride, n_frames, d_rmsd, d_r64, d_err, d_rot, d_cx, st, partials, n_tiles, gpt, prof);
    else launch_rmsd<MODE_RAW>(h, d_store, stride, n_frames, d_rmsd, d_r64, d_err, d_rot, d_cx, st, partials, n_tiles, gpt, prof);
}

}  // namespace

namespace wmd {

void batch_rmsd_enqueue(hipStream_t st, const BatchRmsdTraj *d_traj, int n_traj,
                        long long n_units, long long n_out, const float *d_store,
                        double *d_partials, float *d_out32, double *d_out64, double *d_err) {
    if (n_units <= 0 || n_out <= 0) return;
    const long long blocks = (n_units + (kBlock / 64) - 1) / (kBlock / 64);
    WMD_REQUIRE(blocks < (1ll << 31), WMD_ERR_UNSUPPORTED, "too many RMSD units for one launch");
    hipLaunchKernelGGL(batch_rmsd_moments_kernel, dim3((unsigned)blocks), dim3(kBlock), 0, st, d_traj,
                       n_traj, n_units, d_store, d_partials);
    WMD_HIP_CHECK(hipGetLastError());
    hipLaunchKernelGGL(batch_rmsd_finalize_kernel, dim3((unsigned)((n_out + 63) / 64)), dim3(64), 0, st,
                       d_traj, n_traj, n_out, (const double *)d_partials, d_out32, d_out64, d_err);
    WMD_HIP_CHECK(hipGetLastError());
}

}  // namespace wmd

extern "C" {

wmd_status wmd_rmsd_create(wmd_ctx *ctx, const uint32_t *fit_sel, size_t n_fit,
                           const uint32_t *rmsd_sel, size_t n_rmsd, const float *ref_xyzw,
                           size_t n_ref_atoms, int align, wmd_rmsd **out) {
    WMD_API_BEGIN
    WMD_REQUIRE(ctx && out, WMD_ERR_INVALID, "null context or output handle");
    *out = nullptr;
    WMD_REQUIRE(n_fit == 0 || fit_sel, WMD_ERR_INVALID, "null fit selection");
    WMD_REQUIRE(ref_xyzw || n_ref_atoms == 0, WMD_ERR_INVALID, "null reference");
    bool same = rmsd_sel == nullptr;
    if (same) n_rmsd = n_fit;
    WMD_REQUIRE(same || n_rmsd == 0 || rmsd_sel, WMD_ERR_INVALID, "null rmsd selection");
    for (size_t i = 0; i < n_fit; ++i)
        WMD_REQUIRE(fit_sel[i] < n_ref_atoms, WMD_ERR_MISMATCH, "selection index out of range of the reference");
    if (!same)
        for (size_t i = 0; i < n_rmsd; ++i)
            WMD_REQUIRE(rmsd_sel[i] < n_ref_atoms, WMD_ERR_MISMATCH, "selection index out of range of the reference");
    if (!align && !same) {
        // unfitted: the residual over the RMSD set alone, the fit set unused (rms.rs:664-666)
        fit_sel = rmsd_sel;
        n_fit = n_rmsd;
        same = true;
    }
    bind(ctx);
    auto *h = new wmd_rmsd();
    try {
        h->ctx = ctx;
        h->n_ref_atoms = n_ref_atoms;
        h->mode = !align ? MODE_RAW : (same ? MODE_SAME : MODE_SPLIT);
        std::vector<uint32_t> both, fit_only, rmsd_only;
        if (same) {
            both.assign(fit_sel, fit_sel + n_fit);  // list order, duplicates kept
        } else {
            // per atom: min(mA, mB) copies in both, the rest in the one-sided segments
            std::map<uint32_t, std::pair<size_t, size_t>> mult;
            for (size_t i = 0; i < n_fit; ++i) mult[fit_sel[i]].first += 1;
            for (size_t i = 0; i < n_rmsd; ++i) mult[rmsd_sel[i]].second += 1;
            for (auto &kv : mult) {
                size_t a = kv.second.first, b = kv.second.second, m = std::min(a, b);
                for (size_t k = 0; k < m; ++k) both.push_back(kv.first);
                for (size_t k = m; k < a; ++k) fit_only.push_back(kv.first);
                for (size_t k = m; k < b; ++k) rmsd_only.push_back(kv.first);
            }
        }
        auto pad4 = [](std::vector<uint32_t> &v) { while (v.size() % 4) v.push_back(kPad); };
        pad4(both);
        pad4(fit_only);
        pad4(rmsd_only);
        h->atoms = both;
        h->seg1 = (int)(h->atoms.size() / 4);
        h->atoms.insert(h->atoms.end(), fit_only.begin(), fit_only.end());
        h->seg2 = (int)(h->atoms.size() / 4);
        h->atoms.insert(h->atoms.end(), rmsd_only.begin(), rmsd_only.end());
        h->n_store = h->atoms.size();
        h->stride = std::max<size_t>(4, round_up(h->n_store, 4));
        while (h->atoms.size() < h->stride) h->atoms.push_back(kPad);
        // reference store + reference moments (FP64, exact products)
        std::vector<float> refs(3 * h->stride, 0.0f);
        RefMoments rm{};
        rm.n_a = (double)n_fit;
        rm.n_b = (double)n_rmsd;
        for (size_t k = 0; k < h->stride; ++k) {
            uint32_t a = h->atoms[k];
            if (a == kPad) continue;
            float x = ref_xyzw[4 * (size_t)a], y = ref_xyzw[4 * (size_t)a + 1], z = ref_xyzw[4 * (size_t)a + 2];
            refs[k] = x;
            refs[h->stride + k] = y;
            refs[2 * h->stride + k] = z;
            int seg = ((int)(k / 4) >= h->seg1) + ((int)(k / 4) >= h->seg2);
            double yy = (double)x * x + (double)y * y + (double)z * z;
            if (seg != 2) {
                rm.sy_a[0] += x; rm.sy_a[1] += y; rm.sy_a[2] += z; rm.syy_a += yy;
            }
            if (seg != 1) {
                rm.sy_b[0] += x; rm.sy_b[1] += y; rm.sy_b[2] += z; rm.syy_b += yy;
            }
        }
        h->rm = rm;
        h->d_ref.reserve(refs.size() * 4);
        WMD_HIP_CHECK(hipMemcpy(h->d_ref.ptr, refs.data(), refs.size() * 4, hipMemcpyHostToDevice));
        h->d_atoms.reserve(h->atoms.size() * 4);
        WMD_HIP_CHECK(hipMemcpy(h->d_atoms.ptr, h->atoms.data(), h->atoms.size() * 4, hipMemcpyHostToDevice));
        h->scratch = scratch_address(ctx->device);  // the context's device is current (bind)
    } catch (...) {
        delete h;
        throw;
    }
    *out = h;
    WMD_API_END
}

wmd_status wmd_rmsd_store_layout(const wmd_rmsd *h, uint32_t *atoms, size_t cap, size_t *n_store,
                                 size_t *stride) {
    WMD_API_BEGIN
    WMD_REQUIRE(h != nullptr, WMD_ERR_INVALID, "null handle");
    if (n_store) *n_store = h->n_store;
    if (stride) *stride = h->stride;
    if (atoms) {
        WMD_REQUIRE(cap >= h->stride, WMD_ERR_INVALID, "output buffer too small");
        std::memcpy(atoms, h->atoms.data(), h->stride * 4);
    }
    WMD_API_END
}

wmd_status wmd_rmsd_run(wmd_rmsd *h, const float *coords, size_t n_atoms, size_t n_frames,
                        float *out_rmsd, double *out_err) {
    WMD_API_BEGIN
    WMD_REQUIRE(h != nullptr, WMD_ERR_INVALID, "null handle");
    if (n_frames == 0) return WMD_OK;
    WMD_REQUIRE(coords && out_rmsd, WMD_ERR_INVALID, "null buffer");
    for (uint32_t a : h->atoms)
        WMD_REQUIRE(a == kPad || a < n_atoms, WMD_ERR_MISMATCH,
                    "selection index out of range for the chunk's atom count");
    bind(h->ctx);
    hipStream_t st = h->ctx->stream;
    hipStream_t cs = copy_stream(h->ctx);
    // three overlapping stages: host copies into pinned staging buffer b, the
    // H2D copy of b on the copy stream, the gather and kernels on st
    // (p_copied[b]: the H2D of b is done; p_used[b]: the gather has read it)
    const size_t fbytes = n_atoms * 16;
    const size_t batch = staging_frames(h->ctx, fbytes, n_frames);
    for (int b = 0; b < 2; ++b) {
        if (h->p_stage[b].bytes < batch * fbytes || h->p_soa[b].bytes < batch * 3 * h->stride * 4) {
            WMD_HIP_CHECK(hipStreamSynchronize(st));  // growing frees buffers queued work may use
            WMD_HIP_CHECK(hipStreamSynchronize(cs));
            h->p_stage[b].reserve(batch * fbytes);
            h->p_aos[b].reserve(batch * fbytes);
            h->p_soa[b].reserve(batch * 3 * h->stride * 4);
        }
    }
    if (h->d_out32.bytes < n_frames * 4 || (out_err && h->d_err.bytes < n_frames * 8))
        WMD_HIP_CHECK(hipStreamSynchronize(st));
    float *d_o = (float *)h->d_out32.reserve(n_frames * 4);
    double *d_e = out_err ? (double *)h->d_err.reserve(n_frames * 8) : nullptr;
    WMD_HIP_CHECK(hipEventRecord(h->p_used[0].get(), st));
    WMD_HIP_CHECK(hipEventRecord(h->p_used[1].get(), st));
    size_t k = 0;
    for (size_t f0 = 0; f0 < n_frames; f0 += batch, ++k) {
        const int b = (int)(k & 1);
        const size_t nf = std::min(batch, n_frames - f0);
        h->p_copied[b].wait();  // the copy out of this staging buffer two batches ago is done
        host_copy(h->p_stage[b].ptr, coords + f0 * n_atoms * 4, nf * fbytes);
        WMD_HIP_CHECK(hipStreamWaitEvent(cs, h->p_used[b].get(), 0));
        WMD_HIP_CHECK(hipMemcpyAsync(h->p_aos[b].ptr, h->p_stage[b].ptr, nf * fbytes,
                                     hipMemcpyHostToDevice, cs));
        WMD_HIP_CHECK(hipEventRecord(h->p_copied[b].get(), cs));
        WMD_HIP_CHECK(hipStreamWaitEvent(st, h->p_copied[b].get(), 0));
        float *d_soa = (float *)h->p_soa[b].ptr;
        long long total = (long long)h->stride * nf;
        int g = (int)std::min<long long>(65536, (total + kBlock - 1) / kBlock);
        hipLaunchKernelGGL(rmsd_gather_kernel, dim3(g), dim3(kBlock), 0, st, (const float4 *)h->p_aos[b].ptr,
                           (long long)n_atoms, h->d_atoms.as<unsigned>(), (int)h->stride, (int)nf, d_soa);
        WMD_HIP_CHECK(hipGetLastError());
        WMD_HIP_CHECK(hipEventRecord(h->p_used[b].get(), st));
        run_device(h, d_soa, h->stride, nf, d_o + f0, nullptr, d_e ? d_e + f0 : nullptr, nullptr, nullptr, st);
    }
    WMD_HIP_CHECK(hipMemcpyAsync(out_rmsd, d_o, n_frames * 4, hipMemcpyDeviceToHost, st));
    if (out_err) WMD_HIP_CHECK(hipMemcpyAsync(out_err, d_e, n_frames * 8, hipMemcpyDeviceToHost, st));
    WMD_HIP_CHECK(hipStreamSynchronize(st));
    WMD_API_END
}

wmd_status wmd_rmsd_run_device(wmd_rmsd *h, const float *d_store, size_t stride, size_t n_frames,
                               float *d_rmsd, double *d_rmsd64, double *d_err, double *d_rot,
                               double *d_cx, void *stream) {
    WMD_API_BEGIN
    WMD_REQUIRE(h != nullptr, WMD_ERR_INVALID, "null handle");
    bind(h->ctx);
    run_device(h, d_store, stride, n_frames, d_rmsd, d_rmsd64, d_err, d_rot, d_cx,
               pick_stream(h->ctx, stream));
    WMD_API_END
}

wmd_status wmd_rmsd_run_device_noalloc(wmd_rmsd *h, const float *d_store, size_t stride,
                                       size_t n_frames, float *d_rmsd, double *d_rmsd64,
                                       double *d_err, double *d_rot, double *d_cx, void *stream) {
    WMD_API_BEGIN
    WMD_REQUIRE(h != nullptr, WMD_ERR_INVALID, "null handle");
    if (n_frames == 0) return WMD_OK;
    WMD_REQUIRE(h->scratch != nullptr, WMD_ERR_DEVICE, "RMSD scratch array not resolved");
    int n_tiles = 1, gpt = 1;
    rmsd_tiling(h, n_frames, &n_tiles, &gpt);
    const size_t need = n_frames * (size_t)n_tiles * (size_t)moment_count(h->mode) * sizeof(double);
    WMD_REQUIRE(need <= kScratchBytes, WMD_ERR_UNSUPPORTED,
                "allocation-free RMSD: " + std::to_string(n_frames) + " frames need " +
                    std::to_string(need) + " bytes of scratch, more than the fixed " +
                    std::to_string(kScratchBytes) + "; pass fewer frames per call");
    bind(h->ctx);
    run_device(h, d_store, stride, n_frames, d_rmsd, d_rmsd64, d_err, d_rot, d_cx,
               pick_stream(h->ctx, stream), h->scratch);
    WMD_API_END
}

wmd_status wmd_rmsd_backward_device(wmd_rmsd *h, const float *d_store, size_t stride,
                                    size_t n_frames, const double *d_rmsd64, const double *d_rot,
                                    const double *d_cx, const float *d_gout, float *d_grad,
                                    void *stream) {
    WMD_API_BEGIN
    WMD_REQUIRE(h != nullptr, WMD_ERR_INVALID, "null handle");
    if (n_frames == 0) return WMD_OK;
    WMD_REQUIRE(stride == h->stride, WMD_ERR_INVALID,
                "store stride must equal the plan's stride (" + std::to_string(h->stride) + ")");
    WMD_REQUIRE(d_store && d_rmsd64 && d_gout && d_grad, WMD_ERR_INVALID, "null device buffer");
    WMD_REQUIRE(h->mode == MODE_RAW || (d_rot && d_cx), WMD_ERR_INVALID,
                "the fitted RMSD backward needs the rotations and centroids of the forward run");
    bind(h->ctx);
    hipStream_t st = pick_stream(h->ctx, stream);
    const double n = h->mode == MODE_SAME ? h->rm.n_a : h->rm.n_b;
    double cy[3] = {0, 0, 0};
    if (h->rm.n_a > 0)
        for (int d = 0; d < 3; ++d) cy[d] = h->rm.sy_a[d] / h->rm.n_a;
    const long long total = (long long)n_frames * (long long)stride;
    const int g = (int)std::min<long long>(65536, (total + kBlock - 1) / kBlock);
    ProfScope prof(h->ctx, st, "rmsd_backward");
    if (h->mode == MODE_SPLIT) {
        // moments again, with the RMSD set's second moments, then the
        // per-frame adjoint terms, then the per-slot gradient
        int n_tiles = 1, gpt = 1;
        rmsd_tiling(h, n_frames, &n_tiles, &gpt);
        constexpr int NV = 2 * kMom + 6;
        double *partials = (double *)h->d_bwd_partials.reserve((size_t)n_frames * n_tiles * NV * 8);
        double *coef = (double *)h->d_bwd_coef.reserve((size_t)n_frames * 16 * 8);
        const long long waves = (long long)n_frames * n_tiles;
        const long long blocks = (waves + (kBlock / 64) - 1) / (kBlock / 64);
        WMD_REQUIRE(blocks < (1ll << 31), WMD_ERR_UNSUPPORTED, "too many frames for one launch");
        hipLaunchKernelGGL((rmsd_moments_kernel<MODE_SPLIT, true>), dim3((unsigned)blocks), dim3(kBlock), 0, st,
                           d_store, 3 * (long long)stride, (int)stride, (int)n_frames, h->d_ref.as<float>(),
                           (int)(h->stride / 4), h->seg1, h->seg2, n_tiles, gpt, partials);
        WMD_HIP_CHECK(hipGetLastError());
        hipLaunchKernelGGL(rmsd_split_adjoint_kernel, dim3((unsigned)((n_frames + 63) / 64)), dim3(64), 0, st,
                           (const double *)partials, (int)n_frames, n_tiles, h->rm, d_rot, d_rmsd64, d_gout, coef);
        WMD_HIP_CHECK(hipGetLastError());
        hipLaunchKernelGGL(rmsd_split_backward_kernel, dim3(g), dim3(kBlock), 0, st, d_store, (int)stride,
                           (int)n_frames, h->d_ref.as<float>(), h->d_atoms.as<unsigned>(), h->seg1, h->seg2,
                           cy[0], cy[1], cy[2], d_rot, (const double *)coef, d_grad);
        WMD_HIP_CHECK(hipGetLastError());
        return WMD_OK;
    }
    if (h->mode == MODE_SAME)
        hipLaunchKernelGGL(rmsd_backward_kernel<true>, dim3(g), dim3(kBlock), 0, st, d_store, (int)stride,
                           (int)n_frames, h->d_ref.as<float>(), h->d_atoms.as<unsigned>(), cy[0], cy[1], cy[2],
                           n, d_rmsd64, d_rot, d_cx, d_gout, d_grad);
    else
        hipLaunchKernelGGL(rmsd_backward_kernel<false>, dim3(g), dim3(kBlock), 0, st, d_store, (int)stride,
                           (int)n_frames, h->d_ref.as<float>(), h->d_atoms.as<unsigned>(), cy[0], cy[1], cy[2],
                           n, d_rmsd64, d_rot, d_cx, d_gout, d_grad);
    WMD_HIP_CHECK(hipGetLastError());
    WMD_API_END
}

void wmd_rmsd_destroy(wmd_rmsd *h) {
    if (!h) return;
    (void)hipSetDevice(h->ctx->device);
    // work may still be queued on a caller's stream: wait for the whole device
    (void)hipDeviceSynchronize();
    delete h;
}

wmd_status wmd_qcp_host(const double *m, double gx, double gy, double n, double *rmsd, double *rot,
                        double *lambda, int *iters) {
    WMD_API_BEGIN
    WMD_REQUIRE(m != nullptr, WMD_ERR_INVALID, "null moments");
    QcpResult qr;
    qcp_solve(m, gx, gy, rot != nullptr, &qr);
    double r2 = n > 0 ? (gx + gy - 2.0 * qr.lambda) / n : 0.0;
    if (rmsd) *rmsd = r2 > 0.0 ? std::sqrt(r2) : 0.0;
    if (rot) for (int k = 0; k < 9; ++k) rot[k] = qr.rot[k];
    if (lambda) *lambda = qr.lambda;
    if (iters) *iters = qr.iters;
    WMD_API_END
}

}  // extern "C"
