// AlignPlan and SuperposePlan on the device
// (crates/traj-engine/src/plans/geometry/align/plan_exec.rs:76-465): per frame,
// the optimal rotation R and translation t = cy - R cx that superpose the
// selection onto the reference (kabsch_rotation_weighted,
// geometry/geometry_math.rs:235-297, optionally mass-weighted), and for
// SuperposePlan R x + t applied to every atom of the frame.
//
// The reference's CPU path is all f64 (weighted centroids, H, SVD) and its
// GPU path f32 covariance sums; this engine follows the CPU path: one
// workgroup per frame reduces the weighted centroid and then the weighted
// cross-covariance in f64, and the FP64 QCP solve (wmd_qcp.hpp) supplies R.
// The transform of SuperposePlan is evaluated in f64 in the reference's
// operation order (plan_exec.rs:399-405) and cast to f32, by the same
// workgroup right after the fit: from registers and LDS holding the whole
// frame for frames of up to 20,480 atoms, else from a second read of it.
#include <algorithm>
#include <cstring>
#include <memory>

#include "wmd_internal.hpp"
#include "wmd_qcp.hpp"

namespace {
using namespace wmd;

constexpr int kAlBlock = 256;

// Sum of `CNT` f64 values over a workgroup of WAVES waves; every thread gets
// the totals (waves added in index order).
template <int CNT, int WAVES>
__device__ __forceinline__ void al_block_sum(double *v, double *red) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
    for (int c = 0; c < CNT; ++c)
        for (int off = 32; off > 0; off >>= 1) v[c] += __shfl_down(v[c], off, 64);
    __syncthreads();
    if (lane == 0)
#pragma unroll
        for (int c = 0; c < CNT; ++c) red[w * CNT + c] = v[c];
    __syncthreads();
#pragma unroll
    for (int c = 0; c < CNT; ++c) {
        double t = red[c];
        for (int q = 1; q < WAVES; ++q) t += red[q * CNT + c];
        v[c] = t;
    }
}

// From summed moments m (Sx[3], W, Sxy[9], Sxx): cx, H, R (QCP, or I for
// norotate) and t = cy - R cx, written to frame f's row and to rt_s (R
// row-major, t).  Called by one thread.
__device__ __forceinline__ void al_fit_row(const double *m, bool weighted, int n, double rs0,
                                           double rs1, double rs2, double ref_g, double cy0,
                                           double cy1, double cy2, int norotate, long long f,
                                           double *__restrict__ rt64, float *__restrict__ rt32,
                                           double *rt_s) {
    const double wsum = weighted ? m[3] : (double)n;
    double cx[3] = {0.0, 0.0, 0.0};
    if (wsum != 0.0)
        for (int d = 0; d < 3; ++d) cx[d] = m[d] / wsum;
    const double rs[3] = {rs0, rs1, rs2};
    double H[10];
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int c = 0; c < 3; ++c) H[3 * a + c] = m[4 + 3 * a + c] - cx[a] * rs[c];
    H[9] = m[13] - wsum * (cx[0] * cx[0] + cx[1] * cx[1] + cx[2] * cx[2]);
    double R[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
    if (!norotate && n > 0) {
        QcpResult res;
        qcp_solve(H, H[9] > 0.0 ? H[9] : 0.0, ref_g, true, &res);
        for (int i = 0; i < 9; ++i) R[i] = res.rot[i];
    }
    const double cy[3] = {cy0, cy1, cy2};
    double out[12];
    for (int i = 0; i < 9; ++i) out[i] = R[i];
    for (int a = 0; a < 3; ++a)
        out[9 + a] = cy[a] - (R[3 * a] * cx[0] + R[3 * a + 1] * cx[1] + R[3 * a + 2] * cx[2]);
    for (int i = 0; i < 12; ++i) {
        rt64[f * 12 + i] = out[i];
        if (rt32) rt32[f * 12 + i] = (float)out[i];
        rt_s[i] = out[i];
    }
}

// R x + t for one atom, f64 in the reference's order (plan_exec.rs:399-405),
// cast to f32.
__device__ __forceinline__ void al_transform(const double *q, float x, float y, float z, float *o) {
    const double px = (double)x, py = (double)y, pz = (double)z;
    o[0] = (float)(q[0] * px + q[1] * py + q[2] * pz + q[9]);
    o[1] = (float)(q[3] * px + q[4] * py + q[5] * pz + q[10]);
    o[2] = (float)(q[6] * px + q[7] * py + q[8] * pz + q[11]);
}

// One workgroup per frame (grid-stride), one pass over the selection:
// W = sum w, Sx = sum w x, Sxy = sum w x (y - cy)^T and Sxx = sum w |x|^2 in
// f64, then cx = Sx / W (centroid_weighted, geometry_math.rs:202-233),
// H = sum w (x - cx)(y - cy)^T = Sxy - cx (sum w (y - cy))^T (the last sum is
// a host constant, zero up to rounding), gx = Sxx - W |cx|^2 (only the start of
// the QCP Newton iteration), R (QCP, or I for norotate) and t = cy - R cx.
// aos: [F][n_src] xyzw; sel [n]; wts [n] or NULL; ref_c [n][3] = y - cy
// (host f64, centred), ref_s = sum w (y - cy), ref_g = sum w |y - cy|^2, cy [3].
// rt64 [F][12] (R row-major, t), rt32 [F][12] or NULL; xyz NULL or
// [F][n_src][3] = R x + t for every atom (f64, reference order, cast to f32).
__global__ __launch_bounds__(kAlBlock) void al_fit_kernel(
    const float4 *__restrict__ aos, long long n_src, int n_frames, const unsigned *__restrict__ sel,
    const float *__restrict__ wts, int n, const double *__restrict__ ref_c, double rs0, double rs1,
    double rs2, double ref_g, double cy0, double cy1, double cy2, int norotate,
    double *__restrict__ rt64, float *__restrict__ rt32, float *__restrict__ xyz) {
    __shared__ double red[4 * 14];
    __shared__ double rt_s[12];
    for (int f = blockIdx.x; f < n_frames; f += gridDim.x) {
        const float4 *x = aos + (long long)f * n_src;
        // m: Sx[3], W, Sxy[9], Sxx
        double m[14];
#pragma unroll
        for (int c = 0; c < 14; ++c) m[c] = 0.0;
        for (int k = threadIdx.x; k < n; k += kAlBlock) {
            const float4 p = x[sel[k]];
            const double w = wts ? (double)wts[k] : 1.0;
            const double xv[3] = {(double)p.x, (double)p.y, (double)p.z};
            const double yr[3] = {ref_c[3 * k], ref_c[3 * k + 1], ref_c[3 * k + 2]};
#pragma unroll
            for (int a = 0; a < 3; ++a) m[a] += xv[a] * w;
            m[3] += w;
#pragma unroll
            for (int a = 0; a < 3; ++a)
#pragma unroll
                for (int c = 0; c < 3; ++c) m[4 + 3 * a + c] += xv[a] * yr[c] * w;
            m[13] += (xv[0] * xv[0] + xv[1] * xv[1] + xv[2] * xv[2]) * w;
        }
        al_block_sum<14, kAlBlock / 64>(m, red);
        if (threadIdx.x == 0)
            al_fit_row(m, wts != nullptr, n, rs0, rs1, rs2, ref_g, cy0, cy1, cy2, norotate, f, rt64,
                       rt32, rt_s);
        __syncthreads();
        if (xyz) {
            // SuperposePlan: every atom of this frame while its lines are still cached
            double q[12];
#pragma unroll
            for (int i = 0; i < 12; ++i) q[i] = rt_s[i];
            float *o = xyz + (long long)f * n_src * 3;
            constexpr int U = 4;  // loads in flight per thread
            for (long long a0 = threadIdx.x; a0 < n_src; a0 += U * kAlBlock) {
                float4 p[U];
#pragma unroll
                for (int u = 0; u < U; ++u) {
                    const long long a = a0 + (long long)u * kAlBlock;
                    p[u] = a < n_src ? x[a] : make_float4(0.f, 0.f, 0.f, 0.f);
                }
#pragma unroll
                for (int u = 0; u < U; ++u) {
                    const long long a = a0 + (long long)u * kAlBlock;
                    if (a >= n_src) break;
                    al_transform(q, p[u].x, p[u].y, p[u].z, o + 3 * a);
                }
            }
        }
        __syncthreads();  // red and rt_s are reused by the next frame
    }
}

// SuperposePlan with the frame held on chip: each of the workgroup's
// kRegBlock threads loads atoms t + u kRegBlock (u < UR) into registers and
// atoms t + (UR + v) kRegBlock (v < UL) into LDS (SoA x, y, z planes) before
// the fit, all as coalesced float4 rows; the fit gathers the selection
// exactly as al_fit_kernel does (its lines were just fetched, so the gather
// reads them from L2), and the transform works from registers and LDS: each
// frame crosses HBM once in and once out.  The LDS part keeps the registers
// live across the QCP solve within two waves per SIMD (256 VGPRs).  Same
// moments, rotation and outputs as al_fit_kernel.  One workgroup per CU,
// frames grid-strided; n_src <= (UR + UL) kRegBlock; dynamic LDS
// UL kRegBlock x 12 bytes.
constexpr int kRegBlock = 512;
constexpr int kRegU = 14;  // register-resident atoms per thread
constexpr int kRegSel = 4;  // selected atoms per thread fetched ahead of the frame
// Raw buffer descriptor word 3 (32-bit data format, untyped access); with a
// zero stride the hardware bounds every access by num_records bytes, the
// whole offset being in the VGPR operand.
constexpr int kBufWord3 = 0x00020000;
typedef float v3f __attribute__((ext_vector_type(3)));

template <int UR, int UL>
__global__ __launch_bounds__(kRegBlock, 1) void al_superpose_reg_kernel(
    const float4 *__restrict__ aos, long long n_src, int n_frames, const unsigned *__restrict__ sel,
    const float *__restrict__ wts, int n, const double *__restrict__ ref_c, double rs0, double rs1,
    double rs2, double ref_g, double cy0, double cy1, double cy2, int norotate,
    double *__restrict__ rt64, float *__restrict__ rt32, float *__restrict__ xyz) {
    __shared__ double red[(kRegBlock / 64) * 14];
    __shared__ double rt_s[12];
    extern __shared__ float lxyz[];  // [3][UL * kRegBlock]
    for (int f = blockIdx.x; f < n_frames; f += gridDim.x) {
        const float4 *x = aos + (long long)f * n_src;
        // raw buffer loads bounded by the frame (atoms past n_src read as 0
        // and are never written out): no branches, all loads in flight at once
        const __amdgpu_buffer_rsrc_t in =
            __builtin_amdgcn_make_buffer_rsrc((void *)x, (short)0, (int)(n_src * 16), kBufWord3);
        // the fit's first kRegSel selected atoms per thread are requested
        // before the frame itself: the frame's rows then find those lines in
        // L2 (gathered after the rows, they came from HBM a second time)
        float4 gs[kRegSel];
        if (n > 0) {
            // clamped indices, no branches: the index loads, then the atom
            // loads, each group in flight together
            unsigned si[kRegSel];
#pragma unroll
            for (int u = 0; u < kRegSel; ++u) si[u] = sel[min((int)threadIdx.x + u * kRegBlock, n - 1)];
#pragma unroll
            for (int u = 0; u < kRegSel; ++u) gs[u] = x[si[u]];
        }
        float px[UR], py[UR], pz[UR];
#pragma unroll
        for (int u = 0; u < UR; ++u) {
            const v3f p = __builtin_amdgcn_raw_buffer_load_b96(in, (int)(threadIdx.x + u * kRegBlock) * 16, 0, 0);
            px[u] = p.x;
            py[u] = p.y;
            pz[u] = p.z;
        }
#pragma unroll
        for (int v = 0; v < UL; ++v) {
            const v3f p =
                __builtin_amdgcn_raw_buffer_load_b96(in, (int)(threadIdx.x + (UR + v) * kRegBlock) * 16, 0, 0);
            const int l = v * kRegBlock + threadIdx.x;
            lxyz[l] = p.x;
            lxyz[UL * kRegBlock + l] = p.y;
            lxyz[2 * UL * kRegBlock + l] = p.z;
        }
        double m[14];
#pragma unroll
        for (int c = 0; c < 14; ++c) m[c] = 0.0;
        for (int k = threadIdx.x, u = 0; k < n; k += kRegBlock, ++u) {
            float4 p;
            if (u < kRegSel) {
#pragma unroll
                for (int v = 0; v < kRegSel; ++v)
                    if (v == u) p = gs[v];
            } else {
                p = x[sel[k]];
            }
            const double w = wts ? (double)wts[k] : 1.0;
            const double xv[3] = {(double)p.x, (double)p.y, (double)p.z};
            const double yr[3] = {ref_c[3 * k], ref_c[3 * k + 1], ref_c[3 * k + 2]};
#pragma unroll
            for (int a = 0; a < 3; ++a) m[a] += xv[a] * w;
            m[3] += w;
#pragma unroll
            for (int a = 0; a < 3; ++a)
#pragma unroll
                for (int c = 0; c < 3; ++c) m[4 + 3 * a + c] += xv[a] * yr[c] * w;
            m[13] += (xv[0] * xv[0] + xv[1] * xv[1] + xv[2] * xv[2]) * w;
        }
        // wave sums, then thread 0 adds the waves in index order (a loop that
        // is not unrolled: all 8 x 14 partials live at once would not fit
        // beside the frame)
        const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
        for (int c = 0; c < 14; ++c)
            for (int off = 32; off > 0; off >>= 1) m[c] += __shfl_down(m[c], off, 64);
        if (lane == 0)
#pragma unroll
            for (int c = 0; c < 14; ++c) red[wv * 14 + c] = m[c];
        __syncthreads();
        if (threadIdx.x == 0) {
#pragma unroll
            for (int c = 0; c < 14; ++c) m[c] = red[c];
#pragma unroll 1
            for (int w = 1; w < kRegBlock / 64; ++w)
#pragma unroll
                for (int c = 0; c < 14; ++c) m[c] += red[w * 14 + c];
            al_fit_row(m, wts != nullptr, n, rs0, rs1, rs2, ref_g, cy0, cy1, cy2, norotate, f, rt64,
                       rt32, rt_s);
        }
        __syncthreads();
        double q[12];
#pragma unroll
        for (int i = 0; i < 12; ++i) q[i] = rt_s[i];
        const __amdgpu_buffer_rsrc_t out = __builtin_amdgcn_make_buffer_rsrc(
            (void *)(xyz + (long long)f * n_src * 3), (short)0, (int)(n_src * 12), kBufWord3);
#pragma unroll
        for (int u = 0; u < UR; ++u) {
            float o[3];
            al_transform(q, px[u], py[u], pz[u], o);
            const v3f ov = {o[0], o[1], o[2]};
            __builtin_amdgcn_raw_buffer_store_b96(ov, out, (int)(threadIdx.x + u * kRegBlock) * 12, 0, 0);
        }
#pragma unroll 2
        for (int v = 0; v < UL; ++v) {
            const int l = v * kRegBlock + threadIdx.x;
            float o[3];
            al_transform(q, lxyz[l], lxyz[UL * kRegBlock + l], lxyz[2 * UL * kRegBlock + l], o);
            const v3f ov = {o[0], o[1], o[2]};
            __builtin_amdgcn_raw_buffer_store_b96(ov, out, (int)(threadIdx.x + (UR + v) * kRegBlock) * 12, 0, 0);
        }
        __syncthreads();  // red, rt_s and the LDS frame part are reused by the next frame
    }
}

}  // namespace

struct wmd_align {
    wmd_ctx *ctx = nullptr;
    std::vector<uint32_t> sel;
    std::vector<float> wts;  // empty: unweighted
    bool norotate = false;
    bool have_ref = false;
    double cy[3] = {0, 0, 0};
    double ref_s[3] = {0, 0, 0};  // sum w (y - cy): zero up to rounding
    double ref_g = 0.0;
    DevBuf d_sel, d_wts, d_refc;
    // host-chunk pipeline (wmd_align_run): two of each
    DevBuf p_aos[2], p_rt64[2], p_rt32[2], p_xyz[2];
    HostBuf p_in[2], p_out[2], p_rt[2];
    Event p_copied[2], p_used[2], p_out_ready[2];
};

namespace {

// kabsch_rotation_weighted's reference half (geometry_math.rs:248-249): the
// weighted centroid cy and the centred reference y - cy, in f64 on the host.
void al_set_reference(wmd_align *h, const float *ref_xyzw, size_t ref_stride_atoms,
                      const uint32_t *idx) {
    const size_t n = h->sel.size();
    std::vector<double> yc(std::max<size_t>(n, 1) * 3, 0.0);
    double s[3] = {0, 0, 0}, wsum = 0.0;
    const bool weighted = !h->wts.empty();
    for (size_t i = 0; i < n; ++i) {
        const float *p = ref_xyzw + (size_t)(idx ? idx[i] : i) * 4;
        if (weighted) {
            const double w = (double)h->wts[i];
            for (int d = 0; d < 3; ++d) s[d] += (double)p[d] * w;
            wsum += w;
        } else {
            for (int d = 0; d < 3; ++d) s[d] += (double)p[d];
        }
    }
    if (!weighted) wsum = (double)n;
    for (int d = 0; d < 3; ++d) h->cy[d] = (n == 0 || wsum == 0.0) ? 0.0 : s[d] / wsum;
    double g = 0.0, rs[3] = {0, 0, 0};
    for (size_t i = 0; i < n; ++i) {
        const float *p = ref_xyzw + (size_t)(idx ? idx[i] : i) * 4;
        const double w = weighted ? (double)h->wts[i] : 1.0;
        double e = 0.0;
        for (int d = 0; d < 3; ++d) {
            yc[3 * i + d] = (double)p[d] - h->cy[d];
            e += yc[3 * i + d] * yc[3 * i + d];
            rs[d] += yc[3 * i + d] * w;
        }
        g += e * w;
    }
    for (int d = 0; d < 3; ++d) h->ref_s[d] = rs[d];
    (void)ref_stride_atoms;
    h->ref_g = g;
    h->d_refc.reserve(yc.size() * 8);
    WMD_HIP_CHECK(hipMemcpy(h->d_refc.ptr, yc.data(), yc.size() * 8, hipMemcpyHostToDevice));
    h->have_ref = true;
}

void al_enqueue(wmd_align *h, const float4 *d_aos, size_t n_atoms, size_t F, double *d_rt64,
                float *d_rt32, float *d_xyz, hipStream_t st) {
    for (uint32_t a : h->sel)
        WMD_REQUIRE(a < n_atoms, WMD_ERR_MISMATCH,
                    "selection index " + std::to_string(a) + " out of range for " +
                        std::to_string(n_atoms) + " atoms");
    WMD_REQUIRE(h->have_ref, WMD_ERR_MISMATCH, "reference not initialized");
    if (F == 0) return;
    {
        ProfScope prof(h->ctx, st, d_xyz ? "align_superpose" : "align_fit");
        const float *wts = h->wts.empty() ? nullptr : h->d_wts.as<float>();
        // frame on chip: the smallest variant that holds n_atoms and whose
        // LDS part fits the device
        const size_t per_ul = (size_t)kRegBlock * 12;
        int ul = -1;
        if (d_xyz) {
            for (int c : {0, 8, 16, 26})
                if (n_atoms <= (size_t)(kRegU + c) * kRegBlock && (size_t)c * per_ul <= h->ctx->lds_per_block) {
                    ul = c;
                    break;
                }
        }
        if (ul >= 0) {
            const unsigned grid = (unsigned)std::min<size_t>(F, (size_t)h->ctx->cu_count);
            auto launch = [&](auto kern) {
                hipLaunchKernelGGL(kern, dim3(grid), dim3(kRegBlock), (size_t)ul * per_ul, st, d_aos,
                                   (long long)n_atoms, (int)F, h->d_sel.as<unsigned>(), wts, (int)h->sel.size(),
                                   h->d_refc.as<double>(), h->ref_s[0], h->ref_s[1], h->ref_s[2], h->ref_g,
                                   h->cy[0], h->cy[1], h->cy[2], (int)h->norotate, d_rt64, d_rt32, d_xyz);
            };
            if (n_atoms <= 8u * kRegBlock) launch(al_superpose_reg_kernel<8, 0>);
            else if (ul == 0) launch(al_superpose_reg_kernel<kRegU, 0>);
            else if (ul == 8) launch(al_superpose_reg_kernel<kRegU, 8>);
            else if (ul == 16) launch(al_superpose_reg_kernel<kRegU, 16>);
            else launch(al_superpose_reg_kernel<kRegU, 26>);
        } else {
            hipLaunchKernelGGL(al_fit_kernel, dim3((unsigned)std::min<size_t>(F, 65536)), dim3(kAlBlock), 0,
                               st, d_aos, (long long)n_atoms, (int)F, h->d_sel.as<unsigned>(), wts,
                               (int)h->sel.size(), h->d_refc.as<double>(), h->ref_s[0], h->ref_s[1],
                               h->ref_s[2], h->ref_g, h->cy[0], h->cy[1], h->cy[2], (int)h->norotate,
                               d_rt64, d_rt32, d_xyz);
        }
        WMD_HIP_CHECK(hipGetLastError());
    }
}

}  // namespace

extern "C" {

wmd_status wmd_align_create(wmd_ctx *ctx, const uint32_t *sel, size_t n_sel, const float *masses,
                            const float *ref_xyzw, size_t n_ref_atoms, int norotate,
                            wmd_align **out) {
    WMD_API_BEGIN
    WMD_REQUIRE(ctx && out, WMD_ERR_INVALID, "null context or output handle");
    *out = nullptr;
    WMD_REQUIRE(n_sel == 0 || sel, WMD_ERR_INVALID, "null selection");
    WMD_REQUIRE(n_sel < (1u << 30), WMD_ERR_UNSUPPORTED, "selection too large");
    bind(ctx);
    auto h = std::make_unique<wmd_align>();
    h->ctx = ctx;
    h->norotate = norotate != 0;
    h->sel.assign(sel, sel + n_sel);
    if (masses) h->wts.assign(masses, masses + n_sel);
    const size_t nb = std::max<size_t>(n_sel, 1) * 4;
    h->d_sel.reserve(nb);
    if (n_sel) WMD_HIP_CHECK(hipMemcpy(h->d_sel.ptr, sel, n_sel * 4, hipMemcpyHostToDevice));
    if (masses) {
        h->d_wts.reserve(nb);
        if (n_sel) WMD_HIP_CHECK(hipMemcpy(h->d_wts.ptr, masses, n_sel * 4, hipMemcpyHostToDevice));
    }
    if (ref_xyzw) {
        for (uint32_t a : h->sel)
            WMD_REQUIRE(a < n_ref_atoms, WMD_ERR_MISMATCH, "selection index out of range for the reference");
        al_set_reference(h.get(), ref_xyzw, n_ref_atoms, sel);
    }
    *out = h.release();
    WMD_API_END
}

wmd_status wmd_align_set_reference(wmd_align *h, const float *ref_xyzw, size_t n_ref_atoms) {
    WMD_API_BEGIN
    WMD_REQUIRE(h && ref_xyzw, WMD_ERR_INVALID, "null handle or reference");
    for (uint32_t a : h->sel)
        WMD_REQUIRE(a < n_ref_atoms, WMD_ERR_MISMATCH, "selection index out of range for the reference");
    bind(h->ctx);
    al_set_reference(h, ref_xyzw, n_ref_atoms, h->sel.data());
    WMD_API_END
}

wmd_status wmd_align_run(wmd_align *h, const float *coords_xyzw, size_t n_atoms, size_t n_frames,
                         float *out_rt, float *out_xyz) {
    WMD_API_BEGIN
    WMD_REQUIRE(h != nullptr, WMD_ERR_INVALID, "null handle");
    if (n_frames == 0) return WMD_OK;
    WMD_REQUIRE(coords_xyzw != nullptr, WMD_ERR_INVALID, "null coords");
    bind(h->ctx);
    hipStream_t st = h->ctx->stream;
    hipStream_t cs = copy_stream(h->ctx);
    // Pipeline over batches of one staging buffer of chunk (48 MB unless the
    // context sets another size): host copy into pinned
    // staging b, its H2D on the copy stream, the fit (and transform) and the
    // D2H of the results into pinned staging on the context's stream, and the
    // host copy of batch k - 2's results out while batch k is staged.
    // p_copied[b]: the H2D of b is done; p_used[b]: the kernel has read the
    // device copy of b; p_out_ready[b]: b's results are in pinned memory.
    const size_t fbytes = n_atoms * 16, obytes = n_atoms * 12;
    const size_t batch = staging_frames(h->ctx, fbytes, n_frames);
    bool grow = false;
    for (int b = 0; b < 2; ++b)
        grow |= h->p_in[b].bytes < batch * fbytes || h->p_rt[b].bytes < batch * 48 ||
                (out_xyz && h->p_out[b].bytes < batch * obytes);
    if (grow) {
        WMD_HIP_CHECK(hipStreamSynchronize(st));  // growing frees buffers queued work may use
        WMD_HIP_CHECK(hipStreamSynchronize(cs));
        for (int b = 0; b < 2; ++b) {
            h->p_in[b].reserve(batch * fbytes);
            h->p_aos[b].reserve(batch * fbytes);
            h->p_rt[b].reserve(batch * 48);
            h->p_rt64[b].reserve(batch * 96);
            h->p_rt32[b].reserve(batch * 48);
            if (out_xyz) {
                h->p_out[b].reserve(batch * obytes);
                h->p_xyz[b].reserve(batch * obytes);
            }
        }
    }
    WMD_HIP_CHECK(hipEventRecord(h->p_used[0].get(), st));
    WMD_HIP_CHECK(hipEventRecord(h->p_used[1].get(), st));
    bool pend[2] = {false, false};
    size_t pend_f0[2] = {0, 0}, pend_nf[2] = {0, 0};
    auto drain = [&](int b) {
        if (!pend[b]) return;
        h->p_out_ready[b].wait();
        if (out_rt) std::memcpy(out_rt + pend_f0[b] * 12, h->p_rt[b].ptr, pend_nf[b] * 48);
        if (out_xyz) host_copy(out_xyz + pend_f0[b] * n_atoms * 3, h->p_out[b].ptr, pend_nf[b] * obytes);
        pend[b] = false;
    };
    size_t k = 0;
    for (size_t f0 = 0; f0 < n_frames; f0 += batch, ++k) {
        const int b = (int)(k & 1);
        const size_t nf = std::min(batch, n_frames - f0);
        h->p_copied[b].wait();  // the H2D out of this staging buffer two batches ago is done
        drain(b);               // and that batch's results are copied out
        host_copy(h->p_in[b].ptr, coords_xyzw + f0 * n_atoms * 4, nf * fbytes);
        WMD_HIP_CHECK(hipStreamWaitEvent(cs, h->p_used[b].get(), 0));
        WMD_HIP_CHECK(hipMemcpyAsync(h->p_aos[b].ptr, h->p_in[b].ptr, nf * fbytes, hipMemcpyHostToDevice, cs));
        WMD_HIP_CHECK(hipEventRecord(h->p_copied[b].get(), cs));
        WMD_HIP_CHECK(hipStreamWaitEvent(st, h->p_copied[b].get(), 0));
        al_enqueue(h, (const float4 *)h->p_aos[b].ptr, n_atoms, nf, (double *)h->p_rt64[b].ptr,
                   (float *)h->p_rt32[b].ptr, out_xyz ? (float *)h->p_xyz[b].ptr : nullptr, st);
        WMD_HIP_CHECK(hipEventRecord(h->p_used[b].get(), st));
        WMD_HIP_CHECK(hipMemcpyAsync(h->p_rt[b].ptr, h->p_rt32[b].ptr, nf * 48, hipMemcpyDeviceToHost, st));
        if (out_xyz)
            WMD_HIP_CHECK(hipMemcpyAsync(h->p_out[b].ptr, h->p_xyz[b].ptr, nf * obytes, hipMemcpyDeviceToHost, st));
        WMD_HIP_CHECK(hipEventRecord(h->p_out_ready[b].get(), st));
        pend[b] = true;
        pend_f0[b] = f0;
        pend_nf[b] = nf;
    }
    drain(0);
    drain(1);
    WMD_HIP_CHECK(hipStreamSynchronize(st));
    WMD_API_END
}

wmd_status wmd_align_run_device(wmd_align *h, const float *d_xyzw, size_t n_atoms, size_t n_frames,
                                double *d_rt64, float *d_rt32, float *d_xyz, void *stream) {
    WMD_API_BEGIN
    WMD_REQUIRE(h != nullptr, WMD_ERR_INVALID, "null handle");
    WMD_REQUIRE(n_frames == 0 || (d_xyzw && d_rt64), WMD_ERR_INVALID, "null device buffer");
    WMD_REQUIRE(((uintptr_t)d_xyzw & 15) == 0, WMD_ERR_INVALID, "device frames must be 16-byte aligned");
    bind(h->ctx);
    al_enqueue(h, (const float4 *)d_xyzw, n_atoms, n_frames, d_rt64, d_rt32, d_xyz,
               pick_stream(h->ctx, stream));
    WMD_API_END
}

void wmd_align_destroy(wmd_align *h) {
    if (!h) return;
    int cur = 0;
    if (hipGetDevice(&cur) == hipSuccess && h->ctx) (void)hipSetDevice(h->ctx->device);
    delete h;
    (void)hipSetDevice(cur);
}

}  // extern "C"
