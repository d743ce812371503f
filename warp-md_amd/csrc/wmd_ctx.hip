// Context, error handling, buffers and host-side helpers of the engine.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <stdexcept>
#include <thread>

#include "wmd_internal.hpp"

namespace wmd {

static thread_local std::string g_last_error;

void set_error(const std::string &msg) { g_last_error = msg; }
void clear_error() { g_last_error.clear(); }

void fail(wmd_status st, const std::string &msg) { throw Error{st, msg}; }

DevBuf::~DevBuf() {
    if (ptr) {
        int cur = 0;
        if (hipGetDevice(&cur) == hipSuccess && cur != device) (void)hipSetDevice(device);
        (void)hipFree(ptr);
        if (cur != device) (void)hipSetDevice(cur);
    }
}

void *DevBuf::reserve(size_t n) {
    if (n <= bytes && ptr) return ptr;
    if (ptr) {
        WMD_HIP_CHECK(hipFree(ptr));
        ptr = nullptr;
        bytes = 0;
    }
    size_t want = round_up(n ? n : 256, 256);
    WMD_HIP_CHECK(hipGetDevice(&device));
    WMD_HIP_CHECK(hipMalloc(&ptr, want));
    bytes = want;
    return ptr;
}

Event::~Event() {
    if (ev) (void)hipEventDestroy(ev);
}

hipEvent_t Event::get() {
    if (!ev) WMD_HIP_CHECK(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
    return ev;
}

void Event::wait() {
    if (ev) WMD_HIP_CHECK(hipEventSynchronize(ev));
}

hipStream_t copy_stream(wmd_ctx *ctx) {
    if (!ctx->copy) WMD_HIP_CHECK(hipStreamCreateWithFlags(&ctx->copy, hipStreamNonBlocking));
    return ctx->copy;
}

void host_copy(void *dst, const void *src, size_t bytes) {
    const size_t kMin = 16ull << 20;  // below this one thread is as fast
    unsigned nt = std::thread::hardware_concurrency();
    nt = std::max(1u, std::min(nt, 16u));
    if (bytes < kMin || nt == 1) {
        std::memcpy(dst, src, bytes);
        return;
    }
    const size_t part = round_up((bytes + nt - 1) / nt, 4096);
    std::vector<std::thread> pool;
    for (unsigned t = 0; t < nt; ++t) {
        const size_t b0 = (size_t)t * part;
        if (b0 >= bytes) break;
        const size_t nb = std::min(part, bytes - b0);
        pool.emplace_back([=] { std::memcpy((char *)dst + b0, (const char *)src + b0, nb); });
    }
    for (auto &th : pool) th.join();
}

HostBuf::~HostBuf() {
    if (ptr) (void)hipHostFree(ptr);
}

void *HostBuf::reserve(size_t n) {
    if (n <= bytes && ptr) return ptr;
    if (ptr) {
        WMD_HIP_CHECK(hipHostFree(ptr));
        ptr = nullptr;
        bytes = 0;
    }
    size_t want = round_up(n ? n : 4096, 4096);
    WMD_HIP_CHECK(hipHostMalloc(&ptr, want, hipHostMallocDefault));
    bytes = want;
    return ptr;
}

void bind(const wmd_ctx *ctx) { WMD_HIP_CHECK(hipSetDevice(ctx->device)); }

ProfScope::ProfScope(wmd_ctx *c, hipStream_t s, const char *name) : ctx(c), st(s) {
    if (!ctx || !ctx->profile) return;  // ctx NULL: never record
    wmd_ctx::Mark m;
    m.name = name;
    if (hipEventCreate(&m.start) != hipSuccess) return;
    if (hipEventCreate(&m.stop) != hipSuccess) {
        (void)hipEventDestroy(m.start);
        return;
    }
    (void)hipEventRecord(m.start, st);
    ctx->marks.push_back(m);
    idx = (int)ctx->marks.size() - 1;
}

ProfScope::~ProfScope() {
    if (idx >= 0) (void)hipEventRecord(ctx->marks[idx].stop, st);
}

static void clear_marks(wmd_ctx *ctx) {
    for (auto &m : ctx->marks) {
        (void)hipEventDestroy(m.start);
        (void)hipEventDestroy(m.stop);
    }
    ctx->marks.clear();
}

hipStream_t pick_stream(const wmd_ctx *, void *stream) {
    // the caller's stream as given; NULL is the null (legacy default) stream
    return static_cast<hipStream_t>(stream);
}

// pbc_math.rs:110-164 cell_and_inv_from_box in f64, then the f32 casts used by
// the pair expression (geometry_math.rs:143-166 chunk_cell_mats).
bool make_box(int pbc, const float *m, BoxInfo *out) {
    std::memset(out, 0, sizeof(*out));
    if (pbc == WMD_PBC_NONE) return true;
    if (pbc == WMD_PBC_ORTHO) {
        out->lx = m[0];
        out->ly = m[4];
        out->lz = m[8];
        if (!(out->lx > 0.0f && out->ly > 0.0f && out->lz > 0.0f)) return false;
        out->width[0] = out->lx;
        out->width[1] = out->ly;
        out->width[2] = out->lz;
        for (int i = 0; i < 9; ++i) out->cell64[i] = 0.0, out->inv64[i] = 0.0;
        out->cell64[0] = m[0]; out->cell64[4] = m[4]; out->cell64[8] = m[8];
        out->inv64[0] = 1.0 / m[0]; out->inv64[4] = 1.0 / m[4]; out->inv64[8] = 1.0 / m[8];
        return true;
    }
    double m0 = m[0], m1 = m[1], m2 = m[2], m3 = m[3], m4 = m[4], m5 = m[5], m6 = m[6], m7 = m[7],
           m8 = m[8];
    double det = m0 * (m4 * m8 - m5 * m7) - m1 * (m3 * m8 - m5 * m6) + m2 * (m3 * m7 - m4 * m6);
    if (det == 0.0 || !std::isfinite(det)) return false;
    double inv[9] = {(m4 * m8 - m5 * m7) / det, (m2 * m7 - m1 * m8) / det,
                     (m1 * m5 - m2 * m4) / det, (m5 * m6 - m3 * m8) / det,
                     (m0 * m8 - m2 * m6) / det, (m2 * m3 - m0 * m5) / det,
                     (m3 * m7 - m4 * m6) / det, (m1 * m6 - m0 * m7) / det,
                     (m0 * m4 - m1 * m3) / det};
    for (int i = 0; i < 9; ++i) {
        out->cell64[i] = m[i];
        out->inv64[i] = inv[i];
        out->c[i] = m[i];
        out->r[i] = (float)inv[i];
    }
    // fractional coordinate d is f_d = sum_k inv[k][d] x_k; the spacing of the
    // planes f_d = const is 1 / |column d of inv|.
    for (int d = 0; d < 3; ++d) {
        double nrm = std::sqrt(inv[0 * 3 + d] * inv[0 * 3 + d] + inv[1 * 3 + d] * inv[1 * 3 + d] +
                               inv[2 * 3 + d] * inv[2 * 3 + d]);
        out->width[d] = 1.0 / nrm;
    }
    return true;
}

}  // namespace wmd

using namespace wmd;

extern "C" {

const char *wmd_last_error(void) { return g_last_error.c_str(); }

const char *wmd_version(void) { return "warp-md-amd 0.1.0 (gfx950)"; }

wmd_status wmd_device_count(int *out) {
    clear_error();
    if (!out) {
        set_error("null output");
        return WMD_ERR_INVALID;
    }
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) n = 0;
    *out = n;
    return WMD_OK;
}

wmd_status wmd_ctx_create(int device, wmd_ctx **out) {
    WMD_API_BEGIN
    WMD_REQUIRE(out != nullptr, WMD_ERR_INVALID, "null output handle");
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0)
        fail(WMD_ERR_UNSUPPORTED, "no HIP device available");
    WMD_REQUIRE(device >= 0 && device < n, WMD_ERR_UNSUPPORTED,
                "HIP device index " + std::to_string(device) + " out of range (" +
                    std::to_string(n) + " devices)");
    WMD_HIP_CHECK(hipSetDevice(device));
    hipDeviceProp_t prop;
    WMD_HIP_CHECK(hipGetDeviceProperties(&prop, device));
    std::string arch = prop.gcnArchName;
    if (arch.rfind("gfx950", 0) != 0)
        fail(WMD_ERR_UNSUPPORTED, "this engine is built for gfx950 (MI355X); device is " + arch);
    auto *ctx = new wmd_ctx();
    ctx->device = device;
    ctx->cu_count = prop.multiProcessorCount;
    ctx->lds_per_block = prop.sharedMemPerBlock;
    ctx->arch = arch;
    // default priority; a blocking stream, so work on it is ordered with the
    // null stream that callers (torch included) enqueue their producers on
    hipError_t e = hipStreamCreateWithFlags(&ctx->stream, hipStreamDefault);
    if (e != hipSuccess) {
        delete ctx;
        fail(WMD_ERR_DEVICE, std::string("hipStreamCreateWithFlags: ") + hipGetErrorString(e));
    }
    *out = ctx;
    WMD_API_END
}

void wmd_ctx_destroy(wmd_ctx *ctx) {
    if (!ctx) return;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    (void)hipDeviceSynchronize();
    clear_marks(ctx);
    if (ctx->copy) (void)hipStreamDestroy(ctx->copy);
    (void)hipStreamDestroy(ctx->stream);
    delete ctx;
}

wmd_status wmd_ctx_synchronize(wmd_ctx *ctx) {
    WMD_API_BEGIN
    WMD_REQUIRE(ctx != nullptr, WMD_ERR_INVALID, "null context");
    bind(ctx);
    WMD_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    WMD_API_END
}

void *wmd_ctx_stream(wmd_ctx *ctx) { return ctx ? (void *)ctx->stream : nullptr; }

int wmd_ctx_cu_count(const wmd_ctx *ctx) { return ctx ? ctx->cu_count : 0; }

wmd_status wmd_ctx_set_staging_bytes(wmd_ctx *ctx, size_t bytes, size_t *previous) {
    WMD_API_BEGIN
    WMD_REQUIRE(ctx != nullptr, WMD_ERR_INVALID, "null context");
    if (previous) *previous = ctx->staging_bytes;
    ctx->staging_bytes = bytes ? bytes : (size_t(48) << 20);
    WMD_API_END
}

wmd_status wmd_ctx_profile(wmd_ctx *ctx, int enable) {
    WMD_API_BEGIN
    WMD_REQUIRE(ctx != nullptr, WMD_ERR_INVALID, "null context");
    bind(ctx);
    WMD_HIP_CHECK(hipDeviceSynchronize());
    clear_marks(ctx);
    ctx->profile = enable != 0;
    WMD_API_END
}

wmd_status wmd_ctx_profile_read(wmd_ctx *ctx, const char *name, double *total_ms, int *launches) {
    WMD_API_BEGIN
    WMD_REQUIRE(ctx && name && total_ms && launches, WMD_ERR_INVALID, "null argument");
    bind(ctx);
    WMD_HIP_CHECK(hipDeviceSynchronize());
    double tot = 0.0;
    int cnt = 0;
    for (auto &m : ctx->marks) {
        if (m.name != name) continue;
        float ms = 0.f;
        WMD_HIP_CHECK(hipEventElapsedTime(&ms, m.start, m.stop));
        tot += ms;
        ++cnt;
    }
    *total_ms = tot;
    *launches = cnt;
    WMD_API_END
}

// rdf.rs:943-1016 RdfPlan::finalize with rdf_shell_measure (rdf.rs:487-496), 3-D.
wmd_status wmd_sf_finalize(const uint64_t *counts, size_t bins, double r_max, double n,
                           size_t frames_in, double volume_sum, int pbc, size_t q_bins, double q_max,
                           float *r, float *g_r, float *q, float *s_q) {
    WMD_API_BEGIN
    WMD_REQUIRE((counts && r && g_r) || bins == 0, WMD_ERR_INVALID, "null buffer");
    WMD_REQUIRE((q && s_q) || q_bins == 0, WMD_ERR_INVALID, "null buffer");
    WMD_REQUIRE(bins > 0 && q_bins > 0, WMD_ERR_INVALID, "structure factor requires bins and q_bins > 0");
    const double pi = 3.14159265358979323846;
    const double w = r_max / (double)bins;
    const double frames = (double)(frames_in > 0 ? frames_in : 1);
    double volume = 1.0;
    if (pbc == WMD_PBC_ORTHO) {
        if (volume_sum == 0.0) fail(WMD_ERR_MISMATCH, "structure factor requires box volume");
        volume = volume_sum / frames;
    }
    const double rho = volume > 0.0 ? n / volume : 0.0;
    for (size_t i = 0; i < bins; ++i) {
        const double r0 = ((double)i + 0.5) * w;
        const double shell = 4.0 * pi * r0 * r0 * w;
        const double norm = shell > 0.0 ? rho * shell * n : 1.0;
        r[i] = (float)r0;
        g_r[i] = (float)((double)counts[i] / (norm * frames));
    }
    const double dq = q_bins > 1 ? q_max / ((double)q_bins - 1.0) : 0.0;
    for (size_t k = 0; k < q_bins; ++k) {
        const double qv = dq * (double)k;
        double sum = 0.0;
        for (size_t i = 0; i < bins; ++i) {
            const double r0 = ((double)i + 0.5) * w;
            const double diff = (double)g_r[i] - 1.0;
            const double weight = qv == 0.0 ? 1.0 : std::sin(qv * r0) / (qv * r0);
            sum += diff * weight * r0 * r0 * w;
        }
        q[k] = (float)qv;
        s_q[k] = (float)(1.0 + 4.0 * pi * rho * sum);
    }
    WMD_API_END
}

wmd_status wmd_rdf_finalize(const uint64_t *counts, size_t bins, float r_max, double n_a,
                            double n_b, double common, size_t frames_in, double volume_sum,
                            double density_opt, uint32_t flags, float *r, float *g,
                            float *integral) {
    WMD_API_BEGIN
    WMD_REQUIRE(counts && r && g, WMD_ERR_INVALID, "null buffer");
    const bool want_integral = flags & WMD_RDF_INTEGRAL;
    WMD_REQUIRE(!want_integral || integral, WMD_ERR_INVALID, "null integral buffer");
    if (bins == 0) return WMD_OK;
    const bool use_volume = flags & WMD_RDF_USE_VOLUME;
    const bool raw = flags & WMD_RDF_RAW;
    const bool nd_out = flags & WMD_RDF_NUMBER_DENSITY;
    const bool planar = flags & WMD_RDF_FINALIZE_XY;
    // Arithmetic below is written so that each f32 operation rounds once, in
    // the reference's order (the shell volume is an f32 expression cast to f64).
    volatile float bin_width = r_max / (float)bins;
    const double frames = (double)(frames_in > 1 ? frames_in : 1);
    double total_pairs = n_a * n_b - common;
    if (total_pairs < 0.0) total_pairs = 0.0;
    double density;
    if (use_volume) {
        if (volume_sum <= 0.0)
            fail(WMD_ERR_MISMATCH, "rdf box normalization requires frame box metadata");
        density = total_pairs / (volume_sum / frames);
    } else if (n_a > 0.0) {
        density = density_opt * total_pairs / n_a;
    } else {
        density = 0.0;
    }
    const float four_thirds = 4.0f / 3.0f;
    const float pi32 = 3.14159265358979323846f;
    double acc = 0.0;
    for (size_t i = 0; i < bins; ++i) {
        volatile float r_inner = (float)i * bin_width;
        volatile float r_outer = (float)(i + 1) * bin_width;
        volatile float ro2 = r_outer * r_outer;
        volatile float ro3 = ro2 * r_outer;
        volatile float ri2 = r_inner * r_inner;
        volatile float ri3 = ri2 * r_inner;
        volatile float diff = ro3 - ri3;
        volatile float t = four_thirds * pi32;
        volatile float shell32 = t * diff;
        if (planar) {
            // rdf_shell_measure, PlanarXY (rdf.rs:487-496): PI (r_o^2 - r_i^2) in f32
            volatile float d2 = ro2 - ri2;
            shell32 = pi32 * d2;
        }
        const double shell = (double)shell32;
        const double expected = shell * density;
        const double norm = expected > 0.0 ? frames * expected : 0.0;
        volatile float half = 0.5f * bin_width;
        r[i] = r_inner + half;
        const double count = (double)counts[i];
        const double normalized = norm > 0.0 ? count / norm : 0.0;
        const double nd_norm = frames * n_a * shell;
        const double number_density = nd_norm > 0.0 ? count / nd_norm : 0.0;
        if (want_integral) {
            if (nd_out) {
                acc += number_density * shell;
            } else if (n_b > 0.0) {
                acc += normalized * expected / n_b;
            }
            integral[i] = (float)acc;
        }
        double value = raw ? count : (nd_out ? number_density : normalized);
        g[i] = (float)value;
    }
    WMD_API_END
}

}  // extern "C"
