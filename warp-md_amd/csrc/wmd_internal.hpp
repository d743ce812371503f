// Internal declarations shared by the engine's translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

#include "wmd_hip.h"

namespace wmd {

// ---------------------------------------------------------------------------
// Errors: a thread-local message plus a status code; no exception crosses the
// C ABI (every extern "C" body is wrapped in WMD_API_BEGIN / WMD_API_END).
// ---------------------------------------------------------------------------
void set_error(const std::string &msg);
void clear_error();

struct Error {
    wmd_status status;
    std::string msg;
};

[[noreturn]] void fail(wmd_status st, const std::string &msg);

#define WMD_HIP_CHECK(expr)                                                                   \
    do {                                                                                      \
        hipError_t _e = (expr);                                                               \
        if (_e != hipSuccess) {                                                               \
            ::wmd::fail(_e == hipErrorOutOfMemory ? WMD_ERR_OOM : WMD_ERR_DEVICE,             \
                        std::string(#expr) + ": " + hipGetErrorString(_e));                   \
        }                                                                                     \
    } while (0)

#define WMD_REQUIRE(cond, st, msg)                                                            \
    do {                                                                                      \
        if (!(cond)) ::wmd::fail((st), (msg));                                                \
    } while (0)

#define WMD_API_BEGIN                                                                         \
    ::wmd::clear_error();                                                                     \
    try {
#define WMD_API_END                                                                           \
    }                                                                                         \
    catch (const ::wmd::Error &e) {                                                           \
        ::wmd::set_error(e.msg);                                                              \
        return e.status;                                                                      \
    }                                                                                         \
    catch (const std::exception &e) {                                                         \
        ::wmd::set_error(e.what());                                                           \
        return WMD_ERR_INVALID;                                                               \
    }                                                                                         \
    catch (...) {                                                                             \
        ::wmd::set_error("unknown internal error");                                           \
        return WMD_ERR_INVALID;                                                               \
    }                                                                                         \
    return WMD_OK;

// ---------------------------------------------------------------------------
// Device memory owned by the library: a growable buffer freed on destruction.
// ---------------------------------------------------------------------------
struct DevBuf {
    void *ptr = nullptr;
    size_t bytes = 0;
    int device = 0;
    DevBuf() = default;
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    ~DevBuf();
    // Ensures at least `n` bytes; contents are not preserved when it grows.
    void *reserve(size_t n);
    template <class T> T *as() const { return static_cast<T *>(ptr); }
};

// Pinned host staging buffer (hipHostMalloc), growable.
struct HostBuf {
    void *ptr = nullptr;
    size_t bytes = 0;
    HostBuf() = default;
    HostBuf(const HostBuf &) = delete;
    HostBuf &operator=(const HostBuf &) = delete;
    ~HostBuf();
    void *reserve(size_t n);
};

// A HIP event created on first use (timing disabled), destroyed with its owner.
struct Event {
    hipEvent_t ev = nullptr;
    Event() = default;
    Event(const Event &) = delete;
    Event &operator=(const Event &) = delete;
    ~Event();
    hipEvent_t get();
    // Blocks until the last record has completed (no-op before the first record).
    void wait();
};

// memcpy of host memory with up to 16 threads for large copies (pinned staging).
void host_copy(void *dst, const void *src, size_t bytes);

// ---------------------------------------------------------------------------
// Context (the GpuContext replacement): device, one default-priority stream.
// ---------------------------------------------------------------------------
}  // namespace wmd

struct wmd_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    hipStream_t copy = nullptr;  // host-to-device copies of the host-chunk paths (created on first use)
    int cu_count = 0;
    int xcd_count = 8;
    size_t lds_per_block = 0;
    std::string arch;
    // bytes of host chunk per pinned staging buffer of the host-chunk pipelines
    size_t staging_bytes = size_t(48) << 20;
    // kernel timing (wmd_ctx_profile)
    bool profile = false;
    struct Mark {
        std::string name;
        hipEvent_t start, stop;
    };
    std::vector<Mark> marks;
};

namespace wmd {

// Makes the context's device current for the calling thread.
void bind(const wmd_ctx *ctx);
// The stream a device-buffer entry point runs on: the caller's (NULL = null stream).
hipStream_t pick_stream(const wmd_ctx *ctx, void *stream);
// The context's copy stream (non-blocking, default priority), created on first use.
hipStream_t copy_stream(wmd_ctx *ctx);

inline size_t round_up(size_t x, size_t m) { return (x + m - 1) / m * m; }

// Frames per batch of a host-chunk pipeline: as many frames of `frame_bytes`
// as fit one staging buffer of the context (at least one, at most n_frames).
inline size_t staging_frames(const wmd_ctx *ctx, size_t frame_bytes, size_t n_frames) {
    const size_t per = ctx->staging_bytes / (frame_bytes ? frame_bytes : 1);
    return per < 1 ? 1 : (per < n_frames ? per : (n_frames ? n_frames : 1));
}

// Records HIP events around a launch sequence on `st` when profiling is on
// (c == NULL: records nothing).
struct ProfScope {
    wmd_ctx *ctx;
    hipStream_t st;
    int idx = -1;
    ProfScope(wmd_ctx *c, hipStream_t s, const char *name);
    ~ProfScope();
};

// Batched RMSD to frame 0 over the trajectories of an uneven batch
// (wmd_batch): one descriptor per trajectory in the batch store.
struct BatchRmsdTraj {
    long long store_off;  // float offset of frame 0 ([F][3][stride] SoA)
    long long unit0;      // first moments unit (frame, tile) of this trajectory
    long long out0;       // index of frame 0 in the per-frame outputs
    int stride, n, n_frames, tiles, gpt;  // gpt: float4 groups per tile
    int pad;
};

// Moments (one wave per (trajectory, frame, tile), FP64) and the QCP finalize
// (one thread per output frame): out32 / out64 [n_out] RMSD of every frame to
// its trajectory's frame 0 after optimal superposition, err [n_out] an error
// bound (out64 / err may be NULL).  partials holds n_units x 13 doubles.
void batch_rmsd_enqueue(hipStream_t st, const BatchRmsdTraj *d_traj, int n_traj,
                        long long n_units, long long n_out, const float *d_store,
                        double *d_partials, float *d_out32, double *d_out64, double *d_err);

// Host-side minimum-image support shared by the RDF paths.
struct BoxInfo {
    // f32 parameters the pair kernels use (one per frame).
    float lx, ly, lz;       // orthorhombic lengths
    float c[9];             // triclinic rows (box vectors), f32
    float r[9];             // triclinic inverse rows (f32 casts of the f64 inverse)
    double cell64[9], inv64[9];
    double width[3];        // plane spacings (A) along the three fractional axes
};

// Builds BoxInfo from a row-major box for the given pbc kind.
// Returns false for a singular / non-positive box.
bool make_box(int pbc, const float *m, BoxInfo *out);

}  // namespace wmd
