/*
 * wmd_hip.h -- C ABI of the MI355X (gfx950) engine for warp-md's frame-batched
 * RdfPlan / RmsdPlan hot path.
 *
 * This library replaces the RDF/RMSD slice of crates/traj-gpu (the `GpuContext`
 * handle, crates/traj-gpu/src/cuda.rs:52-159) and the corresponding kernels of
 * crates/traj-kernels (rdf_hist part4.cu:29-70; rmsd_centroid / rmsd_cov /
 * rmsd_raw_accum / rmsd_finalize part1.cu:299-393).  Every entry point takes
 * plain pointers and sizes; no C++ exception crosses it.
 *
 * Conventions
 *  - Host buffers are owned by the caller and only read (or written, for
 *    outputs) during the call.  Handles are owned by the library.
 *  - One context = one device + one blocking HIP stream (host-buffer entry
 *    points run on it and return when done); a context and the handles made
 *    from it are not reentrant.  Distinct contexts may be used from distinct
 *    threads.  Device-buffer entry points run on the stream the caller passes.
 *  - Every function returns a wmd_status; on failure wmd_last_error() returns a
 *    thread-local message.  Status codes mirror traj_core's TrajError variants
 *    (crates/traj-core/src/error.rs) so a Rust shim maps them one to one.
 *  - "xyzw" host coordinates are the reference FrameChunk layout
 *    (crates/traj-core/src/frame.rs:10-20): [n_frames][n_atoms][4] float, w ignored.
 *  - "boxes" are [n_frames][9] row-major box vectors (a, b, c rows).  An
 *    orthorhombic box is diag(lx, ly, lz); the engine reads m[0], m[4], m[8].
 *  - Device-resident stores ("*_device" entry points) are SoA frames:
 *    [n_frames][3][stride] float (x plane, y plane, z plane), stride a
 *    multiple of 4 and >= the plan's store size (see *_store_layout).
 */
#ifndef WMD_HIP_H
#define WMD_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum wmd_status {
    WMD_OK = 0,
    WMD_ERR_UNSUPPORTED = 1, /* TrajError::Unsupported (cuda.rs:384-390) */
    WMD_ERR_MISMATCH = 2,    /* TrajError::Mismatch (runtime_core.rs:61-62, rdf.rs:894-908) */
    WMD_ERR_PARSE = 3,       /* TrajError::Parse (rdf.rs:663-667) */
    WMD_ERR_DEVICE = 4,      /* HIP runtime failure */
    WMD_ERR_OOM = 5,         /* device allocation failure */
    WMD_ERR_INVALID = 6      /* null pointer / index out of range / bad size */
} wmd_status;

enum { WMD_PBC_NONE = 0, WMD_PBC_ORTHO = 1, WMD_PBC_TRICLINIC = 2 };

/* wmd_rdf_create flags */
#define WMD_RDF_EDGE_DIAGNOSTICS 0x1u /* count counted pairs within 1e-5 A of a bin edge */
#define WMD_RDF_FORCE_BRUTE 0x2u      /* never use the cell list (testing / autotuning) */
#define WMD_RDF_PLANAR_XY 0x4u        /* dimension "xy": r = sqrt(dx^2 + dy^2) (rdf.rs:473-478) */
#define WMD_RDF_NO_INTRAMOL 0x8u      /* wmd_rdf_create_groups: skip intramolecular group pairs */

/* wmd_rdf_finalize flags (RdfPlan output options, rdf.rs:204-227) */
#define WMD_RDF_USE_VOLUME 0x1u
#define WMD_RDF_RAW 0x2u
#define WMD_RDF_NUMBER_DENSITY 0x4u
#define WMD_RDF_INTEGRAL 0x8u
#define WMD_RDF_FINALIZE_XY 0x10u /* dimension "xy": shells PI (r_o^2 - r_i^2) */

typedef struct wmd_ctx wmd_ctx;
typedef struct wmd_rdf wmd_rdf;
typedef struct wmd_rmsd wmd_rmsd;
typedef struct wmd_softrdf wmd_softrdf;
typedef struct wmd_pairwise wmd_pairwise;
typedef struct wmd_align wmd_align;

/* ---- library / context --------------------------------------------------- */

/* Thread-local message of the last failing call on this thread ("" if none). */
const char *wmd_last_error(void);
/* Library version string. */
const char *wmd_version(void);
/* Number of visible HIP devices (0 when no GPU / no driver; never fails). */
wmd_status wmd_device_count(int *out);

/* GpuContext::new (crates/traj-gpu/src/cuda.rs:52-159, runtime_core.rs:4-17).
 * Code objects are loaded once per process; nothing is compiled at run time. */
wmd_status wmd_ctx_create(int device, wmd_ctx **out);
void wmd_ctx_destroy(wmd_ctx *ctx);
/* Waits for all work queued on the context's stream. */
wmd_status wmd_ctx_synchronize(wmd_ctx *ctx);
/* The context's hipStream_t (as void*), for callers that enqueue around it. */
void *wmd_ctx_stream(wmd_ctx *ctx);
/* Number of compute units of the context's device. */
int wmd_ctx_cu_count(const wmd_ctx *ctx);
/* Size of each of the two pinned staging buffers through which the host-chunk
 * entry points (wmd_rdf_accumulate[_selected], wmd_rmsd_run, wmd_align_run,
 * wmd_pairwise_append) stream a chunk to the device, a batch of frames at a
 * time: the engine's side of the reference's chunk streaming
 * (crates/traj-engine/src/executor.rs:486-509, chunk sizes from
 * crates/traj-py/src/pack.rs:1447-1466).  bytes = 0 restores the default,
 * 48 MiB; a batch holds at least one frame.  *previous (may be NULL) gets the
 * size in force before the call.  Results do not depend on it. */
wmd_status wmd_ctx_set_staging_bytes(wmd_ctx *ctx, size_t bytes, size_t *previous);
/* Kernel timing with HIP events recorded around the engine's main kernels on
 * the stream they run on (enable = 1 starts recording and clears old marks).
 * wmd_ctx_profile_read synchronises and returns, for kernel family `name`
 * ("rdf_pair", "rdf_gather" (the host-chunk path's selection gather),
 * "rmsd_moments", "rmsd_finalize", "cell_build", ...), the summed
 * duration in ms and the number of launches since profiling was enabled. */
wmd_status wmd_ctx_profile(wmd_ctx *ctx, int enable);
wmd_status wmd_ctx_profile_read(wmd_ctx *ctx, const char *name, double *total_ms, int *launches);

/* ---- RDF ------------------------------------------------------------------ */

/* RdfPlan::init GPU branch (crates/traj-engine/src/plans/rdf.rs:692-714):
 * GpuContext::selection(sel_a), selection(sel_b) (runtime_core.rs:19-39) and
 * alloc_counts + reset_counts (packing_constraints.rs:891-914).
 * sel_a / sel_b: atom indices (duplicates allowed, they weight pairs like the
 * reference's loops).  pbc: WMD_PBC_*.  Pairs with the same atom index are
 * skipped (rdf.rs:917-919).  The IO selection is the order-preserving union of
 * sel_a then sel_b (rdf.rs:498-526).  bins <= 4194304: a histogram larger
 * than the pair kernels' LDS is counted one slice of bins per pass (each
 * pass enumerates the pairs again). */
wmd_status wmd_rdf_create(wmd_ctx *ctx, const uint32_t *sel_a, size_t n_a, const uint32_t *sel_b,
                          size_t n_b, size_t bins, float r_max, int pbc, uint32_t flags,
                          wmd_rdf **out);
/* The IO selection (union of sel_a and sel_b, first-seen order). *n gets its
 * length; out may be NULL to query the length. */
wmd_status wmd_rdf_io_selection(const wmd_rdf *rdf, uint32_t *out, size_t cap, size_t *n);
/* GpuContext::rdf_accum (crates/traj-gpu/src/cuda/structure_polymer.rs:246-313)
 * on a full-system host chunk: coords [n_frames][n_atoms][4], boxes
 * [n_frames][9] (may be NULL for WMD_PBC_NONE).  Counts persist across calls. */
wmd_status wmd_rdf_accumulate(wmd_rdf *rdf, const float *coords_xyzw, size_t n_atoms,
                              size_t n_frames, const float *boxes);
/* RdfPlan::process_chunk_selected (rdf.rs:882-941): the chunk holds only the IO
 * selection atoms, in IO-selection order: coords [n_frames][n_io][4]. */
wmd_status wmd_rdf_accumulate_selected(wmd_rdf *rdf, const float *coords_xyzw, size_t n_frames,
                                       const float *boxes);
/* Hot path on device-resident coordinates: d_soa is [n_frames][3][stride] float
 * holding the IO selection atoms in IO-selection order; boxes are host
 * [n_frames][9].  Work is enqueued on `stream` (a hipStream_t; NULL = the null
 * stream), after the work already queued there, and the call returns before it
 * completes. */
wmd_status wmd_rdf_accumulate_device(wmd_rdf *rdf, const float *d_soa, size_t stride,
                                     size_t n_frames, const float *boxes, void *stream);
/* GpuContext::read_counts (packing_constraints.rs:891-898): copies the u64
 * counts [bins] out (synchronises).  edge: NULL or [1], the number of counted
 * pairs whose f32 distance lies within 1e-5 A of a bin edge (requires
 * WMD_RDF_EDGE_DIAGNOSTICS, else 0). */
wmd_status wmd_rdf_read(wmd_rdf *rdf, uint64_t *counts, uint64_t *edge);
/* Diagnostics with no counterpart in the reference: the number of pair
 * distances the cell-list kernels evaluated since create / reset (unordered
 * pairs; counted only with WMD_RDF_EDGE_DIAGNOSTICS, else 0; the brute-force
 * path, which evaluates every pair of the selections, is not counted).  For
 * one selection paired with itself, (sum of counts / 2) / evaluations is the
 * enumeration's candidate efficiency.  Synchronises. */
wmd_status wmd_rdf_pair_evaluations(wmd_rdf *rdf, uint64_t *evaluations);
/* GpuContext::reset_counts (packing_constraints.rs:900-914). */
wmd_status wmd_rdf_reset(wmd_rdf *rdf);
void wmd_rdf_destroy(wmd_rdf *rdf);
/* Which pair kernel the last accumulate used: 0 none yet, 1 brute force,
 * 2 cell list; *cells_per_side receives the last cell grid (or zeros). */
wmd_status wmd_rdf_last_path(const wmd_rdf *rdf, int *path, int *cells_per_side);
/* Cell-list overrides from an autotuner: k = cells per r_max (1..3, 0 =
 * automatic) and grab = cell units a wave takes per visit of its XCD's work
 * queue (1..64, 0 = automatic: 4 for launches of at least 64 units per wave,
 * else 2).  Take effect on the next accumulate. */
wmd_status wmd_rdf_set_tuning(wmd_rdf *rdf, int cell_k, int grab);

/* RdfPlan::finalize normalisation (rdf.rs:943-1016) on host data: r, g [bins]
 * and, with WMD_RDF_INTEGRAL, integral [bins] (else may be NULL).  flags:
 * WMD_RDF_USE_VOLUME / RAW / NUMBER_DENSITY / INTEGRAL / FINALIZE_XY.  Returns
 * WMD_ERR_MISMATCH for USE_VOLUME with volume_sum <= 0 (rdf.rs:962-966). */
wmd_status wmd_rdf_finalize(const uint64_t *counts, size_t bins, float r_max, double n_a,
                            double n_b, double common, size_t frames, double volume_sum,
                            double density, uint32_t flags, float *r, float *g, float *integral);

/* RdfPlan's grouped modes (rdf.rs:279-376, 808-850; CPU-only in the
 * reference, rdf.rs:697-699): counts over pairs of group positions.
 * The caller builds the groups as rdf_groups does (centre: the whole
 * selection; bymol / byres: one group per molecule / residue; else one per
 * atom) and passes their union, identical atom lists merged:
 * group_offsets [n_groups + 1] into group_atoms (atom indices of each group,
 * in order), group_masses NULL (plain centres) or parallel to group_atoms
 * (mass-weighted centres); groups_a / groups_b: the group ids of each side
 * (a group paired with itself is skipped, rdf.rs:819-821).  With
 * WMD_RDF_NO_INTRAMOL in flags, mol_key / res_key [2 n_groups] (INT64_MIN =
 * none; [g] the keys of group g as a sel_a group, [n_groups + g] as a sel_b
 * group) give same_intramolecular_group (rdf.rs:368-373) and those pairs are
 * skipped.  Centres follow center_of_selection (traj-core/src/centers.rs:3-29)
 * in f64 cast to f32; pair distances, wrap and binning are the plain RDF's, so
 * counts are bit-exact.  WMD_RDF_PLANAR_XY selects the xy dimension.  The
 * handle takes full chunks (wmd_rdf_accumulate) and is read, reset and
 * destroyed with the wmd_rdf_* calls; pbc: none or orthorhombic. */
wmd_status wmd_rdf_create_groups(wmd_ctx *ctx, size_t n_groups, const uint64_t *group_offsets,
                                 const uint32_t *group_atoms, const float *group_masses,
                                 const uint32_t *groups_a, size_t n_a, const uint32_t *groups_b,
                                 size_t n_b, const int64_t *mol_key, const int64_t *res_key,
                                 size_t bins, float r_max, int pbc, uint32_t flags, wmd_rdf **out);

/* StructureFactorPlan's pair counts (crates/traj-engine/src/plans/analysis/
 * structure_factor.rs:160-198; its GPU branch reuses rdf_accum, :128-149): an
 * RDF plan over sel x sel (pairs of an atom with itself skipped) whose pair
 * distances follow the reference's CPU path in f64 -- the f32 coordinate
 * difference cast to f64 and multiplied by length_scale, the orthorhombic
 * wrap in f64 on the scaled box, r = sqrt in f64, r < r_max and
 * (r / (r_max / bins)) as usize in f64 -- so counts are bit-exact with it.
 * pbc: WMD_PBC_NONE or WMD_PBC_ORTHO; bins <= 4096.  The handle is used with
 * the wmd_rdf_* calls above (accumulate*, read, reset, destroy). */
wmd_status wmd_rdf_create_f64(wmd_ctx *ctx, const uint32_t *sel, size_t n_sel, size_t bins,
                              double r_max, int pbc, double length_scale, wmd_rdf **out);
/* StructureFactorPlan::finalize (structure_factor.rs:200-258): r, g_r [bins]
 * at bin centres, q, s_q [q_bins] on q = q_max i / (q_bins - 1).  n: selection
 * size; volume_sum: summed scaled box volumes (required for WMD_PBC_ORTHO). */
wmd_status wmd_sf_finalize(const uint64_t *counts, size_t bins, double r_max, double n,
                           size_t frames, double volume_sum, int pbc, size_t q_bins, double q_max,
                           float *r, float *g_r, float *q, float *s_q);

/* ---- RMSD ----------------------------------------------------------------- */

/* RmsdPlan::init GPU branch (crates/traj-engine/src/plans/rmsd/plan.rs:205-220):
 * GpuContext::selection + GpuContext::reference (runtime_core.rs:127-137).
 * fit_sel: atoms of the superposition fit; rmsd_sel: atoms the RMSD is taken
 * over (NULL => same as fit_sel, the RmsdPlan case; the split form is the
 * fit-on-A / RMSD-over-B composition of rms.rs:643-676).  ref_xyzw:
 * [n_ref_atoms][4] reference coordinates indexed by atom.  align = 0 gives the
 * unfitted RMSD (rmsd/plan.rs:1093-1110) over rmsd_sel (fit_sel when rmsd_sel
 * is NULL); the split form then leaves fit_sel unused, as rms.rs:664-666. */
wmd_status wmd_rmsd_create(wmd_ctx *ctx, const uint32_t *fit_sel, size_t n_fit,
                           const uint32_t *rmsd_sel, size_t n_rmsd, const float *ref_xyzw,
                           size_t n_ref_atoms, int align, wmd_rmsd **out);
/* Device store layout: atoms[slot] = atom index held in store slot `slot`
 * (0xFFFFFFFF = zero padding); *n_store = slots used; *stride = minimum SoA
 * plane stride.  atoms may be NULL to query sizes. */
wmd_status wmd_rmsd_store_layout(const wmd_rmsd *rmsd, uint32_t *atoms, size_t cap,
                                 size_t *n_store, size_t *stride);
/* GpuContext::rmsd_covariance + host rmsd_from_cov, or rmsd_raw
 * (runtime_core.rs:375-535, rmsd/math.rs:7-38) on a host chunk
 * [n_frames][n_atoms][4]: out_rmsd [n_frames] f32 (synchronous).  out_err:
 * NULL or [n_frames] estimated absolute error bound (A) of each RMSD against
 * exact arithmetic on the same f32 inputs. */
wmd_status wmd_rmsd_run(wmd_rmsd *rmsd, const float *coords_xyzw, size_t n_atoms,
                        size_t n_frames, float *out_rmsd, double *out_err);
/* Hot path on a device-resident store [n_frames][3][stride] (see
 * wmd_rmsd_store_layout).  Device outputs (each may be NULL except d_rmsd):
 * d_rmsd [F] f32, d_rmsd64 [F] f64, d_err [F] f64, d_rot [F][9] f64 (row-major
 * R with R (x - cx) ~ (y - cy)), d_cx [F][3] f64 fit centroid of the frame.
 * Enqueued on `stream` (a hipStream_t; NULL = the null stream) after the work
 * already queued there; returns before it completes. */
wmd_status wmd_rmsd_run_device(wmd_rmsd *rmsd, const float *d_store, size_t stride,
                               size_t n_frames, float *d_rmsd, double *d_rmsd64, double *d_err,
                               double *d_rot, double *d_cx, void *stream);
/* Allocation-free variant of wmd_rmsd_run_device.  It has no counterpart in
 * the reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  Same arguments, outputs, results and stream semantics as
 * wmd_rmsd_run_device; the per-(frame, tile) moment partials go to a fixed
 * 16 MiB device array of the library (one per GPU, held until the process
 * ends) instead of a buffer the handle allocates, and the call allocates
 * nothing (kernel timing by wmd_ctx_profile is not recorded for it).
 * Returns WMD_ERR_UNSUPPORTED, before enqueuing anything, when the frames'
 * partials exceed that array: n_frames x tiles x moments x 8 B, with 13
 * moments per (frame, tile) for fit == RMSD selection, 26 for a split fit,
 * 1 unfitted, and one tile per frame beyond 12,288 frames (so at most about
 * 161,000 / 80,000 / 2,000,000 frames per call).  Every call on a GPU
 * shares that GPU's array: calls that overlap on one GPU (other streams,
 * other handles, other threads) race; serialise them. */
wmd_status wmd_rmsd_run_device_noalloc(wmd_rmsd *rmsd, const float *d_store, size_t stride,
                                       size_t n_frames, float *d_rmsd, double *d_rmsd64,
                                       double *d_err, double *d_rot, double *d_cx, void *stream);
/* Backward of wmd_rmsd_run_device (no counterpart in the reference: the adjoint
 * of structure-matching losses, the differentiable companion of RmsdPlan).
 * d_gout [F] f32 = dL/dRMSD per frame; d_rmsd64, d_rot, d_cx: the outputs of
 * the forward run on the same store.  Writes d_grad [F][3][stride] (store
 * layout, 0 in padding slots) = dL/dx: ((x - cx) - R^T (y - cy)) / (n RMSD)
 * per atom for the fitted RMSD over its own fit set (envelope theorem);
 * (x - y) / (n RMSD) unfitted; for a fit on set A and the RMSD over set B,
 * with u = x - cx, v = y - cy, r = R u - v, the adjoint through the rotation:
 *   dRMSD/dx_i = (2 R^T r_i [i in B] - (2 / n_A) R^T sum_B r [i in A]
 *                 + 2 G^T v_i [i in A]) / (2 n_B RMSD),
 * G = [z]x R, z = (tr(P) I - P)^-1 q, P = R H (H the fit set's centred
 * cross-covariance), q the axial vector of R sum_B u r^T (the store is read
 * twice).  The reference coordinates are constants.  Asynchronous on
 * `stream`. */
wmd_status wmd_rmsd_backward_device(wmd_rmsd *rmsd, const float *d_store, size_t stride,
                                    size_t n_frames, const double *d_rmsd64, const double *d_rot,
                                    const double *d_cx, const float *d_gout, float *d_grad,
                                    void *stream);
void wmd_rmsd_destroy(wmd_rmsd *rmsd);

/* ---- Pairwise RMSD ---------------------------------------------------------- */

/* PairwiseRmsdPlan metrics (crates/traj-engine/src/plans/rmsd/plan.rs:586-675). */
enum { WMD_PAIRWISE_RMS = 0, WMD_PAIRWISE_NOFIT = 1, WMD_PAIRWISE_DME = 2 };

/* PairwiseRmsdPlan::new + init (plan.rs:67-73, 514-519): the F x F RMSD matrix
 * of one selection between every two frames, "rms" after optimal superposition
 * (rmsd_aligned, plan.rs:1112-1159: f64 centroids, f64 cross-covariance, the
 * superposition residual) or "nofit" (rmsd_raw, plan.rs:1081-1091).  The
 * matrix is symmetric with a zero diagonal; a selection of fewer than 2 atoms
 * gives the zero matrix (plan.rs:597-603).  The O(F^2 n) cross-covariances are
 * one FP64 Gram product on the matrix cores.  WMD_PAIRWISE_DME is the
 * distance-matrix RMSD of plan.rs:612-650 without PBC (see
 * wmd_pairwise_create_dme). */
wmd_status wmd_pairwise_create(wmd_ctx *ctx, const uint32_t *sel, size_t n_sel, int metric,
                               wmd_pairwise **out);
/* PairwiseRmsdPlan::new(selection, PairwiseMetric::Dme, pbc) (plan.rs:67-73,
 * 612-650): for every two frames sqrt(sum_k (d_i[k] - d_j[k])^2 / n_pairs)
 * over the compact pair-distance vectors of pair_distances_compact
 * (plan.rs:1279-1302: i < j in selection order, the f32 difference cast to
 * f64, apply_pbc on the frame's box lengths for WMD_PBC_ORTHO, sqrt in f64).
 * The sums run in the reference's k order with its roundings.  pbc:
 * WMD_PBC_NONE or WMD_PBC_ORTHO (else WMD_ERR_PARSE); with WMD_PBC_ORTHO the
 * frames must be appended with wmd_pairwise_append_boxes. */
wmd_status wmd_pairwise_create_dme(wmd_ctx *ctx, const uint32_t *sel, size_t n_sel, int pbc,
                                   wmd_pairwise **out);
/* PairwiseRmsdPlan::process_chunk (plan.rs:533-558): appends the selected atoms
 * of a host chunk [n_frames][n_atoms][4] to the plan's device store. */
wmd_status wmd_pairwise_append(wmd_pairwise *pw, const float *coords_xyzw, size_t n_atoms,
                               size_t n_frames);
/* The same with the chunk's boxes [n_frames][9] (FrameChunk.box_, plan.rs:
 * 549-556): a dme handle with WMD_PBC_ORTHO keeps each frame's box lengths
 * (WMD_ERR_MISMATCH "orthorhombic box required for PBC" for a box with
 * off-diagonal terms, as box_lengths, pbc_math.rs:11-15); other handles ignore
 * them.  wmd_pairwise_append on such a handle returns WMD_ERR_MISMATCH. */
wmd_status wmd_pairwise_append_boxes(wmd_pairwise *pw, const float *coords_xyzw, size_t n_atoms,
                                     size_t n_frames, const float *boxes);
wmd_status wmd_pairwise_frames(const wmd_pairwise *pw, size_t *n_frames);
/* PairwiseRmsdPlan::finalize (plan.rs:586-675) over the appended frames:
 * out [F][F] f32, row-major (synchronous). */
wmd_status wmd_pairwise_finalize(wmd_pairwise *pw, float *out);
/* Store stride (floats per coordinate plane) of the appended store and of the
 * stores wmd_pairwise_run_device expects at minimum. */
wmd_status wmd_pairwise_store_stride(const wmd_pairwise *pw, size_t *stride);
/* Hot path on a device store [n_frames][3][stride] holding the selection's
 * atoms in slots 0 .. n_sel - 1 (16-byte aligned): d_out [F][F] f32 and
 * d_out64 (NULL or [F][F] f64, the value before the cast).  Asynchronous on
 * `stream`.  A dme handle with WMD_PBC_ORTHO returns WMD_ERR_UNSUPPORTED
 * (its boxes come with wmd_pairwise_append_boxes). */
wmd_status wmd_pairwise_run_device(wmd_pairwise *pw, const float *d_store, size_t stride,
                                   size_t n_frames, float *d_out, double *d_out64, void *stream);
void wmd_pairwise_destroy(wmd_pairwise *pw);
/* TrajectoryClusterPlan k-means (crates/traj-engine/src/plans/analysis/
 * clustering.rs:190-206, 384-519) over the frames appended to a pairwise handle:
 * features = each frame's selection superposed onto frame 0 in f64
 * (align_to_reference :394-418), then Lloyd iterations seeded by the
 * reference's xorshift64 Fisher-Yates shuffle (:530-550), distances and sums in
 * its order.  n_clusters is clamped to [1, F], max_iter to >= 1, tol to >= 0.
 * labels [F], centroids and sizes [min(n_clusters, F)] (synchronous). */
wmd_status wmd_pairwise_kmeans(wmd_pairwise *pw, size_t n_clusters, size_t max_iter, double tol,
                               uint64_t seed, int32_t *labels, uint32_t *centroids, uint32_t *sizes);
/* TrajectoryClusterPlan DBSCAN (clustering.rs:290-382) on host data: dist is a
 * full [n][n] distance matrix (e.g. wmd_pairwise_finalize's); labels [n]
 * (-1 noise), centroids / sizes [n] of which *n_clusters are written (the
 * medoid of each cluster and its size). */
wmd_status wmd_dbscan(const float *dist, size_t n, double eps, size_t min_samples, int32_t *labels,
                      uint32_t *centroids, uint32_t *sizes, size_t *n_clusters);

/* ---- Align / Superpose ------------------------------------------------------- */

/* AlignPlan / SuperposePlan::new + init (crates/traj-engine/src/plans/geometry/
 * align/plan_exec.rs:3-43, 78-137, 263-315): superposition of `sel` onto a
 * reference with kabsch_rotation_weighted (geometry/geometry_math.rs:235-297;
 * f64 weighted centroids and cross-covariance, the optimal rotation from the
 * FP64 QCP solve).  masses: NULL (unweighted) or [n_sel] weights (mass=True).
 * ref_xyzw: NULL (reference="frame0": call wmd_align_set_reference with the
 * first frame before running) or [n_ref_atoms][4] atom-indexed reference
 * coordinates (reference="topology", positions0).  norotate: R = I. */
wmd_status wmd_align_create(wmd_ctx *ctx, const uint32_t *sel, size_t n_sel, const float *masses,
                            const float *ref_xyzw, size_t n_ref_atoms, int norotate,
                            wmd_align **out);
/* Sets the reference from atom-indexed coordinates [n_ref_atoms][4]
 * (plan_exec.rs:126-139, frame0 mode: the first frame's selected atoms). */
wmd_status wmd_align_set_reference(wmd_align *al, const float *ref_xyzw, size_t n_ref_atoms);
/* AlignPlan / SuperposePlan::process_chunk on a host chunk
 * [n_frames][n_atoms][4] (synchronous).  out_rt: NULL or [n_frames][12] f32 =
 * R row-major then t = cy - R cx (AlignPlan's row, plan_exec.rs:228-243);
 * out_xyz: NULL or [n_frames][n_atoms][3] f32 = R x + t for every atom in f64,
 * cast to f32 (SuperposePlan's series, plan_exec.rs:441-456). */
wmd_status wmd_align_run(wmd_align *al, const float *coords_xyzw, size_t n_atoms, size_t n_frames,
                         float *out_rt, float *out_xyz);
/* The same on device frames [n_frames][n_atoms][4] (16-byte aligned): d_rt64
 * [F][12] f64 (required), d_rt32 [F][12] f32 and d_xyz [F][n_atoms][3] f32
 * (each may be NULL).  Asynchronous on `stream`. */
wmd_status wmd_align_run_device(wmd_align *al, const float *d_xyzw, size_t n_atoms,
                                size_t n_frames, double *d_rt64, float *d_rt32, float *d_xyz,
                                void *stream);
void wmd_align_destroy(wmd_align *al);

/* Gaussian-smoothed RDF (no counterpart in the reference; the differentiable
 * RDF of structure-matching losses such as warp-cg's objectives,
 * crates/warp-cg/src/optimize/objective.rs:271-367).  Over the unordered
 * pairs of one atom set with the RdfPlan minimum-image rule, in
 * particle-mesh form: every ordered pair deposits the quadratic B-spline
 * B(r_ij / h - m) on nodes m h, h = w / 4, w = r_max / bins, giving H_m; then
 *   c_k = sum over |m h - r_k| <= 6 sigma of H_m w phi_sigma(r_k - m h),
 * r_k = (k + 1/2) w, phi_sigma the N(0, sigma^2) density (so each pair
 * contributes the Gaussian convolved with B; c is C1 in the positions).
 * sigma must lie in [w / 4, 4 w].  pbc: WMD_PBC_*. */
wmd_status wmd_softrdf_create(wmd_ctx *ctx, size_t bins, float r_max, float sigma, int pbc,
                              wmd_softrdf **out);
/* d_soa [F][3][stride] f32 device store of n_atoms atoms; boxes_host [F][9]
 * row-major box vectors (NULL without PBC).  Adds c_k summed over the frames
 * to d_counts [bins] (f64, device).  Asynchronous on `stream`. */
wmd_status wmd_softrdf_forward(wmd_softrdf *h, const float *d_soa, size_t stride, size_t n_atoms,
                               size_t n_frames, const float *boxes_host, double *d_counts,
                               void *stream);
/* Adds dL/dx to d_grad [F][3][stride] (f32, device) given d_dcounts [bins]
 * (f64, device) = dL/dc_k.  The pairs are enumerated again (no per-pair term
 * is kept from the forward; see wmd_softrdf_keep_cells for its cell lists).
 * Asynchronous on `stream`. */
wmd_status wmd_softrdf_backward(wmd_softrdf *h, const float *d_soa, size_t stride, size_t n_atoms,
                                size_t n_frames, const float *boxes_host, const double *d_dcounts,
                                float *d_grad, void *stream);
/* Enumeration of the last call: *path 1 = brute force, 2 = cell list (nc [3]). */
wmd_status wmd_softrdf_last_path(const wmd_softrdf *h, int *path, int *nc);
/* cell_k: stencil reach in cells (0 = automatic, 1..3; -1 = brute force). */
wmd_status wmd_softrdf_set_tuning(wmd_softrdf *h, int cell_k);
/* keep != 0: a forward whose cell lists cover all its frames keeps them, and
 * the next backward called with the same d_soa, stride, n_atoms, n_frames and
 * boxes takes them instead of sorting the frames again; the caller promises
 * that the store's contents did not change in between (an autograd wrapper
 * that owns the store).  Default 0.  No counterpart in the reference (the
 * differentiable path is new). */
wmd_status wmd_softrdf_keep_cells(wmd_softrdf *h, int keep);
void wmd_softrdf_destroy(wmd_softrdf *h);

/* ---- uneven batches ---------------------------------------------------------
 * Many independent trajectories (uneven atom counts, own boxes and frame
 * counts) through one RDF pair launch and one RMSD launch, replayed as a
 * captured HIP graph.  No single counterpart in the reference, which runs one
 * trajectory per Plan.run() with a fresh context (crates/traj-py/src/
 * pack.rs:1320-1323) and tunes only its chunk size (pack.rs:1366-1405); per
 * trajectory the results equal RdfPlan's counts over all atoms of the
 * trajectory (rdf.rs:882-941; bit-exact) and RmsdPlan(all atoms,
 * reference="frame0", align=True) (rmsd/plan.rs:336-392). */
#define WMD_BATCH_RDF 0x1u      /* histogram every trajectory (all atoms, A = B) */
#define WMD_BATCH_RMSD 0x2u     /* RMSD of every frame to its trajectory's frame 0 */
#define WMD_BATCH_NO_GRAPH 0x4u /* launch eagerly instead of replaying a HIP graph */

typedef struct wmd_batch wmd_batch;

/* n_atoms, n_frames: [n_traj].  bins <= 4194304 (histograms beyond the pair
 * kernel's per-wave LDS copies are counted one slice of bins per pass),
 * r_max > 0, pbc WMD_PBC_*. */
wmd_status wmd_batch_create(wmd_ctx *ctx, size_t n_traj, const uint32_t *n_atoms,
                            const uint32_t *n_frames, size_t bins, float r_max, int pbc,
                            uint32_t flags, wmd_batch **out);
/* Store layout: trajectory t's frames are SoA planes [n_frames[t]][3][strides[t]]
 * (float) starting at float offset offsets[t] of one device buffer of
 * *total_floats floats; padding slots (atom >= n_atoms[t]) must hold 0.
 * offsets / strides: [n_traj] or NULL. */
wmd_status wmd_batch_store_layout(const wmd_batch *b, uint64_t *offsets, uint32_t *strides,
                                  uint64_t *total_floats);
/* Boxes of every frame, [sum n_frames][9] row-major (trajectory order; NULL for
 * WMD_PBC_NONE).  Plans the enumeration (cell list or brute force per
 * trajectory, units, XCD slices); must precede the first run. */
wmd_status wmd_batch_set_boxes(wmd_batch *b, const float *boxes);
/* Per-trajectory tuning override, each array [n_traj] or NULL (all 0):
 * cell_k[t] 0 automatic, -1 brute force, 1..3 cells per r_max;
 * home_block[t] 0 (64) or a multiple of 64 up to 512: home atoms of one
 * brute-force unit; grab[t] 0 (2) or 1..64: units per queue task.  Re-plans. */
wmd_status wmd_batch_set_tuning(wmd_batch *b, const int *cell_k, const int *home_block,
                                const int *grab);
/* Project autotuner: for every class of trajectories sharing (device arch,
 * pbc, bins, r_max, atom-count bucket, density bucket) it times, on a
 * sub-batch of the class's members (a frame prefix of each, replicated until
 * one run holds about 4e6 atom-frames; reps timed runs after a warm-up):
 * brute force and cells per r_max k = 1, 2, 3; then, for brute force, home
 * blocks of 128 and 256 atoms; then 1, 4 and 8 units per task, keeping the
 * fastest at each step.  Choices are cached per process (see
 * wmd_autotune_export / import) and applied to the batch. */
wmd_status wmd_batch_autotune(wmd_batch *b, const float *d_store, int reps, void *stream);
/* One step over the device store: adds every trajectory's counts to the
 * batch's u64 histograms [n_traj][bins] (WMD_BATCH_RDF) and writes d_rmsd
 * [sum n_frames] f32 (and, if not NULL, d_rmsd64 f64 and d_err f64 error
 * bounds) (WMD_BATCH_RMSD).  Without WMD_BATCH_NO_GRAPH the step is captured
 * into a HIP graph on first use (and again when a pointer changes) and
 * replayed.  Asynchronous on `stream` (NULL = the null stream). */
wmd_status wmd_batch_run_device(wmd_batch *b, const float *d_store, float *d_rmsd,
                                double *d_rmsd64, double *d_err, void *stream);
/* Copies the histograms out, [n_traj][bins] u64 (synchronises). */
wmd_status wmd_batch_read_counts(wmd_batch *b, uint64_t *counts);
/* Zeroes the histograms. */
wmd_status wmd_batch_reset(wmd_batch *b);
/* The plan, arrays [n_traj] or NULL: cell_k[t] (0 = brute force),
 * home_block[t] (brute force; 0 for cells), grab[t] (units per task), and the
 * number of pair units. */
wmd_status wmd_batch_plan_info(const wmd_batch *b, int *cell_k, int *home_block, int *grab,
                               long long *n_units);
/* Wall-clock marks of the last step's pair kernel per queue slice (8 slices,
 * one per XCD): out[8][4] = (first wave start, slice drained, last wave end)
 * in ms from the earliest start (-1 if unset), and the number of the slice's
 * tasks that waves of other slices took.  Synchronises. */
wmd_status wmd_batch_slice_times(wmd_batch *b, double *out);
void wmd_batch_destroy(wmd_batch *b);
/* The autotuner's process-wide cache as text, one choice per line
 * ("arch pbc bins r_max_bits n_bucket density_bucket k home_block grab"); *len gets the full
 * length, buf (cap bytes, NUL-terminated) may be NULL. */
wmd_status wmd_autotune_export(char *buf, size_t cap, size_t *len);
/* Adds lines in the export format to the cache (lines of the first seven
 * fields only keep the default home block and task size). */
wmd_status wmd_autotune_import(const char *text);

/* QCP superposition of one set of centred moments, evaluated on the host with
 * the same code the device runs (for testing without a GPU).
 * m: [9] row-major M = sum (x - cx)(y - cy)^T, gx = sum |x - cx|^2,
 * gy = sum |y - cy|^2, n = atom count.  Outputs: *rmsd (fitted RMSD over the
 * same atoms), rot [9] row-major R, *lambda (largest eigenvalue of the key
 * matrix), *iters (Newton iterations; -1 when the Jacobi fallback ran). */
wmd_status wmd_qcp_host(const double *m, double gx, double gy, double n, double *rmsd,
                        double *rot, double *lambda, int *iters);

#ifdef __cplusplus
}
#endif

#endif /* WMD_HIP_H */
