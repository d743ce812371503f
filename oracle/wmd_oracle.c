/*
 * wmd_oracle.c -- TEST INFRASTRUCTURE ONLY.
 *
 * CPU restatement of warp-md's reference CPU semantics for the RdfPlan /
 * RmsdPlan hot path.  It exists so that tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg can check (and time) the HIP engine.  The product
 * library (warp-md_amd) never links, loads or calls anything in this file.
 *
 * Build: see oracle/Makefile (gcc -O2 -ffp-contract=off, no fast-math, so every
 * f32 operation below rounds exactly once, like the reference's Rust code).
 *
 * Citations are relative to the warp-md 0.4.30 source tree.
 *
 * Function -> what pins it (tests/test_oracle_*.py):
 *   orc_rdf_brute        reference known answers (correlation_structure_tests.rs:1896-2115,
 *                        test_rdf_plan.py:11-45,267-297), lattice shell closed form,
 *                        brute-force python loop on tiny inputs, symmetry invariants.
 *   orc_rdf_cells        bit-identical to orc_rdf_brute on seeded inputs (ortho + triclinic).
 *   orc_rdf_finalize     reference numeric tests (correlation_structure_tests.rs:1916-1976)
 *                        and a float64 numpy restatement of the normalisation.
 *   orc_rmsd_plan        rmsd_plan_align_zero (analysis_smoke_tests.rs:23-37), 90-degree
 *                        rotation recovery (test_rotation_matrix.py:40-59), numpy.linalg.svd
 *                        Kabsch, closed forms (pure rigid motion -> 0, raw RMSD formula).
 *   orc_fit_rmsd         numpy.linalg.svd Kabsch restated independently, rigid-motion zero.
 *   orc_svd3             numpy.linalg.svd on random and rank-deficient matrices.
 *   orc_min_image_f64    pbc_math.rs:166-221 known vectors.
 *   orc_pairdist         reference known answers (correlation_structure_tests.rs:2193-2231,
 *                        2313-2422); equal to orc_rdf_brute where the pair rules coincide.
 *   orc_pairwise_dme     a hand-computed 3-atom case; exact doubling of every coordinate (the
 *                        matrix entry is then the RMS of the frame's pair distances); scipy
 *                        pdist pair order and values; invariance under rigid motions and
 *                        under lattice translations of single atoms with PBC.
 *   orc_pairwise_rmsd    pairwise_rmsd_plan_basic (geometry_matrix_tests.rs:1438-1460),
 *                        equal to orc_rmsd_plan per pair where the two coincide, numpy.linalg.svd
 *                        Kabsch, rigid-motion zero, symmetry, triangle inequality (nofit).
 *   orc_rdf_grouped      the reference's grouped / xy known answers
 *                        (correlation_structure_tests.rs:1979-2008, 2455-2588);
 *                        equal to orc_rdf_brute for one-atom groups without exclusions;
 *                        a pure-Python restatement on small inputs.
 *   orc_align            superpose_plan_basic (analysis_smoke_tests.rs:171-195),
 *                        test_superpose.py:41-67 (rotation + translation recovered),
 *                        numpy.linalg.svd weighted Kabsch restated independently.
 *   orc_sf_counts / orc_sf_finalize
 *                        structure_factor_shapes (correlation_structure_tests.rs:759-780),
 *                        an independent numpy float64 restatement, closed-form lattice shell counts.
 */
#include <math.h>
#include <stddef.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* FP contraction is disabled by the build (-ffp-contract=off); gcc ignores the STDC pragma. */

#define ORC_OK 0
#define ORC_ERR_MISMATCH 2
#define ORC_ERR_PARSE 3
#define ORC_ERR_INVALID 6

enum { PBC_NONE = 0, PBC_ORTHO = 1, PBC_TRICLINIC = 2 };

/* ------------------------------------------------------------------------- */
/* Minimum image                                                              */
/* ------------------------------------------------------------------------- */

/* pbc_math.rs:110-164 cell_and_inv_from_box (Triclinic arm), f64. Rows of
 * `cell` are the box vectors m[0..3], m[3..6], m[6..9]. Returns 0 or
 * ORC_ERR_MISMATCH when the matrix is singular. */
static int cell_and_inv(const float *m, double cell[3][3], double inv[3][3]) {
    double m0 = m[0], m1 = m[1], m2 = m[2], m3 = m[3], m4 = m[4], m5 = m[5],
           m6 = m[6], m7 = m[7], m8 = m[8];
    double det = m0 * (m4 * m8 - m5 * m7) - m1 * (m3 * m8 - m5 * m6) + m2 * (m3 * m7 - m4 * m6);
    if (det == 0.0) return ORC_ERR_MISMATCH;
    inv[0][0] = (m4 * m8 - m5 * m7) / det;
    inv[0][1] = (m2 * m7 - m1 * m8) / det;
    inv[0][2] = (m1 * m5 - m2 * m4) / det;
    inv[1][0] = (m5 * m6 - m3 * m8) / det;
    inv[1][1] = (m0 * m8 - m2 * m6) / det;
    inv[1][2] = (m2 * m3 - m0 * m5) / det;
    inv[2][0] = (m3 * m7 - m4 * m6) / det;
    inv[2][1] = (m1 * m6 - m0 * m7) / det;
    inv[2][2] = (m0 * m4 - m1 * m3) / det;
    cell[0][0] = m0; cell[0][1] = m1; cell[0][2] = m2;
    cell[1][0] = m3; cell[1][1] = m4; cell[1][2] = m5;
    cell[2][0] = m6; cell[2][1] = m7; cell[2][2] = m8;
    return ORC_OK;
}

/* pbc_math.rs:17-27 apply_pbc (f64, with the l > 0 guards). */
static void apply_pbc_f64(double *dx, double *dy, double *dz, double lx, double ly, double lz) {
    if (lx > 0.0) *dx -= round(*dx / lx) * lx;
    if (ly > 0.0) *dy -= round(*dy / ly) * ly;
    if (lz > 0.0) *dz -= round(*dz / lz) * lz;
}

/* pbc_math.rs:92-108 apply_pbc_triclinic (f64). */
static void apply_pbc_triclinic_f64(double *dx, double *dy, double *dz,
                                    double cell[3][3], double inv[3][3]) {
    double fx = inv[0][0] * *dx + inv[1][0] * *dy + inv[2][0] * *dz;
    double fy = inv[0][1] * *dx + inv[1][1] * *dy + inv[2][1] * *dz;
    double fz = inv[0][2] * *dx + inv[1][2] * *dy + inv[2][2] * *dz;
    fx = fx - round(fx);
    fy = fy - round(fy);
    fz = fz - round(fz);
    *dx = fx * cell[0][0] + fy * cell[1][0] + fz * cell[2][0];
    *dy = fx * cell[0][1] + fy * cell[1][1] + fz * cell[2][1];
    *dz = fx * cell[0][2] + fy * cell[1][2] + fz * cell[2][2];
}

/* Exposed for the pbc_math.rs:166-221 known-vector tests.  kind: 0 none,
 * 1 orthorhombic (box = lx,ly,lz in m[0],m[4],m[8]), 2 triclinic (m row-major). */
int orc_min_image_f64(const double *a, const double *b, int kind, const float *m,
                      double length_scale, double *out) {
    double dx = b[0] - a[0], dy = b[1] - a[1], dz = b[2] - a[2];
    if (kind == PBC_ORTHO) {
        apply_pbc_f64(&dx, &dy, &dz, (double)m[0] * length_scale, (double)m[4] * length_scale,
                      (double)m[8] * length_scale);
    } else if (kind == PBC_TRICLINIC) {
        double cell[3][3], inv[3][3];
        if (cell_and_inv(m, cell, inv) != ORC_OK) return ORC_ERR_MISMATCH;
        for (int r = 0; r < 3; ++r)
            for (int c = 0; c < 3; ++c) {
                cell[r][c] *= length_scale;
                inv[r][c] /= length_scale;
            }
        apply_pbc_triclinic_f64(&dx, &dy, &dz, cell, inv);
    }
    out[0] = dx; out[1] = dy; out[2] = dz;
    return ORC_OK;
}

/* Per-frame f32 box data used by the pair expression. */
typedef struct {
    int kind;
    float lx, ly, lz;              /* orthorhombic */
    float c[3][3];                 /* triclinic rows (box vectors), f32 */
    float r[3][3];                 /* triclinic inverse rows, f32 casts of the f64 inverse */
    double cell64[3][3], inv64[3][3];
} frame_box;

static int load_box(int kind, const float *m, frame_box *fb) {
    fb->kind = kind;
    if (kind == PBC_ORTHO) {
        fb->lx = m[0]; fb->ly = m[4]; fb->lz = m[8];
    } else if (kind == PBC_TRICLINIC) {
        if (cell_and_inv(m, fb->cell64, fb->inv64) != ORC_OK) return ORC_ERR_MISMATCH;
        /* geometry_math.rs:143-166 chunk_cell_mats: f32 rows of cell and inverse. */
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) {
                fb->c[i][j] = (float)fb->cell64[i][j];
                fb->r[i][j] = (float)fb->inv64[i][j];
            }
    }
    return ORC_OK;
}

/* The per-pair f32 expression.  Orthorhombic: rdf.rs:921-929 (f32 subtract,
 * dx -= round(dx/lx)*lx with Rust round = C roundf, left-to-right adds in
 * rdf_distance rdf.rs:473-478, IEEE sqrt).  Triclinic (the reference RdfPlan
 * refuses it, rdf.rs:797-802; this order is the engine's own definition): the
 * image n = roundf(f) of the f32 fractional displacement f = inv . d as in the
 * reference device formula wrap_triclinic (traj-kernels part1.cu:12-27), then
 * the lattice vector taken off in Cartesian space, d -= (n0 a + n1 b) + n2 c,
 * each component one f32 sum of f32 products; no FMA contraction.  Returns the
 * f32 distance. */
static inline float pair_r32(const float *pa, const float *pb, const frame_box *fb) {
    float dx = pb[0] - pa[0];
    float dy = pb[1] - pa[1];
    float dz = pb[2] - pa[2];
    if (fb->kind == PBC_ORTHO) {
        dx -= roundf(dx / fb->lx) * fb->lx;
        dy -= roundf(dy / fb->ly) * fb->ly;
        dz -= roundf(dz / fb->lz) * fb->lz;
    } else if (fb->kind == PBC_TRICLINIC) {
        const float n0 = roundf(fb->r[0][0] * dx + fb->r[1][0] * dy + fb->r[2][0] * dz);
        const float n1 = roundf(fb->r[0][1] * dx + fb->r[1][1] * dy + fb->r[2][1] * dz);
        const float n2 = roundf(fb->r[0][2] * dx + fb->r[1][2] * dy + fb->r[2][2] * dz);
        dx -= n0 * fb->c[0][0] + n1 * fb->c[1][0] + n2 * fb->c[2][0];
        dy -= n0 * fb->c[0][1] + n1 * fb->c[1][1] + n2 * fb->c[2][1];
        dz -= n0 * fb->c[0][2] + n1 * fb->c[1][2] + n2 * fb->c[2][2];
    }
    return sqrtf(dx * dx + dy * dy + dz * dz);
}

/* f64 restatement of the same displacement (pbc_math.rs:17-27, 92-108) for the
 * edge diagnostics. */
static inline double pair_r64(const float *pa, const float *pb, const frame_box *fb) {
    double dx = (double)pb[0] - (double)pa[0];
    double dy = (double)pb[1] - (double)pa[1];
    double dz = (double)pb[2] - (double)pa[2];
    if (fb->kind == PBC_ORTHO) {
        apply_pbc_f64(&dx, &dy, &dz, fb->lx, fb->ly, fb->lz);
    } else if (fb->kind == PBC_TRICLINIC) {
        double cell[3][3], inv[3][3];
        memcpy(cell, fb->cell64, sizeof(cell));
        memcpy(inv, fb->inv64, sizeof(inv));
        apply_pbc_triclinic_f64(&dx, &dy, &dz, cell, inv);
    }
    return sqrt(dx * dx + dy * dy + dz * dz);
}

/* rdf.rs:929-935: r < r_max, bin = (r / bin_width) as usize, bin < bins.
 * Returns the bin or -1. */
static inline long bin_of(float r, float r_max, float bin_width, size_t bins) {
    if (!(r < r_max)) return -1;
    float q = r / bin_width;
    size_t bin = (size_t)q; /* r >= 0: truncation toward zero */
    if (bin < bins) return (long)bin;
    return -1;
}

/* Bin-edge diagnostics: edges are k * bin_width (k = 1..bins) in f64. */
static inline int near_edge(double r, double w, size_t bins, double tol) {
    double k = floor(r / w + 0.5);
    if (k < 1.0) k = 1.0;
    if (k > (double)bins) k = (double)bins;
    for (int d = -1; d <= 1; ++d) {
        double kk = k + d;
        if (kk < 1.0 || kk > (double)bins) continue;
        if (fabs(r - kk * w) < tol) return 1;
    }
    return 0;
}

#define EDGE_TOL 1.0e-5

typedef struct {
    uint64_t edge32;   /* counted pairs whose f32 distance lies within 1e-5 A of a bin edge */
    uint64_t edge64;   /* pairs whose f64 distance lies within 1e-5 A of a bin edge */
    uint64_t mismatch; /* pairs whose f32 and f64 bins differ */
} edge_stats;

static inline void pair_account(const float *pa, const float *pb, const frame_box *fb, float r_max,
                                float bin_width, size_t bins, uint64_t *counts, edge_stats *es) {
    float r = pair_r32(pa, pb, fb);
    long bin = bin_of(r, r_max, bin_width, bins);
    if (bin >= 0) counts[bin] += 1;
    if (es) {
        double w = (double)bin_width;
        double r64 = pair_r64(pa, pb, fb);
        if (bin >= 0 && near_edge((double)r, w, bins, EDGE_TOL)) es->edge32 += 1;
        if (r64 < (double)r_max + 2.0 * EDGE_TOL && near_edge(r64, w, bins, EDGE_TOL))
            es->edge64 += 1;
        long b64 = -1;
        if (r64 < (double)r_max) {
            double q = r64 / w;
            if (q < (double)bins) b64 = (long)q;
        }
        if (b64 != bin) es->mismatch += 1;
    }
}

/*
 * RdfPlan::process_chunk pair loop (rdf.rs:851-876; the selected-input twin is
 * rdf.rs:882-941) -- brute force over the ordered pairs (a in sel_a, b in
 * sel_b), skipping a == b by atom index, in the reference loop order.
 *
 * coords: [n_frames][n_atoms][stride] floats (stride 3 or 4, the reference
 * FrameChunk.coords is stride 4).  boxes: [n_frames][9] row-major box vectors
 * (orthorhombic reads m[0], m[4], m[8]); ignored for pbc = none.
 * counts: [bins] u64, accumulated into (not cleared).  edge: NULL or [3].
 */
int orc_rdf_brute(const float *coords, size_t stride, size_t n_atoms, size_t n_frames,
                  const uint32_t *sel_a, size_t n_a, const uint32_t *sel_b, size_t n_b,
                  size_t bins, float r_max, int pbc, const float *boxes, uint64_t *counts,
                  uint64_t *edge) {
    if (bins == 0) return ORC_OK;
    for (size_t i = 0; i < n_a; ++i)
        if (sel_a[i] >= n_atoms) return ORC_ERR_INVALID;
    for (size_t i = 0; i < n_b; ++i)
        if (sel_b[i] >= n_atoms) return ORC_ERR_INVALID;
    float bin_width = r_max / (float)bins; /* rdf.rs:896 */
    edge_stats es = {0, 0, 0};
    for (size_t f = 0; f < n_frames; ++f) {
        frame_box fb;
        if (load_box(pbc, boxes ? boxes + 9 * f : NULL, &fb) != ORC_OK) return ORC_ERR_MISMATCH;
        const float *base = coords + f * n_atoms * stride;
        for (size_t ia = 0; ia < n_a; ++ia) {
            uint32_t a = sel_a[ia];
            const float *pa = base + (size_t)a * stride;
            for (size_t ib = 0; ib < n_b; ++ib) {
                uint32_t b = sel_b[ib];
                if (a == b) continue;
                pair_account(pa, base + (size_t)b * stride, &fb, r_max, bin_width, bins, counts,
                             edge ? &es : NULL);
            }
        }
    }
    if (edge) {
        edge[0] += es.edge32;
        edge[1] += es.edge64;
        edge[2] += es.mismatch;
    }
    return ORC_OK;
}

/*
 * Same ordered-pair multiset and the identical per-pair f32 expression as
 * orc_rdf_brute, enumerated through a periodic cell list so that large frames
 * finish in seconds.  Because counts are integers and every visited pair is
 * evaluated with pair_r32, the result is bit-identical to brute force whenever
 * the cell stencil covers every pair closer than r_max; the cells are sized from
 * the box's plane spacings with a 1e-3 A margin to guarantee that.  Falls back
 * to brute force (return value 1) when the box is too small for a 3x3x3 stencil
 * or pbc is none.
 */
int orc_rdf_cells(const float *coords, size_t stride, size_t n_atoms, size_t n_frames,
                  const uint32_t *sel_a, size_t n_a, const uint32_t *sel_b, size_t n_b,
                  size_t bins, float r_max, int pbc, const float *boxes, uint64_t *counts) {
    if (bins == 0) return ORC_OK;
    if (pbc == PBC_NONE)
        return orc_rdf_brute(coords, stride, n_atoms, n_frames, sel_a, n_a, sel_b, n_b, bins,
                             r_max, pbc, boxes, counts, NULL) == ORC_OK ? 1 : ORC_ERR_INVALID;
    for (size_t i = 0; i < n_a; ++i)
        if (sel_a[i] >= n_atoms) return ORC_ERR_INVALID;
    for (size_t i = 0; i < n_b; ++i)
        if (sel_b[i] >= n_atoms) return ORC_ERR_INVALID;
    float bin_width = r_max / (float)bins;
    int fell_back = 0;
    int *cell_of_b = (int *)malloc(sizeof(int) * (n_b ? n_b : 1));
    for (size_t f = 0; f < n_frames; ++f) {
        const float *m = boxes + 9 * f;
        frame_box fb;
        if (load_box(pbc, m, &fb) != ORC_OK) { free(cell_of_b); return ORC_ERR_MISMATCH; }
        double cell[3][3], inv[3][3];
        if (pbc == PBC_ORTHO) {
            memset(cell, 0, sizeof(cell));
            memset(inv, 0, sizeof(inv));
            cell[0][0] = m[0]; cell[1][1] = m[4]; cell[2][2] = m[8];
            inv[0][0] = 1.0 / m[0]; inv[1][1] = 1.0 / m[4]; inv[2][2] = 1.0 / m[8];
        } else {
            memcpy(cell, fb.cell64, sizeof(cell));
            memcpy(inv, fb.inv64, sizeof(inv));
        }
        /* plane spacing along fractional axis d = 1 / |row d of inv^T| = 1/|column d of inv| */
        int nc[3];
        double reach = (double)r_max + 1e-3;
        int ok = 1;
        for (int d = 0; d < 3; ++d) {
            double nrm = sqrt(inv[0][d] * inv[0][d] + inv[1][d] * inv[1][d] + inv[2][d] * inv[2][d]);
            double width = 1.0 / nrm;
            nc[d] = (int)floor(width / reach);
            if (nc[d] > 64) nc[d] = 64;
            if (nc[d] < 3) ok = 0;
        }
        const float *base = coords + f * n_atoms * stride;
        if (!ok) {
            fell_back = 1;
            if (orc_rdf_brute(base, stride, n_atoms, 1, sel_a, n_a, sel_b, n_b, bins, r_max, pbc,
                              m, counts, NULL) != ORC_OK) {
                free(cell_of_b);
                return ORC_ERR_INVALID;
            }
            continue;
        }
        size_t ncell = (size_t)nc[0] * nc[1] * nc[2];
        size_t *start = (size_t *)calloc(ncell + 1, sizeof(size_t));
        size_t *order = (size_t *)malloc(sizeof(size_t) * (n_b ? n_b : 1));
        for (size_t ib = 0; ib < n_b; ++ib) {
            const float *p = base + (size_t)sel_b[ib] * stride;
            int ci[3];
            for (int d = 0; d < 3; ++d) {
                double fr = inv[0][d] * p[0] + inv[1][d] * p[1] + inv[2][d] * p[2];
                fr -= floor(fr);
                int c = (int)floor(fr * nc[d]);
                if (c < 0) c = 0;
                if (c >= nc[d]) c = nc[d] - 1;
                ci[d] = c;
            }
            int id = (ci[2] * nc[1] + ci[1]) * nc[0] + ci[0];
            cell_of_b[ib] = id;
            start[id + 1] += 1;
        }
        for (size_t c = 0; c < ncell; ++c) start[c + 1] += start[c];
        size_t *fill = (size_t *)malloc(sizeof(size_t) * (ncell ? ncell : 1));
        memcpy(fill, start, sizeof(size_t) * ncell);
        for (size_t ib = 0; ib < n_b; ++ib) order[fill[cell_of_b[ib]]++] = ib;
        free(fill);
        for (size_t ia = 0; ia < n_a; ++ia) {
            uint32_t a = sel_a[ia];
            const float *pa = base + (size_t)a * stride;
            int ci[3];
            for (int d = 0; d < 3; ++d) {
                double fr = inv[0][d] * pa[0] + inv[1][d] * pa[1] + inv[2][d] * pa[2];
                fr -= floor(fr);
                int c = (int)floor(fr * nc[d]);
                if (c < 0) c = 0;
                if (c >= nc[d]) c = nc[d] - 1;
                ci[d] = c;
            }
            for (int oz = -1; oz <= 1; ++oz)
                for (int oy = -1; oy <= 1; ++oy)
                    for (int ox = -1; ox <= 1; ++ox) {
                        int cx = (ci[0] + ox + nc[0]) % nc[0];
                        int cy = (ci[1] + oy + nc[1]) % nc[1];
                        int cz = (ci[2] + oz + nc[2]) % nc[2];
                        int id = (cz * nc[1] + cy) * nc[0] + cx;
                        for (size_t k = start[id]; k < start[id + 1]; ++k) {
                            uint32_t b = sel_b[order[k]];
                            if (a == b) continue;
                            float r = pair_r32(pa, base + (size_t)b * stride, &fb);
                            long bin = bin_of(r, r_max, bin_width, bins);
                            if (bin >= 0) counts[bin] += 1;
                        }
                    }
        }
        free(start);
        free(order);
    }
    free(cell_of_b);
    return fell_back;
}

/*
 * RdfPlan::finalize (rdf.rs:943-1016) with rdf_shell_measure (rdf.rs:487-496)
 * for the 3-D dimension.  The shell measure is evaluated in f32 exactly as the
 * Rust expression ((4.0/3.0) * PI * (ro.powi(3) - ri.powi(3))) as f64, with
 * powi(3) = (x*x)*x.  want_integral selects the intrdf output.
 * Returns ORC_ERR_MISMATCH for use_volume with volume_sum <= 0 (rdf.rs:962-966).
 */
int orc_rdf_finalize(const uint64_t *counts, size_t bins, float r_max, double n_a, double n_b,
                     double common, size_t frames_in, double volume_sum, double density_opt,
                     int use_volume, int raw_output, int number_density_output, int want_integral,
                     float *r_out, float *g_out, float *integral_out) {
    float bin_width = r_max / (float)bins;
    double frames = (double)(frames_in > 1 ? frames_in : 1);
    double total_pairs = n_a * n_b - common;
    if (total_pairs < 0.0) total_pairs = 0.0;
    double density;
    if (use_volume) {
        if (volume_sum <= 0.0) return ORC_ERR_MISMATCH;
        density = total_pairs / (volume_sum / frames);
    } else if (n_a > 0.0) {
        density = density_opt * total_pairs / n_a;
    } else {
        density = 0.0;
    }
    const float four_thirds = 4.0f / 3.0f;
    const float pi32 = 3.14159265358979323846f;
    double integral_sum = 0.0;
    for (size_t i = 0; i < bins; ++i) {
        float r_inner = (float)i * bin_width;
        float r_outer = (float)(i + 1) * bin_width;
        float ro3 = (r_outer * r_outer) * r_outer;
        float ri3 = (r_inner * r_inner) * r_inner;
        float shell32 = four_thirds * pi32 * (ro3 - ri3);
        double shell = (double)shell32;
        double expected = shell * density;
        double norm = expected > 0.0 ? frames * expected : 0.0;
        r_out[i] = r_inner + 0.5f * bin_width;
        double count = (double)counts[i];
        double normalized = norm > 0.0 ? count / norm : 0.0;
        double nd_norm = frames * n_a * shell;
        double number_density = nd_norm > 0.0 ? count / nd_norm : 0.0;
        if (want_integral) {
            if (number_density_output) {
                integral_sum += number_density * shell;
            } else if (n_b > 0.0) {
                integral_sum += normalized * expected / n_b;
            }
            integral_out[i] = (float)integral_sum;
        }
        double value;
        if (raw_output) value = count;
        else if (number_density_output) value = number_density;
        else value = normalized;
        g_out[i] = (float)value;
    }
    return ORC_OK;
}

/* ------------------------------------------------------------------------- */
/* 3x3 SVD (f64) and Kabsch                                                   */
/* ------------------------------------------------------------------------- */

/* One-sided Jacobi (Hestenes) SVD of a 3x3 matrix a = U diag(s) V^T, singular
 * values sorted descending (the order nalgebra 0.32's SVD::new returns, which
 * the reference relies on when it flips row 2 of V^T).  U is completed with a
 * cross product when the third singular value vanishes. */
void orc_svd3(const double *a_in, double *u, double *s, double *v) {
    double b[3][3], vv[3][3];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            b[i][j] = a_in[3 * i + j];
            vv[i][j] = (i == j) ? 1.0 : 0.0;
        }
    for (int sweep = 0; sweep < 60; ++sweep) {
        double off = 0.0;
        for (int p = 0; p < 2; ++p)
            for (int q = p + 1; q < 3; ++q) {
                double alpha = 0, beta = 0, gamma = 0;
                for (int i = 0; i < 3; ++i) {
                    alpha += b[i][p] * b[i][p];
                    beta += b[i][q] * b[i][q];
                    gamma += b[i][p] * b[i][q];
                }
                if (gamma == 0.0) continue;
                double conv = fabs(gamma) / sqrt(alpha * beta);
                if (!(conv > 1e-15)) continue;
                if (conv > off) off = conv;
                double zeta = (beta - alpha) / (2.0 * gamma);
                double t = (zeta >= 0 ? 1.0 : -1.0) / (fabs(zeta) + sqrt(1.0 + zeta * zeta));
                double c = 1.0 / sqrt(1.0 + t * t);
                double sn = c * t;
                for (int i = 0; i < 3; ++i) {
                    double bp = b[i][p], bq = b[i][q];
                    b[i][p] = c * bp - sn * bq;
                    b[i][q] = sn * bp + c * bq;
                    double vp = vv[i][p], vq = vv[i][q];
                    vv[i][p] = c * vp - sn * vq;
                    vv[i][q] = sn * vp + c * vq;
                }
            }
        if (off <= 1e-15) break;
    }
    double sig[3];
    int idx[3] = {0, 1, 2};
    for (int j = 0; j < 3; ++j)
        sig[j] = sqrt(b[0][j] * b[0][j] + b[1][j] * b[1][j] + b[2][j] * b[2][j]);
    for (int i = 0; i < 3; ++i)
        for (int j = i + 1; j < 3; ++j)
            if (sig[idx[j]] > sig[idx[i]]) { int t = idx[i]; idx[i] = idx[j]; idx[j] = t; }
    double uu[3][3];
    double smax = sig[idx[0]];
    for (int k = 0; k < 3; ++k) {
        int j = idx[k];
        s[k] = sig[j];
        for (int i = 0; i < 3; ++i) v[3 * i + k] = vv[i][j];
        if (sig[j] > 1e-300 && sig[j] > smax * 1e-14) {
            for (int i = 0; i < 3; ++i) uu[i][k] = b[i][j] / sig[j];
        } else {
            for (int i = 0; i < 3; ++i) uu[i][k] = 0.0;
        }
    }
    /* complete U for rank-deficient inputs */
    double n0 = uu[0][0] * uu[0][0] + uu[1][0] * uu[1][0] + uu[2][0] * uu[2][0];
    if (n0 == 0.0) { uu[0][0] = 1.0; uu[1][0] = 0.0; uu[2][0] = 0.0; }
    double n1 = uu[0][1] * uu[0][1] + uu[1][1] * uu[1][1] + uu[2][1] * uu[2][1];
    if (n1 == 0.0) {
        /* any unit vector orthogonal to column 0 */
        double ax = fabs(uu[0][0]) < 0.9 ? 1.0 : 0.0, ay = ax == 1.0 ? 0.0 : 1.0;
        double cx = uu[1][0] * 0.0 - uu[2][0] * ay;
        double cy = uu[2][0] * ax - uu[0][0] * 0.0;
        double cz = uu[0][0] * ay - uu[1][0] * ax;
        double nn = sqrt(cx * cx + cy * cy + cz * cz);
        uu[0][1] = cx / nn; uu[1][1] = cy / nn; uu[2][1] = cz / nn;
    }
    double n2 = uu[0][2] * uu[0][2] + uu[1][2] * uu[1][2] + uu[2][2] * uu[2][2];
    if (n2 == 0.0) {
        uu[0][2] = uu[1][0] * uu[2][1] - uu[2][0] * uu[1][1];
        uu[1][2] = uu[2][0] * uu[0][1] - uu[0][0] * uu[2][1];
        uu[2][2] = uu[0][0] * uu[1][1] - uu[1][0] * uu[0][1];
    }
    for (int i = 0; i < 3; ++i)
        for (int k = 0; k < 3; ++k) u[3 * i + k] = uu[i][k];
}

static double det3(const double r[3][3]) {
    return r[0][0] * (r[1][1] * r[2][2] - r[1][2] * r[2][1]) -
           r[0][1] * (r[1][0] * r[2][2] - r[1][2] * r[2][0]) +
           r[0][2] * (r[1][0] * r[2][1] - r[1][1] * r[2][0]);
}

/* R = V U^T with the reflection fix of rmsd/plan.rs:1219-1224 and
 * geometry_math.rs:289-295 (negate row 2 of V^T, i.e. column 2 of V).
 * Returns 0 when the SVD is unusable (never for finite input). */
static int kabsch_from_h(const double h[3][3], double r[3][3]) {
    double hin[9], u[9], s[3], v[9];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) hin[3 * i + j] = h[i][j];
    for (int i = 0; i < 9; ++i)
        if (!isfinite(hin[i])) return 0;
    orc_svd3(hin, u, s, v);
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            double acc = 0.0;
            for (int k = 0; k < 3; ++k) acc += v[3 * i + k] * u[3 * j + k];
            r[i][j] = acc;
        }
    if (det3(r) < 0.0) {
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) {
                double acc = 0.0;
                for (int k = 0; k < 3; ++k) {
                    double vk = v[3 * i + k] * (k == 2 ? -1.0 : 1.0);
                    acc += vk * u[3 * j + k];
                }
                r[i][j] = acc;
            }
    }
    return 1;
}

/* rmsd/plan.rs:1093-1110 rmsd_raw_selected (f64 accumulation, f32 result). */
static double rmsd_raw_selected(const float *frame, size_t n_frame, size_t stride,
                                const float *ref, size_t n_ref, const uint32_t *sel, size_t n_sel) {
    size_t n = n_sel < n_ref ? n_sel : n_ref;
    if (n == 0) return 0.0;
    double sum = 0.0;
    for (size_t i = 0; i < n; ++i) {
        size_t idx = sel[i];
        if (idx >= n_frame) break;
        double dx = (double)frame[idx * stride + 0] - (double)ref[i * stride + 0];
        double dy = (double)frame[idx * stride + 1] - (double)ref[i * stride + 1];
        double dz = (double)frame[idx * stride + 2] - (double)ref[i * stride + 2];
        sum += dx * dx + dy * dy + dz * dz;
    }
    return sqrt(sum / (double)n);
}

/* rmsd/plan.rs:1161-1248 rmsd_aligned_selected: f64 centroids (sum * (1/n)),
 * f64 H = sum (x-cx)(y-cy)^T, SVD Kabsch, f64 residual pass, sqrt(sum/n). */
static double rmsd_aligned_selected(const float *frame, size_t n_frame, size_t stride,
                                    const float *ref, size_t n_ref, const uint32_t *sel,
                                    size_t n_sel) {
    size_t n = n_sel < n_ref ? n_sel : n_ref;
    if (n == 0) return 0.0;
    double cx[3] = {0, 0, 0}, cy[3] = {0, 0, 0};
    for (size_t i = 0; i < n; ++i) {
        size_t idx = sel[i];
        if (idx >= n_frame) return 0.0;
        const float *p = frame + idx * stride;
        const float *q = ref + i * stride;
        cx[0] += (double)p[0]; cx[1] += (double)p[1]; cx[2] += (double)p[2];
        cy[0] += (double)q[0]; cy[1] += (double)q[1]; cy[2] += (double)q[2];
    }
    double inv_n = 1.0 / (double)n;
    for (int d = 0; d < 3; ++d) { cx[d] *= inv_n; cy[d] *= inv_n; }
    double h[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
    for (size_t i = 0; i < n; ++i) {
        const float *p = frame + (size_t)sel[i] * stride;
        const float *q = ref + i * stride;
        double px = (double)p[0] - cx[0], py = (double)p[1] - cx[1], pz = (double)p[2] - cx[2];
        double qx = (double)q[0] - cy[0], qy = (double)q[1] - cy[1], qz = (double)q[2] - cy[2];
        h[0][0] += px * qx; h[0][1] += px * qy; h[0][2] += px * qz;
        h[1][0] += py * qx; h[1][1] += py * qy; h[1][2] += py * qz;
        h[2][0] += pz * qx; h[2][1] += pz * qy; h[2][2] += pz * qz;
    }
    double r[3][3];
    if (!kabsch_from_h(h, r))
        return rmsd_raw_selected(frame, n_frame, stride, ref, n_ref, sel, n_sel);
    double sum = 0.0;
    for (size_t i = 0; i < n; ++i) {
        const float *p = frame + (size_t)sel[i] * stride;
        const float *q = ref + i * stride;
        double px = (double)p[0] - cx[0], py = (double)p[1] - cx[1], pz = (double)p[2] - cx[2];
        double qx = (double)q[0] - cy[0], qy = (double)q[1] - cy[1], qz = (double)q[2] - cy[2];
        double ax = r[0][0] * px + r[0][1] * py + r[0][2] * pz;
        double ay = r[1][0] * px + r[1][1] * py + r[1][2] * pz;
        double az = r[2][0] * px + r[2][1] * py + r[2][2] * pz;
        double dx = ax - qx, dy = ay - qy, dz = az - qz;
        sum += dx * dx + dy * dy + dz * dz;
    }
    return sqrt(sum * inv_n);
}

/*
 * RmsdPlan over a chunk (rmsd/plan.rs:145-162 process_cpu_chunk_out).
 * coords: [n_frames][n_atoms][stride]; sel: global atom indices;
 * reference: [n_ref][stride], one entry per selection position (positions0[sel]
 * for reference="topology", frame 0's selected coords for "frame0").
 * out: f32 per frame (the reference's `as f32`), out64: the f64 before the cast.
 */
int orc_rmsd_plan(const float *coords, size_t stride, size_t n_atoms, size_t n_frames,
                  const uint32_t *sel, size_t n_sel, const float *reference, size_t n_ref,
                  int align, float *out, double *out64) {
    for (size_t f = 0; f < n_frames; ++f) {
        const float *frame = coords + f * n_atoms * stride;
        double v = align ? rmsd_aligned_selected(frame, n_atoms, stride, reference, n_ref, sel, n_sel)
                         : rmsd_raw_selected(frame, n_atoms, stride, reference, n_ref, sel, n_sel);
        out[f] = (float)v;
        if (out64) out64[f] = v;
    }
    return ORC_OK;
}

/*
 * Fit on one selection, RMSD over another: kabsch_rotation
 * (geometry_math.rs:169-174 -> kabsch_rotation_weighted :235-297, unweighted,
 * centroid = sum / n) on the fit atoms, then the residual of rms.rs:656-672 over
 * the rmsd atoms with the reference coordinates of the same atoms.
 * ref_all: [n_atoms][stride] reference coordinates indexed by atom.
 * rot_out: NULL or [n_frames][9] (row-major R), cx_out/cy_out: NULL or [n_frames][3].
 */
int orc_fit_rmsd(const float *coords, size_t stride, size_t n_atoms, size_t n_frames,
                 const uint32_t *fit_sel, size_t n_fit, const uint32_t *rmsd_sel, size_t n_rmsd,
                 const float *ref_all, float *out, double *out64, double *rot_out, double *cx_out,
                 double *cy_out) {
    for (size_t i = 0; i < n_fit; ++i)
        if (fit_sel[i] >= n_atoms) return ORC_ERR_INVALID;
    for (size_t i = 0; i < n_rmsd; ++i)
        if (rmsd_sel[i] >= n_atoms) return ORC_ERR_INVALID;
    for (size_t f = 0; f < n_frames; ++f) {
        const float *frame = coords + f * n_atoms * stride;
        double r[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
        double cx[3] = {0, 0, 0}, cy[3] = {0, 0, 0};
        size_t n = n_fit;
        if (n > 0) {
            for (size_t i = 0; i < n; ++i) {
                const float *p = frame + (size_t)fit_sel[i] * stride;
                const float *q = ref_all + (size_t)fit_sel[i] * stride;
                cx[0] += (double)p[0]; cx[1] += (double)p[1]; cx[2] += (double)p[2];
                cy[0] += (double)q[0]; cy[1] += (double)q[1]; cy[2] += (double)q[2];
            }
            for (int d = 0; d < 3; ++d) { cx[d] /= (double)n; cy[d] /= (double)n; }
            double h[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
            for (size_t i = 0; i < n; ++i) {
                const float *p = frame + (size_t)fit_sel[i] * stride;
                const float *q = ref_all + (size_t)fit_sel[i] * stride;
                double xr[3] = {(double)p[0] - cx[0], (double)p[1] - cx[1], (double)p[2] - cx[2]};
                double yr[3] = {(double)q[0] - cy[0], (double)q[1] - cy[1], (double)q[2] - cy[2]};
                for (int a = 0; a < 3; ++a)
                    for (int b = 0; b < 3; ++b) h[a][b] += xr[a] * yr[b];
            }
            double rr[3][3];
            if (kabsch_from_h(h, rr)) memcpy(r, rr, sizeof(r));
        }
        double sum = 0.0;
        for (size_t i = 0; i < n_rmsd; ++i) {
            const float *p = frame + (size_t)rmsd_sel[i] * stride;
            const float *q = ref_all + (size_t)rmsd_sel[i] * stride;
            double pv[3] = {(double)p[0] - cx[0], (double)p[1] - cx[1], (double)p[2] - cx[2]};
            double rv[3] = {(double)q[0] - cy[0], (double)q[1] - cy[1], (double)q[2] - cy[2]};
            double al[3];
            for (int a = 0; a < 3; ++a) al[a] = r[a][0] * pv[0] + r[a][1] * pv[1] + r[a][2] * pv[2];
            double dx = al[0] - rv[0], dy = al[1] - rv[1], dz = al[2] - rv[2];
            sum += dx * dx + dy * dy + dz * dz;
        }
        double nn = (double)(n_rmsd > 1 ? n_rmsd : 1);
        double v = sqrt(sum / nn);
        out[f] = (float)v;
        if (out64) out64[f] = v;
        if (rot_out)
            for (int a = 0; a < 3; ++a)
                for (int b = 0; b < 3; ++b) rot_out[9 * f + 3 * a + b] = r[a][b];
        if (cx_out) for (int d = 0; d < 3; ++d) cx_out[3 * f + d] = cx[d];
        if (cy_out) for (int d = 0; d < 3; ++d) cy_out[3 * f + d] = cy[d];
    }
    return ORC_OK;
}

/* rmsd/plan.rs:1112-1159 rmsd_aligned(frame, reference): f64 centroids
 * (math.rs:40-46, sum / n), H = sum (x - cx)(y - cy)^T, Kabsch R = V U^T with
 * the reflection fix, residual sum |R (x - cx) - (y - cy)|^2, sqrt(sum / n). */
static double rmsd_aligned_pair(const float *x, const float *y, size_t n, size_t stride) {
    if (n == 0) return 0.0;
    double cx[3] = {0, 0, 0}, cy[3] = {0, 0, 0};
    for (size_t i = 0; i < n; ++i)
        for (int d = 0; d < 3; ++d) {
            cx[d] += (double)x[i * stride + d];
            cy[d] += (double)y[i * stride + d];
        }
    for (int d = 0; d < 3; ++d) {
        cx[d] = cx[d] / (double)n;
        cy[d] = cy[d] / (double)n;
    }
    double h[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
    for (size_t i = 0; i < n; ++i) {
        double xr[3], yr[3];
        for (int d = 0; d < 3; ++d) {
            xr[d] = (double)x[i * stride + d] - cx[d];
            yr[d] = (double)y[i * stride + d] - cy[d];
        }
        for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b) h[a][b] += xr[a] * yr[b];
    }
    double r[3][3];
    if (!kabsch_from_h(h, r)) {
        /* rmsd_raw fallback (plan.rs:1081-1091, 1137-1140) */
        double sum = 0.0;
        for (size_t i = 0; i < n; ++i)
            for (int d = 0; d < 3; ++d) {
                double e = (double)x[i * stride + d] - (double)y[i * stride + d];
                sum += e * e;
            }
        return sqrt(sum / (double)n);
    }
    double sum = 0.0;
    for (size_t i = 0; i < n; ++i) {
        double xr[3], yr[3];
        for (int d = 0; d < 3; ++d) {
            xr[d] = (double)x[i * stride + d] - cx[d];
            yr[d] = (double)y[i * stride + d] - cy[d];
        }
        double dd = 0.0;
        for (int a = 0; a < 3; ++a) {
            double al = r[a][0] * xr[0] + r[a][1] * xr[1] + r[a][2] * xr[2];
            double e = al - yr[a];
            dd += e * e;
        }
        sum += dd;
    }
    return sqrt(sum / (double)n);
}

/*
 * PairwiseRmsdPlan::finalize (rmsd/plan.rs:586-675) for metric 0 = rms
 * (rmsd_aligned) and 1 = nofit (rmsd_raw, :1081-1091, f64 sum, sqrt(sum / n)).
 * coords: [n_frames][n_atoms][stride]; the plan keeps the selected atoms of
 * every frame (:533-558).  out [F][F] f32 (symmetric, zero diagonal; all zero
 * when n_sel < 2), out64 NULL or the f64 values before the cast.
 */
int orc_pairwise_rmsd(const float *coords, size_t stride, size_t n_atoms, size_t n_frames,
                      const uint32_t *sel, size_t n_sel, int metric, float *out, double *out64) {
    if (metric != 0 && metric != 1) return ORC_ERR_INVALID;
    for (size_t i = 0; i < n_sel; ++i)
        if (sel[i] >= n_atoms) return ORC_ERR_INVALID;
    for (size_t t = 0; t < n_frames * n_frames; ++t) {
        out[t] = 0.0f;
        if (out64) out64[t] = 0.0;
    }
    if (n_sel < 2) return ORC_OK;
    float *fs = (float *)malloc(n_frames * n_sel * 3 * sizeof(float));
    if (!fs) return ORC_ERR_INVALID;
    for (size_t f = 0; f < n_frames; ++f)
        for (size_t i = 0; i < n_sel; ++i)
            for (int d = 0; d < 3; ++d)
                fs[(f * n_sel + i) * 3 + d] = coords[(f * n_atoms + sel[i]) * stride + d];
    for (size_t i = 0; i < n_frames; ++i)
        for (size_t j = i + 1; j < n_frames; ++j) {
            const float *x = fs + i * n_sel * 3, *y = fs + j * n_sel * 3;
            double v;
            if (metric == 0) {
                v = rmsd_aligned_pair(x, y, n_sel, 3);
            } else {
                double sum = 0.0;
                for (size_t k = 0; k < n_sel; ++k) {
                    double dx = (double)x[3 * k] - (double)y[3 * k];
                    double dy = (double)x[3 * k + 1] - (double)y[3 * k + 1];
                    double dz = (double)x[3 * k + 2] - (double)y[3 * k + 2];
                    sum += dx * dx + dy * dy + dz * dz;
                }
                v = sqrt(sum / (double)n_sel);
            }
            out[i * n_frames + j] = out[j * n_frames + i] = (float)v;
            if (out64) out64[i * n_frames + j] = out64[j * n_frames + i] = v;
        }
    free(fs);
    return ORC_OK;
}

/*
 * PairwiseRmsdPlan metric "dme" (rmsd/plan.rs:612-650): per frame the
 * compact pair-distance vector of pair_distances_compact (:1279-1302) -- for
 * i < j in selection order, the f32 difference cast to f64, the f64
 * orthorhombic wrap of apply_pbc (pbc_math.rs:17-27) on the frame's box
 * lengths when pbc is ORTHO, sqrt in f64 -- then for every frame pair i < j
 * sqrt(sum_k (d_i[k] - d_j[k])^2 / n_pairs) summed over k in order, cast to
 * f32.  boxes: [F][3] box lengths (f32, read as f64) for PBC_ORTHO, else
 * unused.  out [F][F] f32 (symmetric, zero diagonal; all zero when n_sel < 2),
 * out64 NULL or the f64 values before the cast.
 */
int orc_pairwise_dme(const float *coords, size_t stride, size_t n_atoms, size_t n_frames,
                     const uint32_t *sel, size_t n_sel, int pbc, const float *boxes, float *out,
                     double *out64) {
    if (pbc != PBC_NONE && pbc != PBC_ORTHO) return ORC_ERR_INVALID;
    if (pbc == PBC_ORTHO && !boxes) return ORC_ERR_MISMATCH;
    for (size_t i = 0; i < n_sel; ++i)
        if (sel[i] >= n_atoms) return ORC_ERR_INVALID;
    for (size_t t = 0; t < n_frames * n_frames; ++t) {
        out[t] = 0.0f;
        if (out64) out64[t] = 0.0;
    }
    if (n_sel < 2 || n_frames == 0) return ORC_OK;
    const size_t n_pairs = n_sel * (n_sel - 1) / 2;
    double *dist = (double *)malloc(n_frames * n_pairs * sizeof(double));
    if (!dist) return ORC_ERR_INVALID;
    for (size_t f = 0; f < n_frames; ++f) {
        double lx = 0.0, ly = 0.0, lz = 0.0;
        if (pbc == PBC_ORTHO) {
            lx = boxes[3 * f];
            ly = boxes[3 * f + 1];
            lz = boxes[3 * f + 2];
        }
        double *d = dist + f * n_pairs;
        size_t k = 0;
        for (size_t i = 0; i < n_sel; ++i) {
            const float *pa = coords + (f * n_atoms + sel[i]) * stride;
            for (size_t j = i + 1; j < n_sel; ++j) {
                const float *pb = coords + (f * n_atoms + sel[j]) * stride;
                double dx = (double)(pb[0] - pa[0]);
                double dy = (double)(pb[1] - pa[1]);
                double dz = (double)(pb[2] - pa[2]);
                if (pbc == PBC_ORTHO) apply_pbc_f64(&dx, &dy, &dz, lx, ly, lz);
                d[k++] = sqrt(dx * dx + dy * dy + dz * dz);
            }
        }
    }
    for (size_t i = 0; i < n_frames; ++i)
        for (size_t j = i + 1; j < n_frames; ++j) {
            const double *a = dist + i * n_pairs, *b = dist + j * n_pairs;
            double sum = 0.0;
            for (size_t k = 0; k < n_pairs; ++k) {
                const double diff = a[k] - b[k];
                sum += diff * diff;
            }
            const double v = sqrt(sum / (double)n_pairs);
            out[i * n_frames + j] = out[j * n_frames + i] = (float)v;
            if (out64) out64[i * n_frames + j] = out64[j * n_frames + i] = v;
        }
    free(dist);
    return ORC_OK;
}

/*
 * AlignPlan / SuperposePlan CPU path (geometry/align/plan_exec.rs:215-243 and
 * 422-456): per frame kabsch_rotation_weighted (geometry_math.rs:235-297;
 * centroid_weighted :202-233, H = sum (x - cx)(y - cy)^T [* w], SVD Kabsch
 * with the reflection fix), or R = I and the two centroids when norotate;
 * t = cy - R cx.  coords [n_frames][n_atoms][stride]; ref [n_ref][stride]
 * atom-indexed; weights NULL or [n_sel].  out_rt [F][12] f32 (R row-major,
 * t), out_rt64 NULL or the same in f64, out_xyz NULL or [F][n_atoms][3] f32 =
 * R x + t in f64 for every atom (the SuperposePlan series).
 */
int orc_align(const float *coords, size_t stride, size_t n_atoms, size_t n_frames,
              const uint32_t *sel, size_t n_sel, const float *weights, const float *ref,
              size_t n_ref, int norotate, float *out_rt, double *out_rt64, float *out_xyz) {
    for (size_t i = 0; i < n_sel; ++i)
        if (sel[i] >= n_atoms || sel[i] >= n_ref) return ORC_ERR_INVALID;
    const size_t n = n_sel;
    for (size_t f = 0; f < n_frames; ++f) {
        const float *frame = coords + f * n_atoms * stride;
        double cx[3] = {0, 0, 0}, cy[3] = {0, 0, 0}, r[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
        if (n > 0) {
            double sx[3] = {0, 0, 0}, sy[3] = {0, 0, 0}, wsum = 0.0;
            for (size_t i = 0; i < n; ++i) {
                const float *p = frame + (size_t)sel[i] * stride;
                const float *q = ref + (size_t)sel[i] * stride;
                if (weights) {
                    double wi = (double)weights[i];
                    for (int d = 0; d < 3; ++d) {
                        sx[d] += (double)p[d] * wi;
                        sy[d] += (double)q[d] * wi;
                    }
                    wsum += wi;
                } else {
                    for (int d = 0; d < 3; ++d) {
                        sx[d] += (double)p[d];
                        sy[d] += (double)q[d];
                    }
                }
            }
            if (!weights) wsum = (double)n;
            if (wsum != 0.0)
                for (int d = 0; d < 3; ++d) {
                    cx[d] = sx[d] / wsum;
                    cy[d] = sy[d] / wsum;
                }
            if (!norotate) {
                double h[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
                for (size_t i = 0; i < n; ++i) {
                    const float *p = frame + (size_t)sel[i] * stride;
                    const float *q = ref + (size_t)sel[i] * stride;
                    double xr[3], yr[3];
                    for (int d = 0; d < 3; ++d) {
                        xr[d] = (double)p[d] - cx[d];
                        yr[d] = (double)q[d] - cy[d];
                    }
                    double wi = weights ? (double)weights[i] : 1.0;
                    for (int a = 0; a < 3; ++a)
                        for (int b = 0; b < 3; ++b)
                            h[a][b] += weights ? (xr[a] * yr[b]) * wi : xr[a] * yr[b];
                }
                double rr[3][3];
                if (kabsch_from_h(h, rr)) memcpy(r, rr, sizeof(r));
            }
        }
        double t[3];
        for (int a = 0; a < 3; ++a) t[a] = cy[a] - (r[a][0] * cx[0] + r[a][1] * cx[1] + r[a][2] * cx[2]);
        for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b) {
                out_rt[f * 12 + 3 * a + b] = (float)r[a][b];
                if (out_rt64) out_rt64[f * 12 + 3 * a + b] = r[a][b];
            }
        for (int a = 0; a < 3; ++a) {
            out_rt[f * 12 + 9 + a] = (float)t[a];
            if (out_rt64) out_rt64[f * 12 + 9 + a] = t[a];
        }
        if (out_xyz)
            for (size_t k = 0; k < n_atoms; ++k) {
                const float *p = frame + k * stride;
                double x = (double)p[0], y = (double)p[1], z = (double)p[2];
                for (int a = 0; a < 3; ++a)
                    out_xyz[(f * n_atoms + k) * 3 + a] =
                        (float)(r[a][0] * x + r[a][1] * y + r[a][2] * z + t[a]);
            }
    }
    return ORC_OK;
}

/* group_position (rdf.rs:353-366) over center_of_selection
 * (traj-core/src/centers.rs:3-29): f32 position of one group in one frame. */
static void group_position(const float *frame, size_t stride, const uint32_t *atoms, size_t n,
                           const float *masses, float out[3]) {
    out[0] = out[1] = out[2] = 0.0f;
    if (n == 0) return;
    if (n == 1) {
        const float *p = frame + (size_t)atoms[0] * stride;
        out[0] = p[0]; out[1] = p[1]; out[2] = p[2];
        return;
    }
    double s0 = 0.0, s1 = 0.0, s2 = 0.0, ms = 0.0;
    for (size_t i = 0; i < n; ++i) {
        const float *p = frame + (size_t)atoms[i] * stride;
        double m = masses ? (double)masses[atoms[i]] : 1.0;
        s0 += (double)p[0] * m;
        s1 += (double)p[1] * m;
        s2 += (double)p[2] * m;
        ms += m;
    }
    if (ms == 0.0) return;
    out[0] = (float)(s0 / ms);
    out[1] = (float)(s1 / ms);
    out[2] = (float)(s2 / ms);
}

/*
 * RdfPlan grouped modes and the xy dimension (rdf.rs:808-850, rdf_distance
 * :473-478, same_intramolecular_group :368-373).  Groups are CSR lists of
 * atom indices (off [n + 1], atoms); mol / res: per group key or INT64_MIN
 * (none).  masses: per atom (mass-weighted centres) or NULL.  Pairs of groups
 * with identical atom lists are skipped, and with no_intramol the
 * intramolecular ones; the distance, wrap and binning are the atom loop's.
 */
int orc_rdf_grouped(const float *coords, size_t stride, size_t n_atoms, size_t n_frames,
                    const uint64_t *off_a, const uint32_t *atoms_a, size_t n_ga,
                    const int64_t *mol_a, const int64_t *res_a, const uint64_t *off_b,
                    const uint32_t *atoms_b, size_t n_gb, const int64_t *mol_b,
                    const int64_t *res_b, const float *masses, int no_intramol, int pbc,
                    const float *boxes, int planar, size_t bins, float r_max, uint64_t *counts) {
    const int64_t none = INT64_MIN;
    if (pbc == PBC_TRICLINIC) return ORC_ERR_MISMATCH;
    for (size_t i = 0; i < off_a[n_ga]; ++i)
        if (atoms_a[i] >= n_atoms) return ORC_ERR_INVALID;
    for (size_t i = 0; i < off_b[n_gb]; ++i)
        if (atoms_b[i] >= n_atoms) return ORC_ERR_INVALID;
    const float bin_width = r_max / (float)bins;
    float *pa = (float *)malloc((n_ga + 1) * 3 * sizeof(float));
    float *pb = (float *)malloc((n_gb + 1) * 3 * sizeof(float));
    if (!pa || !pb) { free(pa); free(pb); return ORC_ERR_INVALID; }
    for (size_t f = 0; f < n_frames; ++f) {
        const float *frame = coords + f * n_atoms * stride;
        float lx = 0.f, ly = 0.f, lz = 0.f;
        if (pbc == PBC_ORTHO) {
            const float *m = boxes + 9 * f;
            lx = m[0]; ly = m[4]; lz = m[8];
        }
        for (size_t g = 0; g < n_ga; ++g)
            group_position(frame, stride, atoms_a + off_a[g], off_a[g + 1] - off_a[g], masses, pa + 3 * g);
        for (size_t g = 0; g < n_gb; ++g)
            group_position(frame, stride, atoms_b + off_b[g], off_b[g + 1] - off_b[g], masses, pb + 3 * g);
        for (size_t a = 0; a < n_ga; ++a) {
            for (size_t b = 0; b < n_gb; ++b) {
                size_t la = off_a[a + 1] - off_a[a], lb = off_b[b + 1] - off_b[b];
                if (la == lb && memcmp(atoms_a + off_a[a], atoms_b + off_b[b], la * sizeof(uint32_t)) == 0)
                    continue;
                if (no_intramol) {
                    int same;
                    if (mol_a[a] != none && mol_b[b] != none) same = mol_a[a] == mol_b[b];
                    else same = res_a[a] != none && res_b[b] != none && res_a[a] == res_b[b];
                    if (same) continue;
                }
                float dx = pb[3 * b] - pa[3 * a];
                float dy = pb[3 * b + 1] - pa[3 * a + 1];
                float dz = pb[3 * b + 2] - pa[3 * a + 2];
                if (pbc == PBC_ORTHO) {
                    dx -= roundf(dx / lx) * lx;
                    dy -= roundf(dy / ly) * ly;
                    dz -= roundf(dz / lz) * lz;
                }
                float r = planar ? sqrtf(dx * dx + dy * dy) : sqrtf(dx * dx + dy * dy + dz * dz);
                if (r < r_max) {
                    size_t bin = (size_t)(r / bin_width);
                    if (bin < bins) counts[bin] += 1;
                }
            }
        }
    }
    free(pa);
    free(pb);
    return ORC_OK;
}

/*
 * orc_rdf_brute with the sel_a loop split across OpenMP threads and one private
 * histogram per thread, summed at the end.  Integer counts make the result
 * identical to orc_rdf_brute; this variant exists only to time an all-cores
 * CPU baseline.  n_threads <= 0 uses the OpenMP default.
 */
int orc_rdf_brute_omp(const float *coords, size_t stride, size_t n_atoms, size_t n_frames,
                      const uint32_t *sel_a, size_t n_a, const uint32_t *sel_b, size_t n_b,
                      size_t bins, float r_max, int pbc, const float *boxes, uint64_t *counts,
                      int n_threads) {
    if (bins == 0) return ORC_OK;
    for (size_t i = 0; i < n_a; ++i)
        if (sel_a[i] >= n_atoms) return ORC_ERR_INVALID;
    for (size_t i = 0; i < n_b; ++i)
        if (sel_b[i] >= n_atoms) return ORC_ERR_INVALID;
    float bin_width = r_max / (float)bins;
    int status = ORC_OK;
    for (size_t f = 0; f < n_frames && status == ORC_OK; ++f) {
        frame_box fb;
        if (load_box(pbc, boxes ? boxes + 9 * f : NULL, &fb) != ORC_OK) return ORC_ERR_MISMATCH;
        const float *base = coords + f * n_atoms * stride;
#ifdef _OPENMP
#pragma omp parallel num_threads(n_threads > 0 ? n_threads : omp_get_max_threads())
#endif
        {
            uint64_t *local = (uint64_t *)calloc(bins, sizeof(uint64_t));
#ifdef _OPENMP
#pragma omp for schedule(dynamic, 16)
#endif
            for (long ia = 0; ia < (long)n_a; ++ia) {
                uint32_t a = sel_a[ia];
                const float *pa = base + (size_t)a * stride;
                for (size_t ib = 0; ib < n_b; ++ib) {
                    uint32_t b = sel_b[ib];
                    if (a == b) continue;
                    float r = pair_r32(pa, base + (size_t)b * stride, &fb);
                    long bin = bin_of(r, r_max, bin_width, bins);
                    if (bin >= 0) local[bin] += 1;
                }
            }
#ifdef _OPENMP
#pragma omp critical
#endif
            for (size_t k = 0; k < bins; ++k) counts[k] += local[k];
            free(local);
        }
    }
    return status;
}

int orc_omp_max_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/* ------------------------------------------------------------------------- */
/* PairDistPlan (rdf.rs:1120-1178, CPU loop of process_chunk)                 */
/* ------------------------------------------------------------------------- */
/* Ordered pairs (a_pos, b_pos) of sel_a x sel_b with the reference's skip
 * rules: unique_pairs keeps b_pos > a_pos only; otherwise a == b is skipped
 * when same_sel (Arc::ptr_eq).  Same f32 distance and binning as RdfPlan
 * (orthorhombic or none; other boxes are a Mismatch, rdf.rs:1125-1134).
 * counts: [bins] accumulated; frame_counts: NULL or [n_frames][bins] (the
 * per-frame histograms of output_distribution, rdf.rs:1121-1125). */
int orc_pairdist(const float *coords, size_t stride, size_t n_atoms, size_t n_frames,
                 const uint32_t *sel_a, size_t n_a, const uint32_t *sel_b, size_t n_b,
                 size_t bins, float r_max, int pbc, const float *boxes, int same_sel,
                 int unique_pairs, uint64_t *counts, uint64_t *frame_counts) {
    if (pbc == PBC_TRICLINIC) return ORC_ERR_MISMATCH;
    for (size_t i = 0; i < n_a; ++i)
        if (sel_a[i] >= n_atoms) return ORC_ERR_INVALID;
    for (size_t i = 0; i < n_b; ++i)
        if (sel_b[i] >= n_atoms) return ORC_ERR_INVALID;
    float bin_width = r_max / (float)bins;
    for (size_t f = 0; f < n_frames; ++f) {
        frame_box fb;
        if (load_box(pbc, boxes ? boxes + 9 * f : NULL, &fb) != ORC_OK) return ORC_ERR_MISMATCH;
        const float *base = coords + f * n_atoms * stride;
        uint64_t *fc = frame_counts ? frame_counts + f * bins : NULL;
        for (size_t ia = 0; ia < n_a; ++ia) {
            uint32_t a = sel_a[ia];
            for (size_t ib = 0; ib < n_b; ++ib) {
                uint32_t b = sel_b[ib];
                if (unique_pairs && ib <= ia) continue;
                if (!unique_pairs && a == b && same_sel) continue;
                float r = pair_r32(base + (size_t)a * stride, base + (size_t)b * stride, &fb);
                long bin = bins ? bin_of(r, r_max, bin_width, bins) : -1;
                if (bin >= 0) {
                    counts[bin] += 1;
                    if (fc) fc[bin] += 1;
                }
            }
        }
    }
    return ORC_OK;
}

/* ------------------------------------------------------------------------- */
/* StructureFactorPlan (analysis/structure_factor.rs:160-258)                 */
/* ------------------------------------------------------------------------- */
/* Counts of the CPU path: ordered pairs of one selection skipping a == b,
 * with the displacement (f32 difference) cast to f64 and scaled by
 * length_scale, the orthorhombic wrap in f64 with the scaled box, r = sqrt in
 * f64, r < r_max and bin = (r / (r_max / bins)) as usize in f64.  volume_sum
 * gets the scaled f64 box volume of every frame (orthorhombic). */
int orc_sf_counts(const float *coords, size_t stride, size_t n_atoms, size_t n_frames,
                  const uint32_t *sel, size_t n_sel, size_t bins, double r_max, int pbc,
                  const float *boxes, double length_scale, uint64_t *counts, double *volume_sum) {
    if (pbc == PBC_TRICLINIC) return ORC_ERR_MISMATCH;
    for (size_t i = 0; i < n_sel; ++i)
        if (sel[i] >= n_atoms) return ORC_ERR_INVALID;
    const double w = r_max / (double)bins;
    for (size_t f = 0; f < n_frames; ++f) {
        double lx = 0.0, ly = 0.0, lz = 0.0;
        if (pbc == PBC_ORTHO) {
            const float *m = boxes + 9 * f;
            lx = (double)m[0];
            ly = (double)m[4];
            lz = (double)m[8];
            /* (lx * ly * lz) * scale^3, powi(3) = (s * s) * s */
            *volume_sum += (lx * ly * lz) * (length_scale * length_scale * length_scale);
        }
        const double sx = lx * length_scale, sy = ly * length_scale, sz = lz * length_scale;
        const float *base = coords + f * n_atoms * stride;
        for (size_t ia = 0; ia < n_sel; ++ia) {
            const float *pa = base + (size_t)sel[ia] * stride;
            for (size_t ib = 0; ib < n_sel; ++ib) {
                if (sel[ia] == sel[ib]) continue;
                const float *pb = base + (size_t)sel[ib] * stride;
                double dx = (double)(pb[0] - pa[0]) * length_scale;
                double dy = (double)(pb[1] - pa[1]) * length_scale;
                double dz = (double)(pb[2] - pa[2]) * length_scale;
                if (pbc == PBC_ORTHO) {
                    dx -= round(dx / sx) * sx;
                    dy -= round(dy / sy) * sy;
                    dz -= round(dz / sz) * sz;
                }
                double r = sqrt(dx * dx + dy * dy + dz * dz);
                if (r < r_max) {
                    size_t bin = (size_t)(r / w);
                    if (bin < bins) counts[bin] += 1;
                }
            }
        }
    }
    return ORC_OK;
}

/* StructureFactorPlan::finalize (structure_factor.rs:200-258): g(r) at bin
 * centres (i + 1/2) w with shell 4 pi r0^2 w and rho = n / mean volume (1 without
 * PBC), then S(q) = 1 + 4 pi rho sum (g - 1) sinc(q r0) r0^2 w on q = dq i,
 * dq = q_max / (q_bins - 1).  Returns ORC_ERR_MISMATCH for PBC with no volume. */
int orc_sf_finalize(const uint64_t *counts, size_t bins, double r_max, double n, size_t frames,
                    double volume_sum, int pbc, size_t q_bins, double q_max, float *r, float *g_r,
                    float *q, float *s_q) {
    const double pi = 3.14159265358979323846;
    const double w = r_max / (double)bins;
    const double fr = (double)(frames > 0 ? frames : 1);
    double volume = 1.0;
    if (pbc == PBC_ORTHO) {
        if (volume_sum == 0.0) return ORC_ERR_MISMATCH;
        volume = volume_sum / fr;
    }
    const double rho = volume > 0.0 ? n / volume : 0.0;
    for (size_t i = 0; i < bins; ++i) {
        const double r0 = ((double)i + 0.5) * w;
        const double shell = 4.0 * pi * r0 * r0 * w;
        const double norm = shell > 0.0 ? rho * shell * n : 1.0;
        const double g = (double)counts[i] / (norm * fr);
        r[i] = (float)r0;
        g_r[i] = (float)g;
    }
    const double dq = q_bins > 1 ? q_max / ((double)q_bins - 1.0) : 0.0;
    for (size_t k = 0; k < q_bins; ++k) {
        const double qv = dq * (double)k;
        double sum = 0.0;
        for (size_t i = 0; i < bins; ++i) {
            const double r0 = ((double)i + 0.5) * w;
            const double diff = (double)g_r[i] - 1.0;
            const double weight = qv == 0.0 ? 1.0 : sin(qv * r0) / (qv * r0);
            sum += diff * weight * r0 * r0 * w;
        }
        q[k] = (float)qv;
        s_q[k] = (float)(1.0 + 4.0 * pi * rho * sum);
    }
    return ORC_OK;
}
