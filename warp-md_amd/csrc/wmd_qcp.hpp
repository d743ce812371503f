// QCP superposition in FP64, shared by the device finalize kernel and the
// host test entry point (wmd_qcp_host).
//
// Method: Theobald, Acta Cryst. A61 (2005) 478 -- the largest eigenvalue of
// Horn's 4x4 key matrix K(M) is found by Newton iteration on its characteristic
// polynomial P(l) = l^4 + c2 l^2 + c1 l + c0 started from the upper bound
// E0 = (Gx + Gy) / 2; the rotation follows from the null vector of K - l I
// taken as a column of its adjugate (Liu, Agrafiotis & Theobald, J. Comput.
// Chem. 31 (2010) 1561).  When that column vanishes (degenerate top
// eigenvalue) a cyclic Jacobi eigen-decomposition of K supplies the vector.
//
// Conventions: M[a*3+b] = sum_i (x_i - cx)_a (y_i - cy)_b, x = frame (moved),
// y = reference (fixed).  The returned rotation R (row-major) maximises
// sum_i (y_i - cy) . R (x_i - cx), i.e. R (x - cx) ~ (y - cy), the same R the
// reference's Kabsch builds as V U^T (rmsd/plan.rs:1214-1224).
#pragma once

#include <cmath>

#ifdef __HIPCC__
#define WMD_HD __host__ __device__
#else
#define WMD_HD
#endif

namespace wmd {

struct QcpResult {
    double lambda;  // largest eigenvalue of K
    double e0;      // (gx + gy) / 2
    double q[4];    // unit quaternion (w, x, y, z)
    double rot[9];  // row-major rotation
    int iters;      // Newton iterations (-1: Jacobi fallback supplied q)
};

WMD_HD inline double det3v(double a, double b, double c, double d, double e, double f, double g,
                           double h, double i) {
    return a * (e * i - f * h) - b * (d * i - f * g) + c * (d * h - e * g);
}

// Cofactor C_rc of a symmetric 4x4 matrix A (row-major 16).
WMD_HD inline double cof4(const double *A, int r, int c) {
    double m[9];
    int k = 0;
    for (int i = 0; i < 4; ++i) {
        if (i == r) continue;
        for (int j = 0; j < 4; ++j) {
            if (j == c) continue;
            m[k++] = A[i * 4 + j];
        }
    }
    double d = det3v(m[0], m[1], m[2], m[3], m[4], m[5], m[6], m[7], m[8]);
    return ((r + c) & 1) ? -d : d;
}

WMD_HD inline double det4(const double *A) {
    // expansion along row 0
    return A[0] * cof4(A, 0, 0) + A[1] * cof4(A, 0, 1) + A[2] * cof4(A, 0, 2) +
           A[3] * cof4(A, 0, 3);
}

WMD_HD inline void key_matrix(const double *M, double *K) {
    const double Sxx = M[0], Sxy = M[1], Sxz = M[2];
    const double Syx = M[3], Syy = M[4], Syz = M[5];
    const double Szx = M[6], Szy = M[7], Szz = M[8];
    K[0] = Sxx + Syy + Szz;  K[1] = Syz - Szy;        K[2] = Szx - Sxz;         K[3] = Sxy - Syx;
    K[4] = K[1];             K[5] = Sxx - Syy - Szz;  K[6] = Sxy + Syx;         K[7] = Szx + Sxz;
    K[8] = K[2];             K[9] = K[6];             K[10] = -Sxx + Syy - Szz; K[11] = Syz + Szy;
    K[12] = K[3];            K[13] = K[7];            K[14] = K[11];            K[15] = -Sxx - Syy + Szz;
}

WMD_HD inline void quat_to_rot(const double *q, double *R) {
    const double w = q[0], x = q[1], y = q[2], z = q[3];
    const double ww = w * w, xx = x * x, yy = y * y, zz = z * z;
    R[0] = ww + xx - yy - zz;  R[1] = 2.0 * (x * y - w * z);  R[2] = 2.0 * (x * z + w * y);
    R[3] = 2.0 * (x * y + w * z);  R[4] = ww - xx + yy - zz;  R[5] = 2.0 * (y * z - w * x);
    R[6] = 2.0 * (x * z - w * y);  R[7] = 2.0 * (y * z + w * x);  R[8] = ww - xx - yy + zz;
}

// Cyclic Jacobi eigen-decomposition of a symmetric 4x4; writes the unit
// eigenvector of the largest eigenvalue into q.
WMD_HD inline void jacobi4_top(const double *Kin, double *q) {
    double A[16], V[16];
    for (int i = 0; i < 16; ++i) {
        A[i] = Kin[i];
        V[i] = (i % 5 == 0) ? 1.0 : 0.0;
    }
    for (int sweep = 0; sweep < 50; ++sweep) {
        double off = 0.0, diag = 0.0;
        for (int i = 0; i < 4; ++i) {
            diag += A[i * 5] * A[i * 5];
            for (int j = i + 1; j < 4; ++j) off += A[i * 4 + j] * A[i * 4 + j];
        }
        if (off <= 1e-32 * diag || off == 0.0) break;
        for (int p = 0; p < 3; ++p)
            for (int r = p + 1; r < 4; ++r) {
                double apr = A[p * 4 + r];
                if (apr == 0.0) continue;
                double theta = (A[r * 5] - A[p * 5]) / (2.0 * apr);
                double t = (theta >= 0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
                double c = 1.0 / sqrt(t * t + 1.0), s = t * c;
                for (int k = 0; k < 4; ++k) {
                    double akp = A[k * 4 + p], akr = A[k * 4 + r];
                    A[k * 4 + p] = c * akp - s * akr;
                    A[k * 4 + r] = s * akp + c * akr;
                }
                for (int k = 0; k < 4; ++k) {
                    double apk = A[p * 4 + k], ark = A[r * 4 + k];
                    A[p * 4 + k] = c * apk - s * ark;
                    A[r * 4 + k] = s * apk + c * ark;
                }
                for (int k = 0; k < 4; ++k) {
                    double vkp = V[k * 4 + p], vkr = V[k * 4 + r];
                    V[k * 4 + p] = c * vkp - s * vkr;
                    V[k * 4 + r] = s * vkp + c * vkr;
                }
            }
    }
    int best = 0;
    for (int i = 1; i < 4; ++i)
        if (A[i * 5] > A[best * 5]) best = i;
    double nrm = 0.0;
    for (int k = 0; k < 4; ++k) nrm += V[k * 4 + best] * V[k * 4 + best];
    nrm = sqrt(nrm);
    for (int k = 0; k < 4; ++k) q[k] = V[k * 4 + best] / nrm;
}

WMD_HD inline void qcp_solve(const double *M, double gx, double gy, bool want_rot, QcpResult *out) {
    double K[16];
    key_matrix(M, K);
    const double e0 = 0.5 * (gx + gy);
    out->e0 = e0;
    double ss = 0.0;
    for (int i = 0; i < 9; ++i) ss += M[i] * M[i];
    const double c2 = -2.0 * ss;
    const double c1 = -8.0 * det3v(M[0], M[1], M[2], M[3], M[4], M[5], M[6], M[7], M[8]);
    const double c0 = det4(K);
    double lam = e0;
    int it = 0;
    for (; it < 60; ++it) {
        const double l2 = lam * lam;
        const double p = ((l2 + c2) * lam + c1) * lam + c0;
        const double dp = (4.0 * l2 + 2.0 * c2) * lam + c1;
        if (dp == 0.0) break;
        const double next = lam - p / dp;
        const double step = fabs(next - lam);
        lam = next;
        if (!(step > 1e-15 * fabs(lam))) {
            ++it;
            break;
        }
    }
    out->lambda = lam;
    out->iters = it;
    if (!want_rot) return;
    double A[16];
    for (int i = 0; i < 16; ++i) A[i] = K[i];
    for (int i = 0; i < 4; ++i) A[i * 5] -= lam;
    // adjugate column with the largest diagonal cofactor
    int best = 0;
    double bestv = -1.0;
    for (int j = 0; j < 4; ++j) {
        double v = fabs(cof4(A, j, j));
        if (v > bestv) {
            bestv = v;
            best = j;
        }
    }
    double q[4];
    double nrm = 0.0;
    for (int i = 0; i < 4; ++i) {
        q[i] = cof4(A, i, best);
        nrm += q[i] * q[i];
    }
    nrm = sqrt(nrm);
    // scale of the adjugate entries ~ |K|^3; a vanishing column means the top
    // eigenvalue is (numerically) degenerate
    double scale = fabs(lam) + fabs(e0) + 1e-300;
    if (!(nrm > 1e-11 * scale * scale * scale)) {
        jacobi4_top(K, q);
        out->iters = -1;
    } else {
        for (int i = 0; i < 4; ++i) q[i] /= nrm;
    }
    for (int i = 0; i < 4; ++i) out->q[i] = q[i];
    quat_to_rot(q, out->rot);
}

}  // namespace wmd
