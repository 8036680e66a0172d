// jam_super.h -- FP64 least-squares superposition on the device.
//
// One call superposes n point pairs: molecule points a_i onto template points
// b_i (Superposition_align/rmsd/rotation/centroid, include/jess/super.pxd:1-14).
// The RMSD rule follows patches/Super.c.patch:40-64: eigenvalues e of X^T X
// (X = sum_i (b_i - cb)(a_i - ca)^T) clamped at zero, sumE = sum sqrt(e),
// minus twice the smallest root when det(X) < 1e-8 (reflection), and
// rmsd = sqrt(max(SA + SB - 2 sumE, 0) / n). The eigen solver is a cyclic
// Jacobi (Golub & Van Loan, Alg. 8.5.2-8.5.3); the rotation is
// M = X P diag(sigma_i / sqrt(e_i)) P^T, singular (NaN/inf) when X^T X has
// rank < 3, the case pyjess skips with a warning.
//
// Every operation is plain IEEE double (the library is built with
// -ffp-contract=off): sums are accumulated in point order, no fused
// multiply-add, correctly rounded division and square root. The point
// loaders are template parameters so the search pipeline can gather straight
// from HBM without staging.
#pragma once

#include <hip/hip_runtime.h>
#include <math.h>

namespace jam {

__device__ static inline void jacobi3(double A[3][3], double P[3][3], double e[3])
{
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) P[i][j] = (i == j) ? 1.0 : 0.0;
    for (int sweep = 0; sweep < 32; sweep++) {
        bool rotated = false;
#pragma unroll
        for (int r = 0; r < 3; r++) {
            const int p = (r == 2) ? 1 : 0;
            const int q = (r == 0) ? 1 : 2;
            double apq = A[p][q];
            if (apq == 0.0) continue;
            double scale = fabs(A[p][p]) + fabs(A[q][q]);
            if (fabs(apq) <= 1e-12 * scale) {
                A[p][q] = 0.0;
                A[q][p] = 0.0;
                continue;
            }
            double theta = (A[q][q] - A[p][p]) / (2.0 * apq);
            double t = 1.0 / (fabs(theta) + sqrt(theta * theta + 1.0));
            if (theta < 0.0) t = -t;
            double c = 1.0 / sqrt(t * t + 1.0);
            double s = t * c;
#pragma unroll
            for (int k = 0; k < 3; k++) {
                double akp = A[k][p], akq = A[k][q];
                A[k][p] = c * akp - s * akq;
                A[k][q] = s * akp + c * akq;
            }
#pragma unroll
            for (int k = 0; k < 3; k++) {
                double apk = A[p][k], aqk = A[q][k];
                A[p][k] = c * apk - s * aqk;
                A[q][k] = s * apk + c * aqk;
            }
            A[p][q] = 0.0;
            A[q][p] = 0.0;
#pragma unroll
            for (int k = 0; k < 3; k++) {
                double pkp = P[k][p], pkq = P[k][q];
                P[k][p] = c * pkp - s * pkq;
                P[k][q] = s * pkp + c * pkq;
            }
            rotated = true;
        }
        if (!rotated) break;
    }
#pragma unroll
    for (int i = 0; i < 3; i++) e[i] = A[i][i];
}

// PA(i, d) / PB(i, d): coordinate d of point i of the molecule / template side.
template <class PA, class PB>
__device__ static inline void superpose(int n, PA pa, PB pb, double *rmsd, double rot[9],
                                        double c0[3], double c1[3])
{
    double ca[3] = {0.0, 0.0, 0.0}, cb[3] = {0.0, 0.0, 0.0};
    for (int i = 0; i < n; i++) {
#pragma unroll
        for (int d = 0; d < 3; d++) {
            ca[d] += pa(i, d);
            cb[d] += pb(i, d);
        }
    }
#pragma unroll
    for (int d = 0; d < 3; d++) {
        ca[d] /= (double)n;
        cb[d] /= (double)n;
    }
    double X[3][3] = {{0.0, 0.0, 0.0}, {0.0, 0.0, 0.0}, {0.0, 0.0, 0.0}};
    double sa = 0.0, sb = 0.0;
    for (int i = 0; i < n; i++) {
        double u[3], v[3];
#pragma unroll
        for (int d = 0; d < 3; d++) {
            u[d] = pa(i, d) - ca[d];
            v[d] = pb(i, d) - cb[d];
        }
#pragma unroll
        for (int d = 0; d < 3; d++) {
            sa += u[d] * u[d];
            sb += v[d] * v[d];
        }
#pragma unroll
        for (int j = 0; j < 3; j++)
#pragma unroll
            for (int k = 0; k < 3; k++) X[j][k] += v[j] * u[k];
    }
    double det = X[0][0] * (X[1][1] * X[2][2] - X[1][2] * X[2][1])
               - X[0][1] * (X[1][0] * X[2][2] - X[1][2] * X[2][0])
               + X[0][2] * (X[1][0] * X[2][1] - X[1][1] * X[2][0]);
    double XX[3][3];
#pragma unroll
    for (int j = 0; j < 3; j++)
#pragma unroll
        for (int k = 0; k < 3; k++) {
            double s = 0.0;
#pragma unroll
            for (int l = 0; l < 3; l++) s += X[l][j] * X[l][k];
            XX[j][k] = s;
        }
    double P[3][3], e[3];
    jacobi3(XX, P, e);
#pragma unroll
    for (int i = 0; i < 3; i++) e[i] = fmax(e[i], 0.0);
    int imin = 0;
    if (e[1] < e[imin]) imin = 1;
    if (e[2] < e[imin]) imin = 2;
    double sumE = sqrt(e[0]) + sqrt(e[1]) + sqrt(e[2]);
    double sig[3] = {1.0, 1.0, 1.0};
    if (det < 1e-8) {
        sumE -= 2.0 * sqrt(e[imin]);
        sig[imin] = -1.0;
    }
    double r = sa + sb - 2.0 * sumE;
    *rmsd = sqrt(fmax(r, 0.0) / (double)n);
    double M[3][3] = {{0.0, 0.0, 0.0}, {0.0, 0.0, 0.0}, {0.0, 0.0, 0.0}};
#pragma unroll
    for (int i = 0; i < 3; i++) {
        double w = sig[i] / sqrt(e[i]);
        double xp[3];
#pragma unroll
        for (int j = 0; j < 3; j++) xp[j] = X[j][0] * P[0][i] + X[j][1] * P[1][i] + X[j][2] * P[2][i];
#pragma unroll
        for (int j = 0; j < 3; j++)
#pragma unroll
            for (int k = 0; k < 3; k++) M[j][k] += (w * xp[j]) * P[k][i];
    }
#pragma unroll
    for (int j = 0; j < 3; j++)
#pragma unroll
        for (int k = 0; k < 3; k++) rot[3 * j + k] = M[j][k];
#pragma unroll
    for (int d = 0; d < 3; d++) {
        c0[d] = ca[d];
        c1[d] = cb[d];
    }
}

}  // namespace jam
