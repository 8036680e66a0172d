// jam_grad.h -- differentiable superposition RMSD of many point-set pairs.
//
// rmsd = sqrt(E / n), E = sum_i |B_i - R A_i|^2, A and B the centred points
// and R the proper rotation that minimises E (the least-squares
// superposition of a onto b: the quantity of Superposition_rmsd,
// include/jess/super.pxd:1-14). With X = sum_i B_i A_i^T = U S V^T,
// R = U diag(1, 1, det(U) det(V)) V^T.
//
// R is computed from a one-sided Jacobi SVD of X (Hestenes: plane rotations
// of the columns until they are orthogonal; Demmel & Veselic, "Jacobi's
// method is more accurate than QR", 1992), which keeps the small singular
// values -- nearly planar or nearly collinear sets -- to high relative
// accuracy, with u3 = u1 x u2 so that the frame is proper even when
// sigma3 = 0. E is then summed from the residuals directly: no cancellation,
// and only second order in the error of R because R minimises E. (The
// search's hit decisions use the reference's eigenvalue formula instead,
// jam_super.h; the two agree to rounding on well-conditioned sets.)
//
// Gradient: dE/dR vanishes at the optimum (envelope theorem) and centring
// drops out of the chain rule because centred points sum to zero, so
//     d rmsd / d a_i = (A_i - R^T B_i) / (n rmsd),
//     d rmsd / d b_i = (B_i - R A_i) / (n rmsd);
// taken as zero where rmsd = 0.
#pragma once

#include <hip/hip_runtime.h>
#include <math.h>

#include "jam_wave.h"

namespace jam {

// Optimal proper rotation R (row-major 3x3) for the correlation X.
__device__ static inline void optimal_rotation(const double X[3][3], double R[3][3])
{
    double W[3][3], V[3][3];
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) {
            W[i][j] = X[i][j];
            V[i][j] = (i == j) ? 1.0 : 0.0;
        }
    for (int sweep = 0; sweep < 16; sweep++) {
        bool rotated = false;
#pragma unroll
        for (int r = 0; r < 3; r++) {
            const int p = (r == 2) ? 1 : 0;
            const int q = (r == 0) ? 1 : 2;
            double al = 0.0, be = 0.0, ga = 0.0;
#pragma unroll
            for (int j = 0; j < 3; j++) {
                al += W[j][p] * W[j][p];
                be += W[j][q] * W[j][q];
                ga += W[j][p] * W[j][q];
            }
            if (ga == 0.0 || fabs(ga) <= 1e-15 * sqrt(al * be)) continue;
            const double ze = (be - al) / (2.0 * ga);
            double t = 1.0 / (fabs(ze) + sqrt(1.0 + ze * ze));
            if (ze < 0.0) t = -t;
            const double c = 1.0 / sqrt(1.0 + t * t);
            const double s = c * t;
#pragma unroll
            for (int j = 0; j < 3; j++) {
                const double wp = W[j][p], wq = W[j][q];
                W[j][p] = c * wp - s * wq;
                W[j][q] = s * wp + c * wq;
                const double vp = V[j][p], vq = V[j][q];
                V[j][p] = c * vp - s * vq;
                V[j][q] = s * vp + c * vq;
            }
            rotated = true;
        }
        if (!rotated) break;
    }
    // columns of W are U S: order by singular value
    double sg[3];
#pragma unroll
    for (int i = 0; i < 3; i++) sg[i] = sqrt(W[0][i] * W[0][i] + W[1][i] * W[1][i] + W[2][i] * W[2][i]);
    int i1 = 0, i2 = 1, i3 = 2;
    if (sg[i2] > sg[i1]) { int t = i1; i1 = i2; i2 = t; }
    if (sg[i3] > sg[i2]) { int t = i2; i2 = i3; i3 = t; }
    if (sg[i2] > sg[i1]) { int t = i1; i1 = i2; i2 = t; }
    double u1[3], u2[3], u3[3], v1[3], v2[3], v3[3];
#pragma unroll
    for (int j = 0; j < 3; j++) {
        u1[j] = sg[i1] > 0.0 ? W[j][i1] / sg[i1] : (j == 0 ? 1.0 : 0.0);
        u2[j] = W[j][i2];
        v1[j] = V[j][i1];
        v2[j] = V[j][i2];
        v3[j] = V[j][i3];
    }
    if (sg[i2] > 0.0) {
#pragma unroll
        for (int j = 0; j < 3; j++) u2[j] /= sg[i2];
    } else {
        // rank <= 1: any unit vector orthogonal to u1 (the axis of u1's
        // smallest component, made orthogonal)
        int ax = 0;
        if (fabs(u1[1]) < fabs(u1[ax])) ax = 1;
        if (fabs(u1[2]) < fabs(u1[ax])) ax = 2;
        double e[3] = {0.0, 0.0, 0.0};
        e[ax] = 1.0;
        const double dp = u1[0] * e[0] + u1[1] * e[1] + u1[2] * e[2];
        double w[3] = {e[0] - dp * u1[0], e[1] - dp * u1[1], e[2] - dp * u1[2]};
        const double nw = sqrt(w[0] * w[0] + w[1] * w[1] + w[2] * w[2]);
#pragma unroll
        for (int j = 0; j < 3; j++) u2[j] = w[j] / nw;
    }
    u3[0] = u1[1] * u2[2] - u1[2] * u2[1];
    u3[1] = u1[2] * u2[0] - u1[0] * u2[2];
    u3[2] = u1[0] * u2[1] - u1[1] * u2[0];
    const double detv = v1[0] * (v2[1] * v3[2] - v2[2] * v3[1]) - v1[1] * (v2[0] * v3[2] - v2[2] * v3[0]) +
                        v1[2] * (v2[0] * v3[1] - v2[1] * v3[0]);
    const double sgn = detv < 0.0 ? -1.0 : 1.0;
#pragma unroll
    for (int j = 0; j < 3; j++)
#pragma unroll
        for (int k = 0; k < 3; k++) R[j][k] = u1[j] * v1[k] + u2[j] * v2[k] + sgn * u3[j] * v3[k];
}

// Centroids and the correlation X of points [0, n) of A and B.
__device__ static inline void centre_and_correlate(int n, const double *A, const double *B, double ca[3],
                                                   double cb[3], double X[3][3])
{
#pragma unroll
    for (int d = 0; d < 3; d++) {
        ca[d] = 0.0;
        cb[d] = 0.0;
    }
    for (int i = 0; i < n; i++)
#pragma unroll
        for (int d = 0; d < 3; d++) {
            ca[d] += A[3 * i + d];
            cb[d] += B[3 * i + d];
        }
#pragma unroll
    for (int d = 0; d < 3; d++) {
        ca[d] /= (double)n;
        cb[d] /= (double)n;
    }
#pragma unroll
    for (int j = 0; j < 3; j++)
#pragma unroll
        for (int k = 0; k < 3; k++) X[j][k] = 0.0;
    for (int i = 0; i < n; i++) {
        double u[3], v[3];
#pragma unroll
        for (int d = 0; d < 3; d++) {
            u[d] = A[3 * i + d] - ca[d];
            v[d] = B[3 * i + d] - cb[d];
        }
#pragma unroll
        for (int j = 0; j < 3; j++)
#pragma unroll
            for (int k = 0; k < 3; k++) X[j][k] += v[j] * u[k];
    }
}

// One set: RMSD and optimal rotation (rot may be null).
__device__ __attribute__((always_inline)) static inline void rmsd_one(int n, const double *A, const double *B,
                                                                      double *rmsd, double *rot)
{
    double ca[3], cb[3], X[3][3], R[3][3];
    centre_and_correlate(n, A, B, ca, cb, X);
    optimal_rotation(X, R);
    double e = 0.0;
    for (int i = 0; i < n; i++) {
        double u[3], v[3];
#pragma unroll
        for (int d = 0; d < 3; d++) {
            u[d] = A[3 * i + d] - ca[d];
            v[d] = B[3 * i + d] - cb[d];
        }
#pragma unroll
        for (int d = 0; d < 3; d++) {
            const double ra = R[d][0] * u[0] + R[d][1] * u[1] + R[d][2] * u[2];
            const double r = v[d] - ra;
            e += r * r;
        }
    }
    *rmsd = sqrt(e / (double)n);
    if (rot)
#pragma unroll
        for (int j = 0; j < 3; j++)
#pragma unroll
            for (int k = 0; k < 3; k++) rot[3 * j + k] = R[j][k];
}

// One set: the gradient of g * rmsd, written to GA / GB (which may alias A /
// B: every point is read before its gradient is written). rot: the set's
// optimal rotation from the forward pass, or null to compute it again.
__device__ __attribute__((always_inline)) static inline void rmsd_grad_one(int n, const double *A, const double *B,
                                                                           double r, double g, const double *rot,
                                                                           double *GA, double *GB)
{
    if (!(r > 0.0) || g == 0.0) {
        for (int q = 0; q < 3 * n; q++) {
            if (GA) GA[q] = 0.0;
            if (GB) GB[q] = 0.0;
        }
        return;
    }
    double ca[3], cb[3], R[3][3];
    if (rot) {
        // centroids only: the rotation is the forward pass's
#pragma unroll
        for (int d = 0; d < 3; d++) {
            ca[d] = 0.0;
            cb[d] = 0.0;
        }
        for (int i = 0; i < n; i++)
#pragma unroll
            for (int d = 0; d < 3; d++) {
                ca[d] += A[3 * i + d];
                cb[d] += B[3 * i + d];
            }
#pragma unroll
        for (int d = 0; d < 3; d++) {
            ca[d] /= (double)n;
            cb[d] /= (double)n;
        }
#pragma unroll
        for (int j = 0; j < 3; j++)
#pragma unroll
            for (int k = 0; k < 3; k++) R[j][k] = rot[3 * j + k];
    } else {
        double X[3][3];
        centre_and_correlate(n, A, B, ca, cb, X);
        optimal_rotation(X, R);
    }
    const double scale = g / ((double)n * r);
    for (int i = 0; i < n; i++) {
        double u[3], v[3];
#pragma unroll
        for (int d = 0; d < 3; d++) {
            u[d] = A[3 * i + d] - ca[d];
            v[d] = B[3 * i + d] - cb[d];
        }
        double ga[3], gb[3];
#pragma unroll
        for (int d = 0; d < 3; d++) {
            // (R^T B_i)_d and (R A_i)_d
            const double rtb = R[0][d] * v[0] + R[1][d] * v[1] + R[2][d] * v[2];
            const double ra = R[d][0] * u[0] + R[d][1] * u[1] + R[d][2] * u[2];
            ga[d] = scale * (u[d] - rtb);
            gb[d] = scale * (v[d] - ra);
        }
#pragma unroll
        for (int d = 0; d < 3; d++) {
            if (GA) GA[3 * i + d] = ga[d];
            if (GB) GB[3 * i + d] = gb[d];
        }
    }
}

// One wave per tile of RG_SETS consecutive sets (one set per lane). When the
// tile's points fit, they are staged in LDS with coalesced loads (and the
// gradients leave through LDS with coalesced stores); otherwise each lane
// reads its own set from global memory.
#define RG_SETS 64
#define RG_CAP 768   // points per tile held in LDS (12 per set on average)

// Copies m doubles of a and of b into LDS, coalesced: every lane issues the
// loads of RG_BATCH elements of each array before its LDS stores, so that a
// wave waits for HBM once per batch rather than once per element (the
// kernels run one wave per SIMD, their LDS tiles allowing no more).
#define RG_BATCH 12
__device__ static inline void stage_pair(const double *a, const double *b, int m, double *la, double *lb, int lane)
{
    for (int q0 = 0; q0 < m; q0 += RG_SETS * RG_BATCH) {
        double ra[RG_BATCH], rb[RG_BATCH];
#pragma unroll
        for (int u = 0; u < RG_BATCH; u++) {
            const int q = q0 + lane + RG_SETS * u;
            ra[u] = q < m ? a[q] : 0.0;
            rb[u] = q < m ? b[q] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < RG_BATCH; u++) {
            const int q = q0 + lane + RG_SETS * u;
            if (q < m) {
                la[q] = ra[u];
                lb[q] = rb[u];
            }
        }
    }
}

// Range of set s, clamped into [lo, hi]: with valid offsets (off[0] = 0,
// never decreasing, off[n_sets] = n_points) this is [off[s], off[s+1]);
// malformed offsets give meaningless values, never an access outside a and b.
__device__ static inline void set_range(const int64_t *off, long long s, int64_t lo, int64_t hi, int64_t &p0,
                                        int64_t &p1)
{
    int64_t x = off[s], y = off[s + 1];
    x = x < lo ? lo : (x > hi ? hi : x);
    y = y < x ? x : (y > hi ? hi : y);
    p0 = x;
    p1 = y;
}

// Set s owns points [off[s], off[s+1]) of a and b, 3 doubles per point.
__global__ void __launch_bounds__(RG_SETS) k_rmsd_fwd(const double *a, const double *b, const int64_t *off,
                                                      long long n_sets, long long n_points, double *rmsd,
                                                      double *rot)
{
    __shared__ double la[3 * RG_CAP], lb[3 * RG_CAP];
    const int lane = threadIdx.x;
    const long long s0 = (long long)blockIdx.x * RG_SETS;
    if (s0 >= n_sets) return;
    const long long s1 = s0 + RG_SETS < n_sets ? s0 + RG_SETS : n_sets;
    // the tile's points: from its first set's start to its last set's end
    int64_t P0, P1, q0, q1;
    set_range(off, s0, 0, n_points, P0, q0);
    set_range(off, s1 - 1, P0, n_points, q1, P1);
    const bool staged = P1 - P0 <= RG_CAP;
    if (staged) {
        stage_pair(a + 3 * P0, b + 3 * P0, (int)(3 * (P1 - P0)), la, lb, lane);
        wave_sync();
    }
    const long long s = s0 + lane;
    if (s >= s1) return;
    int64_t p0, p1;
    set_range(off, s, P0, P1, p0, p1);
    const int n = (int)(p1 - p0);
    double *rt = rot ? rot + 9 * s : nullptr;
    if (n <= 0) {
        rmsd[s] = 0.0;
        if (rt)
            for (int q = 0; q < 9; q++) rt[q] = (q % 4 == 0) ? 1.0 : 0.0;
        return;
    }
    double r;
    if (staged) rmsd_one(n, la + 3 * (p0 - P0), lb + 3 * (p0 - P0), &r, rt);
    else rmsd_one(n, a + 3 * p0, b + 3 * p0, &r, rt);
    rmsd[s] = r;
}

__global__ void __launch_bounds__(RG_SETS) k_rmsd_bwd(const double *a, const double *b, const int64_t *off,
                                                      long long n_sets, long long n_points, const double *rmsd,
                                                      const double *rot, const double *grad_rmsd, double *grad_a,
                                                      double *grad_b)
{
    __shared__ double la[3 * RG_CAP], lb[3 * RG_CAP];
    const int lane = threadIdx.x;
    const long long s0 = (long long)blockIdx.x * RG_SETS;
    if (s0 >= n_sets) return;
    const long long s1 = s0 + RG_SETS < n_sets ? s0 + RG_SETS : n_sets;
    int64_t P0, P1, q0, q1;
    set_range(off, s0, 0, n_points, P0, q0);
    set_range(off, s1 - 1, P0, n_points, q1, P1);
    const bool staged = P1 - P0 <= RG_CAP;
    const int m = (int)(3 * (P1 - P0));
    if (staged) {
        stage_pair(a + 3 * P0, b + 3 * P0, m, la, lb, lane);
        wave_sync();
    }
    const long long s = s0 + lane;
    if (s < s1) {
        int64_t p0, p1;
        set_range(off, s, P0, P1, p0, p1);
        const int n = (int)(p1 - p0);
        if (n > 0) {
            const double *rt = rot ? rot + 9 * s : nullptr;
            if (staged) {
                // gradients replace the points in LDS
                double *A = la + 3 * (p0 - P0), *B = lb + 3 * (p0 - P0);
                rmsd_grad_one(n, A, B, rmsd[s], grad_rmsd[s], rt, A, B);
            } else {
                rmsd_grad_one(n, a + 3 * p0, b + 3 * p0, rmsd[s], grad_rmsd[s], rt,
                              grad_a ? grad_a + 3 * p0 : nullptr, grad_b ? grad_b + 3 * p0 : nullptr);
            }
        }
    }
    if (staged) {
        wave_sync();
        for (int q = lane; q < m; q += RG_SETS) {
            if (grad_a) grad_a[3 * P0 + q] = la[q];
            if (grad_b) grad_b[3 * P0 + q] = lb[q];
        }
    }
}

}  // namespace jam
