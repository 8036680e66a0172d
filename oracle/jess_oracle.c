/*
 * jess_oracle.c -- CPU restatement of Jess template matching.
 *
 * TEST INFRASTRUCTURE ONLY: the parity oracle and the CPU baseline. The
 * product library never links or calls this file.
 *
 * Reference: althonos/pyjess v0.10.1 and the Jess engine as patched by it
 * (paths relative to the reference checkout). The Jess C sources themselves are
 * not vendored in the reference checkout; what is restated here follows the
 * visible patch hunks and the pyjess Cython layer, and the hidden parts are
 * pinned by the reference's own fixtures (see tests/test_oracle_fixtures.py):
 *
 *   - residue index order ........ patches/ResIndex.c.patch:18-119
 *   - candidate lists ............ patches/TessTemplate.c.patch:173-235
 *   - match predicate ............ patches/TessAtom.c.patch:148-182 (helpers);
 *                                  the body of TessAtom_match is hidden, modes
 *                                  0, 1 and 3 are inferred from the fixtures
 *   - kd-tree + QuickSelect ...... patches/KdTree.c.patch:160-245,348-451,
 *                                  patches/qselect.h.patch:50-116
 *   - kd query (box prune, leaf annulus test) patches/KdTree.c.patch:226-311
 *   - annulus / box / join ....... patches/Annulus.c.patch:107-140,
 *                                  patches/Annulus.h.patch:54-68,
 *                                  patches/Box.h.patch:63-77,
 *                                  patches/Join.c.patch:154-196
 *   - reorder + Scanner DFS ...... patches/Scanner.c.patch:98-108,115-203,232-334
 *   - check() .................... patches/TessTemplate.c.patch:101-156
 *   - superposition .............. patches/Super.c.patch:35-64 (RMSD formula,
 *                                  reflection rule, eigenvalue clamp); the
 *                                  Jacobi solver and rotation are our own
 *                                  (cyclic Jacobi, Golub & Van Loan 8.5)
 *   - template scan order ........ patches/Jess.c.patch:63-72 (prepend)
 *
 * Compile with -ffp-contract=off: the reference sums squared differences
 * without fused multiply-add.
 */
#include "jess_oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------ */
/* names                                                               */
/* ------------------------------------------------------------------ */

static inline char up(char c) { return (c >= 'a' && c <= 'z') ? (char)(c & ~0x20) : c; }

/* 4-byte case-insensitive compare (TessAtom_compareName, patches/TessAtom.h.patch:74-84). */
static int name4_cmp(const char *a, const char *b)
{
    for (int i = 0; i < 4; i++) {
        int d = (int)up(a[i]) - (int)up(b[i]);
        if (d) return d;
    }
    return 0;
}

/* ------------------------------------------------------------------ */
/* residue index                                                       */
/* ------------------------------------------------------------------ */

typedef struct {
    int32_t nres;      /* distinct residue names */
    int32_t *first;    /* nres+1 offsets into perm */
    int32_t *perm;     /* atoms sorted by (resName, serial) */
} resindex;

static int res_before(const jo_mol *m, int32_t a, int32_t b)
{
    int c = name4_cmp(m->resname + 4 * a, m->resname + 4 * b);
    if (c) return c < 0;
    return m->serial[a] < m->serial[b];
}

/* Stable merge sort by (resName, serial); ties keep molecule order, as the
 * glibc merge-based qsort does for ResNames_compare (patches/ResIndex.c.patch:18-26,60-61). */
static void msort(const jo_mol *m, int32_t *v, int32_t *tmp, int32_t n)
{
    if (n < 2) return;
    int32_t h = n / 2;
    msort(m, v, tmp, h);
    msort(m, v + h, tmp, n - h);
    int32_t i = 0, j = h, k = 0;
    while (i < h && j < n) {
        if (res_before(m, v[j], v[i])) tmp[k++] = v[j++];
        else tmp[k++] = v[i++];
    }
    while (i < h) tmp[k++] = v[i++];
    while (j < n) tmp[k++] = v[j++];
    memcpy(v, tmp, sizeof(int32_t) * n);
}

static int resindex_build(const jo_mol *m, resindex *R)
{
    int32_t n = m->n;
    R->nres = 0;
    R->first = NULL;
    R->perm = (int32_t *)malloc(sizeof(int32_t) * (n > 0 ? n : 1));
    int32_t *tmp = (int32_t *)malloc(sizeof(int32_t) * (n > 0 ? n : 1));
    R->first = (int32_t *)malloc(sizeof(int32_t) * (n + 1));
    if (!R->perm || !tmp || !R->first) { free(tmp); return JO_ENOMEM; }
    for (int32_t i = 0; i < n; i++) R->perm[i] = i;
    msort(m, R->perm, tmp, n);
    free(tmp);
    for (int32_t i = 0; i < n; i++) {
        if (i == 0 || name4_cmp(m->resname + 4 * R->perm[i - 1], m->resname + 4 * R->perm[i]) != 0)
            R->first[R->nres++] = i;
    }
    R->first[R->nres] = n;
    return JO_OK;
}

static int32_t resindex_find(const jo_mol *m, const resindex *R, const char *rn)
{
    int32_t lo = 0, hi = R->nres;
    while (lo < hi) {
        int32_t mid = (lo + hi) / 2;
        int c = name4_cmp(m->resname + 4 * R->perm[R->first[mid]], rn);
        if (c < 0) lo = mid + 1;
        else if (c > 0) hi = mid;
        else return mid;
    }
    return -1;
}

static void resindex_free(resindex *R)
{
    free(R->first);
    free(R->perm);
}

/* ------------------------------------------------------------------ */
/* match predicate                                                     */
/* ------------------------------------------------------------------ */

/* Modes whose semantics are pinned by the reference fixtures:
 *   0: atom name in the template's names, residue name in its residue names;
 *   1: same position letter (name[2]) as a template name, molecule atom not
 *      a carbon (name "_C.."), residue name in the list;
 *   3: same element letters (name[0..1]) as a template name, any position,
 *      residue name in the list.
 * Helpers follow patches/TessAtom.c.patch:148-181. */
static int is_carbon(const char *nm) { return nm[0] == '_' && nm[1] == 'C'; }

int jo_mode_supported(int32_t mode) { return mode == 0 || mode == 1 || mode == 3; }

static int match_resname(const jo_tpls *T, int32_t a, const char *rn)
{
    for (int32_t r = T->resname_off[a]; r < T->resname_off[a + 1]; r++)
        if (name4_cmp(rn, T->resnames + 4 * r) == 0) return 1;
    return 0;
}

static int atom_match(const jo_tpls *T, int32_t a, const jo_mol *m, int32_t i)
{
    const char *nm = m->name + 4 * i;
    if (!match_resname(T, a, m->resname + 4 * i)) return 0;
    int32_t mode = T->mode[a];
    for (int32_t k = T->name_off[a]; k < T->name_off[a + 1]; k++) {
        const char *tn = T->names + 4 * k;
        switch (mode) {
        case 0:
            if (name4_cmp(nm, tn) == 0) return 1;
            break;
        case 1:
            if (up(nm[2]) == up(tn[2]) && !is_carbon(nm)) return 1;
            break;
        case 3:
            if (up(nm[0]) == up(tn[0]) && up(nm[1]) == up(tn[1])) return 1;
            break;
        default:
            return 0;
        }
    }
    return 0;
}

/* Candidate list of template atom a (TessTemplate_candidates,
 * patches/TessTemplate.c.patch:173-235): modes -1..8 walk the residue buckets
 * of the listed residue names (first occurrence of each bucket only), atoms in
 * (resName, serial) order; modes >= 100 scan the molecule in order. */
static int32_t candidates(const jo_mol *m, const resindex *R, const jo_tpls *T, int32_t a,
                          int32_t *out, char *done)
{
    int32_t cnt = 0;
    int32_t mode = T->mode[a];
    if (!jo_mode_supported(mode)) return JO_EMODE;
    if (mode >= -1 && mode <= 8) {
        memset(done, 0, R->nres > 0 ? R->nres : 1);
        for (int32_t r = T->resname_off[a]; r < T->resname_off[a + 1]; r++) {
            int32_t b = resindex_find(m, R, T->resnames + 4 * r);
            if (b < 0 || done[b]) continue;
            done[b] = 1;
            for (int32_t q = R->first[b]; q < R->first[b + 1]; q++)
                if (atom_match(T, a, m, R->perm[q])) out[cnt++] = R->perm[q];
        }
    } else {
        for (int32_t i = 0; i < m->n; i++)
            if (atom_match(T, a, m, i)) out[cnt++] = i;
    }
    return cnt;
}

/* ------------------------------------------------------------------ */
/* kd-tree over one candidate set                                      */
/* ------------------------------------------------------------------ */

typedef struct {
    int32_t left, right;  /* -1 for leaves */
    int32_t pos;          /* leaf: position in the candidate list */
    int32_t depth;
    double lo[3], hi[3];
} kdnode;

typedef struct {
    kdnode *node;
    int32_t count, cap;
    int32_t root;
} kdtree;

typedef struct {
    const double *xyz;
    const int32_t *atoms; /* candidate list -> molecule atom */
    int axis;
} kdcmp_ctx;

static inline double coord(const kdcmp_ctx *c, int32_t pos)
{
    return c->xyz[3 * c->atoms[pos] + c->axis];
}

static inline int kdcmp(const kdcmp_ctx *c, int32_t pa, int32_t pb)
{
    double x = coord(c, pa), y = coord(c, pb);
    return x < y ? -1 : (x > y ? 1 : 0);
}

static inline void swp(int32_t *v, size_t i, size_t j) { int32_t t = v[i]; v[i] = v[j]; v[j] = t; }

/* QuickSelect with median-of-three pivot and Lomuto partition
 * (patches/qselect.h.patch:50-116). The exact swap sequence matters: it
 * decides which of several equal coordinates go left of the median, hence
 * the kd leaf order and the enumeration order of matches. */
static size_t qs_median3(const kdcmp_ctx *c, const int32_t *v, size_t l, size_t r)
{
    size_t mid = (l + r + 1) / 2;
    if ((kdcmp(c, v[l], v[mid]) > 0) != (kdcmp(c, v[l], v[r]) > 0)) return l;
    if ((kdcmp(c, v[mid], v[l]) < 0) != (kdcmp(c, v[mid], v[r]) < 0)) return mid;
    return r;
}

static size_t qs_partition(const kdcmp_ctx *c, int32_t *v, size_t l, size_t r, size_t p)
{
    size_t store = l;
    swp(v, r, p);
    for (size_t i = l; i < r; i++) {
        if (kdcmp(c, v[i], v[r]) <= 0) {
            swp(v, i, store);
            store++;
        }
    }
    swp(v, store, r);
    return store;
}

static size_t qs_select(const kdcmp_ctx *c, int32_t *v, size_t n, size_t k)
{
    size_t l = 0, r = n - 1;
    for (;;) {
        if (l == r) return l;
        size_t p = qs_median3(c, v, l, r);
        p = qs_partition(c, v, l, r, p);
        if (p == k) return p;
        if (k < p) r = p - 1;
        else l = p + 1;
    }
}

static int32_t kd_alloc(kdtree *K)
{
    if (K->count >= K->cap) {
        int32_t nc = K->cap ? K->cap * 2 : 64;
        kdnode *nn = (kdnode *)realloc(K->node, sizeof(kdnode) * nc);
        if (!nn) return -1;
        K->node = nn;
        K->cap = nc;
    }
    return K->count++;
}

/* KdTreeNode_create (patches/KdTree.c.patch:348-451): leaf per point; interior
 * node splits its index range at n/2 after QuickSelect on axis depth%3. */
static int32_t kd_build(kdtree *K, kdcmp_ctx *c, int32_t *idx, int32_t n, int axis)
{
    int32_t k = kd_alloc(K);
    if (k < 0) return -1;
    if (n == 1) {
        kdnode *N = &K->node[k];
        N->left = N->right = -1;
        N->pos = idx[0];
        N->depth = 1;
        for (int d = 0; d < 3; d++) N->lo[d] = N->hi[d] = c->xyz[3 * c->atoms[idx[0]] + d];
        return k;
    }
    c->axis = axis;
    size_t split = qs_select(c, idx, (size_t)n, (size_t)(n / 2));
    int na = (axis + 1) % 3;
    int32_t L = kd_build(K, c, idx, (int32_t)split, na);
    int32_t R = kd_build(K, c, idx + split, n - (int32_t)split, na);
    if (L < 0 || R < 0) return -1;
    kdnode *N = &K->node[k];
    N->left = L;
    N->right = R;
    N->pos = -1;
    N->depth = (K->node[L].depth > K->node[R].depth ? K->node[L].depth : K->node[R].depth) + 1;
    for (int d = 0; d < 3; d++) {
        double a = K->node[L].lo[d], b = K->node[R].lo[d];
        N->lo[d] = a < b ? a : b;
        a = K->node[L].hi[d];
        b = K->node[R].hi[d];
        N->hi[d] = a > b ? a : b;
    }
    return k;
}

static int kdtree_build(kdtree *K, const double *xyz, const int32_t *atoms, int32_t n, int32_t *scratch)
{
    K->count = 0;
    K->root = -1;
    for (int32_t i = 0; i < n; i++) scratch[i] = i;
    kdcmp_ctx c = {xyz, atoms, 0};
    K->root = kd_build(K, &c, scratch, n, 0);
    return K->root < 0 ? JO_ENOMEM : JO_OK;
}

/* Leaf order of a kd-tree built over the candidate list (exported for tests). */
int32_t jo_kd_order(const jo_mol *mol, const int32_t *set_atoms, int32_t n, int32_t *out)
{
    if (n <= 0) return 0;
    for (int32_t i = 0; i < n; i++) out[i] = i;
    kdtree K = {0};
    kdcmp_ctx c = {mol->xyz, set_atoms, 0};
    /* building permutes `out` into leaf order (left subtree first) */
    int32_t r = kd_build(&K, &c, out, n, 0);
    free(K.node);
    return r < 0 ? JO_ENOMEM : n;
}

/* ------------------------------------------------------------------ */
/* annulus, box, join                                                  */
/* ------------------------------------------------------------------ */

typedef struct {
    double c[3];
    double lo[3], hi[3];  /* bounding box c +/- b */
    double min2, max2;
} annulus;

/* Annulus_reuse (patches/Annulus.c.patch:107-140). */
static void annulus_set(annulus *A, const double *u, double a, double b)
{
    if (b < a) { double t = a; a = b; b = t; }
    if (a < 0.0) a = 0.0;
    if (b < 0.0) b = 0.0;
    for (int d = 0; d < 3; d++) {
        A->c[d] = u[d];
        A->lo[d] = u[d] - b;
        A->hi[d] = u[d] + b;
    }
    A->min2 = a * a;
    A->max2 = b * b;
}

/* _Annulus_po (patches/Annulus.h.patch:54-68): squared distance summed x, y, z. */
static inline int annulus_has(const annulus *A, const double *x)
{
    double s = 0.0;
    for (int d = 0; d < 3; d++) {
        double t = A->c[d] - x[d];
        s += t * t;
    }
    return !(s < A->min2 || s > A->max2);
}

/* ------------------------------------------------------------------ */
/* superposition                                                       */
/* ------------------------------------------------------------------ */

/* One cyclic-Jacobi eigen-decomposition of a symmetric 3x3 matrix
 * (Golub & Van Loan, Matrix Computations, Alg. 8.5.2/8.5.3). A is destroyed;
 * e receives the diagonal, P the eigenvectors in its columns. An off-diagonal
 * entry is zeroed without rotation once it is below 1e-12 of the diagonal
 * magnitudes (the PRECISION constant of patches/Super.c.patch:35). */
static void jacobi3(double A[3][3], double P[3][3], double e[3])
{
    static const int PQ[3][2] = {{0, 1}, {0, 2}, {1, 2}};
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) P[i][j] = (i == j) ? 1.0 : 0.0;
    for (int sweep = 0; sweep < 32; sweep++) {
        int rotated = 0;
        for (int r = 0; r < 3; r++) {
            int p = PQ[r][0], q = PQ[r][1];
            double apq = A[p][q];
            if (apq == 0.0) continue;
            double scale = fabs(A[p][p]) + fabs(A[q][q]);
            if (fabs(apq) <= 1e-12 * scale) {
                A[p][q] = A[q][p] = 0.0;
                continue;
            }
            double theta = (A[q][q] - A[p][p]) / (2.0 * apq);
            double t = 1.0 / (fabs(theta) + sqrt(theta * theta + 1.0));
            if (theta < 0.0) t = -t;
            double c = 1.0 / sqrt(t * t + 1.0);
            double s = t * c;
            /* A <- J^T A J with J the (p,q) rotation [c s; -s c] */
            for (int k = 0; k < 3; k++) {
                double akp = A[k][p], akq = A[k][q];
                A[k][p] = c * akp - s * akq;
                A[k][q] = s * akp + c * akq;
            }
            for (int k = 0; k < 3; k++) {
                double apk = A[p][k], aqk = A[q][k];
                A[p][k] = c * apk - s * aqk;
                A[q][k] = s * apk + c * aqk;
            }
            A[p][q] = A[q][p] = 0.0;
            for (int k = 0; k < 3; k++) {
                double pkp = P[k][p], pkq = P[k][q];
                P[k][p] = c * pkp - s * pkq;
                P[k][q] = s * pkp + c * pkq;
            }
            rotated = 1;
        }
        if (!rotated) break;
    }
    for (int i = 0; i < 3; i++) e[i] = A[i][i];
}

/* Least-squares superposition of molecule points a onto template points b.
 * RMSD rule as in patches/Super.c.patch:40-64: eigenvalues of X^T X clamped
 * at 0, sum of their square roots, minus twice the smallest when det(X) <
 * 1e-8 (reflection), rmsd = sqrt(max(SA + SB - 2 sumE, 0) / n). The rotation
 * M = X P diag(sigma_i / sqrt(e_i)) P^T maps centred molecule coordinates to
 * centred template coordinates; it is NaN/inf when X^T X is singular, as the
 * reference's is ("fails if XX has rank<3"). */
void jo_superpose(const double *a, const double *b, int32_t n,
                  double *rmsd, double *rot, double *c0, double *c1)
{
    double ca[3] = {0, 0, 0}, cb[3] = {0, 0, 0};
    for (int32_t i = 0; i < n; i++)
        for (int d = 0; d < 3; d++) {
            ca[d] += a[3 * i + d];
            cb[d] += b[3 * i + d];
        }
    for (int d = 0; d < 3; d++) {
        ca[d] /= (double)n;
        cb[d] /= (double)n;
    }
    double X[3][3] = {{0}}, sa = 0.0, sb = 0.0;
    for (int32_t i = 0; i < n; i++) {
        double u[3], v[3];
        for (int d = 0; d < 3; d++) {
            u[d] = a[3 * i + d] - ca[d];
            v[d] = b[3 * i + d] - cb[d];
        }
        for (int d = 0; d < 3; d++) {
            sa += u[d] * u[d];
            sb += v[d] * v[d];
        }
        for (int j = 0; j < 3; j++)
            for (int k = 0; k < 3; k++) X[j][k] += v[j] * u[k];
    }
    double det = X[0][0] * (X[1][1] * X[2][2] - X[1][2] * X[2][1])
               - X[0][1] * (X[1][0] * X[2][2] - X[1][2] * X[2][0])
               + X[0][2] * (X[1][0] * X[2][1] - X[1][1] * X[2][0]);
    double XX[3][3];
    for (int j = 0; j < 3; j++)
        for (int k = 0; k < 3; k++) {
            double s = 0.0;
            for (int l = 0; l < 3; l++) s += X[l][j] * X[l][k];
            XX[j][k] = s;
        }
    double P[3][3], e[3];
    jacobi3(XX, P, e);
    for (int i = 0; i < 3; i++) e[i] = fmax(e[i], 0.0);
    int imin = 0;
    for (int i = 1; i < 3; i++)
        if (e[i] < e[imin]) imin = i;
    double sumE = sqrt(e[0]) + sqrt(e[1]) + sqrt(e[2]);
    double sig[3] = {1.0, 1.0, 1.0};
    if (det < 1e-8) {
        sumE -= 2.0 * sqrt(e[imin]);
        sig[imin] = -1.0;
    }
    double r = sa + sb - 2.0 * sumE;
    *rmsd = sqrt(fmax(r, 0.0) / (double)n);
    /* M = sum_i w_i (X p_i) p_i^T */
    double M[3][3] = {{0}};
    for (int i = 0; i < 3; i++) {
        double w = sig[i] / sqrt(e[i]);
        double xp[3];
        for (int j = 0; j < 3; j++) xp[j] = X[j][0] * P[0][i] + X[j][1] * P[1][i] + X[j][2] * P[2][i];
        for (int j = 0; j < 3; j++)
            for (int k = 0; k < 3; k++) M[j][k] += (w * xp[j]) * P[k][i];
    }
    for (int j = 0; j < 3; j++)
        for (int k = 0; k < 3; k++) rot[3 * j + k] = M[j][k];
    for (int d = 0; d < 3; d++) {
        c0[d] = ca[d];
        c1[d] = cb[d];
    }
}

/* ------------------------------------------------------------------ */
/* template helpers                                                    */
/* ------------------------------------------------------------------ */

/* TessAtom_distance (patches/TessAtom.c.patch:229-245). */
static double tdist(const jo_tpls *T, int32_t i, int32_t j)
{
    double s = 0.0;
    for (int d = 0; d < 3; d++) {
        double t = T->pos[3 * i + d] - T->pos[3 * j + d];
        s += t * t;
    }
    return sqrt(s);
}

/* TessTemplate_check (patches/TessTemplate.c.patch:101-156): the atom placed
 * at level L (template atom ord[L]) against every earlier level. */
static int check_level(const jo_tpls *T, int32_t base, const jo_mol *m, const int32_t *atom,
                       const int32_t *ord, int32_t L, int32_t ignore)
{
    int32_t p = ord[L];
    for (int32_t i = 0; i < L; i++) {
        int32_t q = ord[i];
        int c = (int)m->chain2[atom[q]] - (int)m->chain2[atom[p]];
        int d = (int)T->chain2[base + q] - (int)T->chain2[base + p];
        if (ignore == 0) {
            if (c == 0 && d != 0) return 0;
            if (c != 0 && d == 0) return 0;
            if (c != 0) continue;
        }
        int64_t ra = (int64_t)m->resseq[atom[q]] - (int64_t)m->resseq[atom[p]];
        int64_t rb = (int64_t)T->resseq[base + q] - (int64_t)T->resseq[base + p];
        if (ra == 0 && rb != 0) return 0;
        if (ra != 0 && rb == 0) return 0;
        if (ignore != 2 && ra == 0 && c != 0) return 0;
    }
    return 1;
}

/* Annulus radii for the pair (template atoms i, j) (patches/Scanner.c.patch:274-292):
 * tau = t + w_i + w_j capped at s; window [d - tau, d + tau], lower end >= 0.5. */
static void pair_window(const jo_tpls *T, int32_t base, int32_t i, int32_t j, const jo_params *p,
                        double *lo, double *hi)
{
    double d = tdist(T, base + i, base + j);
    double tau = p->distance_cutoff + T->weight[base + i] + T->weight[base + j];
    if (tau > p->max_dynamic_distance) tau = p->max_dynamic_distance;
    double a = d - tau, b = d + tau;
    if (a < 0.5) a = 0.5;
    *lo = a;
    *hi = b;
}

/* ------------------------------------------------------------------ */
/* output stream                                                       */
/* ------------------------------------------------------------------ */

struct jo_stream {
    int64_t count, cap, acount, acap;
    int32_t *tpl;
    int64_t *aoff;
    int32_t *atoms;
    double *rmsd, *rot, *c0, *c1;
};

static int stream_push(jo_stream *S, int32_t t, const int32_t *atoms, int32_t n,
                       double rmsd, const double *rot, const double *c0, const double *c1)
{
    if (S->count >= S->cap) {
        int64_t nc = S->cap ? S->cap * 2 : 256;
        int32_t *nt = (int32_t *)realloc(S->tpl, sizeof(int32_t) * nc);
        if (nt) S->tpl = nt;
        int64_t *no = (int64_t *)realloc(S->aoff, sizeof(int64_t) * (nc + 1));
        if (no) S->aoff = no;
        double *nr = (double *)realloc(S->rmsd, sizeof(double) * nc);
        if (nr) S->rmsd = nr;
        double *nm = (double *)realloc(S->rot, sizeof(double) * 9 * nc);
        if (nm) S->rot = nm;
        double *n0 = (double *)realloc(S->c0, sizeof(double) * 3 * nc);
        if (n0) S->c0 = n0;
        double *n1 = (double *)realloc(S->c1, sizeof(double) * 3 * nc);
        if (n1) S->c1 = n1;
        if (!nt || !no || !nr || !nm || !n0 || !n1) return JO_ENOMEM;
        S->cap = nc;
    }
    if (S->acount + n > S->acap) {
        int64_t nc = S->acap ? S->acap * 2 : 1024;
        while (nc < S->acount + n) nc *= 2;
        int32_t *na = (int32_t *)realloc(S->atoms, sizeof(int32_t) * nc);
        if (!na) return JO_ENOMEM;
        S->atoms = na;
        S->acap = nc;
    }
    int64_t i = S->count;
    S->tpl[i] = t;
    S->aoff[i] = S->acount;
    memcpy(S->atoms + S->acount, atoms, sizeof(int32_t) * n);
    S->acount += n;
    S->aoff[i + 1] = S->acount;
    S->rmsd[i] = rmsd;
    memcpy(S->rot + 9 * i, rot, 9 * sizeof(double));
    memcpy(S->c0 + 3 * i, c0, 3 * sizeof(double));
    memcpy(S->c1 + 3 * i, c1, 3 * sizeof(double));
    S->count++;
    return JO_OK;
}

int64_t jo_stream_count(const jo_stream *s) { return s ? s->count : 0; }
int64_t jo_stream_atom_total(const jo_stream *s) { return s ? s->acount : 0; }

void jo_stream_copy(const jo_stream *s, int32_t *tpl, int64_t *atom_off, int32_t *atoms,
                    double *rmsd, double *rot, double *c0, double *c1)
{
    if (!s || s->count == 0) {
        if (atom_off) atom_off[0] = 0;
        return;
    }
    memcpy(tpl, s->tpl, sizeof(int32_t) * s->count);
    memcpy(atom_off, s->aoff, sizeof(int64_t) * (s->count + 1));
    memcpy(atoms, s->atoms, sizeof(int32_t) * s->acount);
    memcpy(rmsd, s->rmsd, sizeof(double) * s->count);
    memcpy(rot, s->rot, sizeof(double) * 9 * s->count);
    memcpy(c0, s->c0, sizeof(double) * 3 * s->count);
    memcpy(c1, s->c1, sizeof(double) * 3 * s->count);
}

void jo_stream_free(jo_stream *s)
{
    if (!s) return;
    free(s->tpl);
    free(s->aoff);
    free(s->atoms);
    free(s->rmsd);
    free(s->rot);
    free(s->c0);
    free(s->c1);
    free(s);
}

/* ------------------------------------------------------------------ */
/* the scanner                                                         */
/* ------------------------------------------------------------------ */

typedef struct {
    int32_t nmax;             /* capacity (template atoms) */
    int32_t setcap;           /* capacity per candidate set */
    int32_t **set;            /* candidate lists */
    int32_t *cnt;
    kdtree *tree;
    int32_t **stack;          /* kd query stacks */
    int32_t *sp;
    annulus **ann;            /* per level: L annuli */
    double (*box)[6];         /* per level join box */
    int32_t *ord, *atom, *cur0;
    int32_t *scratch;
    char *done;
    double *pa, *pb;
    /* work counters (jo_work): sum over opened levels L of L * |set inside join box|,
     * opened levels, complete matches, sum of their sizes */
    int64_t w_inbox, w_levels, w_complete, w_complete_atoms;
} scanner;

static void scanner_free(scanner *S)
{
    for (int32_t k = 0; k < S->nmax; k++) {
        if (S->set) free(S->set[k]);
        if (S->tree) free(S->tree[k].node);
        if (S->stack) free(S->stack[k]);
        if (S->ann) free(S->ann[k]);
    }
    free(S->set);
    free(S->cnt);
    free(S->tree);
    free(S->stack);
    free(S->sp);
    free(S->ann);
    free(S->box);
    free(S->ord);
    free(S->atom);
    free(S->cur0);
    free(S->scratch);
    free(S->done);
    free(S->pa);
    free(S->pb);
    memset(S, 0, sizeof(*S));
}

static int scanner_init(scanner *S, int32_t nmax, int32_t natoms, int32_t nres)
{
    memset(S, 0, sizeof(*S));
    S->w_inbox = -1;
    S->nmax = nmax;
    S->setcap = natoms > 0 ? natoms : 1;
    S->set = (int32_t **)calloc(nmax, sizeof(int32_t *));
    S->cnt = (int32_t *)calloc(nmax, sizeof(int32_t));
    S->tree = (kdtree *)calloc(nmax, sizeof(kdtree));
    S->stack = (int32_t **)calloc(nmax, sizeof(int32_t *));
    S->sp = (int32_t *)calloc(nmax, sizeof(int32_t));
    S->ann = (annulus **)calloc(nmax, sizeof(annulus *));
    S->box = (double(*)[6])calloc(nmax, sizeof(double[6]));
    S->ord = (int32_t *)calloc(nmax, sizeof(int32_t));
    S->atom = (int32_t *)calloc(nmax, sizeof(int32_t));
    S->cur0 = (int32_t *)calloc(nmax, sizeof(int32_t));
    S->scratch = (int32_t *)calloc(S->setcap, sizeof(int32_t));
    S->done = (char *)calloc(nres > 0 ? nres : 1, 1);
    S->pa = (double *)calloc(3 * nmax, sizeof(double));
    S->pb = (double *)calloc(3 * nmax, sizeof(double));
    if (!S->set || !S->cnt || !S->tree || !S->stack || !S->sp || !S->ann || !S->box || !S->ord ||
        !S->atom || !S->cur0 || !S->scratch || !S->done || !S->pa || !S->pb)
        return JO_ENOMEM;
    for (int32_t k = 0; k < nmax; k++) {
        S->set[k] = (int32_t *)malloc(sizeof(int32_t) * S->setcap);
        /* a kd DFS stack never holds more entries than the tree has levels + 1 */
        S->stack[k] = (int32_t *)malloc(sizeof(int32_t) * 2 * (S->setcap + 2));
        S->ann[k] = (annulus *)malloc(sizeof(annulus) * (nmax > 0 ? nmax : 1));
        if (!S->set[k] || !S->stack[k] || !S->ann[k]) return JO_ENOMEM;
    }
    return JO_OK;
}

/* KdTreeQuery_reuse + Join_computeBox (patches/KdTree.c.patch:226-257,
 * patches/Join.c.patch:154-196). */
static void kdq_start(scanner *S, int32_t L)
{
    int32_t k = S->ord[L];
    double *B = S->box[L];
    for (int d = 0; d < 3; d++) {
        B[d] = -INFINITY;
        B[3 + d] = INFINITY;
        for (int32_t j = 0; j < L; j++) {
            if (S->ann[L][j].lo[d] > B[d]) B[d] = S->ann[L][j].lo[d];
            if (S->ann[L][j].hi[d] < B[3 + d]) B[3 + d] = S->ann[L][j].hi[d];
        }
    }
    S->sp[L] = 0;
    S->stack[L][S->sp[L]++] = S->tree[k].root;
}

/* KdTreeQuery_next (patches/KdTree.c.patch:264-311): explicit-stack DFS,
 * interior nodes pruned by the join box, leaves tested against every annulus. */
static int32_t kdq_next(scanner *S, const jo_mol *m, int32_t L, int right_first)
{
    int32_t k = S->ord[L];
    const kdtree *K = &S->tree[k];
    const double *B = S->box[L];
    int32_t *st = S->stack[L];
    while (S->sp[L] > 0) {
        const kdnode *N = &K->node[st[--S->sp[L]]];
        if (N->left < 0) {
            const double *x = m->xyz + 3 * S->set[k][N->pos];
            int in = 1;
            for (int32_t j = 0; j < L && in; j++) in = annulus_has(&S->ann[L][j], x);
            if (in) return N->pos;
            continue;
        }
        int miss = 0;
        for (int d = 0; d < 3; d++)
            if (N->lo[d] > B[3 + d] || B[d] > N->hi[d]) miss = 1;
        if (miss) continue;
        if (right_first) {
            st[S->sp[L]++] = N->left;
            st[S->sp[L]++] = N->right;
        } else {
            st[S->sp[L]++] = N->right;
            st[S->sp[L]++] = N->left;
        }
    }
    return -1;
}

static int emit(scanner *S, const jo_mol *m, const jo_tpls *T, int32_t t, int32_t n, jo_stream *out)
{
    int32_t base = T->atom_off[t];
    for (int32_t i = 0; i < n; i++)
        for (int d = 0; d < 3; d++) {
            S->pa[3 * i + d] = m->xyz[3 * S->atom[i] + d];
            S->pb[3 * i + d] = T->pos[3 * (base + i) + d];
        }
    double rmsd, rot[9], c0[3], c1[3];
    jo_superpose(S->pa, S->pb, n, &rmsd, rot, c0, c1);
    return stream_push(out, t, S->atom, n, rmsd, rot, c0, c1);
}

/* Per-callback sink so the batch driver can count instead of storing. */
typedef int (*match_sink)(void *ctx, scanner *S, const jo_mol *m, const jo_tpls *T, int32_t t, int32_t n);

static int sink_stream(void *ctx, scanner *S, const jo_mol *m, const jo_tpls *T, int32_t t, int32_t n)
{
    return emit(S, m, T, t, n, (jo_stream *)ctx);
}

static int sink_none(void *ctx, scanner *S, const jo_mol *m, const jo_tpls *T, int32_t t, int32_t n)
{
    S->w_complete++;
    S->w_complete_atoms += n;
    return 0;
}

/* Builds the join of level L: one annulus per earlier level, centred on the
 * atom placed there (patches/Scanner.c.patch:274-294), then starts the query. */
static void open_level(scanner *S, const jo_mol *m, const jo_tpls *T, int32_t base,
                       const jo_params *p, int32_t L)
{
    for (int32_t j = 0; j < L; j++) {
        double lo, hi;
        pair_window(T, base, S->ord[j], S->ord[L], p, &lo, &hi);
        annulus_set(&S->ann[L][j], m->xyz + 3 * S->atom[S->ord[j]], lo, hi);
    }
    kdq_start(S, L);
    S->w_levels++;
    if (S->w_inbox >= 0) {
        const double *B = S->box[L];
        const int32_t k = S->ord[L];
        int64_t in = 0;
        for (int32_t i = 0; i < S->cnt[k]; i++) {
            const double *x = m->xyz + 3 * S->set[k][i];
            int ok = 1;
            for (int d = 0; d < 3; d++)
                if (x[d] < B[d] || x[d] > B[3 + d]) ok = 0;
            in += ok;
        }
        S->w_inbox += (int64_t)L * in;
    }
}

/* Scanner_reuse + Scanner_next for one template (patches/Scanner.c.patch:115-203,232-334).
 * Levels are visited depth-first; level 0 walks its candidate list in order,
 * deeper levels walk kd query results in DFS leaf order. For a 1-atom template
 * the reference's first advance steps past list entry 0 before yielding. */
static int scan_template(scanner *S, const jo_mol *m, const resindex *R, const jo_tpls *T, int32_t t,
                         const jo_params *p, match_sink sink, void *ctx)
{
    int32_t base = T->atom_off[t], n = T->atom_off[t + 1] - base;
    if (n <= 0) return JO_OK;
    for (int32_t k = 0; k < n; k++) {
        int32_t c = candidates(m, R, T, base + k, S->set[k], S->done);
        if (c < 0) return c;
        S->cnt[k] = c;
        S->ord[k] = k;
    }
    for (int32_t k = 0; k < n; k++)
        if (S->cnt[k] == 0) return JO_OK;
    for (int32_t k = 0; k < n; k++)
        if (kdtree_build(&S->tree[k], m->xyz, S->set[k], S->cnt[k], S->scratch) != JO_OK) return JO_ENOMEM;
    if (p->reorder) {
        /* stable order by (|S_k|, k): _CandidateSet_size_compare_r (patches/Scanner.c.patch:98-108) */
        for (int32_t i = 1; i < n; i++) {
            int32_t v = S->ord[i], j = i - 1;
            while (j >= 0 && (S->cnt[S->ord[j]] > S->cnt[v] ||
                              (S->cnt[S->ord[j]] == S->cnt[v] && S->ord[j] > v))) {
                S->ord[j + 1] = S->ord[j];
                j--;
            }
            S->ord[j + 1] = v;
        }
    }
    int32_t o0 = S->ord[0];
    for (int32_t i0 = (n == 1 ? 1 : 0); i0 < S->cnt[o0]; i0++) {
        S->atom[o0] = S->set[o0][i0];
        if (n == 1) {
            int r = sink(ctx, S, m, T, t, n);
            if (r) return r;
            continue;
        }
        int32_t L = 1;
        open_level(S, m, T, base, p, L);
        while (L > 0) {
            int32_t pos = kdq_next(S, m, L, p->kd_right_first);
            if (pos < 0) {
                L--;
                continue;
            }
            S->atom[S->ord[L]] = S->set[S->ord[L]][pos];
            if (!check_level(T, base, m, S->atom, S->ord, L, p->ignore_chain)) continue;
            if (L == n - 1) {
                int r = sink(ctx, S, m, T, t, n);
                if (r) return r;
                continue;
            }
            L++;
            open_level(S, m, T, base, p, L);
        }
    }
    return JO_OK;
}

static int32_t max_template_atoms(const jo_tpls *T)
{
    int32_t nmax = 1;
    for (int32_t t = 0; t < T->n_templates; t++) {
        int32_t n = T->atom_off[t + 1] - T->atom_off[t];
        if (n > nmax) nmax = n;
    }
    return nmax;
}

int32_t jo_query(const jo_mol *mol, const jo_tpls *tpls, const jo_params *p,
                 const int32_t *scan_order, int32_t n_scan, jo_stream **out)
{
    *out = (jo_stream *)calloc(1, sizeof(jo_stream));
    if (!*out) return JO_ENOMEM;
    resindex R;
    if (resindex_build(mol, &R) != JO_OK) { resindex_free(&R); return JO_ENOMEM; }
    scanner S;
    int32_t rc = scanner_init(&S, max_template_atoms(tpls), mol->n, R.nres);
    for (int32_t i = 0; rc == JO_OK && i < n_scan; i++)
        rc = scan_template(&S, mol, &R, tpls, scan_order[i], p, sink_stream, *out);
    scanner_free(&S);
    resindex_free(&R);
    return rc;
}

/* Algorithmic work of a query, counted on the reference algorithm: out[0] =
 * sum over opened levels L of L * (candidates of the level's set inside the
 * join box), out[1] = opened levels, out[2] = complete matches, out[3] = sum
 * of their atom counts. */
int32_t jo_work(const jo_mol *mol, const jo_tpls *tpls, const jo_params *p, int64_t *out)
{
    resindex R;
    if (resindex_build(mol, &R) != JO_OK) { resindex_free(&R); return JO_ENOMEM; }
    scanner S;
    int32_t rc = scanner_init(&S, max_template_atoms(tpls), mol->n, R.nres);
    S.w_inbox = 0;
    for (int32_t t = 0; rc == JO_OK && t < tpls->n_templates; t++)
        rc = scan_template(&S, mol, &R, tpls, t, p, sink_none, NULL);
    out[0] = S.w_inbox;
    out[1] = S.w_levels;
    out[2] = S.w_complete;
    out[3] = S.w_complete_atoms;
    scanner_free(&S);
    resindex_free(&R);
    return rc;
}

/* ------------------------------------------------------------------ */
/* candidates / brute force (test helpers)                             */
/* ------------------------------------------------------------------ */

int32_t jo_candidates(const jo_mol *mol, const jo_tpls *tpls, int32_t t, int32_t k,
                      int32_t *out, int32_t cap)
{
    resindex R;
    if (resindex_build(mol, &R) != JO_OK) { resindex_free(&R); return JO_ENOMEM; }
    int32_t *buf = (int32_t *)malloc(sizeof(int32_t) * (mol->n > 0 ? mol->n : 1));
    char *done = (char *)calloc(R.nres > 0 ? R.nres : 1, 1);
    int32_t c = JO_ENOMEM;
    if (buf && done) {
        c = candidates(mol, &R, tpls, tpls->atom_off[t] + k, buf, done);
        for (int32_t i = 0; i < c && i < cap; i++) out[i] = buf[i];
    }
    free(buf);
    free(done);
    resindex_free(&R);
    return c;
}

/* Enumerates the closed form of a complete match directly: one candidate per
 * template atom, every pair (j < L in template order) within its annulus window
 * and check() passing at every level, without kd-trees or boxes. */
int64_t jo_bruteforce(const jo_mol *mol, const jo_tpls *tpls, int32_t t, const jo_params *p,
                      int32_t *out, int64_t cap)
{
    int32_t base = tpls->atom_off[t], n = tpls->atom_off[t + 1] - base;
    if (n <= 0) return 0;
    resindex R;
    if (resindex_build(mol, &R) != JO_OK) { resindex_free(&R); return JO_ENOMEM; }
    int32_t **set = (int32_t **)calloc(n, sizeof(int32_t *));
    int32_t *cnt = (int32_t *)calloc(n, sizeof(int32_t));
    int32_t *cur = (int32_t *)calloc(n, sizeof(int32_t));
    int32_t *atom = (int32_t *)calloc(n, sizeof(int32_t));
    int32_t *ord = (int32_t *)calloc(n, sizeof(int32_t));
    char *done = (char *)calloc(R.nres > 0 ? R.nres : 1, 1);
    int64_t found = 0;
    for (int32_t k = 0; k < n; k++) {
        set[k] = (int32_t *)malloc(sizeof(int32_t) * (mol->n > 0 ? mol->n : 1));
        cnt[k] = candidates(mol, &R, tpls, base + k, set[k], done);
        if (cnt[k] < 0) { found = cnt[k]; goto done; }
        ord[k] = k;
    }
    for (int32_t k = 0; k < n; k++)
        if (cnt[k] == 0) goto done;
    /* odometer over all tuples, template order */
    for (;;) {
        int ok = 1;
        for (int32_t k = 0; k < n; k++) atom[k] = set[k][cur[k]];
        for (int32_t L = 1; L < n && ok; L++) {
            for (int32_t j = 0; j < L && ok; j++) {
                double lo, hi;
                pair_window(tpls, base, j, L, p, &lo, &hi);
                annulus A;
                annulus_set(&A, mol->xyz + 3 * atom[j], lo, hi);
                ok = annulus_has(&A, mol->xyz + 3 * atom[L]);
            }
            if (ok) ok = check_level(tpls, base, mol, atom, ord, L, p->ignore_chain);
        }
        if (ok) {
            if (found < cap) memcpy(out + found * n, atom, sizeof(int32_t) * n);
            found++;
        }
        int32_t k = n - 1;
        while (k >= 0 && ++cur[k] >= cnt[k]) cur[k--] = 0;
        if (k < 0) break;
    }
done:
    for (int32_t k = 0; k < n; k++) free(set[k]);
    free(set);
    free(cnt);
    free(cur);
    free(atom);
    free(ord);
    free(done);
    resindex_free(&R);
    return found;
}

/* ------------------------------------------------------------------ */
/* multi-threaded batch (CPU baseline)                                 */
/* ------------------------------------------------------------------ */

static inline uint64_t mix64(uint64_t x)
{
    x += 0x9e3779b97f4a7c15ULL;
    x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ULL;
    x = (x ^ (x >> 27)) * 0x94d049bb133111ebULL;
    return x ^ (x >> 31);
}

/* Order-independent digest term of one hit: (template, molecule, atoms...). */
static uint64_t hit_digest(int32_t t, int32_t mi, const int32_t *atoms, int32_t n)
{
    uint64_t h = mix64(((uint64_t)(uint32_t)t << 32) ^ (uint32_t)mi);
    for (int32_t i = 0; i < n; i++) h = mix64(h ^ (uint32_t)atoms[i]);
    return h;
}

typedef struct {
    double thr;
    int32_t mol_index;
    int64_t cm, hits;
    uint64_t digest;
} count_ctx;

static int sink_count(void *vctx, scanner *S, const jo_mol *m, const jo_tpls *T, int32_t t, int32_t n)
{
    count_ctx *c = (count_ctx *)vctx;
    int32_t base = T->atom_off[t];
    for (int32_t i = 0; i < n; i++)
        for (int d = 0; d < 3; d++) {
            S->pa[3 * i + d] = m->xyz[3 * S->atom[i] + d];
            S->pb[3 * i + d] = T->pos[3 * (base + i) + d];
        }
    double rmsd, rot[9], c0[3], c1[3];
    jo_superpose(S->pa, S->pb, n, &rmsd, rot, c0, c1);
    c->cm++;
    if (rmsd <= c->thr) {
        int nan = 0;
        for (int i = 0; i < 9; i++) nan |= isnan(rot[i]);
        if (!nan) {
            c->hits++;
            c->digest += hit_digest(t, c->mol_index, S->atom, n);
        }
    }
    return 0;
}

typedef struct {
    const jo_mol *mols;
    int32_t n_mols;
    const jo_tpls *tpls;
    const jo_params *p;
    double thr;
    int64_t *cm, *hits;
    uint64_t *digest;
    resindex *R;           /* one residue index per molecule, built up front */
    int32_t block;         /* templates per tile */
    int32_t n_tiles;
    int32_t next;          /* next tile to take */
    pthread_mutex_t lock;
    int32_t rc;
} batch_ctx;

/* Tiles are (molecule, block of templates), taken from a shared counter: the
 * ThreadPool-over-molecules pattern of the reference README, split finer so
 * that a single large molecule still spreads over all workers. */
static void *batch_worker(void *arg)
{
    batch_ctx *B = (batch_ctx *)arg;
    int32_t T = B->tpls->n_templates;
    int32_t nblk = (T + B->block - 1) / B->block;
    for (;;) {
        pthread_mutex_lock(&B->lock);
        int32_t tile = B->next++;
        pthread_mutex_unlock(&B->lock);
        if (tile >= B->n_tiles) break;
        int32_t mi = tile / nblk, t0 = (tile % nblk) * B->block;
        int32_t t1 = t0 + B->block < T ? t0 + B->block : T;
        const jo_mol *m = &B->mols[mi];
        scanner S;
        int rc = scanner_init(&S, max_template_atoms(B->tpls), m->n, B->R[mi].nres);
        for (int32_t t = t0; rc == JO_OK && t < t1; t++) {
            count_ctx c = {B->thr, mi, 0, 0, 0};
            rc = scan_template(&S, m, &B->R[mi], B->tpls, t, B->p, sink_count, &c);
            int64_t pi = (int64_t)mi * T + t;
            B->cm[pi] = c.cm;
            B->hits[pi] = c.hits;
            B->digest[pi] = c.digest;
        }
        if (rc != JO_OK) B->rc = rc;
        scanner_free(&S);
    }
    return NULL;
}

int32_t jo_batch(const jo_mol *mols, int32_t n_mols, const jo_tpls *tpls, const jo_params *p,
                 double rmsd_threshold, int32_t n_threads,
                 int64_t *cm_count, int64_t *hit_count, uint64_t *digest)
{
    if (n_threads < 1) n_threads = 1;
    int32_t T = tpls->n_templates;
    if (n_mols <= 0 || T <= 0) return JO_OK;
    resindex *R = (resindex *)calloc(n_mols, sizeof(resindex));
    if (!R) return JO_ENOMEM;
    int32_t rc = JO_OK;
    for (int32_t i = 0; i < n_mols && rc == JO_OK; i++) rc = resindex_build(&mols[i], &R[i]);
    int32_t want = 8 * n_threads;                       /* tiles wanted */
    int32_t per_mol = (want + n_mols - 1) / n_mols;     /* blocks per molecule */
    int32_t block = (T + per_mol - 1) / per_mol;
    if (block < 1) block = 1;
    int32_t nblk = (T + block - 1) / block;
    batch_ctx B = {mols, n_mols, tpls, p, rmsd_threshold, cm_count, hit_count, digest, R, block,
                   n_mols * nblk, 0, PTHREAD_MUTEX_INITIALIZER, rc};
    if (rc == JO_OK) {
        pthread_t *th = (pthread_t *)malloc(sizeof(pthread_t) * n_threads);
        if (!th) {
            B.rc = JO_ENOMEM;
        } else {
            for (int32_t i = 0; i < n_threads; i++) pthread_create(&th[i], NULL, batch_worker, &B);
            for (int32_t i = 0; i < n_threads; i++) pthread_join(th[i], NULL);
            free(th);
        }
    }
    for (int32_t i = 0; i < n_mols; i++) resindex_free(&R[i]);
    free(R);
    return B.rc;
}
