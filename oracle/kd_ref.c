/*
 * kd_ref.c -- TEST INFRASTRUCTURE ONLY: the kd-tree leaf order computed with
 * the reference's own QuickSelect.
 *
 * `qselect.h` is not part of this repository: `make ref` (oracle/Makefile)
 * applies the reference's patches/qselect.h.patch (a new file; it needs only
 * <stddef.h> and <stdint.h>) into oracle/_ref/src/ and compiles this file
 * against it into oracle/_ref/libkd_ref.so. The leaf orders it produces are
 * written as golden vectors by tests/golden/make_kd_ref_vectors.py; nothing
 * else loads the library.
 *
 * What this file restates (the reference files are patches of an absent
 * KdTree.c, so the surrounding code cannot be compiled):
 *   - the comparison KdTree_compare_r (patches/KdTree.c.patch:320-329):
 *     coordinate `index` of the two points, -1 / 1 / 0;
 *   - the recursion of KdTreeNode_create (patches/KdTree.c.patch:348-420):
 *     a range of one index is a leaf; otherwise qselect_r(idx, n, sizeof(int),
 *     n/2, ...) on axis `type`, then the left child over idx[0, split) and the
 *     right child over idx[split, n), both on axis (type + 1) % 3, starting
 *     from idx = 0 .. n-1 on axis 0 (KdTree_reuse, patches/KdTree.c.patch:160-185).
 * The leaf order (left subtree first) is the index array after the recursion.
 */
#include <stdint.h>
#include <stdlib.h>

#include "qselect.h"

typedef struct {
    const double **u;
    int index;
} kdref_cmp_data;

static int kdref_compare(const void *pa, const void *pb, void *data)
{
    const kdref_cmp_data *d = (const kdref_cmp_data *)data;
    const int a = *(const int *)pa, b = *(const int *)pb;
    const double c = d->u[a][d->index], e = d->u[b][d->index];
    return c < e ? -1 : (c > e ? 1 : 0);
}

static void kdref_node(int *idx, int n, int type, const double **u)
{
    if (n <= 1) return;
    kdref_cmp_data data = {u, type};
    const size_t split = qselect_r(idx, (size_t)n, sizeof(int), (size_t)(n / 2), kdref_compare, &data);
    const int next = (type + 1) % 3;
    kdref_node(idx, (int)split, next, u);
    kdref_node(idx + split, n - (int)split, next, u);
}

/* Leaf order of the kd-tree over points xyz[3*i .. 3*i+2], i < n: out[p] is
 * the point at leaf p, left subtree first. Returns n, or -1 when out of memory. */
int32_t kdref_leaf_order(const double *xyz, int32_t n, int32_t *out)
{
    if (n <= 0) return 0;
    const double **u = (const double **)malloc(sizeof(double *) * (size_t)n);
    int *idx = (int *)malloc(sizeof(int) * (size_t)n);
    if (!u || !idx) {
        free(u);
        free(idx);
        return -1;
    }
    for (int32_t i = 0; i < n; i++) {
        u[i] = xyz + 3 * (size_t)i;
        idx[i] = i;
    }
    kdref_node(idx, n, 0, u);
    for (int32_t i = 0; i < n; i++) out[i] = idx[i];
    free(u);
    free(idx);
    return n;
}
