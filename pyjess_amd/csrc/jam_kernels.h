// jam_kernels.h -- gfx950 kernels of the Jess search pipeline.
//
//   k_candidates   wave per (molecule, atom type): candidate list in the
//                  reference order (TessTemplate_candidates,
//                  patches/TessTemplate.c.patch:173-235), two passes (count,
//                  fill) with ballot compaction.
//   k_kdorder      workgroup per candidate set: the leaf order of the
//                  reference's QuickSelect kd-tree (patches/KdTree.c.patch:348-451,
//                  patches/qselect.h.patch:50-116) built in parallel in LDS,
//                  reversed because its DFS visits right children first; writes
//                  the set as SoA in that order plus one bounding box per 64
//                  entries.
//   k_pairs        thread per (template, molecule) pair: level order and the
//                  number of level-0 partial matches.
//   k_expand       wave per partial match of level L-1: the candidates of
//                  level L (patches/Scanner.c.patch:232-334) from the 64-entry
//                  blocks of the kd-ordered set whose box meets the join box,
//                  all 64 lanes testing one candidate each against every
//                  annulus and check(); a count pass, a scan, a fill pass keep
//                  the depth-first order of the reference.
//   k_super        thread per complete match: FP64 superposition, RMSD
//                  threshold, event compaction with the match's rank.
#pragma once

#include <hip/hip_runtime.h>

#include "jam_common.h"
#include "jam_super.h"

namespace jam {

__device__ static inline int lane_id() { return (int)(threadIdx.x & (JAM_WAVE - 1)); }

__device__ static inline unsigned long long ballot(bool p) { return __ballot(p); }

// number of set bits of m below this lane
__device__ static inline int rank_below(unsigned long long m)
{
    return (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

__device__ static inline int first_bit(unsigned long long m) { return __ffsll((long long)m) - 1; }

// ---------------------------------------------------------------------------
// match predicate (modes 0, 1, 3; the others are rejected on the host)
// ---------------------------------------------------------------------------

__device__ static inline uint32_t byte_of(uint32_t v, int i) { return (v >> (8 * i)) & 0xFFu; }

__device__ static inline uint32_t up1(uint32_t c) { return (c >= 'a' && c <= 'z') ? (c & ~0x20u) : c; }

__device__ static inline bool type_match_name(const DevTpls &tp, int u, uint32_t nm)
{
    const int mode = tp.type_mode[u];
    const uint32_t unm = jam_upper4(nm);
    const bool carbon = byte_of(nm, 0) == '_' && byte_of(nm, 1) == 'C';
    for (int k = tp.type_name_off[u]; k < tp.type_name_off[u + 1]; k++) {
        const uint32_t tn = tp.type_names[k];
        if (mode == 0) {
            if (unm == jam_upper4(tn)) return true;
        } else if (mode == 1) {
            if (up1(byte_of(nm, 2)) == up1(byte_of(tn, 2)) && !carbon) return true;
        } else if (mode == 3) {
            if (up1(byte_of(nm, 0)) == up1(byte_of(tn, 0)) && up1(byte_of(nm, 1)) == up1(byte_of(tn, 1)))
                return true;
        }
    }
    return false;
}

// bucket of an upper-cased residue name among [b0, b1), or -1
__device__ static inline int find_bucket(const DevMols &ms, int b0, int b1, uint32_t key)
{
    int lo = b0, hi = b1;
    while (lo < hi) {
        int mid = (lo + hi) >> 1;
        uint32_t k = ms.res_key[mid];
        if (k < key) lo = mid + 1;
        else if (k > key) hi = mid;
        else return mid;
    }
    return -1;
}

// ---------------------------------------------------------------------------
// candidates
// ---------------------------------------------------------------------------

template <bool FILL>
__global__ void __launch_bounds__(256) k_candidates(DevMols ms, DevTpls tp, DevSets st)
{
    const int wave = (int)((blockIdx.x * blockDim.x + threadIdx.x) / JAM_WAVE);
    const int lane = lane_id();
    if (wave >= st.n_sets) return;
    const int U = tp.n_types;
    const int m = wave / U, u = wave % U;
    const int abase = ms.atom_off[m];
    const int b0 = ms.res_off[m], b1 = ms.res_off[m + 1];
    int64_t out = 0;
    if (FILL) {
        out = st.off[wave];
        if (*st.total > st.cap) return;   // sets do not fit: the host grows and re-runs
    }
    int count = 0;
    const int r0 = tp.type_rn_off[u], r1 = tp.type_rn_off[u + 1];
    for (int r = r0; r < r1; r++) {
        const int b = find_bucket(ms, b0, b1, tp.type_rns[r]);
        if (b < 0) continue;
        bool seen = false;   // a residue name listed twice walks its bucket once
        for (int q = r0; q < r && !seen; q++) seen = find_bucket(ms, b0, b1, tp.type_rns[q]) == b;
        if (seen) continue;
        const int e = ms.res_end[b];
        for (int q = ms.res_begin[b]; q < e; q += JAM_WAVE) {
            const int pos = q + lane;
            int atom = -1;
            bool ok = false;
            if (pos < e) {
                atom = ms.perm[abase + pos];
                ok = type_match_name(tp, u, ms.name[abase + atom]);
            }
            const unsigned long long mask = ballot(ok);
            if (FILL && ok) st.list[out + rank_below(mask)] = atom;
            const int c = __popcll(mask);
            out += c;
            count += c;
        }
    }
    if (!FILL && lane == 0) st.cnt[wave] = count;
}

// ---------------------------------------------------------------------------
// kd leaf order
// ---------------------------------------------------------------------------

// ---- exact parallel arrangement ------------------------------------------
//
// The reference builds the tree by QuickSelect of the median of every range
// (KdTreeNode_create with qselect_r, patches/KdTree.c.patch:348-451,
// patches/qselect.h.patch:50-116): median-of-three pivot at
// mid = (l + r + 1) / 2, moved to r, Lomuto partition sending "<= pivot" to
// the left, pivot swapped into place, repeat on the side holding the median.
// Where equal coordinates land depends on that exact swap sequence, and it
// decides the leaf order, so the kernel reproduces it, in parallel: all ranges
// of one tree level run their QuickSelect together, one partition round at a
// time, and each Lomuto partition is computed without its serial scan:
//   * the "<= pivot" elements keep their order: a stable compaction;
//   * the others move as a FIFO queue: each is pushed when scanned, and each
//     "<=" element scanned after the first ">" one pops the front and pushes
//     it back; push number t of such a re-push is the q-th pop, i.e. push
//     number q < t. Pointer jumping resolves every push to its element. The
//     final queue fills [l+k+1, r-1] from its second element on, its front
//     goes to r, the pivot to l+k.

// Two workgroup shapes: one wave for sets of up to KD_SMALL entries (LDS
// 15 KB, so many sets per CU), 256 threads for larger ones (LDS up to
// KD_CAP entries, a global scratch area beyond).
#define KD_SMALL 256
#define KD_CAP 1024
#define KD_THREADS 256
// global scratch of sets too large for LDS, per set entry (padded to a power
// of two for the fast path's sorts: 3 doubles, 9 ints and 1 byte per padded
// entry, at most 2 * 61 bytes per entry)
#define KD_SCRATCH_BYTES_PER_ENTRY 128

struct KdWork {
    double *c;               // coordinates: axis a of element e at c[a * stride + e]
    int *arr;                // arrangement, double buffered at arr and arr + stride
    int *pre1, *pre2;        // exclusive prefixes: "<=" flags, re-push flags
    int *src;                // push -> element (>= 0) or -(push index) - 1
    int *rlo, *rhi;          // range of each position at the current level
    int *wl, *wr;            // QuickSelect window of the range starting here; wr < 0: done
    int stride;
};

// in-place exclusive block scan of v[0, n)
template <int NT>
__device__ __attribute__((always_inline)) static inline void kd_scan(int *v, int n, int *partial)
{
    const int tid = threadIdx.x;
    const int per = (n + NT - 1) / NT;
    const int b0 = tid * per, b1 = min(n, b0 + per);
    int cnt = 0;
    for (int p = b0; p < b1; p++) cnt += v[p];
    partial[tid] = cnt;
    __syncthreads();
    if (tid < JAM_WAVE) {
        int x[NT / JAM_WAVE], sum = 0;
#pragma unroll
        for (int q = 0; q < NT / JAM_WAVE; q++) {
            x[q] = partial[(NT / JAM_WAVE) * tid + q];
            sum += x[q];
        }
        int incl = sum;
        for (int sh = 1; sh < JAM_WAVE; sh <<= 1) {
            int o = __shfl_up(incl, sh, JAM_WAVE);
            if (tid >= sh) incl += o;
        }
        int run = incl - sum;
#pragma unroll
        for (int q = 0; q < NT / JAM_WAVE; q++) {
            partial[(NT / JAM_WAVE) * tid + q] = run;
            run += x[q];
        }
    }
    __syncthreads();
    int run = partial[tid];
    for (int p = b0; p < b1; p++) {
        const int x = v[p];
        v[p] = run;
        run += x;
    }
    __syncthreads();
}

__device__ __attribute__((always_inline)) static inline int kd_cmp3(const double *c, int x, int y)
{
    const double u = c[x], v = c[y];
    return u < v ? -1 : (u > v ? 1 : 0);
}

// Arranges elements 0..n-1 into the reference's leaf order, written to out.
template <int NT>
__device__ __attribute__((always_inline)) static inline void kd_exact(const KdWork &w, int n, int *out, int *partial, int *flag)
{
    const int tid = threadIdx.x;
    int cur = 0;
    for (int p = tid; p < n; p += NT) {
        w.arr[p] = p;
        w.rlo[p] = 0;
        w.rhi[p] = n;
    }
    __syncthreads();
    for (int d = 0;; d++) {
        const double *c = w.c + (d % 3) * w.stride;
        if (tid == 0) flag[0] = 0;
        __syncthreads();
        for (int p = tid; p < n; p += NT) {
            if (w.rlo[p] != p) continue;
            const int hi = w.rhi[p];
            if (hi - p >= 2) {
                w.wl[p] = p;
                w.wr[p] = hi - 1;
                flag[0] = 1;
            } else {
                w.wr[p] = -1;
            }
        }
        __syncthreads();
        if (!flag[0]) break;
        for (;;) {
            int *a = w.arr + cur * w.stride;
            int *an = w.arr + (cur ^ 1) * w.stride;
            // pivot: median of three, moved to the right end of the window
            for (int p = tid; p < n; p += NT) {
                if (w.rlo[p] != p || w.wr[p] < 0) continue;
                const int l = w.wl[p], r = w.wr[p], mid = (l + r + 1) / 2;
                int piv;
                if ((kd_cmp3(c, a[l], a[mid]) > 0) != (kd_cmp3(c, a[l], a[r]) > 0)) piv = l;
                else if ((kd_cmp3(c, a[mid], a[l]) < 0) != (kd_cmp3(c, a[mid], a[r]) < 0)) piv = mid;
                else piv = r;
                const int t = a[r];
                a[r] = a[piv];
                a[piv] = t;
            }
            __syncthreads();
            // "<= pivot" flags of the windows
            for (int i = tid; i < n; i += NT) {
                const int R = w.rlo[i];
                const int r = w.wr[R];
                int b = 0;
                if (r >= 0 && i >= w.wl[R] && i < r) b = c[a[i]] <= c[a[r]] ? 1 : 0;
                w.pre1[i] = b;
            }
            __syncthreads();
            kd_scan<NT>(w.pre1, n, partial);
            // re-push flags: a "<=" element after the first ">" one
            for (int i = tid; i < n; i += NT) {
                const int R = w.rlo[i];
                const int l = w.wl[R], r = w.wr[R];
                int rot = 0;
                if (r >= 0 && i >= l && i < r && c[a[i]] <= c[a[r]]) {
                    const int zeros = (i - l) - (w.pre1[i] - w.pre1[l]);
                    rot = zeros > 0 ? 1 : 0;
                }
                w.pre2[i] = rot;
            }
            __syncthreads();
            kd_scan<NT>(w.pre2, n, partial);
            // pushes: ">" elements are new pushes, re-pushes point to push q
            for (int i = tid; i < n; i += NT) {
                const int R = w.rlo[i];
                const int l = w.wl[R], r = w.wr[R];
                if (r < 0 || i < l || i >= r) continue;
                const int ones = w.pre1[i] - w.pre1[l];
                const int zeros = (i - l) - ones;
                const int rots = w.pre2[i] - w.pre2[l];
                const int slot = l + zeros + rots;      // push number - 1, absolute
                if (!(c[a[i]] <= c[a[r]])) w.src[slot] = a[i];
                else if (zeros > 0) w.src[slot] = -(l + rots) - 1;
            }
            __syncthreads();
            // pointer jumping until every push names an element
            for (;;) {
                if (tid == 0) flag[1] = 0;
                __syncthreads();
                for (int i = tid; i < n; i += NT) {
                    const int R = w.rlo[i];
                    const int l = w.wl[R], r = w.wr[R];
                    if (r < 0 || i < l || i >= r) continue;
                    // only the m + nrot pushes of this round are valid; each
                    // re-push points to a lower push, so the chains end
                    const int pushes = (r - l) - (w.pre1[r] - w.pre1[l]) + (w.pre2[r] - w.pre2[l]);
                    if (i - l >= pushes) continue;
                    const int v = w.src[i];
                    if (v < 0) {
                        const int u = w.src[-v - 1];
                        w.src[i] = u;
                        if (u < 0) flag[1] = 1;
                    }
                }
                __syncthreads();
                const int again = flag[1];
                __syncthreads();
                if (!again) break;
            }
            // scatter into the other buffer
            for (int i = tid; i < n; i += NT) {
                const int R = w.rlo[i];
                const int l = w.wl[R], r = w.wr[R];
                if (r < 0 || i < l || i > r) {
                    an[i] = a[i];
                    continue;
                }
                const int k1 = w.pre1[r] - w.pre1[l];
                const int m = (r - l) - k1;
                const int nrot = w.pre2[r] - w.pre2[l];
                if (i == r) {
                    an[l + k1] = a[r];                  // the pivot
                    continue;
                }
                if (c[a[i]] <= c[a[r]]) an[l + (w.pre1[i] - w.pre1[l])] = a[i];
                const int j = i - l;                    // queue slot j+1
                if (j < m) an[j == 0 ? r : l + k1 + j] = w.src[l + nrot + j];
            }
            __syncthreads();
            cur ^= 1;
            // next window of each QuickSelect
            if (tid == 0) flag[1] = 0;
            __syncthreads();
            for (int p = tid; p < n; p += NT) {
                if (w.rlo[p] != p || w.wr[p] < 0) continue;
                const int l = w.wl[p], r = w.wr[p];
                const int pos = l + (w.pre1[r] - w.pre1[l]);
                const int k = p + (w.rhi[p] - p) / 2;
                int nl = l, nr = r;
                if (pos == k) nr = -1;
                else if (k < pos) nr = pos - 1;
                else nl = pos + 1;
                if (nr >= 0 && nl == nr) nr = -1;
                w.wl[p] = nl;
                w.wr[p] = nr;
                if (nr >= 0) flag[1] = 1;
            }
            __syncthreads();
            const int more = flag[1];
            __syncthreads();
            if (!more) break;
        }
        // split every range at its median position
        for (int p = tid; p < n; p += NT) {
            const int lo = w.rlo[p], hi = w.rhi[p];
            if (hi - lo < 2) continue;
            const int m = lo + (hi - lo) / 2;
            if (p < m) w.rhi[p] = m;
            else w.rlo[p] = m;
        }
        __syncthreads();
    }
    const int *fin = w.arr + cur * w.stride;
    for (int p = tid; p < n; p += NT) out[p] = fin[p];
    __syncthreads();
}

// Writes the kd arrays of set s from its arrangement `order` (element ids =
// list positions) and the block boxes: position p of the DFS visit order is
// element n-1-p of the arrangement (right subtrees first).
template <int NT>
__device__ static void kd_emit(const DevMols &ms, const DevSets &st, int s, int n, int64_t off, int abase,
                               const int *order)
{
    const int tid = threadIdx.x;
    const int64_t bbase = off / JAM_WAVE + s;
    const int lane = lane_id();
    const int wave = tid / JAM_WAVE;
    for (int b0 = wave * JAM_WAVE; b0 < n; b0 += NT) {
        const int p = b0 + lane;
        double x = 0.0, y = 0.0, z = 0.0;
        const bool ok = p < n;
        if (ok) {
            const int li = order[n - 1 - p];
            const int atom = st.list[off + li];
            const size_t g = (size_t)(abase + atom);
            x = ms.xyz[3 * g + 0];
            y = ms.xyz[3 * g + 1];
            z = ms.xyz[3 * g + 2];
            st.kx[off + p] = x;
            st.ky[off + p] = y;
            st.kz[off + p] = z;
            st.katom[off + p] = atom;
            st.kres[off + p] = ms.resseq[g];
            st.kch[off + p] = ms.chain[g];
        }
        double blo[3] = {ok ? x : INFINITY, ok ? y : INFINITY, ok ? z : INFINITY};
        double bhi[3] = {ok ? x : -INFINITY, ok ? y : -INFINITY, ok ? z : -INFINITY};
        for (int sh = 32; sh >= 1; sh >>= 1) {
#pragma unroll
            for (int d = 0; d < 3; d++) {
                blo[d] = fmin(blo[d], __shfl_xor(blo[d], sh, JAM_WAVE));
                bhi[d] = fmax(bhi[d], __shfl_xor(bhi[d], sh, JAM_WAVE));
            }
        }
        if (lane == 0) {
            double *bb = st.bbox + 6 * (bbase + b0 / JAM_WAVE);
            bb[0] = blo[0];
            bb[1] = blo[1];
            bb[2] = blo[2];
            bb[3] = bhi[0];
            bb[4] = bhi[1];
            bb[5] = bhi[2];
        }
    }
}

// ---- fast path -------------------------------------------------------------
//
// When no range of the tree has equal coordinates on both sides of its split
// point, the left half of every range is its n/2 smallest elements whatever
// the arrangement, and the leaf order follows from memberships alone: sort
// every axis once, then at level d cut each range of axis d%3 in half and
// stably partition the two other axes by side. A tie across a split point
// sends the set to the exact path (kd_exact).

__device__ static inline int shfl_int(int v, int src) { return __shfl(v, src, JAM_WAVE); }

// value of lane `dst` receives v from this lane (ds_permute: push)
__device__ static inline int push_int(int v, int dst)
{
    return __builtin_amdgcn_ds_permute(dst << 2, v);
}

__device__ static inline bool key_less(double a, int ia, double b, int ib)
{
    return a < b || (a == b && ia < ib);
}

// Sets of at most 64 entries: one element per lane, everything in registers.
// Returns false (nothing written) when a tie straddles a split point.
__device__ static bool kd_wave64(const DevMols &ms, const DevSets &st, int s, int n, int64_t off, int abase)
{
    const int lane = lane_id();
    double c[3] = {INFINITY, INFINITY, INFINITY};
    if (lane < n) {
        const size_t g = (size_t)(abase + st.list[off + lane]);
        c[0] = ms.xyz[3 * g + 0];
        c[1] = ms.xyz[3 * g + 1];
        c[2] = ms.xyz[3 * g + 2];
    }
    // per axis: element at each position, sorted by (coordinate, element)
    int perm[3];
    for (int a = 0; a < 3; a++) {
        double key = c[a];
        int id = lane;
        for (int k = 2; k <= JAM_WAVE; k <<= 1) {
            for (int j = k >> 1; j > 0; j >>= 1) {
                const double ok = __shfl_xor(key, j, JAM_WAVE);
                const int oi = __shfl_xor(id, j, JAM_WAVE);
                const bool up = (lane & k) == 0;
                const bool lower = (lane & j) == 0;
                const bool other_less = key_less(ok, oi, key, id);
                // the lower lane of an ascending pair keeps the smaller key
                const bool take = (lower == up) ? other_less : !other_less && !(ok == key && oi == id);
                if (take) {
                    key = ok;
                    id = oi;
                }
            }
        }
        perm[a] = id;
    }
    int sb = 0, se = n;
    for (int d = 0;; d++) {
        const int a = d % 3;
        const bool act = lane < n && se - sb >= 2;
        const int m = sb + (se - sb) / 2;
        const bool split = act && lane == m;
        if (ballot(act) == 0ull) break;
        // tie across the split point: compare with the previous position
        const int pa = (a == 0) ? perm[0] : (a == 1) ? perm[1] : perm[2];
        const double ca = (a == 0) ? c[0] : (a == 1) ? c[1] : c[2];
        const double here = __shfl(ca, pa & 63, JAM_WAVE);
        const double prev = __shfl_up(here, 1, JAM_WAVE);
        if (ballot(split && prev == here) != 0ull) return false;
        // side of every element, delivered to the lane of the element id
        const int left_pos = act && lane < m ? 1 : 0;
        const int side_by_elem = push_int(left_pos, pa & 63);
        for (int b = 0; b < 3; b++) {
            if (b == a) continue;
            const int pb = (b == 0) ? perm[0] : (b == 1) ? perm[1] : perm[2];
            const int sd = shfl_int(side_by_elem, pb & 63);
            const unsigned long long lm = ballot(act && sd);
            int np = lane;
            if (act) {
                const unsigned long long below = (1ull << lane) - (1ull << sb);
                const int lc = __popcll(lm & below);
                np = sd ? sb + lc : m + (lane - sb - lc);
            }
            const int moved = push_int(pb, np);
            if (b == 0) perm[0] = lane < n ? moved : perm[0];
            else if (b == 1) perm[1] = lane < n ? moved : perm[1];
            else perm[2] = lane < n ? moved : perm[2];
        }
        if (act) {
            if (lane < m) se = m;
            else sb = m;
        }
    }
    // DFS order: position p holds element perm[n-1-p]
    const int src = n - 1 - lane;
    const int e = shfl_int(perm[0], src < 0 ? 0 : src);
    const bool ok = lane < n;
    double x = __shfl(c[0], e & 63, JAM_WAVE);
    double y = __shfl(c[1], e & 63, JAM_WAVE);
    double z = __shfl(c[2], e & 63, JAM_WAVE);
    if (ok) {
        const int atom = st.list[off + e];
        const size_t g = (size_t)(abase + atom);
        st.kx[off + lane] = x;
        st.ky[off + lane] = y;
        st.kz[off + lane] = z;
        st.katom[off + lane] = atom;
        st.kres[off + lane] = ms.resseq[g];
        st.kch[off + lane] = ms.chain[g];
    }
    double blo[3] = {ok ? x : INFINITY, ok ? y : INFINITY, ok ? z : INFINITY};
    double bhi[3] = {ok ? x : -INFINITY, ok ? y : -INFINITY, ok ? z : -INFINITY};
    for (int sh = 32; sh >= 1; sh >>= 1) {
#pragma unroll
        for (int q = 0; q < 3; q++) {
            blo[q] = fmin(blo[q], __shfl_xor(blo[q], sh, JAM_WAVE));
            bhi[q] = fmax(bhi[q], __shfl_xor(bhi[q], sh, JAM_WAVE));
        }
    }
    if (lane == 0) {
        double *bb = st.bbox + 6 * (off / JAM_WAVE + s);
        bb[0] = blo[0];
        bb[1] = blo[1];
        bb[2] = blo[2];
        bb[3] = bhi[0];
        bb[4] = bhi[1];
        bb[5] = bhi[2];
    }
    return true;
}

// Fast path of larger sets in LDS (or global scratch): same algorithm with
// bitonic sorts and block prefix sums. Returns false on a split tie.
struct KdFast {
    double *c;               // coordinates, axis-major, stride entries per axis
    int *perm;               // 3 axes x 2 buffers, stride entries each
    int *pre;                // prefix of left sides
    int *sb, *se;            // range of each position
    unsigned char *side;     // by element
    int stride;
};

template <int NT>
__device__ __attribute__((always_inline)) static inline void kdf_sort(const double *c, int *perm, int n, int n2)
{
    const int tid = threadIdx.x;
    for (int i = tid; i < n2; i += NT) perm[i] = i;
    __syncthreads();
    for (int k = 2; k <= n2; k <<= 1) {
        for (int j = k >> 1; j > 0; j >>= 1) {
            for (int i = tid; i < n2; i += NT) {
                const int ixj = i ^ j;
                if (ixj > i) {
                    const int x = perm[i], y = perm[ixj];
                    const double kx = x < n ? c[x] : INFINITY, ky = y < n ? c[y] : INFINITY;
                    const bool y_less = key_less(ky, y, kx, x);
                    const bool up = (i & k) == 0;
                    if (y_less == up) {
                        perm[i] = y;
                        perm[ixj] = x;
                    }
                }
            }
            __syncthreads();
        }
    }
}

template <int NT>
__device__ __attribute__((always_inline)) static inline bool kd_fast(const KdFast &w, int n, int *out, int *partial, int *flag)
{
    const int tid = threadIdx.x;
    int n2 = 1;
    while (n2 < n) n2 <<= 1;
    const int S = w.stride;
    for (int a = 0; a < 3; a++) kdf_sort<NT>(w.c + a * S, w.perm + 2 * a * S, n, n2);
    for (int p = tid; p < n; p += NT) {
        w.sb[p] = 0;
        w.se[p] = n;
    }
    int cur[3] = {0, 0, 0};
    __syncthreads();
    for (int d = 0;; d++) {
        const int a = d % 3;
        if (tid == 0) {
            flag[0] = 0;
            flag[1] = 0;
        }
        __syncthreads();
        const int *pa = w.perm + (2 * a + cur[a]) * S;
        const double *ca = w.c + a * S;
        for (int p = tid; p < n; p += NT) {
            const int sb = w.sb[p], se = w.se[p];
            if (se - sb < 2) continue;
            const int m = sb + (se - sb) / 2;
            w.side[pa[p]] = p < m ? 1 : 0;
            if (p == m) {
                flag[0] = 1;
                if (ca[pa[m - 1]] == ca[pa[m]]) flag[1] = 1;
            }
        }
        __syncthreads();
        const int splits = flag[0], tie = flag[1];
        __syncthreads();
        if (tie) return false;
        if (!splits) break;
        for (int b = 0; b < 3; b++) {
            if (b == a) continue;
            const int *src = w.perm + (2 * b + cur[b]) * S;
            int *dst = w.perm + (2 * b + (cur[b] ^ 1)) * S;
            for (int p = tid; p < n; p += NT) w.pre[p] = w.side[src[p]];
            __syncthreads();
            kd_scan<NT>(w.pre, n, partial);
            for (int p = tid; p < n; p += NT) {
                const int sb = w.sb[p], se = w.se[p];
                int np = p;
                if (se - sb >= 2) {
                    const int m = sb + (se - sb) / 2;
                    const int lc = w.pre[p] - w.pre[sb];
                    np = w.side[src[p]] ? sb + lc : m + (p - sb - lc);
                }
                dst[np] = src[p];
            }
            __syncthreads();
            cur[b] ^= 1;
        }
        for (int p = tid; p < n; p += NT) {
            const int sb = w.sb[p], se = w.se[p];
            if (se - sb < 2) continue;
            const int m = sb + (se - sb) / 2;
            if (p < m) w.se[p] = m;
            else w.sb[p] = m;
        }
        __syncthreads();
    }
    const int *fin = w.perm + cur[0] * S;
    for (int p = tid; p < n; p += NT) out[p] = fin[p];
    __syncthreads();
    return true;
}

// Fast path of sets of (NT/2, NT] entries (NT = 128 ... 1024), one position
// per thread: the same algorithm as kd_wave64 with the arrangement of every
// axis held in registers (position t: key and element), bitonic stages below
// 64 lanes apart by shuffles and above through LDS, side flags and the
// stable partition's prefix counts by ballot and one LDS word per wave.
// Returns false (nothing written) on a split tie.
template <int NT>
__global__ void __launch_bounds__(NT) k_kdblock(DevMols ms, DevTpls tp, DevSets st, int *tied,
                                                unsigned int *n_tied)
{
    constexpr int NW = NT / JAM_WAVE;
    __shared__ double xkey[NT];
    __shared__ int xid[NT];
    __shared__ int xpre[NT];
    __shared__ int side[NT];
    __shared__ int wsum[NW];
    const int t = threadIdx.x;
    const int lane = lane_id();
    const int wv = t / JAM_WAVE;
    if (*st.total > st.cap) return;
    for (int s = blockIdx.x; s < st.n_sets; s += gridDim.x) {
        const int n = st.cnt[s];
        if (n <= NT / 2 || n > NT) continue;
        const int64_t off = st.off[s];
        const int abase = ms.atom_off[s / tp.n_types];
        double c[3] = {INFINITY, INFINITY, INFINITY};
        if (t < n) {
            const size_t g = (size_t)(abase + st.list[off + t]);
            c[0] = ms.xyz[3 * g + 0];
            c[1] = ms.xyz[3 * g + 1];
            c[2] = ms.xyz[3 * g + 2];
        }
        // per axis: (key, element) at position t, sorted by (key, element)
        double key[3];
        int id[3];
#pragma unroll
        for (int a = 0; a < 3; a++) {
            double k0 = c[a];
            int i0 = t;
            for (int k = 2; k <= NT; k <<= 1) {
                for (int j = k >> 1; j > 0; j >>= 1) {
                    double ok;
                    int oi;
                    if (j >= JAM_WAVE) {
                        xkey[t] = k0;
                        xid[t] = i0;
                        __syncthreads();
                        ok = xkey[t ^ j];
                        oi = xid[t ^ j];
                        __syncthreads();
                    } else {
                        ok = __shfl_xor(k0, j, JAM_WAVE);
                        oi = __shfl_xor(i0, j, JAM_WAVE);
                    }
                    const bool up = (t & k) == 0;
                    const bool lower = (t & j) == 0;
                    const bool other_less = key_less(ok, oi, k0, i0);
                    const bool take = (lower == up) ? other_less : !other_less;
                    if (take) {
                        k0 = ok;
                        i0 = oi;
                    }
                }
            }
            key[a] = k0;
            id[a] = i0;
        }
        int sb = 0, se = n;
        bool tie = false;
        for (int d = 0;; d++) {
            const int a = d % 3;
            const bool act = t < n && se - sb >= 2;
            const int m = sb + (se - sb) / 2;
            if (!__syncthreads_or(act)) break;
            const double ka = (a == 0) ? key[0] : (a == 1) ? key[1] : key[2];
            const int ia = (a == 0) ? id[0] : (a == 1) ? id[1] : id[2];
            xkey[t] = ka;
            side[ia] = (act && t < m) ? 1 : 0;
            __syncthreads();
            const bool my_tie = act && t == m && xkey[t - 1] == ka;
            if (__syncthreads_or(my_tie)) {
                tie = true;
                break;
            }
#pragma unroll
            for (int b = 0; b < 3; b++) {
                if (b == a) continue;
                const int sd = side[id[b]];
                const bool f = act && sd;
                const unsigned long long bal = ballot(f);
                if (lane == 0) wsum[wv] = __popcll(bal);
                __syncthreads();
                int pre = rank_below(bal);
                for (int q = 0; q < wv; q++) pre += wsum[q];
                xpre[t] = pre;
                __syncthreads();
                int np = t;
                if (act) {
                    const int lc = pre - xpre[sb];
                    np = sd ? sb + lc : m + (t - sb - lc);
                }
                xkey[np] = key[b];
                xid[np] = id[b];
                __syncthreads();
                key[b] = xkey[t];
                id[b] = xid[t];
                __syncthreads();
            }
            if (act) {
                if (t < m) se = m;
                else sb = m;
            }
        }
        if (tie) {
            if (t == 0) tied[atomicAdd(n_tied, 1u)] = s;
            __syncthreads();
            continue;
        }
        // DFS order: position p holds element arrangement[n-1-p]
        xid[t] = id[0];
        __syncthreads();
        const bool ok = t < n;
        const int e = ok ? xid[n - 1 - t] : 0;
        double x = 0.0, y = 0.0, z = 0.0;
        if (ok) {
            const int atom = st.list[off + e];
            const size_t g = (size_t)(abase + atom);
            x = ms.xyz[3 * g + 0];
            y = ms.xyz[3 * g + 1];
            z = ms.xyz[3 * g + 2];
            st.kx[off + t] = x;
            st.ky[off + t] = y;
            st.kz[off + t] = z;
            st.katom[off + t] = atom;
            st.kres[off + t] = ms.resseq[g];
            st.kch[off + t] = ms.chain[g];
        }
        double blo[3] = {ok ? x : INFINITY, ok ? y : INFINITY, ok ? z : INFINITY};
        double bhi[3] = {ok ? x : -INFINITY, ok ? y : -INFINITY, ok ? z : -INFINITY};
        for (int sh = 32; sh >= 1; sh >>= 1) {
#pragma unroll
            for (int q = 0; q < 3; q++) {
                blo[q] = fmin(blo[q], __shfl_xor(blo[q], sh, JAM_WAVE));
                bhi[q] = fmax(bhi[q], __shfl_xor(bhi[q], sh, JAM_WAVE));
            }
        }
        if (lane == 0 && wv * JAM_WAVE < n) {
            double *bb = st.bbox + 6 * (off / JAM_WAVE + s + wv);
            bb[0] = blo[0];
            bb[1] = blo[1];
            bb[2] = blo[2];
            bb[3] = bhi[0];
            bb[4] = bhi[1];
            bb[5] = bhi[2];
        }
        __syncthreads();
    }
}

// Loads the coordinates of set s into w, runs the fast path and writes the kd
// arrays. Always inlined, once with w in LDS and once with w in global
// scratch, so that the compiler knows the address space of every access: a
// generic pointer into LDS lets it merge two adjacent ints into one 8-byte
// flat access that is only 4-byte aligned, which LDS does not accept.
template <int NT>
__device__ __attribute__((always_inline)) static inline bool kd_fast_set(const DevMols &ms, const DevSets &st, int s,
                                                                         int n, int64_t off, int abase, KdFast w,
                                                                         int *partial, int *flag)
{
    const int tid = threadIdx.x;
    for (int i = tid; i < n; i += NT) {
        const size_t g = (size_t)(abase + st.list[off + i]);
        w.c[i] = ms.xyz[3 * g + 0];
        w.c[w.stride + i] = ms.xyz[3 * g + 1];
        w.c[2 * w.stride + i] = ms.xyz[3 * g + 2];
    }
    __syncthreads();
    int *order = st.kidx + off;
    const bool done = kd_fast<NT>(w, n, order, partial, flag);
    if (done) kd_emit<NT>(ms, st, s, n, off, abase, order);
    return done;
}

// Fast path of every set; sets with a split tie are appended to `tied` for
// the exact kernel. Tiers: <= 64 entries in registers (one wave), larger in
// LDS (NT threads, up to CAP entries), beyond KD_CAP in global scratch.
template <int NT, int CAP>
__global__ void __launch_bounds__(NT) k_kdfast(DevMols ms, DevTpls tp, DevSets st, unsigned char *scr, int lo,
                                               int hi, int *tied, unsigned int *n_tied)
{
    __shared__ double lc[CAP > JAM_WAVE ? 3 * CAP : 1];
    __shared__ int lint[CAP > JAM_WAVE ? 9 * CAP : 1];
    __shared__ unsigned char lside[CAP > JAM_WAVE ? CAP : 1];
    __shared__ int partial[NT];
    __shared__ int flag[2];
    const int tid = threadIdx.x;
    if (*st.total > st.cap) return;
    for (int s = blockIdx.x; s < st.n_sets; s += gridDim.x) {
        const int n = st.cnt[s];
        if (n <= lo || n > hi) continue;
        const int64_t off = st.off[s];
        const int m = s / tp.n_types;
        const int abase = ms.atom_off[m];
        bool done = true;
        if (CAP <= JAM_WAVE) {
            done = kd_wave64(ms, st, s, n, off, abase);
        } else if (n <= CAP) {
            KdFast w;
            w.c = lc;
            w.perm = lint;
            w.pre = lint + 6 * CAP;
            w.sb = lint + 7 * CAP;
            w.se = lint + 8 * CAP;
            w.side = lside;
            w.stride = CAP;
            done = kd_fast_set<NT>(ms, st, s, n, off, abase, w, partial, flag);
        } else if (CAP == KD_CAP) {
            int n2 = 1;
            while (n2 < n) n2 <<= 1;
            double *cb = (double *)(scr + (size_t)off * KD_SCRATCH_BYTES_PER_ENTRY);
            int *ib = (int *)(cb + 3 * (size_t)n2);
            KdFast w;
            w.c = cb;
            w.perm = ib;
            w.pre = ib + 6 * n2;
            w.sb = ib + 7 * n2;
            w.se = ib + 8 * n2;
            w.side = (unsigned char *)(ib + 9 * n2);
            w.stride = n2;
            done = kd_fast_set<NT>(ms, st, s, n, off, abase, w, partial, flag);
        }
        if (!done && tid == 0) tied[atomicAdd(n_tied, 1u)] = s;
        __syncthreads();
    }
}

// Exact path of one set (coordinates loaded into w, QuickSelect simulation,
// kd arrays written); inlined separately for LDS and global work arrays.
template <int NT>
__device__ __attribute__((always_inline)) static inline void kd_exact_set(const DevMols &ms, const DevSets &st,
                                                                          int s, int n, int64_t off, int abase,
                                                                          KdWork w, int *partial, int *flag)
{
    const int tid = threadIdx.x;
    const size_t stride = (size_t)w.stride;
    for (int i = tid; i < n; i += NT) {
        const int g = abase + st.list[off + i];
        w.c[i] = ms.xyz[3 * (size_t)g + 0];
        w.c[stride + i] = ms.xyz[3 * (size_t)g + 1];
        w.c[2 * stride + i] = ms.xyz[3 * (size_t)g + 2];
    }
    __syncthreads();
    int *order = st.kidx + off;
    kd_exact<NT>(w, n, order, partial, flag);
    kd_emit<NT>(ms, st, s, n, off, abase, order);
}

__device__ static inline KdWork kd_work_at(double *cb, int *ib, size_t stride)
{
    KdWork w;
    w.c = cb;
    w.arr = ib;
    w.pre1 = ib + 2 * stride;
    w.pre2 = ib + 3 * stride;
    w.src = ib + 4 * stride;
    w.rlo = ib + 5 * stride;
    w.rhi = ib + 6 * stride;
    w.wl = ib + 7 * stride;
    w.wr = ib + 8 * stride;
    w.stride = (int)stride;
    return w;
}

// Exact path for the sets of `tied` with n in (lo, hi], grid-stride.
template <int NT, int CAP>
__global__ void __launch_bounds__(NT) k_kdorder(DevMols ms, DevTpls tp, DevSets st, unsigned char *scr,
                                                int lo, int hi, const int *tied, const unsigned int *n_tied)
{
    __shared__ double lc[3 * CAP];
    __shared__ int lint[9 * CAP];
    __shared__ int partial[NT];
    __shared__ int flag[2];
    if (*st.total > st.cap) return;
    const int count = (int)*n_tied;
    for (int q = blockIdx.x; q < count; q += gridDim.x) {
        const int s = tied[q];
        const int n = st.cnt[s];
        if (n <= lo || n > hi) continue;
        const int64_t off = st.off[s];
        const int m = s / tp.n_types;
        const int abase = ms.atom_off[m];
        if (n <= CAP) {
            kd_exact_set<NT>(ms, st, s, n, off, abase, kd_work_at(lc, lint, CAP), partial, flag);
        } else if (CAP == KD_CAP) {
            double *cb = (double *)(scr + (size_t)off * KD_SCRATCH_BYTES_PER_ENTRY);
            int *ib = (int *)(cb + 3 * (size_t)n);
            kd_exact_set<NT>(ms, st, s, n, off, abase, kd_work_at(cb, ib, (size_t)n), partial, flag);
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// pairs: level order of every (template, molecule) pair
// ---------------------------------------------------------------------------

// Thread per pair: candidate-set sizes, the level order (stable by size when
// reordering, _CandidateSet_size_compare_r, patches/Scanner.c.patch:98-108),
// and the number of level-0 partial matches. A pair with an empty set has
// none (Scanner_reuse returns NULL, patches/Scanner.c.patch:162-166).
__global__ void __launch_bounds__(256) k_pairs(DevTpls tp, DevSets st, DevPairs pr, QueryParams qp)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= pr.n_pairs) return;
    const int T = tp.n_templates, U = tp.n_types;
    const int m = p / T, t = p % T;
    const int tbase = tp.atom_off[t];
    const int n = tp.atom_off[t + 1] - tbase;
    unsigned char *ord = pr.ord + (size_t)p * JAM_MAX_LEVELS;
    int chunks = 0;
    if (n > 0 && n <= JAM_MAX_LEVELS) {
        int cnt[JAM_MAX_LEVELS];
        bool empty = false;
        for (int k = 0; k < n; k++) {
            cnt[k] = st.cnt[m * U + tp.atom_type[tbase + k]];
            empty |= cnt[k] == 0;
            ord[k] = (unsigned char)k;
        }
        if (!empty) {
            if (qp.reorder) {
                for (int i = 1; i < n; i++) {
                    const unsigned char v = ord[i];
                    int j = i - 1;
                    while (j >= 0 && (cnt[ord[j]] > cnt[v] || (cnt[ord[j]] == cnt[v] && ord[j] > v))) {
                        ord[j + 1] = ord[j];
                        j--;
                    }
                    ord[j + 1] = v;
                }
            }
            // a one-atom template steps past entry 0 before its first match
            const int c0 = cnt[ord[0]] - (n == 1 ? 1 : 0);
            chunks = c0 > 0 ? c0 : 0;
        }
    }
    pr.nchunk[p] = chunks;
    pr.n1[p] = (n == 1) ? 1 : 0;
}

__global__ void __launch_bounds__(256) k_items(DevPairs pr)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= pr.n_pairs) return;
    const int c = pr.nchunk[p];
    const long long b = pr.item_base[p];
    for (int i = 0; i < c; i++) pr.item_pair[b + i] = p;
}

// ---------------------------------------------------------------------------
// join (the scanner), one tree level at a time
// ---------------------------------------------------------------------------
//
// The reference enumerates complete matches depth first
// (Scanner_next, patches/Scanner.c.patch:232-334): level 0 walks its
// candidate list, level L >= 1 the candidates of its set inside every annulus
// around the atoms placed at levels < L, in kd DFS order, filtered by check().
// Here every partial match of level L-1 is extended to level L by one wave,
// all partials in parallel. Children are written after a scan of their
// counts, so each level's list stays in depth-first order: parents in order,
// the children of a parent in kd order. The complete matches of a pair are
// thus contiguous and in the reference's order.

// Level-0 partials: the candidates of every pair's level-0 list, in order.
__global__ void __launch_bounds__(256) k_level0(DevPairs pr, DevFront fr)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= pr.n_items) return;
    const int p = pr.item_pair[i];
    const int n1 = pr.n1[p];
    fr.out_pair[i] = p;
    fr.out_pos[i * fr.stride] = (int)(i - pr.item_base[p]) + n1;
}

template <int MAXN>
struct ExpandLds {
    double px[MAXN], py[MAXN], pz[MAXN];
    int pch[MAXN], pres[MAXN];
    double r2lo[MAXN], r2hi[MAXN];
    int tch[MAXN], tres[MAXN];
};

// One wave extends partial i of level L-1 by the candidates of level L.
// COUNT: number of children. FILL: write the children (parent row + new kd
// position) to the next level or, when they complete the template, to the
// list of complete matches.
template <int MAXN, bool FILL>
__device__ static void expand(const DevMols &ms, const DevTpls &tp, const DevSets &st, const DevPairs &pr,
                              const DevFront &fr, const QueryParams &qp, ExpandLds<MAXN> &S, long long i)
{
    const int lane = lane_id();
    const int L = fr.level;
    const int p = fr.in_pair[i];
    const int T = tp.n_templates, U = tp.n_types;
    const int m = p / T, t = p % T;
    const int tbase = tp.atom_off[t];
    const int n = tp.atom_off[t + 1] - tbase;
    if (L >= n) {
        // a one-atom template is complete at level 0: the partial itself
        if (FILL) {
            if (lane == 0) {
                const long long r = fr.done_base + fr.off_done[i];
                fr.done_pair[r] = p;
                fr.done_pos[r * fr.stride] = fr.in_pos[i * fr.stride];
            }
        } else if (lane == 0) {
            fr.cnt_next[i] = 0;
            fr.cnt_done[i] = 1;
        }
        return;
    }
    const unsigned char *ord = pr.ord + (size_t)p * JAM_MAX_LEVELS;
    const int kL = ord[L];
    const int setL = m * U + tp.atom_type[tbase + kL];
    const int cntL = st.cnt[setL];
    const long long offL = st.off[setL];
    double lo[3] = {-INFINITY, -INFINITY, -INFINITY}, hi[3] = {INFINITY, INFINITY, INFINITY};
    if (lane < L) {
        // the atom placed at level `lane`
        const int k = ord[lane];
        const int sset = m * U + tp.atom_type[tbase + k];
        const int pos = fr.in_pos[i * fr.stride + lane];
        double x, y, z;
        int ch, res;
        if (lane == 0) {
            const int atom = st.list[st.off[sset] + pos];
            const size_t g = (size_t)(ms.atom_off[m] + atom);
            x = ms.xyz[3 * g + 0];
            y = ms.xyz[3 * g + 1];
            z = ms.xyz[3 * g + 2];
            ch = ms.chain[g];
            res = ms.resseq[g];
        } else {
            const long long e = st.off[sset] + pos;
            x = st.kx[e];
            y = st.ky[e];
            z = st.kz[e];
            ch = st.kch[e];
            res = st.kres[e];
        }
        S.px[lane] = x;
        S.py[lane] = y;
        S.pz[lane] = z;
        S.pch[lane] = ch;
        S.pres[lane] = res;
        S.tch[lane] = tp.chain[tbase + k];
        S.tres[lane] = tp.resseq[tbase + k];
        // annulus window of the pair (k, kL), patches/Scanner.c.patch:274-294,
        // Annulus_reuse patches/Annulus.c.patch:107-140
        const double d = tp.dist[tp.dist_off[t] + k * n + kL];
        double tau = qp.distance_cutoff + tp.weight[tbase + k] + tp.weight[tbase + kL];
        if (tau > qp.max_dynamic_distance) tau = qp.max_dynamic_distance;
        double a = d - tau, b = d + tau;
        if (a < 0.5) a = 0.5;
        if (b < a) {
            double tmp = a;
            a = b;
            b = tmp;
        }
        if (a < 0.0) a = 0.0;
        if (b < 0.0) b = 0.0;
        S.r2lo[lane] = a * a;
        S.r2hi[lane] = b * b;
        lo[0] = x - b;
        lo[1] = y - b;
        lo[2] = z - b;
        hi[0] = x + b;
        hi[1] = y + b;
        hi[2] = z + b;
    }
    // join box (Join_computeBox, patches/Join.c.patch:154-196)
    for (int sh = 32; sh >= 1; sh >>= 1) {
#pragma unroll
        for (int d = 0; d < 3; d++) {
            lo[d] = fmax(lo[d], __shfl_xor(lo[d], sh, JAM_WAVE));
            hi[d] = fmin(hi[d], __shfl_xor(hi[d], sh, JAM_WAVE));
        }
    }
    const int tchL = tp.chain[tbase + kL], tresL = tp.resseq[tbase + kL];
    __syncthreads();
    const int ignore = qp.ignore_chain;
    const int nblk = (cntL + JAM_WAVE - 1) / JAM_WAVE;
    const long long bbase = offL / JAM_WAVE + setL;
    const bool last = (L == n - 1);
    long long wout = 0;
    if (FILL) wout = last ? fr.done_base + fr.off_done[i] : fr.off_next[i];
    int count = 0;
    for (int g = 0; g * JAM_WAVE < nblk; g++) {
        // blocks of this group whose box meets the join box
        unsigned long long bm;
        if (nblk == 1) {
            bm = 1ull;
        } else {
            const int b = g * JAM_WAVE + lane;
            bool ok = false;
            if (b < nblk) {
                const double *bb = st.bbox + 6 * (bbase + b);
                ok = !(bb[0] > hi[0] || lo[0] > bb[3] || bb[1] > hi[1] || lo[1] > bb[4] || bb[2] > hi[2] ||
                       lo[2] > bb[5]);
            }
            bm = ballot(ok);
        }
        while (bm) {
            const int blk = g * JAM_WAVE + first_bit(bm);
            bm &= bm - 1;
            const int q = blk * JAM_WAVE + lane;
            bool ok = q < cntL;
            if (ok) {
                const long long e = offL + q;
                const double x = st.kx[e], y = st.ky[e], z = st.kz[e];
                const int ch = st.kch[e], res = st.kres[e];
                for (int j = 0; j < L; j++) {
                    // _Annulus_po: squared distance summed x, y, z from 0.0
                    double tx = S.px[j] - x, ty = S.py[j] - y, tz = S.pz[j] - z;
                    double s2 = 0.0;
                    s2 += tx * tx;
                    s2 += ty * ty;
                    s2 += tz * tz;
                    if (s2 < S.r2lo[j] || s2 > S.r2hi[j]) ok = false;
                    // TessTemplate_check
                    const bool same_chain = S.pch[j] == ch;
                    const bool tsame_chain = S.tch[j] == tchL;
                    if (ignore == 0) {
                        if (same_chain != tsame_chain) ok = false;
                        if (!same_chain) continue;
                    }
                    const bool same_res = S.pres[j] == res;
                    const bool tsame_res = S.tres[j] == tresL;
                    if (same_res != tsame_res) ok = false;
                    if (ignore != 2 && same_res && !same_chain) ok = false;
                }
            }
            const unsigned long long mask = ballot(ok);
            if (FILL && ok) {
                const long long r = wout + count + rank_below(mask);
                int *row = last ? fr.done_pos + r * fr.stride : fr.out_pos + r * fr.stride;
                for (int l = 0; l < L; l++) row[l] = fr.in_pos[i * fr.stride + l];
                row[L] = q;
                if (last) fr.done_pair[r] = p;
                else fr.out_pair[r] = p;
            }
            count += __popcll(mask);
        }
    }
    if (!FILL && lane == 0) {
        fr.cnt_next[i] = last ? 0 : count;
        fr.cnt_done[i] = last ? count : 0;
    }
}

template <int MAXN, bool FILL>
__global__ void __launch_bounds__(64) k_expand(DevMols ms, DevTpls tp, DevSets st, DevPairs pr, DevFront fr,
                                               QueryParams qp)
{
    __shared__ ExpandLds<MAXN> S;
    const long long i = blockIdx.x;
    if (i >= fr.n_in) return;
    expand<MAXN, FILL>(ms, tp, st, pr, fr, qp, S, i);
}

// ---------------------------------------------------------------------------
// superposition + events
// ---------------------------------------------------------------------------

// first complete match of every pair: the matches of a pair are contiguous
__global__ void __launch_bounds__(256) k_first(const int *done_pair, long long n, long long *first)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (i == 0 || done_pair[i - 1] != done_pair[i]) first[done_pair[i]] = i;
}

__global__ void __launch_bounds__(256) k_super(DevMols ms, DevTpls tp, DevSets st, DevPairs pr, DevFront fr,
                                               DevEvents ev, QueryParams qp, long long n_matches,
                                               const long long *first)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_matches) return;
    const int pair = fr.done_pair[i];
    const int T = tp.n_templates, U = tp.n_types;
    const int m = pair / T, t = pair % T;
    const int tbase = tp.atom_off[t];
    const int n = tp.atom_off[t + 1] - tbase;
    const int abase = ms.atom_off[m];
    const unsigned char *ord = pr.ord + (size_t)pair * JAM_MAX_LEVELS;
    const int *pos = fr.done_pos + i * fr.stride;
    // molecule-local atom of template atom k: placed at the level l with ord[l] == k
    int atoms[JAM_MAX_LEVELS];
    for (int l = 0; l < n; l++) {
        const int k = ord[l];
        const int sset = m * U + tp.atom_type[tbase + k];
        atoms[k] = (l == 0) ? st.list[st.off[sset] + pos[0]] : st.katom[st.off[sset] + pos[l]];
    }
    auto pa = [&](int k, int d) { return ms.xyz[3 * (size_t)(abase + atoms[k]) + d]; };
    auto pb = [&](int k, int d) { return tp.pos[3 * (size_t)(tbase + k) + d]; };
    double rmsd, rot[9], c0[3], c1[3];
    superpose(n, pa, pb, &rmsd, rot, c0, c1);
    if (!(rmsd <= qp.rmsd_threshold)) return;
    bool nan = false;
#pragma unroll
    for (int k = 0; k < 9; k++) nan |= isnan(rot[k]);
    const unsigned long long e = atomicAdd(ev.count, 1ull);
    if (e >= (unsigned long long)ev.cap) return;
    ev.pair[e] = pair;
    ev.rank[e] = i - first[pair];
    ev.flag[e] = nan ? 2 : 1;
    ev.rmsd[e] = rmsd;
#pragma unroll
    for (int k = 0; k < 9; k++) ev.rot[9 * e + k] = rot[k];
#pragma unroll
    for (int d = 0; d < 3; d++) {
        ev.c0[3 * e + d] = c0[d];
        ev.c1[3 * e + d] = c1[d];
    }
    for (int k = 0; k < ev.stride; k++) ev.atoms[e * (size_t)ev.stride + k] = k < n ? atoms[k] : -1;
}

// Standalone batched superposition: sets of k points, a/b as [set][point][xyz].
__global__ void __launch_bounds__(256) k_superpose_batch(const double *a, const double *b, int k,
                                                         long long n_sets, double *rmsd, double *rot,
                                                         double *c0, double *c1)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_sets) return;
    const double *A = a + i * 3 * (size_t)k;
    const double *B = b + i * 3 * (size_t)k;
    auto pa = [&](int q, int d) { return A[3 * q + d]; };
    auto pb = [&](int q, int d) { return B[3 * q + d]; };
    double r, M[9], x0[3], x1[3];
    superpose(k, pa, pb, &r, M, x0, x1);
    rmsd[i] = r;
#pragma unroll
    for (int q = 0; q < 9; q++) rot[9 * i + q] = M[q];
#pragma unroll
    for (int d = 0; d < 3; d++) {
        c0[3 * i + d] = x0[d];
        c1[3 * i + d] = x1[d];
    }
}

}  // namespace jam
