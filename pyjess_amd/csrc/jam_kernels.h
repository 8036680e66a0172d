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
//   k_join         wave per (template, molecule, level-0 range): the Scanner
//                  DFS (patches/Scanner.c.patch:232-334) with an explicit
//                  per-level stack in LDS; each level scans 64-entry blocks of
//                  the kd-ordered set whose box meets the join box, all 64
//                  lanes testing one candidate each against every annulus and
//                  check(), survivors taken lowest lane first (= kd order).
//   k_super        thread per complete match: FP64 superposition, RMSD
//                  threshold, event compaction.
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

#define KD_THREADS 256
#define KD_CAP 1024          // largest set arranged in LDS
// per set entry: 3 coordinates and 9 ints of work arrays (60 bytes), for sets
// too large for LDS
#define KD_SCRATCH_BYTES_PER_ENTRY 64

struct KdWork {
    double *c[3];            // coordinates by element
    int *arr[2];             // arrangement, double buffered
    int *pre1, *pre2;        // exclusive prefixes: "<=" flags, re-push flags
    int *src;                // push -> element (>= 0) or -(push index) - 1
    int *rlo, *rhi;          // range of each position at the current level
    int *wl, *wr;            // QuickSelect window of the range starting here; wr < 0: done
};

// in-place exclusive block scan of v[0, n)
__device__ static void kd_scan(int *v, int n, int *partial)
{
    const int tid = threadIdx.x;
    const int per = (n + KD_THREADS - 1) / KD_THREADS;
    const int b0 = tid * per, b1 = min(n, b0 + per);
    int cnt = 0;
    for (int p = b0; p < b1; p++) cnt += v[p];
    partial[tid] = cnt;
    __syncthreads();
    if (tid < JAM_WAVE) {
        int x[KD_THREADS / JAM_WAVE], sum = 0;
#pragma unroll
        for (int q = 0; q < KD_THREADS / JAM_WAVE; q++) {
            x[q] = partial[(KD_THREADS / JAM_WAVE) * tid + q];
            sum += x[q];
        }
        int incl = sum;
        for (int sh = 1; sh < JAM_WAVE; sh <<= 1) {
            int o = __shfl_up(incl, sh, JAM_WAVE);
            if (tid >= sh) incl += o;
        }
        int run = incl - sum;
#pragma unroll
        for (int q = 0; q < KD_THREADS / JAM_WAVE; q++) {
            partial[(KD_THREADS / JAM_WAVE) * tid + q] = run;
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

__device__ static inline int kd_cmp3(const double *c, int x, int y)
{
    const double u = c[x], v = c[y];
    return u < v ? -1 : (u > v ? 1 : 0);
}

// Arranges elements 0..n-1 into the reference's leaf order, written to out.
__device__ static void kd_exact(const KdWork &w, int n, int *out, int *partial, int *flag)
{
    const int tid = threadIdx.x;
    int cur = 0;
    for (int p = tid; p < n; p += KD_THREADS) {
        w.arr[0][p] = p;
        w.rlo[p] = 0;
        w.rhi[p] = n;
    }
    __syncthreads();
    for (int d = 0;; d++) {
        const double *c = w.c[d % 3];
        if (tid == 0) flag[0] = 0;
        __syncthreads();
        for (int p = tid; p < n; p += KD_THREADS) {
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
            int *a = w.arr[cur];
            int *an = w.arr[cur ^ 1];
            // pivot: median of three, moved to the right end of the window
            for (int p = tid; p < n; p += KD_THREADS) {
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
            for (int i = tid; i < n; i += KD_THREADS) {
                const int R = w.rlo[i];
                const int r = w.wr[R];
                int b = 0;
                if (r >= 0 && i >= w.wl[R] && i < r) b = c[a[i]] <= c[a[r]] ? 1 : 0;
                w.pre1[i] = b;
            }
            __syncthreads();
            kd_scan(w.pre1, n, partial);
            // re-push flags: a "<=" element after the first ">" one
            for (int i = tid; i < n; i += KD_THREADS) {
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
            kd_scan(w.pre2, n, partial);
            // pushes: ">" elements are new pushes, re-pushes point to push q
            for (int i = tid; i < n; i += KD_THREADS) {
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
                for (int i = tid; i < n; i += KD_THREADS) {
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
            for (int i = tid; i < n; i += KD_THREADS) {
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
            for (int p = tid; p < n; p += KD_THREADS) {
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
        for (int p = tid; p < n; p += KD_THREADS) {
            const int lo = w.rlo[p], hi = w.rhi[p];
            if (hi - lo < 2) continue;
            const int m = lo + (hi - lo) / 2;
            if (p < m) w.rhi[p] = m;
            else w.rlo[p] = m;
        }
        __syncthreads();
    }
    const int *fin = w.arr[cur];
    for (int p = tid; p < n; p += KD_THREADS) out[p] = fin[p];
    __syncthreads();
}

// One workgroup per candidate set. Sets of up to KD_CAP entries work in LDS,
// larger ones in a global scratch area of the same layout (scr).
__global__ void __launch_bounds__(KD_THREADS) k_kdorder(DevMols ms, DevTpls tp, DevSets st, unsigned char *scr)
{
    __shared__ double lc[3][KD_CAP];
    __shared__ int lint[9][KD_CAP];
    __shared__ int partial[KD_THREADS];
    __shared__ int flag[2];
    const int s = blockIdx.x;
    const int tid = threadIdx.x;
    if (s >= st.n_sets) return;
    if (*st.total > st.cap) return;
    const int n = st.cnt[s];
    if (n == 0) return;
    const int64_t off = st.off[s];
    const int m = s / tp.n_types;
    const int abase = ms.atom_off[m];
    double *cb;
    int *ib;
    size_t stride;
    if (n <= KD_CAP) {
        cb = &lc[0][0];
        ib = &lint[0][0];
        stride = KD_CAP;
    } else {
        cb = (double *)(scr + (size_t)off * KD_SCRATCH_BYTES_PER_ENTRY);
        ib = (int *)(cb + 3 * (size_t)n);
        stride = (size_t)n;
    }
    KdWork w;
    for (int a = 0; a < 3; a++) w.c[a] = cb + a * stride;
    w.arr[0] = ib;
    w.arr[1] = ib + stride;
    w.pre1 = ib + 2 * stride;
    w.pre2 = ib + 3 * stride;
    w.src = ib + 4 * stride;
    w.rlo = ib + 5 * stride;
    w.rhi = ib + 6 * stride;
    w.wl = ib + 7 * stride;
    w.wr = ib + 8 * stride;
    for (int i = tid; i < n; i += KD_THREADS) {
        const int g = abase + st.list[off + i];
        w.c[0][i] = ms.xyz[3 * (size_t)g + 0];
        w.c[1][i] = ms.xyz[3 * (size_t)g + 1];
        w.c[2][i] = ms.xyz[3 * (size_t)g + 2];
    }
    __syncthreads();
    int *order = st.kidx + off;
    kd_exact(w, n, order, partial, flag);
    // DFS visits right subtrees first: position p of the visit order is
    // element n-1-p of the arrangement.
    const int64_t bbase = off / JAM_WAVE + s;
    const int lane = lane_id();
    const int wave = tid / JAM_WAVE;
    for (int b0 = wave * JAM_WAVE; b0 < n; b0 += KD_THREADS) {
        const int p = b0 + lane;
        double x = 0.0, y = 0.0, z = 0.0;
        const bool ok = p < n;
        if (ok) {
            const int li = order[n - 1 - p];
            const int atom = st.list[off + li];
            const int g = abase + atom;
            x = w.c[0][li];
            y = w.c[1][li];
            z = w.c[2][li];
            st.kx[off + p] = x;
            st.ky[off + p] = y;
            st.kz[off + p] = z;
            st.katom[off + p] = atom;
            st.kres[off + p] = ms.resseq[g];
            st.kch[off + p] = ms.chain[g];
        }
        double lo[3] = {ok ? x : INFINITY, ok ? y : INFINITY, ok ? z : INFINITY};
        double hi[3] = {ok ? x : -INFINITY, ok ? y : -INFINITY, ok ? z : -INFINITY};
        for (int sh = 32; sh >= 1; sh >>= 1) {
#pragma unroll
            for (int d = 0; d < 3; d++) {
                lo[d] = fmin(lo[d], __shfl_xor(lo[d], sh, JAM_WAVE));
                hi[d] = fmax(hi[d], __shfl_xor(hi[d], sh, JAM_WAVE));
            }
        }
        if (lane == 0) {
            double *bb = st.bbox + 6 * (bbase + b0 / JAM_WAVE);
            bb[0] = lo[0];
            bb[1] = lo[1];
            bb[2] = lo[2];
            bb[3] = hi[0];
            bb[4] = hi[1];
            bb[5] = hi[2];
        }
    }
}

// ---------------------------------------------------------------------------
// join (the scanner)
// ---------------------------------------------------------------------------

template <int MAXN>
struct JoinLds {
    int ord[MAXN];            // level -> template atom
    int lvl[MAXN];            // template atom -> level
    int cnt[MAXN];            // per level: set size
    long long off[MAXN];      // per level: set offset
    long long bbase[MAXN];    // per level: first block box
    int tch[MAXN], tres[MAXN];  // per level: template chain / resSeq
    double px[MAXN], py[MAXN], pz[MAXN];  // per level: placed atom
    int pch[MAXN], pres[MAXN], patom[MAXN];
    double box[MAXN][6];
    double r2lo[MAXN * (MAXN - 1) / 2], r2hi[MAXN * (MAXN - 1) / 2];
    unsigned long long blk[MAXN], surv[MAXN];
    int grp[MAXN], cur[MAXN];
};

__device__ static inline int tri(int L) { return L * (L - 1) / 2; }

template <int MAXN>
__device__ static inline unsigned long long test_blocks(const DevSets &st, JoinLds<MAXN> &S, int L, int g)
{
    const int lane = lane_id();
    const int nblk = (S.cnt[L] + JAM_WAVE - 1) / JAM_WAVE;
    const int b = g * JAM_WAVE + lane;
    bool ok = false;
    if (b < nblk) {
        const double *bb = st.bbox + 6 * (S.bbase[L] + b);
        ok = true;
#pragma unroll
        for (int d = 0; d < 3; d++)
            if (bb[d] > S.box[L][3 + d] || S.box[L][d] > bb[3 + d]) ok = false;
    }
    return ballot(ok);
}

template <int MAXN>
__device__ static inline unsigned long long scan_block(const DevSets &st, JoinLds<MAXN> &S, int L, int blk,
                                                       int ignore)
{
    const int lane = lane_id();
    const int p = blk * JAM_WAVE + lane;
    bool ok = p < S.cnt[L];
    if (ok) {
        const long long e = S.off[L] + p;
        const double x = st.kx[e], y = st.ky[e], z = st.kz[e];
        const int ch = st.kch[e], res = st.kres[e];
        const int t0 = tri(L);
        const int tchL = S.tch[L], tresL = S.tres[L];
        for (int j = 0; j < L; j++) {
            // _Annulus_po: squared distance accumulated x, y, z from 0.0
            double tx = S.px[j] - x, ty = S.py[j] - y, tz = S.pz[j] - z;
            double s = 0.0;
            s += tx * tx;
            s += ty * ty;
            s += tz * tz;
            if (s < S.r2lo[t0 + j] || s > S.r2hi[t0 + j]) ok = false;
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
    return ballot(ok);
}

// Opens level L: annulus windows against every earlier level
// (patches/Scanner.c.patch:274-294, Annulus_reuse patches/Annulus.c.patch:107-140),
// the join box (Join_computeBox, patches/Join.c.patch:154-196), block mask.
template <int MAXN>
__device__ static inline void open_level(const DevTpls &tp, const DevSets &st, JoinLds<MAXN> &S, int L,
                                         int t, int tbase, int n, const QueryParams &qp)
{
    const int lane = lane_id();
    double lo[3] = {-INFINITY, -INFINITY, -INFINITY}, hi[3] = {INFINITY, INFINITY, INFINITY};
    if (lane < L) {
        const int i = S.ord[lane], k = S.ord[L];
        const double d = tp.dist[tp.dist_off[t] + i * n + k];
        double tau = qp.distance_cutoff + tp.weight[tbase + i] + tp.weight[tbase + k];
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
        S.r2lo[tri(L) + lane] = a * a;
        S.r2hi[tri(L) + lane] = b * b;
        lo[0] = S.px[lane] - b;
        lo[1] = S.py[lane] - b;
        lo[2] = S.pz[lane] - b;
        hi[0] = S.px[lane] + b;
        hi[1] = S.py[lane] + b;
        hi[2] = S.pz[lane] + b;
    }
    for (int sh = 32; sh >= 1; sh >>= 1) {
#pragma unroll
        for (int d = 0; d < 3; d++) {
            lo[d] = fmax(lo[d], __shfl_xor(lo[d], sh, JAM_WAVE));
            hi[d] = fmin(hi[d], __shfl_xor(hi[d], sh, JAM_WAVE));
        }
    }
    if (lane == 0) {
#pragma unroll
        for (int d = 0; d < 3; d++) {
            S.box[L][d] = lo[d];
            S.box[L][3 + d] = hi[d];
        }
    }
    __syncthreads();
    const unsigned long long bm = test_blocks<MAXN>(st, S, L, 0);
    if (lane == 0) {
        S.grp[L] = 0;
        S.blk[L] = bm;
        S.surv[L] = 0ull;
        S.cur[L] = -1;
    }
    __syncthreads();
}

// Next candidate of level L in kd DFS order, as a position in the set, or -1.
template <int MAXN>
__device__ static inline int next_candidate(const DevSets &st, JoinLds<MAXN> &S, int L, int ignore)
{
    const int lane = lane_id();
    unsigned long long surv = S.surv[L];
    while (surv == 0ull) {
        unsigned long long blk = S.blk[L];
        int g = S.grp[L];
        const int nblk = (S.cnt[L] + JAM_WAVE - 1) / JAM_WAVE;
        while (blk == 0ull) {
            g++;
            if (g * JAM_WAVE >= nblk) {
                if (lane == 0) {
                    S.grp[L] = g;
                    S.blk[L] = 0ull;
                }
                __syncthreads();
                return -1;
            }
            blk = test_blocks<MAXN>(st, S, L, g);
        }
        const int b = first_bit(blk);
        blk &= blk - 1;
        const int cur = g * JAM_WAVE + b;
        surv = scan_block<MAXN>(st, S, L, cur, ignore);
        if (lane == 0) {
            S.grp[L] = g;
            S.blk[L] = blk;
            S.cur[L] = cur;
        }
        __syncthreads();
    }
    const int i = first_bit(surv);
    surv &= surv - 1;
    const int cur = S.cur[L];
    __syncthreads();
    if (lane == 0) S.surv[L] = surv;
    __syncthreads();
    return cur * JAM_WAVE + i;
}

template <int MAXN>
__device__ static inline void place(const DevSets &st, JoinLds<MAXN> &S, int L, int pos)
{
    if (lane_id() == 0) {
        const long long e = S.off[L] + pos;
        S.px[L] = st.kx[e];
        S.py[L] = st.ky[e];
        S.pz[L] = st.kz[e];
        S.pch[L] = st.kch[e];
        S.pres[L] = st.kres[e];
        S.patom[L] = st.katom[e];
    }
    __syncthreads();
}

// One wave per work item = (pair, level-0 range); pair = molecule * T + template.
template <int MAXN>
__global__ void __launch_bounds__(64) k_join(DevMols ms, DevTpls tp, DevSets st, DevMatches mt,
                                             QueryParams qp, int n_items)
{
    __shared__ JoinLds<MAXN> S;
    const int item = blockIdx.x;
    const int lane = lane_id();
    if (item >= n_items) return;
    if (*st.total > st.cap) return;
    const int G = qp.items_per_pair;
    const int pair = item / G, g = item % G;
    const int T = tp.n_templates;
    const int m = pair / T, t = pair % T;
    const int tbase = tp.atom_off[t];
    const int n = tp.atom_off[t + 1] - tbase;
    const int U = tp.n_types;
    int seq = 0;
    if (n <= 0 || n > MAXN) {
        if (lane == 0) mt.item_cm[item] = 0;
        return;
    }
    // candidate-set sizes and the level order (_CandidateSet_size_compare_r,
    // patches/Scanner.c.patch:98-108: by size, then template index)
    int mycnt = 0;
    long long myoff = 0;
    if (lane < n) {
        const int sset = m * U + tp.atom_type[tbase + lane];
        mycnt = st.cnt[sset];
        myoff = st.off[sset];
    }
    const bool empty = ballot(lane < n && mycnt == 0) != 0ull;
    if (empty) {
        if (lane == 0) mt.item_cm[item] = 0;
        return;
    }
    int rank = lane;
    if (qp.reorder) {
        rank = 0;
        for (int j = 0; j < n; j++) {
            const int cj = __shfl(mycnt, j, JAM_WAVE);
            if (lane < n && (cj < mycnt || (cj == mycnt && j < lane))) rank++;
        }
    }
    if (lane < n) {
        S.ord[rank] = lane;
        S.lvl[lane] = rank;
        S.cnt[rank] = mycnt;
        S.off[rank] = myoff;
        S.bbase[rank] = myoff / JAM_WAVE + (m * U + tp.atom_type[tbase + lane]);
        S.tch[rank] = tp.chain[tbase + lane];
        S.tres[rank] = tp.resseq[tbase + lane];
    }
    __syncthreads();
    // level-0 range of this work item
    const int c0 = S.cnt[0];
    int a = (int)(((long long)c0 * g) / G), b = (int)(((long long)c0 * (g + 1)) / G);
    if (n == 1 && a == 0) a = 1;   // the reference's first advance steps past entry 0
    const long long off0 = S.off[0];
    const int abase = ms.atom_off[m];
    const int o0set = m * U + tp.atom_type[tbase + S.ord[0]];
    (void)o0set;
    for (int i0 = a; i0 < b; i0++) {
        if (lane == 0) {
            const int atom = st.list[off0 + i0];
            const size_t gi = (size_t)(abase + atom);
            S.px[0] = ms.xyz[3 * gi + 0];
            S.py[0] = ms.xyz[3 * gi + 1];
            S.pz[0] = ms.xyz[3 * gi + 2];
            S.pch[0] = ms.chain[gi];
            S.pres[0] = ms.resseq[gi];
            S.patom[0] = atom;
        }
        __syncthreads();
        int L = 1;
        bool complete = (n == 1);
        if (!complete) open_level<MAXN>(tp, st, S, 1, t, tbase, n, qp);
        while (true) {
            if (complete) {
                // emit one complete match
                unsigned long long slot = 0;
                if (lane == 0) slot = atomicAdd(mt.count, 1ull);
                slot = __shfl(slot, 0, JAM_WAVE);
                if (slot < (unsigned long long)mt.cap) {
                    if (lane == 0) {
                        mt.item[slot] = item;
                        mt.seq[slot] = seq;
                    }
                    if (lane < n) mt.atoms[slot * (size_t)mt.stride + lane] = S.patom[S.lvl[lane]];
                }
                seq++;
                complete = false;
                if (n == 1) break;
            }
            const int pos = next_candidate<MAXN>(st, S, L, qp.ignore_chain);
            if (pos < 0) {
                if (--L == 0) break;
                continue;
            }
            place<MAXN>(st, S, L, pos);
            if (L == n - 1) {
                complete = true;
                continue;
            }
            L++;
            open_level<MAXN>(tp, st, S, L, t, tbase, n, qp);
        }
    }
    if (lane == 0) mt.item_cm[item] = seq;
}

// ---------------------------------------------------------------------------
// superposition + events
// ---------------------------------------------------------------------------

__global__ void __launch_bounds__(256) k_super(DevMols ms, DevTpls tp, DevMatches mt, DevEvents ev,
                                               QueryParams qp, long long n_matches)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_matches) return;
    const int G = qp.items_per_pair;
    const int item = mt.item[i];
    const int pair = item / G, g = item % G;
    const int T = tp.n_templates;
    const int m = pair / T, t = pair % T;
    const int tbase = tp.atom_off[t];
    const int n = tp.atom_off[t + 1] - tbase;
    const int abase = ms.atom_off[m];
    const int *atoms = mt.atoms + i * (size_t)mt.stride;
    auto pa = [&](int k, int d) { return ms.xyz[3 * (size_t)(abase + atoms[k]) + d]; };
    auto pb = [&](int k, int d) { return tp.pos[3 * (size_t)(tbase + k) + d]; };
    double rmsd, rot[9], c0[3], c1[3];
    superpose(n, pa, pb, &rmsd, rot, c0, c1);
    if (!(rmsd <= qp.rmsd_threshold)) return;
    bool nan = false;
#pragma unroll
    for (int k = 0; k < 9; k++) nan |= isnan(rot[k]);
    long long rank = mt.seq[i];
    for (int h = pair * G; h < item; h++) rank += mt.item_cm[h];
    const unsigned long long e = atomicAdd(ev.count, 1ull);
    if (e >= (unsigned long long)ev.cap) return;
    ev.pair[e] = pair;
    ev.rank[e] = rank;
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
