// jam_kernels.h -- gfx950 kernels of the Jess search pipeline.
//
//   k_candidates   wave per (molecule, atom type): candidate list in the
//                  reference order (TessTemplate_candidates,
//                  patches/TessTemplate.c.patch:173-235), two passes (count,
//                  fill) with ballot compaction.
//   k_kdorder      wave per candidate set: the leaf order of the reference's
//                  QuickSelect kd-tree (patches/KdTree.c.patch:348-451,
//                  patches/qselect.h.patch:50-116), reversed because its DFS
//                  visits right children first; writes the set as SoA in that
//                  order plus one bounding box per 64 entries.
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

struct KdCmp {
    const double *xyz;
    const int *list;
    const int *idx;
    int abase;
    int axis;
    __device__ double at(size_t i) const { return xyz[3 * (size_t)(abase + list[idx[i]]) + axis]; }
    __device__ int cmp(size_t i, size_t j) const
    {
        double x = at(i), y = at(j);
        return x < y ? -1 : (x > y ? 1 : 0);
    }
};

__device__ static inline void iswap(int *v, size_t i, size_t j)
{
    int t = v[i];
    v[i] = v[j];
    v[j] = t;
}

// QuickSelect of the k-th element of idx[0, n) by coordinate `axis`:
// median-of-three pivot (mid = (l + r + 1) / 2), Lomuto partition with
// "<= pivot goes left". The same swap sequence as the reference is needed
// because it decides the side of elements equal to the median.
__device__ static size_t quickselect(KdCmp &c, int *idx, size_t n, size_t k)
{
    size_t l = 0, r = n - 1;
    c.idx = idx;
    for (;;) {
        if (l == r) return l;
        size_t mid = (l + r + 1) / 2, p;
        if ((c.cmp(l, mid) > 0) != (c.cmp(l, r) > 0)) p = l;
        else if ((c.cmp(mid, l) < 0) != (c.cmp(mid, r) < 0)) p = mid;
        else p = r;
        iswap(idx, r, p);
        size_t store = l;
        for (size_t i = l; i < r; i++) {
            if (c.cmp(i, r) <= 0) {
                iswap(idx, i, store);
                store++;
            }
        }
        iswap(idx, store, r);
        p = store;
        if (p == k) return p;
        if (k < p) r = p - 1;
        else l = p + 1;
    }
}

// Recursive median split (left = [0, n/2), right = [n/2, n), axis cycling
// x, y, z) done with an explicit stack of ranges, on lane 0.
__device__ static void kd_arrange(KdCmp &c, int *idx, int n)
{
    struct Rng { int lo, n, axis; };
    Rng stack[64];
    int sp = 0;
    stack[sp++] = {0, n, 0};
    while (sp > 0) {
        Rng r = stack[--sp];
        if (r.n <= 1) continue;
        c.axis = r.axis;
        size_t split = quickselect(c, idx + r.lo, (size_t)r.n, (size_t)(r.n / 2));
        int na = (r.axis + 1) % 3;
        // order of processing does not matter: the ranges are disjoint
        stack[sp++] = {r.lo + (int)split, r.n - (int)split, na};
        stack[sp++] = {r.lo, (int)split, na};
    }
}

__global__ void __launch_bounds__(64) k_kdorder(DevMols ms, DevTpls tp, DevSets st)
{
    const int s = blockIdx.x;
    const int lane = lane_id();
    if (s >= st.n_sets) return;
    if (*st.total > st.cap) return;
    const int cnt = st.cnt[s];
    if (cnt == 0) return;
    const int64_t off = st.off[s];
    const int m = s / tp.n_types;
    const int abase = ms.atom_off[m];
    int *idx = st.kidx + off;
    for (int i = lane; i < cnt; i += JAM_WAVE) idx[i] = i;
    __syncthreads();
    if (lane == 0) {
        KdCmp c{ms.xyz, st.list + off, idx, abase, 0};
        kd_arrange(c, idx, cnt);
    }
    __syncthreads();
    // DFS visits right subtrees first: position p of the visit order is
    // element cnt-1-p of the arranged index array.
    const int64_t bbase = off / JAM_WAVE + s;
    for (int b0 = 0; b0 < cnt; b0 += JAM_WAVE) {
        const int p = b0 + lane;
        double x = 0.0, y = 0.0, z = 0.0;
        const bool ok = p < cnt;
        if (ok) {
            const int li = idx[cnt - 1 - p];
            const int atom = st.list[off + li];
            const int g = abase + atom;
            x = ms.xyz[3 * (size_t)g + 0];
            y = ms.xyz[3 * (size_t)g + 1];
            z = ms.xyz[3 * (size_t)g + 2];
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
