// jam_kernels.h -- gfx950 kernels of the Jess search pipeline.
//
//   k_candidates   wave per (molecule, atom type): candidate list in the
//                  reference order (TessTemplate_candidates,
//                  patches/TessTemplate.c.patch:173-235), two passes (count,
//                  fill) with ballot compaction.
//   k_kd*          (jam_kd.h) the leaf order of the reference's QuickSelect
//                  kd-tree of every candidate set, the set stored as SoA in
//                  that order plus one bounding box per 64 entries.
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
#include "jam_kd.h"
#include "jam_super.h"
#include "jam_wave.h"

namespace jam {

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
    wave_sync();
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
