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
//   k_join         one launch per join level, a team of 16 lanes per partial
//                  match of level L-1 (four per wave): the candidates of
//                  level L (patches/Scanner.c.patch:232-334) from the
//                  64-entry blocks of the kd-ordered set whose box meets the
//                  join box, each lane testing one candidate against every
//                  annulus and check(); children appended in any order.
//   k_sortkey      complete matches sorted by (molecule, template in scan
//   k_rank         order, level positions): the reference's depth-first
//                  order and each match's rank (radix passes, or rank
//                  counting for a few thousand matches).
//   k_super        thread per complete match: FP64 superposition and RMSD
//                  threshold, then (after a scan of the marks) the events in
//                  that same order, each with the match's rank.
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
__global__ void __launch_bounds__(256) k_candidates(DevMols ms, DevTpls tp, DevSets st, DevPairs pr,
                                                    long long *n_items)
{
    if (FILL && blockIdx.x == 0 && threadIdx.x == 0) {
        // level-0 items of all pairs: the end of the scan of the pair chunks
        const int np = pr.n_pairs;
        *n_items = np > 0 ? pr.item_base[np - 1] + pr.nchunk[np - 1] : 0;
    }
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
// Half a wave per pair, lane k for template atom k (templates have at most
// 32 atoms): the level of atom k is its rank by (set size, k) when
// reordering, k otherwise.
__global__ void __launch_bounds__(256) k_pairs(DevTpls tp, DevSets st, DevPairs pr, QueryParams qp)
{
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        // candidate-set entries of the query: the end of the scan of the set sizes
        const int ns = st.n_sets;
        *st.total = ns > 0 ? st.off[ns - 1] + st.cnt[ns - 1] : 0;
    }
    const int lane = lane_id();
    const int team = lane >> 5, k = lane & 31;
    const int p = (int)(((long long)blockIdx.x * blockDim.x + threadIdx.x) / JAM_WAVE) * 2 + team;
    const bool valid = p < pr.n_pairs;
    const int T = tp.n_templates, U = tp.n_types;
    int m = 0, t = 0, tbase = 0, n = 0;
    if (valid) {
        m = p / T;
        t = p % T;
        tbase = tp.atom_off[t];
        n = tp.atom_off[t + 1] - tbase;
    }
    const bool mine = valid && k < n;
    int set = 0, cnt = 0;
    long long off = 0;
    if (mine) {
        set = m * U + tp.atom_type[tbase + k];
        cnt = st.cnt[set];
        off = st.off[set];
    }
    // an empty candidate set means no match for the pair (Scanner_reuse
    // returns NULL, patches/Scanner.c.patch:162-166)
    const unsigned long long empty = ballot(mine && cnt == 0);
    const unsigned long long team_mask = team ? 0xFFFFFFFF00000000ull : 0x00000000FFFFFFFFull;
    const bool any_empty = (empty & team_mask) != 0ull;
    int rank = k;
    if (qp.reorder) {
        rank = 0;
        for (int j = 0; j < 32; j++) {
            const int cj = __shfl(cnt, (team << 5) + j, JAM_WAVE);
            if (j < n && (cj < cnt || (cj == cnt && j < k))) rank++;
        }
    }
    if (mine && !any_empty) {
        pr.ord[(size_t)p * JAM_MAX_LEVELS + rank] = (unsigned char)k;
        PairLevel h;
        h.k = k;
        h.set = set;
        h.cnt = cnt;
        h.off = (int)off;
        pr.lvl[(size_t)p * JAM_MAX_LEVELS + rank] = h;
    }
    if (valid && k == 0) pr.n1[p] = (n == 1) ? 1 : 0;
    // level-0 list size: the atom of rank 0; a one-atom template steps past
    // entry 0 before its first match
    if (valid && (n == 0 || any_empty) && k == 0) pr.nchunk[p] = 0;
    if (mine && !any_empty && rank == 0) {
        const int c0 = cnt - (n == 1 ? 1 : 0);
        pr.nchunk[p] = c0 > 0 ? c0 : 0;
    }
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
// all partials of the level in parallel, one launch per level, and the
// children are appended in any order. Depth-first order is lexicographic
// order of (level-0 list position, level-1 kd position, ...) within a pair,
// so sorting the complete matches by (pair, positions) recovers the
// reference's enumeration order and each match's rank.

// Level-0 rows: the candidates of every pair's level-0 list, enumerated in
// the kd order of its set (consecutive rows then lie close in space, and the
// join's candidate blocks stay in cache from row to row); the ranking of the
// complete matches reads each row's list position back (kidx), so the order
// rows are made in does not matter.
// The buffers were sized before the sizes were known: when the candidate
// sets, the items or the frontier do not fit, no row is written (the join then
// has nothing to do) and the host grows the buffers and runs the query again.
__global__ void __launch_bounds__(256) k_level0(DevSets st, DevPairs pr, DevJoin jn)
{
    const long long n = *pr.n_items;
    const bool fits = *st.total <= st.cap && n <= pr.items_cap && n <= jn.out_cap;
    if (blockIdx.x == 0 && threadIdx.x == 0) *jn.n_out = fits ? (unsigned long long)n : 0ull;
    if (!fits) return;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x) {
        const int p = pr.item_pair[i];
        const int off0 = pr.lvl[(size_t)p * JAM_MAX_LEVELS].off;
        int k = (int)(i - pr.item_base[p]);
        // a one-atom template steps past list entry 0: skip its kd position
        if (pr.n1[p] && k >= st.l2k[off0]) k++;
        jn.out_pair[i] = p;
        jn.out_ent[i] = off0 + k;   // level 0's frontier: one entry per row
    }
}

#ifndef JOIN_WAVES
#define JOIN_WAVES 4
#endif
#ifndef JOIN_STAGE
#define JOIN_STAGE 128
#endif
// waves per SIMD the join's register budget is set for (the LDS of
// JOIN_WAVES-wave workgroups allows five)
#ifndef JOIN_MIN_WAVES
#define JOIN_MIN_WAVES 5
#endif
// lanes per team (one partial match per team; a team tests the JAM_SUB
// candidates of one sub-block at a time)
#ifndef JOIN_TW
#define JOIN_TW 16
#endif
// most waves sharing one row's blocks when the frontier is small
#ifndef JOIN_SPLIT
#define JOIN_SPLIT 16
#endif
// sets of at most JOIN_SMALL blocks skip the block-box test: their candidates
// are loaded at once
#ifndef JOIN_SMALL
#define JOIN_SMALL 4
#endif

// Writes `staged` rows (parent row par[k] + the new entry ent[k]) to a row
// buffer of `ostride` ints per row: one atomic for all of them, then the rows
// by all lanes.
__device__ static void join_write(const DevJoin &jn, const int *par, const int *ent, int &staged,
                                  unsigned long long *count, long long cap, int *out_pair, int *out_ent, int ostride)
{
    if (staged == 0) return;
    wave_sync();
    const int lane = lane_id();
    unsigned long long b0 = 0;
    if (lane == 0) b0 = atomicAdd(count, (unsigned long long)staged);
    const long long base = (long long)__shfl(b0, 0, JAM_WAVE);
    const int L = jn.level;
    // the parent's L entries, then each row's new entry and pair (two
    // passes, so that no load selects between LDS and global memory)
    const int total = staged * L;
    for (int idx = lane; idx < total; idx += JAM_WAVE) {
        const int k = idx / L, l = idx - k * L;
        const long long r = base + k;
        if (r >= cap) continue;
        out_ent[r * ostride + l] = jn.in_ent[(long long)par[k] * jn.in_stride + l];
    }
    for (int k = lane; k < staged; k += JAM_WAVE) {
        const long long r = base + k;
        if (r >= cap) {
            atomicOr(jn.overflow, 1u);
            continue;
        }
        out_ent[r * ostride + L] = ent[k];
        out_pair[r] = jn.in_pair[par[k]];
    }
    wave_sync();
    staged = 0;
}

// ---------------------------------------------------------------------------
// the join, one launch per level: a wave extends JAM_WAVE / TW partial
// matches at once, one per team of TW lanes, so that the dependent loads of
// several rows are in flight together; each team tests its row's candidates
// TW at a time against every annulus and check()
// ---------------------------------------------------------------------------

template <int MAXN>
struct JoinTeamLds {
    double px[MAXN], py[MAXN], pz[MAXN];
    double r2lo[MAXN], r2hi[MAXN];
    // per placed atom: chainID2 and resSeq of the molecule atom, and whether
    // its template atom has the new template atom's chainID2 (bit 0) and
    // resSeq (bit 1)
    int4 pk[MAXN];
};

template <int MAXN, int TW>
struct JoinWaveLds {
    JoinTeamLds<MAXN> tm[JAM_WAVE / TW];
    // children (sp, se) and complete matches (dp, de) waiting to be written:
    // parent row and the new entry
    int sp[JOIN_STAGE], se[JOIN_STAGE];
    int dp[JOIN_STAGE], de[JOIN_STAGE];
};

// What a team keeps of the pair of its last row: the template data of a
// level and the annulus windows of the placed atoms depend only on the pair
// (and the level, fixed in a launch), so the next row of the same pair --
// a team takes runs of nearby rows -- loads none of it again. The windows
// themselves stay in the team's LDS (r2lo, r2hi).
template <int NE>
struct PairCache {
    int p = -1;
    int n = 0, setL = 0, cntL = 0, offL = 0;
    double b[NE];      // outer radius of the window of placed atom tl + jj * TW
    int flags[NE];     // its template atom's chainID2 (bit 0) and resSeq (bit 1) equal the new one's
};

// Row of item `it` when each row is shared by `parts` teams: with one part
// (every frontier of at least one row per team slot) the item is the row;
// otherwise parts <= slots / rows, so the items number at most the team
// slots of the grid and fit in 32 bits: no 64-bit division is needed.
__device__ static inline long long row_of(long long it, int parts)
{
    return parts == 1 ? it : (long long)((unsigned int)it / (unsigned int)parts);
}

// levels up to which a team's lanes read its placed atoms back from LDS to
// form the join box, instead of combining partial boxes by shuffles
#ifndef JOIN_BOX_DIRECT
#define JOIN_BOX_DIRECT 12
#endif
template <int MAXN, int TW>
__device__ static void join_team(const DevMols &ms, const DevTpls &tp, const DevSets &st, const DevPairs &pr,
                                 const DevJoin &jn, const QueryParams &qp, JoinWaveLds<MAXN, TW> &S,
                                 bool active, long long it, int parts, int p, const int (&ent)[(MAXN + TW - 1) / TW],
                                 PairCache<(MAXN + TW - 1) / TW> &pc, int &staged, int &dstaged)
{
    constexpr unsigned long long TMASK = (TW == 64) ? ~0ull : ((1ull << TW) - 1ull);
    const int lane = lane_id();
    const int team = lane / TW, tl = lane % TW;
    JoinTeamLds<MAXN> &M = S.tm[team];
    const int L = jn.level;
    const long long i = active ? row_of(it, parts) : 0;
    const int part = active ? (int)(it - i * parts) : 0;
    constexpr int NE = (MAXN + TW - 1) / TW;
    if (active && p != pc.p) {
        // a new pair (team-uniform): its template data and annulus windows
        const int t = p % tp.n_templates;
        const int tbase = tp.atom_off[t];
        const int n = tp.atom_off[t + 1] - tbase;
        const PairLevel *lv = pr.lvl + (size_t)p * JAM_MAX_LEVELS;
        pc.p = p;
        pc.n = n;
        if (L < n) {
            const PairLevel hL = lv[L];
            const int kL = hL.k;
            pc.setL = hL.set;
            pc.cntL = hL.cnt;
            pc.offL = hL.off;
            const int tchL = tp.chain[tbase + kL], tresL = tp.resseq[tbase + kL];
#pragma unroll
            for (int jj = 0; jj < NE; jj++) {
                const int j = tl + jj * TW;
                if (j >= L) break;
                const int k = lv[j].k;
                pc.flags[jj] = (tp.chain[tbase + k] == tchL ? 1 : 0) | (tp.resseq[tbase + k] == tresL ? 2 : 0);
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
                M.r2lo[j] = a * a;
                M.r2hi[j] = b * b;
                pc.b[jj] = b;
            }
        }
    }
    const int n = pc.n;
    if (active && L >= n && tl == 0 && part == 0) {
        // a one-atom template is complete at level 0
        const unsigned long long r = atomicAdd(jn.n_done, 1ull);
        if ((long long)r < jn.done_cap) {
            jn.done_pair[r] = p;
            jn.done_ent[r * jn.stride] = jn.in_ent[i * jn.in_stride];
        } else {
            atomicOr(jn.overflow, 1u);
        }
    }
    const bool live = active && L < n;
    const int setL = live ? pc.setL : 0, cntL = live ? pc.cntL : 0, offL = live ? pc.offL : 0;
    double lo[3] = {-INFINITY, -INFINITY, -INFINITY}, hi[3] = {INFINITY, INFINITY, INFINITY};
#pragma unroll
    for (int jj = 0; jj < NE; jj++) {
        const int j = tl + jj * TW;
        if (!live || j >= L) break;
        // the atom placed at level j
        const SetEnt en = st.kent[ent[jj]];
        M.px[j] = en.x;
        M.py[j] = en.y;
        M.pz[j] = en.z;
        const double b = pc.b[jj];
        // .w: the outer radius rounded up to float (the join box is a
        // prefilter: a box a little larger finds the same candidates)
        M.pk[j] = make_int4(en.ch, en.res, pc.flags[jj], __float_as_int(__double2float_ru(b)));
        lo[0] = fmax(lo[0], en.x - b);
        lo[1] = fmax(lo[1], en.y - b);
        lo[2] = fmax(lo[2], en.z - b);
        hi[0] = fmin(hi[0], en.x + b);
        hi[1] = fmin(hi[1], en.y + b);
        hi[2] = fmin(hi[2], en.z + b);
    }
    // join box of the team (Join_computeBox, patches/Join.c.patch:154-196):
    // with few placed atoms (the level is the same for every row of a
    // launch) every lane reads them back from the team's LDS, else the lanes
    // combine their partial boxes by shuffles
    if (L <= JOIN_BOX_DIRECT) {
        wave_sync();
        if (live) {
#pragma unroll
            for (int d = 0; d < 3; d++) {
                lo[d] = -INFINITY;
                hi[d] = INFINITY;
            }
            for (int j = 0; j < L; j++) {
                const double x = M.px[j], y = M.py[j], z = M.pz[j];
                const double b = (double)__int_as_float(M.pk[j].w);
                lo[0] = fmax(lo[0], x - b);
                lo[1] = fmax(lo[1], y - b);
                lo[2] = fmax(lo[2], z - b);
                hi[0] = fmin(hi[0], x + b);
                hi[1] = fmin(hi[1], y + b);
                hi[2] = fmin(hi[2], z + b);
            }
        }
    } else {
        for (int sh = TW / 2; sh >= 1; sh >>= 1) {
#pragma unroll
            for (int d = 0; d < 3; d++) {
                lo[d] = fmax(lo[d], __shfl_xor(lo[d], sh, JAM_WAVE));
                hi[d] = fmin(hi[d], __shfl_xor(hi[d], sh, JAM_WAVE));
            }
        }
    }
    wave_sync();
    const int ignore = qp.ignore_chain;
    const long long bbase = (long long)offL / JAM_WAVE + setL;
    const long long sbase = (long long)offL / JAM_SUB + setL;
    const bool last = (L == n - 1);
    const int nblk = live ? (cntL + JAM_WAVE - 1) / JAM_WAVE : 0;
    const int ngroups = (nblk + TW - 1) / TW;
    // a box (lo xyz, hi xyz: three 16-byte loads) meets the join box
    auto meets = [&](const double *box) {
        const double2 *b2 = reinterpret_cast<const double2 *>(box);
        const double2 u = b2[0], v = b2[1], w = b2[2];   // (lo0, lo1), (lo2, hi0), (hi1, hi2)
        return !((u.x > hi[0]) | (lo[0] > v.y) | (u.y > hi[1]) | (lo[1] > w.x) | (v.x > hi[2]) | (lo[2] > w.y));
    };
    // one candidate per lane: the annuli of every placed atom and check(),
    // then the survivors staged (children, or complete matches at the last
    // level). Branch-free: every lane evaluates every test.
    auto test = [&](bool ok, int q) {
        SetEnt c{0.0, 0.0, 0.0, 0, 0};
        if (ok) c = st.kent[offL + q];
        for (int j = 0; j < L; j++) {
            // most candidates fail an early annulus: stop as soon as no lane
            // of the wave is still in
            if (ballot(ok) == 0ull) break;
            // _Annulus_po: squared distance summed x, y, z from 0.0; inside
            // unless below the inner or above the outer squared radius
            const double tx = M.px[j] - c.x, ty = M.py[j] - c.y, tz = M.pz[j] - c.z;
            double s2 = 0.0;
            s2 += tx * tx;
            s2 += ty * ty;
            s2 += tz * tz;
            const bool ann = !(s2 < M.r2lo[j]) & !(s2 > M.r2hi[j]);
            // TessTemplate_check (patches/TessTemplate.c.patch:101-156): chain
            // and residue equality must match the template's; with
            // ignore_chain none the residue test applies within a chain only,
            // with "residues" a residue of another chain does not match
            const int4 k4 = M.pk[j];
            const bool sc = k4.x == c.ch, sr = k4.y == c.res;
            const bool tsc = (k4.z & 1) != 0, tsr = (k4.z & 2) != 0;
            const bool res_ok = sr == tsr;
            const bool chk = ignore == 0   ? ((sc == tsc) & (!sc | res_ok))
                             : ignore == 1 ? (res_ok & !(sr & !sc))
                                           : res_ok;
            ok = ok & ann & chk;
        }
        const unsigned long long cm = ballot(ok && !last), dm = ballot(ok && last);
        if (cm) {
            const int cnt = __popcll(cm);
            if (staged + cnt > JOIN_STAGE) join_write(jn, S.sp, S.se, staged, jn.n_out, jn.out_cap, jn.out_pair,
                                                      jn.out_ent, jn.out_stride);
            if (ok && !last) {
                const int r = rank_below(cm);
                S.sp[staged + r] = (int)i;
                S.se[staged + r] = offL + q;
            }
            staged += cnt;
        }
        if (dm) {
            const int cnt = __popcll(dm);
            if (dstaged + cnt > JOIN_STAGE) join_write(jn, S.dp, S.de, dstaged, jn.n_done, jn.done_cap,
                                                       jn.done_pair, jn.done_ent, jn.stride);
            if (ok && last) {
                const int r = rank_below(dm);
                S.dp[dstaged + r] = (int)i;
                S.de[dstaged + r] = offL + q;
            }
            dstaged += cnt;
        }
    };
    static_assert(TW == JAM_SUB, "a team tests one sub-block of candidates at a time");
    constexpr int SUBS = JAM_WAVE / JAM_SUB;   // sub-blocks per block
    constexpr int BPR = TW / SUBS;             // blocks whose sub-blocks one round tests
    int seen = 0;   // in-box blocks met so far; this team takes every parts-th one
    for (int g = 0;; g++) {
        if (ballot(g < ngroups) == 0ull) break;
        // blocks of this group whose box meets the join box (all of them for
        // small sets: the test is a prefilter only)
        bool inb = false;
        const int b = g * TW + tl;
        if (g < ngroups && b < nblk) inb = nblk <= JOIN_SMALL || meets(st.bbox + 6 * (bbase + b));
        unsigned long long bm = (ballot(inb) >> (team * TW)) & TMASK;
        if (parts > 1) {
            // keep the in-box blocks whose running index is part mod parts
            unsigned long long keep = 0ull, rest = bm;
            while (rest) {
                const int x = first_bit(rest);
                rest &= rest - 1;
                if (seen % parts == part) keep |= 1ull << x;
                seen++;
            }
            bm = keep;
        }
        for (;;) {
            if (ballot(bm != 0ull) == 0ull) break;
            // lane tl tests the box of sub-block tl % SUBS of the
            // (tl / SUBS)-th remaining in-box block
            unsigned long long m = bm;
            for (int k = 0; k < tl / SUBS; k++) m &= m - 1;
            for (int k = 0; k < BPR; k++) bm &= bm - 1;
            const int q0 = m ? (g * TW + first_bit(m)) * JAM_WAVE + (tl % SUBS) * JAM_SUB : 0;
            const bool sin = m && q0 < cntL && meets(st.sbox + 6 * (sbase + q0 / JAM_SUB));
            unsigned long long sm = (ballot(sin) >> (team * TW)) & TMASK;
            // the in-box sub-blocks, one at a time: a candidate per lane
            for (;;) {
                const bool has = sm != 0ull;
                if (ballot(has) == 0ull) break;
                int x = 0;
                if (has) {
                    x = first_bit(sm);
                    sm &= sm - 1;
                }
                const int q = __shfl(q0, team * TW + x, JAM_WAVE) + tl;
                test(has && q < cntL, q);
            }
        }
    }
    wave_sync();
}

// Waves loop over the input rows (the grid is what stays resident on the
// device); the output frontier's counter is also the next level's input
// size, so levels run back to back without the host.
// Workgroups are dealt to the XCDs round-robin (workgroup b runs on XCD
// b % JOIN_XCDS), and each XCD has its own L2: the rows are cut into chunks
// of one XCD's team slots, chunk c going to XCD c % JOIN_XCDS, so that
// neighbouring rows -- mostly the same pair, the same candidate sets --
// share one L2.
#ifndef JOIN_XCDS
#define JOIN_XCDS 8
#endif
#ifndef JOIN_RUN
#define JOIN_RUN 32
#endif
template <int MAXN, int TW>
__global__ void __launch_bounds__(JOIN_WAVES * JAM_WAVE, JOIN_MIN_WAVES) k_join(DevMols ms, DevTpls tp, DevSets st,
                                                                              DevPairs pr, DevJoin jn,
                                                                              QueryParams qp)
{
    constexpr int NTEAM = JAM_WAVE / TW;
    __shared__ JoinWaveLds<MAXN, TW> SW[JOIN_WAVES];
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x / JAM_WAVE);
    const int team = lane_id() / TW;
    long long n_in = (long long)*jn.n_in;
    if (n_in > jn.in_cap) n_in = jn.in_cap;
    int staged = 0, dstaged = 0;
    const long long slots = (long long)gridDim.x * JOIN_WAVES * NTEAM;
    // a frontier smaller than the team slots is spread thinner: each row's
    // in-box blocks are shared by `parts` teams
    const int parts = (n_in > 0 && n_in < slots) ? (int)min((long long)JOIN_SPLIT, slots / n_in) : 1;
    const long long items = n_in * parts;
    const int xcds = (gridDim.x % JOIN_XCDS == 0) ? JOIN_XCDS : 1;
    const int xcd = (int)(blockIdx.x % xcds);
    const long long per = slots / xcds;   // items of one chunk
    const long long tslot = ((long long)(blockIdx.x / xcds) * JOIN_WAVES + w) * NTEAM;
    // the first item of this wave's teams in round i: rounds come in runs of
    // R over consecutive rows (so that a team meets rows of one pair one
    // after the other), when the frontier is large enough for every wave to
    // have several runs
    static_assert((JOIN_RUN & (JOIN_RUN - 1)) == 0, "runs of a power of two rounds");
    const int rsh = items >= 4 * slots * (long long)JOIN_RUN ? __builtin_ctz(JOIN_RUN) : 0;   // log2 of the run
    auto round_base = [&](long long i) {
        const long long k = i >> rsh, j = i & ((1ll << rsh) - 1);
        return (((k * xcds + xcd) * per + tslot) << rsh) + j * NTEAM;
    };
    // the pair and this lane's entries of the team's next row are loaded
    // while the current one is processed
    constexpr int NE = (MAXN + TW - 1) / TW;
    const int tl = lane_id() % TW;
    const int L = jn.level;
    int p_next = 0, e_next[NE];
    auto fetch = [&](long long it) {
#pragma unroll
        for (int jj = 0; jj < NE; jj++) e_next[jj] = 0;
        if (it >= items) return;
        const long long r = row_of(it, parts);
        p_next = jn.in_pair[r];
#pragma unroll
        for (int jj = 0; jj < NE; jj++) {
            const int j = tl + jj * TW;
            if (j < L) e_next[jj] = jn.in_ent[r * jn.in_stride + j];
        }
    };
    PairCache<NE> pc;
    long long base = round_base(0);
    if (base < items) fetch(base + team);
    for (long long i = 0; base < items; i++) {
        const long long it = base + team;
        const int p = p_next;
        int e[NE];
#pragma unroll
        for (int jj = 0; jj < NE; jj++) e[jj] = e_next[jj];
        const long long next = round_base(i + 1);
        if (next < items) fetch(next + team);
        join_team<MAXN, TW>(ms, tp, st, pr, jn, qp, SW[w], it < items, it, parts, p, e, pc, staged, dstaged);
        base = next;
    }
    join_write(jn, SW[w].sp, SW[w].se, staged, jn.n_out, jn.out_cap, jn.out_pair, jn.out_ent, jn.out_stride);
    join_write(jn, SW[w].dp, SW[w].de, dstaged, jn.n_done, jn.done_cap, jn.done_pair, jn.done_ent, jn.stride);
}

// ---------------------------------------------------------------------------
// reference order of the complete matches
// ---------------------------------------------------------------------------

__global__ void __launch_bounds__(256) k_iota(int *v, long long n)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) v[i] = (int)i;
}

// Sort key of complete match i for levels [l0, l0 + per): the positions of
// the match's atoms in their sets (list position at level 0, kd position
// above), `bits` bits each, the earliest level most significant; levels past
// the template's size count as 0 (all matches of a pair have the same size).
__device__ static inline unsigned long long level_key(const DevSets &st, const DevTpls &tp, const DevPairs &pr,
                                                      const DevJoin &jn, int i, int l0, int per, int bits)
{
    const int p = jn.done_pair[i];
    const int t = p % tp.n_templates;
    const int nt = tp.atom_off[t + 1] - tp.atom_off[t];
    const int *row = jn.done_ent + (long long)i * jn.stride;
    unsigned long long k = 0;
    for (int q = 0; q < per; q++) {
        const int l = l0 + q;
        unsigned long long pos = 0;
        if (l == 0) {
            pos = (unsigned long long)st.kidx[row[0]];
        } else if (l < nt) {
            pos = (unsigned long long)(row[l] - pr.lvl[(size_t)p * JAM_MAX_LEVELS + l].off);
        }
        k = (k << bits) | pos;
    }
    return k;
}

// Last sort key: the molecule, then the template in scan order (Jess scans
// templates in reverse insertion order, patches/Jess.c.patch:63-72), so the
// complete matches end up grouped by pair in the order the caller reports
// them.
__device__ static inline unsigned int pair_key(const DevJoin &jn, int n_templates, int i)
{
    const int p = jn.done_pair[i];
    const int t = p % n_templates;
    return (unsigned int)(p - t + (n_templates - 1 - t));
}

__global__ void __launch_bounds__(256) k_sortkey(DevSets st, DevTpls tp, DevPairs pr, DevJoin jn, const int *perm,
                                                 long long n, int l0, int per, int bits, unsigned long long *key)
{
    const long long s = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    key[s] = level_key(st, tp, pr, jn, perm[s], l0, per, bits);
}

// Template size (atoms) of complete match s, the key that orders the matches
// by how many level groups their sort keys reach.
__global__ void __launch_bounds__(256) k_sizekey(DevTpls tp, DevJoin jn, long long n, unsigned int *key)
{
    const long long s = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    const int t = jn.done_pair[s] % tp.n_templates;
    key[s] = (unsigned int)(tp.atom_off[t + 1] - tp.atom_off[t]);
}

// starts[l] (l = 0 ... maxl): how many of the n sorted sizes are at most l.
__global__ void k_size_starts(const unsigned int *sorted, long long n, int maxl, long long *starts)
{
    const int l = (int)threadIdx.x;
    if (l > maxl) return;
    long long lo = 0, hi = n;
    while (lo < hi) {
        const long long mid = lo + (hi - lo) / 2;
        if (sorted[mid] <= (unsigned int)l) lo = mid + 1;
        else hi = mid;
    }
    starts[l] = lo;
}

__global__ void __launch_bounds__(256) k_pairkey(DevJoin jn, int n_templates, const int *perm, long long n,
                                                 unsigned int *key)
{
    const long long s = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    key[s] = pair_key(jn, n_templates, perm[s]);
}

// Small sorts (at most SORT_RANK_MAX complete matches whose full key fits
// SORT_RANK_WORDS words) go by rank counting instead of radix passes: a few
// launches whatever the number of levels.
#ifndef SORT_RANK_MAX
#define SORT_RANK_MAX 8192
#endif
#define SORT_RANK_WORDS 4

// Full sort key of complete match i as a W-word big-endian bit string: the
// pair key (pbits bits), then the level positions of k_sortkey (bits bits
// each, level 0 first). Clears the match's rank.
template <int W>
__global__ void __launch_bounds__(256) k_fullkey(DevSets st, DevTpls tp, DevPairs pr, DevJoin jn, long long n,
                                                 int n_levels, int bits, int pbits, unsigned long long *key,
                                                 int *rank)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    unsigned long long k[W];
#pragma unroll
    for (int w = 0; w < W; w++) k[w] = 0;
    auto push = [&](unsigned long long v, int nb) {
#pragma unroll
        for (int w = 0; w < W - 1; w++) k[w] = (k[w] << nb) | (k[w + 1] >> (64 - nb));
        k[W - 1] = (k[W - 1] << nb) | v;
    };
    push(pair_key(jn, tp.n_templates, (int)i), pbits);
    for (int l = 0; l < n_levels; l++) push(level_key(st, tp, pr, jn, (int)i, l, 1, bits), bits);
#pragma unroll
    for (int w = 0; w < W; w++) key[w * n + i] = k[w];
    rank[i] = 0;
}

// rank[i] += the matches of a SORT_RANK_TJ-wide tile of j that sort before
// match i (smaller key, or equal key and j < i: the order of a stable sort).
// Narrow j tiles give many workgroups; the comparisons are branch-free so
// that the tile reads of several j are in flight at once.
#define SORT_RANK_TJ 64
template <int W>
__global__ void __launch_bounds__(256) k_rank(const unsigned long long *key, long long n, int *rank)
{
    __shared__ unsigned long long tile[W][SORT_RANK_TJ];
    const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
    const long long j0 = (long long)blockIdx.y * SORT_RANK_TJ;
    if (threadIdx.x < SORT_RANK_TJ) {
        const long long jt = j0 + threadIdx.x;
#pragma unroll
        for (int w = 0; w < W; w++) tile[w][threadIdx.x] = jt < n ? key[w * n + jt] : ~0ull;
    }
    __syncthreads();
    if (i >= n) return;
    unsigned long long ki[W];
#pragma unroll
    for (int w = 0; w < W; w++) ki[w] = key[w * n + i];
    // slots past n hold all-ones keys and j >= n: never before i
    const long long lim = n - j0;
    int cnt = 0;
#pragma unroll 8
    for (int jj = 0; jj < SORT_RANK_TJ; jj++) {
        bool lt = jj < lim && j0 + jj < i;   // key equal so far: the index decides
#pragma unroll
        for (int w = W - 1; w >= 0; w--) {
            const unsigned long long kj = tile[w][jj];
            lt = (kj < ki[w]) || (kj == ki[w] && lt);
        }
        cnt += lt ? 1 : 0;
    }
    if (cnt) atomicAdd(&rank[i], cnt);
}

__global__ void __launch_bounds__(256) k_rank_scatter(const int *rank, long long n, int *perm)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) perm[rank[i]] = (int)i;
}

// first sorted position of every pair with complete matches
__global__ void __launch_bounds__(256) k_first(DevJoin jn, const int *perm, long long n, long long *first)
{
    const long long s = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    const int p = jn.done_pair[perm[s]];
    if (s == 0 || jn.done_pair[perm[s - 1]] != p) first[p] = s;
}

// complete matches per pair
__global__ void __launch_bounds__(256) k_pair_matches(DevJoin jn, const int *perm, long long n,
                                                      const long long *first, long long *out)
{
    const long long s = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    const int p = jn.done_pair[perm[s]];
    if (s == n - 1 || jn.done_pair[perm[s + 1]] != p) out[p] = s - first[p] + 1;
}

// ---------------------------------------------------------------------------
// superposition + events
// ---------------------------------------------------------------------------

// Superposition of complete match s (sorted position): pair, molecule-local
// atoms in template order, RMSD, rotation, centroids.
__device__ __attribute__((always_inline)) static inline int super_of(DevMols ms, DevTpls tp, DevSets st,
                                                                     DevPairs pr, DevJoin jn, const int *perm,
                                                                     long long s, int atoms[JAM_MAX_LEVELS],
                                                                     int *n_out, double *rmsd, double rot[9],
                                                                     double c0[3], double c1[3])
{
    const int i = perm[s];
    const int pair = jn.done_pair[i];
    const int T = tp.n_templates;
    const int m = pair / T, t = pair % T;
    const int tbase = tp.atom_off[t];
    const int n = tp.atom_off[t + 1] - tbase;
    const int abase = ms.atom_off[m];
    const unsigned char *ord = pr.ord + (size_t)pair * JAM_MAX_LEVELS;
    const int *row = jn.done_ent + (long long)i * jn.stride;
    // molecule-local atom of template atom k: placed at the level l with ord[l] == k
    for (int l = 0; l < n; l++) atoms[ord[l]] = st.katom[row[l]];
    auto pa = [&](int k, int d) { return ms.xyz[3 * (size_t)(abase + atoms[k]) + d]; };
    auto pb = [&](int k, int d) { return tp.pos[3 * (size_t)(tbase + k) + d]; };
    superpose(n, pa, pb, rmsd, rot, c0, c1);
    *n_out = n;
    return pair;
}

// Thread per complete match, in reference order. Pass 1 (EMIT = false)
// marks the matches within the RMSD threshold (sel[s] = 1); after an
// exclusive scan of sel into pos, pass 2 (EMIT = true) recomputes the
// superposition of the marked ones and writes their events at pos[s], so the
// events keep the order of the sorted matches without a sort of their own.
// Each event carries the match's rank within its pair.
// With keep_d / keep_i (few matches), pass 1 also keeps each superposition
// (16 doubles, then pair, size and atoms as stride + 2 ints per match) and
// pass 2 reads it back instead of computing it again.
template <bool EMIT>
__global__ void __launch_bounds__(256) k_super(DevMols ms, DevTpls tp, DevSets st, DevPairs pr, DevJoin jn,
                                               DevEvents ev, QueryParams qp, const int *perm, long long n_matches,
                                               const long long *first, int *sel, const int *pos, double *keep_d,
                                               int *keep_i)
{
    const long long s = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n_matches) return;
    if (EMIT && !sel[s]) return;
    int atoms[JAM_MAX_LEVELS];
    int n, pair;
    double rmsd, rot[9], c0[3], c1[3];
    double *kd = keep_d ? keep_d + 16 * s : nullptr;
    int *ki = keep_i ? keep_i + (size_t)(ev.stride + 2) * s : nullptr;
    if (EMIT && kd) {
        rmsd = kd[0];
#pragma unroll
        for (int k = 0; k < 9; k++) rot[k] = kd[1 + k];
#pragma unroll
        for (int d = 0; d < 3; d++) {
            c0[d] = kd[10 + d];
            c1[d] = kd[13 + d];
        }
        pair = ki[0];
        n = ki[1];
        for (int k = 0; k < n; k++) atoms[k] = ki[2 + k];
    } else {
        pair = super_of(ms, tp, st, pr, jn, perm, s, atoms, &n, &rmsd, rot, c0, c1);
    }
    if (!EMIT) {
        sel[s] = (rmsd <= qp.rmsd_threshold) ? 1 : 0;
        if (kd) {
            kd[0] = rmsd;
#pragma unroll
            for (int k = 0; k < 9; k++) kd[1 + k] = rot[k];
#pragma unroll
            for (int d = 0; d < 3; d++) {
                kd[10 + d] = c0[d];
                kd[13 + d] = c1[d];
            }
            ki[0] = pair;
            ki[1] = n;
            for (int k = 0; k < n; k++) ki[2 + k] = atoms[k];
        }
        return;
    }
    bool nan = false;
#pragma unroll
    for (int k = 0; k < 9; k++) nan |= isnan(rot[k]);
    const long long e = pos[s];
    if (e >= ev.cap) return;
    const EventLayout L = event_layout((int64_t)*ev.count, ev.stride);
    ((int *)(ev.base + L.pair))[e] = pair;
    ((long long *)(ev.base + L.rank))[e] = s - first[pair];
    ((int *)(ev.base + L.flag))[e] = nan ? 2 : 1;
    ((double *)(ev.base + L.rmsd))[e] = rmsd;
    double *erot = (double *)(ev.base + L.rot) + 9 * e;
    double *ec0 = (double *)(ev.base + L.c0) + 3 * e;
    double *ec1 = (double *)(ev.base + L.c1) + 3 * e;
#pragma unroll
    for (int k = 0; k < 9; k++) erot[k] = rot[k];
#pragma unroll
    for (int d = 0; d < 3; d++) {
        ec0[d] = c0[d];
        ec1[d] = c1[d];
    }
    int *eat = (int *)(ev.base + L.atoms) + e * (size_t)ev.stride;
    for (int k = 0; k < ev.stride; k++) eat[k] = k < n ? atoms[k] : -1;
}

// number of events: the scan's last position plus the last mark
__global__ void k_event_total(const int *sel, const int *pos, long long n, unsigned long long *count)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) *count = n > 0 ? (unsigned long long)(pos[n - 1] + sel[n - 1]) : 0ull;
}

// ---------------------------------------------------------------------------
// hit-set digests (measurement)
// ---------------------------------------------------------------------------

// SplitMix64's finaliser (Steele, Lea & Flood, "Fast splittable pseudorandom
// number generators", 2014): a 64-bit mix whose sum over a set of hits is an
// order-independent digest of that set.
__host__ __device__ static inline unsigned long long mix64(unsigned long long x)
{
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

// Thread per event: hits (flag 1) add one to their pair's count and
// mix64-chain(template << 32 ^ molecule, atoms...) to its digest; the
// molecule is numbered from mol_base. Rows end at the first -1 (the
// template's size).
__global__ void __launch_bounds__(256) k_pair_digest(const unsigned char *base, long long n, int stride,
                                                     int n_templates, int mol_base, unsigned long long *counts,
                                                     unsigned long long *digests)
{
    const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n) return;
    const EventLayout L = event_layout((int64_t)n, stride);
    if (((const int *)(base + L.flag))[e] != 1) return;
    const int p = ((const int *)(base + L.pair))[e];
    const unsigned int t = (unsigned int)(p % n_templates), mi = (unsigned int)(p / n_templates + mol_base);
    unsigned long long h = mix64(((unsigned long long)t << 32) ^ (unsigned long long)mi);
    const int *at = (const int *)(base + L.atoms) + (size_t)e * stride;
    for (int k = 0; k < stride && at[k] >= 0; k++) h = mix64(h ^ (unsigned long long)(unsigned int)at[k]);
    atomicAdd(&counts[p], 1ull);
    atomicAdd(&digests[p], h);
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
