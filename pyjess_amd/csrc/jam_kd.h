// jam_kd.h -- kd-tree leaf order of every candidate set, on gfx950.
//
// The reference builds one kd-tree per candidate set (KdTree_create,
// patches/KdTree.c.patch:348-451): the root splits the set at its median on
// x, the children on y, then z, ..., each median found by QuickSelect
// (qselect_r, patches/qselect.h.patch:50-116): median-of-three pivot at
// mid = (l + r + 1) / 2 moved to r, Lomuto partition sending "<= pivot" to the
// left, pivot swapped into place, repeat on the side holding the median. The
// Join walks the tree depth first, right child first, so the order in which
// candidates are met -- and therefore the order of the hits -- is the tree's
// leaf order read backwards. The kernels here compute that order for every
// set and store the set in it (32-byte entries: x, y, z, chain, resSeq; and
// the atom index) with one bounding box per 64 consecutive entries, which the
// join uses for pruning.
//
// Two algorithms, both exact:
//   fast   when no range of the tree has equal coordinates on both sides of
//          its split point, the left half of every range is its n/2 smallest
//          elements whatever the arrangement, and the leaf order follows from
//          memberships alone: sort every axis once, then at level d cut each
//          range of axis d%3 in half and stably partition the two other axes
//          by side. Detects a tie across a split point and then gives up.
//   exact  for the sets the fast path gave up on: all ranges of a tree level
//          run their QuickSelect together, one partition round at a time,
//          and each Lomuto partition is computed without its serial scan:
//            * the "<= pivot" elements keep their order: a stable compaction;
//            * the others move as a FIFO queue: each is pushed when scanned,
//              and each "<=" element scanned after the first ">" one pops the
//              front and pushes it back; push number t of such a re-push is
//              the q-th pop, i.e. push number q < t. Pointer jumping resolves
//              every push to its element. The final queue fills
//              [l+k+1, r-1] from its second element on, its front goes to r,
//              the pivot to l+k.
//
// Sets are bucketed by size; sets of up to 64 entries run in one wave's
// registers, up to KD_BIG entries with one position per thread (shuffles,
// ballots and LDS exchanges), up to KD_LDS_MAX entries straight to the exact
// pass with 8 positions per thread in LDS (most sets that large have a tie),
// larger ones in global scratch. Every loop of the
// exact pass whose length depends on the data has a bound that a correct
// chain never reaches; a pass that reaches it sets st.kd_err and leaves the
// set unwritten, and the query reports an error.
#pragma once

#include <hip/hip_runtime.h>

#include "jam_common.h"
#include "jam_wave.h"

namespace jam {

#define KD_BIG 1024
#define KD_TIERS 6
#ifndef KD_THREADS
#define KD_THREADS 256
#endif
// threads of the exact pass over a set larger than KD_BIG: its rounds are
// latency-bound loops over the positions, so more threads per set
#define KD_EXACT_THREADS 1024
// global scratch of sets larger than KD_BIG, per set entry (padded to a power
// of two for the fast path's sorts: 3 doubles, 9 ints and 1 byte per padded
// entry, at most 2 * 61 bytes per entry)
#define KD_SCRATCH_BYTES_PER_ENTRY 128

// size tier of a set: (0,64], (64,128], (128,256], (256,512], (512,1024],
// larger; -1 for an empty set
__host__ __device__ static inline int kd_tier(int n)
{
    if (n <= 0) return -1;
    if (n <= 64) return 0;
    if (n <= 128) return 1;
    if (n <= 256) return 2;
    if (n <= 512) return 3;
    if (n <= KD_BIG) return 4;
    return 5;
}

__device__ static inline int shfl_int(int v, int src) { return __shfl(v, src, JAM_WAVE); }

// lane `dst` receives v from this lane (ds_permute: push)
__device__ static inline int push_int(int v, int dst) { return __builtin_amdgcn_ds_permute(dst << 2, v); }

__device__ static inline bool key_less(double a, int ia, double b, int ib)
{
    return a < b || (a == b && ia < ib);
}

// Buckets the sets by size tier: bucket[k * n_sets + i] is the i-th set of
// tier k (in no particular order), nb[k] their number. Each workgroup takes
// one contiguous span of sets: it counts its sets per tier in LDS (one LDS
// atomic per wave and tier), reserves room in each tier's list with one
// global atomic per tier, then places its sets there.
#define KD_BUCKET_SPAN 4096
__global__ void __launch_bounds__(256) k_kdbucket(DevSets st, int *bucket, unsigned int *nb)
{
    __shared__ unsigned int cnt[KD_TIERS], base[KD_TIERS];
    const int lane = lane_id();
    const int t = threadIdx.x;
    for (long long s0 = (long long)blockIdx.x * KD_BUCKET_SPAN; s0 < st.n_sets;
         s0 += (long long)gridDim.x * KD_BUCKET_SPAN) {
        const long long s1 = min(s0 + (long long)KD_BUCKET_SPAN, (long long)st.n_sets);
        if (t < KD_TIERS) cnt[t] = 0u;
        __syncthreads();
        for (long long s = s0 + t; s - lane < s1; s += blockDim.x) {
            const int tier = s < s1 ? kd_tier(st.cnt[s]) : -1;
            for (int k = 0; k < KD_TIERS; k++) {
                const unsigned long long mask = ballot(tier == k);
                if (mask != 0ull && lane == first_bit(mask)) atomicAdd(&cnt[k], (unsigned int)__popcll(mask));
            }
        }
        __syncthreads();
        if (t < KD_TIERS) {
            base[t] = cnt[t] ? atomicAdd(&nb[t], cnt[t]) : 0u;
            cnt[t] = 0u;
        }
        __syncthreads();
        for (long long s = s0 + t; s - lane < s1; s += blockDim.x) {
            const int tier = s < s1 ? kd_tier(st.cnt[s]) : -1;
            for (int k = 0; k < KD_TIERS; k++) {
                const unsigned long long mask = ballot(tier == k);
                if (mask == 0ull) continue;
                const int leader = first_bit(mask);
                unsigned int off = 0u;
                if (lane == leader) off = atomicAdd(&cnt[k], (unsigned int)__popcll(mask));
                off = (unsigned int)shfl_int((int)off, leader);
                if (tier == k) bucket[(size_t)k * st.n_sets + base[k] + off + rank_below(mask)] = (int)s;
            }
        }
        __syncthreads();
    }
}

// Writes position p of set s (element e of its candidate list, p < n) and,
// with the other lanes of the wave holding positions p & ~63 ... p | 63, the
// bounding boxes of that 64-entry block and of its JAM_SUB-entry sub-blocks.
// Every lane of the wave calls it.
__device__ static inline void kd_emit_lane(const DevMols &ms, const DevSets &st, int s, int n, int64_t off,
                                           int abase, int p, int e)
{
    const bool ok = p < n;
    double x = 0.0, y = 0.0, z = 0.0;
    if (ok) {
        const int atom = st.list[off + e];
        const size_t g = (size_t)(abase + atom);
        x = ms.xyz[3 * g + 0];
        y = ms.xyz[3 * g + 1];
        z = ms.xyz[3 * g + 2];
        SetEnt en;
        en.x = x;
        en.y = y;
        en.z = z;
        en.ch = ms.chain[g];
        en.res = ms.resseq[g];
        st.kent[off + p] = en;
        st.katom[off + p] = atom;
        st.kidx[off + p] = e;
        st.l2k[off + e] = p;
    }
    double blo[3] = {ok ? x : INFINITY, ok ? y : INFINITY, ok ? z : INFINITY};
    double bhi[3] = {ok ? x : -INFINITY, ok ? y : -INFINITY, ok ? z : -INFINITY};
    // the box of each JAM_SUB entries, then of the 64
    for (int sh = JAM_SUB / 2; sh >= 1; sh >>= 1) {
#pragma unroll
        for (int q = 0; q < 3; q++) {
            blo[q] = fmin(blo[q], __shfl_xor(blo[q], sh, JAM_WAVE));
            bhi[q] = fmax(bhi[q], __shfl_xor(bhi[q], sh, JAM_WAVE));
        }
    }
    if ((p & (JAM_SUB - 1)) == 0 && p < n) {
        double *sb = st.sbox + 6 * (off / JAM_SUB + s + p / JAM_SUB);
        sb[0] = blo[0];
        sb[1] = blo[1];
        sb[2] = blo[2];
        sb[3] = bhi[0];
        sb[4] = bhi[1];
        sb[5] = bhi[2];
    }
    for (int sh = JAM_WAVE / 2; sh >= JAM_SUB; sh >>= 1) {
#pragma unroll
        for (int q = 0; q < 3; q++) {
            blo[q] = fmin(blo[q], __shfl_xor(blo[q], sh, JAM_WAVE));
            bhi[q] = fmax(bhi[q], __shfl_xor(bhi[q], sh, JAM_WAVE));
        }
    }
    const int b0 = p & ~(JAM_WAVE - 1);
    if ((p & (JAM_WAVE - 1)) == 0 && b0 < n) {
        double *bb = st.bbox + 6 * (off / JAM_WAVE + s + b0 / JAM_WAVE);
        bb[0] = blo[0];
        bb[1] = blo[1];
        bb[2] = blo[2];
        bb[3] = bhi[0];
        bb[4] = bhi[1];
        bb[5] = bhi[2];
    }
}

// ---------------------------------------------------------------------------
// fast path, sets of at most 64 entries: one wave, one element per lane,
// everything in registers
// ---------------------------------------------------------------------------

// Returns true (nothing written) when a tie straddles a split point.
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
                const bool take = (lower == up) ? other_less : !other_less;
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
        if (ballot(split && prev == here) != 0ull) return true;
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
    kd_emit_lane(ms, st, s, n, off, abase, lane, e);
    return false;
}

__global__ void __launch_bounds__(JAM_WAVE) k_kdwave(DevMols ms, DevTpls tp, DevSets st, const int *list,
                                                     const unsigned int *count, int *tied, unsigned int *n_tied)
{
    if (*st.total > st.cap) return;
    const int cnt = (int)*count;
    for (int q = blockIdx.x; q < cnt; q += gridDim.x) {
        const int s = list[q];
        const int n = st.cnt[s];
        const int64_t off = st.off[s];
        const int abase = ms.atom_off[s / tp.n_types];
        if (kd_wave64(ms, st, s, n, off, abase) && lane_id() == 0) tied[atomicAdd(n_tied, 1u)] = s;
    }
}

// ---------------------------------------------------------------------------
// fast path, sets of (NT/2, NT] entries (NT = 128 ... KD_BIG): one position
// per thread, the arrangement of every axis held in registers (position t:
// key and element); bitonic stages less than 64 positions apart by shuffles,
// the others through LDS; side flags and the stable partitions' prefix counts
// by ballots and one LDS word per wave
// ---------------------------------------------------------------------------

template <int NT>
__global__ void __launch_bounds__(NT) k_kdblock(DevMols ms, DevTpls tp, DevSets st, const int *list,
                                                const unsigned int *count, int *tied, unsigned int *n_tied)
{
    constexpr int NW = NT / JAM_WAVE;
    // the two axes other than a level's are partitioned together: one
    // exchange buffer and prefix array each
    __shared__ double xkey[NT], xkey2[NT];
    __shared__ int xid[NT], xid2[NT];
    __shared__ int xpre[NT], xpre2[NT];
    __shared__ int side[NT];
    __shared__ int wsum[NW], wsum2[NW];
    const int t = threadIdx.x;
    const int lane = lane_id();
    const int wv = t / JAM_WAVE;
    if (*st.total > st.cap) return;
    const int cnt = (int)*count;
    for (int q = blockIdx.x; q < cnt; q += gridDim.x) {
        const int s = list[q];
        const int n = st.cnt[s];
        const int64_t off = st.off[s];
        const int abase = ms.atom_off[s / tp.n_types];
        double c[3] = {INFINITY, INFINITY, INFINITY};
        if (t < n) {
            const size_t g = (size_t)(abase + st.list[off + t]);
            c[0] = ms.xyz[3 * g + 0];
            c[1] = ms.xyz[3 * g + 1];
            c[2] = ms.xyz[3 * g + 2];
        }
        // per axis: (key, element) at position t, sorted by (key, element);
        // positions n ... NT-1 hold padding elements with infinite keys
        double key[3];
        int id[3];
        // exchanges across waves go through LDS, the steps taking the two
        // buffer pairs in turn: a buffer is written again only after a
        // barrier that every read of it precedes, so one barrier a step
        bool second = false;
#pragma unroll
        for (int a = 0; a < 3; a++) {
            double k0 = c[a];
            int i0 = t;
            for (int k = 2; k <= NT; k <<= 1) {
                for (int j = k >> 1; j > 0; j >>= 1) {
                    double ok;
                    int oi;
                    if (j >= JAM_WAVE) {
                        double *bk = second ? xkey2 : xkey;
                        int *bi = second ? xid2 : xid;
                        second = !second;
                        bk[t] = k0;
                        bi[t] = i0;
                        __syncthreads();
                        ok = bk[t ^ j];
                        oi = bi[t ^ j];
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
        // Per level, five barriers: the split side of every element; the tie
        // test folded into the publication of both axes' ballot counts;
        // their prefix counts; their stable partitions; and whether any
        // range is still open folded into the last.
        int sb = 0, se = n;
        bool tie = false;
        bool open = __syncthreads_or(t < n && n >= 2);
        for (int d = 0; open; d++) {
            const int a = d % 3;
            const int b1 = a == 0 ? 1 : 0, b2 = a == 2 ? 1 : 2;   // the two other axes
            const bool act = t < n && se - sb >= 2;
            const int m = sb + (se - sb) / 2;
            const double ka = (a == 0) ? key[0] : (a == 1) ? key[1] : key[2];
            const int ia = (a == 0) ? id[0] : (a == 1) ? id[1] : id[2];
            xkey[t] = ka;
            side[ia] = (act && t < m) ? 1 : 0;
            __syncthreads();
            const bool my_tie = act && t == m && xkey[t - 1] == ka;
            const int i1 = (b1 == 0) ? id[0] : id[1];
            const int i2 = (b2 == 1) ? id[1] : id[2];
            const double k1 = (b1 == 0) ? key[0] : key[1];
            const double k2 = (b2 == 1) ? key[1] : key[2];
            const int sd1 = side[i1], sd2 = side[i2];
            const unsigned long long bal1 = ballot(act && sd1), bal2 = ballot(act && sd2);
            if (lane == 0) {
                wsum[wv] = __popcll(bal1);
                wsum2[wv] = __popcll(bal2);
            }
            if (__syncthreads_or(my_tie)) {
                tie = true;
                break;
            }
            int pre1 = rank_below(bal1), pre2 = rank_below(bal2);
            for (int w = 0; w < wv; w++) {
                pre1 += wsum[w];
                pre2 += wsum2[w];
            }
            xpre[t] = pre1;
            xpre2[t] = pre2;
            __syncthreads();
            int np1 = t, np2 = t;
            if (act) {
                const int lc1 = pre1 - xpre[sb], lc2 = pre2 - xpre2[sb];
                np1 = sd1 ? sb + lc1 : m + (t - sb - lc1);
                np2 = sd2 ? sb + lc2 : m + (t - sb - lc2);
            }
            xkey[np1] = k1;
            xid[np1] = i1;
            xkey2[np2] = k2;
            xid2[np2] = i2;
            __syncthreads();
            const double n1 = xkey[t], n2 = xkey2[t];
            const int j1 = xid[t], j2 = xid2[t];
            if (b1 == 0) {
                key[0] = n1;
                id[0] = j1;
            } else {
                key[1] = n1;
                id[1] = j1;
            }
            if (b2 == 1) {
                key[1] = n2;
                id[1] = j2;
            } else {
                key[2] = n2;
                id[2] = j2;
            }
            if (act) {
                if (t < m) se = m;
                else sb = m;
            }
            open = __syncthreads_or(t < n && se - sb >= 2);
        }
        if (tie) {
            if (t == 0) tied[atomicAdd(n_tied, 1u)] = s;
            __syncthreads();
            continue;
        }
        // DFS order: position p holds element arrangement[n-1-p]
        xid[t] = id[0];
        __syncthreads();
        const int e = t < n ? xid[n - 1 - t] : 0;
        kd_emit_lane(ms, st, s, n, off, abase, t, e);
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// exact path, sets of at most NT entries (NT = 64 ... KD_BIG): one position
// per thread. The arrangement, prefix counts and pushes live in LDS; every
// thread keeps its own element, its range and the QuickSelect window of that
// range in registers, and computes the window updates itself, so one
// partition step costs three barriers plus the pointer-jumping rounds.
//
// One partition step of a window [lw, rw] (pivot moved to rw) in the
// reference's order: the "<=" elements keep their order at the front; every
// position from the first ">" element f on pushes once onto a queue, a ">"
// element itself, a "<=" element the element it pops from the queue's front
// (the ">" block rotating by one); the pivot follows the "<=" elements and the
// queue's final contents fill the rest, its front at rw.
// ---------------------------------------------------------------------------

// links of a pop chain followed per thread before the pointer jumping
#ifndef KD_CHASE
#define KD_CHASE 8
#endif

// ceil(log2(n)) for n >= 1
__host__ __device__ constexpr int kd_log2_ceil(int n) { return n <= 1 ? 0 : 1 + kd_log2_ceil((n + 1) / 2); }

// Bounds of the exact pass's data-dependent loops. A window of w elements
// needs at most w partition rounds, a tree of n elements fewer than
// n * (ceil(log2 n) + 1) in all; a pop chain has at most n links, so pointer
// jumping (each round halves every open chain) ends within ceil(log2 n) + 1
// rounds. Reaching a bound means a broken invariant: the pass stops and
// reports it instead of looping.
#define KD_JUMP_BOUND(n) (kd_log2_ceil(n) + 2)
#define KD_ROUND_BOUND(n) ((long long)(n) * (kd_log2_ceil(n) + 2) + 64)

template <int NT>
__global__ void __launch_bounds__(NT) k_kdexact(DevMols ms, DevTpls tp, DevSets st, const int *list,
                                                const unsigned int *count)
{
    constexpr int NW = NT / JAM_WAVE;
    constexpr int JUMPS = KD_JUMP_BOUND(NT);
    __shared__ double cc[3][NT];  // coordinates by element
    __shared__ int xa[NT];        // arrangement: element at each position
    __shared__ int xp1[NT];       // exclusive prefix counts of the "<=" flags
    __shared__ int src[NT];       // queue pushes of the window starting at lw, from lw on
    __shared__ int wf[NT];        // first ">" position of the window of the range starting here
    __shared__ int wsum[NW], wact[NW];
    const int t = threadIdx.x;
    const int lane = lane_id();
    const int wv = t / JAM_WAVE;
    if (*st.total > st.cap) return;
    const int cnt = (int)*count;
    for (int q = blockIdx.x; q < cnt; q += gridDim.x) {
        const int s = list[q];
        const int n = st.cnt[s];
        const int64_t off = st.off[s];
        const int abase = ms.atom_off[s / tp.n_types];
        const bool live = t < n;
        if (live) {
            const size_t g = (size_t)(abase + st.list[off + t]);
            cc[0][t] = ms.xyz[3 * g + 0];
            cc[1][t] = ms.xyz[3 * g + 1];
            cc[2][t] = ms.xyz[3 * g + 2];
            xa[t] = t;
        }
        __syncthreads();
        int a = t;                          // element at position t
        int R = 0, H = n;                   // range of position t
        int lw = 0, rw = n >= 2 ? n - 1 : -1;  // its window; rw < 0: none
        int hmax = n;                       // largest range at this depth
        bool broken = false;                // a loop bound was reached (block-uniform)
        const long long round_bound = KD_ROUND_BOUND(n);
        long long rounds = 0;
        for (int d = 0; hmax >= 2;) {
            if (++rounds > round_bound) {
                broken = true;
                break;
            }
            const double *c = cc[d % 3];
            const bool active = live && rw >= 0;
            double pv = 0.0;
            int piv = -1, er = 0;   // pivot position, and the element it takes in the swap
            if (active) {
                // pivot: median of three, swapped with the window's right end
                const int mid = (lw + rw + 1) / 2;
                const int el = xa[lw], em = xa[mid];
                er = xa[rw];
                const double kl = c[el], km = c[em], kr = c[er];
                if ((kl > km) != (kl > kr)) piv = lw;
                else if ((km < kl) != (km < kr)) piv = mid;
                else piv = rw;
                const int ep = piv == lw ? el : (piv == mid ? em : er);
                pv = c[ep];
                if (t == piv) a = er;
                if (t == rw) a = ep;
                if (t == R) wf[R] = rw;
            }
            const bool inwin = active && t >= lw && t < rw;
            const bool le = inwin && c[a] <= pv;
            const bool gt = inwin && !le;
            const unsigned long long bal = ballot(le);
            const unsigned long long gball = ballot(gt);
            const unsigned long long aball = ballot(active);
            if (lane == 0) {
                wsum[wv] = __popcll(bal);
                wact[wv] = aball != 0ull;
            }
            __syncthreads();
            bool any = false;
            for (int w = 0; w < NW; w++) any |= wact[w] != 0;
            if (!any) {
                // every selection of this depth is done: split the ranges at
                // their medians and open the windows of the next depth
                hmax -= hmax / 2;
                if (live && H - R >= 2) {
                    const int m = R + (H - R) / 2;
                    if (t < m) H = m;
                    else R = m;
                }
                lw = R;
                rw = (live && H - R >= 2) ? H - 1 : -1;
                d++;
                __syncthreads();
                continue;
            }
            int p1 = rank_below(bal);
            for (int w = 0; w < wv; w++) p1 += wsum[w];
            xp1[t] = p1;
            if (gt) {
                // the first ">" of its window among this wave's lanes
                const int lo = max(lw - (t - lane), 0);
                const unsigned long long below = gball & ((1ull << lane) - 1ull) & ~((1ull << lo) - 1ull);
                if (below == 0ull) atomicMin(&wf[R], t);
            }
            __syncthreads();
            int p1l = 0, p1r = 0, f = 0;
            if (active) {
                p1l = xp1[lw];
                p1r = xp1[rw];
                f = wf[R];
            }
            const int ones = p1 - p1l;        // "<=" elements before t in the window
            const int k1 = p1r - p1l;         // "<=" elements of the window
            const int m = (rw - lw) - k1;     // ">" elements
            const int nrot = k1 - (f - lw);   // pops: "<=" elements after f
            const int pushes = rw - f;
            bool again = false;
            if (inwin && t >= f) {
                // push t - f; a "<=" element pushes what pop number
                // r = ones - (f - lw) takes, which is push number r, made at
                // position f + r: that position's element if it is a ">" one
                // (its "<=" flag, read off the prefix counts, is 0), else
                // what that position's own pop took, up to KD_CHASE links;
                // a chain still open then is a link for the pointer jumping
                if (!le) {
                    src[lw + (t - f)] = a;
                } else {
                    int r = ones - (f - lw), q = f + r;
                    bool found = xp1[q + 1] == xp1[q];
                    for (int hop = 1; hop < KD_CHASE && !found; hop++) {
                        r = (xp1[q] - p1l) - (f - lw);
                        q = f + r;
                        found = xp1[q + 1] == xp1[q];
                    }
                    if (found) {
                        src[lw + (t - f)] = q == piv ? er : xa[q];
                    } else {
                        src[lw + (t - f)] = -(lw + r) - 1;
                        again = true;
                    }
                }
            }
            // pointer jumping until every push names an element; each pop
            // names a lower push, so the chains end within JUMPS rounds
            for (int jumps = 0; __syncthreads_or(again); jumps++) {
                if (jumps >= JUMPS) {
                    broken = true;
                    break;
                }
                again = false;
                if (inwin && t - lw < pushes) {
                    const int v = src[t];
                    if (v < 0) {
                        const int u = src[-v - 1];
                        src[t] = u;
                        again = u < 0;
                    }
                }
            }
            if (broken) break;
            // scatter: "<=" elements in order, the pivot after them, the
            // queue from its second element on, its front at the right end
            if (active) {
                if (t == rw) {
                    xa[lw + k1] = a;
                } else if (inwin) {
                    if (le) xa[lw + ones] = a;
                    const int j = t - lw;
                    if (j < m) xa[j == 0 ? rw : lw + k1 + j] = src[lw + nrot + j];
                }
            }
            __syncthreads();
            if (live) a = xa[t];
            if (active) {
                // next window of the range's QuickSelect
                const int pos = lw + k1;
                const int k = R + (H - R) / 2;
                int nl = lw, nr = rw;
                if (pos == k) nr = -1;
                else if (k < pos) nr = pos - 1;
                else nl = pos + 1;
                if (nr >= 0 && nl == nr) nr = -1;
                lw = nl;
                rw = nr;
            }
        }
        __syncthreads();
        if (broken) {
            if (t == 0) atomicOr(st.kd_err, 1u);
        } else {
            // DFS order: position p holds element arrangement[n-1-p]
            const int e = live ? xa[n - 1 - t] : 0;
            kd_emit_lane(ms, st, s, n, off, abase, t, e);
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// exact path, sets of (min_n, NT * P] entries: the partition step of
// k_kdexact with P positions per thread, so that a set of a few thousand
// entries stays in LDS. Wave w holds positions w * 64P ... (w + 1) * 64P - 1,
// its lane l positions w * 64P + k * 64 + l (k < P): P slices of 64
// consecutive positions, so that the lanes of a wave touch consecutive LDS
// entries, and prefix counts within a wave come from ballots. Positions,
// elements and prefix counts are 16-bit, and only the coordinate of the
// current depth's axis is kept (reloaded from HBM at every depth): 18 bytes
// per entry, 144 KB for 8,192 entries. A thread keeps the element, range and
// window of each of its positions in registers; its positions in one range
// share the range's pivot and prefix reads.
// ---------------------------------------------------------------------------

// largest set of the LDS exact pass (1,024 threads); larger ones run in
// global scratch. Sets up to 8 * KD_LDS_SMALL_THREADS entries run with fewer
// threads and half the LDS, two workgroups per CU.
#define KD_LDS_MAX 8192
#define KD_LDS_SMALL_THREADS 512
// a queue push that still names another push (src entries are 16-bit)
#define KD_LINK 0x8000u

template <int NT, int P>
__global__ void __launch_bounds__(NT, 4) k_kdexact_lds(DevMols ms, DevTpls tp, DevSets st, const int *list,
                                                    const unsigned int *count, int min_n)
{
    constexpr int N = NT * P;
    constexpr int NW = NT / JAM_WAVE;
    static_assert(N <= KD_LDS_MAX && N < (int)KD_LINK, "16-bit positions with a link flag");
    __shared__ double cc[N];             // coordinate on the current axis, by element
    __shared__ unsigned short xa[N];     // arrangement: element at each position
    __shared__ unsigned short xp1[N];    // exclusive prefix counts of the "<=" flags
    __shared__ unsigned short src[N];    // queue pushes: element, or KD_LINK | push slot
    __shared__ int wf[N];                // first ">" position of the window of the range starting here
    __shared__ int wsum[NW], wact[NW];
    const int t = threadIdx.x;
    const int lane = lane_id();
    const int wv = t / JAM_WAVE;
    const int p0 = wv * JAM_WAVE * P + lane;   // position of slice k: p0 + k * JAM_WAVE
    if (*st.total > st.cap) return;
    const int cnt = (int)*count;
    for (int qs = blockIdx.x; qs < cnt; qs += gridDim.x) {
        const int s = list[qs];
        const int n = st.cnt[s];
        if (n <= min_n || n > N) continue;
        const int64_t off = st.off[s];
        const int abase = ms.atom_off[s / tp.n_types];
        for (int e = t; e < n; e += NT) {
            cc[e] = ms.xyz[3 * (size_t)(abase + st.list[off + e])];
            xa[e] = (unsigned short)e;
        }
        // per position: element, range [R, H), window [lw, rw] (rw < 0: none)
        int a[P], R[P], H[P], lw[P], rw[P];
#pragma unroll
        for (int k = 0; k < P; k++) {
            a[k] = p0 + k * JAM_WAVE;
            R[k] = 0;
            H[k] = n;
            lw[k] = 0;
            rw[k] = (n >= 2 && p0 + k * JAM_WAVE < n) ? n - 1 : -1;
        }
        __syncthreads();
        int hmax = n;                      // largest range at this depth
        bool broken = false;               // a loop bound was reached (block-uniform)
        const long long round_bound = KD_ROUND_BOUND(n);
        const int jump_bound = KD_JUMP_BOUND(n);
        long long rounds = 0;
        for (int d = 0; hmax >= 2;) {
            if (++rounds > round_bound) {
                broken = true;
                break;
            }
            // pivot of each window: median of three, swapped with its right end
            int piv[P], er[P];
            unsigned lem = 0u, gtm = 0u;   // bit k: position k is "<=" / ">" in its window
            bool act = false;
            {
                int cR = -1, c_piv = 0, c_er = 0;
                double c_pv = 0.0;
#pragma unroll
                for (int k = 0; k < P; k++) {
                    const int p = p0 + k * JAM_WAVE;
                    piv[k] = -1;
                    er[k] = 0;
                    if (rw[k] < 0) continue;
                    act = true;
                    if (R[k] != cR) {
                        const int l = lw[k], r = rw[k], mid = (l + r + 1) / 2;
                        const int el = xa[l], em = xa[mid], e_r = xa[r];
                        const double kl = cc[el], km = cc[em], kr = cc[e_r];
                        if ((kl > km) != (kl > kr)) c_piv = l;
                        else if ((km < kl) != (km < kr)) c_piv = mid;
                        else c_piv = r;
                        c_er = e_r;
                        c_pv = c_piv == l ? kl : (c_piv == mid ? km : kr);
                        cR = R[k];
                    }
                    piv[k] = c_piv;
                    er[k] = c_er;
                    if (p == c_piv) a[k] = c_er;
                    if (p == rw[k]) a[k] = xa[c_piv];
                    if (p == R[k]) wf[p] = rw[k];
                    if (p >= lw[k] && p < rw[k]) {
                        if (cc[a[k]] <= c_pv) lem |= 1u << k;
                        else gtm |= 1u << k;
                    }
                }
            }
            // prefix counts of the "<=" flags: within the wave by slice
            // ballots, across waves through LDS; whether any window is open
            int p1[P];
            int wtot = 0;
#pragma unroll
            for (int k = 0; k < P; k++) {
                const unsigned long long bal = ballot((lem >> k) & 1u);
                p1[k] = wtot + rank_below(bal);
                wtot += __popcll(bal);
            }
            const unsigned long long aball = ballot(act);
            if (lane == 0) {
                wsum[wv] = wtot;
                wact[wv] = aball != 0ull;
            }
            __syncthreads();
            bool any = false;
            int base = 0;
            for (int w = 0; w < NW; w++) {
                any |= wact[w] != 0;
                if (w < wv) base += wsum[w];
            }
            if (!any) {
                // every selection of this depth is done: split the ranges at
                // their medians, open the windows of the next depth and load
                // its axis (every read of cc at this depth came before the
                // barrier above)
                hmax -= hmax / 2;
#pragma unroll
                for (int k = 0; k < P; k++) {
                    const bool live = p0 + k * JAM_WAVE < n;
                    if (live && H[k] - R[k] >= 2) {
                        const int m = R[k] + (H[k] - R[k]) / 2;
                        if (p0 + k * JAM_WAVE < m) H[k] = m;
                        else R[k] = m;
                    }
                    lw[k] = R[k];
                    rw[k] = (live && H[k] - R[k] >= 2) ? H[k] - 1 : -1;
                }
                d++;
                if (hmax >= 2) {
                    const int ax = d % 3;
                    for (int e = t; e < n; e += NT) cc[e] = ms.xyz[3 * (size_t)(abase + st.list[off + e]) + ax];
                }
                __syncthreads();
                continue;
            }
            // prefix counts; the first ">" of each window: per slice, the
            // lowest ">" lane of the window among this wave's lanes
            int ones[P];
#pragma unroll
            for (int k = 0; k < P; k++) {
                const int p = p0 + k * JAM_WAVE;
                ones[k] = base + p1[k];
                if (p < n) xp1[p] = (unsigned short)ones[k];
                const bool g = (gtm >> k) & 1u;
                const unsigned long long gball = ballot(g);
                if (g) {
                    const int lo = max(lw[k] - (p - lane), 0);
                    const unsigned long long below = gball & ((1ull << lane) - 1ull) & ~((1ull << lo) - 1ull);
                    if (below == 0ull) atomicMin(&wf[R[k]], p);
                }
            }
            __syncthreads();
            // queue pushes (as in k_kdexact); ones[k] becomes the "<="
            // elements before the position in its window
            int k1v[P], fv[P];
            bool again = false;
            {
                int cR = -1, c_p1l = 0, c_p1r = 0, c_f = 0;
#pragma unroll
                for (int k = 0; k < P; k++) {
                    const int p = p0 + k * JAM_WAVE;
                    k1v[k] = 0;
                    fv[k] = 0;
                    if (rw[k] < 0) continue;
                    if (R[k] != cR) {
                        c_p1l = xp1[lw[k]];
                        c_p1r = xp1[rw[k]];
                        c_f = wf[R[k]];
                        cR = R[k];
                    }
                    const int l = lw[k], r = rw[k], f = c_f;
                    ones[k] -= c_p1l;
                    k1v[k] = c_p1r - c_p1l;
                    fv[k] = f;
                    if (p >= l && p < r && p >= f) {
                        if (!((lem >> k) & 1u)) {
                            src[l + (p - f)] = (unsigned short)a[k];
                        } else {
                            // pop number rr takes push number rr, made at f + rr
                            int rr = ones[k] - (f - l), q = f + rr;
                            bool found = xp1[q + 1] == xp1[q];
                            for (int hop = 1; hop < KD_CHASE && !found; hop++) {
                                rr = (xp1[q] - c_p1l) - (f - l);
                                q = f + rr;
                                found = xp1[q + 1] == xp1[q];
                            }
                            if (found) {
                                src[l + (p - f)] = (unsigned short)(q == piv[k] ? er[k] : xa[q]);
                            } else {
                                src[l + (p - f)] = (unsigned short)(KD_LINK | (unsigned)(l + rr));
                                again = true;
                            }
                        }
                    }
                }
            }
            // pointer jumping until every push names an element; each pop
            // names a lower push, so the chains end within jump_bound rounds
            for (int jumps = 0; __syncthreads_or(again); jumps++) {
                if (jumps >= jump_bound) {
                    broken = true;
                    break;
                }
                again = false;
#pragma unroll
                for (int k = 0; k < P; k++) {
                    const int p = p0 + k * JAM_WAVE;
                    if (rw[k] < 0 || p < lw[k] || p - lw[k] >= rw[k] - fv[k]) continue;
                    const unsigned v = src[p];
                    if (v & KD_LINK) {
                        const unsigned u = src[v & (KD_LINK - 1u)];
                        src[p] = (unsigned short)u;
                        again |= (u & KD_LINK) != 0u;
                    }
                }
            }
            if (broken) break;
            // scatter: "<=" elements in order, the pivot after them, the
            // queue from its second element on, its front at the right end
#pragma unroll
            for (int k = 0; k < P; k++) {
                const int p = p0 + k * JAM_WAVE;
                if (rw[k] < 0) continue;
                const int l = lw[k], r = rw[k], k1 = k1v[k];
                if (p == r) {
                    xa[l + k1] = (unsigned short)a[k];
                } else if (p >= l && p < r) {
                    if ((lem >> k) & 1u) xa[l + ones[k]] = (unsigned short)a[k];
                    const int j = p - l;
                    const int m = (r - l) - k1;            // ">" elements
                    const int nrot = k1 - (fv[k] - l);     // pops
                    if (j < m) xa[j == 0 ? r : l + k1 + j] = src[l + nrot + j];
                }
            }
            __syncthreads();
#pragma unroll
            for (int k = 0; k < P; k++) {
                if (p0 + k * JAM_WAVE < n) a[k] = xa[p0 + k * JAM_WAVE];
                if (rw[k] < 0) continue;
                // next window of the range's QuickSelect
                const int pos = lw[k] + k1v[k];
                const int kk = R[k] + (H[k] - R[k]) / 2;
                int nl = lw[k], nr = rw[k];
                if (pos == kk) nr = -1;
                else if (kk < pos) nr = pos - 1;
                else nl = pos + 1;
                if (nr >= 0 && nl == nr) nr = -1;
                lw[k] = nl;
                rw[k] = nr;
            }
        }
        __syncthreads();
        if (broken) {
            if (t == 0) atomicOr(st.kd_err, 1u);
        } else {
            // DFS order: position p holds element arrangement[n-1-p]
            for (int b0 = wv * JAM_WAVE; b0 < n; b0 += NT) {
                const int p = b0 + lane;
                kd_emit_lane(ms, st, s, n, off, abase, p, p < n ? (int)xa[n - 1 - p] : 0);
            }
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// sets larger than KD_BIG: both paths on work arrays in global scratch
// (scr + off * KD_SCRATCH_BYTES_PER_ENTRY), KD_THREADS threads looping over
// the positions
// ---------------------------------------------------------------------------

// in-place exclusive block scan of v[0, n)
template <int NT>
__device__ static void kd_scan(int *v, int n, int *partial)
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

// Writes the kd arrays of set s from its arrangement `order` (element ids =
// list positions): position p of the DFS visit order is element n-1-p.
template <int NT>
__device__ static void kd_emit_order(const DevMols &ms, const DevSets &st, int s, int n, int64_t off, int abase,
                                     const int *order)
{
    const int lane = lane_id();
    for (int b0 = (threadIdx.x / JAM_WAVE) * JAM_WAVE; b0 < n; b0 += NT) {
        const int p = b0 + lane;
        kd_emit_lane(ms, st, s, n, off, abase, p, p < n ? order[n - 1 - p] : 0);
    }
}

struct KdFast {
    double *c;               // coordinates, axis-major, stride entries per axis
    int *perm;               // 3 axes x 2 buffers, stride entries each
    int *pre;                // prefix of left sides
    int *sb, *se;            // range of each position
    unsigned char *side;     // by element
    int stride;
};

template <int NT>
__device__ static void kdf_sort(const double *c, int *perm, int n, int n2)
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

// fast path; returns true (nothing written to out) on a split tie
template <int NT>
__device__ static bool kd_fast(const KdFast &w, int n, int *out, int *partial, int *flag)
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
        if (tie) return true;
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
    return false;
}

struct KdWork {
    double *c;               // coordinates: axis a of element e at c[a * stride + e]
    int *arr;                // arrangement, double buffered at arr and arr + stride
    int *pre1, *pre2;        // exclusive prefixes: "<=" flags, re-push flags
    int *src;                // push -> element (>= 0) or -(push index) - 1
    int *rlo, *rhi;          // range of each position at the current level
    int *wl, *wr;            // QuickSelect window of the range starting here; wr < 0: done
    int stride;
};

__device__ static inline int kd_cmp3(const double *c, int x, int y)
{
    const double u = c[x], v = c[y];
    return u < v ? -1 : (u > v ? 1 : 0);
}

// exact path: arranges elements 0..n-1 into the reference's leaf order;
// false when a loop bound was reached (out is then not written)
template <int NT>
__device__ static bool kd_exact(const KdWork &w, int n, int *out, int *partial, int *flag)
{
    const int tid = threadIdx.x;
    const int jump_bound = KD_JUMP_BOUND(n);
    const long long round_bound = KD_ROUND_BOUND(n);
    long long rounds = 0;
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
            if (++rounds > round_bound) return false;
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
            for (int jumps = 0;; jumps++) {
                if (jumps > jump_bound) return false;
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
    return true;
}

__device__ static void kd_load_global(const DevMols &ms, const DevSets &st, int n, int64_t off, int abase,
                                      double *c, size_t stride)
{
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        const size_t g = (size_t)(abase + st.list[off + i]);
        c[i] = ms.xyz[3 * g + 0];
        c[stride + i] = ms.xyz[3 * g + 1];
        c[2 * stride + i] = ms.xyz[3 * g + 2];
    }
    __syncthreads();
}

// fast path of the sets of `list` larger than min_n; sets with a split tie
// go to `tied`
__global__ void __launch_bounds__(KD_THREADS) k_kdbig(DevMols ms, DevTpls tp, DevSets st, unsigned char *scr,
                                                      const int *list, const unsigned int *count, int *tied,
                                                      unsigned int *n_tied, int min_n)
{
    __shared__ int partial[KD_THREADS];
    __shared__ int flag[2];
    if (*st.total > st.cap) return;
    const int cnt = (int)*count;
    for (int q = blockIdx.x; q < cnt; q += gridDim.x) {
        const int s = list[q];
        const int n = st.cnt[s];
        if (n <= min_n) continue;
        const int64_t off = st.off[s];
        const int abase = ms.atom_off[s / tp.n_types];
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
        kd_load_global(ms, st, n, off, abase, w.c, (size_t)n2);
        // the arrangement after the work arrays: 61 * n2 + 4 * n <= 128 * n bytes
        int *order = (int *)((unsigned char *)(ib + 9 * n2) + n2);
        if (!kd_fast<KD_THREADS>(w, n, order, partial, flag)) kd_emit_order<KD_THREADS>(ms, st, s, n, off, abase, order);
        else if (threadIdx.x == 0) tied[atomicAdd(n_tied, 1u)] = s;
        __syncthreads();
    }
}

// exact path of the sets of `list` larger than min_n
__global__ void __launch_bounds__(KD_EXACT_THREADS) k_kdbig_exact(DevMols ms, DevTpls tp, DevSets st,
                                                                  unsigned char *scr, const int *list,
                                                                  const unsigned int *count, int min_n)
{
    __shared__ int partial[KD_EXACT_THREADS];
    __shared__ int flag[2];
    if (*st.total > st.cap) return;
    const int cnt = (int)*count;
    for (int q = blockIdx.x; q < cnt; q += gridDim.x) {
        const int s = list[q];
        const int n = st.cnt[s];
        if (n <= min_n) continue;
        const int64_t off = st.off[s];
        const int abase = ms.atom_off[s / tp.n_types];
        double *cb = (double *)(scr + (size_t)off * KD_SCRATCH_BYTES_PER_ENTRY);
        int *ib = (int *)(cb + 3 * (size_t)n);
        const size_t stride = (size_t)n;
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
        kd_load_global(ms, st, n, off, abase, w.c, stride);
        int *order = ib + 9 * stride;   // the arrangement after the work arrays (64 * n bytes)
        if (kd_exact<KD_EXACT_THREADS>(w, n, order, partial, flag))
            kd_emit_order<KD_EXACT_THREADS>(ms, st, s, n, off, abase, order);
        else if (threadIdx.x == 0)
            atomicOr(st.kd_err, 1u);
        __syncthreads();
    }
}

}  // namespace jam
