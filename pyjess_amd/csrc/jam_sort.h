// jam_sort.h -- stable LSD radix sort of (key, value) pairs, 8 bits per pass.
//
// One pass over bits [shift, shift + 8) of the keys:
//   k_rs_hist     a tile of RS_TILE keys per workgroup: the count of every
//                 digit value in the tile (LDS atomics), written digit-major,
//                 hist[d * n_tiles + tile];
//   (exclusive scan of hist: where each tile's keys of digit d start)
//   k_rs_scatter  the tile again, 256 consecutive keys at a time: each key's
//                 rank among the keys of its digit before it in the chunk
//                 (lanes of a wave with the same digit found by 8 ballots,
//                 earlier waves by their per-digit counts in LDS), plus the
//                 running position of that digit in the tile.
// Keys of equal digit keep their input order, so passes from the least to
// the most significant digit sort by the whole key, ties in input order.
#pragma once

#include <hip/hip_runtime.h>

#include "jam_common.h"
#include "jam_wave.h"

namespace jam {

#define RS_THREADS 256                   // one thread per digit value
#define RS_ITEMS 16
#define RS_TILE (RS_THREADS * RS_ITEMS)  // keys per workgroup
#define RS_WAVES (RS_THREADS / JAM_WAVE)

template <class K>
__global__ void __launch_bounds__(RS_THREADS) k_rs_hist(const K *keys, long long n, int shift, int *hist,
                                                        long long n_tiles)
{
    __shared__ int h[RS_THREADS];
    const int t = threadIdx.x;
    h[t] = 0;
    __syncthreads();
    const long long t0 = (long long)blockIdx.x * RS_TILE;
#pragma unroll 4
    for (int k = 0; k < RS_ITEMS; k++) {
        const long long i = t0 + (long long)k * RS_THREADS + t;
        if (i < n) atomicAdd(&h[(int)((keys[i] >> shift) & 0xFF)], 1);
    }
    __syncthreads();
    hist[(long long)t * n_tiles + blockIdx.x] = h[t];
}

// offs: the exclusive scan of hist (same layout)
template <class K>
__global__ void __launch_bounds__(RS_THREADS) k_rs_scatter(const K *kin, const int *vin, long long n, int shift,
                                                           const int *offs, long long n_tiles, K *kout, int *vout)
{
    __shared__ int run[RS_THREADS];
    __shared__ int wcnt[RS_WAVES][RS_THREADS];
    const int t = threadIdx.x, lane = lane_id(), w = t / JAM_WAVE;
    run[t] = offs[(long long)t * n_tiles + blockIdx.x];
    const long long t0 = (long long)blockIdx.x * RS_TILE;
    const unsigned long long below = (1ull << lane) - 1ull;
    for (int k = 0; k < RS_ITEMS; k++) {
#pragma unroll
        for (int q = 0; q < RS_WAVES; q++) wcnt[q][t] = 0;
        __syncthreads();
        const long long i = t0 + (long long)k * RS_THREADS + t;
        const bool ok = i < n;
        const K key = ok ? kin[i] : (K)0;
        const int val = ok ? vin[i] : 0;
        const int d = (int)((key >> shift) & 0xFF);
        // the lanes of this wave holding a key of digit d
        unsigned long long peers = ballot(ok);
#pragma unroll
        for (int b = 0; b < 8; b++) {
            const bool bit = ((d >> b) & 1) != 0;
            const unsigned long long m = ballot(bit);
            peers &= bit ? m : ~m;
        }
        const int rank = __popcll(peers & below);
        if (ok && rank == 0) wcnt[w][d] = __popcll(peers);
        __syncthreads();
        if (ok) {
            int pos = run[d] + rank;
            for (int q = 0; q < w; q++) pos += wcnt[q][d];
            kout[pos] = key;
            vout[pos] = val;
        }
        __syncthreads();
        int add = 0;
#pragma unroll
        for (int q = 0; q < RS_WAVES; q++) add += wcnt[q][t];
        run[t] += add;
        __syncthreads();
    }
}

}  // namespace jam
