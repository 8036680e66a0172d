// CPU emulation of the HIP device features the kd kernels use, for testing
// them on the host under AddressSanitizer and UBSan (tests/test_kd_emu.py).
// Test infrastructure only: nothing in the package includes it.
//
// A workgroup runs as blockDim.x host threads; workgroups run one after the
// other. __syncthreads is a barrier over the workgroup's threads that have
// not returned (as s_barrier counts only waves still running), wave
// operations (ballot, shuffles, ds_permute, wave barriers) are exchanges
// through a per-wave buffer between barriers over the wave's live lanes, so
// the code must call them uniformly across a wave, as it does on the device.
// A barrier that waits for more than 60 s aborts (a hang or a divergent wave
// operation). __shared__ variables are function-local statics; a linker
// script (lds.ld) gathers them into one section, which the launcher fills
// with 0xFF bytes before every workgroup, so that a read of an LDS slot that
// the workgroup did not write sees garbage, as it may on the device.
#pragma once

#include <chrono>
#include <condition_variable>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <thread>
#include <type_traits>
#include <vector>

#define __global__
#define __device__
#define __host__
#define __shared__ static
#define __launch_bounds__(...)
#define __forceinline__ inline

struct dim3 {
    unsigned x, y, z;
    dim3(unsigned x_ = 1, unsigned y_ = 1, unsigned z_ = 1) : x(x_), y(y_), z(z_) {}
};

namespace emu {

inline thread_local dim3 t_thread, t_block;
inline dim3 g_dim, g_grid;

// a barrier over a changing number of participants, with an OR of a flag
struct Barrier {
    std::mutex m;
    std::condition_variable cv;
    int expected = 0, arrived = 0, acc = 0, result = 0;
    unsigned long gen = 0;

    void release_locked()
    {
        result = acc;
        acc = 0;
        arrived = 0;
        gen++;
        cv.notify_all();
    }
    int arrive(int flag)
    {
        std::unique_lock<std::mutex> lk(m);
        const unsigned long g = gen;
        acc |= flag;
        if (++arrived >= expected) {
            release_locked();
            return result;
        }
        if (!cv.wait_for(lk, std::chrono::seconds(60), [&] { return gen != g; })) {
            std::fprintf(stderr, "emulator: a barrier waited 60 s (hang, or a divergent barrier)\n");
            std::abort();
        }
        return result;
    }
    void drop()
    {
        std::unique_lock<std::mutex> lk(m);
        expected--;
        if (arrived > 0 && arrived >= expected) release_locked();
    }
};

struct Block {
    Barrier all;
    std::vector<Barrier> wave;
    std::vector<uint64_t> xb;   // 64 slots per wave
};
inline Block *g_block = nullptr;

inline int lane() { return (int)(t_thread.x & 63); }
inline int wave_index() { return (int)(t_thread.x / 64); }
inline Barrier &wbar() { return g_block->wave[wave_index()]; }
inline uint64_t *wbuf() { return g_block->xb.data() + 64 * wave_index(); }

template <class T>
inline uint64_t to_bits(T v)
{
    static_assert(sizeof(T) <= 8, "64-bit values at most");
    uint64_t b = 0;
    std::memcpy(&b, &v, sizeof(T));
    return b;
}
template <class T>
inline T from_bits(uint64_t b)
{
    T v;
    std::memcpy(&v, &b, sizeof(T));
    return v;
}

// lane `src`'s value of v (every live lane of the wave calls it)
template <class T>
inline T shfl(T v, int src)
{
    uint64_t *b = wbuf();
    b[lane()] = to_bits(v);
    wbar().arrive(0);
    const T r = from_bits<T>(b[src & 63]);
    wbar().arrive(0);
    return r;
}

// the value the lane whose target is this lane pushed (poison if none)
inline int permute(int dst, int v)
{
    uint64_t *b = wbuf();
    b[lane()] = 0xFFFFFFFFu;
    wbar().arrive(0);
    b[dst & 63] = (uint32_t)v;
    wbar().arrive(0);
    const int r = (int)(uint32_t)b[lane()];
    wbar().arrive(0);
    return r;
}

inline uint64_t ballot(bool p)
{
    uint64_t *b = wbuf();
    b[lane()] = p ? 1u : 0u;
    // a lane that has returned leaves a 0 behind
    wbar().arrive(0);
    uint64_t m = 0;
    const int live = (int)g_dim.x - 64 * wave_index();
    for (int k = 0; k < 64 && k < live; k++)
        if (b[k] == 1u) m |= 1ull << k;
    wbar().arrive(0);
    return m;
}

// Runs kernel(args...) over grid x block threads, one workgroup at a time,
// LDS (every function-local static of the kd kernels) poisoned before each.
extern "C" char __start_jam_lds[] __attribute__((weak));
extern "C" char __stop_jam_lds[] __attribute__((weak));

// 0xFF over the whole LDS section, the AddressSanitizer redzones between
// its arrays included (not instrumented; the redzones stay poisoned in the
// shadow memory, so an overflow of an LDS array is still reported)
__attribute__((no_sanitize("address"))) inline void poison_lds()
{
    for (volatile char *p = __start_jam_lds; p < __stop_jam_lds; ++p) *p = (char)0xFF;
}

template <class... KArgs, class... Args>
void launch(void (*kernel)(KArgs...), dim3 grid, dim3 block, Args... args)
{
    if (!__start_jam_lds || !__stop_jam_lds) {
        std::fprintf(stderr, "emulator: the LDS section is missing (link with lds.ld)\n");
        std::abort();
    }
    g_grid = grid;
    g_dim = block;
    const unsigned nt = block.x;
    for (unsigned bx = 0; bx < grid.x; bx++) {
        poison_lds();
        Block B;
        B.all.expected = (int)nt;
        B.wave = std::vector<Barrier>((nt + 63) / 64);
        for (unsigned w = 0; w < B.wave.size(); w++) B.wave[w].expected = (int)std::min(64u, nt - 64 * w);
        B.xb.assign(64 * B.wave.size(), 0);
        g_block = &B;
        std::vector<std::thread> th;
        th.reserve(nt);
        for (unsigned t = 0; t < nt; t++)
            th.emplace_back([&, t, bx] {
                t_thread = dim3(t, 0, 0);
                t_block = dim3(bx, 0, 0);
                kernel(args...);
                // a returned lane leaves a false ballot vote and its barriers
                g_block->xb[t] = 0;
                g_block->wave[t / 64].drop();
                g_block->all.drop();
            });
        for (auto &x : th) x.join();
        g_block = nullptr;
    }
}

}  // namespace emu

#define threadIdx (emu::t_thread)
#define blockIdx (emu::t_block)
#define blockDim (emu::g_dim)
#define gridDim (emu::g_grid)

inline void __syncthreads() { emu::g_block->all.arrive(0); }
inline int __syncthreads_or(int p) { return emu::g_block->all.arrive(p != 0); }
inline uint64_t __ballot(int p) { return emu::ballot(p != 0); }

template <class T>
inline T __shfl(T v, int src, int width = 64)
{
    return emu::shfl(v, (emu::lane() & ~(width - 1)) + (src & (width - 1)));
}
template <class T>
inline T __shfl_xor(T v, int mask, int width = 64)
{
    const int l = emu::lane();
    const int src = ((l ^ mask) & (width - 1)) | (l & ~(width - 1));
    return emu::shfl(v, src);
}
template <class T>
inline T __shfl_up(T v, unsigned d, int width = 64)
{
    const int l = emu::lane();
    const int base = l & ~(width - 1);
    const int src = (l - base) >= (int)d ? l - (int)d : l;
    return emu::shfl(v, src);
}
template <class T>
inline T __shfl_down(T v, unsigned d, int width = 64)
{
    const int l = emu::lane();
    const int base = l & ~(width - 1);
    const int src = (l - base) + (int)d < width ? l + (int)d : l;
    return emu::shfl(v, src);
}

inline int __builtin_amdgcn_ds_permute(int addr, int v) { return emu::permute(addr >> 2, v); }
inline unsigned __builtin_amdgcn_mbcnt_lo(unsigned m, unsigned acc)
{
    const int l = emu::lane();
    const unsigned below = l >= 32 ? m : (m & ((1u << l) - 1u));
    return acc + (unsigned)__builtin_popcount(below);
}
inline unsigned __builtin_amdgcn_mbcnt_hi(unsigned m, unsigned acc)
{
    const int l = emu::lane();
    const unsigned below = l < 32 ? 0u : (m & ((l == 32) ? 0u : ((1u << (l - 32)) - 1u)));
    return acc + (unsigned)__builtin_popcount(below);
}
inline void __builtin_amdgcn_wave_barrier() { emu::wbar().arrive(0); }
#define __builtin_amdgcn_fence(order, scope) ((void)0)
template <class T>
inline T __builtin_amdgcn_readfirstlane(T v)
{
    return emu::shfl(v, 0);
}

inline int __popc(unsigned v) { return __builtin_popcount(v); }
inline int __popcll(unsigned long long v) { return __builtin_popcountll(v); }
inline int __ffsll(long long v) { return __builtin_ffsll(v); }
inline int __clzll(long long v) { return v == 0 ? 64 : __builtin_clzll((unsigned long long)v); }

template <class T>
inline T atomicAdd(T *p, T v)
{
    return __atomic_fetch_add(p, v, __ATOMIC_SEQ_CST);
}
template <class T>
inline T atomicOr(T *p, T v)
{
    return __atomic_fetch_or(p, v, __ATOMIC_SEQ_CST);
}
template <class T>
inline T atomicMin(T *p, T v)
{
    T old = __atomic_load_n(p, __ATOMIC_SEQ_CST);
    while (v < old && !__atomic_compare_exchange_n(p, &old, v, false, __ATOMIC_SEQ_CST, __ATOMIC_SEQ_CST)) {
    }
    return old;
}
template <class T>
inline T atomicMax(T *p, T v)
{
    T old = __atomic_load_n(p, __ATOMIC_SEQ_CST);
    while (v > old && !__atomic_compare_exchange_n(p, &old, v, false, __ATOMIC_SEQ_CST, __ATOMIC_SEQ_CST)) {
    }
    return old;
}
template <class T>
inline T atomicCAS(T *p, T cmp, T v)
{
    __atomic_compare_exchange_n(p, &cmp, v, false, __ATOMIC_SEQ_CST, __ATOMIC_SEQ_CST);
    return cmp;
}

template <class A, class B>
inline auto min(A a, B b) -> typename std::common_type<A, B>::type
{
    return a < b ? a : b;
}
template <class A, class B>
inline auto max(A a, B b) -> typename std::common_type<A, B>::type
{
    return a > b ? a : b;
}
using std::fmax;
using std::fmin;
