// Shared device/host helpers for the gfx950 tree-gravity library.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdio.h>

#include "../../include/ptg.h"

#define PTG_WAVE 64

// -------------------------------------------------------------------------------------------
// host-side error reporting: every entry point returns a ptg_status and never aborts
// -------------------------------------------------------------------------------------------
void ptg_set_error(const char *fmt, ...);

#define PTG_HIP_CHECK(expr)                                                                       \
    do {                                                                                          \
        hipError_t _e = (expr);                                                                   \
        if (_e != hipSuccess) {                                                                   \
            ptg_set_error("%s failed: %s (%s:%d)", #expr, hipGetErrorString(_e), __FILE__, __LINE__); \
            return PTG_ERR_HIP;                                                                   \
        }                                                                                         \
    } while (0)

#define PTG_LAUNCH_CHECK()                                                                        \
    do {                                                                                          \
        hipError_t _e = hipGetLastError();                                                        \
        if (_e != hipSuccess) {                                                                   \
            ptg_set_error("kernel launch failed: %s (%s:%d)", hipGetErrorString(_e), __FILE__, __LINE__); \
            return PTG_ERR_HIP;                                                                   \
        }                                                                                         \
    } while (0)

static inline int64_t ptg_cdiv(int64_t a, int64_t b) { return (a + b - 1) / b; }
static inline size_t ptg_align256(size_t x) { return (x + 255) & ~(size_t)255; }

// bump allocator over a caller-provided workspace
struct PtgArena {
    char *base;
    size_t size, used;
    __host__ PtgArena(void *b, size_t s) : base((char *)b), size(s), used(0) {}
    template <typename T> __host__ T *take(int64_t count)
    {
        size_t bytes = ptg_align256((size_t)(count > 0 ? count : 1) * sizeof(T));
        if (base == nullptr) { used += bytes; return nullptr; }  // sizing pass
        if (used + bytes > size) { used += bytes; return nullptr; }
        T *p = (T *)(base + used);
        used += bytes;
        return p;
    }
    __host__ bool ok() const { return base == nullptr || used <= size; }
};

// -------------------------------------------------------------------------------------------
// device: wavefront primitives (wave64)
// -------------------------------------------------------------------------------------------
__device__ __forceinline__ int ptg_lane() { return __builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u)); }

__device__ __forceinline__ uint64_t ptg_ballot(bool p) { return __ballot(p); }

__device__ __forceinline__ uint64_t ptg_lanemask_lt() { return (1ull << ptg_lane()) - 1ull; }

// inclusive prefix sum of a 32-bit int across the wave
__device__ __forceinline__ int ptg_wave_incl_scan(int v)
{
    int lane = ptg_lane();
#pragma unroll
    for (int d = 1; d < PTG_WAVE; d <<= 1) {
        int o = __shfl_up(v, d, PTG_WAVE);
        if (lane >= d) v += o;
    }
    return v;
}

__device__ __forceinline__ int ptg_wave_sum(int v)
{
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) v += __shfl_xor(v, d, PTG_WAVE);
    return v;
}

template <typename T> __device__ __forceinline__ T ptg_wave_min(T v)
{
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) {
        T o = __shfl_xor(v, d, PTG_WAVE);
        v = o < v ? o : v;
    }
    return v;
}

template <typename T> __device__ __forceinline__ T ptg_wave_max(T v)
{
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) {
        T o = __shfl_xor(v, d, PTG_WAVE);
        v = o > v ? o : v;
    }
    return v;
}

// -------------------------------------------------------------------------------------------
// Morton digits / keys (63-bit, x in the least significant bit of each 3-bit triplet)
// -------------------------------------------------------------------------------------------
__device__ __forceinline__ uint64_t ptg_spread_bits(uint64_t x)
{
    x &= 0x1FFFFFull;
    x = (x | (x << 32)) & 0x1F00000000FFFFull;
    x = (x | (x << 16)) & 0x1F0000FF0000FFull;
    x = (x | (x << 8)) & 0x100F00F00F00F00Full;
    x = (x | (x << 4)) & 0x10C30C30C30C30C3ull;
    x = (x | (x << 2)) & 0x1249249249249249ull;
    return x;
}

// octant offsets (octree.py:31-42): digit d -> 0.25 * (+-1, +-1, +-1), x from bit 0
__device__ __forceinline__ double ptg_octant_offset(int d, int k) { return ((d >> k) & 1) ? 0.25 : -0.25; }

// number of whole 3-bit levels two 63-bit keys share (0..21)
__device__ __forceinline__ int ptg_common_levels(uint64_t a, uint64_t b)
{
    uint64_t x = a ^ b;
    if (x == 0) return 21;
    int lcp = __clzll((long long)x) - 1;  // bit 63 is zero in both keys
    return lcp / 3;
}

// node count of a tree: nnodes as passed or, for a tree from ptg_build_deferred
// (ptg_tree_arrays.node_count set), the count its build wrote there (0 when the tree did not fit),
// never more than the capacity nnodes
__device__ __forceinline__ int64_t ptg_node_count(int64_t nnodes, const int32_t *count)
{
    if (!count) return nnodes;
    const int64_t c = *count;
    return c < nnodes ? c : nnodes;
}
