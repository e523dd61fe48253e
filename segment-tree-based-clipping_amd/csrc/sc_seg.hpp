// Sorting of many small key groups (segments) at once, each group held at
// [scan[g], scan[g+1]) of a u64 key array.  Keys inside a group must be
// distinct or, where equal, interchangeable.  A group of up to SEG_THREAD
// keys is sorted by a register bitonic network in one thread, up to SEG_WAVE
// by a bitonic sort across one wave, up to SEG_LDS_MAX in one block's LDS;
// a larger group is only reported (the caller then sorts another way).  The
// output is the key's low word as an int (FULL = false) or the whole key.
#pragma once
#include "sc_common.hpp"

namespace sc {
namespace {

constexpr int SEG_BLK = 256;       // threads per block of the launches below
constexpr int SEG_THREAD = 8;      // groups up to this size: a register sorting network per thread
constexpr int SEG_WAVE = 64;       // up to this size: a bitonic sort across one wave
constexpr int SEG_LDS_MAX = 4096;  // up to this size: a bitonic sort in one block's LDS
constexpr int SEG_CTL = 3;         // wave-list count, block-list count, too-large flag

template <bool FULL> __device__ inline void seg_put(void *out, int i, uint64_t v) {
    if (FULL) static_cast<uint64_t *>(out)[i] = v;
    else static_cast<int *>(out)[i] = int(uint32_t(v));
}

// ascending bitonic network on N register values (indices fixed at compile time)
template <int N> __device__ inline void reg_bitonic(uint64_t (&v)[N]) {
#pragma unroll
    for (int k = 2; k <= N; k <<= 1)
#pragma unroll
        for (int j = k >> 1; j > 0; j >>= 1)
#pragma unroll
            for (int i = 0; i < N; ++i) {
                const int l = i ^ j;
                if (l > i) {
                    const uint64_t x = v[i], y = v[l];
                    const bool sw = (i & k) == 0 ? x > y : x < y;
                    v[i] = sw ? y : x;
                    v[l] = sw ? x : y;
                }
            }
}

// one thread per edge group: groups of up to SEG_THREAD sorted in registers,
// larger ones listed for a wave or a block, groups beyond SEG_LDS_MAX reported
template <bool FULL>
__global__ void k_seg_small(int nseg, const int *__restrict__ scan, const uint64_t *__restrict__ K, void *out,
                            uint32_t *wlist, uint32_t *blist, uint32_t *ctl) {
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= nseg) return;
    const int s = scan[g], n = scan[g + 1] - s;
    if (n <= 0) return;
    if (n == 1) {
        seg_put<FULL>(out, s, K[s]);
        return;
    }
    if (n > SEG_THREAD) {
        if (n <= SEG_WAVE) wlist[atomicAdd(&ctl[0], 1u)] = uint32_t(g);
        else if (n <= SEG_LDS_MAX) blist[atomicAdd(&ctl[1], 1u)] = uint32_t(g);
        else atomicOr(&ctl[2], 1u);
        return;
    }
    uint64_t v[SEG_THREAD];
#pragma unroll
    for (int i = 0; i < SEG_THREAD; ++i) v[i] = i < n ? K[s + i] : ~0ull; // keys are below ~0
    reg_bitonic<SEG_THREAD>(v);
#pragma unroll
    for (int i = 0; i < SEG_THREAD; ++i)
        if (i < n) seg_put<FULL>(out, s + i, v[i]);
}

__device__ inline uint64_t shfl_xor_u64(uint64_t v, int m) {
    const int lo = __shfl_xor(int(uint32_t(v)), m, WAVE), hi = __shfl_xor(int(uint32_t(v >> 32)), m, WAVE);
    return (uint64_t(uint32_t(hi)) << 32) | uint32_t(lo);
}

// one wave per listed group of SEG_THREAD+1..SEG_WAVE pairs, one key per lane
template <bool FULL>
__global__ __launch_bounds__(SEG_BLK) void k_seg_wave(const int *__restrict__ scan, const uint64_t *__restrict__ K,
                                                  void *out, const uint32_t *__restrict__ wlist,
                                                  const uint32_t *__restrict__ ctl) {
    const int lane = threadIdx.x & (WAVE - 1);
    const uint32_t nw = ctl[0];
    const uint32_t stride = gridDim.x * (SEG_BLK / WAVE);
    for (uint32_t w = blockIdx.x * (SEG_BLK / WAVE) + threadIdx.x / WAVE; w < nw; w += stride) {
        const int g = int(wlist[w]);
        const int s = scan[g], n = scan[g + 1] - s;
        uint64_t v = lane < n ? K[s + lane] : ~0ull;
#pragma unroll
        for (int k = 2; k <= WAVE; k <<= 1)
#pragma unroll
            for (int j = k >> 1; j > 0; j >>= 1) {
                const uint64_t o = shfl_xor_u64(v, j);
                const bool keep_min = ((lane & j) == 0) == ((lane & k) == 0);
                v = keep_min ? (o < v ? o : v) : (o > v ? o : v);
            }
        if (lane < n) seg_put<FULL>(out, s + lane, v);
    }
}

// one block per listed group of SEG_WAVE+1..SEG_LDS_MAX pairs: bitonic sort in LDS
template <bool FULL>
__global__ __launch_bounds__(SEG_BLK) void k_seg_block(const int *__restrict__ scan, const uint64_t *__restrict__ K,
                                                   void *out, const uint32_t *__restrict__ blist,
                                                   const uint32_t *__restrict__ ctl) {
    __shared__ uint64_t sk[SEG_LDS_MAX];
    const uint32_t nb = ctl[1];
    for (uint32_t b = blockIdx.x; b < nb; b += gridDim.x) {
        const int g = int(blist[b]);
        const int s = scan[g], n = scan[g + 1] - s;
        int m = SEG_WAVE * 2;
        while (m < n) m <<= 1;
        for (int i = threadIdx.x; i < m; i += SEG_BLK) sk[i] = i < n ? K[s + i] : ~0ull;
        __syncthreads();
        for (int k = 2; k <= m; k <<= 1)
            for (int j = k >> 1; j > 0; j >>= 1) {
                // the m/2 compare-exchange pairs (i, i + j), i with bit j clear
                for (int h = threadIdx.x; h < (m >> 1); h += SEG_BLK) {
                    const int i = ((h & ~(j - 1)) << 1) | (h & (j - 1));
                    const int l = i + j;
                    const uint64_t x = sk[i], y = sk[l];
                    if ((i & k) == 0 ? x > y : x < y) {
                        sk[i] = y;
                        sk[l] = x;
                    }
                }
                __syncthreads();
            }
        for (int i = threadIdx.x; i < n; i += SEG_BLK) seg_put<FULL>(out, s + i, sk[i]);
        __syncthreads();
    }
}


// the three launches for groups 0..nseg-1; wl and bl hold the listed groups
// (capacity: more than SEG_THREAD, resp. SEG_WAVE keys per listed group),
// ctl[0..SEG_CTL) is zero before the call
template <bool FULL>
inline void seg_sort(int nseg, const int *scan, const uint64_t *K, void *out, uint32_t *wl, uint32_t *bl,
                     uint32_t *ctl, hipStream_t s) {
    if (nseg <= 0) return;
    k_seg_small<FULL><<<unsigned((nseg + SEG_BLK - 1) / SEG_BLK), SEG_BLK, 0, s>>>(nseg, scan, K, out, wl, bl, ctl);
    k_seg_wave<FULL><<<1024, SEG_BLK, 0, s>>>(scan, K, out, wl, ctl);
    k_seg_block<FULL><<<1280, SEG_BLK, 0, s>>>(scan, K, out, bl, ctl); // five 32 KB blocks per CU
    SC_HIP(hipGetLastError());
}

} // namespace
} // namespace sc
