// Sorting of many small key groups (segments) at once, each group held at
// [scan[g], scan[g+1]) of a u64 key array.  Keys inside a group must be
// distinct or, where equal, interchangeable.  A group of up to SEG_THREAD
// keys is sorted by a register bitonic network in one thread, up to
// SEG_WAVE_MAX by a bitonic sort in one wave's registers (1, 2, 4 or 8 keys
// per lane, no barriers), up to SEG_LDS_MAX in one block's LDS; a larger
// group is only reported (the caller then sorts another way).  The output is
// the key's low word as an int (FULL = false) or the whole key.
#pragma once
#include "sc_common.hpp"

namespace sc {
namespace {

constexpr int SEG_BLK = 256;       // threads per block of the launches below
constexpr int SEG_THREAD = 8;      // groups up to this size: a register sorting network per thread
constexpr int SEG_WAVE = 64;       // lanes of the wave sort
constexpr int SEG_WAVE_MAX = 8 * SEG_WAVE; // up to this size: a bitonic sort in one wave's registers
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
        if (n <= SEG_WAVE_MAX) wlist[atomicAdd(&ctl[0], 1u)] = uint32_t(g);
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

// Ascending bitonic sort of the n <= R * SEG_WAVE keys at K[s..s+n) by one
// wave: lane holds elements lane*R .. lane*R+R-1 (missing ones as ~0, above
// every key).  Compare-exchange distances below R stay in a lane's registers,
// larger ones pair lane with lane ^ (j / R) through shuffles.
template <int R, bool FULL>
__device__ inline void wave_sort(const uint64_t *__restrict__ K, int s, int n, void *out, int lane) {
    uint64_t v[R];
    const int base = lane * R;
#pragma unroll
    for (int r = 0; r < R; ++r) v[r] = base + r < n ? K[s + base + r] : ~0ull;
#pragma unroll
    for (int k = 2; k <= R * SEG_WAVE; k <<= 1)
#pragma unroll
        for (int j = k >> 1; j > 0; j >>= 1) {
            if (j < R) {
#pragma unroll
                for (int r = 0; r < R; ++r)
                    if ((r & j) == 0) {
                        const int l = r | j;
                        const bool asc = ((base | r) & k) == 0;
                        const uint64_t x = v[r], y = v[l];
                        const bool sw = asc ? x > y : x < y;
                        v[r] = sw ? y : x;
                        v[l] = sw ? x : y;
                    }
            } else {
                // k > j >= R: the direction bit and the partner bit are lane bits
                const int jl = j / R;
                const bool keep_min = ((lane & jl) == 0) == ((base & k) == 0);
#pragma unroll
                for (int r = 0; r < R; ++r) {
                    const uint64_t o = shfl_xor_u64(v[r], jl);
                    v[r] = keep_min ? (o < v[r] ? o : v[r]) : (o > v[r] ? o : v[r]);
                }
            }
        }
#pragma unroll
    for (int r = 0; r < R; ++r)
        if (base + r < n) seg_put<FULL>(out, s + base + r, v[r]);
}

// one wave per listed group of SEG_THREAD+1..SEG_WAVE_MAX keys, with the
// fewest keys per lane that hold the group
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
        if (n <= SEG_WAVE) wave_sort<1, FULL>(K, s, n, out, lane);
        else if (n <= 2 * SEG_WAVE) wave_sort<2, FULL>(K, s, n, out, lane);
        else if (n <= 4 * SEG_WAVE) wave_sort<4, FULL>(K, s, n, out, lane);
        else wave_sort<8, FULL>(K, s, n, out, lane);
    }
}

// one block per listed group of SEG_WAVE_MAX+1..SEG_LDS_MAX keys: bitonic sort in LDS
template <bool FULL>
__global__ __launch_bounds__(SEG_BLK) void k_seg_block(const int *__restrict__ scan, const uint64_t *__restrict__ K,
                                                   void *out, const uint32_t *__restrict__ blist,
                                                   const uint32_t *__restrict__ ctl) {
    __shared__ uint64_t sk[SEG_LDS_MAX];
    const uint32_t nb = ctl[1];
    for (uint32_t b = blockIdx.x; b < nb; b += gridDim.x) {
        const int g = int(blist[b]);
        const int s = scan[g], n = scan[g + 1] - s;
        int m = SEG_WAVE_MAX * 2;
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
// (capacity: more than SEG_THREAD, resp. SEG_WAVE_MAX keys per listed group),
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
