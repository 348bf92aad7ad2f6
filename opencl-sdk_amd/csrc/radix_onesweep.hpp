// The onesweep pass kernel of the plain radix sort, its tile shapes, and a
// launcher that picks a shape at run time. Instantiated per key type in
// radix_os_*.hip so the many specialisations compile in parallel.
#pragma once

#include "context.hpp"
#include "radix_tile.hpp"

namespace clc {
namespace radix {

//! Tile shapes (workgroup size x keys per thread). The sort tuner picks one
//! per key width and values flag; shape 0 is the default.
constexpr int NOS = 6;
constexpr int OS_BLOCK[NOS] = { 256, 256, 512, 512, 1024, 256 };
constexpr int OS_ITEMS[NOS] = { 16, 8, 16, 8, 8, 24 };

inline uint32_t os_tile(int cfg) { return uint32_t(OS_BLOCK[cfg]) * OS_ITEMS[cfg]; }

//! Dynamic LDS bytes of shape cfg.
inline std::size_t os_lds(int cfg, std::size_t kbytes, bool values)
{
    const std::size_t tile = os_tile(cfg);
    return std::size_t(OS_BLOCK[cfg] / 64) * 1024 + tile * kbytes
        + (values ? tile * 4 : 0) + RADIX * 4 + 64;
}

struct onesweep_args
{
    const void* keys_in;
    void* keys_out;
    const uint32_t* vals_in;
    uint32_t* vals_out;
    uint32_t n;
    uint32_t shift;
    const uint32_t* digit_base;
    void* status; // uint32_t words, or uint64_t when wide_status
    void* status_next;
    uint32_t* tile_counter;
};

//! Launches one onesweep pass with shape cfg on stream s (tiles = ceil(n /
//! os_tile(cfg)) workgroups). Defined in radix_os_*.hip.
template <typename K, int KIND>
void launch_onesweep(int cfg, bool descending, bool values, bool wide_status,
                     const onesweep_args& a, hipStream_t s);

template <typename K, typename S, int KIND, bool DESC, bool VALUES, int BLOCK,
          int ITEMS>
__global__ __launch_bounds__(BLOCK) void onesweep_kernel(
    const K* __restrict__ keys_in, K* __restrict__ keys_out,
    const uint32_t* __restrict__ vals_in, uint32_t* __restrict__ vals_out,
    uint32_t n, uint32_t shift, const uint32_t* __restrict__ digit_base,
    S* status, S* status_next, uint32_t* tile_counter)
{
    static_assert(BLOCK >= 256 && BLOCK % WAVE == 0, "block shape");
    constexpr int W = BLOCK / WAVE;
    constexpr uint32_t TILE = BLOCK * ITEMS;
    extern __shared__ __align__(16) unsigned char smem[];
    uint32_t* s_hist = reinterpret_cast<uint32_t*>(smem); // per-wave digit counters
    uint32_t* s_gbase = s_hist + W * RADIX; // global destination - tile digit start
    K* s_keys = reinterpret_cast<K*>(s_gbase + RADIX); // tile reordered by digit
    uint32_t* s_vals = reinterpret_cast<uint32_t*>(s_keys + TILE);
    __shared__ uint32_t s_tmp[4];
    __shared__ uint32_t s_tile;

    const uint32_t tid = threadIdx.x;
    if (tid == 0) s_tile = fetch_add_agent(tile_counter, 1u);
    for (uint32_t i = tid; i < W * RADIX; i += BLOCK) s_hist[i] = 0;
    __syncthreads();
    const uint32_t tile = s_tile;
    // Clear this tile's status words for the next digit place; that pass
    // starts after this one has finished, so nobody reads them meanwhile.
    if (status_next != nullptr && tid < RADIX)
        status_next[static_cast<size_t>(tile) * RADIX + tid] = 0;

    const key_digit<K, KIND, DESC> dig{ shift };
    if (tile * TILE + TILE <= n)
        onesweep_tile<K, S, VALUES, BLOCK, ITEMS, true>(
            keys_in, keys_out, vals_in, vals_out, n, dig, digit_base, status, tile,
            s_hist, s_keys, s_vals, s_gbase, s_tmp);
    else
        onesweep_tile<K, S, VALUES, BLOCK, ITEMS, false>(
            keys_in, keys_out, vals_in, vals_out, n, dig, digit_base, status, tile,
            s_hist, s_keys, s_vals, s_gbase, s_tmp);
}

namespace os_detail {

template <typename K, typename S, int KIND, bool DESC, bool VALUES, int C>
void launch(const onesweep_args& a, hipStream_t s)
{
    constexpr int BLOCK = OS_BLOCK[C], ITEMS = OS_ITEMS[C];
    auto kernel = onesweep_kernel<K, S, KIND, DESC, VALUES, BLOCK, ITEMS>;
    const std::size_t lds = os_lds(C, sizeof(K), VALUES);
    static bool opted = false; // benign race: the attribute call is idempotent
    if (!opted && lds > 65536)
    {
        CLC_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(kernel),
                                    hipFuncAttributeMaxDynamicSharedMemorySize,
                                    static_cast<int>(lds)));
        opted = true;
    }
    const uint32_t tiles = (a.n + os_tile(C) - 1) / os_tile(C);
    hipLaunchKernelGGL(kernel, dim3(tiles), dim3(BLOCK), lds, s,
                       static_cast<const K*>(a.keys_in), static_cast<K*>(a.keys_out),
                       a.vals_in, a.vals_out, a.n, a.shift, a.digit_base,
                       static_cast<S*>(a.status), static_cast<S*>(a.status_next),
                       a.tile_counter);
}

template <typename K, typename S, int KIND, bool DESC, bool VALUES>
void by_cfg(int cfg, const onesweep_args& a, hipStream_t s)
{
    switch (cfg)
    {
        case 0: launch<K, S, KIND, DESC, VALUES, 0>(a, s); return;
        case 1: launch<K, S, KIND, DESC, VALUES, 1>(a, s); return;
        case 2: launch<K, S, KIND, DESC, VALUES, 2>(a, s); return;
        case 3: launch<K, S, KIND, DESC, VALUES, 3>(a, s); return;
        case 4: launch<K, S, KIND, DESC, VALUES, 4>(a, s); return;
        case 5: launch<K, S, KIND, DESC, VALUES, 5>(a, s); return;
    }
    throw std::invalid_argument("radix sort: unknown tile shape");
}

} // namespace os_detail

template <typename K, int KIND>
void launch_onesweep(int cfg, bool descending, bool values, bool wide_status,
                     const onesweep_args& a, hipStream_t s)
{
    using namespace os_detail;
#define CLC_OS(S, D, V) by_cfg<K, S, KIND, D, V>(cfg, a, s)
    if (wide_status)
    {
        if (descending) { if (values) CLC_OS(uint64_t, true, true); else CLC_OS(uint64_t, true, false); }
        else { if (values) CLC_OS(uint64_t, false, true); else CLC_OS(uint64_t, false, false); }
    }
    else
    {
        if (descending) { if (values) CLC_OS(uint32_t, true, true); else CLC_OS(uint32_t, true, false); }
        else { if (values) CLC_OS(uint32_t, false, true); else CLC_OS(uint32_t, false, false); }
    }
#undef CLC_OS
}

} // namespace radix
} // namespace clc
