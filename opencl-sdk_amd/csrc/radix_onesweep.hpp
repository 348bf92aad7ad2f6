// The onesweep pass kernel of the plain radix sort, its tile shapes, and a
// launcher that picks a shape at run time. Instantiated per key type in
// radix_os_*.hip so the many specialisations compile in parallel.
#pragma once

#include "context.hpp"
#include "radix_tile.hpp"

namespace clc {
namespace radix {

//! Tile shapes: workgroup size x keys per thread, split staging, and the
//! waves per SIMD the register allocation must leave room for (4 with
//! 512-thread groups = two groups per CU when their LDS fits, so one ranks
//! while the other waits on its look-back or on HBM). The sort tuner picks
//! one per key width and values flag; os_default() is used untuned.
constexpr int NOS = 8;
constexpr int OS_BLOCK[NOS] = { 1024, 1024, 1024, 1024, 512, 512, 1024, 1024 };
constexpr int OS_ITEMS[NOS] = { 16, 24, 28, 16, 24, 16, 20, 12 };
constexpr bool OS_SPLIT[NOS] = { true, true, true, false, true, true, true, true };
constexpr int OS_WAVES[NOS] = { 4, 4, 4, 4, 4, 4, 4, 4 };
//! LDS a workgroup may use on gfx950.
constexpr std::size_t OS_LDS_LIMIT = 160 * 1024;

inline uint32_t os_tile(int cfg) { return uint32_t(OS_BLOCK[cfg]) * OS_ITEMS[cfg]; }

//! Dynamic LDS bytes of shape cfg: per-wave counters, destination bases and
//! the staging buffer (keys and values together, or one after the other).
inline std::size_t os_lds(int cfg, std::size_t kbytes, bool values)
{
    const std::size_t tile = os_tile(cfg);
    const std::size_t stage =
        OS_SPLIT[cfg] ? tile * kbytes : tile * kbytes + (values ? tile * 4 : 0);
    return std::size_t(OS_BLOCK[cfg] / 64) * 1024 + RADIX * 4 + stage + 64;
}

//! Static LDS of the pass kernel (digit totals, scan scratch), with margin.
constexpr std::size_t OS_STATIC_LDS = 2304;

//! Whether shape cfg fits in LDS for this key width (shape 0 always does).
inline bool os_feasible(int cfg, std::size_t kbytes, bool values)
{
    return cfg >= 0 && cfg < NOS
        && os_lds(cfg, kbytes, values) + OS_STATIC_LDS <= OS_LDS_LIMIT;
}

//! Untuned shape: the largest tiles that fit one workgroup per CU. Measured
//! on MI355X, 2^28 u32 pairs: shape 6 (1024 x 20, split) 58.1 GKeys/s,
//! shape 3 56.3, shape 0 53.4, every two-groups-per-CU shape 42.8-51.8;
//! 2^27 u64 pairs: shape 0 (1024 x 16, split) 27.3, the rest 19.2-24.1.
inline int os_default(std::size_t kbytes, bool values)
{
    return kbytes == 4 && os_feasible(6, kbytes, values) ? 6 : 0;
}

//! One digit place. The keys and values live in two buffers (a = the
//! caller's, b = scratch); which one the pass reads is decided on the device:
//! passes whose digit is the same for every key (trivial[p] != 0, set by the
//! digit scan) are skipped, and each executed pass reads the buffer the
//! previous executed pass wrote.
struct onesweep_args
{
    void* keys_a;
    void* keys_b;
    uint32_t* vals_a;
    uint32_t* vals_b;
    uint32_t n;
    uint32_t pass;
    const uint32_t* trivial; // one flag per digit place
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

//! One tile per workgroup, taken in launch order from an agent-scope
//! counter, so a tile only waits on tiles that are already running (no
//! co-residency is assumed). Not persistent on purpose: on gfx9 a wave's
//! stores and loads retire in one in-order count, so a workgroup that went
//! on to a next tile would wait for its previous tile's stores before it
//! could use the new keys (measured 20% slower than fresh workgroups).
template <typename K, typename S, int KIND, bool DESC, bool VALUES, int BLOCK,
          int ITEMS, bool SPLIT, int WPE>
__global__ __launch_bounds__(BLOCK, WPE) void onesweep_kernel(
    K* keys_a, K* keys_b, uint32_t* vals_a, uint32_t* vals_b, uint32_t n,
    uint32_t pass, const uint32_t* __restrict__ trivial,
    const uint32_t* __restrict__ digit_base, S* status, S* status_next,
    uint32_t* tile_counter)
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
    // The tile index and every place's flag are requested together (a
    // skipped pass wastes an index of its own counter, which nothing reads).
    if (tid == 0) s_tile = fetch_add_agent(tile_counter, 1u);
    uint32_t skip = 0; // bit p: digit place p is constant
#pragma unroll
    for (int q = 0; q < static_cast<int>(sizeof(K)) / 4; ++q)
    {
        const uint4 f = reinterpret_cast<const uint4*>(trivial)[q];
        skip |= ((f.x != 0 ? 1u : 0u) | (f.y != 0 ? 2u : 0u) | (f.z != 0 ? 4u : 0u)
                 | (f.w != 0 ? 8u : 0u)) << (4 * q);
    }
    if ((skip >> pass) & 1u)
    {
        // Every key has the same digit here: the pass would be the identity.
        // Only clear the next pass's status words (one row per workgroup).
        if (status_next != nullptr && tid < RADIX)
            status_next[static_cast<size_t>(blockIdx.x) * RADIX + tid] = 0;
        return;
    }
    // Executed passes before this one, mod 2, choose the buffers.
    const uint32_t parity = __popc(~skip & ((1u << pass) - 1u)) & 1u;
    const K* keys_in = parity ? keys_b : keys_a;
    K* keys_out = parity ? keys_a : keys_b;
    const uint32_t* vals_in = parity ? vals_b : vals_a;
    uint32_t* vals_out = parity ? vals_a : vals_b;

    for (uint32_t i = tid; i < W * RADIX; i += BLOCK) s_hist[i] = 0;
    __syncthreads();
    const uint32_t tile = s_tile;
    // Clear this tile's status words for the next digit place; that pass
    // starts after this one has finished, so nobody reads them meanwhile.
    if (status_next != nullptr && tid < RADIX)
        status_next[static_cast<size_t>(tile) * RADIX + tid] = 0;

    const key_digit<K, KIND, DESC> dig{ 8u * pass };
    if (tile * TILE + TILE <= n)
        onesweep_tile<K, S, VALUES, BLOCK, ITEMS, true, SPLIT>(
            keys_in, keys_out, vals_in, vals_out, n, dig, digit_base, status, tile,
            s_hist, s_keys, s_vals, s_gbase, s_tmp);
    else
        onesweep_tile<K, S, VALUES, BLOCK, ITEMS, false, SPLIT>(
            keys_in, keys_out, vals_in, vals_out, n, dig, digit_base, status, tile,
            s_hist, s_keys, s_vals, s_gbase, s_tmp);
}

namespace os_detail {

template <typename K, typename S, int KIND, bool DESC, bool VALUES, int C>
void launch(const onesweep_args& a, hipStream_t s)
{
    constexpr int BLOCK = OS_BLOCK[C], ITEMS = OS_ITEMS[C];
    auto kernel = onesweep_kernel<K, S, KIND, DESC, VALUES, BLOCK, ITEMS, OS_SPLIT[C],
                                  OS_WAVES[C]>;
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
                       static_cast<K*>(a.keys_a), static_cast<K*>(a.keys_b), a.vals_a,
                       a.vals_b, a.n, a.pass, a.trivial, a.digit_base,
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
        case 6: launch<K, S, KIND, DESC, VALUES, 6>(a, s); return;
        case 7: launch<K, S, KIND, DESC, VALUES, 7>(a, s); return;
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

// Instantiated once each, in radix_os_<type>.hip; other translation units
// only call them.
#define CLC_OS_EXTERN(K, KIND)                                                       \
    extern template void launch_onesweep<K, KIND>(int, bool, bool, bool,             \
                                                  const onesweep_args&, hipStream_t)
CLC_OS_EXTERN(uint32_t, dev::KEY_UNSIGNED);
CLC_OS_EXTERN(uint32_t, dev::KEY_SIGNED);
CLC_OS_EXTERN(uint32_t, dev::KEY_FLOAT);
CLC_OS_EXTERN(uint64_t, dev::KEY_UNSIGNED);
CLC_OS_EXTERN(uint64_t, dev::KEY_SIGNED);
CLC_OS_EXTERN(uint64_t, dev::KEY_FLOAT);
#undef CLC_OS_EXTERN

} // namespace radix
} // namespace clc
