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
//! Shapes 8-10 run the persistent pipelined kernel (OS_PIPE; keys and values
//! both staged, one workgroup per CU): 8 for 32-bit key-value pairs (and
//! 64-bit keys alone), 9 for 64-bit key-value pairs, 10 for 32-bit keys
//! alone.
constexpr int NOS = 11;
constexpr int OS_BLOCK[NOS] = { 1024, 1024, 1024, 1024, 512, 512, 1024, 1024, 1024, 1024, 1024 };
constexpr int OS_ITEMS[NOS] = { 16, 24, 26, 16, 24, 16, 20, 12, 16, 11, 20 };
constexpr bool OS_SPLIT[NOS] = { true, true, true, false, true, true, true, true,
                                 false, false, false };
constexpr int OS_WAVES[NOS] = { 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4 };
constexpr bool OS_PIPE[NOS] = { false, false, false, false, false, false, false, false,
                                true, true, true };
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

//! Untuned shape: the largest tiles that fit one workgroup per CU without
//! spilling. Measured on MI355X, 2^28 u32 pairs, ms per pass: shape 2
//! (1024 x 26, split) 0.962-0.966, shape 1 (1024 x 24) 0.967-0.968, shape 6
//! (1024 x 20) 0.994-0.995, shape 0 (1024 x 16) 1.053 (1024 x 28 spilled
//! three VGPRs: 0.964); 2^27 u64 pairs: shape 0 (1024 x 16, split) 0.746,
//! the others that fit no faster.
inline int os_default(std::size_t kbytes, bool values)
{
    return kbytes == 4 && os_feasible(2, kbytes, values) ? 2 : 0;
}

//! One digit place. The keys and values live in two or three buffers (a =
//! the caller's, b and c = scratch, c optional: null); which one the pass
//! reads is decided on the device (pass_route): passes whose digit is the
//! same for every key (trivial[p] != 0, set by the digit scan) are skipped,
//! and each executed pass reads the buffer the previous executed pass wrote.
struct onesweep_args
{
    void* keys_a;
    void* keys_b;
    void* keys_c;
    uint32_t* vals_a;
    uint32_t* vals_b;
    uint32_t* vals_c;
    uint32_t n;
    uint32_t pass;
    const uint32_t* trivial; // one flag per digit place
    const uint32_t* digit_base;
    void* status; // uint32_t words, or uint64_t when wide_status
    void* status_next;
    uint32_t* tile_counter;
    uint32_t max_groups; // workgroups of a persistent (OS_PIPE) shape: one per CU
};

//! Launches one onesweep pass with shape cfg on stream s (tiles = ceil(n /
//! os_tile(cfg)) workgroups). Defined in radix_os_*.hip.
template <typename K, int KIND>
void launch_onesweep(int cfg, bool descending, bool values, bool wide_status,
                     const onesweep_args& a, hipStream_t s);

//! The buffers an executed pass reads and writes (0 = a, 1 = b, 2 = c), from
//! the number of executed passes before it (j) and in all (e). Two buffers
//! alternate a -> b -> a ..., which leaves the result of an odd count in b
//! (the caller then copies it back). With c, an odd count of at least three
//! goes a -> b, then b -> c and c -> b in turn, and last c -> a, so the
//! result lands in a: no copy after the passes.
struct pass_route
{
    uint32_t src, dst;
    __device__ __forceinline__ pass_route(uint32_t j, uint32_t e, bool three)
    {
        if (!three || (e & 1u) == 0 || e == 1) { src = j & 1u; dst = src ^ 1u; }
        else if (j == 0) { src = 0; dst = 1; }
        else if (j + 1 == e) { src = 2; dst = 0; }
        else { src = (j & 1u) ? 1u : 2u; dst = 3u - src; }
    }
    //! Buffer i of (a, b, c), by selects (no indexed private array).
    template <typename T>
    __device__ __forceinline__ static T* pick(uint32_t i, T* a, T* b, T* c)
    {
        return i == 0 ? a : (i == 1 ? b : c);
    }
};

//! One tile per workgroup, taken in launch order from an agent-scope
//! counter, so a tile only waits on tiles that are already running (no
//! co-residency is assumed). Not persistent on purpose: on gfx9 a wave's
//! stores and loads retire in one in-order count, so a workgroup that went
//! on to a next tile would wait for its previous tile's stores before it
//! could use the new keys (measured 20% slower than fresh workgroups).
template <typename K, typename S, int KIND, bool DESC, bool VALUES, int BLOCK,
          int ITEMS, bool SPLIT, int WPE>
__global__ __launch_bounds__(BLOCK, WPE) void onesweep_kernel(
    K* keys_a, K* keys_b, K* keys_c, uint32_t* vals_a, uint32_t* vals_b, uint32_t* vals_c,
    uint32_t n, uint32_t pass, const uint32_t* __restrict__ trivial,
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
    // Executed passes before this one and in all choose the buffers.
    const uint32_t ran = ~skip & ((1u << sizeof(K)) - 1u);
    const pass_route route(__popc(ran & ((1u << pass) - 1u)), __popc(ran), keys_c != nullptr);
    const K* keys_in = pass_route::pick(route.src, keys_a, keys_b, keys_c);
    K* keys_out = pass_route::pick(route.dst, keys_a, keys_b, keys_c);
    const uint32_t* vals_in = pass_route::pick(route.src, vals_a, vals_b, vals_c);
    uint32_t* vals_out = pass_route::pick(route.dst, vals_a, vals_b, vals_c);

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

// ---- persistent, software-pipelined pass ------------------------------------
//
// One workgroup per CU keeps taking tiles in order from the counter. A tile's
// keys and values are staged in LDS by digit before its look-back (both
// arrays: no split staging), so the registers that held them are free while
// the look-back and the write-out run. The waves that do not poll load the
// next tile into those registers at once; PIPE_LB_WAVES waves run the
// look-back and only then load their part, because a wave waits for its
// memory operations in issue order and a poll queued behind a tile's loads
// would wait for all of them. The write-out stores are issued after those
// loads, so they drain while the next tile is ranked (ranking waits only for
// the loads, which are older). Ranking is VALU-bound and the transfers are
// memory-bound; the pipeline lets one tile's memory traffic run under the
// next tile's ranking on the same CU.

//! Waves that run the look-back (one digit per lane).
constexpr int PIPE_LB_WAVES = RADIX / WAVE;
//! Write-out rows the scheduler may overlap.
constexpr int PIPE_WRITE_ROWS = 4;

//! Ranks the tile held in k[] (see onesweep_rank_scatter: round by round,
//! lane order within a round), then publishes the tile's digit totals,
//! scans them, and leaves per-wave slot bases in s_hist, the totals in
//! s_total and digit_base - digit start in s_gbase.
template <typename K, typename S, int BLOCK, int ITEMS, bool FULL, typename DF>
__device__ __forceinline__ void pipe_rank(const K (&k)[ITEMS], uint32_t (&rank)[ITEMS],
                                          uint32_t n, DF dig,
                                          const uint32_t* __restrict__ digit_base,
                                          S* status, uint32_t tile, uint32_t* s_hist,
                                          uint32_t* s_total, uint32_t* s_gbase,
                                          uint32_t* s_tmp)
{
    using SB = status_bits<S>;
    constexpr int W = BLOCK / WAVE;
    constexpr uint32_t TILE = BLOCK * ITEMS;
    const uint32_t tid = threadIdx.x, lane = tid % WAVE, wave = tid / WAVE;
    const uint32_t wave_base = tile * TILE + wave * (WAVE * ITEMS);
    uint32_t* my_hist = s_hist + wave * RADIX;
    bool skewed = false;
#pragma unroll
    for (int r = 0; r < ITEMS; ++r)
    {
        const uint32_t d = dig(k[r]);
        if (r == 0) skewed = round_is_skewed(d);
        rank[r] = rank_digit(my_hist, d, FULL || wave_base + r * WAVE + lane < n, skewed);
    }
    __syncthreads();
    uint32_t total = 0;
    if (tid < RADIX)
    {
#pragma unroll
        for (int w = 0; w < W; ++w)
        {
            const uint32_t c = s_hist[w * RADIX + tid];
            s_hist[w * RADIX + tid] = total;
            total += c;
        }
        store_agent(status + static_cast<size_t>(tile) * RADIX + tid,
                    (tile == 0 ? SB::PREFIX : SB::AGG) | S(total));
    }
    const uint32_t digit_start = block256_exclusive_add<BLOCK>(total, s_tmp);
    if (tid < RADIX)
    {
#pragma unroll
        for (int w = 0; w < W; ++w) s_hist[w * RADIX + tid] += digit_start;
        s_total[tid] = total;
        s_gbase[tid] = digit_base[tid] - digit_start;
    }
    __syncthreads();
}

//! Loads one tile of a pipelined pass into registers (wave-striped, as
//! onesweep_load). With 32-bit elements and offsets (BUF): raw buffer loads
//! from the thread's offset in the tile plus the tile's offset, with the row
//! in the immediate; the buffer's range check returns 0 past n, so the last
//! tile needs no other code path.
//! Otherwise global loads, with a guarded form for the last tile.
template <typename T, int BLOCK, int ITEMS, bool BUF> struct pipe_loader
{
    const T* base;
    uint32_t n;
    __amdgpu_buffer_rsrc_t rsrc;
    uint32_t voff;
    __device__ __forceinline__ pipe_loader(const T* p, uint32_t n_) : base(p), n(n_)
    {
        if constexpr (BUF)
        {
            rsrc = __builtin_amdgcn_make_buffer_rsrc(const_cast<T*>(p), 0,
                                                     static_cast<int>(n_ * sizeof(T)),
                                                     0x00020000);
            const uint32_t lane = threadIdx.x % WAVE, wave = threadIdx.x / WAVE;
            voff = (wave * (WAVE * ITEMS) + lane) * sizeof(T);
        }
    }
    __device__ __forceinline__ void operator()(uint32_t tile, T (&x)[ITEMS]) const
    {
        constexpr uint32_t TILE = BLOCK * ITEMS;
        if constexpr (BUF)
        {
            static_assert(sizeof(T) == 4, "buffer loads of 32-bit elements");
            // The tile's offset goes in the per-lane offset, which the range
            // check covers (the scalar offset is not checked), and the row's
            // in the instruction's immediate offset (rows past 4 KiB add it).
            const uint32_t vt = voff + tile * TILE * 4;
#pragma unroll
            for (int r = 0; r < ITEMS; ++r)
                x[r] = static_cast<T>(__builtin_amdgcn_raw_buffer_load_b32(
                    rsrc, vt + r * WAVE * 4, 0, 0));
        }
        else if (tile * TILE + TILE <= n)
            onesweep_load<T, BLOCK, ITEMS, true>(base, n, tile, x);
        else
            onesweep_load<T, BLOCK, ITEMS, false>(base, n, tile, x);
    }
};

template <typename K, typename S, int KIND, bool DESC, bool VALUES, int BLOCK, int ITEMS>
__global__ __launch_bounds__(BLOCK, BLOCK / 256) void onesweep_pipe_kernel(
    K* keys_a, K* keys_b, K* keys_c, uint32_t* vals_a, uint32_t* vals_b, uint32_t* vals_c,
    uint32_t n, uint32_t pass, const uint32_t* __restrict__ trivial,
    const uint32_t* __restrict__ digit_base, S* status, S* status_next,
    uint32_t* tile_counter)
{
    static_assert(BLOCK >= 256 && BLOCK % WAVE == 0, "block shape");
    constexpr int W = BLOCK / WAVE;
    constexpr uint32_t TILE = BLOCK * ITEMS;
    // 32-bit keys below 2^30 (32-bit status words): every byte offset of
    // the key and value arrays fits 32 bits.
    constexpr bool OFF32 = sizeof(K) == 4 && sizeof(S) == 4;
    extern __shared__ __align__(16) unsigned char smem[];
    uint32_t* s_hist = reinterpret_cast<uint32_t*>(smem); // per-wave digit counters
    uint32_t* s_gbase = s_hist + W * RADIX;
    K* s_keys = reinterpret_cast<K*>(s_gbase + RADIX);
    uint32_t* s_vals = reinterpret_cast<uint32_t*>(s_keys + TILE);
    // 32-bit keys with values are staged as (key, value) pairs: one LDS
    // access per element each way, and rows that two base registers reach
    // with immediate offsets (separate arrays put the values past the 64 KiB
    // an LDS instruction's offset field spans).
    constexpr bool PAIRED = VALUES && sizeof(K) == 4;
    uint2* s_pairs = reinterpret_cast<uint2*>(s_gbase + RADIX);
    __shared__ uint32_t s_total[RADIX];
    __shared__ uint32_t s_tmp[4];
    __shared__ uint32_t s_idx[2];

    const uint32_t tid = threadIdx.x, lane = tid % WAVE, wave = tid / WAVE;
    const uint32_t tiles = (n + TILE - 1) / TILE;
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
        // The pass would be the identity: only clear the next pass's status.
        if (status_next != nullptr)
            for (size_t i = size_t(blockIdx.x) * BLOCK + tid; i < size_t(tiles) * RADIX;
                 i += size_t(gridDim.x) * BLOCK)
                status_next[i] = 0;
        return;
    }
    // Executed passes before this one and in all choose the buffers.
    const uint32_t ran = ~skip & ((1u << sizeof(K)) - 1u);
    const pass_route route(__popc(ran & ((1u << pass) - 1u)), __popc(ran), keys_c != nullptr);
    const K* keys_in = pass_route::pick(route.src, keys_a, keys_b, keys_c);
    K* keys_out = pass_route::pick(route.dst, keys_a, keys_b, keys_c);
    const uint32_t* vals_in = pass_route::pick(route.src, vals_a, vals_b, vals_c);
    uint32_t* vals_out = pass_route::pick(route.dst, vals_a, vals_b, vals_c);
    const key_digit<K, KIND, DESC> dig{ 8u * pass };
    const pipe_loader<K, BLOCK, ITEMS, OFF32> load_keys(keys_in, n);
    const pipe_loader<uint32_t, BLOCK, ITEMS, OFF32> load_vals(vals_in, n);

    for (uint32_t i = tid; i < W * RADIX; i += BLOCK) s_hist[i] = 0;

    // Iteration i ranks, stages and writes tile i-1 (loaded by iteration
    // i-1) and loads tile i; iteration 0 only loads. One load site keeps the
    // registers the loads fill the same on every path, so no copy of a
    // register still being loaded forces a wait. Tile indices come one
    // iteration ahead: thread 0 (a look-back wave) requests the index that
    // iteration i+1 loads while its own look-back polls, and leaves it in
    // s_idx[(i+1) & 1]; a request made where its result is needed at once
    // would wait for every load and store its wave has in flight.
    if (tid == 0) s_idx[0] = inc_agent(tile_counter);
    constexpr uint32_t NONE = 0xFFFFFFFFu;
    uint32_t tile = NONE;
    uint32_t iter = 0;
    K k[ITEMS];
    uint32_t v[VALUES ? ITEMS : 1];
    __syncthreads();
    for (;; ++iter)
    {
        const bool have = tile != NONE; // uniform
        const bool full = have && tile * TILE + TILE <= n;
        const uint32_t wave_base = tile * TILE + wave * (WAVE * ITEMS);
        uint32_t* my_hist = s_hist + wave * RADIX;
        if (have)
        {
            if (status_next != nullptr && tid < RADIX)
                status_next[static_cast<size_t>(tile) * RADIX + tid] = 0;
            uint32_t rank[ITEMS];
            if (full)
                pipe_rank<K, S, BLOCK, ITEMS, true>(k, rank, n, dig, digit_base, status, tile,
                                                    s_hist, s_total, s_gbase, s_tmp);
            else
                pipe_rank<K, S, BLOCK, ITEMS, false>(k, rank, n, dig, digit_base, status,
                                                     tile, s_hist, s_total, s_gbase, s_tmp);
            // Stage the tile by digit: slot = digit start + earlier waves +
            // rank.
#pragma unroll
            for (int r = 0; r < ITEMS; ++r)
            {
                if (full || wave_base + r * WAVE + lane < n)
                {
                    const uint32_t slot = my_hist[dig(k[r])] + rank[r];
                    if constexpr (PAIRED)
                        s_pairs[slot] = make_uint2(k[r], v[r]);
                    else
                    {
                        s_keys[slot] = k[r];
                        if constexpr (VALUES) s_vals[slot] = v[r];
                    }
                }
            }
            __syncthreads(); // slots written, bases read
#pragma unroll
            for (int q = 0; q < static_cast<int>(RADIX / WAVE); ++q)
                my_hist[q * WAVE + lane] = 0;
        }
        const uint32_t next = s_idx[iter & 1];
        const bool more = next < tiles;
        // The look-back waves poll before they load; the others load at once.
        if (wave < PIPE_LB_WAVES)
        {
            uint32_t req = 0;
            if (tid == 0 && more) req = inc_agent(tile_counter);
            if (have) lookback_window<S, RADIX>(status, tile, s_total, s_gbase);
            if (tid == 0 && more) s_idx[(iter + 1) & 1] = req;
        }
        if (more)
        {
            load_keys(next, k);
            if constexpr (VALUES) load_vals(next, v);
        }
        __syncthreads(); // s_gbase final, the next index visible
        if (have)
        {
            const uint32_t count = full ? TILE : n - tile * TILE;
            // Row j of the staged tile is element j * BLOCK + tid. An LDS
            // instruction's immediate offset spans 64 KiB, so rows are
            // addressed from one base per GROUP rows; the bases are made
            // opaque here, in the loop, so the compiler computes them once a
            // tile instead of hoisting one address per row out of the loop
            // into registers.
            constexpr int ELEM = PAIRED ? 8 : static_cast<int>(sizeof(K));
            constexpr int GROUP = 65536 / (BLOCK * ELEM) > 0 ? 65536 / (BLOCK * ELEM) : 1;
            constexpr int NGROUP = (ITEMS + GROUP - 1) / GROUP;
            uint32_t row_base[NGROUP];
#pragma unroll
            for (int g = 0; g < NGROUP; ++g)
            {
                row_base[g] = g * GROUP * BLOCK + tid;
                asm volatile("" : "+v"(row_base[g]));
            }
#pragma unroll
            for (int j = 0; j < ITEMS; ++j)
            {
                // At most PIPE_WRITE_ROWS rows in flight: the scheduler would
                // otherwise hoist every row's LDS reads and run out of
                // registers (the next tile's are live here).
                if (j % PIPE_WRITE_ROWS == 0) __builtin_amdgcn_sched_barrier(0);
                const uint32_t i = j * BLOCK + tid;
                const uint32_t at = row_base[j / GROUP] + (j % GROUP) * BLOCK; // == i
                if (full || i < count)
                {
                    K key;
                    uint32_t val = 0;
                    if constexpr (PAIRED)
                    {
                        const uint2 kv = s_pairs[at];
                        key = kv.x;
                        val = kv.y;
                    }
                    else
                    {
                        key = s_keys[at];
                        if constexpr (VALUES) val = s_vals[at];
                    }
                    const uint32_t dst = s_gbase[dig(key)] + i;
                    if constexpr (OFF32)
                    {
                        // From the uniform base with a 32-bit byte offset: one
                        // register, no 64-bit add.
                        const uint32_t off = dst * 4u;
                        *reinterpret_cast<K*>(reinterpret_cast<char*>(keys_out) + off) = key;
                        if constexpr (VALUES)
                            *reinterpret_cast<uint32_t*>(reinterpret_cast<char*>(vals_out)
                                                         + off) = val;
                    }
                    else
                    {
                        keys_out[dst] = key;
                        if constexpr (VALUES) vals_out[dst] = val;
                    }
                }
            }
        }
        if (!more) break;
        tile = next;
        // s_pairs / s_keys / s_vals / s_gbase are rewritten only after the
        // barriers in the next tile's ranking, which every thread reaches
        // after its write-out reads.
    }
}

namespace os_detail {

template <typename K, typename S, int KIND, bool DESC, bool VALUES, int C>
void launch(const onesweep_args& a, hipStream_t s)
{
    constexpr int BLOCK = OS_BLOCK[C], ITEMS = OS_ITEMS[C];
    const std::size_t lds = os_lds(C, sizeof(K), VALUES);
    const uint32_t tiles = (a.n + os_tile(C) - 1) / os_tile(C);
    if constexpr (OS_PIPE[C])
    {
        auto kernel = onesweep_pipe_kernel<K, S, KIND, DESC, VALUES, BLOCK, ITEMS>;
        static bool opted = false; // benign race: the attribute call is idempotent
        if (!opted)
        {
            CLC_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(kernel),
                                        hipFuncAttributeMaxDynamicSharedMemorySize,
                                        static_cast<int>(lds)));
            opted = true;
        }
        const uint32_t groups = std::max(1u, std::min(tiles, a.max_groups));
        hipLaunchKernelGGL(kernel, dim3(groups), dim3(BLOCK), lds, s,
                           static_cast<K*>(a.keys_a), static_cast<K*>(a.keys_b),
                           static_cast<K*>(a.keys_c), a.vals_a, a.vals_b, a.vals_c, a.n,
                           a.pass, a.trivial, a.digit_base,
                           static_cast<S*>(a.status), static_cast<S*>(a.status_next),
                           a.tile_counter);
        return;
    }
    auto kernel = onesweep_kernel<K, S, KIND, DESC, VALUES, BLOCK, ITEMS, OS_SPLIT[C],
                                  OS_WAVES[C]>;
    static bool opted = false; // benign race: the attribute call is idempotent
    if (!opted && lds > 65536)
    {
        CLC_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(kernel),
                                    hipFuncAttributeMaxDynamicSharedMemorySize,
                                    static_cast<int>(lds)));
        opted = true;
    }
    hipLaunchKernelGGL(kernel, dim3(tiles), dim3(BLOCK), lds, s,
                       static_cast<K*>(a.keys_a), static_cast<K*>(a.keys_b),
                       static_cast<K*>(a.keys_c), a.vals_a, a.vals_b, a.vals_c, a.n, a.pass,
                       a.trivial, a.digit_base,
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
        case 8: launch<K, S, KIND, DESC, VALUES, 8>(a, s); return;
        case 9: launch<K, S, KIND, DESC, VALUES, 9>(a, s); return;
        case 10: launch<K, S, KIND, DESC, VALUES, 10>(a, s); return;
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
