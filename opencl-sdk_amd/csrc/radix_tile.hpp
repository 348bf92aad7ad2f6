// Building blocks of the gfx950 radix sorts, shared by the plain sort
// (radix_sort.hip) and the segmented sort (segmented_sort.hip): the
// look-back status words, the stable ranking of 8-bit digits (LDS atomics,
// or a wave64 ballot match for rounds of mostly equal digits), and the
// processing of one onesweep tile.
#pragma once

#include "device_common.hpp"

namespace clc {
namespace radix {

using namespace dev;

constexpr uint32_t RADIX = 256;

template <typename S> struct status_bits
{
    static constexpr int SHIFT = sizeof(S) * 8 - 2;
    static constexpr S AGG = S(1) << SHIFT; // tile's own count published
    static constexpr S PREFIX = S(2) << SHIFT; // count of tiles 0..t published
    static constexpr S FLAGS = S(3) << SHIFT;
    static constexpr S VALUE = ~FLAGS;
};

//! Lanes of the wave whose digit equals this lane's: 8 ballots, one per
//! digit bit. m is the lane's bit b sign-extended to all ones or zero, so
//! ballot ^ m has a 1 exactly at the lanes that differ in bit b; OR over the
//! bits leaves 0 at the peers.
__device__ __forceinline__ uint64_t match_digit(uint32_t d)
{
    // All eight ballots are taken before they are folded, so each lands in
    // an SGPR pair of its own and the folds read ballots written a few
    // instructions earlier: a ballot folded at once (through VCC) costs
    // hazard wait states on gfx950 (s_nop in the code per bit: 512 -> 92 in
    // the u32 pair pass).
    uint32_t m[8];
    uint64_t bal[8];
#pragma unroll
    for (int b = 0; b < 8; ++b)
    {
        m[b] = static_cast<uint32_t>(static_cast<int32_t>(d << (31 - b)) >> 31);
        // Opaque to the optimiser, so the ballot compares m itself (one
        // v_bfe_i32 + one v_cmp per bit, instead of re-deriving the bit with
        // a second shift) and each half folds into one three-input op.
        asm("" : "+v"(m[b]));
        bal[b] = __ballot(m[b] != 0);
    }
    uint32_t diff_lo = 0, diff_hi = 0;
#pragma unroll
    for (int b = 0; b < 7; ++b)
    {
        diff_lo |= static_cast<uint32_t>(bal[b]) ^ m[b];
        diff_hi |= static_cast<uint32_t>(bal[b] >> 32) ^ m[b];
        // Keeps the accumulation a chain of diff | (bal ^ m) steps, each one
        // v_bitop3, rather than XORs re-associated into OR trees.
        asm("" : "+v"(diff_lo), "+v"(diff_hi));
    }
    // The last fold also complements: ~(diff | (bal ^ m)) is one v_bitop3
    // per half (a separate NOT cost two instructions a row: 2% of the pass).
    const uint32_t lo = ~(diff_lo | (static_cast<uint32_t>(bal[7]) ^ m[7]));
    const uint32_t hi = ~(diff_hi | (static_cast<uint32_t>(bal[7] >> 32) ^ m[7]));
    return (static_cast<uint64_t>(hi) << 32) | lo;
}

//! Lanes of a round on one digit from which a wave ranks by the ballot match
//! instead of LDS atomics: an LDS atomic serves the lanes that hit one
//! address one after another, so rounds of mostly equal digits (the
//! constant-ish high bytes of skewed keys, runs of duplicates) cost about
//! one LDS cycle a lane there, while the match costs the same for any
//! digits.
constexpr uint32_t SKEW_LANES = 16;

//! Whether a wave ranks its onesweep tile by the match: decided once, from
//! its first round, by the lanes sharing the first or the last lane's digit
//! (a test every round cost 0.5% of the uniform u32 pass).
__device__ __forceinline__ bool round_is_skewed(uint32_t d)
{
    const uint32_t a = __builtin_amdgcn_readfirstlane(d);
    const uint32_t b = __builtin_amdgcn_readlane(d, WAVE - 1);
    return __popcll(__ballot(d == a)) >= SKEW_LANES || __popcll(__ballot(d == b)) >= SKEW_LANES;
}

//! Stable rank of this lane's digit d among the wave's keys so far, and the
//! wave's counter of d advanced past this round: the number of earlier
//! keys of the wave with digit d (my_hist[d] before the round) plus the
//! valid lanes below this one with digit d. Invalid lanes neither count
//! nor get a rank. With skewed (wave-uniform) the ballot match ranks the
//! round; otherwise one LDS atomic add per lane, which gfx950 serves in
//! lane order for lanes that hit the same address (so equal digits rank in
//! lane order; a wave's rounds follow in program order).
__device__ __forceinline__ uint32_t rank_digit(uint32_t* my_hist, uint32_t d, bool valid,
                                               bool skewed)
{
    if (!skewed)
        return valid ? __hip_atomic_fetch_add(&my_hist[d], 1u, __ATOMIC_RELAXED,
                                              __HIP_MEMORY_SCOPE_WORKGROUP)
                     : 0u;
    const uint64_t peers = match_digit(d) & __ballot(valid);
    const uint32_t pre = count_below(peers);
    const uint32_t base = my_hist[d];
    if (valid && pre == 0) my_hist[d] = base + __popcll(peers);
    return base + pre;
}

//! Digit of a key for one pass: byte shift/8 of the key's order image.
template <typename K, int KIND, bool DESC> struct key_digit
{
    //! Cheap to recompute: no digits are staged in LDS (see
    //! onesweep_rank_scatter for functors that stage them).
    static constexpr bool STAGED = false;
    uint32_t shift;
    __device__ __forceinline__ uint32_t operator()(K k) const
    {
        return digit_of<K, KIND, DESC>(k, shift);
    }
};

// ---- one digit place --------------------------------------------------------

//! Decoupled look-back of one tile over a window of predecessors. Wave w
//! owns digits [w * DPW, (w + 1) * DPW); lane (dd, j) = (lane / J, lane % J)
//! reads digit w * DPW + dd of the tiles `consumed + q * J + j + 1` places
//! back for q < R, so one round of R parallel loads per lane covers the
//! WIN = J * R = 16 nearest unsummed predecessors of every digit (a
//! single-thread walk would take WIN round trips). A digit sums the words
//! before the first unpublished one and stops at its first PREFIX; it then
//! publishes this tile's inclusive PREFIX (total from s_total) and adds its
//! exclusive prefix to s_gbase. Every wave of the block calls it.
template <typename S, int BLOCK>
__device__ __forceinline__ void lookback_window(S* status, uint32_t tile,
                                                const uint32_t* s_total, uint32_t* s_gbase)
{
    using SB = status_bits<S>;
    constexpr uint32_t J = BLOCK / RADIX;
    constexpr uint32_t DPW = WAVE / J;
    constexpr uint32_t R = 16 / J;
    constexpr uint32_t WIN = J * R;
    static_assert(J * RADIX == BLOCK && (J == 1 || J == 2 || J == 4), "block of 256/512/1024");
    const uint32_t lane = threadIdx.x % WAVE, wave = threadIdx.x / WAVE;
    const uint32_t dd = lane / J, j = lane % J;
    const uint32_t d = wave * DPW + dd;
    S excl = 0;
    uint32_t consumed = 0; // predecessors already summed; the same in a digit's J lanes
    bool done = tile == 0;
    while (__ballot(!done) != 0)
    {
        S w[R];
#pragma unroll
        for (uint32_t q = 0; q < R; ++q)
        {
            w[q] = SB::PREFIX; // before tile 0: an empty prefix
            const uint32_t back = consumed + q * J + j + 1;
            if (!done && back <= tile)
                w[q] = load_agent(status + static_cast<size_t>(tile - back) * RADIX + d);
        }
        // Window position p = q * J + j: bit p of the digit's masks.
        uint32_t gp = 0, gz = 0;
#pragma unroll
        for (uint32_t q = 0; q < R; ++q)
        {
            const uint64_t pm = __ballot(!done && (w[q] & SB::PREFIX) != 0);
            const uint64_t zm = __ballot(!done && (w[q] & SB::FLAGS) == 0);
            gp |= (static_cast<uint32_t>(pm >> (dd * J)) & ((1u << J) - 1)) << (q * J);
            gz |= (static_cast<uint32_t>(zm >> (dd * J)) & ((1u << J) - 1)) << (q * J);
        }
        const uint32_t jp = gp ? __builtin_ctz(gp) : WIN; // first PREFIX in the window
        const uint32_t jz = gz ? __builtin_ctz(gz) : WIN; // first unpublished word
        const uint32_t take = jz < jp ? jz : (jp < WIN ? jp + 1 : WIN);
        S part = 0;
#pragma unroll
        for (uint32_t q = 0; q < R; ++q)
            if (q * J + j < take) part += w[q] & SB::VALUE;
        if constexpr (J >= 2) part += __shfl_xor(part, 1);
        if constexpr (J >= 4) part += __shfl_xor(part, 2);
        if (!done)
        {
            excl += part;
            consumed += take;
            done = jp < jz;
        }
        // A digit stopped at an unpublished word: back off briefly.
        if (__ballot(!done && jz < WIN) != 0) __builtin_amdgcn_s_sleep(1);
    }
    if (j == 0)
    {
        if (tile > 0)
            store_agent(status + static_cast<size_t>(tile) * RADIX + d,
                        SB::PREFIX | (excl + S(s_total[d])));
        s_gbase[d] += static_cast<uint32_t>(excl);
    }
}

//! Loads one tile's keys into registers, wave-striped: round r of a wave
//! covers 64 consecutive keys, so ranking round by round ranks in input
//! order. In the guarded (last) tile, lanes past n get 0.
template <typename K, int BLOCK, int ITEMS, bool FULL>
__device__ __forceinline__ void onesweep_load(const K* __restrict__ in, uint32_t n,
                                              uint32_t tile, K (&x)[ITEMS])
{
    constexpr uint32_t TILE = BLOCK * ITEMS;
    const uint32_t lane = threadIdx.x % WAVE, wave = threadIdx.x / WAVE;
    const uint32_t wave_base = tile * TILE + wave * (WAVE * ITEMS);
    // One 64-bit base per thread and constant row offsets: indexing with
    // the 32-bit idx of each row makes the compiler widen every row's
    // address on its own (it cannot rule out a 32-bit wrap), two registers
    // and a 64-bit add a row.
    const K* p = in + (static_cast<size_t>(wave_base) + lane);
#pragma unroll
    for (int r = 0; r < ITEMS; ++r)
    {
        if constexpr (FULL)
            x[r] = p[r * WAVE];
        else
            x[r] = wave_base + r * WAVE + lane < n ? p[r * WAVE] : K(0);
    }
}

//! One tile of one digit place whose keys (and values) are already in
//! registers (k, v, loaded by onesweep_load). FULL tiles skip every bounds
//! check; only the last tile of the array runs the guarded form. With SPLIT,
//! keys and values are reordered one after the other through the same LDS
//! buffer (s_vals unused; each thread keeps its slots, and for 32-bit keys
//! the digits of the slots it writes, in registers), which halves the
//! staging LDS of key-value tiles so larger tiles (longer contiguous digit
//! runs in HBM) fit.
//!
//! db is the digit place's global start of digit tid (threads tid < 256;
//! loaded by the caller before the tile, so its latency is hidden).
//!
//! A digit functor with STAGED = true is costly to evaluate: each key's
//! digit is then computed once, kept (four to a register) for the reorder,
//! and stored as a byte at its slot in dig.s_dig (TILE bytes of LDS) for
//! the write-out, instead of being recomputed from the key three times.
template <typename K, typename S, bool VALUES, int BLOCK, int ITEMS, bool FULL,
          bool SPLIT, typename DF>
__device__ __forceinline__ void onesweep_rank_scatter(
    K (&k)[ITEMS], uint32_t (&v)[VALUES ? ITEMS : 1], K* __restrict__ keys_out,
    uint32_t* __restrict__ vals_out, uint32_t n, DF dig, uint32_t db, S* status,
    uint32_t tile, uint32_t* s_hist, K* s_keys, uint32_t* s_vals, uint32_t* s_gbase,
    uint32_t* s_tmp)
{
    using SB = status_bits<S>;
    constexpr int W = BLOCK / WAVE;
    constexpr uint32_t TILE = BLOCK * ITEMS;
    const uint32_t tid = threadIdx.x, lane = tid % WAVE, wave = tid / WAVE;
    const uint32_t tile_base = tile * TILE;
    const uint32_t wave_base = tile_base + wave * (WAVE * ITEMS);

    // Rank inside the wave: round by round, each key's rank among equal
    // digits is the per-wave running count of its digit (rank_digit).
    constexpr bool STAGED = DF::STAGED;
    uint32_t rank[ITEMS];
    uint32_t dk[STAGED ? (ITEMS + 3) / 4 : 1]; // digits of k[], four per word
    uint32_t* my_hist = s_hist + wave * RADIX;
    bool skewed = false;
#pragma unroll
    for (int r = 0; r < ITEMS; ++r)
    {
        const uint32_t d = dig(k[r]);
        if constexpr (STAGED)
            dk[r / 4] = r % 4 == 0 ? d : dk[r / 4] | (d << (8 * (r % 4)));
        if (r == 0) skewed = round_is_skewed(d);
        rank[r] = rank_digit(my_hist, d, FULL || wave_base + r * WAVE + lane < n, skewed);
    }
    __syncthreads();

    // Per digit (thread d): tile total, wave-exclusive prefixes, publish.
    __shared__ uint32_t s_total[RADIX];
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
        S* mine = status + static_cast<size_t>(tile) * RADIX + tid;
        store_agent(mine, (tile == 0 ? SB::PREFIX : SB::AGG) | S(total));
    }
    const uint32_t digit_start = block256_exclusive_add<BLOCK>(total, s_tmp);
    if (tid < RADIX)
    {
#pragma unroll
        for (int w = 0; w < W; ++w) s_hist[w * RADIX + tid] += digit_start;
        s_total[tid] = total;
        s_gbase[tid] = db - digit_start; // the look-back adds the prefix
    }
    __syncthreads();

    // Reorder the tile in LDS: slot = digit start + earlier waves + rank.
#pragma unroll
    for (int r = 0; r < ITEMS; ++r)
    {
        if (FULL || wave_base + r * WAVE + lane < n)
        {
            uint32_t d;
            if constexpr (STAGED) d = (dk[r / 4] >> (8 * (r % 4))) & 0xFFu;
            else d = dig(k[r]);
            const uint32_t slot = my_hist[d] + rank[r];
            if constexpr (STAGED) dig.s_dig[slot] = static_cast<uint8_t>(d);
            s_keys[slot] = k[r];
            if constexpr (VALUES && !SPLIT) s_vals[slot] = v[r];
            if constexpr (SPLIT) rank[r] = slot; // kept for the value pass
        }
    }
    // Nothing reads this wave's counters any more: clear them for the tile
    // a persistent caller processes next.
#pragma unroll
    for (int q = 0; q < static_cast<int>(RADIX / WAVE); ++q) my_hist[q * WAVE + lane] = 0;

    lookback_window<S, BLOCK>(status, tile, s_total, s_gbase);
    __syncthreads();

    // Write the reordered tile: consecutive threads take consecutive slots,
    // which land in contiguous runs of one digit each.
    const uint32_t count = FULL ? TILE : n - tile_base;
    auto digit_at = [&](uint32_t i, K key) -> uint32_t {
        if constexpr (STAGED) return dig.s_dig[i];
        else return dig(key);
    };
    // 32-bit keys with 32-bit status words (fewer than 2^30 keys): every byte
    // offset into the key and value arrays fits 32 bits, and a store from the
    // uniform base with a 32-bit vector offset needs no 64-bit address add.
    constexpr bool OFF32 = sizeof(K) == 4 && sizeof(S) == 4;
    auto store_off = [](auto* base, uint32_t off, auto x) {
        *reinterpret_cast<decltype(x)*>(reinterpret_cast<char*>(base) + off) = x;
    };
    if constexpr (SPLIT && VALUES && sizeof(K) == 8)
    {
        // An 8-byte slot holds two words once its key is written out: the
        // key's destination (low) and, after the value pass, the value that
        // belongs there (high). No destinations stay in registers.
        uint32_t* s_w = reinterpret_cast<uint32_t*>(s_keys);
#pragma unroll
        for (int j = 0; j < ITEMS; ++j)
        {
            const uint32_t i = j * BLOCK + tid;
            if (FULL || i < count)
            {
                const K key = s_keys[i];
                const uint32_t dst = s_gbase[digit_at(i, key)] + i;
                keys_out[dst] = key;
                s_w[2 * i] = dst;
            }
        }
        __syncthreads(); // every destination is in LDS
#pragma unroll
        for (int r = 0; r < ITEMS; ++r)
            if (FULL || wave_base + r * WAVE + lane < n) s_w[2 * rank[r] + 1] = v[r];
        __syncthreads();
#pragma unroll
        for (int j = 0; j < ITEMS; ++j)
        {
            const uint32_t i = j * BLOCK + tid;
            if (FULL || i < count)
            {
                const uint2 dv = reinterpret_cast<const uint2*>(s_keys)[i];
                vals_out[dv.x] = dv.y;
            }
        }
    }
    else if constexpr (SPLIT && VALUES && OFF32)
    {
        // A thread writes the same slots (j * BLOCK + tid) in both loops: it
        // keeps their byte offsets for the value pass (one register each).
        uint32_t off[ITEMS];
#pragma unroll
        for (int j = 0; j < ITEMS; ++j)
        {
            const uint32_t i = j * BLOCK + tid;
            if (FULL || i < count)
            {
                const K key = s_keys[i];
                off[j] = (s_gbase[digit_at(i, key)] + i) * 4u;
                store_off(keys_out, off[j], key);
            }
        }
        __syncthreads(); // keys out of the buffer: reuse it for the values
        uint32_t* s_v = reinterpret_cast<uint32_t*>(s_keys);
#pragma unroll
        for (int r = 0; r < ITEMS; ++r)
            if (FULL || wave_base + r * WAVE + lane < n) s_v[rank[r]] = v[r];
        __syncthreads();
#pragma unroll
        for (int j = 0; j < ITEMS; ++j)
        {
            const uint32_t i = j * BLOCK + tid;
            if (FULL || i < count) store_off(vals_out, off[j], s_v[i]);
        }
    }
    else if constexpr (SPLIT && VALUES)
    {
        // A thread writes the same slots (j * BLOCK + tid) in both loops:
        // it keeps their digits, four to a register, for the value pass.
        uint32_t dg[STAGED ? 1 : (ITEMS + 3) / 4];
#pragma unroll
        for (int q = 0; q < (STAGED ? 1 : (ITEMS + 3) / 4); ++q) dg[q] = 0;
#pragma unroll
        for (int j = 0; j < ITEMS; ++j)
        {
            const uint32_t i = j * BLOCK + tid;
            if (FULL || i < count)
            {
                const K key = s_keys[i];
                const uint32_t d = digit_at(i, key);
                if constexpr (!STAGED) dg[j / 4] |= d << (8 * (j % 4));
                keys_out[s_gbase[d] + i] = key;
            }
        }
        __syncthreads(); // keys out of the buffer: reuse it for the values
        uint32_t* s_v = reinterpret_cast<uint32_t*>(s_keys);
#pragma unroll
        for (int r = 0; r < ITEMS; ++r)
            if (FULL || wave_base + r * WAVE + lane < n) s_v[rank[r]] = v[r];
        __syncthreads();
#pragma unroll
        for (int j = 0; j < ITEMS; ++j)
        {
            const uint32_t i = j * BLOCK + tid;
            if (FULL || i < count)
            {
                uint32_t d;
                if constexpr (STAGED) d = dig.s_dig[i];
                else d = (dg[j / 4] >> (8 * (j % 4))) & 0xFFu;
                vals_out[s_gbase[d] + i] = s_v[i];
            }
        }
    }
    else
    {
#pragma unroll
        for (int j = 0; j < ITEMS; ++j)
        {
            const uint32_t i = j * BLOCK + tid;
            if (FULL || i < count)
            {
                const K key = s_keys[i];
                const uint32_t dst = s_gbase[digit_at(i, key)] + i;
                if constexpr (OFF32)
                {
                    store_off(keys_out, dst * 4u, key);
                    if constexpr (VALUES) store_off(vals_out, dst * 4u, s_vals[i]);
                }
                else
                {
                    keys_out[dst] = key;
                    if constexpr (VALUES) vals_out[dst] = s_vals[i];
                }
            }
        }
    }
}

//! One tile of one digit place, loads included (the non-persistent form).
template <typename K, typename S, bool VALUES, int BLOCK, int ITEMS, bool FULL,
          bool SPLIT, typename DF>
__device__ __forceinline__ void onesweep_tile(
    const K* __restrict__ keys_in, K* __restrict__ keys_out,
    const uint32_t* __restrict__ vals_in, uint32_t* __restrict__ vals_out,
    uint32_t n, DF dig, const uint32_t* __restrict__ digit_base,
    S* status, uint32_t tile, uint32_t* s_hist, K* s_keys, uint32_t* s_vals,
    uint32_t* s_gbase, uint32_t* s_tmp)
{
    // Issued before the tile's loads: the digit scan, which uses it, then
    // finds it long arrived (read there, its latency stalled the scan).
    const uint32_t db = threadIdx.x < RADIX ? digit_base[threadIdx.x] : 0u;
    K k[ITEMS];
    uint32_t v[VALUES ? ITEMS : 1];
    onesweep_load<K, BLOCK, ITEMS, FULL>(keys_in, n, tile, k);
    // Values are loaded with the keys: loaded later, their latency shows
    // (measured slower), and any load issued before the look-back delays its
    // polls (a wave waits for its vector memory operations in issue order).
    if constexpr (VALUES) onesweep_load<uint32_t, BLOCK, ITEMS, FULL>(vals_in, n, tile, v);
    onesweep_rank_scatter<K, S, VALUES, BLOCK, ITEMS, FULL, SPLIT>(
        k, v, keys_out, vals_out, n, dig, db, status, tile, s_hist, s_keys, s_vals, s_gbase,
        s_tmp);
}

// ---- whole segment in LDS ----------------------------------------------------

//! Exclusive scan of the 256 digit totals held in LDS by a block of BLOCK
//! threads (64..1024). Each of the first min(BLOCK, 256) threads owns
//! 256 / min(BLOCK, 256) consecutive digits. Turns the per-wave counters
//! s_hist[w][d] into slot bases: digit start + counts of earlier waves.
template <int BLOCK>
__device__ __forceinline__ void digit_bases(uint32_t* s_hist, uint32_t* s_tmp)
{
    constexpr int W = BLOCK / WAVE;
    constexpr int T = BLOCK < 256 ? BLOCK : 256; // threads owning digits
    constexpr int DPT = RADIX / T; // digits per owning thread
    constexpr int TW = T / WAVE; // waves owning digits
    const uint32_t tid = threadIdx.x, lane = tid % WAVE, wave = tid / WAVE;
    uint32_t tot[DPT];
    uint32_t local = 0;
    if (tid < T)
    {
#pragma unroll
        for (int j = 0; j < DPT; ++j)
        {
            const uint32_t d = tid * DPT + j;
            uint32_t run = 0;
#pragma unroll
            for (int w = 0; w < W; ++w)
            {
                const uint32_t c = s_hist[w * RADIX + d];
                s_hist[w * RADIX + d] = run;
                run += c;
            }
            tot[j] = run;
            local += run;
        }
    }
    const uint32_t inc = wave_inclusive_add(local);
    // One wave owns every digit when the block is one wave: no exchange.
    if constexpr (TW > 1)
    {
        if (tid < T && lane == WAVE - 1) s_tmp[wave] = inc;
        __syncthreads();
    }
    if (tid < T)
    {
        uint32_t start = inc - local;
#pragma unroll
        for (int w = 0; w < TW; ++w)
            if (w < static_cast<int>(wave)) start += s_tmp[w];
#pragma unroll
        for (int j = 0; j < DPT; ++j)
        {
            const uint32_t d = tid * DPT + j;
#pragma unroll
            for (int w = 0; w < W; ++w) s_hist[w * RADIX + d] += start;
            start += tot[j];
        }
    }
    __syncthreads();
}

//! One stable LDS counting pass of a block over words w[] (row r < rows of
//! wave `wave` holds position wave * 64 * rows + r * 64 + lane) by the
//! nb-bit digit (w >> shift) & (2^nb - 1): rank_digit ranks,
//! per-wave digit counters, an LDS reorder. Only positions below m are
//! keys; rows of a wave that hold none are skipped, and the other positions
//! past m take the largest digit, so they end behind every key (stability).
//! With MOVE set, values v[] follow the words through s_v in the same
//! reorder. The counters must be zero on entry; each wave clears its own
//! after reading its rows back, so they are all zero again once every wave
//! has returned (a barrier after the call, or the next pass's first one).
//! The reordered words stay in s_w, complete on return.
template <typename W, int MOVE, int BLOCK, int ITEMS>
__device__ __forceinline__ void lds_pass(W (&w)[ITEMS], uint32_t (&v)[ITEMS], uint32_t shift,
                                         uint32_t nb, uint32_t m, uint32_t rows,
                                         uint32_t* s_hist, W* s_w, uint32_t* s_v,
                                         uint32_t* s_tmp)
{
    constexpr int NW = BLOCK / WAVE;
    const uint32_t tid = threadIdx.x, lane = tid % WAVE, wave = tid / WAVE;
    const uint32_t wave_base = wave * (WAVE * rows);
    const uint32_t live = m > wave_base ? min(rows, (m - wave_base + WAVE - 1) / WAVE) : 0u;
    const uint32_t mask = (1u << nb) - 1u;
    uint32_t* my_hist = s_hist + wave * RADIX;
    uint32_t rank[ITEMS];
    uint32_t dg[(ITEMS + 3) / 4]; // digits, four to a register
#pragma unroll
    for (int r = 0; r < ITEMS; ++r)
        if (static_cast<uint32_t>(r) < live)
        {
            const bool valid = wave_base + r * WAVE + lane < m;
            const uint32_t d = valid ? static_cast<uint32_t>(w[r] >> shift) & mask : mask;
            dg[r / 4] = r % 4 == 0 ? d : dg[r / 4] | (d << (8 * (r % 4)));
            // LDS atomics only: the match's registers cost the 128 x 8 and
            // 256 x 8 shapes a wave per SIMD (the segmented graph 1.05 ->
            // 1.10 ms), and a segment's keys are relative to its minimum.
            rank[r] = rank_digit(my_hist, d, true, false);
        }
    __syncthreads();
    digit_bases<BLOCK>(s_hist, s_tmp);
#pragma unroll
    for (int r = 0; r < ITEMS; ++r)
        if (static_cast<uint32_t>(r) < live)
        {
            rank[r] += my_hist[(dg[r / 4] >> (8 * (r % 4))) & 0xFFu]; // the slot
            s_w[rank[r]] = w[r];
            if constexpr (MOVE == 1) s_v[rank[r]] = v[r];
        }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < ITEMS; ++r)
        if (static_cast<uint32_t>(r) < live)
        {
            w[r] = s_w[wave_base + r * WAVE + lane];
            if constexpr (MOVE == 1) v[r] = s_v[wave_base + r * WAVE + lane];
        }
    // Each wave clears its own counters: nothing reads them before the next
    // pass's first barrier, which also orders these read-backs before the
    // next reorder, so the pass needs no closing barrier.
#pragma unroll
    for (int q = 0; q < static_cast<int>(RADIX / WAVE); ++q) my_hist[q * WAVE + lane] = 0;
}

//! Stable LSD passes of lds_pass over bits [lo, lo + span) of the words:
//! ceil(span / nbmax) passes of nbmax (1..8) bits, the last one narrower.
template <typename W, int MOVE, int BLOCK, int ITEMS>
__device__ __forceinline__ void lds_passes(W (&w)[ITEMS], uint32_t (&v)[ITEMS], uint32_t lo,
                                           uint32_t span, uint32_t nbmax, uint32_t m,
                                           uint32_t rows, uint32_t* s_hist, W* s_w,
                                           uint32_t* s_v, uint32_t* s_tmp)
{
    const uint32_t passes = (span + nbmax - 1) / nbmax;
    uint32_t shift = lo;
#pragma unroll 1
    for (uint32_t q = 0; q < passes; ++q)
    {
        const uint32_t nb = min(nbmax, span - nbmax * q);
        lds_pass<W, MOVE, BLOCK, ITEMS>(w, v, shift, nb, m, rows, s_hist, s_w, s_v, s_tmp);
        shift += nb;
    }
}

//! LDS bytes per key of lds_sort_segment's staging buffer: 64-bit keys
//! with values stage packed words (values move once, outside LDS), 32-bit
//! keys stage values beside them.
constexpr std::size_t lds_segment_bytes_per_key(std::size_t kbytes, bool values)
{
    return kbytes == 8 ? 8 : 4 + (values ? 4 : 0);
}

//! Stable sort of m <= BLOCK * ITEMS keys (and values) by a single block,
//! entirely in LDS, over s_buf (BLOCK * ITEMS * lds_segment_bytes_per_key
//! bytes). The keys' order images are taken relative to their minimum:
//! only the low bits = ceil(log2(max - min + 1)) bits can differ, sorted in
//! ceil(bits / nbmax) passes of at most nbmax (1..8) bits (none for one
//! repeated key, which is left as it is). 64-bit keys
//! with values are sorted as packed words, (digits << 16) | original
//! position, 48 bits of the relative image at a time: the passes move 8
//! bytes a key and no value, and each value is fetched once at the end by
//! position (and each key too when the range passes 48 bits). Then one
//! contiguous write back. LDS (s_hist, s_buf, s_tmp) may be reused after
//! the caller's next barrier. A workgroup sorting segments one after
//! another passes first for its first one only and a barrier between
//! them: each call leaves the digit counters zero and the range words
//! reset for the next (the first call sets both up itself).
template <typename K, int KIND, bool DESC, bool VALUES, int BLOCK, int ITEMS>
__device__ __forceinline__ void lds_sort_segment(
    K* __restrict__ keys, uint32_t* __restrict__ vals, uint32_t m, uint32_t nbmax,
    uint32_t* s_hist, void* s_buf, uint32_t* s_tmp, bool first)
{
    constexpr bool WIDE = sizeof(K) == 8;
    constexpr uint32_t TILE = BLOCK * ITEMS;
    constexpr uint32_t NW = BLOCK / WAVE;
    static_assert(TILE <= 65536, "positions are packed in 16 bits");
    const uint32_t tid = threadIdx.x, lane = tid % WAVE, wave = tid / WAVE;
    // The segment is spread evenly over the waves, `rows` rows of 64 keys
    // each (the waves wait on each other at every pass: a 4097-key segment
    // in a 6144-key shape filled as 24 + 24 + 17 + 0 rows ran its passes at
    // the pace of 24 rows; spread, 17 + 17 + 17 + 15). Rows r >= rows of a
    // wave hold no positions.
    const uint32_t rows = min(static_cast<uint32_t>(ITEMS),
                              max(1u, (m + NW * WAVE - 1) / (NW * WAVE)));
    const uint32_t wave_base = wave * (WAVE * rows);

    using U = typename std::conditional<WIDE, unsigned long long, unsigned int>::type;
    __shared__ U s_min, s_max;
    // Resets the range words for the next segment; called once every
    // thread has read them and passed a barrier since.
    auto reset_range = [&] {
        if (tid == 0)
        {
            s_min = ~U(0);
            s_max = 0;
        }
    };
    if (first) reset_range();
    const uint32_t live = m > wave_base ? min(rows, (m - wave_base + WAVE - 1) / WAVE) : 0u;
    // Loads first, unconditionally (positions past m read key m - 1), so
    // they are all in flight together: no load waits on a lane or row test
    // (row tests here also cost the compiler registers).
    // 32-bit keys load their values with them (one round trip); 64-bit
    // keys fetch values once, by position, after sorting.
    constexpr bool EARLY_VALUES = false;
    K k[ITEMS];
    uint32_t v[ITEMS];
#pragma unroll
    for (int r = 0; r < ITEMS; ++r)
    {
        const uint32_t at = min(wave_base + r * WAVE + lane, m - 1);
        k[r] = keys[at];
        v[r] = 0;
        if constexpr (EARLY_VALUES) v[r] = vals[at];
    }
    U lo = ~U(0), hi = 0;
#pragma unroll
    for (int r = 0; r < ITEMS; ++r)
    {
        k[r] = order_bits<K, KIND, DESC>(k[r]);
        if (static_cast<uint32_t>(r) < rows && wave_base + r * WAVE + lane < m)
        {
            lo = min(lo, static_cast<U>(k[r]));
            hi = max(hi, static_cast<U>(k[r]));
        }
    }
#pragma unroll
    for (int o = WAVE / 2; o > 0; o >>= 1)
    {
        lo = min(lo, static_cast<U>(__shfl_xor(lo, o, WAVE)));
        hi = max(hi, static_cast<U>(__shfl_xor(hi, o, WAVE)));
    }
    if (first)
    {
        for (uint32_t i = tid; i < (BLOCK / WAVE) * RADIX; i += BLOCK) s_hist[i] = 0;
        __syncthreads(); // s_min / s_max initialised, digit counters zero
    }
    if (lane == 0 && wave_base < m)
    {
        atomicMin(&s_min, lo);
        atomicMax(&s_max, hi);
    }
    __syncthreads();
    const K kmin = static_cast<K>(s_min), range = static_cast<K>(s_max - s_min);
    if (range == 0)
    {
        __syncthreads(); // everyone has read s_min / s_max
        reset_range();
        return;
    }
    // The same in every lane (read from LDS, which the compiler cannot
    // know): made a scalar, so the pass loop is a plain loop rather than a
    // lane-masked one.
    const uint32_t bits = __builtin_amdgcn_readfirstlane(
        WIDE ? 64u - __clzll(static_cast<unsigned long long>(range))
             : 32u - __clz(static_cast<uint32_t>(range)));

    if constexpr (WIDE && VALUES)
    {
        // Words (digits << 16) | original position. The relative image's
        // low 48 bits are sorted first; a wider range then re-packs each
        // word with the image's high 16 bits, fetched by position from the
        // (not yet overwritten) input, and sorts those.
        uint64_t* s_w = static_cast<uint64_t*>(s_buf);
        uint32_t unused[ITEMS];
#pragma unroll
        for (int r = 0; r < ITEMS; ++r)
            k[r] = (static_cast<uint64_t>(k[r] - kmin) << 16) | (wave_base + r * WAVE + lane);
        const bool wide = bits > 48;
        lds_passes<uint64_t, 0, BLOCK, ITEMS>(k, unused, 16, min(bits, 48u), nbmax, m, rows,
                                              s_hist, s_w, nullptr, s_tmp);
        if (wide)
        {
            // Positions past m fetch position 0 (in bounds, never used).
            K hi16[ITEMS];
#pragma unroll
            for (int r = 0; r < ITEMS; ++r)
                if (static_cast<uint32_t>(r) < live)
                {
                    const uint32_t at = wave_base + r * WAVE + lane < m
                        ? static_cast<uint32_t>(k[r]) & 0xFFFFu : 0u;
                    hi16[r] = keys[at];
                }
#pragma unroll
            for (int r = 0; r < ITEMS; ++r)
                if (static_cast<uint32_t>(r) < live)
                {
                    const uint64_t rel = order_bits<K, KIND, DESC>(hi16[r]) - kmin;
                    k[r] = ((rel >> 48) << 16) | (k[r] & 0xFFFFu);
                }
            lds_passes<uint64_t, 0, BLOCK, ITEMS>(k, unused, 16, bits - 48, nbmax, m, rows,
                                                  s_hist, s_w, nullptr, s_tmp);
        }
        // Each value (and, for a wide range, each key) by its original
        // position; all are read before any is overwritten in place (the
        // wait and the barrier).
#pragma unroll
        for (int r = 0; r < ITEMS; ++r)
            if (static_cast<uint32_t>(r) < live)
            {
                const uint32_t at = wave_base + r * WAVE + lane < m
                    ? static_cast<uint32_t>(k[r]) & 0xFFFFu : 0u;
                v[r] = vals[at];
                if (wide) k[r] = keys[at];
                else k[r] = from_order_bits<K, KIND, DESC>(static_cast<K>(kmin + (k[r] >> 16)));
            }
        __builtin_amdgcn_s_waitcnt(0);
        __syncthreads();
#pragma unroll
        for (int r = 0; r < ITEMS; ++r)
        {
            const uint32_t pos = wave_base + r * WAVE + lane;
            if (static_cast<uint32_t>(r) < rows && pos < m)
            {
                keys[pos] = k[r];
                vals[pos] = v[r];
            }
        }
        reset_range();
        return; // no LDS use after the last pass's barrier

    }
    if constexpr (!WIDE && VALUES)
    {
        // 32-bit keys with values: 64-bit words (relative key << 16) |
        // position, as for 64-bit keys (the relative key has at most 32
        // bits): one 8-byte LDS write and read a key and pass instead of
        // two 4-byte ones for key and value, and the values, unneeded by
        // the passes, fetched once by position at the end.
        uint64_t* s_w = static_cast<uint64_t*>(s_buf);
        uint32_t unused[ITEMS];
        uint64_t w[ITEMS];
#pragma unroll
        for (int r = 0; r < ITEMS; ++r)
            w[r] = (static_cast<uint64_t>(k[r] - kmin) << 16) | (wave_base + r * WAVE + lane);
        lds_passes<uint64_t, 0, BLOCK, ITEMS>(w, unused, 16, bits, nbmax, m, rows, s_hist, s_w,
                                              nullptr, s_tmp);
#pragma unroll
        for (int r = 0; r < ITEMS; ++r)
            if (static_cast<uint32_t>(r) < live)
            {
                const uint32_t at = wave_base + r * WAVE + lane < m
                    ? static_cast<uint32_t>(w[r]) & 0xFFFFu : 0u;
                v[r] = vals[at];
                k[r] = from_order_bits<K, KIND, DESC>(static_cast<K>(kmin + (w[r] >> 16)));
            }
        __builtin_amdgcn_s_waitcnt(0);
        __syncthreads(); // every value read before any is overwritten in place
#pragma unroll
        for (int r = 0; r < ITEMS; ++r)
        {
            const uint32_t pos = wave_base + r * WAVE + lane;
            if (static_cast<uint32_t>(r) < rows && pos < m)
            {
                keys[pos] = k[r];
                vals[pos] = v[r];
            }
        }
        reset_range();
        return;
    }
    constexpr int MOVE = 0; // 32-bit keys without values
    K* s_w = static_cast<K*>(s_buf);
    uint32_t* s_v = reinterpret_cast<uint32_t*>(s_w + TILE);
#pragma unroll
    for (int r = 0; r < ITEMS; ++r) k[r] -= kmin;
    lds_passes<K, MOVE, BLOCK, ITEMS>(k, v, 0, bits, nbmax, m, rows, s_hist, s_w, s_v, s_tmp);
#pragma unroll
    for (int r = 0; r < ITEMS; ++r)
    {
        const uint32_t pos = wave_base + r * WAVE + lane;
        if (static_cast<uint32_t>(r) < rows && pos < m)
        {
            keys[pos] = from_order_bits<K, KIND, DESC>(static_cast<K>(k[r] + kmin));
            if constexpr (VALUES) vals[pos] = v[r];
        }
    }
    reset_range(); // the passes' barriers follow every read of the range
}

} // namespace radix
} // namespace clc
