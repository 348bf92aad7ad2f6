// Host side of the segmented sort: size-class tables (default and
// autotuned, with an on-disk cache), the entry point, and the hipGraph
// executor that captures the whole launch sequence once and replays it.
#include "segmented_sort.hpp"

#include "ops.hpp"

#include <chrono>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <memory>
#include <sstream>
#include <unistd.h>

namespace clc {
namespace seg {

table default_table(std::size_t kbytes, bool values, std::size_t lds_limit)
{
    table t;
    for (int b = 0; b < NBUCKET; ++b)
    {
        t.cls[b] = LARGE;
        t.digit_bits[b] = MAX_DIGIT_BITS;
        // Smallest LDS configuration that holds the whole bucket.
        for (int c = 0; c < NCFG; ++c)
        {
            if ((std::size_t(1) << b) > cfg_capacity(c)) continue;
            if (cfg_lds(c, kbytes, values) > lds_limit) continue;
            t.cls[b] = c;
            break;
        }
    }
    return t;
}

namespace {

using sort_fn = std::size_t (*)(context&, bool, void*, bool, void*, std::size_t,
                                const uint64_t*, std::size_t, bool, const table&, bool);

sort_fn pick(int key_type)
{
    switch (key_type)
    {
        case TY_U32: return &sort_typed<uint32_t, dev::KEY_UNSIGNED>;
        case TY_I32: return &sort_typed<uint32_t, dev::KEY_SIGNED>;
        case TY_F32: return &sort_typed<uint32_t, dev::KEY_FLOAT>;
        case TY_U64: return &sort_typed<uint64_t, dev::KEY_UNSIGNED>;
        case TY_I64: return &sort_typed<uint64_t, dev::KEY_SIGNED>;
        case TY_F64: return &sort_typed<uint64_t, dev::KEY_FLOAT>;
    }
    throw std::invalid_argument(
        "segmented sort: keys must be I32, U32, F32, I64, U64 or F64");
}

std::size_t key_bytes(int key_type)
{
    return (key_type == TY_U64 || key_type == TY_I64 || key_type == TY_F64) ? 8 : 4;
}

std::size_t lds_limit(const context& ctx)
{
    const std::string& a = ctx.facts().arch;
    return a.compare(0, 5, "gfx95") == 0 ? 160 * 1024 : 64 * 1024;
}

int tune_slot(int key_type, bool values) { return 1000 + key_type * 2 + (values ? 1 : 0); }

// Tuned tables persist through the context's tuning cache (load_tuning /
// store_tuning), one entry per key type and values flag: the 32 classes,
// then the 32 digit widths (an entry of the classes alone takes 8-bit
// digits). An entry that names a class infeasible for this key type, or a
// width outside 1..8, is treated as absent.

std::vector<int> table_values(const table& t)
{
    std::vector<int> v(t.cls, t.cls + NBUCKET);
    v.insert(v.end(), t.digit_bits, t.digit_bits + NBUCKET);
    return v;
}

//! The table a value list holds (classes, then optionally digit widths),
//! unchecked; false when the length fits neither form.
bool table_from(const std::vector<int>& v, table& t)
{
    if (v.size() != std::size_t(NBUCKET) && v.size() != std::size_t(2 * NBUCKET)) return false;
    for (int b = 0; b < NBUCKET; ++b)
    {
        t.cls[b] = v[b];
        t.digit_bits[b] = v.size() == std::size_t(2 * NBUCKET) ? v[NBUCKET + b] : MAX_DIGIT_BITS;
    }
    return true;
}

std::string table_name(int key_type, bool values)
{
    return "segsort_t" + std::to_string(key_type) + (values ? "_kv" : "_k");
}

bool load_table(const context& ctx, int key_type, bool values, table& t)
{
    std::vector<int> v;
    table r;
    if (!load_tuning(ctx, table_name(key_type, values), v) || !table_from(v, r)) return false;
    for (int b = 0; b < NBUCKET; ++b)
    {
        if (r.cls[b] < 0 || r.cls[b] > LARGE) return false;
        if (r.digit_bits[b] < 1 || r.digit_bits[b] > MAX_DIGIT_BITS) return false;
        if (r.cls[b] < NCFG
            && (cfg_lds(r.cls[b], key_bytes(key_type), values) > lds_limit(ctx)
                || (b > 0 && (std::size_t(1) << b) > cfg_capacity(r.cls[b]))))
            return false;
    }
    t = r;
    return true;
}

void store_table(const context& ctx, int key_type, bool values, const table& t)
{
    store_tuning(ctx, table_name(key_type, values), table_values(t));
}

table table_for(context& ctx, int key_type, bool values)
{
    auto& tuned = ctx.tuned();
    const int slot = tune_slot(key_type, values);
    auto it = tuned.find(slot);
    table t;
    if (it != tuned.end() && table_from(it->second, t)) return t;
    if (!load_table(ctx, key_type, values, t))
        t = default_table(key_bytes(key_type), values, lds_limit(ctx));
    tuned[slot] = table_values(t);
    return t;
}

// ---- autotuner ------------------------------------------------------------------

__global__ void tune_fill_kernel(uint32_t* words, std::size_t n, uint64_t seed)
{
    for (std::size_t i = size_t(blockIdx.x) * 256 + threadIdx.x; i < n;
         i += size_t(gridDim.x) * 256)
    {
        // splitmix64 finaliser of (seed, index)
        uint64_t z = seed + (i + 1) * 0x9E3779B97F4A7C15ull;
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        words[i] = static_cast<uint32_t>(z ^ (z >> 31));
    }
}

//! Times every feasible class for each size bucket on synthetic segments of
//! that bucket and keeps the fastest; for an LDS class then the widest
//! digit of its passes, from 8 bits down to 4 (a narrower digit means more
//! passes over fewer distinct counters).
table run_tuner(context& ctx, int key_type, bool values)
{
    const std::size_t kb = key_bytes(key_type);
    const sort_fn fn = pick(key_type);
    const std::size_t limit = lds_limit(ctx);
    table best = default_table(kb, values, limit);
    constexpr std::size_t TOTAL = std::size_t(1) << 22;
    constexpr int MAX_BUCKET = 22;
    temp_buffer keys(ctx.resource(), TOTAL * kb, ctx.stream());
    temp_buffer vals(ctx.resource(), values ? TOTAL * 4 : 0, ctx.stream());
    temp_buffer offs(ctx.resource(), (TOTAL + 1) * 8, ctx.stream());
    std::vector<uint64_t> host_off;
    hipEvent_t e0, e1;
    CLC_HIP(hipEventCreate(&e0));
    CLC_HIP(hipEventCreate(&e1));
    struct events
    {
        hipEvent_t a, b;
        ~events()
        {
            (void)hipEventDestroy(a);
            (void)hipEventDestroy(b);
        }
    } guard{ e0, e1 };
    hipStream_t s = ctx.stream();
    const bool profiling = ctx.profiler().enabled();
    ctx.profiler().enable(false);
    for (int b = 1; b <= MAX_BUCKET; ++b)
    {
        const std::size_t m = b == 1 ? 2 : (std::size_t(3) << (b - 2)); // in (2^(b-1), 2^b]
        const std::size_t nseg = std::max<std::size_t>(1, TOTAL / m);
        const std::size_t n = nseg * m;
        host_off.resize(nseg + 1);
        for (std::size_t i = 0; i <= nseg; ++i) host_off[i] = i * m;
        CLC_HIP(hipMemcpyAsync(offs.get(), host_off.data(), (nseg + 1) * 8,
                               hipMemcpyHostToDevice, s));
        // Best of three timed runs after one warm-up, on fresh keys each.
        auto time_table = [&](const table& t) {
            float run_best = 1e30f;
            for (int rep = 0; rep < 4; ++rep)
            {
                hipLaunchKernelGGL(tune_fill_kernel, dim3(1024), dim3(256), 0, s,
                                   keys.as<uint32_t>(), n * kb / 4,
                                   uint64_t(b * 131 + rep));
                CLC_HIP(hipEventRecord(e0, s));
                fn(ctx, false, keys.get(), values, vals.get(), n,
                   offs.as<const uint64_t>(), nseg, false, t, false);
                CLC_HIP(hipEventRecord(e1, s));
                CLC_HIP(hipEventSynchronize(e1));
                float ms = 0;
                CLC_HIP(hipEventElapsedTime(&ms, e0, e1));
                if (rep > 0) run_best = std::min(run_best, ms);
            }
            return run_best;
        };
        std::vector<int> cands;
        for (int c = 0; c < NCFG; ++c)
            if (cfg_capacity(c) >= (std::size_t(1) << b) && cfg_lds(c, kb, values) <= limit)
                cands.push_back(c);
        if (b >= 12) cands.push_back(LARGE);
        float best_ms = 1e30f;
        for (int c : cands)
        {
            table t = best;
            t.cls[b] = c;
            t.digit_bits[b] = MAX_DIGIT_BITS;
            const float ms = time_table(t);
            if (ms < best_ms)
            {
                best_ms = ms;
                best.cls[b] = c;
                best.digit_bits[b] = MAX_DIGIT_BITS;
            }
        }
        if (best.cls[b] < NCFG)
            for (int nb = MAX_DIGIT_BITS - 1; nb >= 4; --nb)
            {
                table t = best;
                t.digit_bits[b] = nb;
                const float ms = time_table(t);
                if (ms < best_ms)
                {
                    best_ms = ms;
                    best.digit_bits[b] = nb;
                }
            }
    }
    for (int b = MAX_BUCKET + 1; b < NBUCKET; ++b) best.cls[b] = LARGE;
    ctx.profiler().enable(profiling);
    CLC_HIP(hipStreamSynchronize(s));
    return best;
}

//! Scratch source over one block the graph holds: every launch recorded
//! during the capture takes its scratch from it, and nothing is recycled
//! while the graph can be replayed.
class fixed_resource final : public scratch_source {
public:
    fixed_resource(void* p, std::size_t bytes) : p_(p), bytes_(bytes) {}

    void* take(std::size_t bytes, stream_ref) override
    {
        if (used_ || bytes > bytes_)
            throw std::logic_error("segmented sort graph: scratch already taken");
        used_ = true;
        return p_;
    }
    void give(void*, std::size_t, stream_ref) noexcept override {}

private:
    void* p_;
    std::size_t bytes_;
    bool used_ = false;
};

} // namespace

// ---- graph executor ---------------------------------------------------------------

graph::~graph()
{
    if (exec) (void)hipGraphExecDestroy(exec);
    if (g) (void)hipGraphDestroy(g);
    if (capture_stream) (void)hipStreamDestroy(capture_stream);
}

void graph::launch()
{
    ctx->activate();
    CLC_HIP(hipGraphLaunch(exec, ctx->stream()));
}

} // namespace seg

namespace detail {

namespace {
void check_args(int key_type, const void* keys, const uint64_t* offsets,
                std::size_t num_items, std::size_t num_segments)
{
    seg::pick(key_type);
    if (num_segments > 0xFFFFFFFEu)
        throw std::invalid_argument("segmented sort: at most 2^32-2 segments");
    if (num_segments > 0 && (keys == nullptr || offsets == nullptr))
        throw std::invalid_argument("segmented sort: null keys or offsets");
    (void)num_items;
}
} // namespace

void segmented_sort(context& ctx, void* keys, int key_type, void* values,
                    std::size_t num_items, const uint64_t* offsets,
                    std::size_t num_segments, bool descending, bool overlap_large)
{
    check_args(key_type, keys, offsets, num_items, num_segments);
    if (num_segments == 0 || num_items < 2) return;
    ctx.activate();
    const seg::table t = seg::table_for(ctx, key_type, values != nullptr);
    seg::pick(key_type)(ctx, false, keys, values != nullptr, values, num_items, offsets,
                        num_segments, descending, t, overlap_large);
}

void segmented_sort_tune(context& ctx, int key_type, bool with_values)
{
    seg::pick(key_type);
    ctx.activate();
    const seg::table t = seg::run_tuner(ctx, key_type, with_values);
    ctx.tuned()[seg::tune_slot(key_type, with_values)] = seg::table_values(t);
    seg::store_table(ctx, key_type, with_values, t);
}

std::vector<int> segmented_sort_table(context& ctx, int key_type, bool with_values)
{
    seg::pick(key_type);
    ctx.activate();
    const seg::table t = seg::table_for(ctx, key_type, with_values);
    return std::vector<int>(t.cls, t.cls + seg::NBUCKET);
}

std::vector<int> segmented_sort_digit_bits(context& ctx, int key_type, bool with_values)
{
    seg::pick(key_type);
    ctx.activate();
    const seg::table t = seg::table_for(ctx, key_type, with_values);
    return std::vector<int>(t.digit_bits, t.digit_bits + seg::NBUCKET);
}

std::unique_ptr<seg::graph> segmented_sort_graph(context& ctx, void* keys, int key_type,
                                                 void* values, std::size_t num_items,
                                                 const uint64_t* offsets,
                                                 std::size_t num_segments, bool descending)
{
    check_args(key_type, keys, offsets, num_items, num_segments);
    ctx.activate();
    const seg::table t = seg::table_for(ctx, key_type, values != nullptr);
    const seg::sort_fn fn = seg::pick(key_type);
    std::unique_ptr<seg::graph> g(new seg::graph);
    g->ctx = &ctx;
    const std::size_t bytes = fn(ctx, true, keys, values != nullptr, values, num_items,
                                 offsets, num_segments, descending, t, true);
    g->scratch.reset(new temp_buffer(ctx.resource(), bytes, ctx.stream()));
    CLC_HIP(hipStreamCreateWithFlags(&g->capture_stream, hipStreamNonBlocking));
    // Launch everything on a private stream under capture, drawing scratch
    // from the graph's own block and with event profiling off.
    const hipStream_t old_stream = ctx.swap_stream(g->capture_stream);
    auto old_res =
        ctx.swap_resource(std::make_shared<seg::fixed_resource>(g->scratch->get(), bytes));
    const bool profiling = ctx.profiler().enabled();
    ctx.profiler().enable(false);
    struct restore
    {
        context& c;
        hipStream_t s;
        std::shared_ptr<scratch_source> r;
        bool p;
        ~restore()
        {
            c.swap_stream(s);
            c.swap_resource(r);
            c.profiler().enable(p);
        }
    } guard{ ctx, old_stream, old_res, profiling };
    CLC_HIP(hipStreamBeginCapture(g->capture_stream, hipStreamCaptureModeThreadLocal));
    try
    {
        if (num_segments > 0 && num_items >= 2)
            fn(ctx, false, keys, values != nullptr, values, num_items, offsets,
               num_segments, descending, t, true);
    } catch (...)
    {
        hipGraph_t broken = nullptr;
        (void)hipStreamEndCapture(g->capture_stream, &broken);
        if (broken) (void)hipGraphDestroy(broken);
        throw;
    }
    CLC_HIP(hipStreamEndCapture(g->capture_stream, &g->g));
    CLC_HIP(hipGraphInstantiate(&g->exec, g->g, nullptr, nullptr, 0));
    return g;
}

} // namespace detail
} // namespace clc
