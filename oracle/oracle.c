/*
 * CPU oracle for the sort / scan / reduce / histogram path.
 *
 * TEST INFRASTRUCTURE ONLY. Nothing in the shipped library links, loads or
 * calls this file; only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg use it, as the checker (and, in bench.py, as the timed CPU
 * baseline). It restates the reference algorithms in plain C:
 *
 *   pcg32                lib/src/SDK/Random.c:13-22, :43-50
 *   bounded ints         lib/src/SDK/Random.c:73-102
 *   radix sort           lib/src/Compute/Algorithm.cpp:129-213 (kernels),
 *                        :453-527 (host driver): stable LSD, 4-bit digits,
 *                        per-group digit histogram (digit-major), exclusive
 *                        scan of the histogram, in-order scatter, 8 passes
 *   key transform        lib/src/Compute/Algorithm.cpp:459-469
 *   scan                 lib/src/Compute/Algorithm.cpp:76-127, :280-318
 *   reduce               lib/src/Compute/Algorithm.cpp:58-74, :400-442
 *   histogram_even       lib/src/Compute/Algorithm.cpp:240-273, :569-615
 *   operators            lib/include/CL/Compute/Algorithm.hpp:53-65
 *
 * 64-bit keys are an extension (the reference rejects them,
 * Algorithm.hpp:206,217; CApi.cpp:522-524): the same LSD algorithm runs 16
 * four-bit passes over the 64-bit transformed key.
 *
 * How each function is pinned is stated in the test files that check it
 * (tests/test_oracle.py): radix sort against numpy's stable argsort and the
 * reference tests' own std::stable_sort results, pcg32 against the published
 * demo sequence, scan/reduce against closed forms, histogram against the
 * reference test's integer bin formula.
 */
#include <pthread.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

/* ---- pcg32 (Random.c:13-22, :43-50) ------------------------------------ */

typedef struct { uint64_t state, inc; } oracle_pcg32;

static uint32_t pcg32_next(oracle_pcg32* r)
{
    uint64_t old = r->state;
    r->state = old * 6364136223846793005ULL + (r->inc | 1u);
    uint32_t xs = (uint32_t)(((old >> 18u) ^ old) >> 27u);
    uint32_t rot = (uint32_t)(old >> 59u);
    return (xs >> rot) | (xs << ((0u - rot) & 31u));
}

static void pcg32_seed(oracle_pcg32* r, uint64_t initstate, uint64_t initseq)
{
    r->state = 0u;
    r->inc = (initseq << 1u) | 1u;
    pcg32_next(r);
    r->state += initstate;
    pcg32_next(r);
}

/* Jump the generator ahead by delta steps in O(log delta): the published
 * pcg32_advance (O'Neill, "PCG: A Family of Simple Fast Space-Efficient
 * Statistically Good Algorithms", section 4.3.1; Brown, "Random Number
 * Generation with Arbitrary Strides"). Lets threads fill disjoint chunks of
 * one stream. */
static void pcg32_advance(oracle_pcg32* r, uint64_t delta)
{
    uint64_t cur_mult = 6364136223846793005ULL, cur_plus = r->inc | 1u;
    uint64_t acc_mult = 1u, acc_plus = 0u;
    while (delta > 0)
    {
        if (delta & 1u)
        {
            acc_mult *= cur_mult;
            acc_plus = acc_plus * cur_mult + cur_plus;
        }
        cur_plus = (cur_mult + 1u) * cur_plus;
        cur_mult *= cur_mult;
        delta >>= 1u;
    }
    r->state = acc_mult * r->state + acc_plus;
}

void oracle_pcg32_u32(uint64_t initstate, uint64_t initseq, uint64_t skip,
                      uint32_t* out, size_t n)
{
    oracle_pcg32 r;
    pcg32_seed(&r, initstate, initseq);
    pcg32_advance(&r, skip);
    for (size_t i = 0; i < n; ++i) out[i] = pcg32_next(&r);
}

/* Bounded integers in [low, hi] by bit-slice rejection (Random.c:73-102). */
void oracle_ints_range(uint64_t initstate, uint64_t initseq, int32_t* out,
                       size_t n, int32_t low, int32_t hi)
{
    oracle_pcg32 r;
    pcg32_seed(&r, initstate, initseq);
    const uint32_t diff = (uint32_t)hi - (uint32_t)low;
    uint32_t bits = 1;
    if (diff)
    {
        bits = 0;
        for (uint32_t v = diff; v; v >>= 1) ++bits; /* floor(log2)+1 */
    }
    const uint32_t mask = diff ? (bits >= 32 ? 0xFFFFFFFFu : (1u << bits) - 1u)
                               : 0u;
    for (size_t i = 0; i < n; ++i)
    {
        uint32_t res = 0;
        int taken = 0;
        while (!taken)
        {
            res = pcg32_next(&r);
            for (uint32_t k = 0; k < 32u / bits; ++k)
            {
                if ((res & mask) <= diff)
                {
                    res &= mask;
                    taken = 1;
                    break;
                }
                res >>= bits;
            }
        }
        out[i] = (int32_t)((uint32_t)low + res);
    }
}

/* ---- key transforms (Algorithm.cpp:459-469) ---------------------------- */

/* Type codes follow clc_type (compute_c.h:44-56). */
enum { T_I8, T_U8, T_I16, T_U16, T_I32, T_U32, T_I64, T_U64, T_F32, T_F64 };

static uint32_t xform32(uint32_t k, int type, int desc)
{
    uint32_t t = k;
    if (type == T_I32) t = k ^ 0x80000000u;
    else if (type == T_F32) t = (k & 0x80000000u) ? ~k : (k | 0x80000000u);
    return desc ? ~t : t;
}

static uint64_t xform64(uint64_t k, int type, int desc)
{
    const uint64_t top = 0x8000000000000000ULL;
    uint64_t t = k;
    if (type == T_I64) t = k ^ top;
    else if (type == T_F64) t = (k & top) ? ~k : (k | top);
    return desc ? ~t : t;
}

/* ---- radix sort (Algorithm.cpp:153-212, :453-527) ----------------------- */

#define RADIX 16u
#define MAX_THREADS 256

typedef struct
{
    int nthreads;
    int wide; /* 64-bit keys */
    int type, desc;
    size_t n;
    void* key_buf[2];
    uint32_t* val_buf[2];
    unsigned passes;
    size_t* counts; /* [RADIX][nthreads], digit-major as clc_radix_histogram */
    pthread_barrier_t barrier;
} sort_job;

typedef struct
{
    sort_job* job;
    int id;
} sort_arg;

static unsigned digit_of(const sort_job* j, const void* keys, size_t i,
                         unsigned shift)
{
    if (j->wide)
        return (unsigned)((xform64(((const uint64_t*)keys)[i], j->type,
                                   j->desc) >> shift) & (RADIX - 1u));
    return (xform32(((const uint32_t*)keys)[i], j->type, j->desc) >> shift)
        & (RADIX - 1u);
}

static void* sort_worker(void* p)
{
    sort_arg* a = (sort_arg*)p;
    sort_job* j = a->job;
    const int t = a->id, T = j->nthreads;
    const size_t chunk = (j->n + (size_t)T - 1) / (size_t)T;
    const size_t lo = (size_t)t * chunk < j->n ? (size_t)t * chunk : j->n;
    const size_t hi = lo + chunk < j->n ? lo + chunk : j->n;
    for (unsigned pass = 0; pass < j->passes; ++pass)
    {
        const unsigned shift = pass * 4u;
        const int src = (int)(pass % 2u), dst = 1 - src;
        const void* kin = j->key_buf[src];
        /* Phase 1: this group's digit counts (clc_radix_histogram). */
        size_t local[RADIX] = { 0 };
        for (size_t i = lo; i < hi; ++i) ++local[digit_of(j, kin, i, shift)];
        for (unsigned d = 0; d < RADIX; ++d)
            j->counts[(size_t)d * (size_t)T + (size_t)t] = local[d];
        pthread_barrier_wait(&j->barrier);
        /* Phase 2: exclusive scan over the digit-major counts
         * (scan_impl(counts -> offsets), Algorithm.cpp:510-511). */
        if (t == 0)
        {
            size_t run = 0;
            for (size_t i = 0; i < (size_t)RADIX * (size_t)T; ++i)
            {
                size_t c = j->counts[i];
                j->counts[i] = run;
                run += c;
            }
        }
        pthread_barrier_wait(&j->barrier);
        /* Phase 3: in-order scatter (clc_radix_scatter). */
        size_t base[RADIX];
        for (unsigned d = 0; d < RADIX; ++d)
            base[d] = j->counts[(size_t)d * (size_t)T + (size_t)t];
        for (size_t i = lo; i < hi; ++i)
        {
            const size_t o = base[digit_of(j, kin, i, shift)]++;
            if (j->wide)
                ((uint64_t*)j->key_buf[dst])[o] = ((const uint64_t*)kin)[i];
            else
                ((uint32_t*)j->key_buf[dst])[o] = ((const uint32_t*)kin)[i];
            if (j->val_buf[0]) j->val_buf[dst][o] = j->val_buf[src][i];
        }
        pthread_barrier_wait(&j->barrier);
    }
    return NULL;
}

/* Stable in-place sort of n keys (4- or 8-byte, by clc_type code) with
 * optional 32-bit values. Returns 0, or -1 for an unsupported key type, -2
 * when out of host memory. threads <= 0 means one thread. */
int oracle_radix_sort(int key_type, void* keys, uint32_t* values, size_t n,
                      int descending, int threads)
{
    int wide;
    if (key_type == T_U32 || key_type == T_I32 || key_type == T_F32) wide = 0;
    else if (key_type == T_U64 || key_type == T_I64 || key_type == T_F64)
        wide = 1;
    else
        return -1;
    if (n <= 1) return 0;
    if (threads <= 0) threads = 1;
    if (threads > MAX_THREADS) threads = MAX_THREADS;
    if ((size_t)threads > n) threads = (int)n;

    sort_job j;
    memset(&j, 0, sizeof j);
    j.nthreads = threads;
    j.wide = wide;
    j.type = key_type;
    j.desc = descending != 0;
    j.n = n;
    j.passes = wide ? 16u : 8u;
    const size_t kbytes = n * (wide ? 8u : 4u);
    void* alt = malloc(kbytes);
    uint32_t* valt = values ? (uint32_t*)malloc(n * 4u) : NULL;
    j.counts = (size_t*)malloc(sizeof(size_t) * RADIX * (size_t)threads);
    if (!alt || (values && !valt) || !j.counts)
    {
        free(alt);
        free(valt);
        free(j.counts);
        return -2;
    }
    j.key_buf[0] = keys;
    j.key_buf[1] = alt;
    j.val_buf[0] = values;
    j.val_buf[1] = valt;
    pthread_barrier_init(&j.barrier, NULL, (unsigned)threads);
    pthread_t tid[MAX_THREADS];
    sort_arg args[MAX_THREADS];
    for (int t = 0; t < threads; ++t)
    {
        args[t].job = &j;
        args[t].id = t;
        if (t > 0) pthread_create(&tid[t], NULL, sort_worker, &args[t]);
    }
    sort_worker(&args[0]);
    for (int t = 1; t < threads; ++t) pthread_join(tid[t], NULL);
    pthread_barrier_destroy(&j.barrier);
    /* An even pass count leaves the result in the caller's buffers
     * (Algorithm.cpp:526). */
    free(alt);
    free(valt);
    free(j.counts);
    return 0;
}

/* ---- scan and reduce (Algorithm.cpp:58-127, :280-318, :400-451) -------- */

/* Operator codes: the canonical ops:: of Algorithm.hpp:53-65 plus the two
 * projections the reference tests use ("b" in main.cpp:231-238). */
enum { OP_PLUS, OP_MUL, OP_AND, OP_OR, OP_XOR, OP_MIN, OP_MAX, OP_RIGHT,
       OP_LEFT };

#define APPLY(T, op, a, b, out, bad)                                          \
    switch (op)                                                               \
    {                                                                         \
        case OP_PLUS: out = (T)((a) + (b)); break;                            \
        case OP_MUL: out = (T)((a) * (b)); break;                             \
        case OP_MIN: out = ((b) < (a)) ? (b) : (a); break;                    \
        case OP_MAX: out = ((a) < (b)) ? (b) : (a); break;                    \
        case OP_RIGHT: out = (b); break;                                      \
        case OP_LEFT: out = (a); break;                                       \
        default: bad = 1; out = (a); break;                                   \
    }

#define APPLY_INT(T, op, a, b, out, bad)                                      \
    switch (op)                                                               \
    {                                                                         \
        case OP_AND: out = (T)((a) & (b)); break;                             \
        case OP_OR: out = (T)((a) | (b)); break;                              \
        case OP_XOR: out = (T)((a) ^ (b)); break;                             \
        default: APPLY(T, op, a, b, out, bad)                                 \
    }

/* Unsigned arithmetic for the integer plus/multiply so that wrap-around is
 * defined, as in OpenCL C. */
#define SCAN_INT(T, UT)                                                       \
    static int scan_##T(int op, T id, const T* in, T* out, size_t n, int ex)  \
    {                                                                         \
        int bad = 0;                                                          \
        T run = id;                                                           \
        for (size_t i = 0; i < n; ++i)                                        \
        {                                                                     \
            const T x = in[i];                                                \
            T next;                                                           \
            if (op == OP_PLUS) next = (T)((UT)run + (UT)x);                   \
            else if (op == OP_MUL) next = (T)((UT)run * (UT)x);               \
            else { APPLY_INT(T, op, run, x, next, bad) }                      \
            out[i] = ex ? run : next;                                         \
            run = next;                                                       \
        }                                                                     \
        return bad ? -1 : 0;                                                  \
    }                                                                         \
    static int fold_##T(int op, T id, const T* in, size_t n, T* r)            \
    {                                                                         \
        int bad = 0;                                                          \
        T acc = id;                                                           \
        for (size_t i = 0; i < n; ++i)                                        \
        {                                                                     \
            T next;                                                           \
            if (op == OP_PLUS) next = (T)((UT)acc + (UT)in[i]);               \
            else if (op == OP_MUL) next = (T)((UT)acc * (UT)in[i]);           \
            else { APPLY_INT(T, op, acc, in[i], next, bad) }                  \
            acc = next;                                                       \
        }                                                                     \
        *r = acc;                                                             \
        return bad ? -1 : 0;                                                  \
    }

#define SCAN_FP(T)                                                            \
    static int scan_##T(int op, T id, const T* in, T* out, size_t n, int ex)  \
    {                                                                         \
        int bad = 0;                                                          \
        T run = id;                                                           \
        for (size_t i = 0; i < n; ++i)                                        \
        {                                                                     \
            const T x = in[i];                                                \
            T next;                                                           \
            APPLY(T, op, run, x, next, bad)                                   \
            out[i] = ex ? run : next;                                         \
            run = next;                                                       \
        }                                                                     \
        return bad ? -1 : 0;                                                  \
    }                                                                         \
    static int fold_##T(int op, T id, const T* in, size_t n, T* r)            \
    {                                                                         \
        int bad = 0;                                                          \
        T acc = id;                                                           \
        for (size_t i = 0; i < n; ++i)                                        \
        {                                                                     \
            T next;                                                           \
            APPLY(T, op, acc, in[i], next, bad)                               \
            acc = next;                                                       \
        }                                                                     \
        *r = acc;                                                             \
        return bad ? -1 : 0;                                                  \
    }

SCAN_INT(int8_t, uint8_t)
SCAN_INT(uint8_t, uint8_t)
SCAN_INT(int16_t, uint16_t)
SCAN_INT(uint16_t, uint16_t)
SCAN_INT(int32_t, uint32_t)
SCAN_INT(uint32_t, uint32_t)
SCAN_INT(int64_t, uint64_t)
SCAN_INT(uint64_t, uint64_t)
typedef float float_t_;
typedef double double_t_;
SCAN_FP(float_t_)
SCAN_FP(double_t_)

/* out[i] = in[0] op ... op in[i] (inclusive) or identity op ... op in[i-1]
 * (exclusive), combined strictly left to right. Returns -1 for an operator
 * the type does not have. */
int oracle_scan(int type, int op, const void* identity, const void* in,
                void* out, size_t n, int exclusive)
{
    switch (type)
    {
#define CASE(code, T)                                                         \
    case code:                                                                \
        return scan_##T(op, *(const T*)identity, (const T*)in, (T*)out, n,    \
                        exclusive);
        CASE(T_I8, int8_t)
        CASE(T_U8, uint8_t)
        CASE(T_I16, int16_t)
        CASE(T_U16, uint16_t)
        CASE(T_I32, int32_t)
        CASE(T_U32, uint32_t)
        CASE(T_I64, int64_t)
        CASE(T_U64, uint64_t)
        CASE(T_F32, float_t_)
        CASE(T_F64, double_t_)
#undef CASE
    }
    return -1;
}

/* result = op(init, identity op in[0] op ... op in[n-1]); n == 0 gives init
 * (Algorithm.cpp:406-410, :66-72). */
int oracle_reduce(int type, int op, const void* identity, const void* init,
                  const void* in, size_t n, void* result)
{
    switch (type)
    {
#define CASE(code, T)                                                         \
    case code:                                                                \
    {                                                                         \
        T r, fin;                                                             \
        int bad = 0;                                                          \
        if (n == 0)                                                           \
        {                                                                     \
            *(T*)result = *(const T*)init;                                    \
            return 0;                                                         \
        }                                                                     \
        if (fold_##T(op, *(const T*)identity, (const T*)in, n, &r)) return -1; \
        T two[2] = { *(const T*)init, r };                                    \
        if (fold_##T(op, two[0], two + 1, 1, &fin)) bad = 1;                  \
        *(T*)result = fin;                                                    \
        return bad ? -1 : 0;                                                  \
    }
        CASE(T_I8, int8_t)
        CASE(T_U8, uint8_t)
        CASE(T_I16, int16_t)
        CASE(T_U16, uint16_t)
        CASE(T_I32, int32_t)
        CASE(T_U32, uint32_t)
        CASE(T_I64, int64_t)
        CASE(T_U64, uint64_t)
        CASE(T_F32, float_t_)
        CASE(T_F64, double_t_)
#undef CASE
    }
    return -1;
}

/* ---- histogram_even (Algorithm.cpp:240-273, :569-615) ------------------ */

/* Zeroes bins counters, then counts samples with lower <= x < upper into
 * bin (uint)((x - lower) * scale), scale = (ACC)bins / (upper - lower),
 * clamped to bins - 1. ACC is float unless the samples are double, and
 * lower/upper are first rounded to ACC (Algorithm.cpp:601-610). Returns -1
 * when !(lower < upper) (invalid_argument in the reference). */
int oracle_histogram_even(int type, const void* in, size_t n, uint32_t* hist,
                          size_t bins, double lower, double upper)
{
    if (bins == 0) return 0;
    if (!(lower < upper)) return -1;
    memset(hist, 0, bins * sizeof(uint32_t));
    if (type == T_F64)
    {
        const double lo = lower, up = upper;
        const double scale = (double)bins / (up - lo);
        for (size_t i = 0; i < n; ++i)
        {
            const double x = ((const double*)in)[i];
            if (x >= lo && x < up)
            {
                uint32_t b = (uint32_t)((x - lo) * scale);
                if (b >= bins) b = (uint32_t)bins - 1u;
                ++hist[b];
            }
        }
        return 0;
    }
    const float lo = (float)lower, up = (float)upper;
    const float scale = (float)bins / (up - lo);
    for (size_t i = 0; i < n; ++i)
    {
        float x;
        switch (type)
        {
            case T_I8: x = (float)((const int8_t*)in)[i]; break;
            case T_U8: x = (float)((const uint8_t*)in)[i]; break;
            case T_I16: x = (float)((const int16_t*)in)[i]; break;
            case T_U16: x = (float)((const uint16_t*)in)[i]; break;
            case T_I32: x = (float)((const int32_t*)in)[i]; break;
            case T_U32: x = (float)((const uint32_t*)in)[i]; break;
            case T_I64: x = (float)((const int64_t*)in)[i]; break;
            case T_U64: x = (float)((const uint64_t*)in)[i]; break;
            case T_F32: x = ((const float*)in)[i]; break;
            default: return -1;
        }
        if (x >= lo && x < up)
        {
            const float f = (x - lo) * scale;
            uint32_t b = (uint32_t)f;
            if (b >= bins) b = (uint32_t)bins - 1u;
            ++hist[b];
        }
    }
    return 0;
}

/* ---- segmented sort (no counterpart in the reference) -------------------- */

typedef struct
{
    int key_type, desc, id, nthreads;
    char* keys;
    uint32_t* values;
    size_t kbytes;
    const uint64_t* off;
    size_t nseg;
    int rc;
} seg_job;

static void* seg_worker(void* p)
{
    seg_job* j = (seg_job*)p;
    for (size_t i = (size_t)j->id; i < j->nseg; i += (size_t)j->nthreads)
    {
        const uint64_t b = j->off[i], e = j->off[i + 1];
        if (e <= b + 1) continue;
        int rc = oracle_radix_sort(j->key_type, j->keys + b * j->kbytes,
                                   j->values ? j->values + b : NULL, (size_t)(e - b),
                                   j->desc, 1);
        if (rc) j->rc = rc;
    }
    return NULL;
}

/* Every segment [off[i], off[i+1]) sorted on its own by the restated LSD
 * sort above: the per-segment loop of clc_sort calls a caller of the
 * reference would write. Segments with end <= begin + 1 are left as they
 * are. */
int oracle_segmented_sort(int key_type, void* keys, uint32_t* values,
                          const uint64_t* off, size_t nseg, int descending,
                          int threads)
{
    size_t kbytes;
    if (key_type == T_U32 || key_type == T_I32 || key_type == T_F32) kbytes = 4;
    else if (key_type == T_U64 || key_type == T_I64 || key_type == T_F64) kbytes = 8;
    else return -1;
    if (threads <= 0) threads = 1;
    if (threads > MAX_THREADS) threads = MAX_THREADS;
    seg_job jobs[MAX_THREADS];
    pthread_t tid[MAX_THREADS];
    for (int t = 0; t < threads; ++t)
    {
        seg_job jb = { key_type, descending != 0, t, threads, (char*)keys, values,
                       kbytes, off, nseg, 0 };
        jobs[t] = jb;
        if (t > 0) pthread_create(&tid[t], NULL, seg_worker, &jobs[t]);
    }
    seg_worker(&jobs[0]);
    int rc = jobs[0].rc;
    for (int t = 1; t < threads; ++t)
    {
        pthread_join(tid[t], NULL);
        if (jobs[t].rc) rc = jobs[t].rc;
    }
    return rc;
}
