// engine.hip -- device context, pooled HBM, and the prefix-doubling pipeline.
//
// Pipeline for one text (all on one stream):
//   histogram (freq, alphabet)  ->  packed keys of the first k symbols
//   ->  onesweep sort  ->  groups (ISA = group head, SA for singletons, active list)
//   ->  rounds h = k, 2k, 4k, ...: key (group, rank of i+h) -> onesweep sort
//       -> regroup, retire singletons, compact the active list   until none left
//   ->  optional BWT gather, PLCP (irreducible + prefix max), LCP gather.
#include "engine.h"
#include "../../include/sagpu.h"

#include <algorithm>
#include <atomic>
#include <condition_variable>
#include <map>
#include <mutex>
#include <new>
#include <pthread.h>
#include <stdio.h>
#include <string.h>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#define SAGPU_HIDDEN __attribute__((visibility("hidden")))

// ------------------------------------------------------------------ errors
static thread_local std::string g_err;

void sagpu_set_error(const char *msg, const char *file, int line)
{
    char buf[512];
    snprintf(buf, sizeof buf, "%s (%s:%d)", msg, file, line);
    g_err = buf;
}

// ------------------------------------------------------------------ pool
namespace {

// Cap on the live bytes of each context's pool (sagpu_set_memory_limit; 0 = none).
std::atomic<u64> g_mem_limit{0};
// Device and pinned-host allocations the engine has made for its working memory
// (pools, staging); reported in sagpu_stats.device_allocs.
std::atomic<long long> g_device_allocs{0};

void trim_idle_slots();

// Device memory of one context.  Buffers are cached by size and reused by later
// calls (stream-ordered: a buffer returned while kernels on the context's stream
// still use it is only handed out again to work queued behind them on the same
// stream).
struct Pool {
    std::multimap<size_t, void *> cache;
    std::unordered_map<void *, size_t> live;
    size_t live_bytes = 0;
    size_t peak_bytes = 0;   // largest live_bytes since the last reset_peak()

    // Arena mode (the allocation-free build): buffers are carved from one fixed
    // device region, bottom up, and never come from hipMalloc.  A freed block is
    // reclaimed once every block above it is freed too (the engine frees its
    // buffers in scopes, so the region works as a stack).
    uint8_t *arena = nullptr;
    size_t arena_cap = 0, arena_top = 0;
    struct Block { size_t off, end; bool freed; };
    std::vector<Block> arena_blocks;

    static size_t round(size_t b)
    {
        if (b < (1u << 20)) return (b + 255) & ~(size_t)255;
        return (b + (2u << 20) - 1) & ~(size_t)((2u << 20) - 1);
    }

    void reset_peak() { peak_bytes = live_bytes; }

    void *arena_get(size_t bytes)
    {
        const size_t off = (arena_top + 255) & ~(size_t)255;
        if (off > arena_cap || bytes > arena_cap - off) {
            sagpu_set_error("allocation-free build: the static scratch is too small for this call",
                            __FILE__, __LINE__);
            return nullptr;
        }
        arena_blocks.push_back({off, off + bytes, false});
        arena_top = off + bytes;
        live_bytes = arena_top;
        if (live_bytes > peak_bytes) peak_bytes = live_bytes;
        return arena + off;
    }

    void arena_put(void *p)
    {
        const size_t off = (size_t)((uint8_t *)p - arena);
        for (auto it = arena_blocks.rbegin(); it != arena_blocks.rend(); ++it)
            if (it->off == off && !it->freed) {
                it->freed = true;
                break;
            }
        while (!arena_blocks.empty() && arena_blocks.back().freed) arena_blocks.pop_back();
        arena_top = arena_blocks.empty() ? 0 : arena_blocks.back().end;
        live_bytes = arena_top;
    }

    void *get(size_t bytes)
    {
        if (arena) return arena_get(bytes ? bytes : 1);
        bytes = round(bytes ? bytes : 1);
        const u64 limit = g_mem_limit.load();
        if (limit && live_bytes + bytes > limit) {
            sagpu_set_error("device memory limit reached", __FILE__, __LINE__);
            return nullptr;
        }
        auto it = cache.lower_bound(bytes);
        if (it != cache.end() && it->first <= bytes + bytes / 4 + (4u << 20)) {
            void *p = it->second;
            size_t sz = it->first;
            cache.erase(it);
            live[p] = sz;
            live_bytes += sz;
            if (live_bytes > peak_bytes) peak_bytes = live_bytes;
            return p;
        }
        void *p = nullptr;
        g_device_allocs++;
        if (hipMalloc(&p, bytes) != hipSuccess) {
            (void)hipGetLastError();
            trim();
            trim_idle_slots();
            if (hipMalloc(&p, bytes) != hipSuccess) {
                (void)hipGetLastError();
                sagpu_set_error("hipMalloc failed", __FILE__, __LINE__);
                return nullptr;
            }
        }
        live[p] = bytes;
        live_bytes += bytes;
        if (live_bytes > peak_bytes) peak_bytes = live_bytes;
        return p;
    }

    void put(void *p)
    {
        if (!p) return;
        if (arena) {
            arena_put(p);
            return;
        }
        auto it = live.find(p);
        if (it == live.end()) return;
        cache.emplace(it->second, p);
        live_bytes -= it->second;
        live.erase(it);
    }

    void trim()
    {
        for (auto &kv : cache) (void)hipFree(kv.second);
        cache.clear();
    }
};

struct Timer {
    bool on = false;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> free_ev;
    struct Rec { int family; hipEvent_t a, b; u64 bytes; };
    std::vector<Rec> recs;
    std::pair<hipEvent_t, hipEvent_t> open[KI_COUNT] = {};   // kernel spans begun, not ended
    u64 hook_n = 0;   // elements of the sort whose passes the radix hook brackets
};

// ------------------------------------------------------ host copies
// Bulk copies between caller host memory (pageable) and the device run through
// pinned slots: the host side of each chunk is copied by a pool of threads while
// the DMA of the neighbouring chunk runs.  On the MI355X host, pageable hipMemcpy
// moves 22-29 GB/s to the device and 4 GB/s into pages the caller has not touched
// yet (first-touch faults taken by one thread); the staged form 54 GB/s and
// 27-29 GB/s with 16 copy threads.
constexpr int XS_SLOTS = 4;
// 1: copies of at least 4 MiB staged in 32 MiB chunks (default); 0: every copy
// direct; 2: every copy staged in 64 KiB chunks (exercises the chunk logic on
// small inputs)
int g_host_staging = 1;
int g_copy_threads = 0;   // 0: min(16, hardware threads)

int copy_threads()
{
    if (g_copy_threads > 0) return g_copy_threads;
    const unsigned hw = std::thread::hardware_concurrency();
    return hw == 0 ? 4 : (int)(hw < 16 ? hw : 16);
}

// memcpy split over worker threads; the calling thread takes parts too.  Workers
// are detached and live until the process ends; a forked child starts a new pool.
class CopyPool {
    struct Job {
        uint8_t *d;
        const uint8_t *s;
        u64 len, per;
        int parts, next, left;
    };
    std::mutex m;
    std::condition_variable work_cv, done_cv;
    std::vector<Job *> open;   // jobs with parts not taken yet
    int workers = 0;

    static void run(Job *j, int p)
    {
        const u64 a = (u64)p * j->per;
        if (a >= j->len) return;
        const u64 b = a + j->per < j->len ? a + j->per : j->len;
        memcpy(j->d + a, j->s + a, b - a);
    }
    // Takes the next part of the first open job; m is held.
    Job *take(int *p)
    {
        Job *j = open.front();
        *p = j->next++;
        if (j->next == j->parts) open.erase(open.begin());
        return j;
    }
    void loop()
    {
        std::unique_lock<std::mutex> lk(m);
        for (;;) {
            work_cv.wait(lk, [&] { return !open.empty(); });
            int p;
            Job *j = take(&p);
            lk.unlock();
            run(j, p);
            lk.lock();
            if (--j->left == 0) done_cv.notify_all();
        }
    }

  public:
    void copy(void *d, const void *s, u64 len, int threads)
    {
        if (threads <= 1 || len < (1u << 20)) {
            memcpy(d, s, len);
            return;
        }
        // parts of whole pages, at least 1 MiB
        u64 per = (len + threads - 1) / threads;
        per = per < (1u << 20) ? (1u << 20) : (per + 4095) & ~(u64)4095;
        Job j{(uint8_t *)d, (const uint8_t *)s, len, per, (int)((len + per - 1) / per), 0, 0};
        j.left = j.parts;
        std::unique_lock<std::mutex> lk(m);
        while (workers < threads - 1) {
            std::thread([this] { loop(); }).detach();
            workers++;
        }
        open.push_back(&j);
        work_cv.notify_all();
        while (j.next < j.parts) {
            int p = j.next++;
            if (j.next == j.parts) open.erase(std::find(open.begin(), open.end(), &j));
            lk.unlock();
            run(&j, p);
            lk.lock();
            j.left--;
        }
        done_cv.wait(lk, [&] { return j.left == 0; });
    }
    void lock() { m.lock(); }
    void unlock() { m.unlock(); }
};

CopyPool *g_copies = nullptr;
std::once_flag g_copies_once;

CopyPool &copies()
{
    std::call_once(g_copies_once, [] {
        g_copies = new CopyPool();
        // fork: hold the pool's lock across it; the child leaves the old pool (its
        // workers do not exist there) and makes a new one
        pthread_atfork([] { g_copies->lock(); }, [] { g_copies->unlock(); },
                       [] { g_copies = new CopyPool(); });
    });
    return *g_copies;
}

enum Family { F_ONESWEEP = 0, F_ONESWEEP32, F_INITIAL, F_ROUNDS, F_BWT, F_PLCP, F_LCP, F_TOTAL, F_PLCP_PAIRS, F_COUNT };
constexpr int F_KERNEL0 = 64;   // + KernelId: one kernel's launches

constexpr int ROUND_SIG = 31;   // words of a round's launch signature (Ctx::round_sig)

struct Ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    Pool pool;
    RadixScratch rs{};
    bool stats_fresh = false;   // stats were reset by the current call (copied out on release)
    u64 *pinned = nullptr;      // small pinned host staging (counts, histograms)
    uint8_t *stage = nullptr;   // large pinned staging for batched jobs
    size_t stage_cap = 0;
    // pinned slots for copies from and to caller host memory (h2d / d2h)
    uint8_t *xs = nullptr;
    u64 xs_chunk = 0;
    hipEvent_t xs_ev[XS_SLOTS] = {};
    bool xs_busy[XS_SLOTS] = {};
    int xs_next = 0;
    Timer timer;
    sagpu_stats stats{};
    bool sort_u32 = false;  // the radix passes being timed are u32-key passes
    // Executable graphs of a doubling round, one per buffer parity (rounds 2, 4, ...
    // and 3, 5, ...), kept across builds: reused as they are while every argument and
    // launch size of the round is the same (round_sig), else captured again and
    // updated in place.
    hipGraphExec_t round_exec[2] = {nullptr, nullptr};
    u64 round_sig[2][ROUND_SIG] = {};
};

// The engine context of the calling thread's current call (for ktime_*).
thread_local Ctx *t_ctx = nullptr;

std::mutex g_opt_mu;   // serialises option writes
bool g_timing = false;
int g_radix_variant = RADIX_ONESWEEP;
int g_radix_items = 32;
int g_radix_threads = 512;
int g_radix_bits = 8;
// Texts longer than this use 64-bit rank keys and PLCP scan values.
u64 g_wide_from = 0x7FFFFFFFull;
// 0: seeded from the initial sort where the symbol code fits beside the suffix
// index, else gather T[SA[j]-1] in SA order; 1: scatter in text order; 2: gather
int g_bwt_variant = 0;
// 1: byte texts (no GSA) sort keys generated from a packed code stream inside the
// first radix pass; 0: keys and values written by k_pack_keys32 first
int g_key_stream = 1;
// 1: with a code stream, the first round orders groups by the packed key of i + h
// instead of its rank ISA[i + h] (the same order, read from a smaller array)
int g_raw_round1 = 1;
// 1: the PLCP pass compares suffixes on the code stream when there is one
int g_stream_compare = 1;
// 1: round 1 on raw keys records exact lcps across its splits; 0: lower bounds
int g_exact_round1 = 1;
// 1: the fused build with LCP packs the exact lcp of each suffix alone in its
// initial group into its ISA word (texts up to 2^28), so the PLCP pass skips its
// pair-word read; 0: every position reads its pair word
int g_plcp_hint = 1;
// 1: Φ of a caller's suffix array by the bucketed two-pass scatter; 0: one pass
int g_bucketed_phi = 1;
// 1: the fused PLCP pass reads one pair word per position (k_plcp_pairs); 0: marks,
// SA and text gathered separately
int g_plcp_pairs = 1;
// 1: doubling rounds from the second on replay a captured graph (round parameters
// in device memory); 0: kernels launched one by one.  Timing mode launches them one
// by one (its events bracket single kernels).
int g_round_graphs = 1;
// 1: texts over 2^22 write the initial ranks through a radix pass into windows and
// a scatter in which each XCD writes only its own windows; 0: one direct scatter
int g_windowed_scatter = 1;
// Big groups of rounds after the first take up to BP frequent keys per group out of
// the sort when they hold at least this many members in all (0: never)
int g_big_pivot_min = 1 << 20;
// 1: texts of at most 2^31 suffixes flag group words whose members rank at another
// slot of the group's range (stored at the first slot + 1), so the subgroup whose
// slots hold that rank keeps its ISA entries when the group splits; 0: every group
// ranks at its first slot
int g_group_flags = 1;
// 1: the radix pass that groups the ISA pairs into windows takes its digit histogram
// from the closed form of a permutation's (the "perm_hist" option)
int g_perm_hist = 1;
// radix sorts of fewer keys use the smallest tiles (2048 keys), whatever radix_items
// and radix_threads say
int g_radix_small_n = 1 << 20;

// A context is one stream with its own pooled buffers and scratch; calls on one
// context are serialised, different contexts run concurrently on the device.
void apply_options(Ctx &c)
{
    c.rs.variant = g_radix_variant;
    c.rs.items = g_radix_items;
    c.rs.threads = g_radix_threads;
    c.rs.small_n = (u64)g_radix_small_n;
    // the allocation-free build's look-back scratch is sized for 8-bit digits
    c.rs.bits = c.pool.arena ? 8 : g_radix_bits;
}

void ctx_destroy(Ctx *c);

int ctx_create(Ctx **out, int device = 0)
{
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) {
        (void)hipGetLastError();
        sagpu_set_error("no HIP device available", __FILE__, __LINE__);
        return SAGPU_EDEVICE;
    }
    Ctx *c = new Ctx();
    c->device = device;
    int rc = SAGPU_OK;
    if (hipSetDevice(c->device) != hipSuccess || hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
        sagpu_set_error("stream creation failed", __FILE__, __LINE__);
        rc = SAGPU_EDEVICE;
    } else {
        g_device_allocs++;
        if (hipHostMalloc((void **)&c->pinned, 256 * 1024, hipHostMallocDefault) != hipSuccess) {
            sagpu_set_error("pinned host allocation failed", __FILE__, __LINE__);
            c->pinned = nullptr;
            rc = SAGPU_ENOMEM;
        }
    }
    if (rc != SAGPU_OK) {
        (void)hipGetLastError();
        ctx_destroy(c);
        return rc;
    }
    c->rs.hist_host = c->pinned + 16384;   // RADIX_HIST_WORDS u64 = 96 KiB
    apply_options(*c);
    *out = c;
    return SAGPU_OK;
}

void ctx_destroy(Ctx *c)
{
    if (!c) return;
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    if (c->rs.hist) c->pool.put(c->rs.hist);
    if (c->rs.tile_ctr) c->pool.put(c->rs.tile_ctr);
    if (c->rs.lookback) c->pool.put(c->rs.lookback);
    if (c->rs.counts) c->pool.put(c->rs.counts);
    if (c->rs.scan.partials) c->pool.put(c->rs.scan.partials);
    c->pool.trim();
    for (auto &kv : c->pool.live) (void)hipFree(kv.first);
    for (auto &r : c->timer.recs) {
        (void)hipEventDestroy(r.a);
        (void)hipEventDestroy(r.b);
    }
    for (auto &e : c->timer.free_ev) {
        (void)hipEventDestroy(e.first);
        (void)hipEventDestroy(e.second);
    }
    for (auto &e : c->round_exec)
        if (e) (void)hipGraphExecDestroy(e);
    if (c->stage) (void)hipHostFree(c->stage);
    if (c->xs) (void)hipHostFree(c->xs);
    for (auto &e : c->xs_ev)
        if (e) (void)hipEventDestroy(e);
    if (c->pinned) (void)hipHostFree(c->pinned);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

// ------------------------------------------------------------------ leases
// Every host entry point runs on one context for the whole call (a "lease").
//  - Calls made through a libsais context (libsais_ctx and the like) use that
//    context's own engine context: a stream and a pool of its own, created on the
//    context's first call.  Two threads with two contexts run concurrently on the
//    device, as two threads with two contexts do on the CPU in the reference
//    (include/libsais.h:54, src/libsais.c:7008-7016).
//  - Calls without a context take a free slot context (created on demand, up to
//    g_api_slots of them) and wait only when all are busy.
std::mutex g_slot_mu;
std::condition_variable g_slot_cv;
std::vector<Ctx *> g_slots;
std::vector<char> g_slot_busy;
int g_api_slots = 8;

struct Token {
    u32 magic = 0x5341474Bu;
    std::mutex mu;
    Ctx *c = nullptr;
};
thread_local Token *t_bound = nullptr;

std::mutex g_stats_mu;
sagpu_stats g_last_stats{};

// Engine contexts of freed libsais context handles, kept for the next handle's
// first call (a stream, pinned staging and a warm pool each), so that callers that
// create and free a context per call do not pay for stream and pool set-up.
std::vector<Ctx *> g_parked;
constexpr size_t PARKED_MAX = 16;

// Releases the pooled buffers of slot contexts no call holds (their streams are
// idle: every call synchronises its stream before it ends) and of parked contexts.
void trim_idle_slots()
{
    std::lock_guard<std::mutex> lk(g_slot_mu);
    for (size_t q = 0; q < g_slots.size(); q++)
        if (!g_slot_busy[q]) g_slots[q]->pool.trim();
    for (Ctx *c : g_parked) c->pool.trim();
}

// A context for a libsais context handle: a parked one, or a new one.
int ctx_for_token(Ctx **out)
{
    {
        std::lock_guard<std::mutex> lk(g_slot_mu);
        if (!g_parked.empty()) {
            *out = g_parked.back();
            g_parked.pop_back();
            return SAGPU_OK;
        }
    }
    return ctx_create(out);
}

// Parks the context of a freed handle, or destroys it when enough are parked.
void ctx_park(Ctx *c)
{
    if (!c) return;
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    {
        std::lock_guard<std::mutex> lk(g_slot_mu);
        if (g_parked.size() < PARKED_MAX) {
            g_parked.push_back(c);
            return;
        }
    }
    ctx_destroy(c);
}

struct Lease {
    Ctx *ctx = nullptr;
    Token *tok = nullptr;
    int slot = -1;
    Ctx *prev_ctx = nullptr;

    int acquire()
    {
        if (t_bound) {
            tok = t_bound;
            tok->mu.lock();
            if (!tok->c) {
                const int rc = ctx_for_token(&tok->c);
                if (rc != SAGPU_OK) {
                    tok->c = nullptr;
                    tok->mu.unlock();
                    tok = nullptr;
                    return rc;
                }
            }
            ctx = tok->c;
        } else {
            std::unique_lock<std::mutex> lk(g_slot_mu);
            for (;;) {
                for (size_t q = 0; q < g_slots.size() && slot < 0; q++)
                    if (!g_slot_busy[q]) slot = (int)q;
                if (slot >= 0) break;
                if ((int)g_slots.size() < g_api_slots) {
                    Ctx *c = nullptr;
                    SAGPU_TRY(ctx_create(&c));
                    g_slots.push_back(c);
                    g_slot_busy.push_back(0);
                    slot = (int)g_slots.size() - 1;
                    break;
                }
                g_slot_cv.wait(lk);
            }
            g_slot_busy[slot] = 1;
            ctx = g_slots[slot];
        }
        (void)hipSetDevice(ctx->device);
        apply_options(*ctx);
        ctx->stats_fresh = false;
        ctx->pool.reset_peak();
        prev_ctx = t_ctx;
        t_ctx = ctx;
        return SAGPU_OK;
    }

    ~Lease()
    {
        if (!ctx) return;
        t_ctx = prev_ctx;
        if (ctx->stats_fresh) {
            ctx->stats.peak_bytes = (int64_t)ctx->pool.peak_bytes;
            std::lock_guard<std::mutex> lk(g_stats_mu);
            g_last_stats = ctx->stats;
        }
        if (tok) {
            tok->mu.unlock();
        } else {
            {
                std::lock_guard<std::mutex> lk(g_slot_mu);
                g_slot_busy[slot] = 0;
            }
            g_slot_cv.notify_one();
        }
    }
};

#define LEASE(c)                   \
    Lease lease_;                  \
    SAGPU_TRY(lease_.acquire());   \
    Ctx *c = lease_.ctx

// Per-call scope: every buffer taken from the pool returns to it on exit.
struct Scope {
    Ctx &c;
    std::vector<void *> held;
    explicit Scope(Ctx &ctx) : c(ctx) {}
    ~Scope()
    {
        for (void *p : held) c.pool.put(p);
    }
    template <typename T> T *get(u64 count)
    {
        void *p = c.pool.get((size_t)count * sizeof(T));
        if (p) held.push_back(p);
        return (T *)p;
    }
};

#define GET(var, T, count)                             \
    T *var = sc.get<T>(count);                         \
    if (!var) return SAGPU_ENOMEM

// ------------------------------------------------------------------ timing
struct Span {
    Ctx &c;
    int fam;
    hipEvent_t a = nullptr, b = nullptr;
    Span(Ctx &ctx, int family) : c(ctx), fam(family)
    {
        if (!c.timer.on) return;
        if (!c.timer.free_ev.empty()) {
            a = c.timer.free_ev.back().first;
            b = c.timer.free_ev.back().second;
            c.timer.free_ev.pop_back();
        } else {
            (void)hipEventCreate(&a);
            (void)hipEventCreate(&b);
        }
        (void)hipEventRecord(a, c.stream);
    }
    ~Span()
    {
        if (!a) return;
        (void)hipEventRecord(b, c.stream);
        c.timer.recs.push_back({fam, a, b, 0});
    }
};

void timer_collect(Ctx &c)
{
    if (c.timer.recs.empty()) return;
    (void)hipStreamSynchronize(c.stream);
    for (auto &r : c.timer.recs) {
        float ms = 0;
        (void)hipEventElapsedTime(&ms, r.a, r.b);
        switch (r.family) {
        case F_ONESWEEP: c.stats.ms_onesweep += ms; c.stats.onesweep_launches++; break;
        case F_ONESWEEP32:
            c.stats.ms_onesweep += ms;
            c.stats.onesweep_launches++;
            c.stats.ms_onesweep_u32 += ms;
            c.stats.onesweep_launches_u32++;
            break;
        case F_INITIAL: c.stats.ms_initial += ms; break;
        case F_ROUNDS: c.stats.ms_rounds += ms; break;
        case F_BWT: c.stats.ms_bwt += ms; break;
        case F_PLCP: c.stats.ms_plcp += ms; break;
        case F_LCP: c.stats.ms_lcp += ms; break;
        case F_TOTAL: c.stats.ms_total += ms; break;
        case F_PLCP_PAIRS: c.stats.ms_plcp_pairs_pass += ms; c.stats.plcp_pairs_launches++; break;
        default:
            if (r.family >= F_KERNEL0 && r.family < F_KERNEL0 + KI_COUNT) {
                const int id = r.family - F_KERNEL0;
                c.stats.kernel_ms[id] += ms;
                c.stats.kernel_launches[id]++;
                c.stats.kernel_bytes[id] += (int64_t)r.bytes;
            }
        }
        c.timer.free_ev.push_back({r.a, r.b});
    }
    c.timer.recs.clear();
}

} // namespace

void ktime_begin(int id)
{
    Ctx *c = t_ctx;
    if (!c || !c->timer.on || id < 0 || id >= KI_COUNT) return;
    hipEvent_t a = nullptr, b = nullptr;
    if (!c->timer.free_ev.empty()) {
        a = c->timer.free_ev.back().first;
        b = c->timer.free_ev.back().second;
        c->timer.free_ev.pop_back();
    } else if (hipEventCreate(&a) != hipSuccess || hipEventCreate(&b) != hipSuccess) {
        (void)hipGetLastError();
        return;
    }
    (void)hipEventRecord(a, c->stream);
    c->timer.open[id] = {a, b};
}

void ktime_end(int id, u64 bytes)
{
    Ctx *c = t_ctx;
    if (!c || !c->timer.on || id < 0 || id >= KI_COUNT || !c->timer.open[id].first) return;
    auto ev = c->timer.open[id];
    c->timer.open[id] = {};
    (void)hipEventRecord(ev.second, c->stream);
    c->timer.recs.push_back({F_KERNEL0 + id, ev.first, ev.second, bytes});
}

namespace {

// ------------------------------------------------------------------ pieces
int bits_for(u64 v)
{
    int b = 0;
    while (b < 64 && (v >> b) != 0) b++;
    return b;
}

int ensure_radix_scratch(Ctx &c, Scope &sc, u64 n)
{
    u64 tiles = (n + RS_TILE - 1) / RS_TILE + 1;
    if (!c.rs.hist) {
        c.rs.hist = (u64 *)c.pool.get(2 * RADIX_HIST_WORDS * sizeof(u64));
        c.rs.tile_ctr = (u32 *)c.pool.get(64 * sizeof(u32));
        if (!c.rs.hist || !c.rs.tile_ctr) return SAGPU_ENOMEM;
    }
    if (c.rs.counts_cap < tiles * 256) {
        if (c.rs.counts) c.pool.put(c.rs.counts);
        if (c.rs.scan.partials) c.pool.put(c.rs.scan.partials);
        c.rs.counts = (u32 *)c.pool.get(tiles * 256 * sizeof(u32));
        c.rs.scan.cap = tiles * 256 / 4096 + 2;
        c.rs.scan.partials = (u32 *)c.pool.get(c.rs.scan.cap * sizeof(u32));
        if (!c.rs.counts || !c.rs.scan.partials) { c.rs.counts_cap = 0; return SAGPU_ENOMEM; }
        c.rs.counts_cap = tiles * 256;
    }
    const u64 words = tiles * ((u64)1 << (c.rs.bits > 8 ? c.rs.bits : 8));   // look-back words per tile: one per digit
    if (c.rs.lb_words < words) {
        if (c.rs.lookback) c.pool.put(c.rs.lookback);
        c.rs.lookback = (u64 *)c.pool.get(words * sizeof(u64));
        if (!c.rs.lookback) { c.rs.lb_words = 0; return SAGPU_ENOMEM; }
        c.rs.lb_words = words;
        SAGPU_TRY(launch_fill_u64(c.rs.lookback, words, 0, c.stream));
        c.rs.epoch = 0;
    }
    (void)sc;
    return SAGPU_OK;
}

// Brackets every onesweep pass launch with a pair of events (timing mode only).
void onesweep_hook(void *user, int phase)
{
    Ctx &c = *(Ctx *)user;
    static thread_local Span *open = nullptr;
    static thread_local int kid = 0;
    if (phase == 0) {
        open = new Span(c, c.sort_u32 && !c.rs.gen_pass ? F_ONESWEEP32 : F_ONESWEEP);
        kid = !c.sort_u32 ? KI_RADIX64 : c.rs.gen_pass ? KI_RADIX_GEN : KI_RADIX32;
        ktime_begin(kid);
    } else {
        const u64 n = c.timer.hook_n;
        // u32 pass: key + value read and written; generating pass: the code stream
        // read, key + value written; u64 pass: 8-byte key + value read and written
        const u64 bytes = kid == KI_RADIX32 ? 16 * n
                          : kid == KI_RADIX_GEN ? 8 * n + n * (u64)(c.rs.gen ? c.rs.gen->bits : 0) / 8
                                                : 24 * n;
        ktime_end(kid, bytes);
        delete open;
        open = nullptr;
    }
}

// Brackets the pair-word PLCP pass (timing mode only); nested inside the F_PLCP span.
void plcp_pairs_hook(void *user, int phase)
{
    Ctx &c = *(Ctx *)user;
    static thread_local Span *open = nullptr;
    if (phase == 0) {
        open = new Span(c, F_PLCP_PAIRS);
        ktime_begin(KI_PLCP_PASS);
    } else {
        // ISA 4, the preceding byte 1, one pair word 8, the value written 4
        ktime_end(KI_PLCP_PASS, 17 * c.timer.hook_n);
        delete open;
        open = nullptr;
    }
}

int sort_u64(Ctx &c, u64 *k0, u32 *v0, u64 *k1, u32 *v1, u64 n, int begin_bit, int end_bit, int *alt)
{
    int passes = 0;
    c.rs.hook = c.timer.on ? onesweep_hook : nullptr;
    c.rs.hook_user = &c;
    c.sort_u32 = false;
    c.timer.hook_n = n;
    int rc = radix_sort_u64(c.rs, k0, v0, k1, v1, n, begin_bit, end_bit, c.stream, alt, &passes);
    c.stats.sort_passes += passes;
    c.stats.sorted_elems += (int64_t)passes * (int64_t)n;
    c.stats.sorted_bytes += (int64_t)passes * (int64_t)n * 24;
    return rc;
}

int sort_u32(Ctx &c, u32 *k0, u32 *v0, u32 *k1, u32 *v1, u64 n, int begin_bit, int end_bit, int *alt)
{
    int passes = 0;
    c.rs.hook = c.timer.on ? onesweep_hook : nullptr;
    c.rs.hook_user = &c;
    c.sort_u32 = true;
    c.timer.hook_n = n;
    int rc = radix_sort_u32(c.rs, k0, v0, k1, v1, n, begin_bit, end_bit, c.stream, alt, &passes);
    c.sort_u32 = false;
    c.stats.sort_passes += passes;
    c.stats.sorted_elems += (int64_t)passes * (int64_t)n;
    // a generating first pass reads the code stream and writes 8 bytes per key; it
    // is counted with all passes, not with the array-to-array u32 passes
    const bool gen = c.rs.gen && c.rs.variant != RADIX_REDUCE_THEN_SCAN && passes > 0;
    const int plain = passes - (gen ? 1 : 0);
    c.stats.sorted_bytes += (int64_t)plain * (int64_t)n * 16;
    if (gen) c.stats.sorted_bytes += (int64_t)n * 8 + (int64_t)(n * (u64)c.rs.gen->bits / 8);
    c.stats.sorted_bytes_u32 += (int64_t)plain * (int64_t)n * 16;
    return rc;
}

// ISA from the (suffix, rank) pairs of the initial grouping, bucketed by suffix >> cw
// (bucket-major).  Texts over 2^22 with at most 8 buckets: one radix pass on bits
// [cw - 8, cw) groups the pairs into windows of 2^(cw - 8) suffixes, and each XCD
// writes only its own windows (launch_scatter_windows); otherwise one pass over the
// pairs in bucket order.  K1 / V1: free buffers of n words for the radix pass.
int isa_from_pairs(Ctx &c, Scope &sc, const u32 *PV, const u32 *PG, u64 n, u32 *ISA, u32 *K1, u32 *V1)
{
    const int cw = isa_bucket_log2(n);
    const u64 nc = ((n - 1) >> cw) + 1;
    if (g_windowed_scatter && n > (1ull << 22) && nc <= 8 && cw >= 16) {
        int alt = 0;
        // PV holds every suffix once: the partition pass's histogram has a closed form
        c.rs.perm_keys = g_perm_hist;
        const int src = sort_u32(c, (u32 *)PV, (u32 *)PG, K1, V1, n, cw - 8, cw, &alt);
        c.rs.perm_keys = 0;
        SAGPU_TRY(src);
        GET(bounds, u64, 256 * nc + 1);
        GET(ctr, u32, 256);
        ktime_begin(KI_ISA_SCATTER);
        SAGPU_TRY(launch_scatter_windows(alt ? K1 : PV, alt ? V1 : PG, n, cw - 8, (u32)nc, cw, ISA, bounds, ctr,
                                         c.stream));
        ktime_end(KI_ISA_SCATTER, 12 * n);   // pair read, rank written
        return SAGPU_OK;
    }
    ktime_begin(KI_ISA_SCATTER);
    SAGPU_TRY(launch_isa_scatter(PV, PG, n, ISA, c.stream));
    ktime_end(KI_ISA_SCATTER, 12 * n);
    return SAGPU_OK;
}

int read_u32(Ctx &c, const u32 *d, u32 *out)
{
    SAGPU_CHECK(hipMemcpyAsync(c.pinned, d, sizeof(u32), hipMemcpyDeviceToHost, c.stream));
    SAGPU_CHECK(hipStreamSynchronize(c.stream));
    *out = *(u32 *)c.pinned;
    return SAGPU_OK;
}

// Byte alphabet of a text: freq (optional, host), the GSA validity test, and the
// dense symbol codes (device LUT) with their width and the symbols per 32-bit key.
int byte_alphabet(Ctx &c, Scope &sc, const uint8_t *dT, u64 n, int gsa, u64 *freq_out, int *bits_out, int *ksym_out,
                  uint8_t **lut_out, uint8_t **inv_out = nullptr)
{
    hipStream_t s = c.stream;
    GET(d_hist, u64, 256 + 8);
    u32 *d_bad = (u32 *)(d_hist + 256);
    u64 freq[256];
    {
        Span sp(c, F_INITIAL);
        SAGPU_TRY(launch_fill_u64(d_hist, 256 + 8, 0, s));
        SAGPU_TRY(launch_histogram(dT, n, gsa, d_hist, d_bad, s));
    }
    SAGPU_CHECK(hipMemcpyAsync(c.pinned, d_hist, (256 + 1) * sizeof(u64), hipMemcpyDeviceToHost, s));
    SAGPU_CHECK(hipStreamSynchronize(s));
    memcpy(freq, c.pinned, sizeof freq);
    u32 bad = *(u32 *)(c.pinned + 256);
    if (freq_out) memcpy(freq_out, freq, sizeof freq);
    if (gsa && bad) return SAGPU_EINVAL;

    uint8_t *lut_h = (uint8_t *)(c.pinned + 4096);
    int sigma = 0;
    for (int ch = 0; ch < 256; ch++) {
        lut_h[ch] = (uint8_t)sigma;
        if (freq[ch]) sigma++;
    }
    *bits_out = sigma <= 2 ? 1 : bits_for((u64)sigma - 1);
    *ksym_out = 32 / *bits_out;
    GET(d_lut, uint8_t, 256);
    SAGPU_CHECK(hipMemcpyAsync(d_lut, lut_h, 256, hipMemcpyHostToDevice, s));
    *lut_out = d_lut;
    if (inv_out) {  // code -> byte
        uint8_t *inv_h = lut_h + 256;
        memset(inv_h, 0, 256);
        for (int ch = 0; ch < 256; ch++)
            if (freq[ch]) inv_h[lut_h[ch]] = (uint8_t)ch;
        GET(d_inv, uint8_t, 256);
        SAGPU_CHECK(hipMemcpyAsync(d_inv, inv_h, 256, hipMemcpyHostToDevice, s));
        *inv_out = d_inv;
    }
    return SAGPU_OK;
}

// Suffix array of the padded device text dT[0..n) (2 <= n <= 2^32; above
// g_wide_from, INT32_MAX by default, the rounds use 64-bit rank keys).  On
// return dSA holds the suffix array and dISA its inverse.  freq_out (host, 256)
// optional.  dLCP (optional) receives, per SA slot j >= 1, the exact lcp with the
// previous suffix when the initial keys separate them, else LCP_UNKNOWN | bound.
//
// seed (optional, byte texts without GSA): when the BWT symbol's code fits beside
// the suffix index in the sort values, the build leaves BWT seeds there for
// launch_bwt_seeded (seed->cb > 0 on return); otherwise seed->cb stays 0.
int sa_build(Ctx &c, Scope &sc, const uint8_t *dT, u64 n, int gsa, u32 *dSA, u32 *dISA, u64 *freq_out, u32 *dLCP,
             const u32 *sid = nullptr, u32 nstrings = 0, const u32 *dSym = nullptr, int sym_bits = 0,
             BwtSeed *seed = nullptr, CodeStream *cs = nullptr, PlcpHint *hint = nullptr)
{
    hipStream_t s = c.stream;
    sagpu_stats &st = c.stats;
    int bits, ksym;
    uint8_t *d_lut = nullptr, *d_inv = nullptr;
    const bool want_seed = seed && !dSym && !sid && !gsa;
    if (seed) *seed = BwtSeed{};
    if (dSym) {
        // integer alphabet: symbols already below 2^sym_bits
        bits = sym_bits < 1 ? 1 : sym_bits;
        ksym = 32 / bits;
    } else {
        SAGPU_TRY(byte_alphabet(c, sc, dT, n, gsa, freq_out, &bits, &ksym, &d_lut, want_seed ? &d_inv : nullptr));
    }
    if (want_seed && ((n - 1) >> (32 - bits)) == 0) {
        GET(codes, uint8_t, n);
        *seed = BwtSeed{bits, codes, d_inv};
    }
    const int cb = seed ? seed->cb : 0;
    st.initial_symbols = ksym;
    st.bits_per_symbol = bits;


    // Buffer plan (each block is reused once its first use is over):
    //   kA  [2n u32, 4n for batched u64 keys]: initial sort keys (ping-pong), then
    //       the per-slot group heads, then the second active list
    //   vA  [2n u32]: initial sort values (ping-pong), then key2 (u32 or u64)
    //   aA  [2n u32]: first active list (suffix, group head)
    // Groups too big for one workgroup get their own sort buffers per round.
    const bool wide = n > g_wide_from;
    GET(kA, u32, sid ? 4 * n : 2 * n);
    GET(vA, u32, 2 * n);
    GET(aA, u32, 2 * n);
    const u64 big_cap = n / (RT_SMALL + 1) + 16;
    GET(big_head, u32, big_cap);
    GET(big_g, u32, big_cap);
    GET(big_size, u32, big_cap);
    GET(d_cnt, u64, 8);
    GET(d_boff, u64, big_cap + 1);
    // the tile kernel's big heads, in one segment per XCD before k_big_sizes joins them
    const u64 bseg_cap = round_bigseg_cap(n);
    GET(bseg_h, u32, 8 * bseg_cap);
    GET(bseg_g, u32, 8 * bseg_cap);
    u64 *out_counter = d_cnt;
    u32 *big_counter = (u32 *)(d_cnt + 4), *err = (u32 *)(d_cnt + 4) + 1;
    u32 *act_i[2] = {aA, kA}, *act_g[2] = {aA + n, kA + n};
    void *key2 = vA;

    WsScratch ws{};
    ws.tiles_cap = n / (n < INIT_SMALL_N ? 512 : 2048) + 2;   // tiles of the initial grouping
    {
        // the tile aggregates, then the aggregates of their 8192-tile chunks
        GET(tagg, u64, (ws.tiles_cap + ws.tiles_cap / 8192 + 2) * 2);
        ws.tile_agg = tagg;
        ws.total = d_cnt + 2;
        ws.base = d_cnt + 3;
    }
    SAGPU_TRY(ensure_radix_scratch(c, sc, n));
    {
        // bucket counts of the initial grouping: at most 64 buckets per tile
        ws.bcnt_cap = 64 * ws.tiles_cap;
        GET(bcnt, u32, ws.bcnt_cap);
        ws.bcnt = bcnt;
        ws.scan = c.rs.scan;
    }

    int alt = 0;
    bool a_known = false;   // the first list's length is on the host already
    RawKey1 rk1{};   // first round from the code stream (when there is one)
    ExactLcp ex1{};  // and its exact lcps across the splits
    if (!sid) {
        u32 *k0 = kA, *k1 = kA + n, *v0 = vA, *v1 = vA + n;
        KeyGen gen{};
        bool use_gen = false;
        {
            Span sp(c, F_INITIAL);
            if (dSym) {
                SAGPU_TRY(launch_pack_int(dSym, n, bits, ksym, gsa, k0, v0, s));
            } else if (!gsa && g_key_stream && n <= (1ull << 31)) {
                // keys and values are generated by the sort's first pass from a
                // packed code stream (n * bits / 8 bytes) instead of written here
                GET(pw, u32, pack_stream_words(n, bits));
                SAGPU_TRY(launch_pack_stream(dT, n, d_lut, bits, pw, s));
                gen = KeyGen{pw, pack_stream_words(n, bits), bits, ksym * bits, cb};
                use_gen = true;
                if (cs) *cs = CodeStream{pw, bits};
            } else {
                SAGPU_TRY(launch_pack_keys32(dT, n, d_lut, bits, ksym, gsa, cb, k0, v0, s));
            }
        }
        c.rs.gen = use_gen ? &gen : nullptr;
        const int src = sort_u32(c, k0, v0, k1, v1, n, 0, ksym * bits, &alt);
        c.rs.gen = nullptr;
        SAGPU_TRY(src);
        Span sp(c, F_INITIAL);
        if (hint) {   // byte texts without separators, up to 2^28 suffixes
            hint->on = hint->on && !gsa && !dSym && n <= (1ull << 28);
            // the 15 hint values cover marks [off, off + 14]: the largest marks of the
            // initial grouping (below ksym) and the first round's exact marks (ksym
            // and a few above)
            hint->off = ksym > 10 ? ksym - 10 : 0;
        }
        // the unsorted key buffer is free now: it holds the per-slot group heads
        SAGPU_TRY(group_init(ws, alt ? k1 : k0, alt ? v1 : v0, n, bits, ksym, gsa, dSA, dISA, act_i[0], act_g[0], dLCP,
                             alt ? k0 : k1, alt ? v0 : v1, s, cb ? seed : nullptr, hint));
        // the sorted keys and values are dead now: they take the pairs' radix pass
        SAGPU_TRY(isa_from_pairs(c, sc, alt ? k0 : k1, alt ? v0 : v1, n, dISA, alt ? k1 : k0, alt ? v1 : v0));
        if (use_gen && g_raw_round1) {
            // group heads of the last 2 ksym suffixes: those of the last ksym (shorter
            // than ksym) keep ranks in round 1; none of them gets exact round-1 lcps
            const u64 cnt2 = n < 2 * (u64)ksym ? n : 2 * (u64)ksym;
            const u64 cnt = n < (u64)ksym ? n : (u64)ksym;
            u32 *hs = (u32 *)(c.pinned + 6144);   // past the histogram and LUT staging
            SAGPU_CHECK(hipMemcpyAsync(hs, dISA + (n - cnt2), cnt2 * sizeof(u32), hipMemcpyDeviceToHost, s));
            SAGPU_CHECK(hipMemcpyAsync(c.pinned, ws.total, sizeof(u64), hipMemcpyDeviceToHost, s));   // the list length
            SAGPU_CHECK(hipStreamSynchronize(s));
            a_known = true;
            rk1.pw = gen.pw;
            rk1.bits = bits;
            rk1.kbits = ksym * bits;
            rk1.nspecial = (int)cnt;
            for (u64 t = 0; t < cnt2; t++) hs[t] = isa_rank(hs[t], n);
            for (u64 t = 0; t < cnt; t++) rk1.special[t] = hs[cnt2 - cnt + t];
            if (g_exact_round1) {
                ex1.on = 1;
                ex1.bits = bits;
                ex1.kbits = ksym * bits;
                ex1.nskip = (int)cnt2;
                for (u64 t = 0; t < cnt2; t++) ex1.skip[t] = hs[t];
                // suffixes placed by the first round with an exact mark carry it as a
                // PLCP hint too (their ISA words are written anyway)
                ex1.hint = hint && hint->on ? 1 : 0;
                ex1.hint_off = hint ? hint->off : 0;
            }
        }
    } else {
        // batched GSA: (string id, symbols) keys, so every group stays inside one string
        u64 *k0 = (u64 *)kA, *k1 = (u64 *)kA + n;
        u32 *v0 = vA, *v1 = vA + n;
        {
            Span sp(c, F_INITIAL);
            SAGPU_TRY(launch_pack_keys64_sid(dT, n, d_lut, bits, ksym, sid, k0, v0, s));
        }
        SAGPU_TRY(sort_u64(c, k0, v0, k1, v1, n, 0, 32 + bits_for(nstrings ? nstrings - 1 : 0), &alt));
        Span sp(c, F_INITIAL);
        SAGPU_TRY(group_init64(ws, alt ? k1 : k0, alt ? v1 : v0, n, bits, ksym, dSA, dISA, act_i[0], act_g[0], dLCP,
                               (u32 *)(alt ? k0 : k1), (u32 *)(alt ? k0 : k1) + n, s));
        SAGPU_TRY(isa_from_pairs(c, sc, (u32 *)(alt ? k0 : k1), (u32 *)(alt ? k0 : k1) + n, n, dISA,
                                 (u32 *)(alt ? k1 : k0), (u32 *)(alt ? k1 : k0) + n));
    }
    u64 a = 0;
    // Short texts without a code-stream round skip the wait for the first list's
    // length: the first round is launched for n entries and reads the length from
    // the device (its kernels stop at the list's end); the host learns it with the
    // round's counters.
    const bool lazy_a = !a_known && n < INIT_SMALL_N && !rk1.pw;
    u64 *h_a1 = c.pinned + 7204;   // past the round parameters and the pivot regions
    if (lazy_a) {
        a = n;
    } else {
        if (!a_known) {
            SAGPU_CHECK(hipMemcpyAsync(c.pinned, ws.total, sizeof(u64), hipMemcpyDeviceToHost, s));
            SAGPU_CHECK(hipStreamSynchronize(s));
        }
        a = c.pinned[0];
    }

    // list entries carry BWT codes when seeded; group words carry the stored-rank flag
    // when every slot fits 31 bits
    const ActCode ac{cb, cb ? seed->codes : nullptr, g_group_flags && n <= (1ull << 31) ? 0x80000000u : 0u};
    u64 h = (u64)ksym;
    int cur = 0;
    st.rounds = 0;
    // The round's list length and offset reach the kernels through device memory
    // (d_rp), copied from pinned memory as the round's first operation, so rounds
    // from the second on replay one captured graph per buffer parity: parameter copy,
    // gather, counter reset, tile kernel, big-group sizes, counter copy back.  The
    // graph's launch geometry is sized for the round it was captured in; later
    // rounds of the same parity have shorter lists.
    GET(d_rp, u64, 2);
    u64 *h_rp = c.pinned + 7168;   // past the histogram, LUT and tail-rank staging
    // per-XCD survivor segments of the tile kernel, sized for the first (longest) list
    const u64 seg_words = 8 * round_seg_cap(a);
    GET(seg_i, u32, seg_words);
    GET(seg_g, u32, seg_words);
    GET(seg_ctr, u64, 128);
    // (not in the allocation-free build: instantiating a graph allocates)
    const bool graphs = g_round_graphs && !c.timer.on && !c.pool.arena;
    bool captured[2] = {false, false};
    while (a > 0) {
        if (st.rounds < 64) st.active[st.rounds] = (int64_t)a;
        st.rounds++;
        const int nxt = cur ^ 1;
        h_rp[0] = a;
        h_rp[1] = h;
        const bool raw = st.rounds == 1 && rk1.pw;
        const bool lazy_first = lazy_a && st.rounds == 1;
        // one round's work on the stream (launch geometry for a_grid entries)
        auto enqueue = [&](u64 a_grid, bool timed) -> int {
            SAGPU_CHECK(hipMemcpyAsync(d_rp, h_rp, 2 * sizeof(u64), hipMemcpyHostToDevice, s));
            if (lazy_first) SAGPU_CHECK(hipMemcpyAsync(d_rp, ws.total, sizeof(u64), hipMemcpyDeviceToDevice, s));
            if (timed) ktime_begin(KI_ROUND_GATHER);
            SAGPU_TRY(round_gather(act_i[cur], act_g[cur], a_grid, d_rp, dISA, n, key2, wide, s, raw ? &rk1 : nullptr,
                                   ac.cb));
            // suffix read, one random 4-byte read (rank, or code-stream window), key
            // written (+ the group read for the first round's form)
            if (timed) ktime_end(KI_ROUND_GATHER, a * ((wide ? 16 : 12) + (raw ? 4 : 0)));
            SAGPU_TRY(launch_fill_u64(d_cnt, 8, 0, s));
            // (the tile kernel is timed inside round_tile; the timer is off while a
            // round is captured)
            SAGPU_TRY(round_tile(act_i[cur], act_g[cur], key2, wide, a_grid, d_rp, dSA, dISA, act_i[nxt], act_g[nxt],
                                 out_counter, err, big_head, big_g, big_size, big_counter, big_cap, dLCP, s, ac,
                                 raw ? &ex1 : nullptr, seg_i, seg_g, seg_ctr, bseg_h, bseg_g, bseg_cap, d_boff,
                                 d_cnt + 5));
            SAGPU_CHECK(hipMemcpyAsync(c.pinned, d_cnt, 6 * sizeof(u64), hipMemcpyDeviceToHost, s));
            if (lazy_first) SAGPU_CHECK(hipMemcpyAsync(h_a1, d_rp, sizeof(u64), hipMemcpyDeviceToHost, s));
            return SAGPU_OK;
        };
        {
            Span sp(c, F_ROUNDS);
            if (graphs && !raw && st.rounds >= 2) {
                hipGraphExec_t &exec = c.round_exec[cur];
                // every argument and launch size the round's operations take
                const u64 sig[ROUND_SIG] = {(u64)act_i[cur], (u64)act_g[cur], (u64)act_i[nxt], (u64)act_g[nxt],
                                            (u64)key2, (u64)dISA, (u64)dSA, (u64)dLCP, (u64)d_cnt, (u64)big_head,
                                            (u64)big_g, (u64)big_size, (u64)ac.codes, (u64)d_rp, (u64)c.pinned,
                                            (u64)seg_i, (u64)seg_g, (u64)seg_ctr, (u64)ac.cb, a, n, (u64)wide,
                                            (u64)g_tile_range, big_cap, (u64)(dLCP != nullptr), (u64)bseg_h,
                                            (u64)bseg_g, bseg_cap, (u64)ac.gf, (u64)d_boff, (u64)exec};
                if (!captured[cur] && exec && !memcmp(sig, c.round_sig[cur], sizeof sig)) captured[cur] = true;
                if (!captured[cur]) {
                    // capture this parity's round once per build; the executable graph of
                    // the previous build is updated when its topology still matches
                    SAGPU_CHECK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
                    const int rc = enqueue(a, false);
                    hipGraph_t graph = nullptr;
                    const hipError_t ec = hipStreamEndCapture(s, &graph);
                    if (rc != SAGPU_OK || ec != hipSuccess) {
                        if (graph) (void)hipGraphDestroy(graph);
                        (void)hipGetLastError();
                        if (rc != SAGPU_OK) return rc;
                        sagpu_set_error("round graph capture failed", __FILE__, __LINE__);
                        return SAGPU_EDEVICE;
                    }
                    bool ok = false;
                    if (exec) {
                        hipGraphNode_t bad = nullptr;
                        hipGraphExecUpdateResult ur;
                        ok = hipGraphExecUpdate(exec, graph, &bad, &ur) == hipSuccess &&
                             ur == hipGraphExecUpdateSuccess;
                        if (!ok) {
                            (void)hipGetLastError();
                            (void)hipGraphExecDestroy(exec);
                            exec = nullptr;
                        }
                    }
                    if (!ok && hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0) != hipSuccess) {
                        (void)hipGetLastError();
                        exec = nullptr;
                        (void)hipGraphDestroy(graph);
                        sagpu_set_error("round graph instantiation failed", __FILE__, __LINE__);
                        return SAGPU_EDEVICE;
                    }
                    (void)hipGraphDestroy(graph);
                    captured[cur] = true;
                    st.round_graphs++;
                    memcpy(c.round_sig[cur], sig, sizeof sig);
                    c.round_sig[cur][ROUND_SIG - 1] = (u64)exec;   // the executable graph now serving them
                }
                SAGPU_CHECK(hipGraphLaunch(exec, s));
                st.round_graph_launches++;
            } else {
                SAGPU_TRY(enqueue(a, true));
            }
        }
        SAGPU_CHECK(hipStreamSynchronize(s));
        if (lazy_first) {
            a = *h_a1;
            st.active[0] = (int64_t)a;
            if (a == 0) {   // nothing was grouped: no round ran
                st.rounds = 0;
                break;
            }
        }
        u64 out_n = c.pinned[0];
        const u32 nb = ((u32 *)(c.pinned + 4))[0];
        if (nb > big_cap || ((u32 *)(c.pinned + 4))[1] != 0 || out_n > a) {
            sagpu_set_error("round bookkeeping inconsistent", __FILE__, __LINE__);
            return SAGPU_EDEVICE;
        }
        if (nb > 0) {
            // Big groups in the order k_big_sizes listed them (the outputs do not
            // depend on the order of the groups), with their offsets in the
            // concatenation of their members from the round's scan (d_boff) and the
            // member count among the round's counters
            const u64 total = c.pinned[5];
            if (total > a) {
                sagpu_set_error("big-group list inconsistent", __FILE__, __LINE__);
                return SAGPU_EDEVICE;
            }
            const u32 *d_bhead = big_head, *d_bslot = big_g;   // first list entries, group words
            Scope rsc(c);  // this round's big-group sort buffers
            // key2 < h + n in rank form; the first round's raw keys have rk1.kbits bits
            int kb = bits_for(h + n - 1);
            if (st.rounds == 1 && rk1.pw && rk1.kbits > kb) kb = rk1.kbits;
            const int cbits = kb + bits_for(nb - 1);
            // index and key2 in 32 bits (the few big groups of degenerate texts): u32
            // sort keys, 16 B per element and pass instead of 24
            const bool comp32 = cbits <= 32;
            // Few-key partition (rounds in rank form with many big-group members):
            // members equal to a frequent key of their group skip the sort.
            const bool pivots = !raw && g_big_pivot_min > 0 && total >= (u64)g_big_pivot_min &&
                                total < (1ull << 31) && (BP + 1) * big_part_tiles(total) <= ws.bcnt_cap;
            BigPivots bp{};
            PivRegions reg{};
            u64 nrest = total;
            u32 *pq = nullptr, *d_rk = nullptr, *d_rv = nullptr, *d_ccnt = ws.bcnt;
            // flagged groups: the rank their members hold and the key of the sorted
            // members' subgroup whose slots hold it
            u32 *d_grank = nullptr;
            u64 *d_keepkey = nullptr;
            if (ac.gf) {
                d_grank = rsc.get<u32>(nb);
                d_keepkey = rsc.get<u64>(nb);
                if (!d_grank || !d_keepkey) return SAGPU_ENOMEM;
            }
            if (pivots) {
                u64 *d_piv = rsc.get<u64>((u64)nb * BP);
                u32 *d_np = rsc.get<u32>(nb), *d_eq = rsc.get<u32>((u64)nb * BP);
                d_rk = rsc.get<u32>((u64)nb * BP);
                d_rv = rsc.get<u32>((u64)nb * BP);
                if (!d_piv || !d_np || !d_eq || !d_rk || !d_rv) return SAGPU_ENOMEM;
                const u64 ptiles = big_part_tiles(total);
                // per class (and per group, for few groups) and tile
                const u64 ncnt = (BP + 1) * (u64)(big_part_groups(nb) ? big_part_groups(nb) : 1u) * ptiles;
                if (ncnt > ws.bcnt_cap) {
                    d_ccnt = rsc.get<u32>(ncnt);
                    if (!d_ccnt) return SAGPU_ENOMEM;
                }
                {
                    Span sp(c, F_ROUNDS);
                    SAGPU_TRY(big_pivots(key2, wide, d_bhead, d_boff, nb, d_piv, d_np, s));
                    SAGPU_TRY(big_part_count(key2, wide, d_bhead, d_boff, nb, total, d_piv, d_np, d_ccnt, s));
                    SAGPU_TRY(scan_exclusive_u32(ws.scan, d_ccnt, ncnt, s));
                }
                u32 *preg = (u32 *)(c.pinned + 7200);   // past the round parameters
                u32 *d_starts = rsc.get<u32>(BP + 1);
                if (!d_starts) return SAGPU_ENOMEM;
                SAGPU_TRY(big_part_starts(d_ccnt, total, nb, d_starts, s));
                SAGPU_CHECK(hipMemcpyAsync(preg, d_starts, (BP + 1) * sizeof(u32), hipMemcpyDeviceToHost, s));
                SAGPU_CHECK(hipStreamSynchronize(s));
                for (int b = 0; b <= BP; b++) reg.start[b] = preg[b];
                for (int b = 0; b < BP; b++)
                    if (reg.start[b] > reg.start[b + 1] || reg.start[BP] > total) {
                        sagpu_set_error("pivot partition inconsistent", __FILE__, __LINE__);
                        return SAGPU_EDEVICE;
                    }
                nrest = total - reg.start[BP];
                if (!big_part_groups(nb)) {   // few groups: their counts replace the index
                    pq = rsc.get<u32>(reg.start[BP] + 1);
                    if (!pq) return SAGPU_ENOMEM;
                }
                bp = BigPivots{d_piv, d_np, d_eq};
                st.big_pivot_rounds++;
                st.big_pivot_elems += (int64_t)reg.start[BP];
            }
            u64 *b0 = rsc.get<u64>(nrest + 1), *b1 = rsc.get<u64>(nrest + 1);
            u32 *bv0 = rsc.get<u32>(nrest + 1), *bv1 = rsc.get<u32>(nrest + 1);
            if (!b0 || !b1 || !bv0 || !bv1) return SAGPU_ENOMEM;
            {
                Span sp(c, F_ROUNDS);
                // the pivot members' list entries go straight to their range of the next
                // list, behind the tile kernel's out_n survivors
                if (pivots)
                    SAGPU_TRY(big_part_scatter(act_i[cur], key2, wide, d_bhead, d_boff, nb, total, kb, comp32,
                                               bp.piv, bp.npiv, d_ccnt, reg.start[BP], b0, bv0, pq,
                                               act_i[nxt] + out_n, s));
                else
                    SAGPU_TRY(big_extract(act_i[cur], key2, wide, d_bhead, d_boff, nb, total, kb, comp32, b0, bv0, s));
            }
            alt = 0;
            if (comp32) SAGPU_TRY(sort_u32(c, (u32 *)b0, bv0, (u32 *)b1, bv1, nrest, 0, cbits, &alt));
            else SAGPU_TRY(sort_u64(c, b0, bv0, b1, bv1, nrest, 0, cbits, &alt));
            {
                Span sp(c, F_ROUNDS);
                if (pivots)
                    SAGPU_TRY(big_piv_ranks(alt ? b1 : b0, comp32, kb, nrest, pq, reg, bp.piv, bp.npiv, d_bslot, ac.gf,
                                            dSA, d_grank, d_keepkey, nb, (u32 *)bp.eq, d_rk, d_rv, out_counter,
                                            out_n, err, d_ccnt, total, s));
                else
                    SAGPU_TRY(big_keep(alt ? b1 : b0, comp32, d_boff, kb, nb, d_bslot, ac.gf, dSA, d_grank, d_keepkey,
                                       s));
                if (nrest > 0)
                    SAGPU_TRY(big_emit(ws, alt ? b1 : b0, comp32, kb, alt ? bv1 : bv0, d_bslot, nrest, h, dSA, dISA,
                                       act_i[nxt], act_g[nxt], out_counter, dLCP, s, ac, d_grank, d_keepkey,
                                       st.rounds == 1 && rk1.pw ? &ex1 : nullptr, pivots ? &bp : nullptr));
                if (pivots)
                    SAGPU_TRY(big_piv_emit(pq, reg, d_rk, d_rv, d_bslot, ac.gf, out_n, h, dISA, act_i[nxt],
                                           act_g[nxt], dLCP, ac.cb, d_ccnt, nb, total, s));
            }
            SAGPU_CHECK(hipMemcpyAsync(c.pinned, d_cnt, 6 * sizeof(u64), hipMemcpyDeviceToHost, s));
            SAGPU_CHECK(hipStreamSynchronize(s));
            out_n = c.pinned[0];
            if (((u32 *)(c.pinned + 4))[1] != 0 || out_n > a) {
                sagpu_set_error("big-group bookkeeping inconsistent", __FILE__, __LINE__);
                return SAGPU_EDEVICE;
            }
            st.big_rounds++;
            st.big_elems += (int64_t)total;
        }
        a = out_n;
        cur = nxt;
        h *= 2;
        if (h > 4 * n + 64) {
            sagpu_set_error("doubling did not converge", __FILE__, __LINE__);
            return SAGPU_EDEVICE;
        }
    }
    return SAGPU_OK;
}

// Φ of a suffix array given by the caller: bucketed two-pass scatter (launch_phi_bucketed)
// or, with the bucketed_phi option off, one unordered scatter.
int phi_run(Ctx &c, Scope &sc, const u32 *dSA, u64 n, u32 *dPHI)
{
    if (n == 0) return SAGPU_OK;
    if (!g_bucketed_phi) return launch_phi(dSA, n, dPHI, c.stream);
    GET(pv, u32, n);
    GET(pg, u32, n);
    GET(bcur, u32, 64);
    return launch_phi_bucketed(dSA, n, dPHI, pv, pg, bcur, isa_bucket_log2(n), c.stream);
}

int plcp_run(Ctx &c, Scope &sc, const uint8_t *dT, u64 n, int gsa, const u32 *PHI, const u32 *ISA, const u32 *SA,
             u32 *PLCP, const u32 *marks, int sym_bytes = 1, const uint8_t *dU = nullptr,
             const CodeStream *cs = nullptr, int hint_off = 0)
{
    LcpScratch ls{};
    ls.queue_cap = n / 16 + 4096;
    ls.tiles_cap = n / 4096 + 2;
    GET(q, u32, ls.queue_cap);
    GET(ql, u32, ls.queue_cap);
    GET(qn, u32, 64);
    GET(tm, u64, ls.tiles_cap + ls.tiles_cap / 8192 + 2);   // tile maxima, then chunk maxima
    GET(vqp, u32, 256);
    GET(vqo, u64, 256);
    GET(vqr, u64, 256);
    ls.vq_pos = vqp;
    ls.vq_off = vqo;
    ls.vq_res = vqr;
    ls.queue = q;
    ls.queue_l = ql;
    ls.queue_n = qn;
    ls.tile_max = tm;
    ls.hook = c.timer.on ? plcp_pairs_hook : nullptr;
    ls.hook_user = &c;
    c.timer.hook_n = n;
    u64 *V64 = nullptr;
    if (n > g_wide_from) {  // PLCP[i] + i may reach 2^32
        V64 = sc.get<u64>(n);
        if (!V64) return SAGPU_ENOMEM;
    }
    u64 *pairs = nullptr;
    if (marks && sym_bytes == 1 && g_plcp_pairs) {
        pairs = sc.get<u64>(n);
        if (!pairs) return SAGPU_ENOMEM;
    }
    return launch_plcp(ls, dT, n, gsa, PHI, ISA, SA, PLCP, V64, marks, c.stream, sym_bytes, pairs, dU,
                       cs && g_stream_compare ? cs->pw : nullptr, cs ? cs->bits : 0, hint_off);
}

// Pinned slots of `chunk` bytes for the staged copies; false if the mode asks for
// direct copies of this size (or pinned memory is not available).
bool xs_ready(Ctx &c, u64 bytes)
{
    if (g_host_staging == 0 || (g_host_staging == 1 && bytes < (4u << 20))) return false;
    const u64 chunk = g_host_staging == 2 ? (64u << 10) : (32u << 20);
    if (c.xs && c.xs_chunk == chunk) return true;
    if (c.xs) {
        for (int q = 0; q < XS_SLOTS; q++)
            if (c.xs_busy[q]) (void)hipEventSynchronize(c.xs_ev[q]);
        (void)hipHostFree(c.xs);
        c.xs = nullptr;
    }
    for (int q = 0; q < XS_SLOTS; q++) {
        c.xs_busy[q] = false;
        if (!c.xs_ev[q] && hipEventCreateWithFlags(&c.xs_ev[q], hipEventDisableTiming) != hipSuccess) {
            (void)hipGetLastError();
            return false;
        }
    }
    if (hipHostMalloc((void **)&c.xs, XS_SLOTS * chunk, hipHostMallocDefault) != hipSuccess) {
        (void)hipGetLastError();
        c.xs = nullptr;
        return false;
    }
    c.xs_chunk = chunk;
    return true;
}

// Host to device on the context's stream.  Returns once every byte of src has
// been read, so the caller may then overwrite it (outputs that alias inputs);
// the DMA of the last chunks may still be running.
int h2d(Ctx &c, void *dst, const void *src, u64 bytes)
{
    if (!bytes) return SAGPU_OK;
    if (!xs_ready(c, bytes)) {
        SAGPU_CHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, c.stream));
        return SAGPU_OK;
    }
    const int th = copy_threads();
    for (u64 off = 0; off < bytes; off += c.xs_chunk) {
        const int q = c.xs_next;
        c.xs_next = (q + 1) % XS_SLOTS;
        if (c.xs_busy[q]) {
            SAGPU_CHECK(hipEventSynchronize(c.xs_ev[q]));
            c.xs_busy[q] = false;
        }
        const u64 len = bytes - off < c.xs_chunk ? bytes - off : c.xs_chunk;
        uint8_t *slot = c.xs + (u64)q * c.xs_chunk;
        copies().copy(slot, (const uint8_t *)src + off, len, th);
        SAGPU_CHECK(hipMemcpyAsync((uint8_t *)dst + off, slot, len, hipMemcpyHostToDevice, c.stream));
        SAGPU_CHECK(hipEventRecord(c.xs_ev[q], c.stream));
        c.xs_busy[q] = true;
    }
    return SAGPU_OK;
}

// Device to host after the work already queued on the context's stream.  With
// staging it returns with every byte in dst; the direct form leaves the copy to
// the caller's stream synchronisation.
int d2h(Ctx &c, void *dst, const void *src, u64 bytes)
{
    if (!bytes) return SAGPU_OK;
    if (!xs_ready(c, bytes)) {
        SAGPU_CHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, c.stream));
        return SAGPU_OK;
    }
    const int th = copy_threads();
    const u64 ch = c.xs_chunk, nch = (bytes + ch - 1) / ch;
    auto issue = [&](u64 k) -> int {
        const int q = (int)(k % XS_SLOTS);
        if (c.xs_busy[q]) {
            SAGPU_CHECK(hipEventSynchronize(c.xs_ev[q]));
            c.xs_busy[q] = false;
        }
        const u64 len = bytes - k * ch < ch ? bytes - k * ch : ch;
        SAGPU_CHECK(hipMemcpyAsync(c.xs + (u64)q * ch, (const uint8_t *)src + k * ch, len, hipMemcpyDeviceToHost,
                                   c.stream));
        SAGPU_CHECK(hipEventRecord(c.xs_ev[q], c.stream));
        c.xs_busy[q] = true;
        return SAGPU_OK;
    };
    for (u64 k = 0; k < nch && k < (u64)XS_SLOTS; k++) SAGPU_TRY(issue(k));
    for (u64 k = 0; k < nch; k++) {
        const int q = (int)(k % XS_SLOTS);
        SAGPU_CHECK(hipEventSynchronize(c.xs_ev[q]));
        c.xs_busy[q] = false;
        const u64 len = bytes - k * ch < ch ? bytes - k * ch : ch;
        copies().copy((uint8_t *)dst + k * ch, c.xs + (u64)q * ch, len, th);
        if (k + XS_SLOTS < nch) SAGPU_TRY(issue(k + XS_SLOTS));
    }
    return SAGPU_OK;
}

// Where a text comes from: device memory, caller host memory, or pinned memory
// of the library (copied directly).
enum { SRC_DEVICE = 0, SRC_HOST = 1, SRC_PINNED = 2 };

// Copies a text into a pool buffer with 64 zero bytes of padding.
int padded_text(Ctx &c, Scope &sc, const uint8_t *src, u64 n, int src_kind, uint8_t **out)
{
    GET(dT, uint8_t, n + 64);
    SAGPU_TRY(launch_fill_u8(dT + n, 64, 0, c.stream));
    if (src_kind == SRC_HOST) {
        SAGPU_TRY(h2d(c, dT, src, n));
    } else if (n) {
        SAGPU_CHECK(hipMemcpyAsync(dT, src, n, src_kind == SRC_DEVICE ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice,
                                   c.stream));
    }
    *out = dT;
    return SAGPU_OK;
}

void stats_reset(Ctx &c)
{
    c.stats = sagpu_stats{};
    c.timer.on = g_timing;
    c.stats_fresh = true;
}

// Fused device pipeline.  Returns 0 or the primary index when U is requested.
int64_t build_dev(Ctx &c, Scope &sc, const uint8_t *dTpad, u64 n, int gsa, u32 *dSA, uint8_t *dU, u32 *dPLCP,
                  u32 *dLCP, u64 *freq, u32 *dISA = nullptr)
{
    if (!dSA) {
        dSA = sc.get<u32>(n);
        if (!dSA) return SAGPU_ENOMEM;
    }
    const bool own_isa = !dISA;
    if (!dISA) {
        dISA = sc.get<u32>(n);
        if (!dISA) return SAGPU_ENOMEM;
    }
    const bool want_lcp = dPLCP || dLCP;
    u32 *marks = dLCP;
    if (want_lcp && !marks) {
        marks = sc.get<u32>(n);
        if (!marks) return SAGPU_ENOMEM;
    }
    if (want_lcp && !dPLCP) {
        dPLCP = sc.get<u32>(n);
        if (!dPLCP) return SAGPU_ENOMEM;
    }
    int rc;
    {
        Span total(c, F_TOTAL);
        Scope work(c);  // construction buffers go back to the pool before the caller's downloads
        BwtSeed seed{};
        CodeStream cs{};
        // the PLCP hints ride in the ISA words only when the ISA stays inside this call
        PlcpHint hint;
        hint.on = want_lcp && own_isa && g_plcp_hint;
        rc = sa_build(c, work, dTpad, n, gsa, dSA, dISA, freq, want_lcp ? marks : nullptr, nullptr, 0, nullptr, 0,
                      dU && g_bwt_variant == 0 ? &seed : nullptr, &cs, &hint);
        if (rc != SAGPU_OK) return rc;
        if (dU) {
            Span sp(c, F_BWT);
            if (seed.cb) {
                ktime_begin(KI_BWT);
                SAGPU_TRY(launch_bwt_seeded(dTpad, dSA, dISA, seed, n, dU, c.stream));
                // seed byte and bitmap bit read, BWT byte written
                ktime_end(KI_BWT, 2 * n + n / 8);
            }
            else if (g_bwt_variant == 1) SAGPU_TRY(launch_bwt_scatter(dTpad, dISA, n, dU, c.stream));
            else SAGPU_TRY(launch_bwt(dTpad, dSA, dISA, n, dU, c.stream));
        }
        if (want_lcp) {
            {
                Span sp(c, F_PLCP);
                SAGPU_TRY(plcp_run(c, work, dTpad, n, gsa, nullptr, dISA, dSA, dPLCP, marks, 1, dU,
                                   cs.pw ? &cs : nullptr, hint.on ? hint.off : 0));
            }
            if (dLCP) {
                Span sp(c, F_LCP);
                ktime_begin(KI_LCP_FIX);
                SAGPU_TRY(launch_lcp_fix(dPLCP, dSA, n, dLCP, c.stream));
                ktime_end(KI_LCP_FIX, 12 * n);   // mark and SA read, LCP written
            }
        }
    }
    int64_t ret = 0;
    if (dU) {
        u32 j0 = 0;
        SAGPU_TRY(read_u32(c, dISA, &j0));
        ret = (int64_t)isa_rank(j0, n) + 1;
    }
    SAGPU_CHECK(hipStreamSynchronize(c.stream));
    timer_collect(c);
    return ret;
}

__global__ void k_widen(const u32 *in, int64_t *out, u64 n)
{
    const u64 stride = (u64)gridDim.x * blockDim.x;
    for (u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) out[i] = (int64_t)in[i];
}

__global__ void k_narrow(const int64_t *in, u32 *out, u64 n)
{
    const u64 stride = (u64)gridDim.x * blockDim.x;
    for (u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) out[i] = (u32)in[i];
}

unsigned grid_n(u64 n)
{
    u64 b = (n + 2047) / 2048;
    return (unsigned)(b < 1 ? 1 : b > 65535 ? 65535 : b);
}

// Download u32 device array into an int32 or int64 host array.
int download_idx(Ctx &c, Scope &sc, const u32 *d, void *h, u64 n, int width)
{
    if (width == 4) {
        SAGPU_TRY(d2h(c, h, d, n * 4));
    } else {
        GET(w, int64_t, n);
        hipLaunchKernelGGL(k_widen, dim3(grid_n(n)), dim3(256), 0, c.stream, d, w, n);
        SAGPU_TRY(d2h(c, h, w, n * 8));
    }
    return SAGPU_OK;
}

int upload_idx(Ctx &c, Scope &sc, const void *h, u64 n, int width, u32 **out)
{
    GET(d, u32, n);
    if (width == 4) {
        SAGPU_TRY(h2d(c, d, h, n * 4));
    } else {
        GET(w, int64_t, n);
        SAGPU_TRY(h2d(c, w, h, n * 8));
        hipLaunchKernelGGL(k_narrow, dim3(grid_n(n)), dim3(256), 0, c.stream, w, d, n);
    }
    *out = d;
    return SAGPU_OK;
}

void store_freq(const u64 *f, void *freq, int width)
{
    if (!freq) return;
    for (int ch = 0; ch < 256; ch++) {
        if (width == 4) ((int32_t *)freq)[ch] = (int32_t)f[ch];
        else ((int64_t *)freq)[ch] = (int64_t)f[ch];
    }
}

} // namespace

// ======================================================== libsais contexts

extern "C" SAGPU_HIDDEN void *sagpu_host_token_new(void) { return new (std::nothrow) Token(); }

extern "C" SAGPU_HIDDEN void sagpu_host_token_free(void *token)
{
    if (!token) return;
    Token *t = (Token *)token;
    {
        std::lock_guard<std::mutex> lk(t->mu);   // waits for a call still running on it
        ctx_park(t->c);
        t->c = nullptr;
        t->magic = 0;
    }
    delete t;
}

extern "C" SAGPU_HIDDEN void *sagpu_host_bind(const void *token)
{
    Token *prev = t_bound;
    t_bound = (Token *)token;
    return prev;
}

// ======================================================== host-buffer layer
// Used by the libsais-compatible wrappers (api.c).  Arguments are already
// validated there; n >= 2.  width = 4 (int32 arrays) or 8 (int64 arrays).

extern "C" SAGPU_HIDDEN int64_t sagpu_host_sa(const uint8_t *T, void *SA, int64_t n, int32_t gsa, void *freq,
                                              int32_t width)
{
    LEASE(c);
    if ((u64)n > (1ull << 32)) {
        sagpu_set_error("n > 2^32 is not supported (indices are kept as 32-bit words)", __FILE__, __LINE__);
        return SAGPU_ENOMEM;
    }
    stats_reset(*c);
    int64_t rc;
    {
        Scope sc(*c);
        uint8_t *dT;
        SAGPU_TRY(padded_text(*c, sc, T, (u64)n, SRC_HOST, &dT));
        u32 *dSA = sc.get<u32>((u64)n);
        if (!dSA) return SAGPU_ENOMEM;
        u64 f[256];
        rc = build_dev(*c, sc, dT, (u64)n, gsa, dSA, nullptr, nullptr, nullptr, f);
        if (rc < 0) return rc;
        SAGPU_TRY(download_idx(*c, sc, dSA, SA, (u64)n, width));
        SAGPU_CHECK(hipStreamSynchronize(c->stream));
        store_freq(f, freq, width);
    }
    return 0;
}

extern "C" SAGPU_HIDDEN int64_t sagpu_host_bwt(const uint8_t *T, uint8_t *U, int64_t n, void *freq, int32_t width)
{
    LEASE(c);
    if ((u64)n > (1ull << 32)) {
        sagpu_set_error("n > 2^32 is not supported (indices are kept as 32-bit words)", __FILE__, __LINE__);
        return SAGPU_ENOMEM;
    }
    stats_reset(*c);
    Scope sc(*c);
    uint8_t *dT;
    SAGPU_TRY(padded_text(*c, sc, T, (u64)n, SRC_HOST, &dT));
    uint8_t *dU = sc.get<uint8_t>((u64)n);
    if (!dU) return SAGPU_ENOMEM;
    u64 f[256];
    int64_t primary = build_dev(*c, sc, dT, (u64)n, 0, nullptr, dU, nullptr, nullptr, f);
    if (primary < 0) return primary;
    SAGPU_TRY(d2h(*c, U, dU, (u64)n));
    SAGPU_CHECK(hipStreamSynchronize(c->stream));
    store_freq(f, freq, width);
    return primary;
}

// Integer alphabet (libsais_int, libsais64_long): T holds int32 (width 4) or int64
// (width 8) symbols, expected in [0, k).  Symbols below 2^16 are packed as they
// are; wider ones are replaced by their dense ranks first (one sort by value), so
// the initial keys still hold as many symbols as the alphabet allows.
extern "C" SAGPU_HIDDEN int64_t sagpu_host_sa_int(const void *T, void *SA, int64_t n, int32_t width)
{
    LEASE(c);
    if ((u64)n > (1ull << 32)) {
        sagpu_set_error("n > 2^32 is not supported (indices are kept as 32-bit words)", __FILE__, __LINE__);
        return SAGPU_ENOMEM;
    }
    stats_reset(*c);
    const u64 N = (u64)n;
    hipStream_t s = c->stream;
    Scope sc(*c);
    GET(dSym, u32, N);
    int bits = 1;
    {
        Scope in(*c);  // the uploaded symbols and ranking buffers go back before the build
        uint8_t *dTin = in.get<uint8_t>(N * (u64)width);
        u64 *d_max = in.get<u64>(1);
        if (!dTin || !d_max) return SAGPU_ENOMEM;
        SAGPU_TRY(h2d(*c, dTin, T, N * (u64)width));
        SAGPU_TRY(launch_fill_u64(d_max, 1, 0, s));
        SAGPU_TRY(launch_max_sym(dTin, width, N, d_max, s));
        SAGPU_CHECK(hipMemcpyAsync(c->pinned, d_max, 8, hipMemcpyDeviceToHost, s));
        SAGPU_CHECK(hipStreamSynchronize(s));
        u64 maxv = c->pinned[0];
        if (maxv < (1ull << 16)) {
            SAGPU_TRY(launch_sym_narrow(dTin, width, N, dSym, s));
            bits = maxv ? bits_for(maxv) : 1;
        } else {
            u64 *k0 = in.get<u64>(N), *k1 = in.get<u64>(N);
            u32 *v0 = in.get<u32>(N), *v1 = in.get<u32>(N), *flags = in.get<u32>(N);
            const u64 pcap = N / 4096 + 2;
            u32 *part = in.get<u32>(pcap);
            if (!k0 || !k1 || !v0 || !v1 || !flags || !part) return SAGPU_ENOMEM;
            SAGPU_TRY(launch_sym_keys64(dTin, width, N, k0, v0, s));
            SAGPU_TRY(ensure_radix_scratch(*c, in, N));
            int alt = 0;
            SAGPU_TRY(sort_u64(*c, k0, v0, k1, v1, N, 0, bits_for(maxv), &alt));
            const u64 *sk = alt ? k1 : k0;
            const u32 *sv = alt ? v1 : v0;
            SAGPU_TRY(launch_run_flags64(sk, N, flags, s));
            ScanScratch ssc{part, pcap};
            SAGPU_TRY(scan_exclusive_u32(ssc, flags, N, s));
            SAGPU_TRY(launch_dense_scatter64(sk, sv, flags, N, dSym, s));
            SAGPU_TRY(launch_fill_u64(d_max, 1, 0, s));
            SAGPU_TRY(launch_max_sym(dSym, 4, N, d_max, s));
            SAGPU_CHECK(hipMemcpyAsync(c->pinned, d_max, 8, hipMemcpyDeviceToHost, s));
            SAGPU_CHECK(hipStreamSynchronize(s));
            const u64 maxrank = c->pinned[0];
            bits = maxrank ? bits_for(maxrank) : 1;
        }
    }
    GET(dSA, u32, N);
    GET(dISA, u32, N);
    {
        Scope work(*c);
        SAGPU_TRY(sa_build(*c, work, nullptr, N, 0, dSA, dISA, nullptr, nullptr, nullptr, 0, dSym, bits));
    }
    SAGPU_TRY(download_idx(*c, sc, dSA, SA, N, width));
    SAGPU_CHECK(hipStreamSynchronize(s));
    return 0;
}

// PLCP of an integer string (libsais_plcp_int): int32 symbols compared 4 bytes at a
// time by the byte kernels.
extern "C" SAGPU_HIDDEN int64_t sagpu_host_plcp_int(const int32_t *T, const void *SA, void *PLCP, int64_t n,
                                                    int32_t width)
{
    LEASE(c);
    if ((u64)n > (1ull << 32)) {
        sagpu_set_error("n > 2^32 is not supported (indices are kept as 32-bit words)", __FILE__, __LINE__);
        return SAGPU_ENOMEM;
    }
    stats_reset(*c);
    Scope sc(*c);
    uint8_t *dT;
    SAGPU_TRY(padded_text(*c, sc, (const uint8_t *)T, 4 * (u64)n, SRC_HOST, &dT));
    u32 *dSA;
    SAGPU_TRY(upload_idx(*c, sc, SA, (u64)n, width, &dSA));
    u32 *dPHI = sc.get<u32>((u64)n);
    u32 *dP = sc.get<u32>((u64)n);
    if (!dPHI || !dP) return SAGPU_ENOMEM;
    SAGPU_TRY(phi_run(*c, sc, dSA, (u64)n, dPHI));
    SAGPU_TRY(plcp_run(*c, sc, dT, (u64)n, 0, dPHI, nullptr, nullptr, dP, nullptr, 4));
    SAGPU_TRY(download_idx(*c, sc, dP, PLCP, (u64)n, width));
    SAGPU_CHECK(hipStreamSynchronize(c->stream));
    return 0;
}

// BWT with auxiliary indexes: I[k] = 1 + SA slot of suffix k*r (libsais_bwt_aux).
extern "C" SAGPU_HIDDEN int64_t sagpu_host_bwt_aux(const uint8_t *T, uint8_t *U, int64_t n, void *freq, int64_t r,
                                                   void *I, int32_t width)
{
    LEASE(c);
    if ((u64)n > (1ull << 32)) {
        sagpu_set_error("n > 2^32 is not supported (indices are kept as 32-bit words)", __FILE__, __LINE__);
        return SAGPU_ENOMEM;
    }
    stats_reset(*c);
    Scope sc(*c);
    uint8_t *dT;
    SAGPU_TRY(padded_text(*c, sc, T, (u64)n, SRC_HOST, &dT));
    uint8_t *dU = sc.get<uint8_t>((u64)n);
    u32 *dISA = sc.get<u32>((u64)n);
    const u64 count = ((u64)n - 1) / (u64)r + 1;
    u32 *dS = sc.get<u32>(count);
    if (!dU || !dISA || !dS) return SAGPU_ENOMEM;
    u64 f[256];
    int64_t primary = build_dev(*c, sc, dT, (u64)n, 0, nullptr, dU, nullptr, nullptr, f, dISA);
    if (primary < 0) return primary;
    SAGPU_TRY(launch_aux_samples(dISA, (u64)n, (u64)r, count, dS, c->stream));
    std::vector<u32> hs(count);
    SAGPU_CHECK(hipMemcpyAsync(hs.data(), dS, count * 4, hipMemcpyDeviceToHost, c->stream));
    SAGPU_TRY(d2h(*c, U, dU, (u64)n));
    SAGPU_CHECK(hipStreamSynchronize(c->stream));
    for (u64 k = 0; k < count; k++) {
        if (width == 4) ((int32_t *)I)[k] = (int32_t)(hs[k] + 1);
        else ((int64_t *)I)[k] = (int64_t)hs[k] + 1;
    }
    store_freq(f, freq, width);
    return 0;
}

// Inverse BWT (libsais_unbwt / libsais_unbwt_aux and the 16-bit forms).  T = the
// BWT (sym_bytes per symbol), U = the output text, I[0..(n-1)/r] already
// range-checked by the caller, r == n for plain unbwt.
int unbwt_host(Ctx &c, const void *T, void *U, u64 N, u64 r, const void *I, int width, int sym_bytes)
{
    const u64 chunks = (N - 1) / r + 1;
    std::vector<u64> rows(chunks);
    for (u64 k = 0; k < chunks; k++)
        rows[k] = width == 4 ? (u64)((const int32_t *)I)[k] : (u64)((const int64_t *)I)[k];
    Scope sc(c);
    hipStream_t s = c.stream;
    const u64 bytes = N * (u64)sym_bytes;
    GET(dB, uint8_t, bytes);
    GET(dOut, uint8_t, bytes);
    GET(S, u32, N);
    GET(d_err, u32, 64);
    SAGPU_TRY(h2d(c, dB, T, bytes));
    SAGPU_CHECK(hipMemsetAsync(dOut, 0, bytes, s));
    SAGPU_TRY(launch_fill_u32(d_err, 2, 0, s));
    const u64 pcap = N / 4096 + 2;
    if (sym_bytes == 1) {
        const u64 slots = unbwt_count_slots(N);
        GET(counts, u32, slots);
        const u64 scap = slots / 4096 + 2;
        GET(part, u32, scap);
        ScanScratch ssc{part, scap};
        SAGPU_TRY(launch_unbwt_succ(dB, N, rows[0], counts, ssc, S, s));
    } else {
        // 16-bit symbols: stable sort of (symbol, position); slot k holds LF row k + 1
        Scope rs(c);
        u32 *k0 = rs.get<u32>(N), *k1 = rs.get<u32>(N), *v0 = rs.get<u32>(N), *v1 = rs.get<u32>(N);
        if (!k0 || !k1 || !v0 || !v1) return SAGPU_ENOMEM;
        SAGPU_TRY(launch_unbwt_succ16((const uint16_t *)dB, N, rows[0], k0, v0, S, s));
        SAGPU_TRY(ensure_radix_scratch(c, rs, N));
        int alt = 0;
        SAGPU_TRY(sort_u32(c, k0, v0, k1, v1, N, 0, 16, &alt));
        SAGPU_TRY(launch_unbwt_succ_sorted(alt ? v1 : v0, N, rows[0], S, s));
        SAGPU_CHECK(hipStreamSynchronize(s));
    }
    if (r < N && chunks >= 32768) {
        GET(d_rows, u64, chunks);
        SAGPU_CHECK(hipMemcpyAsync(d_rows, rows.data(), chunks * 8, hipMemcpyHostToDevice, s));
        SAGPU_TRY(launch_unbwt_aux_walk(dB, sym_bytes, S, N, rows[0], r, d_rows, chunks, dOut, d_err, s));
    } else {
        GET(part, u32, pcap);
        ScanScratch ssc{part, pcap};
        GET(sid, u32, N);
        SAGPU_TRY(launch_unbwt_flags(N, sid, s));
        SAGPU_TRY(scan_exclusive_u32(ssc, sid, N, s));
        SAGPU_TRY(launch_unbwt_sample_total(sid, N, d_err + 1, s));
        u32 m = 0;
        SAGPU_TRY(read_u32(c, d_err + 1, &m));
        if (m == 0 || m > N) {
            sagpu_set_error("inverse BWT: sample count inconsistent", __FILE__, __LINE__);
            return SAGPU_EDEVICE;
        }
        GET(samp_q, u32, m);
        GET(nxt0, u32, m);
        GET(nxt1, u32, m);
        GET(dist0, u64, m);
        GET(dist1, u64, m);
        GET(gap, u64, m);
        u32 *nxt[2] = {nxt0, nxt1};
        u64 *dist[2] = {dist0, dist1};
        SAGPU_TRY(launch_unbwt_rank(dB, sym_bytes, S, sid, N, m, samp_q, nxt, dist, gap, dOut, d_err, s));
    }
    u32 err = 0;
    SAGPU_TRY(read_u32(c, d_err, &err));
    if (err & 2u) {
        sagpu_set_error("inverse BWT: a sample gap exceeded its step limit", __FILE__, __LINE__);
        return SAGPU_EDEVICE;
    }
    SAGPU_TRY(d2h(c, U, dOut, bytes));
    SAGPU_CHECK(hipStreamSynchronize(s));
    return 0;
}

extern "C" SAGPU_HIDDEN int64_t sagpu_host_unbwt(const uint8_t *T, uint8_t *U, int64_t n, int64_t r, const void *I,
                                                 int32_t width)
{
    LEASE(c);
    if ((u64)n >= (1ull << 32)) {
        sagpu_set_error("inverse BWT: n >= 2^32 is not supported", __FILE__, __LINE__);
        return SAGPU_ENOMEM;
    }
    stats_reset(*c);
    return unbwt_host(*c, T, U, (u64)n, (u64)r, I, width, 1);
}

// ------------------------------------------------------------ 16-bit alphabet
// (libsais16 / libsais16x64): the 65536-bin histogram gives freq, the GSA check and
// dense symbol codes; from there the integer-mode pipeline builds SA and ISA.
int sa16_build(Ctx &c, Scope &sc, const uint16_t *dT, u64 n, int gsa, u32 *dSA, u32 *dISA, u64 *freq)
{
    hipStream_t s = c.stream;
    GET(d_hist, u64, 65536 + 8);
    u32 *d_bad = (u32 *)(d_hist + 65536);
    SAGPU_TRY(launch_fill_u64(d_hist, 65536 + 8, 0, s));
    SAGPU_TRY(launch_hist16(dT, n, gsa, d_hist, d_bad, s));
    std::vector<u64> h(65536 + 1);
    SAGPU_CHECK(hipMemcpyAsync(h.data(), d_hist, (65536 + 1) * sizeof(u64), hipMemcpyDeviceToHost, s));
    SAGPU_CHECK(hipStreamSynchronize(s));
    if (freq) memcpy(freq, h.data(), 65536 * sizeof(u64));
    if (gsa && (u32)h[65536] != 0) return SAGPU_EINVAL;
    std::vector<u32> lut(65536);
    u32 sigma = 0;
    for (u32 v = 0; v < 65536; v++) {
        lut[v] = sigma;
        if (h[v]) sigma++;
    }
    const int bits = sigma <= 2 ? 1 : bits_for((u64)sigma - 1);
    GET(d_lut, u32, 65536);
    SAGPU_CHECK(hipMemcpyAsync(d_lut, lut.data(), 65536 * 4, hipMemcpyHostToDevice, s));
    GET(dSym, u32, n);
    SAGPU_TRY(launch_map16(dT, n, d_lut, dSym, s));
    Scope work(c);
    SAGPU_TRY(sa_build(c, work, nullptr, n, gsa, dSA, dISA, nullptr, nullptr, nullptr, 0, dSym, bits));
    SAGPU_CHECK(hipStreamSynchronize(s));
    return SAGPU_OK;
}

void store_freq16(const u64 *f, void *freq, int width)
{
    if (!freq) return;
    for (u32 v = 0; v < 65536; v++) {
        if (width == 4) ((int32_t *)freq)[v] = (int32_t)f[v];
        else ((int64_t *)freq)[v] = (int64_t)f[v];
    }
}

extern "C" SAGPU_HIDDEN int64_t sagpu_host_sa16(const uint16_t *T, void *SA, int64_t n, int32_t gsa, void *freq,
                                                int32_t width)
{
    LEASE(c);
    if ((u64)n > (1ull << 32)) {
        sagpu_set_error("n > 2^32 is not supported (indices are kept as 32-bit words)", __FILE__, __LINE__);
        return SAGPU_ENOMEM;
    }
    stats_reset(*c);
    const u64 N = (u64)n;
    Scope sc(*c);
    GET(dT, uint16_t, N);
    SAGPU_TRY(h2d(*c, dT, T, 2 * N));
    GET(dSA, u32, N);
    GET(dISA, u32, N);
    std::vector<u64> f(65536);
    SAGPU_TRY(sa16_build(*c, sc, dT, N, gsa, dSA, dISA, f.data()));
    SAGPU_TRY(download_idx(*c, sc, dSA, SA, N, width));
    SAGPU_CHECK(hipStreamSynchronize(c->stream));
    store_freq16(f.data(), freq, width);
    return 0;
}

// BWT of 16-bit symbols; r > 0 also writes the auxiliary indexes I (libsais16_bwt_aux)
// and returns 0, r == 0 returns the primary index (libsais16_bwt).
extern "C" SAGPU_HIDDEN int64_t sagpu_host_bwt16(const uint16_t *T, uint16_t *U, int64_t n, void *freq, int64_t r,
                                                 void *I, int32_t width)
{
    LEASE(c);
    if ((u64)n > (1ull << 32)) {
        sagpu_set_error("n > 2^32 is not supported (indices are kept as 32-bit words)", __FILE__, __LINE__);
        return SAGPU_ENOMEM;
    }
    stats_reset(*c);
    const u64 N = (u64)n;
    hipStream_t s = c->stream;
    Scope sc(*c);
    GET(dT, uint16_t, N);
    SAGPU_TRY(h2d(*c, dT, T, 2 * N));
    GET(dSA, u32, N);
    GET(dISA, u32, N);
    GET(dU, uint16_t, N);
    std::vector<u64> f(65536);
    SAGPU_TRY(sa16_build(*c, sc, dT, N, 0, dSA, dISA, f.data()));
    SAGPU_TRY(launch_bwt16(dT, dSA, dISA, N, dU, s));
    u32 j0 = 0;
    SAGPU_TRY(read_u32(*c, dISA, &j0));
    j0 = isa_rank(j0, N);
    if (r > 0) {
        const u64 count = (N - 1) / (u64)r + 1;
        GET(dS, u32, count);
        SAGPU_TRY(launch_aux_samples(dISA, N, (u64)r, count, dS, s));
        std::vector<u32> hs(count);
        SAGPU_CHECK(hipMemcpyAsync(hs.data(), dS, count * 4, hipMemcpyDeviceToHost, s));
        SAGPU_CHECK(hipStreamSynchronize(s));
        for (u64 k = 0; k < count; k++) {
            if (width == 4) ((int32_t *)I)[k] = (int32_t)(hs[k] + 1);
            else ((int64_t *)I)[k] = (int64_t)hs[k] + 1;
        }
    }
    SAGPU_TRY(d2h(*c, U, dU, 2 * N));
    SAGPU_CHECK(hipStreamSynchronize(s));
    store_freq16(f.data(), freq, width);
    return r > 0 ? 0 : (int64_t)j0 + 1;
}

extern "C" SAGPU_HIDDEN int64_t sagpu_host_plcp16(const uint16_t *T, const void *SA, void *PLCP, int64_t n, int32_t gsa,
                                                  int32_t width)
{
    LEASE(c);
    if ((u64)n > (1ull << 32)) {
        sagpu_set_error("n > 2^32 is not supported (indices are kept as 32-bit words)", __FILE__, __LINE__);
        return SAGPU_ENOMEM;
    }
    stats_reset(*c);
    Scope sc(*c);
    uint8_t *dT;
    SAGPU_TRY(padded_text(*c, sc, (const uint8_t *)T, 2 * (u64)n, SRC_HOST, &dT));
    u32 *dSA;
    SAGPU_TRY(upload_idx(*c, sc, SA, (u64)n, width, &dSA));
    u32 *dPHI = sc.get<u32>((u64)n);
    u32 *dP = sc.get<u32>((u64)n);
    if (!dPHI || !dP) return SAGPU_ENOMEM;
    SAGPU_TRY(phi_run(*c, sc, dSA, (u64)n, dPHI));
    SAGPU_TRY(plcp_run(*c, sc, dT, (u64)n, gsa, dPHI, nullptr, nullptr, dP, nullptr, 2));
    SAGPU_TRY(download_idx(*c, sc, dP, PLCP, (u64)n, width));
    SAGPU_CHECK(hipStreamSynchronize(c->stream));
    return 0;
}

extern "C" SAGPU_HIDDEN int64_t sagpu_host_unbwt16(const uint16_t *T, uint16_t *U, int64_t n, int64_t r,
                                                   const void *I, int32_t width)
{
    LEASE(c);
    if ((u64)n >= (1ull << 32)) {
        sagpu_set_error("inverse BWT: n >= 2^32 is not supported", __FILE__, __LINE__);
        return SAGPU_ENOMEM;
    }
    stats_reset(*c);
    return unbwt_host(*c, T, U, (u64)n, (u64)r, I, width, 2);
}

extern "C" SAGPU_HIDDEN int64_t sagpu_host_plcp(const uint8_t *T, const void *SA, void *PLCP, int64_t n, int32_t gsa,
                                                int32_t width)
{
    LEASE(c);
    if ((u64)n > (1ull << 32)) {
        sagpu_set_error("n > 2^32 is not supported (indices are kept as 32-bit words)", __FILE__, __LINE__);
        return SAGPU_ENOMEM;
    }
    stats_reset(*c);
    Scope sc(*c);
    uint8_t *dT;
    SAGPU_TRY(padded_text(*c, sc, T, (u64)n, SRC_HOST, &dT));
    u32 *dSA;
    SAGPU_TRY(upload_idx(*c, sc, SA, (u64)n, width, &dSA));
    u32 *dPHI = sc.get<u32>((u64)n);
    u32 *dP = sc.get<u32>((u64)n);
    if (!dPHI || !dP) return SAGPU_ENOMEM;
    SAGPU_TRY(phi_run(*c, sc, dSA, (u64)n, dPHI));
    SAGPU_TRY(plcp_run(*c, sc, dT, (u64)n, gsa, dPHI, nullptr, nullptr, dP, nullptr));
    SAGPU_TRY(download_idx(*c, sc, dP, PLCP, (u64)n, width));
    SAGPU_CHECK(hipStreamSynchronize(c->stream));
    return 0;
}

extern "C" SAGPU_HIDDEN int64_t sagpu_host_lcp(const void *PLCP, const void *SA, void *LCP, int64_t n, int32_t width)
{
    LEASE(c);
    if ((u64)n > (1ull << 32)) {
        sagpu_set_error("n > 2^32 is not supported (indices are kept as 32-bit words)", __FILE__, __LINE__);
        return SAGPU_ENOMEM;
    }
    Scope sc(*c);
    u32 *dSA, *dP;
    SAGPU_TRY(upload_idx(*c, sc, SA, (u64)n, width, &dSA));
    SAGPU_TRY(upload_idx(*c, sc, PLCP, (u64)n, width, &dP));
    u32 *dL = sc.get<u32>((u64)n);
    if (!dL) return SAGPU_ENOMEM;
    SAGPU_TRY(launch_lcp(dP, dSA, (u64)n, dL, c->stream));
    SAGPU_TRY(download_idx(*c, sc, dL, LCP, (u64)n, width));
    SAGPU_CHECK(hipStreamSynchronize(c->stream));
    return 0;
}

// ======================================================== public device ABI

namespace {
int noalloc_init_current();
} // namespace

// Creates the first slot context (otherwise created on first use) and sets up the
// allocation-free build on the current device.
extern "C" int32_t sagpu_init(void)
{
    {
        LEASE(c);
        (void)c;
    }
    return noalloc_init_current();
}

// Destroys the slot contexts (after the calls that hold them end).  Contexts of
// live libsais context handles and servers stay.
extern "C" void sagpu_shutdown(void)
{
    {
        std::unique_lock<std::mutex> lk(g_slot_mu);
        g_slot_cv.wait(lk, [] {
            for (char b : g_slot_busy)
                if (b) return false;
            return true;
        });
        for (Ctx *c : g_slots) ctx_destroy(c);
        g_slots.clear();
        g_slot_busy.clear();
        for (Ctx *c : g_parked) ctx_destroy(c);
        g_parked.clear();
    }
}

extern "C" const char *sagpu_last_error(void) { return g_err.c_str(); }

extern "C" void *sagpu_malloc(size_t bytes)
{
    void *p = nullptr;
    if (hipMalloc(&p, bytes ? bytes : 1) != hipSuccess) {
        (void)hipGetLastError();
        return nullptr;
    }
    return p;
}

extern "C" void sagpu_free(void *p)
{
    if (p) (void)hipFree(p);
}

extern "C" int32_t sagpu_memcpy_h2d(void *d, const void *h, size_t bytes)
{
    return hipMemcpy(d, h, bytes, hipMemcpyHostToDevice) == hipSuccess ? 0 : SAGPU_EDEVICE;
}

extern "C" int32_t sagpu_memcpy_d2h(void *h, const void *d, size_t bytes)
{
    return hipMemcpy(h, d, bytes, hipMemcpyDeviceToHost) == hipSuccess ? 0 : SAGPU_EDEVICE;
}

extern "C" int32_t sagpu_host_memcpy(void *h, const void *s, size_t bytes, int32_t threads)
{
    if (threads < 0) return SAGPU_EINVAL;
    if (bytes) copies().copy(h, s, bytes, threads ? threads : copy_threads());
    return 0;
}

extern "C" int32_t sagpu_memset(void *d, int value, size_t bytes)
{
    return hipMemset(d, value, bytes) == hipSuccess ? 0 : SAGPU_EDEVICE;
}

extern "C" int32_t sagpu_memcpy_d2d(void *d_dst, const void *d_src, size_t bytes)
{
    if ((!d_dst || !d_src) && bytes) return SAGPU_EINVAL;
    if (!bytes) return 0;
    LEASE(c);
    SAGPU_TRY(launch_copy(d_dst, d_src, bytes, c->stream));
    SAGPU_CHECK(hipStreamSynchronize(c->stream));
    return 0;
}

extern "C" int32_t sagpu_synchronize(void)
{
    return hipDeviceSynchronize() == hipSuccess ? 0 : SAGPU_EDEVICE;
}

extern "C" void sagpu_set_memory_limit(uint64_t bytes) { g_mem_limit.store(bytes); }

extern "C" void sagpu_trim(void) { trim_idle_slots(); }

extern "C" void sagpu_set_timing(int32_t enabled) { g_timing = enabled != 0; }

extern "C" int32_t sagpu_set_option(const char *name, int64_t value)
{
    if (!name) return SAGPU_EINVAL;
    std::lock_guard<std::mutex> lk(g_opt_mu);
    if (!strcmp(name, "radix_items")) {
        if (value != 8 && value != 16 && value != 32) return SAGPU_EINVAL;
        g_radix_items = (int)value;
        return 0;
    }
    if (!strcmp(name, "radix_threads")) {
        if (value != 256 && value != 512) return SAGPU_EINVAL;
        g_radix_threads = (int)value;
        return 0;
    }
    if (!strcmp(name, "radix_variant")) {
        if (value != RADIX_ONESWEEP && value != RADIX_REDUCE_THEN_SCAN) return SAGPU_EINVAL;
        g_radix_variant = (int)value;
        return 0;
    }
    if (!strcmp(name, "api_slots")) {
        if (value < 1 || value > 16) return SAGPU_EINVAL;
        g_api_slots = (int)value;
        return 0;
    }
    if (!strcmp(name, "host_staging")) {
        if (value < 0 || value > 2) return SAGPU_EINVAL;
        g_host_staging = (int)value;
        return 0;
    }
    if (!strcmp(name, "copy_threads")) {
        if (value < 0 || value > 64) return SAGPU_EINVAL;
        g_copy_threads = (int)value;
        return 0;
    }
    if (!strcmp(name, "bucketed_phi")) {
        if (value != 0 && value != 1) return SAGPU_EINVAL;
        g_bucketed_phi = (int)value;
        return 0;
    }
    if (!strcmp(name, "plcp_hint")) {
        if (value != 0 && value != 1) return SAGPU_EINVAL;
        g_plcp_hint = (int)value;
        return 0;
    }
    if (!strcmp(name, "exact_round1")) {
        if (value != 0 && value != 1) return SAGPU_EINVAL;
        g_exact_round1 = (int)value;
        return 0;
    }
    if (!strcmp(name, "stream_compare")) {
        if (value != 0 && value != 1) return SAGPU_EINVAL;
        g_stream_compare = (int)value;
        return 0;
    }
    if (!strcmp(name, "raw_round1")) {
        if (value != 0 && value != 1) return SAGPU_EINVAL;
        g_raw_round1 = (int)value;
        return 0;
    }
    if (!strcmp(name, "key_stream")) {
        if (value != 0 && value != 1) return SAGPU_EINVAL;
        g_key_stream = (int)value;
        return 0;
    }
    if (!strcmp(name, "radix_bits")) {
        if (value != 4 && value != 8 && value != 11) return SAGPU_EINVAL;
        g_radix_bits = (int)value;
        return 0;
    }
    if (!strcmp(name, "round_graphs")) {
        if (value != 0 && value != 1) return SAGPU_EINVAL;
        g_round_graphs = (int)value;
        return 0;
    }
    if (!strcmp(name, "radix_small_n")) {
        if (value < 0 || value > 0x7FFFFFFF) return SAGPU_EINVAL;
        g_radix_small_n = (int)value;
        return 0;
    }
    if (!strcmp(name, "big_pivot_min")) {
        if (value < 0 || value > 0x7FFFFFFF) return SAGPU_EINVAL;
        g_big_pivot_min = (int)value;
        return 0;
    }
    if (!strcmp(name, "perm_hist")) {
        if (value != 0 && value != 1) return SAGPU_EINVAL;
        g_perm_hist = (int)value;
        return 0;
    }
    if (!strcmp(name, "group_flags")) {
        if (value != 0 && value != 1) return SAGPU_EINVAL;
        g_group_flags = (int)value;
        return 0;
    }
    if (!strcmp(name, "windowed_scatter")) {
        if (value != 0 && value != 1) return SAGPU_EINVAL;
        g_windowed_scatter = (int)value;
        return 0;
    }
    if (!strcmp(name, "plcp_pairs")) {
        if (value != 0 && value != 1) return SAGPU_EINVAL;
        g_plcp_pairs = (int)value;
        return 0;
    }
    if (!strcmp(name, "bwt_variant")) {
        if (value < 0 || value > 2) return SAGPU_EINVAL;
        g_bwt_variant = (int)value;
        return 0;
    }
    if (!strcmp(name, "hist_windows")) {
        if (value != 0 && value != 1) return SAGPU_EINVAL;
        g_hist_windows = (int)value;
        return 0;
    }
    if (!strcmp(name, "init_rows")) {
        if (value != 8 && value != 16) return SAGPU_EINVAL;
        g_init_rows = (int)value;
        return 0;
    }
    if (!strcmp(name, "tile_range")) {
        if (value != 256 && value != 512 && value != 1024) return SAGPU_EINVAL;
        set_tile_range((int)value);
        return 0;
    }
    if (!strcmp(name, "isa_window_log2")) {
        if (value < 10 || value > 32) return SAGPU_EINVAL;
        set_isa_window_log2((int)value);
        return 0;
    }
    if (!strcmp(name, "wide_from")) {
        if (value < 0 || value > 0x7FFFFFFF) return SAGPU_EINVAL;
        g_wide_from = (u64)value;
        return 0;
    }
    return SAGPU_EINVAL;
}

namespace {

template <typename K>
int sort_pairs_dev(K *d_keys, u32 *d_vals, int64_t n, int begin_bit, int end_bit)
{
    if (!d_keys || !d_vals || n < 0 || begin_bit < 0 || end_bit > (int)(8 * sizeof(K)) || begin_bit > end_bit)
        return SAGPU_EINVAL;
    if (n <= 1) return 0;
    LEASE(c);
    Scope sc(*c);
    K *k1 = sc.get<K>((u64)n);
    u32 *v1 = sc.get<u32>((u64)n);
    if (!k1 || !v1) return SAGPU_ENOMEM;
    SAGPU_TRY(ensure_radix_scratch(*c, sc, (u64)n));
    int alt = 0, passes = 0;
    c->rs.hook = nullptr;
    int rc = sizeof(K) == 8
                 ? radix_sort_u64(c->rs, (u64 *)d_keys, d_vals, (u64 *)k1, v1, (u64)n, begin_bit, end_bit, c->stream,
                                  &alt, &passes)
                 : radix_sort_u32(c->rs, (u32 *)d_keys, d_vals, (u32 *)k1, v1, (u64)n, begin_bit, end_bit, c->stream,
                                  &alt, &passes);
    if (rc) return rc;
    if (alt) {
        SAGPU_CHECK(hipMemcpyAsync(d_keys, k1, (size_t)n * sizeof(K), hipMemcpyDeviceToDevice, c->stream));
        SAGPU_CHECK(hipMemcpyAsync(d_vals, v1, (size_t)n * sizeof(u32), hipMemcpyDeviceToDevice, c->stream));
    }
    SAGPU_CHECK(hipStreamSynchronize(c->stream));
    return 0;
}

__global__ void k_bench_keys(u64 *k64, u32 *k32, u32 *v, u64 n, u64 seed, u64 mask)
{
    const u64 stride = (u64)gridDim.x * blockDim.x;
    for (u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        u64 z = (i + seed) * 0x9E3779B97F4A7C15ull;
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        z ^= z >> 31;
        if (k64) k64[i] = z & mask;
        else k32[i] = (u32)(z & mask);
        v[i] = (u32)i;
    }
}

} // namespace

extern "C" int32_t sagpu_sort_pairs_u32(uint32_t *d_keys, uint32_t *d_vals, int64_t n, int32_t b, int32_t e)
{
    return sort_pairs_dev<u32>(d_keys, d_vals, n, b, e);
}

extern "C" int32_t sagpu_sort_pairs_u64(uint64_t *d_keys, uint32_t *d_vals, int64_t n, int32_t b, int32_t e)
{
    return sort_pairs_dev<u64>((u64 *)d_keys, d_vals, n, b, e);
}

extern "C" int32_t sagpu_sort_bench(int64_t n, int32_t key_bytes, int32_t key_bits, int32_t reps, double *ms_per_sort,
                                    double *ms_per_pass)
{
    if (n < 2 || (key_bytes != 4 && key_bytes != 8) || key_bits < 1 || key_bits > 8 * key_bytes || reps < 1)
        return SAGPU_EINVAL;
    LEASE(c);
    Scope sc(*c);
    u64 *k64a = nullptr, *k64b = nullptr;
    u32 *k32a = nullptr, *k32b = nullptr;
    if (key_bytes == 8) {
        k64a = sc.get<u64>((u64)n);
        k64b = sc.get<u64>((u64)n);
        if (!k64a || !k64b) return SAGPU_ENOMEM;
    } else {
        k32a = sc.get<u32>((u64)n);
        k32b = sc.get<u32>((u64)n);
        if (!k32a || !k32b) return SAGPU_ENOMEM;
    }
    u32 *va = sc.get<u32>((u64)n), *vb = sc.get<u32>((u64)n);
    if (!va || !vb) return SAGPU_ENOMEM;
    SAGPU_TRY(ensure_radix_scratch(*c, sc, (u64)n));
    const u64 mask = key_bits >= 64 ? ~0ull : ((1ull << key_bits) - 1ull);
    hipEvent_t e0, e1;
    SAGPU_CHECK(hipEventCreate(&e0));
    SAGPU_CHECK(hipEventCreate(&e1));
    double total = 0;
    int passes_total = 0;
    c->rs.hook = nullptr;
    for (int r = 0; r < reps + 1; r++) {
        hipLaunchKernelGGL(k_bench_keys, dim3(4096), dim3(256), 0, c->stream, k64a, k32a, va, (u64)n,
                           (u64)r * 7919u, mask);
        SAGPU_CHECK(hipEventRecord(e0, c->stream));
        int alt = 0, passes = 0;
        int rc = key_bytes == 8 ? radix_sort_u64(c->rs, k64a, va, k64b, vb, (u64)n, 0, key_bits, c->stream, &alt,
                                                 &passes)
                                : radix_sort_u32(c->rs, k32a, va, k32b, vb, (u64)n, 0, key_bits, c->stream, &alt,
                                                 &passes);
        if (rc) return rc;
        SAGPU_CHECK(hipEventRecord(e1, c->stream));
        SAGPU_CHECK(hipEventSynchronize(e1));
        float ms = 0;
        SAGPU_CHECK(hipEventElapsedTime(&ms, e0, e1));
        if (r > 0) {
            total += ms;
            passes_total += passes;
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (ms_per_sort) *ms_per_sort = total / reps;
    if (ms_per_pass) *ms_per_pass = passes_total ? total / passes_total : 0.0;
    return 0;
}

extern "C" int32_t sagpu_get_stats(sagpu_stats *out)
{
    if (!out) return SAGPU_EINVAL;
    std::lock_guard<std::mutex> lk(g_stats_mu);
    *out = g_last_stats;
    out->device_allocs = g_device_allocs.load();
    return 0;
}

namespace {
// sagpu_build32 on one context.
int64_t build32_on(Ctx &c, const uint8_t *d_T, u64 n, int32_t flags, int32_t *d_SA, uint8_t *d_U, int32_t *d_PLCP,
                   int32_t *d_LCP, int64_t *h_freq)
{
    stats_reset(c);
    Scope sc(c);
    uint8_t *dT;
    SAGPU_TRY(padded_text(c, sc, d_T, n, SRC_DEVICE, &dT));
    u64 f[256];
    int64_t rc = build_dev(c, sc, dT, n, (flags & SAGPU_FLAG_GSA) ? 1 : 0, (u32 *)d_SA, d_U, (u32 *)d_PLCP,
                           (u32 *)d_LCP, f);
    if (rc >= 0 && h_freq)
        for (int ch = 0; ch < 256; ch++) h_freq[ch] = (int64_t)f[ch];
    return rc;
}
} // namespace

extern "C" int64_t sagpu_build32(const uint8_t *d_T, int64_t n, int32_t flags, int32_t *d_SA, uint8_t *d_U,
                                 int32_t *d_PLCP, int32_t *d_LCP, int64_t *h_freq)
{
    if (!d_T || n < 2 || n > INT32_MAX) return SAGPU_EINVAL;
    LEASE(c);
    return build32_on(*c, d_T, (u64)n, flags, d_SA, d_U, d_PLCP, d_LCP, h_freq);
}

// ------------------------------------------------------ allocation-free build
// The working buffers of sagpu_build32_noalloc: one static array per GPU (each
// device holds its own copy, at its own address, from module load until the
// process ends).  Host code never passes this name to a kernel or a copy: the
// device address comes from hipGetSymbolAddress at set-up (NoAllocDev::ptr).
static __device__ __align__(256) uint8_t scratch[SAGPU_NOALLOC_SCRATCH_BYTES];

namespace {

constexpr int NOALLOC_DEVICES = 64;

// Per-device state of the allocation-free build: the scratch address and an engine
// context whose pool carves every buffer from it (arena mode).  The radix sort's
// persistent scratch is carved once, at the bottom, for SAGPU_NOALLOC_MAX_N.
struct NoAllocDev {
    std::mutex mu;        // one call at a time per device: calls share the array
    Ctx *c = nullptr;
    uint8_t *ptr = nullptr;
};
NoAllocDev g_noalloc[NOALLOC_DEVICES];

// Set-up for `dev` (the current device); caller holds g_noalloc[dev].mu.
int noalloc_setup(int dev)
{
    NoAllocDev &d = g_noalloc[dev];
    if (d.c) return SAGPU_OK;
    void *ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr) {
        (void)hipGetLastError();
        sagpu_set_error("hipGetSymbolAddress of the allocation-free scratch failed", __FILE__, __LINE__);
        return SAGPU_EDEVICE;
    }
    Ctx *c = nullptr;
    SAGPU_TRY(ctx_create(&c, dev));
    c->pool.arena = (uint8_t *)ptr;
    // the top 64 KiB stay unused, so a vector load that runs a few bytes past the
    // last buffer still reads inside the array
    c->pool.arena_cap = SAGPU_NOALLOC_SCRATCH_BYTES - (64u << 10);
    Scope keep(*c);   // nothing is put back: the radix scratch stays at the bottom
    int rc = ensure_radix_scratch(*c, keep, SAGPU_NOALLOC_MAX_N);
    keep.held.clear();
    if (rc == SAGPU_OK && hipStreamSynchronize(c->stream) != hipSuccess) rc = SAGPU_EDEVICE;
    if (rc != SAGPU_OK) {
        (void)hipGetLastError();
        ctx_destroy(c);
        return rc;
    }
    d.ptr = (uint8_t *)ptr;
    d.c = c;
    return SAGPU_OK;
}

int current_device(int *dev)
{
    if (hipGetDevice(dev) != hipSuccess || *dev < 0 || *dev >= NOALLOC_DEVICES) {
        (void)hipGetLastError();
        sagpu_set_error("no usable HIP device", __FILE__, __LINE__);
        return SAGPU_EDEVICE;
    }
    return SAGPU_OK;
}

int noalloc_init_current()
{
    int dev = 0;
    SAGPU_TRY(current_device(&dev));
    std::lock_guard<std::mutex> lk(g_noalloc[dev].mu);
    return noalloc_setup(dev);
}

} // namespace

extern "C" int64_t sagpu_build32_noalloc(const uint8_t *d_T, int64_t n, int32_t flags, int32_t *d_SA,
                                         uint8_t *d_U, int32_t *d_PLCP, int32_t *d_LCP, int64_t *h_freq)
{
    if (!d_T || n < 2) return SAGPU_EINVAL;
    if (n > SAGPU_NOALLOC_MAX_N) {
        sagpu_set_error("sagpu_build32_noalloc: n exceeds SAGPU_NOALLOC_MAX_N (the static scratch holds "
                        "the working buffers of at most that many positions)", __FILE__, __LINE__);
        return SAGPU_EINVAL;
    }
    int dev = 0;
    SAGPU_TRY(current_device(&dev));
    NoAllocDev &d = g_noalloc[dev];
    std::lock_guard<std::mutex> lk(d.mu);
    SAGPU_TRY(noalloc_setup(dev));
    Ctx &c = *d.c;
    apply_options(c);
    c.pool.reset_peak();
    Ctx *prev = t_ctx;
    t_ctx = &c;
    const int64_t rc = build32_on(c, d_T, (u64)n, flags, d_SA, d_U, d_PLCP, d_LCP, h_freq);
    t_ctx = prev;
    c.stats.peak_bytes = (int64_t)c.pool.peak_bytes;
    {
        std::lock_guard<std::mutex> sl(g_stats_mu);
        g_last_stats = c.stats;
    }
    return rc;
}

extern "C" int32_t sagpu_plcp32(const uint8_t *d_T, const int32_t *d_SA, int32_t *d_PLCP, int64_t n, int32_t flags)
{
    if (!d_T || !d_SA || !d_PLCP || n < 2 || n > INT32_MAX) return SAGPU_EINVAL;
    LEASE(c);
    Scope sc(*c);
    uint8_t *dT;
    SAGPU_TRY(padded_text(*c, sc, d_T, (u64)n, SRC_DEVICE, &dT));
    u32 *dPHI = sc.get<u32>((u64)n);
    if (!dPHI) return SAGPU_ENOMEM;
    SAGPU_TRY(phi_run(*c, sc, (const u32 *)d_SA, (u64)n, dPHI));
    SAGPU_TRY(plcp_run(*c, sc, dT, (u64)n, (flags & SAGPU_FLAG_GSA) ? 1 : 0, dPHI, nullptr, nullptr, (u32 *)d_PLCP, nullptr));
    SAGPU_CHECK(hipStreamSynchronize(c->stream));
    return 0;
}

extern "C" int32_t sagpu_lcp32(const int32_t *d_PLCP, const int32_t *d_SA, int32_t *d_LCP, int64_t n)
{
    if (!d_PLCP || !d_SA || !d_LCP || n < 1 || n > INT32_MAX) return SAGPU_EINVAL;
    LEASE(c);
    SAGPU_TRY(launch_lcp((const u32 *)d_PLCP, (const u32 *)d_SA, (u64)n, (u32 *)d_LCP, c->stream));
    SAGPU_CHECK(hipStreamSynchronize(c->stream));
    return 0;
}

extern "C" int32_t sagpu_gsa_batch32(const uint8_t *d_T, int64_t n, const int64_t *h_offsets, int32_t count,
                                     int32_t *d_SA, int32_t *d_LCP)
{
    if (!d_T || !h_offsets || !d_SA || count < 1 || n < 2 || n > INT32_MAX) return SAGPU_EINVAL;
    if (h_offsets[0] != 0 || h_offsets[count] != n) return SAGPU_EINVAL;
    for (int32_t q = 0; q < count; q++)
        if (h_offsets[q + 1] <= h_offsets[q] + 1) return SAGPU_EINVAL;
    LEASE(c);
    stats_reset(*c);
    Scope sc(*c);
    hipStream_t s = c->stream;
    uint8_t *dT;
    SAGPU_TRY(padded_text(*c, sc, d_T, (u64)n, SRC_DEVICE, &dT));
    u32 *sid = sc.get<u32>((u64)n);
    int64_t *d_off = sc.get<int64_t>((u64)count + 1);
    u32 *d_bad = sc.get<u32>(64);
    u32 *part = sc.get<u32>((u64)n / 4096 + 2);
    if (!sid || !d_off || !d_bad || !part) return SAGPU_ENOMEM;
    ScanScratch ssc{part, (u64)n / 4096 + 2};
    SAGPU_TRY(launch_string_ids(dT, (u64)n, sid, ssc, s));
    SAGPU_CHECK(hipMemcpyAsync(d_off, h_offsets, ((size_t)count + 1) * sizeof(int64_t), hipMemcpyHostToDevice, s));
    SAGPU_TRY(launch_fill_u32(d_bad, 1, 0, s));
    SAGPU_TRY(launch_check_offsets(dT, sid, d_off, (u32)count, d_bad, s));
    u32 bad = 0;
    SAGPU_TRY(read_u32(*c, d_bad, &bad));
    if (bad) return SAGPU_EINVAL;

    u32 *gSA = sc.get<u32>((u64)n), *ISA = sc.get<u32>((u64)n);
    if (!gSA || !ISA) return SAGPU_ENOMEM;
    u32 *marks = nullptr, *P = nullptr;
    if (d_LCP) {
        marks = (u32 *)d_LCP;
        P = sc.get<u32>((u64)n);
        if (!P) return SAGPU_ENOMEM;
    }
    {
        Span total(*c, F_TOTAL);
        int rc = sa_build(*c, sc, dT, (u64)n, 1, gSA, ISA, nullptr, marks, sid, (u32)count);
        if (rc != SAGPU_OK) return rc;
        if (d_LCP) {
            {
                Span sp(*c, F_PLCP);
                SAGPU_TRY(plcp_run(*c, sc, dT, (u64)n, 1, nullptr, ISA, gSA, P, marks));
            }
            Span sp(*c, F_LCP);
            SAGPU_TRY(launch_lcp_fix(P, gSA, (u64)n, (u32 *)d_LCP, s));
        }
        SAGPU_TRY(launch_localize(gSA, sid, d_off, (u64)n, (u32 *)d_SA, s));
    }
    SAGPU_CHECK(hipStreamSynchronize(s));
    timer_collect(*c);
    return 0;
}

// ======================================================== jobs and serving

#include <chrono>
#include <condition_variable>
#include <deque>
#include <thread>

namespace {

double now_s()
{
    using namespace std::chrono;
    return duration<double>(steady_clock::now().time_since_epoch()).count();
}

// Device bytes one build of an n-byte text may hold at once, used for admission:
// text, SA, ISA, LCP marks and PLCP (17 n), the three 8 n sort/list buffers (24 n),
// code stream, BWT seeds and radix look-back (about 3 n), PLCP pair words (8.5 n),
// and up to 24 n of big-group sort buffers in one round; plus staging.
u64 job_footprint(u64 n) { return 80ull * n + (64ull << 20); }

int run_tiny(sagpu_job *j)
{
    if (j->n == 1) {
        j->SA[0] = 0;
        if (j->kind == SAGPU_JOB_BWT) { j->U[0] = j->T[0]; j->primary = 1; }
        if (j->kind == SAGPU_JOB_LCP) j->LCP[0] = 0;
    } else if (j->kind == SAGPU_JOB_BWT) {
        j->primary = 0;
    }
    return 0;
}

// One job on one context: SA (+ BWT or LCP) of a host text into host buffers.
int run_single(Ctx &c, sagpu_job *j)
{
    if (j->n <= 1) return run_tiny(j);
    const u64 n = (u64)j->n;
    if (n > (u64)INT32_MAX) return SAGPU_ENOMEM;
    stats_reset(c);
    Scope sc(c);
    uint8_t *dT;
    SAGPU_TRY(padded_text(c, sc, j->T, n, SRC_HOST, &dT));
    GET(dSA, u32, n);
    uint8_t *dU = nullptr;
    u32 *dL = nullptr;
    if (j->kind == SAGPU_JOB_BWT) {
        dU = sc.get<uint8_t>(n);
        if (!dU) return SAGPU_ENOMEM;
    }
    if (j->kind == SAGPU_JOB_LCP) {
        dL = sc.get<u32>(n);
        if (!dL) return SAGPU_ENOMEM;
    }
    int64_t r = build_dev(c, sc, dT, n, 0, dSA, dU, nullptr, dL, nullptr);
    if (r < 0) return (int)r;
    SAGPU_TRY(d2h(c, j->SA, dSA, n * 4));
    if (dU) SAGPU_TRY(d2h(c, j->U, dU, n));
    if (dL) SAGPU_TRY(d2h(c, j->LCP, dL, n * 4));
    SAGPU_CHECK(hipStreamSynchronize(c.stream));
    if (dU) j->primary = r;
    return 0;
}

int ensure_stage(Ctx &c, size_t bytes)
{
    if (c.stage_cap >= bytes) return SAGPU_OK;
    if (c.stage) (void)hipHostFree(c.stage);
    c.stage = nullptr;
    c.stage_cap = 0;
    size_t cap = bytes + (bytes >> 2);
    SAGPU_CHECK(hipHostMalloc((void **)&c.stage, cap, hipHostMallocDefault));
    c.stage_cap = cap;
    return SAGPU_OK;
}

// Several small jobs (texts without 0 bytes) as one batched build: every text is
// followed by a separator and the keys carry the job id, so the batched suffix
// array of job k, minus its separator suffix, is exactly job k's suffix array;
// lcps stop at the separator, which is where job k's text ends.  Returns -1 when
// a text turns out to contain a 0 byte (caller then runs the jobs one by one).
int run_batch(Ctx &c, sagpu_job **jobs, int count)
{
    std::vector<int64_t> off((size_t)count + 1);
    u64 N = 0;
    bool any_u = false, any_l = false;
    for (int q = 0; q < count; q++) {
        off[q] = (int64_t)N;
        N += (u64)jobs[q]->n + 1;
        any_u |= jobs[q]->kind == SAGPU_JOB_BWT;
        any_l |= jobs[q]->kind == SAGPU_JOB_LCP;
    }
    off[count] = (int64_t)N;
    if (N > (u64)INT32_MAX) return SAGPU_ENOMEM;
    const u64 M = N - (u64)count;  // packed output entries
    SAGPU_TRY(ensure_stage(c, (size_t)(N + 8 * M + 64)));
    uint8_t *hin = c.stage;
    for (int q = 0; q < count; q++) {
        memcpy(hin + off[q], jobs[q]->T, (size_t)jobs[q]->n);
        hin[off[q + 1] - 1] = 0;
    }
    stats_reset(c);
    Scope sc(c);
    hipStream_t s = c.stream;
    uint8_t *dT;
    SAGPU_TRY(padded_text(c, sc, hin, N, SRC_PINNED, &dT));
    GET(sid, u32, N);
    GET(d_off, int64_t, (u64)count + 1);
    GET(d_bad, u32, 64);
    GET(part, u32, N / 4096 + 2);
    ScanScratch ssc{part, N / 4096 + 2};
    SAGPU_TRY(launch_string_ids(dT, N, sid, ssc, s));
    SAGPU_CHECK(hipMemcpyAsync(d_off, off.data(), ((size_t)count + 1) * sizeof(int64_t), hipMemcpyHostToDevice, s));
    SAGPU_TRY(launch_fill_u32(d_bad, 1, 0, s));
    SAGPU_TRY(launch_check_offsets(dT, sid, d_off, (u32)count, d_bad, s));
    u32 bad = 0;
    SAGPU_TRY(read_u32(c, d_bad, &bad));
    if (bad) return SAGPU_EINVAL;

    GET(gSA, u32, N);
    GET(ISA, u32, N);
    u32 *marks = nullptr, *P = nullptr, *gL = nullptr;
    if (any_l) {
        marks = sc.get<u32>(N);
        P = sc.get<u32>(N);
        gL = marks;
        if (!marks || !P) return SAGPU_ENOMEM;
    }
    SAGPU_TRY(sa_build(c, sc, dT, N, 1, gSA, ISA, nullptr, marks, sid, (u32)count));
    if (any_l) {
        SAGPU_TRY(plcp_run(c, sc, dT, N, 1, nullptr, ISA, gSA, P, marks));
        SAGPU_TRY(launch_lcp_fix(P, gSA, N, gL, s));
    }
    GET(oSA, u32, M);
    u32 *oL = nullptr;
    uint8_t *oU = nullptr;
    if (any_l) {
        oL = sc.get<u32>(M);
        if (!oL) return SAGPU_ENOMEM;
    }
    if (any_u) {
        oU = sc.get<uint8_t>(M);
        if (!oU) return SAGPU_ENOMEM;
    }
    GET(prim, u32, (u64)count);
    SAGPU_TRY(launch_batch_extract(gSA, gL, sid, d_off, ISA, dT, N, oSA, oL, oU, s));
    SAGPU_TRY(launch_batch_primary(ISA, d_off, (u32)count, N, prim, s));
    uint8_t *hSA = c.stage + ((N + 63) & ~(u64)63);
    uint8_t *hL = hSA + 4 * M;
    uint8_t *hU = c.stage;  // input text no longer needed on the host side
    std::vector<u32> hprim((size_t)count);
    SAGPU_CHECK(hipMemcpyAsync(hSA, oSA, 4 * M, hipMemcpyDeviceToHost, s));
    if (oL) SAGPU_CHECK(hipMemcpyAsync(hL, oL, 4 * M, hipMemcpyDeviceToHost, s));
    if (oU) SAGPU_CHECK(hipMemcpyAsync(hU, oU, M, hipMemcpyDeviceToHost, s));
    SAGPU_CHECK(hipMemcpyAsync(hprim.data(), prim, (size_t)count * 4, hipMemcpyDeviceToHost, s));
    SAGPU_CHECK(hipStreamSynchronize(s));
    for (int q = 0; q < count; q++) {
        sagpu_job *j = jobs[q];
        const u64 po = (u64)off[q] - (u64)q, nk = (u64)j->n;
        memcpy(j->SA, hSA + 4 * po, 4 * nk);
        if (j->kind == SAGPU_JOB_LCP) memcpy(j->LCP, hL + 4 * po, 4 * nk);
        if (j->kind == SAGPU_JOB_BWT) {
            memcpy(j->U, hU + po, nk);
            j->primary = hprim[q];
        }
        j->batched = 1;
    }
    return 0;
}

struct Server {
    std::vector<Ctx *> ctxs;
    std::vector<std::thread> workers;
    std::mutex mu;
    std::condition_variable cv;
    std::deque<sagpu_job *> queue;
    u64 budget = 0, reserved = 0;
    u64 batch_bytes = 0;
    u64 small_max = 1ull << 20;
    int batch_max_jobs = 8192;
    bool stop = false;
};

void finish(Server &sv, std::vector<sagpu_job *> &take, int rc)
{
    double t = now_s();
    for (sagpu_job *j : take) {
        j->t_done = t;
        __atomic_store_n(&j->status, rc, __ATOMIC_RELEASE);
    }
    sv.cv.notify_all();
}

void worker(Server *svp, int slot)
{
    Server &sv = *svp;
    Ctx &c = *sv.ctxs[slot];
    (void)hipSetDevice(c.device);
    std::vector<sagpu_job *> take;
    for (;;) {
        u64 need = 0;
        {
            std::unique_lock<std::mutex> lk(sv.mu);
            for (;;) {
                if (sv.stop && sv.queue.empty()) return;
                if (sv.queue.empty()) {
                    sv.cv.wait(lk);
                    continue;
                }
                take.clear();
                sagpu_job *f = sv.queue.front();
                u64 bytes = 0;
                if (sv.batch_bytes && f->batchable && (u64)f->n < sv.small_max) {
                    for (sagpu_job *j : sv.queue) {
                        if (!j->batchable || (u64)j->n >= sv.small_max) break;
                        if (!take.empty() && (bytes + (u64)j->n + 1 > sv.batch_bytes ||
                                              (int)take.size() >= sv.batch_max_jobs))
                            break;
                        take.push_back(j);
                        bytes += (u64)j->n + 1;
                    }
                } else {
                    take.push_back(f);
                    bytes = (u64)f->n + 1;
                }
                need = job_footprint(bytes);
                if (need > sv.budget) {
                    // can never be admitted: fail the oldest job alone
                    sv.queue.pop_front();
                    take.assign(1, f);
                    lk.unlock();
                    finish(sv, take, SAGPU_ENOMEM);
                    lk.lock();
                    continue;
                }
                if (sv.reserved + need > sv.budget) {
                    sv.cv.wait(lk);
                    continue;
                }
                for (size_t q = 0; q < take.size(); q++) sv.queue.pop_front();
                sv.reserved += need;
                break;
            }
        }
        double t0 = now_s();
        for (sagpu_job *j : take) j->t_start = t0;
        apply_options(c);
        int rc;
        if (take.size() == 1) {
            rc = run_single(c, take[0]);
        } else {
            rc = run_batch(c, take.data(), (int)take.size());
            if (rc == SAGPU_EINVAL) {
                rc = 0;
                for (sagpu_job *j : take) {
                    int r = run_single(c, j);
                    if (r) rc = r;
                }
            }
        }
        {
            std::lock_guard<std::mutex> lk(sv.mu);
            sv.reserved -= need;
        }
        finish(sv, take, rc < -2 ? SAGPU_ENOMEM : rc);
    }
}

} // namespace

extern "C" double sagpu_clock(void) { return now_s(); }

extern "C" int32_t sagpu_run_job(sagpu_job *job)
{
    if (!job || !job->T || !job->SA || job->n < 0 || job->kind < 0 || job->kind > 2 ||
        (job->kind == SAGPU_JOB_BWT && !job->U) || (job->kind == SAGPU_JOB_LCP && !job->LCP))
        return SAGPU_EINVAL;
    LEASE(c);
    job->batched = 0;
    job->t_start = now_s();
    int rc = run_single(*c, job);
    job->t_done = now_s();
    job->status = rc < -2 ? SAGPU_ENOMEM : rc;
    return job->status;
}

extern "C" void *sagpu_server_create(int32_t slots, uint64_t hbm_budget, int64_t batch_bytes)
{
    if (slots < 1 || slots > 16 || batch_bytes < 0) return nullptr;
    Server *sv = new Server();
    for (int q = 0; q < slots; q++) {
        Ctx *c = nullptr;
        if (ctx_create(&c) != SAGPU_OK) {
            for (Ctx *x : sv->ctxs) ctx_destroy(x);
            delete sv;
            return nullptr;
        }
        sv->ctxs.push_back(c);
    }
    if (!hbm_budget) {
        size_t fr = 0, tot = 0;
        if (hipMemGetInfo(&fr, &tot) != hipSuccess) fr = 16ull << 30;
        hbm_budget = (uint64_t)(0.9 * (double)fr);
    }
    sv->budget = hbm_budget;
    sv->batch_bytes = (u64)batch_bytes;
    for (int q = 0; q < slots; q++) sv->workers.emplace_back(worker, sv, q);
    return sv;
}

extern "C" int32_t sagpu_server_submit(void *server, sagpu_job *job)
{
    if (!server || !job || !job->T || !job->SA || job->n < 0 || job->kind < 0 || job->kind > 2 ||
        (job->kind == SAGPU_JOB_BWT && !job->U) || (job->kind == SAGPU_JOB_LCP && !job->LCP))
        return SAGPU_EINVAL;
    Server &sv = *(Server *)server;
    job->batched = 0;
    job->primary = 0;
    job->batchable = job->n >= 2 && memchr(job->T, 0, (size_t)job->n) == nullptr;
    job->t_submit = now_s();
    job->t_start = job->t_done = 0;
    __atomic_store_n(&job->status, 1, __ATOMIC_RELEASE);
    {
        std::lock_guard<std::mutex> lk(sv.mu);
        if (sv.stop) return SAGPU_EINVAL;
        sv.queue.push_back(job);
    }
    sv.cv.notify_all();
    return 0;
}

extern "C" int32_t sagpu_server_wait(void *server, sagpu_job *job)
{
    if (!server || !job) return SAGPU_EINVAL;
    Server &sv = *(Server *)server;
    std::unique_lock<std::mutex> lk(sv.mu);
    sv.cv.wait(lk, [&] { return __atomic_load_n(&job->status, __ATOMIC_ACQUIRE) != 1; });
    return job->status;
}

extern "C" int32_t sagpu_server_drain(void *server)
{
    if (!server) return SAGPU_EINVAL;
    Server &sv = *(Server *)server;
    std::unique_lock<std::mutex> lk(sv.mu);
    sv.cv.wait(lk, [&] { return sv.queue.empty() && sv.reserved == 0; });
    return 0;
}

extern "C" void sagpu_server_destroy(void *server)
{
    if (!server) return;
    Server *sv = (Server *)server;
    {
        std::lock_guard<std::mutex> lk(sv->mu);
        sv->stop = true;
    }
    sv->cv.notify_all();
    for (auto &t : sv->workers) t.join();
    for (Ctx *c : sv->ctxs) ctx_destroy(c);
    delete sv;
}
