// psacak.hip -- C ABI of psacak_amd (include/psacak.h): contexts, argument
// checks, host/device entry points, batching and the device-side SA checker.
//
// Argument checks run before any HIP call, so they work (and are tested)
// without a GPU.  Every entry point converts failures into status codes.
#include <chrono>
#include <memory>
#include <mutex>

#include "psk_level8.hpp"
#include "psk_server.hpp"
#include "psk_xfer.hpp"

using namespace psk;

struct psk_ctx {
    int device = -1;
    bool ready = false;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    hipEvent_t ev2 = nullptr;  // end of the translation of a psacak32_raw call
    Arena arena;
    u32 *pinned = nullptr;
    u64 budget = 0;
    std::string err;
    Runtime rt;
    Prof prof;
    Xfer xfer;
    double last_ms[4] = {0, 0, 0, 0};
    u64 last_syncs = 0;  // host synchronisations inside the last timed call
};

namespace {

void ensure_ready(psk_ctx *c) {
    if (c->ready) return;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count == 0) {
        (void)hipGetLastError();
        fail(PSK_EDEVICE, "no GPU device present");
    }
    PSK_HIP(hipGetDevice(&c->device));
    PSK_HIP(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    PSK_HIP(hipEventCreate(&c->ev0));
    PSK_HIP(hipEventCreate(&c->ev1));
    PSK_HIP(hipEventCreate(&c->ev2));
    PSK_HIP(hipHostMalloc((void **)&c->pinned, 16384, hipHostMallocDefault));
    PSK_HIP(hipMalloc((void **)&c->rt.d.err, 256));
    PSK_HIP(hipMemset(c->rt.d.err, 0, 256));
    c->arena.set_budget(c->budget);
    c->rt.d.s = c->stream;
    c->rt.d.a = &c->arena;
    c->rt.d.pinned = c->pinned;
    c->rt.d.prof = &c->prof;
    {
        const char *tr = getenv("PSACAK_AMD_TRACE");
        c->rt.d.trace = tr && atoi(tr) > 0;
    }
    c->ready = true;
}

template <class F>
int guarded(psk_ctx *c, F f) {
    try {
        if (c) c->err.clear();
        f();
        return PSK_OK;
    } catch (const Error &e) {
        if (c) c->err = e.msg;
        return e.code;
    } catch (const std::bad_alloc &) {
        if (c) c->err = "host allocation failed";
        return PSK_ENOMEM;
    } catch (...) {
        if (c) c->err = "unknown failure";
        return PSK_EDEVICE;
    }
}

psk_ctx *default_ctx() {
    thread_local std::unique_ptr<psk_ctx> c;
    if (!c) c.reset(new psk_ctx());
    return c.get();
}

int check8(const void *text, u64 n, const void *sa, u64 sa_len) {
    if (n == 0) return PSK_OK;
    if (!text || !sa) return PSK_EINVAL;
    if (sa_len < n) return PSK_EINVAL;       // lib.rs:78
    if (n >= 0xFFFFFFFFull) return PSK_ETOOLONG;  // lib.rs:65,79
    return PSK_OK;
}

int check32(const void *text, u64 n, const void *sa, u64 sa_len) {
    if (n == 0) return PSK_OK;
    if (!text || !sa) return PSK_EINVAL;
    if (sa_len < n) return PSK_EINVAL;
    if (n > 0x7FFFFFFFull) return PSK_ETOOLONG;  // psacak32.rs:25
    return PSK_OK;
}

struct Timer {
    psk_ctx *c;
    u64 syncs0;
    bool split = false;  // ev2 marks the end of a first phase (translation)
    explicit Timer(psk_ctx *cc) : c(cc), syncs0(cc->rt.d.syncs) {
        c->rt.d.zero(c->rt.d.err, sizeof(u32));
        PSK_HIP(hipEventRecord(c->ev0, c->stream));
    }
    void mark() {
        PSK_HIP(hipEventRecord(c->ev2, c->stream));
        split = true;
    }
    void stop() {
        PSK_HIP(hipEventRecord(c->ev1, c->stream));
        PSK_HIP(hipEventSynchronize(c->ev1));
        float ms = 0, ms0 = 0;
        PSK_HIP(hipEventElapsedTime(&ms, c->ev0, c->ev1));
        if (split) PSK_HIP(hipEventElapsedTime(&ms0, c->ev0, c->ev2));
        c->last_ms[0] = ms;
        c->last_ms[2] = ms0;
        c->last_syncs = c->rt.d.syncs - syncs0;
        if (c->rt.d.trace)
            fprintf(stderr, "[psk] call: %.3f ms, %llu host synchronisations\n", ms,
                    (unsigned long long)c->last_syncs);
        // the index-check flag and each level's first LMS position, in one read
        u32 w[64];
        c->rt.d.read(w, c->rt.d.err, sizeof(w));
        if (w[0]) fail(PSK_EDEVICE, "internal index check failed (flags " + std::to_string(w[0]) + ")");
        for (size_t i = 0; i < c->rt.levels.size() && i < 62; i++) {
            LevelStat &s = c->rt.levels[i];
            if (s.m1 && !s.sumlen) s.sumlen = s.m - w[1 + i] + s.m1;
        }
    }
};

// every symbol < n? (bucket-pointer precondition, checked on the device)
void check_precond32(psk_ctx *c, const u32 *d_text, u64 n) {
    Dev &d = c->rt.d;
    ArenaScope mk(*d.a);
    u32 *bad = d.a->get<u32>(1);
    d.zero(bad, sizeof(u32));
    const u64 nn = n;
    for_each(d, n, [=] __device__(u64 i) {
        if (d_text[i] >= nn) atomicOr(bad, 1u);
    });
    u32 b = d.read_u32(bad);
    mk.reset();
    if (b) fail(PSK_EPRECOND, "psacak32: a symbol is >= n (not a bucket-pointer text)");
}

void run8_dev(psk_ctx *c, const u8 *d_text, u64 n, u32 *d_sa) {
    Timer t(c);
    psacak8_run(c->rt, d_text, n, d_sa);
    t.stop();
}

void run32_dev(psk_ctx *c, const u32 *d_text, u64 n, u32 *d_sa, bool raw) {
    if (!raw) check_precond32(c, d_text, n);
    Timer t(c);
    if (raw) {
        ArenaScope mk(c->arena);
        u32 *tr = c->arena.get<u32>(n);
        const bool go = translate32_run(c->rt, d_text, n, tr, d_sa);
        t.mark();
        if (go) psacak32_run(c->rt, tr, n, d_sa);
        mk.reset();
    } else {
        psacak32_run(c->rt, d_text, n, d_sa);
    }
    t.stop();
}

// The SA check of bin/psacak.rs:73-80 in linear time (Burkhardt and
// Kaerkkaeinen): sa is a permutation of [0, n) (every slot claims its suffix
// once), and for every adjacent pair (a, b) either T[a] < T[b], or T[a] == T[b]
// and the suffix after a ranks below the one after b (the empty suffix
// lowest).  bad: 1 = not a permutation, 2 = order violated.  rank: n u32.
template <class C>
void check_sa_dev(Dev &d, const C *text, u64 n, const u32 *sa, u32 *rank, u32 *bad) {
    const u64 nn = n;
    d.fill_u32(rank, 0xFFFFFFFFu, n);
    for_each(d, n, [=] __device__(u64 i) {
        u32 x = sa[i];
        if (x >= nn) { atomicOr(bad, 1u); return; }
        u32 old = atomicExch(&rank[x], (u32)i);
        if (old != 0xFFFFFFFFu) atomicOr(bad, 1u);
    });
    for_each(d, n - 1, [=] __device__(u64 i) {
        u32 a = sa[i], b = sa[i + 1];
        if (a >= nn || b >= nn) return;
        C ca = text[a], cb = text[b];
        if (ca > cb) { atomicOr(bad, 2u); return; }
        if (ca == cb) {
            i64 ra = a + 1 < nn ? (i64)rank[a + 1] : -1;
            i64 rb = b + 1 < nn ? (i64)rank[b + 1] : -1;
            if (ra >= rb) atomicOr(bad, 2u);
        }
    });
}

hipStream_t pick_stream(psk_ctx *c, void *stream) {
    return stream ? (hipStream_t)stream : c->stream;
}

// Run f with the context's Dev pointed at `stream`.
// Every device entry point runs through here: the arena scope returns the
// context's workspace stack to where the call found it, on success and on
// failure alike.
template <class F>
void on_stream(psk_ctx *c, void *stream, F f) {
    ArenaScope outer(c->arena);
    hipStream_t saved = c->rt.d.s;
    c->rt.d.s = pick_stream(c, stream);
    hipStream_t saved_ctx = c->stream;
    c->stream = c->rt.d.s;
    try {
        f();
    } catch (...) {
        c->rt.d.s = saved;
        c->stream = saved_ctx;
        throw;
    }
    c->rt.d.s = saved;
    c->stream = saved_ctx;
}

}  // namespace

// ---- zero-allocation entry point ----------------------------------------------
// psk_psacak8_dev_noalloc takes every byte of device workspace from one
// file-scope array (each GPU the library runs on holds its own copy until the
// process ends) instead of the context arenas' hipMalloc'd chunks.  The
// array's name is a host-side stand-in for the device symbol: host code only
// ever uses the per-device address that hipGetSymbolAddress returned for it.
static __device__ __align__(256) unsigned char scratch[PSK_NOALLOC_SCRATCH_BYTES];

namespace {

constexpr int kNoAllocMaxDevices = 64;
constexpr u64 kNoAllocFlagBytes = 256;  // internal index-check flag, then the workspace stack

struct NoAllocState {
    std::mutex mu;          // one call at a time per device (they share the array)
    bool ready = false;
    void *scratch = nullptr;  // this device's copy of the array
    Arena arena;            // fixed: allocates from the array only
    Runtime rt;
    u32 host_buf[4096];     // host staging of small readbacks (pageable)
};

NoAllocState g_noalloc[kNoAllocMaxDevices];

// Set-up on the first call on a device, with that device current.
void noalloc_setup(NoAllocState &g) {
    if (g.ready) return;
    void *p = nullptr;
    PSK_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    g.scratch = p;
    g.arena.attach((char *)p + kNoAllocFlagBytes, PSK_NOALLOC_SCRATCH_BYTES - kNoAllocFlagBytes);
    g.rt.d.a = &g.arena;
    g.rt.d.err = (u32 *)p;
    g.rt.d.pinned = g.host_buf;
    g.rt.d.prof = nullptr;
    // per-call workspace proportional to the text: no fixed-size naming
    // dictionary, no placement records (results are identical)
    g.rt.dict_naming = false;
    g.rt.lms_records = false;
    g.rt.levels.reserve(64);
    g.ready = true;
}

}  // namespace

extern "C" {

const char *psk_strerror(int status) {
    switch (status) {
        case PSK_OK: return "ok";
        case PSK_EINVAL: return "invalid argument";
        case PSK_ETOOLONG: return "text too long";
        case PSK_EPRECOND: return "psacak32 precondition violated (symbol >= n)";
        case PSK_ENOMEM: return "out of memory";
        case PSK_EBUDGET: return "HBM budget exceeded";
        case PSK_EDEVICE: return "device error";
        case PSK_EBUSY: return "busy";
        case PSK_ENOTFOUND: return "not found";
        default: return "unknown status";
    }
}

const char *psk_version(void) { return "psacak_amd 0.1.0 (gfx950)"; }

int psk_ctx_create(psk_ctx **out, uint64_t hbm_budget_bytes) {
    if (!out) return PSK_EINVAL;
    psk_ctx *c = new (std::nothrow) psk_ctx();
    if (!c) return PSK_ENOMEM;
    c->budget = hbm_budget_bytes;
    *out = c;
    return PSK_OK;
}

void psk_ctx_destroy(psk_ctx *ctx) {
    if (!ctx) return;
    if (ctx->ready) {
        (void)hipStreamSynchronize(ctx->stream);
        ctx->arena.release();
        ctx->xfer.release();
        (void)hipEventDestroy(ctx->ev0);
        (void)hipEventDestroy(ctx->ev1);
        (void)hipEventDestroy(ctx->ev2);
        (void)hipStreamDestroy(ctx->stream);
        (void)hipHostFree(ctx->pinned);
        (void)hipFree(ctx->rt.d.err);
        ctx->prof.release();
    }
    delete ctx;
}

const char *psk_last_error(const psk_ctx *ctx) { return ctx ? ctx->err.c_str() : ""; }

int psk_ctx_set_direct_rounds(psk_ctx *ctx, int rounds) {
    if (!ctx || rounds < 0 || rounds > 64) return PSK_EINVAL;
    ctx->rt.direct_rounds = rounds;
    return PSK_OK;
}

int psk_ctx_set_doubling(psk_ctx *ctx, int distinct_per_1024) {
    if (!ctx || distinct_per_1024 < 0 || distinct_per_1024 > 1024) return PSK_EINVAL;
    ctx->rt.doubling_distinct = (u32)distinct_per_1024;
    return PSK_OK;
}

int psk_ctx_set_chain_induce(psk_ctx *ctx, int min_distinct_bytes) {
    if (!ctx || min_distinct_bytes < 0 || min_distinct_bytes > 256) return PSK_EINVAL;
    ctx->rt.chain_sigma = ctx->rt.chain_sigma_small = (u32)min_distinct_bytes;
    return PSK_OK;
}

int psk_ctx_set_dict_naming(psk_ctx *ctx, int on) {
    if (!ctx || on < 0 || on > 1) return PSK_EINVAL;
    ctx->rt.dict_naming = on != 0;
    return PSK_OK;
}

int psk_ctx_trim(psk_ctx *ctx) {
    if (!ctx) ctx = default_ctx();
    return guarded(ctx, [&] {
        if (ctx->ready) {
            PSK_HIP(hipStreamSynchronize(ctx->stream));
            ctx->arena.release();
        }
    });
}

int psk_ctx_workspace(const psk_ctx *ctx, uint64_t *reserved, uint64_t *peak) {
    if (!ctx) ctx = default_ctx();
    if (reserved) *reserved = ctx->arena.reserved();
    if (peak) *peak = ctx->arena.peak();
    return PSK_OK;
}

uint64_t psk_workspace_bytes32(uint64_t n) { return psk::workspace_bytes32(n); }

uint64_t psk_workspace_bytes8(uint64_t n) { return psk::workspace_bytes8(n); }

int psk_psacak_inplace(psk_ctx *ctx, const uint8_t *text, uint64_t n, uint32_t *sa,
                       uint64_t sa_len) {
    int rc = check8(text, n, sa, sa_len);
    if (rc != PSK_OK || n == 0) return rc;
    psk_ctx *c = ctx ? ctx : default_ctx();
    return guarded(c, [&] {
        ensure_ready(c);
        ArenaScope mk(c->arena);
        u8 *dt = c->arena.get<u8>(n);
        u32 *ds = c->arena.get<u32>(n);
        upload_segs(c->rt.d, c->xfer, {Seg{(char *)text, n}}, dt);
        run8_dev(c, dt, n, ds);
        download_segs(c->rt.d, c->xfer, ds, {Seg{(char *)sa, n * sizeof(u32)}});
    });
}

int psk_psacak(psk_ctx *ctx, const uint8_t *text, uint64_t n, uint32_t *sa) {
    return psk_psacak_inplace(ctx, text, n, sa, n);
}

static int host32(psk_ctx *ctx, const uint32_t *text, uint64_t n, uint32_t *sa, uint64_t sa_len,
                  bool raw) {
    int rc = check32(text, n, sa, sa_len);
    if (rc != PSK_OK || n == 0) return rc;
    psk_ctx *c = ctx ? ctx : default_ctx();
    return guarded(c, [&] {
        ensure_ready(c);
        ArenaScope mk(c->arena);
        u32 *dt = c->arena.get<u32>(n);
        u32 *ds = c->arena.get<u32>(n);
        upload_segs(c->rt.d, c->xfer, {Seg{(char *)text, n * sizeof(u32)}}, dt);
        run32_dev(c, dt, n, ds, raw);
        download_segs(c->rt.d, c->xfer, ds, {Seg{(char *)sa, n * sizeof(u32)}});
    });
}

int psk_psacak32(psk_ctx *ctx, const uint32_t *text, uint64_t n, uint32_t *sa, uint64_t sa_len) {
    return host32(ctx, text, n, sa, sa_len, false);
}

int psk_psacak32_raw(psk_ctx *ctx, const uint32_t *text, uint64_t n, uint32_t *sa,
                     uint64_t sa_len) {
    return host32(ctx, text, n, sa, sa_len, true);
}

int psk_psacak8_dev(psk_ctx *ctx, const uint8_t *d_text, uint64_t n, uint32_t *d_sa,
                    void *stream) {
    int rc = check8(d_text, n, d_sa, n);
    if (rc != PSK_OK || n == 0) return rc;
    psk_ctx *c = ctx ? ctx : default_ctx();
    return guarded(c, [&] {
        ensure_ready(c);
        on_stream(c, stream, [&] { run8_dev(c, d_text, n, d_sa); });
    });
}

int psk_psacak32_dev(psk_ctx *ctx, const uint32_t *d_text, uint64_t n, uint32_t *d_sa,
                     void *stream) {
    int rc = check32(d_text, n, d_sa, n);
    if (rc != PSK_OK || n == 0) return rc;
    psk_ctx *c = ctx ? ctx : default_ctx();
    return guarded(c, [&] {
        ensure_ready(c);
        on_stream(c, stream, [&] { run32_dev(c, d_text, n, d_sa, false); });
    });
}

int psk_psacak32_raw_dev(psk_ctx *ctx, const uint32_t *d_text, uint64_t n, uint32_t *d_sa,
                         void *stream) {
    int rc = check32(d_text, n, d_sa, n);
    if (rc != PSK_OK || n == 0) return rc;
    psk_ctx *c = ctx ? ctx : default_ctx();
    return guarded(c, [&] {
        ensure_ready(c);
        on_stream(c, stream, [&] { run32_dev(c, d_text, n, d_sa, true); });
    });
}

int psk_last_levels(const psk_ctx *ctx, uint64_t *rows, int max_levels) {
    if (!ctx || !rows || max_levels < 0) return PSK_EINVAL;
    int n = std::min<int>(max_levels, (int)ctx->rt.levels.size());
    for (int i = 0; i < n; i++) {
        const LevelStat &s = ctx->rt.levels[i];
        uint64_t *r = rows + 6 * i;
        r[0] = s.m; r[1] = s.m1; r[2] = s.mL; r[3] = s.k; r[4] = s.sumlen; r[5] = s.csize;
    }
    return n;
}

int psk_ctx_set_capture(psk_ctx *ctx, int level) {
    if (!ctx || level < -1 || level > 63) return PSK_EINVAL;
    ctx->rt.capture_depth = level;
    ctx->rt.captured.clear();
    ctx->rt.captured_ok = false;
    return PSK_OK;
}

int psk_last_capture(const psk_ctx *ctx, uint32_t *out, uint64_t cap, uint64_t *count) {
    if (!ctx || !count) return PSK_EINVAL;
    if (!ctx->rt.captured_ok) return PSK_ENOTFOUND;
    const u64 m = ctx->rt.captured.size();
    *count = m;
    if (out) {
        if (cap < m) return PSK_EINVAL;
        if (m) memcpy(out, ctx->rt.captured.data(), m * sizeof(u32));
    }
    return PSK_OK;
}

int psk_last_timing(const psk_ctx *ctx, double *ms, int max_entries) {
    if (!ctx || !ms) return PSK_EINVAL;
    int n = std::min(max_entries, 4);
    const double v[4] = {ctx->last_ms[0], (double)ctx->last_syncs, ctx->last_ms[2], 0.0};
    for (int i = 0; i < n; i++) ms[i] = v[i];
    return n;
}

int psk_profile(psk_ctx *ctx, int on) {
    if (!ctx) return PSK_EINVAL;
    return guarded(ctx, [&] {
        ensure_ready(ctx);
        if (on) ctx->prof.clear();
        else ctx->prof.flush();
        ctx->prof.on = on != 0;
    });
}

int psk_profile_only(psk_ctx *ctx, const char *names) {
    if (!ctx) return PSK_EINVAL;
    return guarded(ctx, [&] {
        ctx->prof.only.clear();
        std::string cur;
        for (const char *p = names ? names : ""; ; p++) {
            if (*p == ',' || *p == '\0') {
                if (!cur.empty()) ctx->prof.only.push_back(cur);
                cur.clear();
                if (*p == '\0') break;
            } else {
                cur += *p;
            }
        }
    });
}

int psk_profile_stats(psk_ctx *ctx, char *names, uint64_t names_len, uint64_t *launches,
                      double *ms, double *bytes, int max) {
    if (!ctx || !names || names_len == 0 || max < 0) return PSK_EINVAL;
    int count = 0;
    int rc = guarded(ctx, [&] {
        if (!ctx->ready) { names[0] = 0; return; }
        const auto &t = ctx->prof.totals();
        std::string all;
        for (size_t i = 0; i < t.size() && (int)i < max; i++) {
            if (i) all += "\n";
            all += t[i].name;
            if (launches) launches[i] = t[i].launches;
            if (ms) ms[i] = t[i].ms;
            if (bytes) bytes[i] = t[i].bytes;
            count++;
        }
        if (all.size() + 1 > names_len) fail(PSK_EINVAL, "names buffer too small");
        memcpy(names, all.c_str(), all.size() + 1);
    });
    return rc == PSK_OK ? count : rc;
}

// Texts are grouped into sub-batches of at most kBatchCap bytes (each one
// level-0 run over their concatenation); a text above the cap runs alone.
static const u64 kBatchCap = 1ull << 30;

static void batch_groups(const uint64_t *lens, uint64_t count,
                         std::vector<std::pair<u64, u64>> &groups) {
    u64 i = 0;
    while (i < count) {
        u64 j = i, tot = 0;
        while (j < count && (j == i || tot + lens[j] <= kBatchCap)) tot += lens[j++];
        groups.push_back({i, j});
        i = j;
    }
}

int psk_batch8(psk_ctx *ctx, const uint8_t *const *texts, const uint64_t *lens,
               uint32_t *const *sas, uint64_t count) {
    if (count && (!texts || !lens || !sas)) return PSK_EINVAL;
    for (uint64_t i = 0; i < count; i++) {
        int rc = check8(texts[i], lens[i], sas[i], lens[i]);
        if (rc != PSK_OK) return rc;
    }
    if (count == 0) return PSK_OK;
    psk_ctx *c = ctx ? ctx : default_ctx();
    return guarded(c, [&] {
        ensure_ready(c);
        std::vector<std::pair<u64, u64>> groups;
        batch_groups(lens, count, groups);
        Timer t(c);
        for (auto &g : groups) {
            std::vector<u64> off(g.second - g.first + 1, 0);
            for (u64 i = g.first; i < g.second; i++) off[i - g.first + 1] = off[i - g.first] + lens[i];
            const u64 N = off.back(), K = g.second - g.first;
            if (N == 0) continue;
            ArenaScope mk(c->arena);
            u8 *dt = c->arena.get<u8>(N);
            u32 *ds = c->arena.get<u32>(N);
            std::vector<Seg> tin, sout;
            for (u64 i = 0; i < K; i++) {
                tin.push_back(Seg{(char *)texts[g.first + i], lens[g.first + i]});
                sout.push_back(Seg{(char *)sas[g.first + i], lens[g.first + i] * sizeof(u32)});
            }
            upload_segs(c->rt.d, c->xfer, tin, dt);
            if (K == 1) psacak8_run(c->rt, dt, N, ds);
            else batch8_run(c->rt, dt, off.data(), K, ds);
            download_segs(c->rt.d, c->xfer, ds, sout);
        }
        t.stop();
    });
}

int psk_batch8_dev(psk_ctx *ctx, const uint8_t *d_texts, const uint64_t *offsets, uint64_t count,
                   uint32_t *d_sas, void *stream) {
    if (count && (!d_texts || !offsets || !d_sas)) return PSK_EINVAL;
    std::vector<uint64_t> lens(count);
    for (uint64_t i = 0; i < count; i++) {
        if (offsets[i + 1] < offsets[i]) return PSK_EINVAL;
        lens[i] = offsets[i + 1] - offsets[i];
        if (lens[i] >= 0xFFFFFFFFull) return PSK_ETOOLONG;
    }
    if (count == 0) return PSK_OK;
    psk_ctx *c = ctx ? ctx : default_ctx();
    return guarded(c, [&] {
        ensure_ready(c);
        on_stream(c, stream, [&] {
            std::vector<std::pair<u64, u64>> groups;
            batch_groups(lens.data(), count, groups);
            Timer t(c);
            for (auto &g : groups) {
                const u64 base = offsets[g.first], K = g.second - g.first;
                std::vector<u64> off(K + 1);
                for (u64 i = 0; i <= K; i++) off[i] = offsets[g.first + i] - base;
                if (off[K] == 0) continue;
                if (K == 1) psacak8_run(c->rt, d_texts + base, off[K], d_sas + base);
                else batch8_run(c->rt, d_texts + base, off.data(), K, d_sas + base);
            }
            t.stop();
        });
    });
}

int psk_check8_dev(psk_ctx *ctx, const uint8_t *d_text, uint64_t n, const uint32_t *d_sa,
                   int *ok, void *stream) {
    if (!ok) return PSK_EINVAL;
    *ok = 0;
    int rc = check8(d_text, n, d_sa, n);
    if (rc != PSK_OK) return rc;
    if (n == 0) { *ok = 1; return PSK_OK; }
    psk_ctx *c = ctx ? ctx : default_ctx();
    return guarded(c, [&] {
        ensure_ready(c);
        on_stream(c, stream, [&] {
            Dev &d = c->rt.d;
            ArenaScope mk(*d.a);
            u32 *rank = d.a->get<u32>(n);
            u32 *bad = d.a->get<u32>(1);
            d.zero(bad, sizeof(u32));
            check_sa_dev(d, d_text, n, d_sa, rank, bad);
            *ok = d.read_u32(bad) == 0;
        });
    });
}

// ---- zero-allocation entry point (state above the C ABI block) -------------
int psk_psacak8_dev_noalloc(const uint8_t *d_text, uint64_t n, uint32_t *d_sa,
                                       void *stream) {
    int rc = check8(d_text, n, d_sa, n);
    if (rc != PSK_OK || n == 0) return rc;
    if (n > PSK_NOALLOC_MAX_N) return PSK_ETOOLONG;
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess) {
        (void)hipGetLastError();
        return PSK_EDEVICE;
    }
    if (dev < 0 || dev >= kNoAllocMaxDevices) return PSK_EDEVICE;
    NoAllocState &g = g_noalloc[dev];
    std::lock_guard<std::mutex> lk(g.mu);
    return guarded(nullptr, [&] {
        noalloc_setup(g);
        Dev &d = g.rt.d;
        d.s = (hipStream_t)stream;
        ArenaScope sc(g.arena);
        // the array keeps the last call's contents: the flag starts at zero
        d.zero(d.err, sizeof(u32));
        psacak8_run(g.rt, d_text, n, d_sa);
        const u32 e = d.read_u32(d.err);
        if (e) fail(PSK_EDEVICE, "internal index check failed (flags " + std::to_string(e) + ")");
    });
}

int psk_check32_dev(psk_ctx *ctx, const uint32_t *d_text, uint64_t n, const uint32_t *d_sa,
                    int *ok, void *stream) {
    if (!ok) return PSK_EINVAL;
    *ok = 0;
    int rc = check32(d_text, n, d_sa, n);
    if (rc != PSK_OK) return rc;
    if (n == 0) { *ok = 1; return PSK_OK; }
    psk_ctx *c = ctx ? ctx : default_ctx();
    return guarded(c, [&] {
        ensure_ready(c);
        on_stream(c, stream, [&] {
            Dev &d = c->rt.d;
            ArenaScope mk(*d.a);
            u32 *rank = d.a->get<u32>(n);
            u32 *bad = d.a->get<u32>(1);
            d.zero(bad, sizeof(u32));
            check_sa_dev(d, d_text, n, d_sa, rank, bad);
            *ok = d.read_u32(bad) == 0;
        });
    });
}

// ---- serving ----------------------------------------------------------------

uint64_t psk_server_footprint(uint64_t bytes, int batched) {
    // text + SA + level workspace (the workers' contexts are lean), plus the
    // text-id partition of a batch
    return psk::lean_bytes8(bytes) + 5 * bytes + (batched ? 24 * bytes : 0);
}

// Worker loop.  psk_ctx_trim frees device memory, and hipFree waits for the
// whole device: it is always called with the server lock released, so the
// other workers keep admitting, finishing and launching jobs meanwhile.
static void server_worker(psk_server *s, int w) {
    psk_ctx *c = s->ctxs[w];
    for (;;) {
        std::vector<SrvReq *> job;
        u64 foot = 0;
        bool trim = false;  // give back this worker's cached chunks first
        {
            std::unique_lock<std::mutex> lk(s->mu);
            for (;;) {
                if (s->stop) return;
                if (!s->pending.empty()) {
                    SrvReq *head = s->reqs[s->pending.front()].get();
                    std::vector<SrvReq *> cand{head};
                    u64 tot = head->n;
                    if (head->n <= psk_server::kSmall) {
                        for (size_t i = 1; i < s->pending.size(); i++) {
                            SrvReq *r = s->reqs[s->pending[i]].get();
                            if (r->n > psk_server::kSmall || tot + r->n > psk_server::kBatchBytes) break;
                            cand.push_back(r);
                            tot += r->n;
                        }
                    }
                    u64 f = psk_server_footprint(tot, cand.size() > 1);
                    // device memory the other workers hold: their arenas, or
                    // their running job's footprint if larger; idle_cache: what
                    // idle workers keep cached
                    u64 others = 0, idle_cache = 0;
                    for (int v = 0; v < s->nworkers; v++) {
                        if (v == w) continue;
                        const u64 r = s->ctxs[v]->arena.reserved();
                        others += std::max(r, s->wfoot[v]);
                        if (!s->wfoot[v]) idle_cache += r;
                    }
                    const u64 mine = c->arena.reserved();
                    bool admit = s->budget == 0 || others + std::max(mine, f) <= s->budget;
                    if (!admit && mine && others + f <= s->budget) {
                        // this worker's cached chunks are what stands in the
                        // way: take the job, trim before running it
                        trim = true;
                        admit = true;
                    }
                    if (admit) {
                        for (size_t i = 0; i < cand.size(); i++) s->pending.pop_front();
                        for (SrvReq *r : cand) {
                            if (r->id <= s->last_admitted) s->order_violations++;
                            s->last_admitted = r->id;
                        }
                        s->n_jobs++;
                        s->n_admitted += cand.size();
                        s->inflight += f;
                        s->max_inflight = std::max(s->max_inflight, s->inflight);
                        s->wfoot[w] = f;
                        foot = f;
                        job.swap(cand);
                        break;
                    }
                    s->n_blocked++;
                    if (mine) {
                        // idle and blocked: give back the cache (below, with
                        // the lock released), then look again
                        trim = true;
                        break;
                    }
                    // idle workers' caches can be given back: wake them
                    if (idle_cache) s->cv_work.notify_all();
                }
                s->cv_work.wait(lk);
            }
        }
        if (trim) {
            {
                std::lock_guard<std::mutex> lk(s->mu);
                s->n_trims++;
            }
            (void)psk_ctx_trim(c);
            if (job.empty()) {
                s->cv_work.notify_all();
                continue;
            }
        }
        int rc;
        if (job.size() == 1) {
            SrvReq *r = job[0];
            rc = psk_psacak_inplace(c, r->tp, r->n, r->sp, r->n);
        } else {
            std::vector<const uint8_t *> tp;
            std::vector<uint64_t> lp;
            std::vector<uint32_t *> sp;
            for (SrvReq *r : job) {
                tp.push_back(r->tp);
                lp.push_back(r->n);
                sp.push_back(r->sp);
            }
            rc = psk_batch8(c, tp.data(), lp.data(), sp.data(), job.size());
        }
        double t = now_ms();
        bool trim_after = false;
        {
            std::lock_guard<std::mutex> lk(s->mu);
            for (SrvReq *r : job) {
                r->status = rc;
                r->done = true;
                r->t_done = t;
                std::vector<u8>().swap(r->text);
            }
            s->inflight -= foot;
            s->wfoot[w] = 0;
            u64 res = 0;
            for (psk_ctx *x : s->ctxs) res += x->arena.reserved();
            s->peak = std::max(s->peak, res);
            // cached chunks of finished jobs are not in flight: a worker keeps
            // at most its share of the budget between jobs
            trim_after = s->budget && (res > s->budget || c->arena.reserved() > s->budget / s->nworkers);
        }
        s->cv_done.notify_all();
        if (trim_after) (void)psk_ctx_trim(c);
        s->cv_work.notify_all();
    }
}

int psk_server_create(psk_server **out, uint64_t hbm_budget_bytes, int max_workers) {
    if (!out || max_workers < 1 || max_workers > 8) return PSK_EINVAL;
    psk_server *s = new (std::nothrow) psk_server();
    if (!s) return PSK_ENOMEM;
    s->budget = hbm_budget_bytes;
    s->nworkers = max_workers;
    for (int w = 0; w < max_workers; w++) {
        psk_ctx *c = nullptr;
        int rc = psk_ctx_create(&c, hbm_budget_bytes);
        if (rc != PSK_OK) {
            for (psk_ctx *x : s->ctxs) psk_ctx_destroy(x);
            delete s;
            return rc;
        }
        // lean contexts: no placement records.  Workers take jobs of every
        // size one after another, and each size lays its chunks out
        // differently, so every byte a job does not need is a byte of
        // fragmentation across jobs (and trims, which wait for the device).
        c->rt.lms_records = false;
        s->ctxs.push_back(c);
    }
    s->wfoot.assign(max_workers, 0);
    for (int w = 0; w < max_workers; w++) s->workers.emplace_back(server_worker, s, w);
    *out = s;
    return PSK_OK;
}

static int submit_req(psk_server *s, const uint8_t *text, uint64_t n, uint32_t *sa, uint64_t *id) {
    if (!s || !id || (n && !text)) return PSK_EINVAL;
    if (n >= 0xFFFFFFFFull) return PSK_ETOOLONG;
    u64 f = psk_server_footprint(n, 0);
    if (s->budget && f > s->budget) return PSK_EBUDGET;
    std::unique_ptr<SrvReq> r(new (std::nothrow) SrvReq());
    if (!r) return PSK_ENOMEM;
    if (sa) {
        r->tp = text;
        r->sp = sa;
        r->owned = false;
    } else {
        try {
            r->text.assign(text, text + n);
            r->sa.resize(n);
        } catch (...) {
            return PSK_ENOMEM;
        }
        r->tp = r->text.data();
        r->sp = r->sa.data();
    }
    r->n = n;
    r->t_submit = now_ms();
    {
        std::lock_guard<std::mutex> lk(s->mu);
        r->id = s->next_id++;
        *id = r->id;
        s->pending.push_back(r->id);
        s->reqs[r->id] = std::move(r);
    }
    s->cv_work.notify_all();
    return PSK_OK;
}

int psk_server_submit(psk_server *s, const uint8_t *text, uint64_t n, uint64_t *id) {
    return submit_req(s, text, n, nullptr, id);
}

int psk_server_submit_into(psk_server *s, const uint8_t *text, uint64_t n, uint32_t *sa,
                           uint64_t *id) {
    if (n && !sa) return PSK_EINVAL;
    return submit_req(s, text, n, n ? sa : nullptr, id);
}

static int take_result(psk_server *s, std::unique_ptr<SrvReq> &r, uint32_t *sa, double *lat) {
    if (sa && r->owned && r->n) memcpy(sa, r->sa.data(), r->n * sizeof(u32));
    if (lat) *lat = r->t_done - r->t_submit;
    int st = r->status;
    s->reqs.erase(r->id);
    s->cv_done.notify_all();
    return st;
}

int psk_server_poll(psk_server *s, uint64_t id, int *done, uint32_t *sa, double *latency_ms) {
    if (!s || !done) return PSK_EINVAL;
    *done = 0;
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->reqs.find(id);
    if (it == s->reqs.end()) return PSK_ENOTFOUND;
    *done = it->second->done ? 1 : 0;
    if (!*done) return PSK_OK;
    int st = take_result(s, it->second, sa, latency_ms);
    // waiters on the same id wake up and find it gone
    s->cv_done.notify_all();
    return st;
}

int psk_server_wait(psk_server *s, uint64_t id, uint32_t *sa, double *latency_ms,
                    double timeout_ms) {
    if (!s) return PSK_EINVAL;
    std::unique_lock<std::mutex> lk(s->mu);
    if (s->reqs.find(id) == s->reqs.end()) return PSK_ENOTFOUND;
    // the request is looked up again on every wake-up: a concurrent poll or
    // wait on the same id may have taken its result (and released it) while
    // this caller slept
    auto ready = [&] {
        auto it = s->reqs.find(id);
        return it == s->reqs.end() || it->second->done;
    };
    if (timeout_ms < 0) s->cv_done.wait(lk, ready);
    else if (!s->cv_done.wait_for(lk, std::chrono::duration<double, std::milli>(timeout_ms), ready))
        return PSK_EBUSY;
    auto it = s->reqs.find(id);
    if (it == s->reqs.end()) return PSK_ENOTFOUND;
    return take_result(s, it->second, sa, latency_ms);
}

uint64_t psk_server_peak_bytes(psk_server *s) {
    if (!s) return 0;
    std::lock_guard<std::mutex> lk(s->mu);
    return s->peak;
}

int psk_server_stats(psk_server *s, uint64_t *out, int n) {
    if (!s || !out || n < 0) return PSK_EINVAL;
    std::lock_guard<std::mutex> lk(s->mu);
    const u64 v[7] = {s->n_jobs, s->n_admitted, s->n_blocked, s->n_trims,
                      s->max_inflight, s->order_violations, s->peak};
    const int m = std::min(n, 7);
    for (int i = 0; i < m; i++) out[i] = v[i];
    return m;
}

uint64_t psk_server_pending(psk_server *s) {
    if (!s) return 0;
    std::lock_guard<std::mutex> lk(s->mu);
    u64 k = 0;
    for (auto &kv : s->reqs) k += !kv.second->done;
    return k;
}

void psk_server_destroy(psk_server *s) {
    if (!s) return;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
    }
    s->cv_work.notify_all();
    for (auto &t : s->workers) t.join();
    for (psk_ctx *c : s->ctxs) psk_ctx_destroy(c);
    delete s;
}

}  // extern "C"
