// psk_common.hpp -- shared types, error handling, wave64 helpers and the
// device-memory arena of psacak_amd (gfx950).
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <type_traits>
#include <vector>

#include "../../include/psacak.h"

namespace psk {

typedef uint8_t u8;
typedef uint32_t u32;
typedef uint64_t u64;
typedef int64_t i64;

constexpr int WAVE = 64;     // CDNA wavefront width
constexpr int BLOCK = 256;   // 4 waves per workgroup
constexpr int NWAVES = BLOCK / WAVE;

struct Error {
    int code;
    std::string msg;
};

[[noreturn]] inline void fail(int code, const std::string &msg) { throw Error{code, msg}; }

#define PSK_HIP(x)                                                                          \
    do {                                                                                    \
        hipError_t e_ = (x);                                                                \
        if (e_ != hipSuccess)                                                               \
            ::psk::fail(PSK_EDEVICE, std::string(#x) + " failed: " + hipGetErrorString(e_)); \
    } while (0)

#define PSK_LAUNCH_CHECK() PSK_HIP(hipGetLastError())

inline u32 ceil_div(u64 a, u64 b) { return (u32)((a + b - 1) / b); }

// ---- wave64 helpers --------------------------------------------------------
__device__ __forceinline__ int lane_id() { return (int)(threadIdx.x & (WAVE - 1)); }
__device__ __forceinline__ int wave_id() { return (int)(threadIdx.x / WAVE); }
__device__ __forceinline__ u64 lanemask_lt() { return (1ull << lane_id()) - 1ull; }
__device__ __forceinline__ u64 ballot(bool p) { return (u64)__ballot(p); }

// Mask of lanes whose (valid, 8-bit key) equals this lane's: 8 ballots.
__device__ __forceinline__ u64 match_key8(bool valid, u32 key) {
    u64 m = ballot(valid);
    #pragma unroll
    for (int b = 0; b < 8; b++) {
        bool bit = (key >> b) & 1u;
        u64 v = ballot(bit);
        m &= bit ? v : ~v;
    }
    return valid ? m : 0ull;
}

template <typename T>
__device__ __forceinline__ T wave_inclusive_sum(T x) {
    const int l = lane_id();
    #pragma unroll
    for (int o = 1; o < WAVE; o <<= 1) {
        T y = __shfl_up(x, o, WAVE);
        if (l >= o) x += y;
    }
    return x;
}

// Maximum over the lanes of the wave (all lanes must be active).
__device__ __forceinline__ u32 wave_max(u32 x) {
    #pragma unroll
    for (int o = WAVE / 2; o > 0; o >>= 1) {
        u32 y = __shfl_xor(x, o, WAVE);
        x = x > y ? x : y;
    }
    return x;
}

template <typename T>
__device__ __forceinline__ T wave_inclusive_max(T x) {
    const int l = lane_id();
    #pragma unroll
    for (int o = 1; o < WAVE; o <<= 1) {
        T y = __shfl_up(x, o, WAVE);
        if (l >= o) x = x > y ? x : y;
    }
    return x;
}

// Block-wide exclusive sum of one value per thread.  lds needs NWAVES+1
// entries.  Returns the exclusive prefix; *total receives the block sum.
template <typename T>
__device__ __forceinline__ T block_exclusive_sum(T v, T *lds, T *total) {
    T inc = wave_inclusive_sum(v);
    if (lane_id() == WAVE - 1) lds[wave_id()] = inc;
    __syncthreads();
    if (threadIdx.x == 0) {
        T acc = 0;
        for (int w = 0; w < NWAVES; w++) { T t = lds[w]; lds[w] = acc; acc += t; }
        lds[NWAVES] = acc;
    }
    __syncthreads();
    T r = inc - v + lds[wave_id()];
    *total = lds[NWAVES];
    __syncthreads();
    return r;
}

// ---- fault injection ------------------------------------------------------------
// PSACAK_AMD_FAIL_ALLOC=k (k >= 1) in the environment makes every k-th
// workspace chunk allocation of the process fail as if hipMalloc had (the
// count restarts whenever the value changes): for tests that every failure
// surfaces as a status code and leaves the context usable.  Unset or 0: off.
inline bool inject_alloc_failure() {
    static std::atomic<long> count{0};
    static std::atomic<long> period{0};
    const char *e = getenv("PSACAK_AMD_FAIL_ALLOC");
    const long k = e ? atol(e) : 0;
    if (k != period.load()) {
        period.store(k);
        count.store(0);
    }
    if (k <= 0) return false;
    return (count.fetch_add(1) + 1) % k == 0;
}

// ---- device arena -------------------------------------------------------------
// Stack allocator over large hipMalloc'd chunks that are kept for reuse.
// mark()/reset() pop everything allocated after the mark; ArenaScope does it
// when the scope ends, on every path out of it (exceptions included), so a
// failed call leaves the arena where the call found it.
// A fixed arena (attach) allocates from one caller-provided region and never
// calls hipMalloc: an allocation that does not fit fails with PSK_EBUDGET.
class Arena {
  public:
    struct Mark { size_t chunk, top; };

    ~Arena() { release(); }

    void set_budget(u64 b) { budget_ = b; }
    // readable from other threads (a server sums its workers' arenas)
    u64 reserved() const { return reserved_.load(std::memory_order_relaxed); }
    u64 peak() const { return peak_; }
    // highest stack top reached in a fixed arena (bytes from its base)
    u64 fixed_peak() const { return fixed_peak_; }
    void clear_fixed_peak() { fixed_peak_ = 0; }

    Mark mark() const { return {cur_, top_}; }
    void reset(Mark m) { cur_ = m.chunk; top_ = m.top; }

    // allocate from [base, base + cap) only (device memory the caller owns)
    void attach(void *base, u64 cap) {
        release();
        fixed_ = true;
        chunks_.push_back({(char *)base, cap});
        cur_ = top_ = 0;
    }

    void *alloc(u64 bytes) {
        bytes = (bytes + 255) & ~(u64)255;
        if (bytes == 0) bytes = 256;
        if (fixed_) {
            if (chunks_.empty() || top_ + bytes > chunks_[0].cap)
                fail(PSK_EBUDGET, "device workspace exceeds the fixed scratch region (" +
                                      std::to_string(chunks_.empty() ? 0 : chunks_[0].cap) +
                                      " bytes)");
            void *p = chunks_[0].p + top_;
            top_ += bytes;
            fixed_peak_ = std::max(fixed_peak_, top_);
            return p;
        }
        while (cur_ < chunks_.size()) {
            if (top_ + bytes <= chunks_[cur_].cap) {
                void *p = chunks_[cur_].p + top_;
                top_ += bytes;
                return p;
            }
            cur_++;
            top_ = 0;
        }
        u64 cap = std::max<u64>(bytes, min_chunk_);
        const u64 res = reserved();
        if (budget_ && res + cap > budget_) {
            cap = bytes;
            if (res + cap > budget_)
                fail(PSK_EBUDGET, "device workspace exceeds the HBM budget");
        }
        char *p = nullptr;
        if (inject_alloc_failure())
            fail(PSK_ENOMEM, "hipMalloc of " + std::to_string(cap) + " bytes failed (injected)");
        if (hipMalloc((void **)&p, cap) != hipSuccess) {
            (void)hipGetLastError();
            fail(PSK_ENOMEM, "hipMalloc of " + std::to_string(cap) + " bytes failed");
        }
        chunks_.push_back({p, cap});
        reserved_.store(res + cap, std::memory_order_relaxed);
        peak_ = std::max(peak_, res + cap);
        cur_ = chunks_.size() - 1;
        top_ = bytes;
        return p;
    }

    template <typename T>
    T *get(u64 count) { return (T *)alloc(count * sizeof(T)); }

    void release() {
        if (!fixed_)
            for (auto &c : chunks_) (void)hipFree(c.p);
        chunks_.clear();
        reserved_.store(0, std::memory_order_relaxed);
        cur_ = top_ = 0;
        fixed_ = false;
    }

  private:
    struct Chunk { char *p; u64 cap; };
    std::vector<Chunk> chunks_;
    size_t cur_ = 0;
    u64 top_ = 0;
    std::atomic<u64> reserved_{0};
    u64 peak_ = 0, budget_ = 0;
    u64 min_chunk_ = 64ull << 20;
    bool fixed_ = false;
    u64 fixed_peak_ = 0;
};

// Pops the arena back to where it stood at construction when the scope ends
// (or at reset(), which may be called early; the destructor's reset is then a
// no-op unless more was allocated since).
class ArenaScope {
  public:
    explicit ArenaScope(Arena &a) : a_(a), m_(a.mark()) {}
    ~ArenaScope() { a_.reset(m_); }
    void reset() { a_.reset(m_); }
    ArenaScope(const ArenaScope &) = delete;
    ArenaScope &operator=(const ArenaScope &) = delete;

  private:
    Arena &a_;
    Arena::Mark m_;
};

// ---- kernel profiler ---------------------------------------------------------
// When enabled, tagged launches are bracketed by HIP events on their stream;
// each record carries the launch's algorithmic bytes (static, known on the
// host) plus a device-side u64 slot that kernels with data-dependent traffic
// add to.  Nothing synchronises until stats are collected.
class Prof {
  public:
    struct Total { std::string name; u64 launches = 0; double ms = 0, bytes = 0; };

    bool on = false;

    ~Prof() { release(); }

    // tags to record (empty: all); a launch of another tag takes no events
    std::vector<std::string> only;

    // returns the record index, or -1 when profiling is off
    int begin(hipStream_t s, const char *name, double static_bytes) {
        if (!on) return -1;
        if (!only.empty() && std::find(only.begin(), only.end(), std::string(name)) == only.end())
            return -1;
        if (!dyn_) {
            if (hipMalloc((void **)&dyn_, kMaxRecs * sizeof(u64)) != hipSuccess) { on = false; return -1; }
            if (hipMemset(dyn_, 0, kMaxRecs * sizeof(u64)) != hipSuccess) { on = false; return -1; }
        }
        if (recs_.size() >= kMaxRecs) flush();
        Rec r;
        r.name = name;
        r.bytes = static_bytes;
        r.a = event();
        r.b = event();
        r.s = s;
        (void)hipEventRecord(r.a, s);
        recs_.push_back(r);
        return (int)recs_.size() - 1;
    }
    void end(int rec) {
        if (rec < 0) return;
        (void)hipEventRecord(recs_[rec].b, recs_[rec].s);
    }
    u64 *dyn_slot(int rec) { return rec < 0 ? nullptr : dyn_ + rec; }
    // bytes of a pending record known on the host only after its launch
    void add(int rec, double bytes) {
        if (rec >= 0 && (size_t)rec < recs_.size()) recs_[rec].bytes += bytes;
    }

    // synchronise and fold the pending records into the totals
    void flush() {
        if (recs_.empty()) return;
        std::vector<u64> dyn(recs_.size(), 0);
        // the streams, not only the end events: data-dependent byte slots are
        // filled by kernels queued after a record's end event
        std::vector<hipStream_t> seen;
        for (auto &r : recs_)
            if (std::find(seen.begin(), seen.end(), r.s) == seen.end()) seen.push_back(r.s);
        for (auto s : seen) (void)hipStreamSynchronize(s);
        (void)hipMemcpy(dyn.data(), dyn_, recs_.size() * sizeof(u64), hipMemcpyDeviceToHost);
        (void)hipMemset(dyn_, 0, recs_.size() * sizeof(u64));
        for (size_t i = 0; i < recs_.size(); i++) {
            Rec &r = recs_[i];
            float ms = 0;
            (void)hipEventElapsedTime(&ms, r.a, r.b);
            Total &t = total(r.name);
            t.launches++;
            t.ms += ms;
            t.bytes += r.bytes + (double)dyn[i];
            pool_.push_back(r.a);
            pool_.push_back(r.b);
        }
        recs_.clear();
    }
    const std::vector<Total> &totals() { flush(); return totals_; }
    void clear() { flush(); totals_.clear(); }
    void release() {
        recs_.clear();
        for (auto e : pool_) (void)hipEventDestroy(e);
        pool_.clear();
        if (dyn_) (void)hipFree(dyn_);
        dyn_ = nullptr;
    }

  private:
    static constexpr size_t kMaxRecs = 8192;
    struct Rec { std::string name; double bytes; hipEvent_t a, b; hipStream_t s; };
    hipEvent_t event() {
        if (!pool_.empty()) { hipEvent_t e = pool_.back(); pool_.pop_back(); return e; }
        hipEvent_t e = nullptr;
        (void)hipEventCreate(&e);
        return e;
    }
    Total &total(const std::string &name) {
        for (auto &t : totals_) if (t.name == name) return t;
        totals_.push_back(Total{name, 0, 0, 0});
        return totals_.back();
    }
    std::vector<Rec> recs_;
    std::vector<hipEvent_t> pool_;
    std::vector<Total> totals_;
    u64 *dyn_ = nullptr;
};

}  // namespace psk
