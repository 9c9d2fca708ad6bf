// psk_xfer.hpp -- host <-> device transfers of the host-buffer entry points.
//
// The caller's buffers are ordinary (pageable) memory.  Large transfers go
// through two pinned staging chunks per context: while the DMA engine moves
// one chunk, host threads copy the next (H2D) or previous (D2H) chunk between
// the caller's buffers and the other staging chunk.  A transfer is a list of
// host segments that are contiguous on the device (the texts of a batch).
#pragma once

#include <thread>

#include "psk_prim.hpp"

namespace psk {

struct Seg {
    char *host;
    u64 bytes;
};

constexpr u64 XFER_CHUNK = 64ull << 20;
constexpr u64 XFER_DIRECT = 8ull << 20;  // below this a plain copy is cheaper
constexpr int XFER_THREADS = 4;

struct Xfer {
    char *stage[2] = {nullptr, nullptr};
    hipEvent_t ev[2] = {nullptr, nullptr};

    void release() {
        for (int i = 0; i < 2; i++) {
            if (stage[i]) (void)hipHostFree(stage[i]);
            if (ev[i]) (void)hipEventDestroy(ev[i]);
            stage[i] = nullptr;
            ev[i] = nullptr;
        }
    }
    void ready() {
        for (int i = 0; i < 2; i++) {
            if (!stage[i]) {
                if (inject_alloc_failure()) fail(PSK_ENOMEM, "pinned staging allocation failed (injected)");
                PSK_HIP(hipHostMalloc((void **)&stage[i], XFER_CHUNK, hipHostMallocDefault));
            }
            if (!ev[i]) PSK_HIP(hipEventCreateWithFlags(&ev[i], hipEventDisableTiming));
        }
    }
};

// Copy the bytes [off, off + len) of the concatenation of segs to (to_host)
// or from the contiguous buffer buf, split over XFER_THREADS threads.
inline void seg_copy(const std::vector<Seg> &segs, u64 off, u64 len, char *buf, bool to_host) {
    auto part = [&](u64 a, u64 b) {  // bytes [a, b) of the range
        u64 pos = 0;
        for (const Seg &s : segs) {
            const u64 s0 = pos, s1 = pos + s.bytes;
            pos = s1;
            const u64 lo = std::max(s0, off + a), hi = std::min(s1, off + b);
            if (lo >= hi) continue;
            char *h = s.host + (lo - s0);
            char *g = buf + (lo - off);
            if (to_host) memcpy(h, g, hi - lo);
            else memcpy(g, h, hi - lo);
            if (s1 >= off + b) break;
        }
    };
    if (len < (4ull << 20)) {
        part(0, len);
        return;
    }
    std::thread th[XFER_THREADS - 1];
    const u64 step = ((len + XFER_THREADS - 1) / XFER_THREADS + 4095) & ~4095ull;
    for (int i = 1; i < XFER_THREADS; i++) {
        const u64 a = std::min(len, i * step), b = std::min(len, (i + 1) * step);
        th[i - 1] = std::thread(part, a, b);
    }
    part(0, std::min(len, step));
    for (auto &t : th) t.join();
}

inline u64 seg_total(const std::vector<Seg> &segs) {
    u64 t = 0;
    for (const Seg &s : segs) t += s.bytes;
    return t;
}

// Host segments -> device buffer d (synchronous on return).
inline void upload_segs(Dev &dv, Xfer &x, const std::vector<Seg> &segs, void *d) {
    const u64 n = seg_total(segs);
    if (n == 0) return;
    if (n < XFER_DIRECT) {
        u64 pos = 0;
        for (const Seg &s : segs) {
            if (s.bytes)
                PSK_HIP(hipMemcpyAsync((char *)d + pos, s.host, s.bytes, hipMemcpyHostToDevice, dv.s));
            pos += s.bytes;
        }
        PSK_HIP(hipStreamSynchronize(dv.s));
        return;
    }
    x.ready();
    u64 k = 0;
    for (u64 off = 0; off < n; off += XFER_CHUNK, k++) {
        const int b = (int)(k & 1);
        const u64 len = std::min(XFER_CHUNK, n - off);
        if (k >= 2) PSK_HIP(hipEventSynchronize(x.ev[b]));  // chunk k-2 has left this buffer
        seg_copy(segs, off, len, x.stage[b], false);
        PSK_HIP(hipMemcpyAsync((char *)d + off, x.stage[b], len, hipMemcpyHostToDevice, dv.s));
        PSK_HIP(hipEventRecord(x.ev[b], dv.s));
    }
    PSK_HIP(hipStreamSynchronize(dv.s));
}

// Device buffer d -> host segments (synchronous on return; waits for the
// work queued on the stream before it).
inline void download_segs(Dev &dv, Xfer &x, const void *d, const std::vector<Seg> &segs) {
    const u64 n = seg_total(segs);
    if (n == 0) {
        PSK_HIP(hipStreamSynchronize(dv.s));
        return;
    }
    if (n < XFER_DIRECT) {
        u64 pos = 0;
        for (const Seg &s : segs) {
            if (s.bytes)
                PSK_HIP(hipMemcpyAsync(s.host, (const char *)d + pos, s.bytes, hipMemcpyDeviceToHost,
                                       dv.s));
            pos += s.bytes;
        }
        PSK_HIP(hipStreamSynchronize(dv.s));
        return;
    }
    x.ready();
    const u64 nk = (n + XFER_CHUNK - 1) / XFER_CHUNK;
    for (u64 k = 0; k <= nk; k++) {
        if (k < nk) {
            // chunk k into buffer k&1: its previous contents (chunk k-2) were
            // copied out in iteration k-1
            const u64 off = k * XFER_CHUNK, len = std::min(XFER_CHUNK, n - off);
            PSK_HIP(hipMemcpyAsync(x.stage[k & 1], (const char *)d + off, len, hipMemcpyDeviceToHost,
                                   dv.s));
            PSK_HIP(hipEventRecord(x.ev[k & 1], dv.s));
        }
        if (k >= 1) {
            const u64 j = k - 1, off = j * XFER_CHUNK, len = std::min(XFER_CHUNK, n - off);
            PSK_HIP(hipEventSynchronize(x.ev[j & 1]));
            seg_copy(segs, off, len, x.stage[j & 1], true);
        }
    }
}

}  // namespace psk
