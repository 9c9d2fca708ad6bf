// psk_server.hpp -- continuous-batching request server (C ABI psk_server_*).
//
// Requests arrive at any time (psk_server_submit copies the text).  Worker
// threads, each with its own context and HIP stream, take admitted work in
// arrival order: consecutive small requests are coalesced into one batch
// (one shared level-0 run, batch8_run), a large request runs alone.  A job is
// admitted only while the estimated device footprint of all jobs in flight
// stays inside the HBM budget, so requests that arrive while others are on
// the GPU start as soon as a worker and memory are free, and jobs of
// different workers overlap on the device.  Results are identical to a
// single call's: batching never changes an SA (every text keeps its own
// sentinel).
#pragma once

#include <chrono>
#include <condition_variable>
#include <deque>
#include <memory>
#include <mutex>
#include <thread>
#include <unordered_map>

struct psk_server;

namespace psk {

inline double now_ms() {
    using namespace std::chrono;
    return duration<double, std::milli>(steady_clock::now().time_since_epoch()).count();
}

struct SrvReq {
    u64 id = 0, n = 0, foot = 0;
    std::vector<u8> text;  // owned copies (psk_server_submit)
    std::vector<u32> sa;
    const u8 *tp = nullptr;  // the text and SA buffers the worker uses: the
    u32 *sp = nullptr;       // owned copies, or the caller's (submit_into)
    bool owned = true;
    double t_submit = 0, t_done = 0;
    int status = PSK_OK;
    bool done = false;
};

}  // namespace psk

struct psk_server {
    psk::u64 budget = 0;
    int nworkers = 1;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<psk::u64> pending;
    std::unordered_map<psk::u64, std::unique_ptr<psk::SrvReq>> reqs;
    psk::u64 next_id = 1, inflight = 0, peak = 0;
    bool stop = false;
    std::vector<std::thread> workers;
    std::vector<psk_ctx *> ctxs;
    std::vector<psk::u64> wfoot;  // footprint of each worker's running job (0: idle)
    // counters (psk_server_stats): jobs admitted, requests in them, times a
    // worker found the head request blocked by the budget, trims, the most
    // footprint in flight at once, and admissions out of arrival order
    psk::u64 n_jobs = 0, n_admitted = 0, n_blocked = 0, n_trims = 0, max_inflight = 0,
             order_violations = 0, last_admitted = 0;
    // requests up to this size are coalesced; a coalesced batch holds at
    // most kBatchBytes of text
    static constexpr psk::u64 kSmall = 16ull << 20;
#ifndef PSK_SRV_BATCH_MB
#define PSK_SRV_BATCH_MB 512
#endif
    static constexpr psk::u64 kBatchBytes = (psk::u64)PSK_SRV_BATCH_MB << 20;
};
