// Host-side exercise of the C ABI for sanitizer builds (AddressSanitizer and
// UndefinedBehaviorSanitizer on the host code of the library, linked into
// this executable).  Without a GPU every device call must come back as
// PSK_EDEVICE; argument checks, contexts, capture settings, workspace bounds
// and the server (worker threads, admission, concurrent waits) run on the
// host alone.  Exit status 0 when every expectation holds.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <thread>
#include <vector>

#include "../../include/psacak.h"

static int failures = 0;
#define EXPECT(c)                                                        \
    do {                                                                 \
        if (!(c)) {                                                      \
            std::fprintf(stderr, "%s:%d: expectation failed: %s\n", __FILE__, __LINE__, #c); \
            failures++;                                                  \
        }                                                                \
    } while (0)

int main() {
    std::vector<uint8_t> text(4096);
    for (size_t i = 0; i < text.size(); i++) text[i] = (uint8_t)("ACGT"[(i * 7919u) % 4]);
    std::vector<uint32_t> sa(text.size() + 8, 7u);
    const int no_gpu = psk_psacak_inplace(nullptr, text.data(), text.size(), sa.data(), sa.size());
    EXPECT(no_gpu == PSK_OK || no_gpu == PSK_EDEVICE);
    EXPECT(sa[text.size()] == 7u);  // sa[n..) untouched
    // argument checks (lib.rs:65,78-79; psacak32.rs:25)
    EXPECT(psk_psacak_inplace(nullptr, text.data(), 16, sa.data(), 15) == PSK_EINVAL);
    EXPECT(psk_psacak_inplace(nullptr, text.data(), 0xFFFFFFFFull, sa.data(), 1ull << 40) ==
           PSK_ETOOLONG);
    EXPECT(psk_psacak32(nullptr, sa.data(), 1ull << 31, sa.data(), 1ull << 31) == PSK_ETOOLONG);
    EXPECT(psk_psacak(nullptr, nullptr, 4, sa.data()) == PSK_EINVAL);
    EXPECT(psk_psacak8_dev_noalloc(nullptr, 5, nullptr, nullptr) == PSK_EINVAL);
    EXPECT(psk_psacak8_dev_noalloc(text.data(), PSK_NOALLOC_MAX_N + 1, sa.data(), nullptr) ==
           PSK_ETOOLONG);
    for (int code = -8; code <= 0; code++) EXPECT(std::strlen(psk_strerror(code)) > 0);
    // contexts and settings
    psk_ctx *c = nullptr;
    EXPECT(psk_ctx_create(&c, 1ull << 30) == PSK_OK);
    EXPECT(psk_ctx_set_direct_rounds(c, 65) == PSK_EINVAL);
    EXPECT(psk_ctx_set_doubling(c, 1025) == PSK_EINVAL);
    EXPECT(psk_ctx_set_capture(c, 0) == PSK_OK);
    uint64_t cnt = 0;
    EXPECT(psk_last_capture(c, nullptr, 0, &cnt) == PSK_ENOTFOUND);
    uint64_t res = 1, peak = 1;
    EXPECT(psk_ctx_workspace(c, &res, &peak) == PSK_OK && res == 0);
    uint64_t rows[6 * 4];
    EXPECT(psk_last_levels(c, rows, 4) == 0);
    const int rc = psk_psacak_inplace(c, text.data(), text.size(), sa.data(), text.size());
    EXPECT(rc == PSK_OK || rc == PSK_EDEVICE);
    if (rc != PSK_OK) EXPECT(std::strlen(psk_last_error(c)) > 0);
    EXPECT(psk_ctx_trim(c) == PSK_OK);
    psk_ctx_destroy(c);
    EXPECT(psk_ctx_trim(nullptr) == PSK_OK);  // the thread's default context
    // workspace bounds are monotone
    uint64_t prev = 0;
    for (uint64_t n = 1; n < (1ull << 33); n <<= 3) {
        EXPECT(psk_workspace_bytes8(n) >= prev);
        prev = psk_workspace_bytes8(n);
    }
    // the server: admission, a request larger than the budget, two threads
    // waiting on one request (exactly one receives it)
    psk_server *s = nullptr;
    EXPECT(psk_server_create(&s, 1ull << 31, 2) == PSK_OK);
    uint64_t id = 0;
    EXPECT(psk_server_submit(s, text.data(), 1ull << 31, &id) == PSK_EBUDGET);
    std::vector<uint64_t> ids;
    for (int i = 0; i < 6; i++) {
        EXPECT(psk_server_submit(s, text.data(), 1000 + 500 * i, &id) == PSK_OK);
        ids.push_back(id);
    }
    int got[2] = {0, 0};
    std::thread a([&] { got[0] = psk_server_wait(s, ids[0], nullptr, nullptr, 60000.0); });
    std::thread b([&] { got[1] = psk_server_wait(s, ids[0], nullptr, nullptr, 60000.0); });
    a.join();
    b.join();
    EXPECT((got[0] == PSK_ENOTFOUND) != (got[1] == PSK_ENOTFOUND));
    for (size_t i = 1; i < ids.size(); i++) {
        std::vector<uint32_t> out(1000 + 500 * i);
        double ms = 0;
        const int w = psk_server_wait(s, ids[i], out.data(), &ms, 60000.0);
        EXPECT(w == PSK_OK || w == PSK_EDEVICE);
    }
    uint64_t st[7];
    EXPECT(psk_server_stats(s, st, 7) == 7);
    EXPECT(st[5] == 0);  // no admission out of arrival order
    EXPECT(psk_server_pending(s) == 0);
    psk_server_destroy(s);
    if (failures) std::fprintf(stderr, "%d expectation(s) failed\n", failures);
    else std::printf("abi driver: all expectations hold\n");
    return failures ? 1 : 0;
}
