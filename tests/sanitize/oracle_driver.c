/* The CPU oracle (test infrastructure) under AddressSanitizer and
 * UndefinedBehaviorSanitizer: the serial and multi-threaded restatements on
 * random and adversarial texts, checked against brute force and the SA
 * checker.  Exit status 0 when all agree. */
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "../../oracle/psk_oracle.h"

static uint64_t rng = 0x9E3779B97F4A7C15ull;
static uint32_t next(void) {
    rng ^= rng << 13;
    rng ^= rng >> 7;
    rng ^= rng << 17;
    return (uint32_t)(rng >> 11);
}

int main(void) {
    int bad = 0;
    for (int it = 0; it < 300; it++) {
        uint32_t n = 1 + next() % (it < 250 ? 300 : 20000);
        uint32_t sigma = 1 + next() % (it % 3 ? 4 : 256);
        uint8_t *t = malloc(n);
        uint32_t *u = malloc(n * sizeof(uint32_t)), *tr = malloc(n * sizeof(uint32_t));
        uint32_t *sa = malloc(n * sizeof(uint32_t)), *ref = malloc(n * sizeof(uint32_t));
        for (uint32_t i = 0; i < n; i++) {
            t[i] = (uint8_t)(it % 7 == 0 ? (i % 3 == 0 ? 'a' : 'b') : 'A' + next() % sigma);
            u[i] = next() % (sigma * 4);
        }
        if (or_sacak8(t, n, sa) != 0) bad++;
        if (n <= 2000) {
            or_saca_naive8(t, n, ref);
            if (memcmp(sa, ref, n * sizeof(uint32_t))) bad++;
        }
        if (or_check_sa8(t, n, sa) != 0) bad++;
        if (or_par_sacak8(t, n, ref) != 0 || memcmp(sa, ref, n * sizeof(uint32_t))) bad++;
        if (or_translate32(u, n, tr)) {
            if (or_sacak32(tr, n, sa) != 0) bad++;
            if (or_check_sa32(u, n, sa) != 0) bad++;
            if (or_par_sacak32(tr, n, ref) != 0 || memcmp(sa, ref, n * sizeof(uint32_t))) bad++;
        }
        uint64_t st[5 * 16];
        (void)or_level_stats8(t, n, st, 16);
        free(t); free(u); free(tr); free(sa); free(ref);
    }
    if (bad) fprintf(stderr, "%d disagreement(s)\n", bad);
    else printf("oracle driver: all agree\n");
    return bad ? 1 : 0;
}
