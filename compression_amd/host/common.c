/* Host-side utilities: the seeded PRNG of the synthetic workloads
 * (splitmix64-seeded xoshiro256**) and a small pthread pool. */
#define _GNU_SOURCE
#include "mdz_host.h"

#include <pthread.h>
#include <sched.h>
#include <stdlib.h>
#include <stdatomic.h>

static uint64_t splitmix64(uint64_t* x) {
    uint64_t z = (*x += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

void mdz_rng_seed(mdz_rng* r, uint64_t seed) {
    uint64_t x = seed;
    for (int i = 0; i < 4; ++i) r->s[i] = splitmix64(&x);
}

static inline uint64_t rotl64(uint64_t v, int k) { return (v << k) | (v >> (64 - k)); }

uint64_t mdz_rng_next(mdz_rng* r) {
    uint64_t* s = r->s;
    uint64_t result = rotl64(s[1] * 5, 7) * 9;
    uint64_t t = s[1] << 17;
    s[2] ^= s[0];
    s[3] ^= s[1];
    s[1] ^= s[2];
    s[0] ^= s[3];
    s[2] ^= t;
    s[3] = rotl64(s[3], 45);
    return result;
}

uint64_t mdz_rng_below(mdz_rng* r, uint64_t n) {
    return (uint64_t)(((unsigned __int128)mdz_rng_next(r) * n) >> 64);
}

double mdz_rng_unit(mdz_rng* r) {
    return (double)(mdz_rng_next(r) >> 11) * (1.0 / 9007199254740992.0);
}

/* ------------------------------------------------------------ thread pool */
typedef struct {
    _Atomic uint64_t next;
    uint64_t n;
    mdz_task_fn fn;
    void* ctx;
} pf_state;

static void* pf_worker(void* arg) {
    pf_state* st = (pf_state*)arg;
    for (;;) {
        uint64_t i = atomic_fetch_add(&st->next, 1);
        if (i >= st->n) break;
        st->fn(st->ctx, i);
    }
    return NULL;
}

void mdz_parallel_for(uint64_t n, int threads, mdz_task_fn fn, void* ctx) {
    if (threads <= 1 || n <= 1) {
        for (uint64_t i = 0; i < n; ++i) fn(ctx, i);
        return;
    }
    if ((uint64_t)threads > n) threads = (int)n;
    pf_state st;
    atomic_init(&st.next, 0);
    st.n = n;
    st.fn = fn;
    st.ctx = ctx;
    pthread_t* tids = (pthread_t*)malloc(sizeof(pthread_t) * (size_t)threads);
    int started = 0;
    for (int t = 0; t < threads - 1; ++t)
        if (pthread_create(&tids[started], NULL, pf_worker, &st) == 0) ++started;
    pf_worker(&st);
    for (int t = 0; t < started; ++t) pthread_join(tids[t], NULL);
    free(tids);
}

int mdz_host_cores(void) {
    cpu_set_t set;
    CPU_ZERO(&set);
    if (sched_getaffinity(0, sizeof(set), &set) != 0) return 1;
    return CPU_COUNT(&set);
}

void mdz_free(void* p) { free(p); }
