/* Philox-4x32-10 counter-based generator (Salmon, Moraes, Dror, Shaw, "Parallel random
 * numbers: as easy as 1, 2, 3", SC'11), written from the paper's description:
 * each round maps (c0,c1,c2,c3) with multipliers M0, M1 and the key (k0,k1):
 *   (hi0,lo0) = M0*c0, (hi1,lo1) = M1*c2
 *   c' = (hi1 ^ c1 ^ k0, lo1, hi0 ^ c3 ^ k1, lo0)
 * and the key is bumped by the Weyl constants W0, W1 between rounds.
 *
 * ORACLE COPY (test infrastructure): the product has its own device implementation;
 * this one is pinned by the published known-answer vectors in tests/test_oracle_pins.py.
 */
#ifndef OR_PHILOX_H
#define OR_PHILOX_H
#include <stdint.h>

typedef struct { uint32_t v[4]; } or_u32x4;

static inline or_u32x4 or_philox4x32_10(or_u32x4 c, uint32_t k0, uint32_t k1)
{
    const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u;
    const uint32_t W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
    for (int r = 0; r < 10; ++r) {
        uint64_t p0 = (uint64_t)M0 * c.v[0];
        uint64_t p1 = (uint64_t)M1 * c.v[2];
        uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        or_u32x4 n;
        n.v[0] = hi1 ^ c.v[1] ^ k0;
        n.v[1] = lo1;
        n.v[2] = hi0 ^ c.v[3] ^ k1;
        n.v[3] = lo0;
        c = n;
        k0 += W0;
        k1 += W1;
    }
    return c;
}

/* 53-bit uniform strictly inside (0,1) from two 32-bit words. */
static inline double or_u01(uint32_t a, uint32_t b)
{
    uint64_t m = ((uint64_t)(a >> 5) << 26) | (uint64_t)(b >> 6);
    return ((double)m + 0.5) * (1.0 / 9007199254740992.0);
}

/* 42-bit uniform strictly inside (0,1): (hi * 2^10 + lo + 0.5) * 2^-42. */
static inline double or_u42(uint32_t hi, uint32_t lo)
{
    uint64_t m = ((uint64_t)hi << 10) | (uint64_t)(lo & 0x3FFu);
    return ((double)m + 0.5) * (1.0 / 4398046511104.0);
}

/* Per-particle stream: key = (seed, iteration).  Sequential mode (sources): draw k uses
 * counter (particle, k/2, phase, 0); words (0,1) serve even draws, (2,3) odd ones.  Stepped
 * mode (transport): draw j of step s uses counter (particle, s, phase, j/3), three 42-bit
 * uniforms per block: word j%3 above the 10 bits of word 3 starting at bit 10(j%3). */
typedef struct {
    uint32_t k0, k1, pid, phase;
    uint32_t ndraw, step;
    int stepped;
    or_u32x4 buf;
} or_stream;

static inline void or_stream_init(or_stream *s, uint32_t seed, uint32_t iter,
                                  uint32_t pid, uint32_t phase)
{
    s->k0 = seed; s->k1 = iter; s->pid = pid; s->phase = phase; s->ndraw = 0;
    s->step = 0; s->stepped = 0;
}

/* start transport step `step` (1-based event count of the particle) */
static inline void or_stream_step(or_stream *s, uint32_t step)
{
    s->stepped = 1; s->step = step; s->ndraw = 0;
}

static inline double or_draw(or_stream *s)
{
    uint32_t k = s->ndraw++;
    if (s->stepped) {
        uint32_t j = k % 3u;
        if (j == 0) {
            or_u32x4 c = {{s->pid, s->step, s->phase, k / 3u}};
            s->buf = or_philox4x32_10(c, s->k0, s->k1);
        }
        return or_u42(s->buf.v[j], s->buf.v[3] >> (10u * j));
    }
    if ((k & 1u) == 0) {
        or_u32x4 c = {{s->pid, k >> 1, s->phase, 0u}};
        s->buf = or_philox4x32_10(c, s->k0, s->k1);
        return or_u01(s->buf.v[0], s->buf.v[1]);
    }
    return or_u01(s->buf.v[2], s->buf.v[3]);
}

#endif
