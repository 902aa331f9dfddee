// Known-answer tests of include/voronoids_rand.hpp (CPU only).
//   - the quarter round and the block function against RFC 8439 sections 2.1.1 and 2.3.2;
//   - ChaCha20Rng with the all-zero seed against the keystream of RFC 7539 appendix A.1 (test
//     vectors 1 and 2: counter 0 and 1), which pins the key / counter / stream layout and the
//     block order of the buffered output;
//   - the buffer's refill boundary for next_u32 / next_u64 against direct block calls.
//   - the PCG32 LCG step and XSH-RR output function against the published output of O'Neill's
//     pcg32 demo (pcg-c-basic, pcg32_srandom_r(42, 54)), which pins the multiplier, the shifts
//     and the rotation direction used by the seed expansion;
//   - the float mapping on chosen next_u64 values, where the bit construction fixes the result.
// ChaCha12 (StdRng) and rand_core's increment constant have no published vectors: StdRng is
// exercised for determinism and range.
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "voronoids_rand.hpp"

using namespace voronoids::rand_compat;

static int failures = 0;
#define CHECK(cond)                                                                       \
    do {                                                                                  \
        if (!(cond)) {                                                                    \
            std::fprintf(stderr, "%s:%d: CHECK failed: %s\n", __FILE__, __LINE__, #cond); \
            ++failures;                                                                   \
        }                                                                                 \
    } while (0)

int main() {
    {   // RFC 8439 2.1.1
        uint32_t a = 0x11111111u, b = 0x01020304u, c = 0x9b8d6f43u, d = 0x01234567u;
        quarter_round(a, b, c, d);
        CHECK(a == 0xea2a92f4u && b == 0xcb1cf8ceu && c == 0x4581472eu && d == 0x5881c4bbu);
    }
    {   // RFC 8439 2.3.2: key 00 01 .. 1f, block count 1, nonce 00 00 00 09 00 00 00 4a 00 00 00 00
        uint32_t key[8];
        for (int i = 0; i < 8; ++i)
            key[i] = (uint32_t)(4 * i) | (uint32_t)(4 * i + 1) << 8 | (uint32_t)(4 * i + 2) << 16 | (uint32_t)(4 * i + 3) << 24;
        uint32_t out[16];
        chacha_block(key, 1ull | (0x09000000ull << 32), 0x4a000000ull, 20, out);
        const uint32_t want[16] = {0xe4e7f110u, 0x15593bd1u, 0x1fdd0f50u, 0xc47120a3u, 0xc7f4d1c7u, 0x0368c033u,
                                   0x9aaa2204u, 0x4e6cd4c3u, 0x466482d2u, 0x09aa9f07u, 0x05d7c214u, 0xa2028bd9u,
                                   0xd19c12b5u, 0xb94e16deu, 0xe883d0cbu, 0x4e3c50a2u};
        for (int i = 0; i < 16; ++i) CHECK(out[i] == want[i]);
    }
    {   // RFC 7539 A.1 #1 and #2: all-zero key and nonce, counters 0 and 1
        ChaCha20Rng rng = ChaCha20Rng::from_seed({});
        const uint32_t want[32] = {
            0xade0b876u, 0x903df1a0u, 0xe56a5d40u, 0x28bd8653u, 0xb819d2bdu, 0x1aed8da0u, 0xccef36a8u, 0xc70d778bu,
            0x7c5941dau, 0x8d485751u, 0x3fe02477u, 0x374ad8b8u, 0xf4b8436au, 0x1ca11815u, 0x69b687c3u, 0x8665eeb2u,
            0xbee7079fu, 0x7a385155u, 0x7c97ba98u, 0x0d082d73u, 0xa0290fcbu, 0x6965e348u, 0x3e53c612u, 0xed7aee32u,
            0x7621b729u, 0x434ee69cu, 0xb03371d5u, 0xd539d874u, 0x281fed31u, 0x45fb0a51u, 0x1f0ae1acu, 0x6f4d794bu};
        for (int i = 0; i < 32; ++i) CHECK(rng.next_u32() == want[i]);
    }
    {   // refill boundaries: words 0..319 of StdRng::seed_from_u64(0) drawn as u32, as u64 pairs, and
        // as u64 straddling a refill, against direct block calls
        StdRng::Seed seed{};
        {
            uint64_t s = 0;
            for (int i = 0; i < 8; ++i) {
                s = s * 6364136223846793005ull + 11634580027462260723ull;
                const uint32_t xs = (uint32_t)(((s >> 18) ^ s) >> 27);
                const int rot = (int)(s >> 59);
                const uint32_t x = (xs >> rot) | (xs << ((32 - rot) & 31));
                for (int b = 0; b < 4; ++b) seed[4 * i + b] = (uint8_t)(x >> (8 * b));
            }
        }
        uint32_t key[8];
        for (int i = 0; i < 8; ++i)
            key[i] = (uint32_t)seed[4 * i] | (uint32_t)seed[4 * i + 1] << 8 | (uint32_t)seed[4 * i + 2] << 16 |
                     (uint32_t)seed[4 * i + 3] << 24;
        std::vector<uint32_t> words(320);
        for (int blk = 0; blk < 20; ++blk) chacha_block(key, (uint64_t)blk, 0, 12, words.data() + 16 * blk);
        StdRng a = StdRng::seed_from_u64(0);
        for (int i = 0; i < 320; ++i) CHECK(a.next_u32() == words[i]);
        StdRng b = StdRng::seed_from_u64(0);
        for (int i = 0; i < 160; ++i) CHECK(b.next_u64() == ((uint64_t)words[2 * i] | (uint64_t)words[2 * i + 1] << 32));
        StdRng c = StdRng::seed_from_u64(0);
        for (int i = 0; i < 63; ++i) c.next_u32();
        CHECK(c.next_u64() == ((uint64_t)words[63] | (uint64_t)words[64] << 32));   // low word from the old buffer
        CHECK(c.next_u32() == words[65]);
    }
    {   // O'Neill's pcg32 demo: srandom(initstate 42, initseq 54) = {state 0, inc 2*54+1; step;
        // state += 42; step}; each draw outputs from the state before its step
        const uint64_t inc = (54ull << 1) | 1u;
        uint64_t st = pcg32_advance(0, inc);
        st = pcg32_advance(st + 42u, inc);
        const uint32_t want[6] = {0xa15c02b7u, 0x7b47f409u, 0xba1d3330u, 0x83d2f293u, 0xbfa4784bu, 0xcbed606eu};
        for (int i = 0; i < 6; ++i) {
            CHECK(pcg32_output(st) == want[i]);
            st = pcg32_advance(st, inc);
        }
    }
    {   // the float mapping: next_u64() >> 12 is the mantissa of value1_2 in [1, 2)
        struct Fixed {
            uint64_t v;
            uint64_t next_u64() { return v; }
        };
        const UniformF64 unit(0.0, 1.0), wide(-2.0, 3.0);
        Fixed zero{0}, half{1ull << 63}, top{~0ull}, low_bits{0xfffull};
        CHECK(unit.sample(zero) == 0.0 && wide.sample(zero) == -2.0);
        CHECK(unit.sample(half) == 0.5 && wide.sample(half) == 0.5);
        CHECK(unit.sample(top) == 1.0 - 0x1p-52);
        CHECK(wide.sample(top) == 3.0 - 0x1p-50);   // 5 (1 - 2^-52) rounds to 5 - 2^-50
        CHECK(unit.sample(low_bits) == 0.0);         // the low 12 bits are dropped
    }
    {   // Uniform(0, 1): in range, deterministic, mean near 1/2
        StdRng r1 = StdRng::seed_from_u64(0), r2 = StdRng::seed_from_u64(0);
        const UniformF64 dist(0.0, 1.0);
        double sum = 0.0;
        for (int i = 0; i < 100000; ++i) {
            const double x = dist.sample(r1);
            CHECK(x >= 0.0 && x < 1.0);
            CHECK(x == dist.sample(r2));
            sum += x;
        }
        CHECK(std::fabs(sum / 100000 - 0.5) < 0.01);
        const UniformF64 shifted(-2.0, 3.0);
        StdRng r3 = StdRng::seed_from_u64(7);
        for (int i = 0; i < 1000; ++i) {
            const double x = shifted.sample(r3);
            CHECK(x >= -2.0 && x < 3.0);
        }
    }
    std::printf("rand_compat: %s\n", failures ? "FAILED" : "ok");
    return failures ? 1 : 0;
}
