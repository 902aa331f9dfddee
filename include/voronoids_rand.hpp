/*
 * voronoids_rand.hpp -- the random streams of the reference's tests, in C++ (header only).
 *
 * The reference's tests draw their points with rand 0.8.5 (tests/test_delaunay_tree.rs:9-17):
 *   let mut rng = StdRng::seed_from_u64(0);  let dist = Uniform::from(0.0..1.0);
 *   dist.sample(&mut rng)
 * rand and its dependencies are not part of the reference's tree, so this header restates their
 * published algorithms:
 *   - StdRng = ChaCha12Rng (rand_chacha 0.3): the ChaCha block function (RFC 8439 section 2.3)
 *     with 12 rounds, key = the 32-byte seed as eight little-endian words, a 64-bit block counter
 *     in words 12-13 starting at 0, a 64-bit stream id in words 14-15 (0); output words in block
 *     order, buffered four blocks at a time (rand_core's BlockRng: next_u64 = low word first).
 *   - seed_from_u64 (rand_core 0.6): PCG32 (multiplier 6364136223846793005, increment
 *     11634580027462260723, state advanced before each output, XSH-RR output) fills the seed, four
 *     little-endian bytes per output.
 *   - Uniform<f64>::from(low..high) (rand 0.8 UniformFloat): a draw takes next_u64() >> 12 as the
 *     mantissa of a double in [1, 2), subtracts 1 and returns value0_1 * scale + low, where scale
 *     is high - low reduced until low + scale * (1 - 2^-52) < high.
 * The ChaCha core and the buffering are checked against RFC 8439 / RFC 7539 known answers, and
 * the PCG32 step and output function against the published output of O'Neill's pcg32 demo
 * (pcg32_srandom_r(42, 54)).  What no published vector fixes: rand_core's increment constant
 * and its advance-before-output order, and the float mapping (checked for range, determinism
 * and the exact values its bit construction gives).
 */
#ifndef VORONOIDS_RAND_HPP
#define VORONOIDS_RAND_HPP

#include <array>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <stdexcept>

namespace voronoids {
namespace rand_compat {

inline uint32_t rotl32(uint32_t x, int r) { return (x << r) | (x >> (32 - r)); }

/* PCG32 (O'Neill 2014, pcg-c-basic): the 64-bit LCG step and the XSH-RR output function */
constexpr uint64_t kPcgMul = 6364136223846793005ull;
inline uint64_t pcg32_advance(uint64_t state, uint64_t inc) { return state * kPcgMul + inc; }
inline uint32_t pcg32_output(uint64_t state) {
    const uint32_t xorshifted = (uint32_t)(((state >> 18) ^ state) >> 27);
    const int rot = (int)(state >> 59);
    return (xorshifted >> rot) | (xorshifted << ((32 - rot) & 31));
}

/* the ChaCha quarter round (RFC 8439 section 2.1) on four words of the state */
inline void quarter_round(uint32_t& a, uint32_t& b, uint32_t& c, uint32_t& d) {
    a += b; d ^= a; d = rotl32(d, 16);
    c += d; b ^= c; b = rotl32(b, 12);
    a += b; d ^= a; d = rotl32(d, 8);
    c += d; b ^= c; b = rotl32(b, 7);
}

/* one block: the input state (constants, key, 64-bit counter, 64-bit stream id) after `rounds`
 * rounds (double rounds of column and diagonal quarter rounds), added to itself */
inline void chacha_block(const uint32_t key[8], uint64_t counter, uint64_t stream, int rounds, uint32_t out[16]) {
    uint32_t in[16] = {0x61707865u, 0x3320646eu, 0x79622d32u, 0x6b206574u,
                       key[0], key[1], key[2], key[3], key[4], key[5], key[6], key[7],
                       (uint32_t)counter, (uint32_t)(counter >> 32), (uint32_t)stream, (uint32_t)(stream >> 32)};
    uint32_t x[16];
    std::memcpy(x, in, sizeof(x));
    for (int r = 0; r < rounds; r += 2) {
        quarter_round(x[0], x[4], x[8], x[12]);
        quarter_round(x[1], x[5], x[9], x[13]);
        quarter_round(x[2], x[6], x[10], x[14]);
        quarter_round(x[3], x[7], x[11], x[15]);
        quarter_round(x[0], x[5], x[10], x[15]);
        quarter_round(x[1], x[6], x[11], x[12]);
        quarter_round(x[2], x[7], x[8], x[13]);
        quarter_round(x[3], x[4], x[9], x[14]);
    }
    for (int i = 0; i < 16; ++i) out[i] = x[i] + in[i];
}

/* ChaCha<rounds>Rng of rand_chacha 0.3 behind rand_core's BlockRng (four blocks per refill) */
template <int Rounds>
class ChaChaRng {
public:
    using Seed = std::array<uint8_t, 32>;

    static ChaChaRng from_seed(const Seed& seed) {
        ChaChaRng r;
        for (int i = 0; i < 8; ++i)
            r.key_[i] = (uint32_t)seed[4 * i] | (uint32_t)seed[4 * i + 1] << 8 | (uint32_t)seed[4 * i + 2] << 16 |
                        (uint32_t)seed[4 * i + 3] << 24;
        return r;
    }

    /* rand_core 0.6 SeedableRng::seed_from_u64: PCG32 output words fill the seed */
    static ChaChaRng seed_from_u64(uint64_t state) {
        Seed seed{};
        for (int i = 0; i < 8; ++i) {
            state = pcg32_advance(state, 11634580027462260723ull);   // advanced before the output
            const uint32_t x = pcg32_output(state);
            for (int b = 0; b < 4; ++b) seed[4 * i + b] = (uint8_t)(x >> (8 * b));
        }
        return from_seed(seed);
    }

    void set_stream(uint64_t stream) {
        stream_ = stream;
        index_ = kWords;   // the buffer refills from the current counter with the new stream
    }

    uint32_t next_u32() {
        if (index_ >= kWords) refill();
        return buf_[index_++];
    }

    uint64_t next_u64() {
        if (index_ < kWords - 1) {
            const uint64_t v = (uint64_t)buf_[index_] | (uint64_t)buf_[index_ + 1] << 32;
            index_ += 2;
            return v;
        }
        if (index_ >= kWords) {
            refill();
            index_ = 2;
            return (uint64_t)buf_[0] | (uint64_t)buf_[1] << 32;
        }
        const uint64_t lo = buf_[kWords - 1];   // one word left: it is the low half
        refill();
        index_ = 1;
        return lo | (uint64_t)buf_[0] << 32;
    }

private:
    static constexpr int kBlocks = 4, kWords = 16 * kBlocks;

    void refill() {
        for (int b = 0; b < kBlocks; ++b) chacha_block(key_, counter_ + (uint64_t)b, stream_, Rounds, buf_ + 16 * b);
        counter_ += kBlocks;
        index_ = 0;
    }

    uint32_t key_[8] = {};
    uint64_t counter_ = 0, stream_ = 0;
    uint32_t buf_[kWords] = {};
    int index_ = kWords;
};

using ChaCha12Rng = ChaChaRng<12>;
using ChaCha20Rng = ChaChaRng<20>;
/* rand 0.8: StdRng is ChaCha12Rng */
using StdRng = ChaCha12Rng;

/* rand 0.8 Uniform<f64> over [low, high) (UniformFloat::new and ::sample) */
class UniformF64 {
public:
    UniformF64(double low, double high) : low_(low) {
        if (!(low < high) || !std::isfinite(low) || !std::isfinite(high))
            throw std::invalid_argument("Uniform: low < high, both finite");
        const double max_rand = 1.0 - 0x1p-52;    // the largest value0_1 a draw can give
        scale_ = high - low;
        while (scale_ * max_rand + low >= high) scale_ = std::nextafter(scale_, 0.0);
    }
    template <class Rng>
    double sample(Rng& rng) const {
        const uint64_t bits = (rng.next_u64() >> 12) | 0x3ff0000000000000ull;
        double value1_2;
        std::memcpy(&value1_2, &bits, sizeof(double));
        const double value0_1 = value1_2 - 1.0;
        return value0_1 * scale_ + low_;
    }

private:
    double low_, scale_;
};

}  // namespace rand_compat
}  // namespace voronoids

#endif
