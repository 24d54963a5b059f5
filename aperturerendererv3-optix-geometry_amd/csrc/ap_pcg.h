// PCG32 (XSH-RR variant, 64-bit state, O'Neill 2014, "PCG: A Family of
// Simple Fast Space-Efficient Statistically Good Algorithms for Random
// Number Generation"), host and device. The core uses it in place of the
// reference's curand XORWOW streams (core/random.cuh:31-53): a stream per
// pixel (or per ray) selected by the increment, seeded by the pixel index.
#pragma once

#include <cstdint>

#include "ap_vec.h"

namespace ap {

struct Pcg32 {
    uint64_t state;
    uint64_t inc;
};

AP_HD uint32_t pcg32_next(Pcg32& r) {
    const uint64_t old = r.state;
    r.state = old * 6364136223846793005ULL + r.inc;
    const uint32_t xorshifted = static_cast<uint32_t>(((old >> 18u) ^ old) >> 27u);
    const uint32_t rot = static_cast<uint32_t>(old >> 59u);
    return (xorshifted >> rot) | (xorshifted << ((32u - rot) & 31u));
}

// pcg32_srandom_r(initstate, initseq).
AP_HD Pcg32 pcg32_seed(uint64_t initstate, uint64_t initseq) {
    Pcg32 r{0u, (initseq << 1u) | 1u};
    pcg32_next(r);
    r.state += initstate;
    pcg32_next(r);
    return r;
}

// SplitMix64 finaliser (Steele, Lea, Flood 2014): a bijective 64-bit mix.
AP_HD uint64_t splitmix64(uint64_t x) {
    x += 0x9E3779B97F4A7C15ULL;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ULL;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBULL;
    return x ^ (x >> 31);
}

// Hash of a radiance-cache frame (seed, frame index): its training-path
// offset, and, with kNrcTrainSalt, the seed of its training draws.
AP_HD uint64_t nrc_frame_hash(uint64_t seed, int32_t frame) {
    return splitmix64(splitmix64(static_cast<uint64_t>(frame)) ^ seed ^ 0x6e7263ull);
}
constexpr uint64_t kNrcTrainSalt = 0x747261696eull;

// The stream of one path sample: stream `pixel`, start state hashed from
// (seed, frame, pixel). Seeding consecutive frames with the raw frame number
// (pcg32_seed(pixel, frame)) leaves their start states on an arithmetic
// lattice, and the first three draws of neighbouring frames are then visibly
// correlated; the hash removes that.
AP_HD Pcg32 pcg32_path(uint64_t pixel, uint64_t frame, uint64_t seed) {
    const uint64_t k = splitmix64(seed ^ splitmix64(frame));
    return pcg32_seed(splitmix64(k ^ pixel), pixel);
}

// Uniform float in [0, 1) from the top 24 bits.
AP_HD float pcg32_float(Pcg32& r) { return static_cast<float>(pcg32_next(r) >> 8) * 0x1p-24f; }

}  // namespace ap
