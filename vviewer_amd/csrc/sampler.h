/*
 * sampler.h - the low-discrepancy sampler of the extended mode, compiled for
 * the host (vvpt_sample_sequence with a NULL context) and for the gfx950
 * kernels, bit-exact between the two (integer arithmetic only).
 *
 * Padded 2-D Sobol points with Owen scrambling:
 *   - dimensions are used in pairs (2j, 2j+1); pair j of sample n of a pixel is
 *     point idx_j of the 2-D Sobol (0,2)-sequence (van der Corput in x, the
 *     generator v_0 = 2^31, v_{i+1} = v_i ^ (v_i >> 1) in y; Kollig and Keller,
 *     "Efficient Multidimensional Sampling", 2002), where idx_j is n under a
 *     nested uniform scramble keyed by (pixel, j) so that pairs decorrelate
 *     ("shuffled" Sobol);
 *   - each coordinate then gets its own nested uniform (Owen) scramble, by the
 *     hash-based permutation of Burley, "Practical Hash-based Owen Scrambling",
 *     JCGT 9(4), 2020 (reverse bits, Laine-Karras style multiply-xor steps,
 *     reverse back).
 * The first 2^k samples of every pair form a (0,k,2)-net in [0,1)^2 (Owen
 * scrambling keeps the net property), and each coordinate is stratified in
 * 2^k intervals.  Keys come from the lowbias32 integer hash (C. Wellons).
 *
 * The reference draws every random number from one xorshift stream per (pixel,
 * batch) (rng_def.glsl); parity mode keeps that stream, and this sampler is used
 * only in the extended mode.
 */
#pragma once
#include <stdint.h>

#if defined(__HIPCC__)
#define VVPT_SHD __host__ __device__ __forceinline__
#else
#define VVPT_SHD inline
#endif

namespace vvpt {

VVPT_SHD uint32_t rev_bits32(uint32_t x) {
    x = (x << 16) | (x >> 16);
    x = ((x & 0x00ff00ffu) << 8) | ((x & 0xff00ff00u) >> 8);
    x = ((x & 0x0f0f0f0fu) << 4) | ((x & 0xf0f0f0f0u) >> 4);
    x = ((x & 0x33333333u) << 2) | ((x & 0xccccccccu) >> 2);
    x = ((x & 0x55555555u) << 1) | ((x & 0xaaaaaaaau) >> 1);
    return x;
}

VVPT_SHD uint32_t lowbias32(uint32_t x) {
    x ^= x >> 16;
    x *= 0x7feb352du;
    x ^= x >> 15;
    x *= 0x846ca68bu;
    x ^= x >> 16;
    return x;
}

VVPT_SHD uint32_t hash_mix(uint32_t a, uint32_t b) { return lowbias32(a ^ lowbias32(b ^ 0x9e3779b9u)); }

/* Burley 2020, listing 2: a permutation whose output bit i depends only on input bits 0..i. */
VVPT_SHD uint32_t lk_permute(uint32_t x, uint32_t seed) {
    x += seed;
    x ^= x * 0x6c50b47cu;
    x ^= x * 0xb82f1e52u;
    x ^= x * 0xc7afe638u;
    x ^= x * 0x8d22f6e6u;
    return x;
}

/* Nested uniform scramble: counting bits from the most significant one, output bit j depends only on input
 * bits 0..j, so every aligned block of 2^k values maps onto an aligned block of 2^k values. */
VVPT_SHD uint32_t owen_scramble(uint32_t x, uint32_t seed) { return rev_bits32(lk_permute(rev_bits32(x), seed)); }

VVPT_SHD uint32_t sobol_x(uint32_t n) { return rev_bits32(n); }

VVPT_SHD uint32_t sobol_y(uint32_t n) {
    uint32_t r = 0;
    for (uint32_t v = 1u << 31; n != 0u; n >>= 1, v ^= v >> 1)
        if (n & 1u) r ^= v;
    return r;
}

/* Key of a pixel's sample sequence. */
VVPT_SHD uint32_t sampler_key(uint32_t pixel, uint32_t seed) { return hash_mix(pixel, seed ^ 0x5bd1e995u); }

/* Dimension `dim` of sample `n` as 32 scrambled bits. */
VVPT_SHD uint32_t sampler_bits(uint32_t key, uint32_t n, uint32_t dim) {
    const uint32_t pk = hash_mix(key, dim >> 1);
    /* bits of the shuffled index above 24 only reach output bits below 8, which the float drops */
    const uint32_t idx = owen_scramble(n, pk) & 0x00ffffffu;
    const uint32_t x = (dim & 1u) ? sobol_y(idx) : sobol_x(idx);
    return owen_scramble(x, hash_mix(pk, 1u + (dim & 1u)));
}

/* The same as a float in [0, 1): the top 24 bits. */
VVPT_SHD float sampler_float(uint32_t key, uint32_t n, uint32_t dim) {
    return (float)(sampler_bits(key, n, dim) >> 8) * (1.0f / 16777216.0f);
}

}  // namespace vvpt
