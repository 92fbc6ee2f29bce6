/*
 * ORACLE (test infrastructure only): stimulus and observation definitions the
 * oracle harnesses use.  The HIP library carries its own independent copy of
 * these definitions; tests check that both produce the same streams.
 *
 * Random input bits use the SplitMix64 finaliser (Steele, Lea, Flood,
 * "Fast splittable pseudorandom number generators", OOPSLA 2014) as a
 * counter-based hash.  Bit b of input signal s at cycle c for lane l is bit
 * (l mod 64) of  stim_word(seed, (s << 8) | b, c, l / 64).
 *
 * Observation digests:
 *   - bit-sliced (64 lanes per u64): a 64-word signature per lane word; per
 *     cycle c every observed word v_j is folded in as S[(j + c) & 63] ^= v_j,
 *     then one nonlinear round S'[p] = S[p] ^ (S[(p+17)&63] & ~S[(p+41)&63]);
 *   - word-level (one lane per thread): a 64-bit chain per lane,
 *     d = mix64(d ^ limb ^ ((j*64 + k + 1) << 32)) per observed limb.
 */
#ifndef RRTL_ORACLE_STIM_H
#define RRTL_ORACLE_STIM_H
#include <stdint.h>

static inline uint64_t stim_mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
static inline uint64_t stim_stream_key(uint64_t seed, uint64_t stream, uint64_t cycle) {
    return stim_mix64(seed ^ ((stream + 1) * 0x9E3779B97F4A7C15ull) ^ ((cycle + 1) * 0xD1B54A32D192ED03ull));
}
static inline uint64_t stim_word_from_key(uint64_t key, uint64_t word) {
    return stim_mix64(key + (word + 1) * 0x9E3779B97F4A7C15ull);
}
static inline uint64_t stim_word(uint64_t seed, uint64_t stream, uint64_t cycle, uint64_t word) {
    return stim_word_from_key(stim_stream_key(seed, stream, cycle), word);
}
/* one cycle of the bit-sliced signature for one lane word; S strided by `stride` */
static inline void sig_digest_cycle(uint64_t *S, uint64_t stride, uint64_t cycle, const uint64_t *v, uint32_t n_obs, uint64_t vstride) {
    uint64_t T[64];
    for (uint32_t j = 0; j < n_obs; j++) S[((j + cycle) & 63) * stride] ^= v[(uint64_t)j * vstride];
    for (uint32_t p = 0; p < 64; p++) T[p] = S[p * stride] ^ (S[((p + 17) & 63) * stride] & ~S[((p + 41) & 63) * stride]);
    for (uint32_t p = 0; p < 64; p++) S[p * stride] = T[p];
}
static inline uint64_t lane_digest_update(uint64_t d, uint32_t limb, uint32_t j, uint32_t k) {
    return stim_mix64(d ^ (uint64_t)limb ^ ((uint64_t)(j * 64u + k + 1u) << 32));
}
#endif
