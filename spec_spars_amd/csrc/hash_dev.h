// Counter-based random sources evaluated in registers.
//
// sign_hash: CityHash64WithSeed over the 16 bytes LE64(edge) || LE64(column),
//   the key Sparsifier::hash_with_inputs feeds fasthash's CityHasher
//   (src/sparsifier.rs:397-420).  Google CityHash's published algorithm:
//   v1.1 HashLen0to16 (variant 0, default) or v1.0.3 (variant 1), then
//   HashLen16(h - k2, seed).
// coin: the e-th f64 draw of rand 0.8.5 StdRng::seed_from_u64(seed)
//   (src/sparsifier.rs:625-633): ChaCha12 keystream words 2e, 2e+1 of block
//   e/8 (64-bit block counter, stream 0), mapped (u >> 12) * 2^-52.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace ss {

__host__ __device__ __forceinline__ uint64_t ch_rotr(uint64_t x, unsigned r) {
    return (x >> r) | (x << (64u - r));
}

__host__ __device__ __forceinline__ uint64_t ch_mix128(uint64_t lo, uint64_t hi) {
    const uint64_t K = 0x9ddfea08eb382d69ULL;
    uint64_t a = (lo ^ hi) * K;
    a ^= a >> 47;
    uint64_t b = (hi ^ a) * K;
    b ^= b >> 47;
    return b * K;
}

__host__ __device__ __forceinline__ uint64_t sign_hash(uint64_t e, uint64_t j, uint64_t seed,
                                                       int variant) {
    const uint64_t K2 = 0x9ae16a3b2f90404fULL;
    uint64_t h;
    if (variant == 0) {
        const uint64_t mul = K2 + 32u;           // k2 + 2 * len
        uint64_t a = e + K2;
        uint64_t c = ch_rotr(j, 37) * mul + a;
        uint64_t d = (ch_rotr(a, 25) + j) * mul;
        uint64_t x = (c ^ d) * mul;
        x ^= x >> 47;
        uint64_t y = (d ^ x) * mul;
        y ^= y >> 47;
        h = y * mul;
    } else {
        h = ch_mix128(e, ch_rotr(j + 16u, 16)) ^ j;
    }
    return ch_mix128(h - K2, seed);
}

// Q[e,j] before scaling: Rademacher +-1 from bit 31 of the low word, or the
// open-interval uniform map 2((x + 0.5) / 2^32) - 1.
__host__ __device__ __forceinline__ double sign_value(uint64_t e, uint64_t j, uint64_t seed,
                                                      int uniform, int variant) {
    uint32_t x = (uint32_t)sign_hash(e, j, seed, variant);
    if (uniform) return 2.0 * (((double)x + 0.5) / 4294967296.0) - 1.0;
    return (x & 0x80000000u) ? 1.0 : -1.0;
}

// ---- ChaCha12 ------------------------------------------------------------
__host__ __device__ __forceinline__ uint32_t cc_rotl(uint32_t v, unsigned c) {
    return (v << c) | (v >> (32u - c));
}

#define SS_CC_QR(a, b, c, d)                 \
    a += b; d ^= a; d = cc_rotl(d, 16);      \
    c += d; b ^= c; b = cc_rotl(b, 12);      \
    a += b; d ^= a; d = cc_rotl(d, 8);       \
    c += d; b ^= c; b = cc_rotl(b, 7);

struct CoinKey {
    uint32_t k[8];
};

// rand_core 0.6 SeedableRng::seed_from_u64: PCG32 output words fill the key.
inline CoinKey coin_key_from_u64(uint64_t seed) {
    CoinKey ck;
    uint64_t st = seed;
    for (int i = 0; i < 8; i++) {
        st = st * 6364136223846793005ULL + 11634580027462260723ULL;
        uint32_t xs = (uint32_t)(((st >> 18) ^ st) >> 27);
        uint32_t rot = (uint32_t)(st >> 59);
        ck.k[i] = (xs >> rot) | (xs << ((32u - rot) & 31u));
    }
    return ck;
}

// Words 2*slot and 2*slot+1 of ChaCha12 block `block` (stream 0).
__device__ __forceinline__ uint64_t chacha12_u64(const CoinKey &ck, uint64_t block, int slot) {
    uint32_t s[16], x[16];
    s[0] = 0x61707865u; s[1] = 0x3320646eu; s[2] = 0x79622d32u; s[3] = 0x6b206574u;
#pragma unroll
    for (int i = 0; i < 8; i++) s[4 + i] = ck.k[i];
    s[12] = (uint32_t)block; s[13] = (uint32_t)(block >> 32); s[14] = 0u; s[15] = 0u;
#pragma unroll
    for (int i = 0; i < 16; i++) x[i] = s[i];
#pragma unroll
    for (int r = 0; r < 6; r++) {
        SS_CC_QR(x[0], x[4], x[8], x[12]);
        SS_CC_QR(x[1], x[5], x[9], x[13]);
        SS_CC_QR(x[2], x[6], x[10], x[14]);
        SS_CC_QR(x[3], x[7], x[11], x[15]);
        SS_CC_QR(x[0], x[5], x[10], x[15]);
        SS_CC_QR(x[1], x[6], x[11], x[12]);
        SS_CC_QR(x[2], x[7], x[8], x[13]);
        SS_CC_QR(x[3], x[4], x[9], x[14]);
    }
    // select the two words without dynamic register indexing
    uint32_t lo = 0, hi = 0;
#pragma unroll
    for (int i = 0; i < 8; i++) {
        if (slot == i) {
            lo = x[2 * i] + s[2 * i];
            hi = x[2 * i + 1] + s[2 * i + 1];
        }
    }
    return (uint64_t)lo | ((uint64_t)hi << 32);
}

__device__ __forceinline__ double coin_value(const CoinKey &ck, uint64_t e) {
    uint64_t u = chacha12_u64(ck, e >> 3, (int)(e & 7u));
    return (double)(u >> 12) * (1.0 / 4503599627370496.0);
}

}  // namespace ss
