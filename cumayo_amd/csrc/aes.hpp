// AES-128 encryption (FIPS 197) for the public-key expansion keystream.
//
// Table form: one 256-entry round table T0 (column-mix of the S-box output),
// staged in LDS per workgroup in a bank-conflict-free replicated layout; the
// other three tables are byte rotations of T0 (v_alignbit) and the S-box is
// a byte of T0. The tables are produced on the host from the field
// definition and uploaded once per device.
#pragma once
#include <cstdint>

namespace aes {

__device__ __forceinline__ uint32_t ror(uint32_t x, int s) {
    return __builtin_amdgcn_alignbit(x, x, s);
}

// Round keys as 44 big-endian words (FIPS 197 section 5.2).
__device__ __forceinline__ void expand_key(const uint8_t *key, uint32_t *rk, const uint32_t *S) {
    for (int i = 0; i < 4; ++i)
        rk[i] = ((uint32_t)key[4 * i] << 24) | ((uint32_t)key[4 * i + 1] << 16) |
                ((uint32_t)key[4 * i + 2] << 8) | (uint32_t)key[4 * i + 3];
    uint32_t rcon = 0x01;
    for (int i = 4; i < 44; ++i) {
        uint32_t t = rk[i - 1];
        if ((i & 3) == 0) {
            t = (t << 8) | (t >> 24);
            t = (S[t >> 24] << 24) | (S[(t >> 16) & 0xFF] << 16) | (S[(t >> 8) & 0xFF] << 8) |
                S[t & 0xFF];
            t ^= rcon << 24;
            rcon = (rcon << 1) ^ ((rcon & 0x80) ? 0x11B : 0);
        }
        rk[i] = rk[i - 4] ^ t;
    }
}

// Table layout: T0 replicated 32 times, entry e of copy c at word
// 32 e + c, and lane l reads copy l % 32, so the 32 lanes of a half-wave
// always hit 32 different LDS banks whatever the indices. The S-box bytes of
// the last round come from the same table (byte 2 of T0[x] is S[x]).
struct Replicated {
    uint32_t T0[256 * 32];
};

__device__ __forceinline__ uint32_t rt(const Replicated &tb, uint32_t idx, uint32_t lane32) {
    return tb.T0[(idx << 5) | lane32];
}

__device__ __forceinline__ uint32_t xor3(uint32_t a, uint32_t b, uint32_t c) {
    return __builtin_amdgcn_bitop3_b32(a, b, c, 0x96);
}

// One output column of a middle round: T0[a] ^ ror8(T0[b]) ^ ror16(T0[c]) ^ ror24(T0[d]) ^ k.
__device__ __forceinline__ uint32_t round_col(const Replicated &tb, uint32_t l, uint32_t a, uint32_t b,
                                              uint32_t c, uint32_t d, uint32_t k) {
    return xor3(xor3(rt(tb, a, l), ror(rt(tb, b, l), 8), ror(rt(tb, c, l), 16)), ror(rt(tb, d, l), 24), k);
}

// Encrypt the CTR block for block index blk: input 0^64 || BE64(blk).
// Output words are big-endian (byte 0 in bits 31..24 of o[0]).
__device__ __forceinline__ void ctr_block(const uint32_t *rk, const Replicated &tb, uint64_t blk,
                                          uint32_t o[4]) {
    const uint32_t l = threadIdx.x & 31;
    uint32_t s0 = rk[0], s1 = rk[1];
    uint32_t s2 = (uint32_t)(blk >> 32) ^ rk[2];
    uint32_t s3 = (uint32_t)blk ^ rk[3];
#pragma unroll
    for (int r = 1; r < 10; ++r) {
        const uint32_t t0 = round_col(tb, l, s0 >> 24, (s1 >> 16) & 0xFF, (s2 >> 8) & 0xFF, s3 & 0xFF, rk[4 * r]);
        const uint32_t t1 = round_col(tb, l, s1 >> 24, (s2 >> 16) & 0xFF, (s3 >> 8) & 0xFF, s0 & 0xFF, rk[4 * r + 1]);
        const uint32_t t2 = round_col(tb, l, s2 >> 24, (s3 >> 16) & 0xFF, (s0 >> 8) & 0xFF, s1 & 0xFF, rk[4 * r + 2]);
        const uint32_t t3 = round_col(tb, l, s3 >> 24, (s0 >> 16) & 0xFF, (s1 >> 8) & 0xFF, s2 & 0xFF, rk[4 * r + 3]);
        s0 = t0; s1 = t1; s2 = t2; s3 = t3;
    }
    // last round: S-box bytes (bits 16..23 of T0) placed by byte position
    auto sb = [&](uint32_t idx) { return (rt(tb, idx, l) >> 16) & 0xFFu; };
    o[0] = (sb(s0 >> 24) << 24 | sb((s1 >> 16) & 0xFF) << 16 | sb((s2 >> 8) & 0xFF) << 8 | sb(s3 & 0xFF)) ^ rk[40];
    o[1] = (sb(s1 >> 24) << 24 | sb((s2 >> 16) & 0xFF) << 16 | sb((s3 >> 8) & 0xFF) << 8 | sb(s0 & 0xFF)) ^ rk[41];
    o[2] = (sb(s2 >> 24) << 24 | sb((s3 >> 16) & 0xFF) << 16 | sb((s0 >> 8) & 0xFF) << 8 | sb(s1 & 0xFF)) ^ rk[42];
    o[3] = (sb(s3 >> 24) << 24 | sb((s0 >> 16) & 0xFF) << 16 | sb((s1 >> 8) & 0xFF) << 8 | sb(s2 & 0xFF)) ^ rk[43];
}

}  // namespace aes
