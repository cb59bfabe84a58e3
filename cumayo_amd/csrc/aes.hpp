// AES-128 encryption (FIPS 197) for the public-key expansion keystream.
//
// Table form: one 256-entry round table T0 (column-mix of the S-box output)
// and the S-box itself, staged in LDS per workgroup; the other three tables
// are byte rotations of T0 (v_alignbit). The tables are produced on the host
// from the field definition and uploaded once per device.
#pragma once
#include <cstdint>

namespace aes {

struct Tables {
    uint32_t T0[256];
    uint32_t S[256];
};

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

// Encrypt the CTR block for block index blk: input 0^64 || BE64(blk).
// Output words are big-endian (byte 0 in bits 31..24 of o[0]).
__device__ __forceinline__ void ctr_block(const uint32_t *rk, const Tables &tb, uint64_t blk,
                                          uint32_t o[4]) {
    uint32_t s0 = rk[0], s1 = rk[1];
    uint32_t s2 = (uint32_t)(blk >> 32) ^ rk[2];
    uint32_t s3 = (uint32_t)blk ^ rk[3];
#pragma unroll
    for (int r = 1; r < 10; ++r) {
        uint32_t t0 = tb.T0[s0 >> 24] ^ ror(tb.T0[(s1 >> 16) & 0xFF], 8) ^
                      ror(tb.T0[(s2 >> 8) & 0xFF], 16) ^ ror(tb.T0[s3 & 0xFF], 24) ^ rk[4 * r];
        uint32_t t1 = tb.T0[s1 >> 24] ^ ror(tb.T0[(s2 >> 16) & 0xFF], 8) ^
                      ror(tb.T0[(s3 >> 8) & 0xFF], 16) ^ ror(tb.T0[s0 & 0xFF], 24) ^ rk[4 * r + 1];
        uint32_t t2 = tb.T0[s2 >> 24] ^ ror(tb.T0[(s3 >> 16) & 0xFF], 8) ^
                      ror(tb.T0[(s0 >> 8) & 0xFF], 16) ^ ror(tb.T0[s1 & 0xFF], 24) ^ rk[4 * r + 2];
        uint32_t t3 = tb.T0[s3 >> 24] ^ ror(tb.T0[(s0 >> 16) & 0xFF], 8) ^
                      ror(tb.T0[(s1 >> 8) & 0xFF], 16) ^ ror(tb.T0[s2 & 0xFF], 24) ^ rk[4 * r + 3];
        s0 = t0; s1 = t1; s2 = t2; s3 = t3;
    }
    o[0] = (tb.S[s0 >> 24] << 24 | tb.S[(s1 >> 16) & 0xFF] << 16 | tb.S[(s2 >> 8) & 0xFF] << 8 |
            tb.S[s3 & 0xFF]) ^ rk[40];
    o[1] = (tb.S[s1 >> 24] << 24 | tb.S[(s2 >> 16) & 0xFF] << 16 | tb.S[(s3 >> 8) & 0xFF] << 8 |
            tb.S[s0 & 0xFF]) ^ rk[41];
    o[2] = (tb.S[s2 >> 24] << 24 | tb.S[(s3 >> 16) & 0xFF] << 16 | tb.S[(s0 >> 8) & 0xFF] << 8 |
            tb.S[s1 & 0xFF]) ^ rk[42];
    o[3] = (tb.S[s3 >> 24] << 24 | tb.S[(s0 >> 16) & 0xFF] << 16 | tb.S[(s1 >> 8) & 0xFF] << 8 |
            tb.S[s2 & 0xFF]) ^ rk[43];
}

}  // namespace aes
