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

// Table layout: T0 replicated 64 times, entry e of copy c at byte
// 256 e + 4 c, and lane l reads copy l % 64, so the lanes of a half-wave
// always hit different LDS banks whatever the indices. A lookup address is
// then one v_perm_b32 of the state word and the lane's copy offset
// (4 (l % 64)): the index byte goes to byte 1, the offset to byte 0. The
// S-box bytes of the last round come from the same table (byte 2 of T0[x]
// is S[x]).
struct Replicated {
    static constexpr int COPIES = 64;
    uint32_t T0[256 * COPIES];
};

__device__ __forceinline__ uint32_t copy_offset() { return (threadIdx.x & 63u) << 2; }

// T0[byte B of s] from this lane's copy (off = copy_offset())
template <int B>
__device__ __forceinline__ uint32_t rt(const Replicated &tb, uint32_t s, uint32_t off) {
    // v_perm_b32 selectors: 4..7 = bytes of s, 0..3 = bytes of off, 0x0C = zero
    constexpr uint32_t sel = 0x0C0C0000u | ((4u + B) << 8);
    return *(const uint32_t *)((const uint8_t *)tb.T0 + __builtin_amdgcn_perm(s, off, sel));
}

__device__ __forceinline__ uint32_t xor3(uint32_t a, uint32_t b, uint32_t c) {
    return __builtin_amdgcn_bitop3_b32(a, b, c, 0x96);
}

// One output column of a middle round from state words a, b, c, d:
// T0[a.b3] ^ ror8(T0[b.b2]) ^ ror16(T0[c.b1]) ^ ror24(T0[d.b0]) ^ k.
__device__ __forceinline__ uint32_t round_col(const Replicated &tb, uint32_t off, uint32_t a, uint32_t b,
                                              uint32_t c, uint32_t d, uint32_t k) {
    return xor3(xor3(rt<3>(tb, a, off), ror(rt<2>(tb, b, off), 8), ror(rt<1>(tb, c, off), 16)),
                ror(rt<0>(tb, d, off), 24), k);
}

// Last-round column: the S-box bytes (byte 2 of the T0 entries) of a.b3,
// b.b2, c.b1, d.b0 in bytes 3, 2, 1, 0, plus k.
__device__ __forceinline__ uint32_t last_col(const Replicated &tb, uint32_t off, uint32_t a, uint32_t b,
                                             uint32_t c, uint32_t d, uint32_t k) {
    const uint32_t hi = __builtin_amdgcn_perm(rt<3>(tb, a, off), rt<2>(tb, b, off), 0x06020C0Cu);
    const uint32_t lo = __builtin_amdgcn_perm(rt<1>(tb, c, off), rt<0>(tb, d, off), 0x0C0C0602u);
    return (hi | lo) ^ k;
}

// Encrypt the CTR block for block index blk: input 0^64 || BE64(blk).
// Output words are big-endian (byte 0 in bits 31..24 of o[0]).
__device__ __forceinline__ void ctr_block(const uint32_t *rk, const Replicated &tb, uint64_t blk,
                                          uint32_t o[4]) {
    const uint32_t off = copy_offset();
    uint32_t s0 = rk[0], s1 = rk[1];
    uint32_t s2 = (uint32_t)(blk >> 32) ^ rk[2];
    uint32_t s3 = (uint32_t)blk ^ rk[3];
#pragma unroll
    for (int r = 1; r < 10; ++r) {
        const uint32_t t0 = round_col(tb, off, s0, s1, s2, s3, rk[4 * r]);
        const uint32_t t1 = round_col(tb, off, s1, s2, s3, s0, rk[4 * r + 1]);
        const uint32_t t2 = round_col(tb, off, s2, s3, s0, s1, rk[4 * r + 2]);
        const uint32_t t3 = round_col(tb, off, s3, s0, s1, s2, rk[4 * r + 3]);
        s0 = t0; s1 = t1; s2 = t2; s3 = t3;
    }
    o[0] = last_col(tb, off, s0, s1, s2, s3, rk[40]);
    o[1] = last_col(tb, off, s1, s2, s3, s0, rk[41]);
    o[2] = last_col(tb, off, s2, s3, s0, s1, rk[42]);
    o[3] = last_col(tb, off, s3, s0, s1, s2, rk[43]);
}

// ---- counter blocks of one key, a wavefront at a time ----------------------
// The counter is 0^64 || BE64(blk) with blk < 2^16, so the state words after
// AddRoundKey are s0, s1, s2 = key words and s3 = rk[3] ^ blk. Across the 64
// blocks base + lane of one wavefront (base a multiple of 64), bytes 1..3 of
// s3 are uniform: after round 1 only column 0 varies (through T3[s3.b0]),
// column 1 changes with blk's byte 1, columns 2 and 3 are fixed per key; in
// round 2 each column takes one lookup of column 0 and three uniform ones.
// The uniform lookups run on the scalar unit from the constant T0 (scalar
// loads and rotations), once per key or per wavefront step; a block then
// costs 1 + 4 lookups for rounds 1 and 2 instead of 4 + 16.
__device__ __forceinline__ uint32_t ror_u(uint32_t x, int s) { return s ? (x >> s) | (x << (32 - s)) : x; }

// Tt[byte B of x] for a wave-uniform x, from the constant table T0
template <int T, int B>
__device__ __forceinline__ uint32_t ut(const uint32_t *T0, uint32_t x) {
    return ror_u(T0[(x >> (8 * B)) & 0xFFu], 8 * T);
}

struct CtrKey {
    const uint32_t *rk;      // 44 round keys (SGPRs)
    uint32_t a0, a1, t2, t3; // round-1 parts fixed per key
    __device__ __forceinline__ CtrKey(const uint32_t *rk_, const uint32_t *T0) : rk(rk_) {
        const uint32_t s0 = rk[0], s1 = rk[1], s2 = rk[2], s3 = rk[3];  // bytes 2, 3 of s3 only
        a0 = ut<0, 3>(T0, s0) ^ ut<1, 2>(T0, s1) ^ ut<2, 1>(T0, s2) ^ rk[4];
        a1 = ut<0, 3>(T0, s1) ^ ut<1, 2>(T0, s2) ^ ut<3, 0>(T0, s0) ^ rk[5];
        t2 = ut<0, 3>(T0, s2) ^ ut<1, 2>(T0, s3) ^ ut<2, 1>(T0, s0) ^ ut<3, 0>(T0, s1) ^ rk[6];
        t3 = ut<0, 3>(T0, s3) ^ ut<1, 2>(T0, s0) ^ ut<2, 1>(T0, s1) ^ ut<3, 0>(T0, s2) ^ rk[7];
    }
};

// round-2 parts of one wavefront step (blocks base .. base + 63)
struct CtrStep {
    uint32_t c[4];
    __device__ __forceinline__ CtrStep(const CtrKey &k, const uint32_t *T0, uint32_t base) {
        const uint32_t t1 = k.a1 ^ ut<2, 1>(T0, k.rk[3] ^ base);
        c[0] = ut<1, 2>(T0, t1) ^ ut<2, 1>(T0, k.t2) ^ ut<3, 0>(T0, k.t3) ^ k.rk[8];
        c[1] = ut<0, 3>(T0, t1) ^ ut<1, 2>(T0, k.t2) ^ ut<2, 1>(T0, k.t3) ^ k.rk[9];
        c[2] = ut<0, 3>(T0, k.t2) ^ ut<1, 2>(T0, k.t3) ^ ut<3, 0>(T0, t1) ^ k.rk[10];
        c[3] = ut<0, 3>(T0, k.t3) ^ ut<2, 1>(T0, t1) ^ ut<3, 0>(T0, k.t2) ^ k.rk[11];
    }
};

// The block base + lane of the step: rounds 1 and 2 from the uniform parts,
// then rounds 3..10 as ctr_block.
__device__ __forceinline__ void ctr_block_step(const CtrKey &k, const CtrStep &st, const Replicated &tb,
                                               uint32_t blk, uint32_t o[4]) {
    const uint32_t off = copy_offset();
    const uint32_t *rk = k.rk;
    const uint32_t t0 = k.a0 ^ ror(rt<0>(tb, rk[3] ^ blk, off), 24);  // T3[s3.b0]
    uint32_t s0 = st.c[0] ^ rt<3>(tb, t0, off);                        // T0[t0.b3]
    uint32_t s1 = st.c[1] ^ ror(rt<0>(tb, t0, off), 24);               // T3[t0.b0]
    uint32_t s2 = st.c[2] ^ ror(rt<1>(tb, t0, off), 16);               // T2[t0.b1]
    uint32_t s3 = st.c[3] ^ ror(rt<2>(tb, t0, off), 8);                // T1[t0.b2]
#pragma unroll
    for (int r = 3; r < 10; ++r) {
        const uint32_t t0_ = round_col(tb, off, s0, s1, s2, s3, rk[4 * r]);
        const uint32_t t1_ = round_col(tb, off, s1, s2, s3, s0, rk[4 * r + 1]);
        const uint32_t t2_ = round_col(tb, off, s2, s3, s0, s1, rk[4 * r + 2]);
        const uint32_t t3_ = round_col(tb, off, s3, s0, s1, s2, rk[4 * r + 3]);
        s0 = t0_; s1 = t1_; s2 = t2_; s3 = t3_;
    }
    o[0] = last_col(tb, off, s0, s1, s2, s3, rk[40]);
    o[1] = last_col(tb, off, s1, s2, s3, s0, rk[41]);
    o[2] = last_col(tb, off, s2, s3, s0, s1, rk[42]);
    o[3] = last_col(tb, off, s3, s0, s1, s2, rk[43]);
}

}  // namespace aes
