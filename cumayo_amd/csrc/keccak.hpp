// Keccak-f[1600] and SHAKE, one state per lane (FIPS 202).
//
// Each lane owns one complete 25-word state in registers, so a wavefront
// hashes 64 independent items at once; that is how the engine batches the
// digest / salt / target / vinegar derivations of many signatures. The round
// constants and rotation offsets are the FIPS 202 values (derived there from
// the rc LFSR and the (x, y) walk).
#pragma once
#include <cstddef>
#include <cstdint>

namespace keccak {

__device__ __forceinline__ uint64_t rotl(uint64_t a, int s) {
    return (a << s) | (a >> (64 - s));
}

__constant__ uint64_t KECCAK_RC[24] = {
    0x0000000000000001ull, 0x0000000000008082ull, 0x800000000000808aull,
    0x8000000080008000ull, 0x000000000000808bull, 0x0000000080000001ull,
    0x8000000080008081ull, 0x8000000000008009ull, 0x000000000000008aull,
    0x0000000000000088ull, 0x0000000080008009ull, 0x000000008000000aull,
    0x000000008000808bull, 0x800000000000008bull, 0x8000000000008089ull,
    0x8000000000008003ull, 0x8000000000008002ull, 0x8000000000000080ull,
    0x000000000000800aull, 0x800000008000000aull, 0x8000000080008081ull,
    0x8000000000008080ull, 0x0000000080000001ull, 0x8000000080008008ull};

__device__ __forceinline__ void f1600(uint64_t *A) {
#pragma unroll 1
    for (int round = 0; round < 24; ++round) {
        uint64_t C0 = A[0] ^ A[5] ^ A[10] ^ A[15] ^ A[20];
        uint64_t C1 = A[1] ^ A[6] ^ A[11] ^ A[16] ^ A[21];
        uint64_t C2 = A[2] ^ A[7] ^ A[12] ^ A[17] ^ A[22];
        uint64_t C3 = A[3] ^ A[8] ^ A[13] ^ A[18] ^ A[23];
        uint64_t C4 = A[4] ^ A[9] ^ A[14] ^ A[19] ^ A[24];
        uint64_t D0 = C4 ^ rotl(C1, 1), D1 = C0 ^ rotl(C2, 1), D2 = C1 ^ rotl(C3, 1);
        uint64_t D3 = C2 ^ rotl(C4, 1), D4 = C3 ^ rotl(C0, 1);
        // theta, then rho and pi: B[y + 5*((2x+3y)%5)] = rotl(A[x+5y], r[x][y])
        uint64_t B0 = A[0] ^ D0;
        uint64_t B10 = rotl(A[1] ^ D1, 1);
        uint64_t B20 = rotl(A[2] ^ D2, 62);
        uint64_t B5 = rotl(A[3] ^ D3, 28);
        uint64_t B15 = rotl(A[4] ^ D4, 27);
        uint64_t B16 = rotl(A[5] ^ D0, 36);
        uint64_t B1 = rotl(A[6] ^ D1, 44);
        uint64_t B11 = rotl(A[7] ^ D2, 6);
        uint64_t B21 = rotl(A[8] ^ D3, 55);
        uint64_t B6 = rotl(A[9] ^ D4, 20);
        uint64_t B7 = rotl(A[10] ^ D0, 3);
        uint64_t B17 = rotl(A[11] ^ D1, 10);
        uint64_t B2 = rotl(A[12] ^ D2, 43);
        uint64_t B12 = rotl(A[13] ^ D3, 25);
        uint64_t B22 = rotl(A[14] ^ D4, 39);
        uint64_t B23 = rotl(A[15] ^ D0, 41);
        uint64_t B8 = rotl(A[16] ^ D1, 45);
        uint64_t B18 = rotl(A[17] ^ D2, 15);
        uint64_t B3 = rotl(A[18] ^ D3, 21);
        uint64_t B13 = rotl(A[19] ^ D4, 8);
        uint64_t B14 = rotl(A[20] ^ D0, 18);
        uint64_t B24 = rotl(A[21] ^ D1, 2);
        uint64_t B9 = rotl(A[22] ^ D2, 61);
        uint64_t B19 = rotl(A[23] ^ D3, 56);
        uint64_t B4 = rotl(A[24] ^ D4, 14);
        // chi
        A[0] = B0 ^ (~B1 & B2);   A[1] = B1 ^ (~B2 & B3);   A[2] = B2 ^ (~B3 & B4);
        A[3] = B3 ^ (~B4 & B0);   A[4] = B4 ^ (~B0 & B1);
        A[5] = B5 ^ (~B6 & B7);   A[6] = B6 ^ (~B7 & B8);   A[7] = B7 ^ (~B8 & B9);
        A[8] = B8 ^ (~B9 & B5);   A[9] = B9 ^ (~B5 & B6);
        A[10] = B10 ^ (~B11 & B12); A[11] = B11 ^ (~B12 & B13); A[12] = B12 ^ (~B13 & B14);
        A[13] = B13 ^ (~B14 & B10); A[14] = B14 ^ (~B10 & B11);
        A[15] = B15 ^ (~B16 & B17); A[16] = B16 ^ (~B17 & B18); A[17] = B17 ^ (~B18 & B19);
        A[18] = B18 ^ (~B19 & B15); A[19] = B19 ^ (~B15 & B16);
        A[20] = B20 ^ (~B21 & B22); A[21] = B21 ^ (~B22 & B23); A[22] = B22 ^ (~B23 & B24);
        A[23] = B23 ^ (~B24 & B20); A[24] = B24 ^ (~B20 & B21);
        A[0] ^= KECCAK_RC[round];
    }
}

// Incremental SHAKE over byte pieces held anywhere the lane can address.
// Full rate-sized blocks take a fast path with static state indices; only the
// ragged edges go through the dynamic-index helpers.
template <int RATE>
struct Shake {
    static constexpr int RW = RATE / 8;
    uint64_t st[25];
    int pos;  // bytes absorbed into the current block

    __device__ __forceinline__ void init() {
#pragma unroll
        for (int i = 0; i < 25; ++i) st[i] = 0;
        pos = 0;
    }
    __device__ __forceinline__ static uint64_t load64(const uint8_t *p) {
        uint64_t x = 0;
#pragma unroll
        for (int b = 0; b < 8; ++b) x |= (uint64_t)p[b] << (8 * b);
        return x;
    }
    __device__ __forceinline__ void xor_byte(int at, uint32_t b) {
#pragma unroll
        for (int w = 0; w < RW; ++w)
            if (w == (at >> 3)) st[w] ^= (uint64_t)(b & 0xFF) << (8 * (at & 7));
    }
    __device__ __forceinline__ void xor_word(int w, uint64_t x) {
#pragma unroll
        for (int i = 0; i < RW; ++i)
            if (i == w) st[i] ^= x;
    }
    // Absorb len bytes from p (any address space, any alignment).
    __device__ void absorb(const uint8_t *p, size_t len) {
        while (len && (pos & 7)) {
            xor_byte(pos, *p++);
            --len;
            if (++pos == RATE) { f1600(st); pos = 0; }
        }
        if (pos == 0) {
            while (len >= (size_t)RATE) {
#pragma unroll
                for (int i = 0; i < RW; ++i) st[i] ^= load64(p + 8 * i);
                f1600(st);
                p += RATE; len -= RATE;
            }
        }
        while (len >= 8) {
            xor_word(pos >> 3, load64(p));
            p += 8; len -= 8; pos += 8;
            if (pos == RATE) { f1600(st); pos = 0; }
        }
        while (len) {
            xor_byte(pos, *p++);
            --len;
            if (++pos == RATE) { f1600(st); pos = 0; }
        }
    }
    __device__ __forceinline__ void absorb_byte(uint32_t b) {
        xor_byte(pos, b);
        if (++pos == RATE) { f1600(st); pos = 0; }
    }
    // Pad (0x1F ... 0x80) and permute.
    __device__ __forceinline__ void finalize() {
        xor_byte(pos, 0x1F);
        st[RW - 1] ^= 0x8000000000000000ull;
        f1600(st);
    }
    // Squeeze len bytes (a single call after finalize) into out.
    __device__ void squeeze(uint8_t *out, int len) {
        for (int base = 0;; base += RATE) {
#pragma unroll
            for (int i = 0; i < RW; ++i) {
#pragma unroll
                for (int b = 0; b < 8; ++b) {
                    int off = base + 8 * i + b;
                    if (off < len) out[off] = (uint8_t)(st[i] >> (8 * b));
                }
            }
            if (base + RATE >= len) break;
            f1600(st);
        }
    }
};

}  // namespace keccak
