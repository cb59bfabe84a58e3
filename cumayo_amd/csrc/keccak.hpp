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

// 64-bit rotation by a constant on (lo, hi) halves: two v_alignbit_b32.
template <int R>
__device__ __forceinline__ void rot64(uint32_t lo, uint32_t hi, uint32_t &olo, uint32_t &ohi) {
    if constexpr (R == 0) {
        olo = lo; ohi = hi;
    } else if constexpr (R < 32) {
        ohi = __builtin_amdgcn_alignbit(hi, lo, 32 - R);
        olo = __builtin_amdgcn_alignbit(lo, hi, 32 - R);
    } else if constexpr (R == 32) {
        olo = hi; ohi = lo;
    } else {
        ohi = __builtin_amdgcn_alignbit(lo, hi, 64 - R);
        olo = __builtin_amdgcn_alignbit(hi, lo, 64 - R);
    }
}

// a ^ (~b & c) in one v_bitop3_b32 (truth table 0xD2 over (a, b, c)).
__device__ __forceinline__ uint32_t chi(uint32_t a, uint32_t b, uint32_t c) {
    return __builtin_amdgcn_bitop3_b32(a, b, c, 0xD2);
}
// a ^ b ^ c in one v_bitop3_b32 (truth table 0x96).
__device__ __forceinline__ uint32_t xor3(uint32_t a, uint32_t b, uint32_t c) {
    return __builtin_amdgcn_bitop3_b32(a, b, c, 0x96);
}

// Keccak-f[1600] on the state as 32-bit halves: theta with 3-input XORs
// (bitop3), rho as alignbit pairs, pi by renaming, chi as one bitop3 per half.
__device__ __forceinline__ void f1600(uint64_t *A) {
    uint32_t L[25], H[25];
#pragma unroll
    for (int i = 0; i < 25; ++i) { L[i] = (uint32_t)A[i]; H[i] = (uint32_t)(A[i] >> 32); }
#pragma unroll 1
    for (int round = 0; round < 24; ++round) {
        uint32_t CL[5], CH[5], DL[5], DH[5];
#pragma unroll
        for (int x = 0; x < 5; ++x) {
            CL[x] = xor3(xor3(L[x], L[x + 5], L[x + 10]), L[x + 15], L[x + 20]);
            CH[x] = xor3(xor3(H[x], H[x + 5], H[x + 10]), H[x + 15], H[x + 20]);
        }
#pragma unroll
        for (int x = 0; x < 5; ++x) {
            uint32_t rl, rh;
            rot64<1>(CL[(x + 1) % 5], CH[(x + 1) % 5], rl, rh);
            DL[x] = CL[(x + 4) % 5] ^ rl;
            DH[x] = CH[(x + 4) % 5] ^ rh;
        }
        // theta, then rho and pi: B[y + 5*((2x+3y)%5)] = rotl(A[x+5y] ^ D[x], r[x][y])
        uint32_t BL[25], BH[25];
#define KECCAK_RP(src, dst, r) rot64<r>(L[src] ^ DL[(src) % 5], H[src] ^ DH[(src) % 5], BL[dst], BH[dst])
        KECCAK_RP(0, 0, 0);   KECCAK_RP(1, 10, 1);  KECCAK_RP(2, 20, 62); KECCAK_RP(3, 5, 28);
        KECCAK_RP(4, 15, 27); KECCAK_RP(5, 16, 36); KECCAK_RP(6, 1, 44);  KECCAK_RP(7, 11, 6);
        KECCAK_RP(8, 21, 55); KECCAK_RP(9, 6, 20);  KECCAK_RP(10, 7, 3);  KECCAK_RP(11, 17, 10);
        KECCAK_RP(12, 2, 43); KECCAK_RP(13, 12, 25); KECCAK_RP(14, 22, 39); KECCAK_RP(15, 23, 41);
        KECCAK_RP(16, 8, 45); KECCAK_RP(17, 18, 15); KECCAK_RP(18, 3, 21); KECCAK_RP(19, 13, 8);
        KECCAK_RP(20, 14, 18); KECCAK_RP(21, 24, 2); KECCAK_RP(22, 9, 61); KECCAK_RP(23, 19, 56);
        KECCAK_RP(24, 4, 14);
#undef KECCAK_RP
        // chi
#pragma unroll
        for (int y = 0; y < 25; y += 5)
#pragma unroll
            for (int x = 0; x < 5; ++x) {
                L[y + x] = chi(BL[y + x], BL[y + (x + 1) % 5], BL[y + (x + 2) % 5]);
                H[y + x] = chi(BH[y + x], BH[y + (x + 1) % 5], BH[y + (x + 2) % 5]);
            }
        const uint64_t rc = KECCAK_RC[round];
        L[0] ^= (uint32_t)rc;
        H[0] ^= (uint32_t)(rc >> 32);
    }
#pragma unroll
    for (int i = 0; i < 25; ++i) A[i] = (uint64_t)L[i] | ((uint64_t)H[i] << 32);
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
    // ---- register-only forms: every state index is a compile-time value, so
    // the state stays in VGPRs (the dynamic-position helpers above put the
    // rate words on the stack) ----

    // Absorb the len-byte message at p into a fresh state (position 0): full
    // blocks, then the tail word by word (bytes past the end are neither
    // read nor absorbed). Returns the tail length, the position to finalize at.
    __device__ __forceinline__ int absorb_msg(const uint8_t *p, size_t len) {
        while (len >= (size_t)RATE) {
#pragma unroll
            for (int i = 0; i < RW; ++i) st[i] ^= load64(p + 8 * i);
            f1600(st);
            p += RATE; len -= RATE;
        }
        const int tail = (int)len;
#pragma unroll
        for (int w = 0; w < RW; ++w) {
            const int nb = tail - 8 * w;
            uint64_t x = 0;
            if (nb >= 8) {
                x = load64(p + 8 * w);
            } else if (nb > 0) {
#pragma unroll
                for (int b = 0; b < 8; ++b)
                    if (b < nb) x |= (uint64_t)p[8 * w + b] << (8 * b);
            }
            st[w] ^= x;
        }
        return tail;
    }
    // Pad at position pos (< RATE, a runtime value) and permute.
    __device__ __forceinline__ void finalize_at(int pos) {
#pragma unroll
        for (int w = 0; w < RW; ++w) st[w] ^= w == (pos >> 3) ? 0x1Full << (8 * (pos & 7)) : 0ull;
        st[RW - 1] ^= 0x8000000000000000ull;
        f1600(st);
    }
    // Absorb LEN bytes from p as the input bytes [POS, POS + LEN) (the
    // stream so far being POS bytes long), permuting after each full block.
    template <int POS, int LEN>
    __device__ __forceinline__ void absorb_at(const uint8_t *p) {
#pragma unroll
        for (int w = POS / 8; w <= (POS + LEN - 1) / 8; ++w) {
            uint64_t x = 0;
#pragma unroll
            for (int b = 0; b < 8; ++b) {
                const int at = 8 * w + b;
                if (at >= POS && at < POS + LEN) x |= (uint64_t)p[at - POS] << (8 * b);
            }
            st[w % RW] ^= x;
            if (w % RW == RW - 1 && 8 * w + 7 < POS + LEN) f1600(st);
        }
    }
    // Absorb NW 64-bit words (little-endian byte order) as the input bytes
    // [POS, POS + 8 NW), POS a multiple of 8, permuting after each full
    // block: inputs held in registers (digests, salts, seeds).
    template <int POS, int NW>
    __device__ __forceinline__ void absorb_words_at(const uint64_t *x) {
        static_assert(POS % 8 == 0, "word-aligned position");
#pragma unroll
        for (int k = 0; k < NW; ++k) {
            constexpr int w0 = POS / 8;
            const int w = w0 + k;
            st[w % RW] ^= x[k];
            if (w % RW == RW - 1) f1600(st);
        }
    }
    // One byte b at stream position POS.
    template <int POS>
    __device__ __forceinline__ void absorb_byte_at(uint32_t b) {
        constexpr int at = POS % RATE;
        st[at >> 3] ^= (uint64_t)(b & 0xFF) << (8 * (at & 7));
        if constexpr (at == RATE - 1) f1600(st);
    }
    // Pad after a POS-byte stream and permute.
    template <int POS>
    __device__ __forceinline__ void finalize_static() {
        constexpr int at = POS % RATE;
        st[at >> 3] ^= 0x1Full << (8 * (at & 7));
        st[RW - 1] ^= 0x8000000000000000ull;
        f1600(st);
    }

    // The first NW words of the output (a single call after finalize, NW
    // within the rate) into registers.
    template <int NW>
    __device__ __forceinline__ void squeeze_words(uint64_t *x) const {
        static_assert(NW <= RW, "one block");
#pragma unroll
        for (int k = 0; k < NW; ++k) x[k] = st[k];
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
