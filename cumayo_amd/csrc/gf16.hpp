// GF(16) = GF(2)[x]/(x^4 + x + 1) on nibble-packed 32-bit words.
//
// A word holds 8 field elements, element e in bits [4e, 4e+4) -- the byte
// order of the reference's m-vectors (low nibble first, code.h:7-11) read as
// a little-endian word. A scalar-times-vector product is built from the four
// multiples x, a*x, a^2*x, a^3*x of the vector (a = the class of x in the
// field) and one mask per scalar bit, so each multiply-accumulate costs four
// AND-XOR operations per word; the multiples of one vector are shared by
// every scalar that multiplies it.
#pragma once
#include <cstdint>

namespace gf16 {

// Multiply each of the 8 nibbles by a (x^4 = x + 1).
__device__ __forceinline__ uint32_t mul_a(uint32_t x) {
    uint32_t h = (x & 0x88888888u) >> 3;
    return ((x << 1) & 0xEEEEEEEEu) ^ h ^ (h << 1);
}

// All-ones if bit b of s is set, else zero.
__device__ __forceinline__ uint32_t bitmask(uint32_t s, int b) {
    return (uint32_t)(-(int32_t)((s >> b) & 1u));
}

struct Multiples {
    uint32_t x0, x1, x2, x3;
    __device__ __forceinline__ explicit Multiples(uint32_t x) {
        x0 = x; x1 = mul_a(x0); x2 = mul_a(x1); x3 = mul_a(x2);
    }
    // s * x for a scalar nibble s
    __device__ __forceinline__ uint32_t times(uint32_t s) const {
        return (x0 & bitmask(s, 0)) ^ (x1 & bitmask(s, 1)) ^ (x2 & bitmask(s, 2)) ^
               (x3 & bitmask(s, 3));
    }
};

// acc ^ (x & m) in one v_bitop3_b32 (truth table 0x78 over (acc, x, m)).
__device__ __forceinline__ uint32_t axm(uint32_t acc, uint32_t x, uint32_t m) {
    return __builtin_amdgcn_bitop3_b32(acc, x, m, 0x78);
}

// acc ^= s * x given the multiples of x and the four bit masks of s.
__device__ __forceinline__ uint32_t mac(uint32_t acc, uint32_t x0, uint32_t x1, uint32_t x2,
                                        uint32_t x3, uint32_t m0, uint32_t m1, uint32_t m2,
                                        uint32_t m3) {
    acc = axm(acc, x0, m0);
    acc = axm(acc, x1, m1);
    acc = axm(acc, x2, m2);
    return axm(acc, x3, m3);
}

// Scalar product of two nibbles.
__device__ __forceinline__ uint32_t mul(uint32_t a, uint32_t b) {
    Multiples m(a & 0xF);
    return m.times(b & 0xF) & 0xF;
}

// Inverse via a^14 (a^15 = 1 for a != 0); 0 maps to 0.
__device__ __forceinline__ uint32_t inv(uint32_t a) {
    uint32_t a2 = mul(a, a), a4 = mul(a2, a2), a8 = mul(a4, a4);
    return mul(mul(a8, a4), a2);
}

// Replicate a nibble into all 8 positions of a word.
__device__ __forceinline__ uint32_t splat(uint32_t s) { return (s & 0xF) * 0x11111111u; }

// ---- planar words ----
// The same 8 elements with their bits grouped by plane: bit d of element e
// at bit 8d + e, so byte d is bit plane d. Sums and masked sums are the same
// bitwise operations in either form, but multiplication by a is a byte
// rotation plus one XOR: planes (P0, P1, P2, P3) -> (P3, P0 + P3, P1, P2).
namespace planar {

// exchange the bits selected by m with the bits D positions above them
template <int D>
__device__ __forceinline__ uint32_t delta_swap(uint32_t x, uint32_t m) {
    const uint32_t t = (x ^ (x >> D)) & m;
    return x ^ t ^ (t << D);
}

// nibble-packed -> planar: bit 4e + d -> bit 8d + e, four index-bit swaps
__device__ __forceinline__ uint32_t from_nibbles(uint32_t x) {
    x = delta_swap<14>(x, 0x0000CCCCu);
    x = delta_swap<7>(x, 0x00AA00AAu);
    x = delta_swap<2>(x, 0x0C0C0C0Cu);
    return delta_swap<1>(x, 0x22222222u);
}

// planar -> nibble-packed (the same swaps in reverse order)
__device__ __forceinline__ uint32_t to_nibbles(uint32_t x) {
    x = delta_swap<1>(x, 0x22222222u);
    x = delta_swap<2>(x, 0x0C0C0C0Cu);
    x = delta_swap<7>(x, 0x00AA00AAu);
    return delta_swap<14>(x, 0x0000CCCCu);
}

__device__ __forceinline__ uint32_t mul_a(uint32_t w) {
    return __builtin_amdgcn_alignbit(w, w, 24) ^ ((w >> 16) & 0xFF00u);
}

// a * acc + t (one Horner step)
__device__ __forceinline__ uint32_t mul_a_add(uint32_t acc, uint32_t t) {
    return __builtin_amdgcn_bitop3_b32(__builtin_amdgcn_alignbit(acc, acc, 24), (acc >> 16) & 0xFF00u, t, 0x96);
}

struct Multiples {
    uint32_t x0, x1, x2, x3;
    __device__ __forceinline__ explicit Multiples(uint32_t x) {
        x0 = x; x1 = mul_a(x0); x2 = mul_a(x1); x3 = mul_a(x2);
    }
};

}  // namespace planar
}  // namespace gf16
