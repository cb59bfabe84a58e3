// m-vector word access and the E (multiply by z mod f) operator on
// nibble-packed 32-bit words.
#pragma once
#include <cstdint>
#include "gf16.hpp"
#include "mayo_params.hpp"

namespace mvec {

using namespace mayo;

// Number of bytes of an m-vector that fall in word w.
__device__ __forceinline__ int word_valid(int w) {
    int v = MB - 4 * w;
    return v >= 4 ? 4 : v;
}

// Load nvalid (1..4) bytes starting at p as the low bytes of a word.
// Reads the aligned dwords that contain those bytes and nothing beyond.
__device__ __forceinline__ uint32_t load_word(const uint8_t *p, int nvalid) {
    uintptr_t a = (uintptr_t)p;
    uint32_t sh = (uint32_t)(a & 3);
    const uint32_t *base = (const uint32_t *)(a - sh);
    uint32_t lo = base[0];
    uint32_t hi = (sh + (uint32_t)nvalid > 4) ? base[1] : 0u;
    uint32_t v = __builtin_amdgcn_alignbyte(hi, lo, sh);
    if (nvalid < 4) v &= (1u << (8 * nvalid)) - 1u;
    return v;
}

// Word w of the m-vector that starts at byte address e.
__device__ __forceinline__ uint32_t load_mw(const uint8_t *e, int w) {
    return load_word(e + 4 * w, word_valid(w));
}

// Store the valid bytes of word w of the m-vector at e.
__device__ __forceinline__ void store_mw(uint8_t *e, int w, uint32_t v) {
    uint8_t *p = e + 4 * w;
    int nv = word_valid(w);
    if (nv == 4 && ((uintptr_t)p & 3) == 0) {
        *(uint32_t *)p = v;
    } else {
        for (int b = 0; b < nv; ++b) p[b] = (uint8_t)(v >> (8 * b));
    }
}

// Nibble e of a packed byte array.
__device__ __forceinline__ uint32_t nib(const uint8_t *p, int e) {
    return (p[e >> 1] >> ((e & 1) * 4)) & 0xFu;
}

// Mask of the valid nibbles of the last m-vector word.
constexpr uint32_t LAST_MASK = (M % 8) ? ((1u << (4 * (M % 8))) - 1u) : 0xFFFFFFFFu;

// y <- z * y mod f(z) on an m-vector held in MW registers (poly.cuh:11-69):
// every coefficient moves up one position and the coefficient that leaves
// position m-1 re-enters through z^m = F0 + F1 z + F2 z^2 + F3 z^3.
__device__ __forceinline__ void E_inplace(uint32_t (&y)[MW]) {
    constexpr int TOPW = (M - 1) / 8, TOPS = 4 * ((M - 1) % 8);
    uint32_t top = (y[TOPW] >> TOPS) & 0xFu;
#pragma unroll
    for (int w = MW - 1; w > 0; --w) y[w] = (y[w] << 4) | (y[w - 1] >> 28);
    y[0] <<= 4;
    y[MW - 1] &= LAST_MASK;
    y[0] ^= gf16::mul(top, F0) | (gf16::mul(top, F1) << 4) | (gf16::mul(top, F2) << 8) |
            (gf16::mul(top, F3) << 12);
}

// acc ^= z^ell * src mod f(z), src an m-vector in memory (MW words, the
// nibbles past m zero). Valid for ell + 3 < m, which every parameter set
// meets for ell < k(k+1)/2 (sign.cu:340-388 fold the same overflow).
__device__ __forceinline__ void add_E_pow(uint32_t (&acc)[MW], const uint32_t *src, int ell) {
    const int ws = ell >> 3, ns = 4 * (ell & 7);
    // shifted part, truncated to m nibbles
#pragma unroll
    for (int w = 0; w < MW; ++w) {
        int hiw = w - ws, low = w - ws - 1;
        uint32_t hi = (hiw >= 0) ? src[hiw] : 0u;
        uint32_t lo = (low >= 0 && ns) ? src[low] : 0u;
        uint32_t val = (hi << ns) | (ns ? (lo >> (32 - ns)) : 0u);
        if (w == MW - 1) val &= LAST_MASK;
        acc[w] ^= val;
    }
    if (ell == 0) return;
    // overflow part: ov nibble t = src nibble m - ell + t, t < ell
    const int q = M - ell, qs = q >> 3, qn = 4 * (q & 7);
    uint32_t ov[MW];
#pragma unroll
    for (int w = 0; w < MW; ++w) {
        int a = w + qs, b = w + qs + 1;
        uint32_t lo = (a < MW) ? src[a] : 0u;
        uint32_t hi = (b < MW && qn) ? src[b] : 0u;
        ov[w] = (lo >> qn) | (qn ? (hi << (32 - qn)) : 0u);
    }
    // acc += F_n * ov shifted up by n nibbles, n = 0..3
#pragma unroll
    for (int w = 0; w < MW; ++w) {
        gf16::Multiples mo(ov[w]);
        uint32_t t0 = mo.times(F0), t1 = mo.times(F1), t2 = mo.times(F2), t3 = mo.times(F3);
        acc[w] ^= t0 ^ (t1 << 4) ^ (t2 << 8) ^ (t3 << 12);
        if (w + 1 < MW) acc[w + 1] ^= (t1 >> 28) ^ (t2 >> 24) ^ (t3 >> 20);
    }
}

}  // namespace mvec
