// Wavefront-cooperative Keccak-f[1600] and SHAKE256 (FIPS 202).
//
// Lane i < 25 holds state word A[x + 5y] with x = i % 5, y = i / 5; the
// theta column sums, the rho-pi move and the chi neighbours are fetched with
// ds_bpermute (the LDS crossbar, no LDS storage). Used inside the
// wave-per-item kernels where only one item needs hashing at a time (the
// in-kernel verification target and the rare signing retry), so that no lane
// has to hold a whole 1600-bit state in registers.
#pragma once
#include <cstdint>

#include "keccak.hpp"

namespace keccak {

__device__ __forceinline__ uint32_t bperm(uint32_t v, int lane) {
    return (uint32_t)__builtin_amdgcn_ds_bpermute(lane << 2, (int)v);
}
__device__ __forceinline__ uint64_t bperm64(uint64_t v, int lane) {
    return ((uint64_t)bperm((uint32_t)(v >> 32), lane) << 32) | bperm((uint32_t)v, lane);
}
__device__ __forceinline__ uint64_t rotl_var(uint64_t a, uint32_t s) {
    return s ? (a << s) | (a >> (64 - s)) : a;
}

// rho offsets r[x + 5 y] (FIPS 202 table 2)
__constant__ uint8_t KECCAK_RHO[25] = {0,  1,  62, 28, 27, 36, 44, 6,  55, 20, 3,  10, 43,
                                        25, 39, 41, 45, 15, 21, 8,  18, 2,  61, 56, 14};

struct WaveKeccak {
    int x, y, sx0;     // own coordinates; theta column base
    int dm1, dp1;      // lanes holding C[x-1], C[x+1]
    int src;           // rho-pi source lane for this position
    uint32_t rho;      // rotation this lane applies before sending
    int c1, c2;        // chi neighbours

    __device__ __forceinline__ WaveKeccak() {
        const int lane = threadIdx.x & 63;
        const int i = lane < 25 ? lane : 0;
        x = i % 5; y = i / 5;
        sx0 = x;
        dm1 = (x + 4) % 5;
        dp1 = (x + 1) % 5;
        // B[X][Y] with X = y_src, Y = 2 x_src + 3 y_src  =>  source of (x, y):
        // y_src = x, x_src = 3 (y - 3 x) mod 5
        const int ys = x, xs = ((3 * (y - 3 * x)) % 5 + 5) % 5;
        src = xs + 5 * ys;
        rho = KECCAK_RHO[i];
        c1 = (x + 1) % 5 + 5 * y;
        c2 = (x + 2) % 5 + 5 * y;
    }

    __device__ __forceinline__ uint64_t permute(uint64_t a) const {
        const int lane = threadIdx.x & 63;
#pragma unroll 1
        for (int round = 0; round < 24; ++round) {
            uint64_t C = bperm64(a, sx0) ^ bperm64(a, sx0 + 5) ^ bperm64(a, sx0 + 10) ^
                         bperm64(a, sx0 + 15) ^ bperm64(a, sx0 + 20);
            uint64_t D = bperm64(C, dm1) ^ rotl_var(bperm64(C, dp1), 1);
            a ^= D;
            uint64_t B = bperm64(rotl_var(a, rho), src);
            a = B ^ (~bperm64(B, c1) & bperm64(B, c2));
            if (lane == 0) a ^= KECCAK_RC[round];
        }
        return a;
    }
};

// Little-endian 8-byte read of p[0..n), n <= 8, any alignment.
__device__ __forceinline__ uint64_t load_le(const uint8_t *p, int n) {
    uint64_t v = 0;
    for (int b = 0; b < n; ++b) v |= (uint64_t)p[b] << (8 * b);
    return v;
}

// SHAKE256 of the contiguous input in[0, len) into out[0, outlen), computed
// by the whole wavefront (every lane must call it).
__device__ __forceinline__ void wave_shake256(uint8_t *out, int outlen, const uint8_t *in,
                                              uint32_t len) {
    constexpr int RATE = 136, RW = RATE / 8;
    const int lane = threadIdx.x & 63;
    WaveKeccak k;
    uint64_t a = 0;
    while (len >= (uint32_t)RATE) {
        if (lane < RW) a ^= load_le(in + 8 * lane, 8);
        a = k.permute(a);
        in += RATE;
        len -= RATE;
    }
    if (lane < RW) {
        const int off = 8 * lane;
        const int n = (int)len - off;
        if (n > 0) a ^= load_le(in + off, n > 8 ? 8 : n);
        if ((int)len / 8 == lane) a ^= (uint64_t)0x1F << (8 * (len % 8));
        if (lane == RW - 1) a ^= 0x8000000000000000ull;
    }
    a = k.permute(a);
    for (int base = 0;; base += RATE) {
        if (lane < RW) {
            for (int b = 0; b < 8; ++b) {
                const int o = base + 8 * lane + b;
                if (o < outlen) out[o] = (uint8_t)(a >> (8 * b));
            }
        }
        if (base + RATE >= outlen) break;
        a = k.permute(a);
    }
}

}  // namespace keccak
