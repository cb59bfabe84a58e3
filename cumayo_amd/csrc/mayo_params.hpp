// Compile-time MAYO parameter set for one library build (-DMAYO_VERSION=1|2|3|5).
// Values follow params.h:10-96 of the reference; derived sizes are computed.
#pragma once
#include <cstddef>
#include <cstdint>

#ifndef MAYO_VERSION
#error "build with -DMAYO_VERSION=1, 2, 3 or 5"
#endif

namespace mayo {

#if MAYO_VERSION == 1
constexpr int N = 86, M = 78, O = 8, K = 10, SK = 24, DIGEST = 32;
constexpr uint32_t F0 = 0x8, F1 = 0x1, F2 = 0x1, F3 = 0x0;
constexpr const char *NAME = "MAYO_1";
#elif MAYO_VERSION == 2
constexpr int N = 81, M = 64, O = 17, K = 4, SK = 24, DIGEST = 32;
constexpr uint32_t F0 = 0x8, F1 = 0x0, F2 = 0x2, F3 = 0x8;
constexpr const char *NAME = "MAYO_2";
#elif MAYO_VERSION == 3
constexpr int N = 118, M = 108, O = 10, K = 11, SK = 32, DIGEST = 48;
constexpr uint32_t F0 = 0x8, F1 = 0x0, F2 = 0x1, F3 = 0x7;
constexpr const char *NAME = "MAYO_3";
#elif MAYO_VERSION == 5
constexpr int N = 154, M = 142, O = 12, K = 12, SK = 40, DIGEST = 64;
constexpr uint32_t F0 = 0x4, F1 = 0x0, F2 = 0x8, F3 = 0x1;
constexpr const char *NAME = "MAYO_5";
#else
#error "MAYO_VERSION must be 1, 2, 3 or 5"
#endif

constexpr int V = N - O;                    // vinegar variables (== M in every set)
static_assert(V == M, "the reference's layouts assume v == m");
static_assert(M % 2 == 0, "m-vectors are whole bytes");
constexpr int MB = M / 2;                   // bytes per m-vector
constexpr int MW = (MB + 3) / 4;            // 32-bit words per m-vector (nibble packed)
constexpr int KO = K * O;
constexpr int PAIRS = K * (K + 1) / 2;
constexpr int O_BYTES = (V * O + 1) / 2;
constexpr int P1_ENTRIES = V * (V + 1) / 2;
constexpr int P2_ENTRIES = V * O;
constexpr int P3_ENTRIES = O * (O + 1) / 2;
constexpr int P1_BYTES = P1_ENTRIES * MB;
constexpr int P2_BYTES = P2_ENTRIES * MB;
constexpr int P3_BYTES = P3_ENTRIES * MB;
constexpr int P_BYTES = P1_BYTES + P2_BYTES;
constexpr int PK_BYTES = 16 + P3_BYTES;
constexpr int SIG_S_BYTES = K * N / 2;      // packed s
constexpr int SIG_BYTES = SIG_S_BYTES + SK; // s || salt
constexpr int ESK_BYTES = SK + O_BYTES + P_BYTES;
constexpr int EPK_BYTES = P_BYTES + P3_BYTES;
constexpr int VR_BYTES = K * N / 2;         // V || r squeezed per attempt
constexpr int V_BYTES = V / 2;              // one V_i
constexpr int R_OFF = K * V_BYTES;          // r follows the k vinegar vectors
constexpr int AES_BLOCKS = (P_BYTES + 15) / 16;
constexpr int MAX_ATTEMPTS = 255;           // ctr in [0, 255), sign.cu:244

constexpr size_t align128(size_t x) { return (x + 127) / 128 * 128; }

// Pair index of (i, j), i <= j (code.h:36-39).
__host__ __device__ constexpr int pair_index(int i, int j) {
    return i * K - i * (i - 1) / 2 + (j - i);
}
// Upper-triangular entry (r, c), c >= r, of a dim x dim matrix (code.h:15-21).
__host__ __device__ constexpr int tri_entry(int r, int c, int dim) {
    return (c - r) + dim * r - r * (r - 1) / 2;
}

}  // namespace mayo
