// The shared-key P1 contraction of signing, T_j[r] = sum_{c >= r} P1[r][c] V_j[c]
// (sign.cu:396-418: every row r of the upper-triangular P1, every vinegar
// vector j), computed two ways over the same data and timed with HIP events:
//
//  * VALU: the bitsliced column loop of the signing kernels. One wavefront
//    per signature; lanes = (word w of an m-vector, row slot), two rows per
//    slot and chunk; per column one AND-XOR per entry word and scalar bit
//    with the scalar bits as SGPR masks; T_j = sum_b a^b T_jb per row.
//  * MFMA: the GF(2)-plane form as an int8 matrix product,
//      C[(r,e)][(s,j,i)] = sum_{(c,b)} A[(r,e)][(c,b)] B[(c,b)][(s,j,i)] mod 2,
//    A[(r,e)][(c,b)] = bit b of element e of P1[r][c] (stored once per key
//    in v_mfma_i32_32x32x32_i8 A-operand order, zero below the diagonal),
//    B[(c,b)][(s,j,i)] = bit i of a^b V_{s,j}[c] (built per signature in
//    registers), so that the parity of each int32 is bit i of
//    T_{s,j}[r][e]. A workgroup takes G = 8 signatures of one key; each of
//    its 10 waves holds one 32-column tile of B (0.8 signatures, 40 VGPRs)
//    and multiplies it with the key's A blocks, which the workgroup stages
//    in LDS one M tile ahead (10 KB per tile); the parities are packed into
//    words with one v_alignbit per accumulator. (16 signatures per
//    workgroup with two tiles per wave runs one workgroup per CU and took
//    twice as long.)
//
//  * LDS tables: a workgroup of TW = 8 signatures of one key builds, in LDS,
//    the 16 multiples s P1[r][c] (s = 0..15) of the entries of two columns
//    of a 12-row chunk at a time (double-buffered, one barrier per table),
//    and each signature's wavefront takes every product S_j[c] P1[r][c] as
//    one 8-byte LDS read at an address set by its (uniform) scalar, plus
//    two XORs; lanes = (word pair, row), one row per lane and chunk. This is
//    the column loop of the shared-key signing kernels, without their fold.
//
// All kernels reduce what they compute into per-wave checksums, and all
// write T in full for the first 16 signatures, which the program compares
// with a CPU restatement (the reference's loops). MAYO_1 sizes (v = m = 78,
// k = 10): 65,536 signatures under 64 keys, 1,024 signatures per key. Prints
// one JSON line with the kernels' times and the correctness check.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <cstring>
#include <type_traits>
#include <vector>

#include "gf16.hpp"

namespace {

constexpr int VV = 78, MM = 78, KK = 10;     // MAYO_1: v, m, k
constexpr int MW = (MM / 2 + 3) / 4;        // words per m-vector (10)
constexpr int NSIG = 65536, NKEY = 64, SIGS_PER_KEY = NSIG / NKEY;
constexpr int P1E = VV * (VV + 1) / 2;      // 3081 upper-triangular entries
constexpr int CHECK = 16;                   // signatures whose T is written in full (a multiple of G)

__host__ __device__ constexpr int tri(int r, int c) { return (c - r) + VV * r - r * (r - 1) / 2; }

// ---- MFMA layout ----
constexpr int KB = VV * 4;                  // K = (c, b): 312
constexpr int KSTEPS = (KB + 31) / 32;      // 10
constexpr int MROWS = VV * MM;              // M = (r, e): 6084
constexpr int MTILES = (MROWS + 31) / 32;   // 191
#ifndef CB_G
#define CB_G 8
#endif
#ifndef CB_TPW
#define CB_TPW 1
#endif
constexpr int G = CB_G;                     // signatures per workgroup
constexpr int NCOLS = G * KK * 4;           // 640 columns (s, j, i)
constexpr int NTILES = NCOLS / 32;          // 20
constexpr int TPW = CB_TPW;                 // N tiles per wave
constexpr int WAVES = NTILES / TPW;         // 10
static_assert(NTILES % TPW == 0, "tiles per wave");

// first K step of M tile t: rows r >= r_min need columns c >= r_min
__host__ __device__ constexpr int ks_start(int t) { return ((32 * t) / MM) / 8; }
// first M tile whose first K step is k
__host__ __device__ constexpr int first_tile(int k) {
    int t = 0;
    while (t < MTILES && ks_start(t) < k) ++t;
    return t;
}
template <int N, int I = 0, class F>
__device__ __forceinline__ void static_for_k(F &&f) {
    if constexpr (I < N) {
        f(std::integral_constant<int, I>{});
        static_for_k<N, I + 1>(f);
    }
}

typedef int v4i __attribute__((ext_vector_type(4)));
typedef int v16i __attribute__((ext_vector_type(16)));

// ---- VALU kernel ----
constexpr int SL = 64 / MW;                 // 6 row slots
constexpr int RT = 2;                       // rows per slot and chunk
constexpr int RPC = RT * SL;                // 12
constexpr int NCH = (VV + RPC - 1) / RPC;   // 7

// P1 in planar words (the prepared-key layout): entry e at e * MW words.
// scol[c] of a signature: nibble j = V_j[c]. One wave per signature.
__global__ void __launch_bounds__(64, 3) k_valu(const uint32_t *__restrict__ p1, const uint64_t *__restrict__ scol_all,
                                                uint32_t *__restrict__ cs_out, uint8_t *__restrict__ t_out) {
    const int sig = blockIdx.x, key = sig / SIGS_PER_KEY;
    const uint32_t *P = p1 + (size_t)key * P1E * MW;
    const uint64_t *scol = scol_all + (size_t)sig * VV;
    const int lane = threadIdx.x, w = lane % MW, slot = lane / MW;
    const bool active = slot < SL;
    uint32_t cs = 0;
#pragma unroll 1
    for (int ch = 0; ch < NCH; ++ch) {
        const int r0 = ch * RPC;
        int rr[RT];
        uint32_t off[RT];
        uint32_t Tb[RT][KK][4];
#pragma unroll
        for (int tt = 0; tt < RT; ++tt) {
            rr[tt] = r0 + tt * SL + slot;
            const int r = (active && rr[tt] < VV) ? rr[tt] : VV;  // VV: no row
            // entry (r, c) at tri(r, c) = tri(r, r) + (c - r); fetched from c = r0 on, masked below r
            off[tt] = r < VV ? (uint32_t)((tri(r, r) - (r - r0)) * MW + w) : 0u;
#pragma unroll
            for (int j = 0; j < KK; ++j)
#pragma unroll
                for (int b = 0; b < 4; ++b) Tb[tt][j][b] = 0;
        }
#pragma unroll 2
        for (int c = r0; c < VV; ++c) {
            const uint64_t sc = scol[c];  // uniform
            const uint32_t sh[2] = {(uint32_t)__builtin_amdgcn_readfirstlane((uint32_t)sc),
                                    (uint32_t)__builtin_amdgcn_readfirstlane((uint32_t)(sc >> 32))};
            uint32_t x[RT];
#pragma unroll
            for (int tt = 0; tt < RT; ++tt) {
                x[tt] = (active && rr[tt] < VV && c >= rr[tt]) ? P[off[tt]] : 0u;
                off[tt] += MW;
            }
#pragma unroll
            for (int j = 0; j < KK; ++j)
#pragma unroll
                for (int b = 0; b < 4; ++b) {
                    const int bit = 4 * j + b;
                    const uint32_t m = gf16::bitmask(sh[bit >> 5], bit & 31);
#pragma unroll
                    for (int tt = 0; tt < RT; ++tt) Tb[tt][j][b] = gf16::axm(Tb[tt][j][b], x[tt], m);
                }
        }
#pragma unroll
        for (int tt = 0; tt < RT; ++tt) {
#pragma unroll
            for (int j = 0; j < KK; ++j) {
                using namespace gf16::planar;
                const uint32_t T = mul_a_add(mul_a_add(mul_a_add(Tb[tt][j][3], Tb[tt][j][2]), Tb[tt][j][1]), Tb[tt][j][0]);
                cs ^= __builtin_amdgcn_alignbit(T, T, (uint32_t)(rr[tt] + j) & 31u);
                if (sig < CHECK && active && rr[tt] < VV) {
                    // element e = 8w + q: bit d of it at bit 8d + q (planar)
                    for (int q = 0; q < 8 && 8 * w + q < MM; ++q) {
                        uint32_t nib = 0;
#pragma unroll
                        for (int d = 0; d < 4; ++d) nib |= ((T >> (8 * d + q)) & 1u) << d;
                        t_out[(((size_t)sig * KK + j) * VV + rr[tt]) * MM + 8 * w + q] = (uint8_t)nib;
                    }
                }
            }
        }
    }
    cs_out[(size_t)sig * 64 + lane] = cs;
}

// ---- MFMA kernel ----
// A blocks of one key: block (t, ks) for ks >= ks_start(t), in order of t
// then ks, each 64 lanes x 16 bytes; blk_off[t] = index of block (t, ks_start(t)).
// B tiles built per signature from vnib[s][j][c] (one nibble per byte).
__global__ void __launch_bounds__(64 * WAVES) k_mfma(const v4i *__restrict__ ablk, const int *__restrict__ blk_off,
                                                        int blocks_per_key, const uint8_t *__restrict__ vnib,
                                                        uint32_t *__restrict__ cs_out, uint32_t *__restrict__ t_out) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int s0 = blockIdx.x * G, key = s0 / SIGS_PER_KEY;
    const v4i *A = ablk + (size_t)key * blocks_per_key * 64;
    // B: tile nt of this wave, column n = 32 nt + lane % 32 = (s, j, i);
    // k = 32 ks + 16 (lane / 32) + q = (c, b): bit i of a^b V_{s,j}[c]
    v4i B[TPW][KSTEPS];
#pragma unroll
    for (int u = 0; u < TPW; ++u) {
        const int n = 32 * (wave * TPW + u) + (lane & 31);
        const int s = n / (KK * 4), j = (n / 4) % KK, i = n % 4;
        const uint8_t *v = vnib + ((size_t)(s0 + s) * KK + j) * VV;
#pragma unroll
        for (int ks = 0; ks < KSTEPS; ++ks) {
            // dword q4 of the lane: k = 32 ks + 16 (lane / 32) + 4 q4 + b, i.e.
            // column c = 8 ks + 4 (lane / 32) + q4 and bits b = 0..3
            uint32_t wd[4];
#pragma unroll
            for (int q4 = 0; q4 < 4; ++q4) {
                const int c = 8 * ks + 4 * (lane >> 5) + q4;
                uint32_t x = c < VV ? v[c] : 0u, o = 0;
#pragma unroll
                for (int b = 0; b < 4; ++b) {
                    o |= ((x >> i) & 1u) << (8 * b);
                    x = gf16::mul_a(x) & 0xFu;
                }
                wd[q4] = o;
            }
            B[u][ks] = v4i{(int)wd[0], (int)wd[1], (int)wd[2], (int)wd[3]};
        }
    }
    uint32_t cs[TPW] = {};
    // A blocks of the current tile staged in LDS once for all waves, the
    // next tile's loaded from global memory meanwhile (double buffer)
    __shared__ v4i abuf[2][KSTEPS * 64];
    const int tid = threadIdx.x;
    auto nblk = [&](int t) { return t < MTILES ? (KSTEPS - ks_start(t)) * 64 : 0; };
    static_assert(64 * WAVES >= KSTEPS * 64, "one 16-byte load per thread stages a whole tile");
    if (tid < nblk(0)) abuf[0][tid] = A[tid];
    __syncthreads();
    int t = 0;
    // M tiles grouped by their first K step k0 (non-decreasing in t), each
    // group with its K range [k0, KSTEPS) fixed at compile time
    static_for_k<KSTEPS>([&](auto Kc) {
        constexpr int k0 = decltype(Kc)::value;
        const int t_end = k0 + 1 < KSTEPS ? first_tile(k0 + 1) : MTILES;
#pragma unroll 1
        for (; t < t_end; ++t) {
            const bool pre = tid < nblk(t + 1);
            v4i nxt{};
            if (pre) nxt = A[(size_t)blk_off[t + 1] * 64 + tid];
            const v4i *a = abuf[t & 1] + lane;
            v16i acc[TPW];
#pragma unroll
            for (int u = 0; u < TPW; ++u) acc[u] = v16i{};
#pragma unroll
            for (int ks = k0; ks < KSTEPS; ++ks) {
                const v4i av = a[(ks - k0) * 64];
#pragma unroll
                for (int u = 0; u < TPW; ++u) acc[u] = __builtin_amdgcn_mfma_i32_32x32x32_i8(av, B[u][ks], acc[u], 0, 0, 0);
            }
#pragma unroll
            for (int u = 0; u < TPW; ++u) {
                // parities of the 16 rows into the top 16 bits, one alignbit each
                uint32_t wd = 0;
#pragma unroll
                for (int v = 0; v < 16; ++v) wd = __builtin_amdgcn_alignbit((uint32_t)acc[u][v], wd, 1);
                cs[u] ^= __builtin_amdgcn_alignbit(wd, wd, (uint32_t)t & 31u);
                if (s0 < CHECK) {
                    const int n = 32 * (wave * TPW + u) + (lane & 31);
                    const int s = n / (KK * 4), j = (n / 4) % KK, i = n % 4;
                    for (int v = 0; v < 16; ++v) {
                        const int row = 32 * t + 8 * (v / 4) + 4 * (lane >> 5) + (v % 4);
                        if (row < MROWS) {
                            const int r = row / MM, e = row % MM;
                            const uint32_t bit = (wd >> (16 + v)) & 1u;
                            atomicOr(t_out + (((size_t)(s0 + s) * KK + j) * VV + r) * MM + e, bit << i);
                        }
                    }
                }
            }
            if (pre) abuf[(t + 1) & 1][tid] = nxt;
            __syncthreads();
        }
    });
#pragma unroll
    for (int u = 0; u < TPW; ++u) cs_out[((size_t)blockIdx.x * WAVES + wave) * 128 + u * 64 + lane] = cs[u];
}

// ---- LDS-table kernel ----
constexpr int TW = 8;                       // signatures per workgroup (one key)
constexpr int TWP = MW / 2;                 // word pairs per m-vector (5)
constexpr int TRPC = 64 / TWP;              // rows per chunk, one per lane slot (12)
constexpr int TNCH = (VV + TRPC - 1) / TRPC;
constexpr int TCG = 2;                      // columns per table
constexpr int TCSTR = TRPC * MW * 4;        // bytes of one column's entries (480)
constexpr int TSSTR = 1024;                 // bytes per multiple plane (>= TCG * TCSTR + idle lanes)
constexpr int TBUF = 16 * TSSTR;            // one table: 16 planes
constexpr int TNQ = TRPC * TCG * MW;        // entry words per table (240)
static_assert(TCG * TCSTR + 8 * (64 - TRPC * TWP) <= TSSTR, "idle lanes read inside the plane");
static_assert(TNQ <= 64 * TW, "one word per producing thread");

__device__ __forceinline__ void tab_next(int &ch, int &c0) {
    c0 += TCG;
    if (c0 >= VV) {
        ++ch;
        c0 = ch * TRPC;
    }
}

__global__ void __launch_bounds__(64 * TW) k_ldst(const uint32_t *__restrict__ p1,
                                                  const uint64_t *__restrict__ scol_all,
                                                  uint32_t *__restrict__ cs_out, uint8_t *__restrict__ t_out) {
    __shared__ __attribute__((aligned(16))) uint8_t tab[2 * TBUF];
    __shared__ uint64_t scol_s[TW][VV + TCG];
    const int tid = threadIdx.x, lane = tid & 63, wave = (int)__builtin_amdgcn_readfirstlane(tid >> 6);
    const int s0 = blockIdx.x * TW, key = s0 / SIGS_PER_KEY;
    const int sig = s0 + wave;
    const uint32_t *P = p1 + (size_t)key * P1E * MW;
    uint64_t *scol = scol_s[wave];
    for (int c = lane; c < VV + TCG; c += 64) scol[c] = c < VV ? scol_all[(size_t)sig * VV + c] : 0;
    for (int q = tid; q < TSSTR / 4; q += 64 * TW) {  // the zero multiple of both tables
        ((uint32_t *)tab)[q] = 0;
        ((uint32_t *)(tab + TBUF))[q] = 0;
    }
    // consumer lane: word pair g of row slot `slot`; producer thread: word q
    const int g = lane % TWP, slot = lane / TWP;
    const bool active = slot < TRPC;
    const int qcl = tid / (TRPC * MW), qrow = (tid / MW) % TRPC, qw = tid % MW;
    auto fetch = [&](int ch, int c0) -> uint32_t {
        const int r = ch * TRPC + qrow, c = c0 + qcl;
        return tid < TNQ && ch < TNCH && r < VV && c < VV && c >= r ? P[tri(r, c) * MW + qw] : 0u;
    };
    auto produce = [&](int buf, uint32_t x1) {
        if (tid >= TNQ) return;
        using namespace gf16::planar;
        const uint32_t x2 = mul_a(x1), x4 = mul_a(x2), x8 = mul_a(x4);
        const uint32_t x3 = x1 ^ x2, x5 = x1 ^ x4, x6 = x2 ^ x4, x7 = x3 ^ x4;
        uint32_t *d = (uint32_t *)(tab + buf * TBUF) + tid;
        constexpr int S4 = TSSTR / 4;
        d[1 * S4] = x1; d[2 * S4] = x2; d[3 * S4] = x3; d[4 * S4] = x4;
        d[5 * S4] = x5; d[6 * S4] = x6; d[7 * S4] = x7; d[8 * S4] = x8;
        d[9 * S4] = x1 ^ x8; d[10 * S4] = x2 ^ x8; d[11 * S4] = x3 ^ x8; d[12 * S4] = x4 ^ x8;
        d[13 * S4] = x5 ^ x8; d[14 * S4] = x6 ^ x8; d[15 * S4] = x7 ^ x8;
    };
    uint32_t acc[KK][2];
#pragma unroll
    for (int j = 0; j < KK; ++j) acc[j][0] = acc[j][1] = 0;
    uint32_t cs = 0;
    int ch = 0, c0 = 0, pch = 0, pc0 = 0;
    produce(0, fetch(0, 0));
    tab_next(pch, pc0);
    uint32_t x = fetch(pch, pc0);
    __syncthreads();
    int buf = 0;
#pragma unroll 1
    for (;;) {
        if (pch < TNCH) {
            produce(buf ^ 1, x);
            tab_next(pch, pc0);
            x = fetch(pch, pc0);
        }
        const uint8_t *bb = tab + buf * TBUF + slot * MW * 4 + g * 8;
        uint64_t scv[TCG];
#pragma unroll
        for (int cl = 0; cl < TCG; ++cl) scv[cl] = scol[c0 + cl];
#pragma unroll
        for (int cl = 0; cl < TCG; ++cl) {
            if (c0 + cl >= VV) break;
            const uint64_t sc = scv[cl];
            const uint32_t sh[2] = {(uint32_t)__builtin_amdgcn_readfirstlane((uint32_t)sc),
                                    (uint32_t)__builtin_amdgcn_readfirstlane((uint32_t)(sc >> 32))};
#pragma unroll
            for (int j = 0; j < KK; ++j) {
                const int bit = 4 * j;
                const uint32_t s = bit < 32 ? ((sh[0] >> bit) | (bit > 28 ? sh[1] << (32 - bit) : 0u)) & 15u
                                            : (sh[1] >> (bit - 32)) & 15u;
                const uint2 v = *(const uint2 *)(bb + s * TSSTR + cl * TCSTR);
                acc[j][0] ^= v.x;
                acc[j][1] ^= v.y;
            }
        }
        int nch = ch, nc0 = c0;
        tab_next(nch, nc0);
        if (nch != ch) {
            // chunk done: row ch * TRPC + slot of every T_j, words 2g, 2g + 1
            const int r = ch * TRPC + slot;
#pragma unroll
            for (int j = 0; j < KK; ++j) {
#pragma unroll
                for (int h = 0; h < 2; ++h) {
                    const uint32_t T = acc[j][h];
                    cs ^= __builtin_amdgcn_alignbit(T, T, (uint32_t)(r + j + h) & 31u);
                    const int w = 2 * g + h;
                    if (sig < CHECK && active && r < VV) {
                        for (int q = 0; q < 8 && 8 * w + q < MM; ++q) {
                            uint32_t nib = 0;
#pragma unroll
                            for (int d = 0; d < 4; ++d) nib |= ((T >> (8 * d + q)) & 1u) << d;
                            t_out[(((size_t)sig * KK + j) * VV + r) * MM + 8 * w + q] = (uint8_t)nib;
                        }
                    }
                    acc[j][h] = 0;
                }
            }
        }
        ch = nch;
        c0 = nc0;
        __syncthreads();
        if (ch >= TNCH) break;
        buf ^= 1;
    }
    cs_out[(size_t)sig * 64 + lane] = cs;
}

// ---- host: data, CPU restatement, layouts ----
uint32_t rng_state = 0x9E3779B9u;
uint32_t rnd() {
    rng_state ^= rng_state << 13;
    rng_state ^= rng_state >> 17;
    rng_state ^= rng_state << 5;
    return rng_state;
}
uint32_t gmul(uint32_t a, uint32_t b) {  // GF(16) mod x^4 + x + 1
    uint32_t r = 0;
    for (int i = 0; i < 4; ++i)
        if ((b >> i) & 1) r ^= a << i;
    for (int i = 6; i >= 4; --i)
        if ((r >> i) & 1) r ^= 0x13u << (i - 4);
    return r;
}
uint32_t planar_of(const uint8_t *nib8) {  // element q bit d -> bit 8d + q
    uint32_t w = 0;
    for (int q = 0; q < 8; ++q)
        for (int d = 0; d < 4; ++d) w |= ((nib8[q] >> d) & 1u) << (8 * d + q);
    return w;
}

}  // namespace

int main() {
    // P1 of every key: element e of entry (r, c), one nibble per byte
    std::vector<uint8_t> p1n((size_t)NKEY * P1E * MM);
    for (auto &x : p1n) x = rnd() & 0xF;
    std::vector<uint8_t> vn((size_t)NSIG * KK * VV);  // V_{s,j}[c]
    for (auto &x : vn) x = rnd() & 0xF;
    // VALU layout: planar words, entry e at e * MW
    std::vector<uint32_t> p1w((size_t)NKEY * P1E * MW, 0);
    for (int k = 0; k < NKEY; ++k)
        for (int e = 0; e < P1E; ++e)
            for (int w = 0; w < MW; ++w) {
                uint8_t n8[8] = {};
                for (int q = 0; q < 8 && 8 * w + q < MM; ++q) n8[q] = p1n[((size_t)k * P1E + e) * MM + 8 * w + q];
                p1w[((size_t)k * P1E + e) * MW + w] = planar_of(n8);
            }
    std::vector<uint64_t> scol((size_t)NSIG * VV);
    for (int s = 0; s < NSIG; ++s)
        for (int c = 0; c < VV; ++c) {
            uint64_t v = 0;
            for (int j = 0; j < KK; ++j) v |= (uint64_t)vn[((size_t)s * KK + j) * VV + c] << (4 * j);
            scol[(size_t)s * VV + c] = v;
        }
    // MFMA A blocks
    std::vector<int> boff(MTILES);
    int nb = 0;
    for (int t = 0; t < MTILES; ++t) {
        boff[t] = nb;
        nb += KSTEPS - ks_start(t);
    }
    const int bpk = nb;
    std::vector<uint8_t> ab((size_t)NKEY * bpk * 1024, 0);
    for (int k = 0; k < NKEY; ++k)
        for (int t = 0; t < MTILES; ++t)
            for (int ks = ks_start(t); ks < KSTEPS; ++ks) {
                uint8_t *blk = &ab[((size_t)k * bpk + boff[t] + (ks - ks_start(t))) * 1024];
                for (int l = 0; l < 64; ++l)
                    for (int q = 0; q < 16; ++q) {
                        const int row = 32 * t + (l & 31), kk = 32 * ks + 16 * (l >> 5) + q;
                        if (row >= MROWS || kk >= KB) continue;
                        const int r = row / MM, e = row % MM, c = kk >> 2, b = kk & 3;
                        if (c < r) continue;
                        blk[l * 16 + q] = (p1n[((size_t)k * P1E + tri(r, c)) * MM + e] >> b) & 1u;
                    }
            }
    // CPU restatement of T for the checked signatures (sign.cu:409-418)
    std::vector<uint8_t> tref((size_t)CHECK * KK * VV * MM, 0);
    for (int s = 0; s < CHECK; ++s) {
        const int k = s / SIGS_PER_KEY;
        for (int j = 0; j < KK; ++j)
            for (int r = 0; r < VV; ++r)
                for (int e = 0; e < MM; ++e) {
                    uint32_t acc = 0;
                    for (int c = r; c < VV; ++c)
                        acc ^= gmul(p1n[((size_t)k * P1E + tri(r, c)) * MM + e], vn[((size_t)s * KK + j) * VV + c]);
                    tref[(((size_t)s * KK + j) * VV + r) * MM + e] = (uint8_t)acc;
                }
    }

    uint32_t *d_p1w, *d_csv, *d_csm;
    uint64_t *d_scol;
    uint8_t *d_ab, *d_vn, *d_tv, *d_tm;
    int *d_boff;
    const size_t tbytes = (size_t)CHECK * KK * VV * MM;
    const int mblocks = NSIG / G;
    if (hipMalloc(&d_p1w, p1w.size() * 4) || hipMalloc(&d_scol, scol.size() * 8) || hipMalloc(&d_ab, ab.size()) ||
        hipMalloc(&d_vn, vn.size()) || hipMalloc(&d_boff, MTILES * 4) || hipMalloc(&d_csv, (size_t)NSIG * 64 * 4) ||
        hipMalloc(&d_csm, (size_t)mblocks * WAVES * 128 * 4) || hipMalloc(&d_tv, tbytes) ||
        hipMalloc(&d_tm, tbytes * 4)) {
        printf("{\"error\": \"allocation\"}\n");
        return 1;
    }
    (void)hipMemcpy(d_p1w, p1w.data(), p1w.size() * 4, hipMemcpyHostToDevice);
    (void)hipMemcpy(d_scol, scol.data(), scol.size() * 8, hipMemcpyHostToDevice);
    (void)hipMemcpy(d_ab, ab.data(), ab.size(), hipMemcpyHostToDevice);
    (void)hipMemcpy(d_vn, vn.data(), vn.size(), hipMemcpyHostToDevice);
    (void)hipMemcpy(d_boff, boff.data(), MTILES * 4, hipMemcpyHostToDevice);
    hipEvent_t e0, e1;
    (void)hipEventCreate(&e0);
    (void)hipEventCreate(&e1);
    auto time_ms = [&](auto launch) {
        float best = 1e30f;
        for (int rep = 0; rep < 4; ++rep) {
            (void)hipEventRecord(e0, nullptr);
            launch();
            (void)hipEventRecord(e1, nullptr);
            (void)hipEventSynchronize(e1);
            float ms = 0;
            (void)hipEventElapsedTime(&ms, e0, e1);
            if (rep > 0 && ms < best) best = ms;  // first run untimed
        }
        return best;
    };
    (void)hipMemset(d_tv, 0, tbytes);
    (void)hipMemset(d_tm, 0, tbytes * 4);
    const float valu_ms = time_ms([&] { k_valu<<<NSIG, 64>>>(d_p1w, d_scol, d_csv, d_tv); });
    const float mfma_ms = time_ms([&] {
        k_mfma<<<mblocks, 64 * WAVES>>>((const v4i *)d_ab, d_boff, bpk, d_vn, d_csm, (uint32_t *)d_tm);
    });
    uint8_t *d_tl;
    uint32_t *d_csl;
    if (hipMalloc(&d_tl, tbytes) || hipMalloc(&d_csl, (size_t)NSIG * 64 * 4)) {
        printf("{\"error\": \"allocation\"}\n");
        return 1;
    }
    (void)hipMemset(d_tl, 0, tbytes);
    const float lds_ms = time_ms([&] { k_ldst<<<NSIG / TW, 64 * TW>>>(d_p1w, d_scol, d_csl, d_tl); });
    const int err = (int)hipDeviceSynchronize() | (int)hipGetLastError();
    std::vector<uint8_t> tv(tbytes), tl(tbytes);
    (void)hipMemcpy(tl.data(), d_tl, tbytes, hipMemcpyDeviceToHost);
    std::vector<uint32_t> tm(tbytes);
    (void)hipMemcpy(tv.data(), d_tv, tbytes, hipMemcpyDeviceToHost);
    (void)hipMemcpy(tm.data(), d_tm, tbytes * 4, hipMemcpyDeviceToHost);
    size_t bad_v = 0, bad_m = 0, bad_l = 0;
    for (size_t x = 0; x < tbytes; ++x) {
        bad_v += tv[x] != tref[x];
        bad_m += (tm[x] & 0xFu) != tref[x];  // each of the four timed runs ORs the same bits in
        bad_l += tl[x] != tref[x];
    }
    const double macs = (double)NSIG * P1E * KK * MM;  // GF(16) MACs of the contraction
    printf("{\"contraction\": \"T_j[r] = sum_c P1[r][c] V_j[c], MAYO_1, %d signatures, %d keys\", "
           "\"valu_ms\": %.4f, \"mfma_ms\": %.4f, \"lds_table_ms\": %.4f, \"valu_tmacs\": %.2f, "
           "\"mfma_tmacs\": %.2f, \"lds_table_tmacs\": %.2f, "
           "\"mfma_instructions_per_signature\": %.1f, \"checked_signatures\": %d, "
           "\"valu_mismatches\": %zu, \"mfma_mismatches\": %zu, \"lds_table_mismatches\": %zu, "
           "\"hip_error\": %d}\n",
           NSIG, NKEY, valu_ms, mfma_ms, lds_ms, macs / (valu_ms * 1e-3) / 1e12, macs / (mfma_ms * 1e-3) / 1e12,
           macs / (lds_ms * 1e-3) / 1e12, (double)bpk * NTILES / G, CHECK, bad_v, bad_m, bad_l, err);
    return err != 0 || bad_v || bad_m || bad_l;
}
