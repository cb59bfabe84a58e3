// Wave-per-item signing and verification kernels.
//
// One wavefront (64 lanes, one workgroup) owns one signature end to end, so
// the per-signature scalars (the vinegar / signature nibbles) are uniform
// across the wave and their bit masks live in SGPRs. Phases:
//
//  M      M_i[.][x] = sum_r V_i[r] L[r][x]: lanes = (column x, word w) items,
//         r uniform.
//  U      U_ij = V_i^T P1 V_j + V_j^T P1 V_i: lanes = (word w, row slot);
//         the column index c is uniform, every lane accumulates T_j[r] for a
//         chunk of its rows, then folds T into the pair accumulators.
//  y, A   sums of z^l-shifted m-vectors, accumulated unreduced (degree
//         < m + pairs) and reduced once by z^m = F0 + F1 z + F2 z^2 + F3 z^3.
//  solve  Gauss-Jordan on the columns of [A | y - A r], each column held
//         bitsliced (4 bit-planes over the rows) in one lane's registers; the
//         pivot column is broadcast through SGPRs. The pivot is the lowest
//         unused row with a nonzero entry; x = r + z with z supported on the
//         pivot columns, identical to the reference's echelon form plus
//         back-substitution because that solution is unique.
#pragma once
#include <type_traits>

#include "keccak_wave.hpp"
#include "mayo_kernels.hpp"

namespace mayo {
namespace wv {

constexpr int SL = 64 / MW;                    // quad-form row slots per wave
constexpr int NG = (M + 31) / 32;              // 32-row groups of a bitsliced column
constexpr int NCOL = KO + 1;                   // columns of [A | y']
constexpr int NCL = (NCOL + 63) / 64;          // columns held per lane
constexpr int OVW = (PAIRS + 7) / 8 + 1;       // overflow words of an unreduced sum
constexpr int UW = MW + OVW;                   // words of an unreduced sum
constexpr int YG_MAX = 64 / UW;                // pair groups of the y sum (wave_y)
// quad-form rows per lane per chunk, and the waves per SIMD the quad-form
// kernels are compiled for (VGPR budget 512 / QF_OCC for the pair
// accumulators and QF_RT x 4K bit-plane accumulators). Two rows and three
// waves were the fastest setting measured for every parameter set; with one
// row the 4K scalar masks per column make the SALU the bottleneck for k >= 11.
#ifndef MAYO_QF_RT
#define MAYO_QF_RT 0
#endif
#ifndef MAYO_QF_OCC
#define MAYO_QF_OCC 0
#endif
constexpr int QF_RT = MAYO_QF_RT ? MAYO_QF_RT : 2;
constexpr int QF_OCC = MAYO_QF_OCC ? MAYO_QF_OCC : (K <= 4 ? 4 : 3);
// waves per SIMD of the solve kernel (the bitsliced columns of [A | y'])
constexpr int SOLVE_OCC = M <= 80 ? 8 : 3;
// Launch variants of the batch kernels over prepared keys that the launch
// autotuner chooses among (entry 0 is the default above). Quadratic-form
// kernels: code 10 x rows per lane + waves per SIMD they are built for (23 =
// two rows per lane per chunk, three waves per SIMD). One row halves the
// bit-plane accumulators (more waves) but shares each column's scalar masks
// over half as many rows; three rows share them over more rows at fewer
// waves and longer masked diagonals; two rows at one wave less avoid the
// spills the default has for MAYO_3. Solve kernel: waves per SIMD. Every
// variant computes the same bytes.
constexpr int qf_code(int rt, int occ) { return 10 * rt + occ; }
inline constexpr int PREP_OCCS[4] = {qf_code(QF_RT, QF_OCC), qf_code(1, K >= 12 ? 3 : QF_OCC + 1), qf_code(3, 2),
                                     qf_code(QF_RT, QF_OCC - 1)};
// Solve kernel codes: 100 (signatures per workgroup - 1) + 10 schedule +
// waves per SIMD. Schedule 0 searches the pivot row on the scalar unit from
// the broadcast pivot column, 1 has every lane search its own column on the
// vector unit (a VALU for SALU trade); with two signatures per workgroup the
// two waves keep their own LDS and never wait for each other.
inline constexpr int SOLVE_OCCS[5] = {SOLVE_OCC, M <= 80 ? 6 : 2, 4, 10 + SOLVE_OCC, 100 + SOLVE_OCC};
constexpr int solve_occ(int code) { return code % 10; }
constexpr int solve_schedule(int code) { return (code / 10) % 10; }
constexpr int solve_sigs(int code) { return code / 100 + 1; }
inline constexpr int VERIFY_OCCS[4] = {qf_code(QF_RT, QF_OCC), qf_code(QF_RT, QF_OCC - 1),
                                       qf_code(1, K >= 12 ? 3 : QF_OCC + 1), qf_code(3, 2)};

// Item of this workgroup: position base + blockIdx.x of the batch, or the
// item at that position of the XCD-aware order (k_key_order) when `ordered`
// is set. A launch over part of a batch passes the part's first position.
template <class Acc>
__device__ __forceinline__ size_t item_of(const Acc &acc, int ordered, size_t base = 0) {
    const size_t k = base + blockIdx.x;
    return ordered ? acc.order(k) : k;
}
// column c lives in lane col_lane(c), set col_set(c); the first SHIFT columns
// go to the last set so that only one set stays active for most pivots
constexpr int SHIFT = NCOL - 64 * (NCL - 1);

__host__ __device__ constexpr int col_of(int set, int lane) {
    // set NCL-1 holds columns [0, SHIFT) (its other lanes hold nothing: an
    // index past NCOL); sets 0..NCL-2 hold [SHIFT + 64 s, SHIFT + 64 s + 64)
    return set == NCL - 1 ? (lane < SHIFT ? lane : NCOL + lane) : SHIFT + 64 * set + lane;
}
// order in which the sets' columns come: set NCL-1 first, then 0, 1, ...
__host__ __device__ constexpr int set_rank(int set) { return set == NCL - 1 ? 0 : set + 1; }
// the y' column (index KO = NCOL - 1) is the last lane of set NCL-2
constexpr int Y_SET = NCL - 2;
constexpr int Y_LANE = 63;
static_assert(NCL >= 2 && col_of(Y_SET, Y_LANE) == KO, "y' column placement");

__device__ __forceinline__ uint32_t uni(uint32_t x) { return __builtin_amdgcn_readfirstlane(x); }
__device__ __forceinline__ uint64_t uni64(uint64_t x) {
    return ((uint64_t)uni((uint32_t)(x >> 32)) << 32) | (uint64_t)uni((uint32_t)x);
}

// The lane within the wavefront (threadIdx.x mod 64) behind an empty asm
// statement. The phases of the whole-loop signing kernels derive per-lane
// constants from it; taken this way inside each phase they are recomputed
// there, instead of being hoisted out of the attempt loop and held (spilled)
// across the quadratic form.
__device__ __forceinline__ int lane_here() {
    int l = (int)(threadIdx.x & 63u);
    asm volatile("" : "+v"(l));
    return l;
}

// f(integral_constant<int, i>) for i = I .. N-1, in order
template <int N, int I = 0, class F>
__device__ __forceinline__ void static_for(F &&f) {
    if constexpr (I < N) {
        f(std::integral_constant<int, I>{});
        static_for<N, I + 1>(f);
    }
}

// Scalar-unit ops on wave-uniform values. Written out so that the compiler
// keeps them on the SALU: left to itself it fuses and-not / xor chains of
// uniform values into v_bitop3 on the vector unit, which is the busier port
// of the elimination.
__device__ __forceinline__ uint32_t s_xor(uint32_t a, uint32_t b) {
    uint32_t r;
    asm("s_xor_b32 %0, %1, %2" : "=s"(r) : "s"(a), "s"(b));
    return r;
}
__device__ __forceinline__ uint32_t s_andn2(uint32_t a, uint32_t b) {
    uint32_t r;
    asm("s_andn2_b32 %0, %1, %2" : "=s"(r) : "s"(a), "s"(b));
    return r;
}

// Barrier policies of the helpers below: the whole workgroup, or only the
// one wave that runs them while the others wait at the caller's next
// barrier. A wave's LDS accesses complete in program order, so one wave
// needs only a fence (no compiler reordering across it, own LDS operations
// complete).
struct BlockSync {
    __device__ static void sync() { __syncthreads(); }
};
struct WaveSync {
    __device__ static void sync() { __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup"); }
};

// XOR of x over the 64 lanes of the wave (uniform result; every lane of the
// wave must be active): within each row of 16 lanes by DPP (quad swaps, then
// rotations by 4 and 8), then across the four rows on the scalar unit.
__device__ __forceinline__ uint32_t wave_xor(uint32_t x) {
    x ^= (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0xB1, 0xF, 0xF, false);   // quad_perm [1,0,3,2]
    x ^= (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x4E, 0xF, 0xF, false);   // quad_perm [2,3,0,1]
    x ^= (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x124, 0xF, 0xF, false);  // row_ror:4
    x ^= (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x128, 0xF, 0xF, false);  // row_ror:8
    return __builtin_amdgcn_readlane(x, 0) ^ __builtin_amdgcn_readlane(x, 16) ^
           __builtin_amdgcn_readlane(x, 32) ^ __builtin_amdgcn_readlane(x, 48);
}

// Multiply a bitsliced nibble vector (planes p0..p3) by a.
__device__ __forceinline__ void bs_mul_a(uint32_t &p0, uint32_t &p1, uint32_t &p2, uint32_t &p3) {
    uint32_t t = p3;
    p3 = p2; p2 = p1; p1 = p0 ^ t; p0 = t;
}

// Bit d of the 8 nibbles of w, packed into 8 consecutive bits.
__device__ __forceinline__ uint32_t nib_plane8(uint32_t w, int d) {
    uint32_t x = (w >> d) & 0x11111111u;
    x = (x | (x >> 3)) & 0x03030303u;
    x = (x | (x >> 6)) & 0x000F000Fu;
    return (x | (x >> 12)) & 0xFFu;
}

struct Shared {
    uint8_t vr[VR_BYTES];
    uint64_t scol[N];            // scol[c] nibble j = S_j[c]
    uint32_t Mv[K * O * MW];     // M_i column x as an m-vector
    union {
        uint32_t U[PAIRS * MW];       // pair sums (quadratic form, then y)
        uint32_t Mb[KO * 4 * NG];     // bit planes of the M_j[x] (columns of A)
    };
    uint32_t yu[YG_MAX * UW];    // unreduced y (partial sums of the pair groups)
    uint32_t y[MW];              // y = t + sum E^l u_l
    uint32_t t[MW];
    uint32_t yplanes[4 * NG];    // final y' column, bitsliced
    int16_t prow[KO];            // pivot row of each column, -1 if free
    uint8_t pinv[KO];            // inverse of each column's pivot value
    uint8_t x[KO];
    uint8_t mul[256];            // GF(16) product table, a*16 + b
    uint8_t hbuf[DIGEST + 2 * SK + 1];  // hash input staging
};

// ---- key-row layouts ----------------------------------------------------------
// Reference rows (esk = sk || O || P1 || L, epk = P1 || P2 || P3): entries
// MB bytes apart at any alignment. Prepared rows: the same blocks with every
// entry padded with zero bytes to MBP (a multiple of 4) and each block
// 16-byte aligned, so that a word of an entry is one aligned dword load, and
// every entry word stored planar (gf16::planar), which makes the per-row
// multiplications by a of the quadratic form a third as costly.
constexpr int MBP = (MB + 3) & ~3;

// AHEAD: the columns the quad-form column loop loads ahead. Loading 1, 2 and
// 3 ahead took MAYO_1 stage 2 over prepared keys (64 shared) from 5.58 to
// 5.41 to 5.36 ms and MAYO_2 over 65,536 distinct keys from 3.01 to 2.71 to
// 2.66 ms; the larger sets' verification kernel spills with 3, and with 4
// every set's does. A reference-layout load takes three registers, so those
// rows load one ahead, except in pitched pools, where every item reads its
// own key row from HBM: there MAYO_2, whose kernels have registers to spare,
// loads three ahead (drop-in signing and verification over 65,536 distinct
// keys: 4.50 -> 4.62 M/s and 9.27 -> 9.41 M/s; the same depth made tables of
// 64 shared keys slower).
struct RefKeys {
    static constexpr bool ALIGNED = false;
    static constexpr bool PLANAR = false;
    static constexpr int EB = MB;
    static constexpr int AHEAD = 1;
    static constexpr uint32_t ESK_P1 = SK + O_BYTES, ESK_L = ESK_P1 + P1_BYTES;
    static constexpr uint32_t EPK_P1 = 0, EPK_P2 = P1_BYTES, EPK_P3 = P1_BYTES + P2_BYTES;
};
struct PitchedRefKeys : RefKeys {
    static constexpr int AHEAD = M <= 64 ? 3 : 1;
};
// Reference rows whose entries all start on 4-byte boundaries: whole-word
// entries (MB a multiple of 4, MAYO_2), every block offset a multiple of 4,
// and a pool whose base and pitch are (checked at launch). A word of an
// entry is then one aligned dword load, as in prepared rows.
constexpr bool REF_WORD_ALIGNED = MB % 4 == 0 && (SK + O_BYTES) % 4 == 0 && P1_BYTES % 4 == 0 && P2_BYTES % 4 == 0;
struct PitchedAlignedKeys : PitchedRefKeys {
    static constexpr bool ALIGNED = true;
};
struct PrepKeys {
    static constexpr bool ALIGNED = true;
    static constexpr bool PLANAR = true;
    static constexpr int AHEAD = M <= 80 ? 3 : 2;
    static constexpr int EB = MBP;
    static constexpr uint32_t ESK_P1 = (SK + O_BYTES + 15) / 16 * 16;
    static constexpr uint32_t ESK_L = ESK_P1 + P1_ENTRIES * MBP;
    static constexpr uint32_t ESK_BYTES = (uint32_t)align128(ESK_L + P2_ENTRIES * MBP);
    static constexpr uint32_t EPK_P1 = 0, EPK_P2 = P1_ENTRIES * MBP;
    static constexpr uint32_t EPK_P3 = EPK_P2 + P2_ENTRIES * MBP;
    static constexpr uint32_t EPK_BYTES = (uint32_t)align128(EPK_P3 + P3_ENTRIES * MBP);
};

// Rewrite reference key rows in the prepared layout. Each destination dword
// is assembled from the source bytes it maps to: the header bytes [0, hdr)
// are copied, region r maps entry e (MB bytes at src + e MB) to dst + e MBP
// with zero padding and each of its dwords made planar, everything else is
// zero. Grid: (dwords per row / 256 rounded up) x keys.
struct PrepRegion {
    uint32_t src, dst, count;
};
__global__ void __launch_bounds__(256) k_prepare_keys(const uint8_t *__restrict__ src, size_t src_stride,
                                                      uint8_t *__restrict__ dst, size_t dst_stride,
                                                      uint32_t dst_bytes, uint32_t hdr, PrepRegion r0,
                                                      PrepRegion r1, PrepRegion r2) {
    const uint32_t o = 4 * (blockIdx.x * 256 + threadIdx.x);
    if (o >= dst_bytes) return;
    const uint8_t *s = src + (size_t)blockIdx.y * src_stride;
    // region starts are 16-byte aligned and MBP a multiple of 4: a dword
    // lies in one entry of one region or in none
    const PrepRegion rg[3] = {r0, r1, r2};
    bool in_region = false;
#pragma unroll
    for (int r = 0; r < 3; ++r) in_region |= o >= rg[r].dst && o < rg[r].dst + rg[r].count * (uint32_t)MBP;
    uint32_t v = 0;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const uint32_t ob = o + i;
        uint32_t byte = 0;
        if (ob < hdr) {
            byte = s[ob];
        } else {
#pragma unroll
            for (int r = 0; r < 3; ++r) {
                if (ob >= rg[r].dst && ob < rg[r].dst + rg[r].count * (uint32_t)MBP) {
                    const uint32_t e = (ob - rg[r].dst) / MBP, q = (ob - rg[r].dst) % MBP;
                    if (q < (uint32_t)MB) byte = s[rg[r].src + e * MB + q];
                }
            }
        }
        v |= byte << (8 * i);
    }
    if (in_region) v = gf16::planar::from_nibbles(v);
    *(uint32_t *)(dst + (size_t)blockIdx.y * dst_stride + o) = v;
}

// Word w of a key entry (byte offset off from base, nv valid bytes), in two
// steps so that the loads of the next column are in flight while the
// current one is used: fetch issues the loads, word() assembles the value.
template <class KL>
struct EntryLoad {
    uint32_t lo, hi, sh;
    __device__ __forceinline__ void fetch(const uint8_t *base, uint32_t off, int nv) {
        if constexpr (KL::ALIGNED) {
            lo = *(const uint32_t *)(base + off);  // padding bytes are zero
        } else {
            const uintptr_t a = (uintptr_t)(base + off);
            sh = (uint32_t)(a & 3);
            const uint32_t *q = (const uint32_t *)(a - sh);
            lo = q[0];
            hi = (sh + (uint32_t)nv > 4) ? q[1] : 0u;  // only the dwords holding the entry's bytes
        }
    }
    __device__ __forceinline__ uint32_t word(int nv) const {
        if constexpr (KL::ALIGNED) {
            return lo;
        } else {
            const uint32_t v = __builtin_amdgcn_alignbyte(hi, lo, sh);
            return nv < 4 ? v & ((1u << (8 * nv)) - 1u) : v;
        }
    }
};

// ---- the oil space O, packed by rows ---------------------------------------------
constexpr int OWD = (O + 7) / 8;            // words of one packed row of O

// ocol[c * OWD + q]: nibbles x = 8q .. 8q+7 of row c of O.
__device__ __forceinline__ void load_ocol(const uint8_t *Onib, uint32_t *ocol) {
    for (int c = threadIdx.x; c < V; c += 64)
#pragma unroll
        for (int q = 0; q < OWD; ++q) {
            uint32_t v = 0;
            for (int x = 8 * q; x < O && x < 8 * q + 8; ++x) v |= mvec::nib(Onib, c * O + x) << (4 * (x - 8 * q));
            ocol[c * OWD + q] = v;
        }
}

// ---- quadratic form -------------------------------------------------------------
// T_j[row] = sum_c P[row][c] S_j[c] for every row a lane owns, with c uniform
// across the wave. Writing S_j[c] = sum_b s_jb(c) a^b (bits s_jb), T_j =
// sum_b a^b T_jb with T_jb = XOR of the entries whose scalar has bit b set:
// one AND-XOR per entry word and scalar bit (SGPR masks), and the powers of
// a are applied once per row at the end (Horner), not per entry.
//
// Rows: with NLR > 0 (signing) the first NLR rows are the columns x of L,
// entries L[c][x] over every c, so T_j of row x is M_j[x] (sign.cu:256-330);
// then the rows r of P1 (entries c >= r; sign.cu:395-437). With FULL
// (verification) the rows are those of [[P1, P2], [0, P3]] over c < n
// (verify.cu:72-182). Each P row then adds S_a[r] T_j[r] into the pair
// accumulators U[pair(a, j)] (ordered (a, j) both ways), XOR-reduced into
// Ulds at the end; each L row stores M_j[x] into Mv. Words keep the key's
// form (planar for KL::PLANAR) throughout; Ulds is returned nibble-packed,
// Mv in the key's form.
//
// Lanes = (word w, row slot); a chunk gives each slot RT rows, rows in
// ascending start column so that every row of a chunk starts within a few
// columns of the first (those columns are masked per lane).
//
// LF (signing without L, from a key row that holds P2 where L would be):
// the first NLR rows read P2[c][x], which gives the P2 V_j part of M_j[x],
// and the rest, sum_r O[r][x] W_j[r] with W_j = (P1 + P1^T) S_j, is added
// per chunk. Once a P1 row's T_j[r] has gone into U, the row goes on over
// the entries P1[c][r], c <= r (the column r of P1; the diagonal, counted
// twice, cancels), so its accumulators then hold W_j[r]; the chunk's W rows
// go to Wc (LDS, RT SL K MW words), and lanes (j, w) fold them into Mv with
// the SGPR masks of O[r][x] (ocol, packed rows of O). One wave only.
// Costs K (P1 entries + V O) MACs per signature, against V^2 O for L.
template <int NLR, bool FULL, int RT = QF_RT>
constexpr int qf_chunks() {
    return (NLR + (FULL ? N : V) + RT * SL - 1) / (RT * SL);
}
template <class KL, int NLR, bool FULL, int NW = 1, int RT = QF_RT, bool LF = false>
__device__ __forceinline__ void quad_form(const uint8_t *key, const uint64_t *scol, uint32_t *Ulds,
                                          uint32_t *Mv, const uint32_t *ocol = nullptr,
                                          uint32_t *Wc = nullptr) {
    static_assert(!LF || (NW == 1 && !FULL && NLR == O), "the L-free form is one wave's signing");
    constexpr int NP = FULL ? N : V;          // P rows and columns
    constexpr int NVR = NLR + NP;             // rows of the schedule
    constexpr int RPC = RT * SL;           // rows per chunk
    constexpr int NCH = qf_chunks<NLR, FULL, RT>();
    constexpr int EB = KL::EB;
    constexpr uint32_t P1o = FULL ? KL::EPK_P1 : KL::ESK_P1;
    // NW waves of the workgroup share the chunks (chunk ch goes to wave ch % NW)
    const int lane = lane_here() & 63;
    const int wave = NW > 1 ? (int)(threadIdx.x >> 6) : 0;
    const int w = lane % MW, slot = lane / MW;
    const bool active = slot < SL;
    const int nv = mvec::word_valid(w);
    auto start_of = [](int vr) { return vr < NLR ? 0 : vr - NLR; };
    uint32_t U[PAIRS];
#pragma unroll
    for (int p = 0; p < PAIRS; ++p) U[p] = 0;
#pragma unroll 1
    for (int ch = wave; ch < NCH; ch += NW) {
        const int v0 = ch * RPC;
        const int vlast = (v0 + RPC < NVR ? v0 + RPC : NVR) - 1;
        const int c_lo = start_of(v0), c_free = start_of(vlast);
        int st[RT];
        uint32_t off[RT], stride[RT];
        uint32_t Tb[RT][K][4];
#pragma unroll
        for (int tt = 0; tt < RT; ++tt) {
            const int vr = v0 + tt * SL + slot;
            const bool valid = active && vr < NVR;
            st[tt] = valid ? start_of(vr) : NP;
            off[tt] = 0;
            stride[tt] = 0;
            if (valid) {
                if (vr < NLR) {
                    off[tt] = KL::ESK_L + (uint32_t)((c_lo * O + vr) * EB);
                    stride[tt] = O * EB;
                } else {
                    const int r = vr - NLR;
                    if (r < V) {
                        off[tt] = P1o + (uint32_t)((tri_entry(r, r, V) - (r - c_lo)) * EB);
                        stride[tt] = EB;
                    }
                }
            }
            off[tt] += 4 * w;
#pragma unroll
            for (int j = 0; j < K; ++j)
#pragma unroll
                for (int b = 0; b < 4; ++b) Tb[tt][j][b] = 0;
        }
        // one column: every row of the chunk takes its entry at c, fetched
        // into load set S during an earlier column; off[tt] is the next
        // column to fetch
        constexpr int NSET = KL::AHEAD;  // load sets
        EntryLoad<KL> ld[NSET][RT];
        using Set0 = std::integral_constant<int, 0>;
        using Set1 = std::integral_constant<int, (NSET > 1 ? 1 : 0)>;
        using Set2 = std::integral_constant<int, (NSET > 2 ? 2 : 0)>;
        auto fetch = [&](auto S) {
#pragma unroll
            for (int tt = 0; tt < RT; ++tt) {
                ld[decltype(S)::value][tt].fetch(key, off[tt], nv);
                off[tt] += stride[tt];
            }
        };
        // the rows of slot group tt start at column tt_start[tt] or later
        // (uniform): left of the last group's start fewer groups are summed
        int tt_start[RT];
#pragma unroll
        for (int tt = 0; tt < RT; ++tt) tt_start[tt] = v0 + tt * SL < NVR ? start_of(v0 + tt * SL) : NP;
        // groups 0 .. NTT-1 take column c
        auto column = [&](auto NT, auto S, int c, bool masked, bool more) {
            constexpr int NTT = decltype(NT)::value;
            const uint64_t sc = uni64(scol[c]);
            // masks from the 32-bit halves: one scalar bit-field extract each
            const uint32_t sh[2] = {(uint32_t)sc, (uint32_t)(sc >> 32)};
            uint32_t x[RT];
#pragma unroll
            for (int tt = 0; tt < RT; ++tt) {
                x[tt] = ld[decltype(S)::value][tt].word(nv);
                if (masked) x[tt] = c >= st[tt] ? x[tt] : 0u;
            }
            if (more) fetch(S);
#pragma unroll
            for (int j = 0; j < K; ++j)
#pragma unroll
                for (int b = 0; b < 4; ++b) {
                    const int bit = 4 * j + b;
                    const uint32_t m = gf16::bitmask(sh[bit >> 5], bit & 31);
#pragma unroll
                    for (int tt = 0; tt < NTT; ++tt) Tb[tt][j][b] = gf16::axm(Tb[tt][j][b], x[tt], m);
                }
        };
        using All = std::integral_constant<int, RT>;
        // columns [c0, cend), masked below cm, in three loops: only group 0
        // started (RT = 2), all groups masked (loads one column ahead), and
        // unmasked, the long one, in runs of NSET = KL::AHEAD columns over the
        // NSET load sets, loading NSET columns ahead (set s holds column c + s,
        // then takes c + s + NSET).
        auto segment = [&](int c0, int cm, int cend) {
            if (c0 < cend) fetch(Set0{});
            int c = c0;
            if constexpr (RT == 2) {
                const int c1 = tt_start[1] < cm ? tt_start[1] : cm;
#pragma unroll 1
                for (; c < c1; ++c) column(std::integral_constant<int, 1>{}, Set0{}, c, true, c + 1 < cend);
            }
#pragma unroll 1
            for (; c < cm; ++c) column(All{}, Set0{}, c, true, c + 1 < cend);
            if constexpr (NSET == 3) {
                // set 0 holds column c; sets 1 and 2 take c + 1 and c + 2
                if (c + 1 < cend) fetch(Set1{});
                if (c + 2 < cend) fetch(Set2{});
#pragma unroll 1
                for (; c + 2 < cend; c += 3) {
                    column(All{}, Set0{}, c, false, c + 3 < cend);
                    column(All{}, Set1{}, c + 1, false, c + 4 < cend);
                    column(All{}, Set2{}, c + 2, false, c + 5 < cend);
                }
                // fewer than three columns left, already loaded
                if (c < cend) column(All{}, Set0{}, c, false, false);
                if (c + 1 < cend) column(All{}, Set1{}, c + 1, false, false);
            } else if constexpr (NSET == 2) {
                if (c + 1 < cend) fetch(Set1{});
#pragma unroll 1
                for (; c + 1 < cend; c += 2) {
                    column(All{}, Set0{}, c, false, c + 2 < cend);
                    column(All{}, Set1{}, c + 1, false, c + 3 < cend);
                }
                if (c < cend) column(All{}, Set0{}, c, false, false);
            } else {
#pragma unroll 1
                for (; c < cend; ++c) column(All{}, Set0{}, c, false, c + 1 < cend);
            }
        };
        // columns < V (P1 and L entries)
        if (c_lo < V) segment(c_lo, c_free < V ? c_free : V, V);
        if constexpr (FULL) {
            // columns >= V: P2 for rows < V, P3 for rows >= V
            const int cb = c_lo > V ? c_lo : V;
#pragma unroll
            for (int tt = 0; tt < RT; ++tt) {
                const int r = v0 + tt * SL + slot;
                if (active && r < NP) {
                    if (r < V) {
                        off[tt] = KL::EPK_P2 + (uint32_t)((r * O + (cb - V)) * EB) + 4 * w;
                    } else {
                        const int r3 = r - V;
                        off[tt] = KL::EPK_P3 +
                                  (uint32_t)((tri_entry(r3, r3, O) - (r3 - (cb - V))) * EB) + 4 * w;
                    }
                    stride[tt] = EB;
                }
            }
            segment(cb, c_free > cb ? c_free : cb, N);
        }
        // rows done: T_j = sum_b a^b T_jb
#pragma unroll
        for (int tt = 0; tt < RT; ++tt) {
            const int vr = v0 + tt * SL + slot;
            uint32_t T[K];
#pragma unroll
            for (int j = 0; j < K; ++j) {
                if constexpr (KL::PLANAR) {
                    using namespace gf16::planar;
                    T[j] = mul_a_add(mul_a_add(mul_a_add(Tb[tt][j][3], Tb[tt][j][2]), Tb[tt][j][1]), Tb[tt][j][0]);
                } else {
                    T[j] = gf16::mul_a(gf16::mul_a(gf16::mul_a(Tb[tt][j][3]) ^ Tb[tt][j][2]) ^ Tb[tt][j][1]) ^
                           Tb[tt][j][0];
                }
            }
            if (!active || vr >= NVR) continue;
            if (vr < NLR) {
#pragma unroll
                for (int j = 0; j < K; ++j) Mv[(j * O + vr) * MW + w] = T[j];
            } else {
                const uint64_t sr = scol[vr - NLR];
                uint32_t am[K][4];
#pragma unroll
                for (int a = 0; a < K; ++a)
#pragma unroll
                    for (int b = 0; b < 4; ++b) am[a][b] = gf16::bitmask((uint32_t)(sr >> (4 * a)), b);
#pragma unroll
                for (int j = 0; j < K; ++j) {
                    const std::conditional_t<KL::PLANAR, gf16::planar::Multiples, gf16::Multiples> mt(T[j]);
#pragma unroll
                    for (int a = 0; a < K; ++a) {
                        const int p = a <= j ? pair_index(a, j) : pair_index(j, a);
                        U[p] = gf16::mac(U[p], mt.x0, mt.x1, mt.x2, mt.x3, am[a][0], am[a][1], am[a][2],
                                         am[a][3]);
                    }
                }
            }
        }
        if constexpr (LF) {
            // P1 rows of the chunk: r_first .. r_last
            const int vp = v0 > NLR ? v0 : NLR;
            if (vp <= vlast) {
                const int r_first = vp - NLR, r_last = vlast - NLR;
                int rl[RT];
                uint32_t lbase[RT];
#pragma unroll
                for (int tt = 0; tt < RT; ++tt) {
                    const int vr = v0 + tt * SL + slot;
                    rl[tt] = active && vr >= NLR && vr <= vlast ? vr - NLR : -1;
                    // entry (c, r) = tri(c, c) - c + r; rows without a P1 row
                    // read (c, 0)'s address (in P1) and are not used
                    lbase[tt] = P1o + (uint32_t)((rl[tt] > 0 ? rl[tt] : 0) * EB) + 4 * w;
                }
                constexpr int NSET = KL::AHEAD;
                EntryLoad<KL> ld[NSET][RT];
                auto lfetch = [&](auto S, int c) {
                    const uint32_t lb = (uint32_t)((c * (V - 1) - c * (c - 1) / 2) * EB);
#pragma unroll
                    for (int tt = 0; tt < RT; ++tt) ld[decltype(S)::value][tt].fetch(key, lbase[tt] + lb, nv);
                };
                // columns 0 .. r_last; below r_first every row of the chunk
                // takes the entry, from r_first on only rows r >= c
                static_for<NSET>([&](auto S) {
                    if (decltype(S)::value <= r_last) lfetch(S, decltype(S)::value);
                });
#pragma unroll 1
                for (int c0 = 0; c0 <= r_last; c0 += NSET) {
                    static_for<NSET>([&](auto S) {
                        constexpr int s = decltype(S)::value;
                        const int c = c0 + s;
                        if (c <= r_last) {
                            const uint64_t sc = uni64(scol[c]);
                            const uint32_t sh[2] = {(uint32_t)sc, (uint32_t)(sc >> 32)};
                            uint32_t x[RT];
#pragma unroll
                            for (int tt = 0; tt < RT; ++tt) x[tt] = ld[s][tt].word(nv);
                            if (c >= r_first) {
#pragma unroll
                                for (int tt = 0; tt < RT; ++tt) x[tt] = c <= rl[tt] ? x[tt] : 0u;
                            }
                            if (c + NSET <= r_last) lfetch(S, c + NSET);
#pragma unroll
                            for (int j = 0; j < K; ++j)
#pragma unroll
                                for (int b = 0; b < 4; ++b) {
                                    const int bit = 4 * j + b;
                                    const uint32_t m = gf16::bitmask(sh[bit >> 5], bit & 31);
#pragma unroll
                                    for (int tt = 0; tt < RT; ++tt) Tb[tt][j][b] = gf16::axm(Tb[tt][j][b], x[tt], m);
                                }
                        }
                    });
                }
                // W_j[r] of the chunk's rows into Wc (row index vr - v0)
#pragma unroll
                for (int tt = 0; tt < RT; ++tt) {
                    if (rl[tt] < 0) continue;
#pragma unroll
                    for (int j = 0; j < K; ++j) {
                        uint32_t Wj;
                        if constexpr (KL::PLANAR) {
                            using namespace gf16::planar;
                            Wj = mul_a_add(mul_a_add(mul_a_add(Tb[tt][j][3], Tb[tt][j][2]), Tb[tt][j][1]), Tb[tt][j][0]);
                        } else {
                            Wj = gf16::mul_a(gf16::mul_a(gf16::mul_a(Tb[tt][j][3]) ^ Tb[tt][j][2]) ^ Tb[tt][j][1]) ^
                                 Tb[tt][j][0];
                        }
                        Wc[((tt * SL + slot) * K + j) * MW + w] = Wj;
                    }
                }
                WaveSync::sync();
                // M_j[x] += sum_r O[r][x] W_j[r] over the chunk's rows
                for (int it = lane; it < K * MW; it += 64) {
                    const int j = it / MW, ww = it % MW;
                    uint32_t acc[O];
#pragma unroll
                    for (int x = 0; x < O; ++x) acc[x] = 0;
#pragma unroll 1
                    for (int vr = vp; vr <= vlast; ++vr) {
                        const int r = vr - NLR;
                        const std::conditional_t<KL::PLANAR, gf16::planar::Multiples, gf16::Multiples> mt(
                            Wc[((vr - v0) * K + j) * MW + ww]);
                        uint32_t oc[OWD];
#pragma unroll
                        for (int q = 0; q < OWD; ++q) oc[q] = uni(ocol[r * OWD + q]);
#pragma unroll
                        for (int x = 0; x < O; ++x) {
                            const uint32_t ox = oc[x >> 3] >> (4 * (x & 7));
                            acc[x] = gf16::mac(acc[x], mt.x0, mt.x1, mt.x2, mt.x3, gf16::bitmask(ox, 0),
                                               gf16::bitmask(ox, 1), gf16::bitmask(ox, 2), gf16::bitmask(ox, 3));
                        }
                    }
#pragma unroll
                    for (int x = 0; x < O; ++x) Mv[(j * O + x) * MW + ww] ^= acc[x];
                }
                WaveSync::sync();  // Wc is rewritten by the next chunk
            }
        }
    }
    if (active) {
#pragma unroll
        for (int p = 0; p < PAIRS; ++p) atomicXor(&Ulds[p * MW + w], U[p]);
    }
    if constexpr (KL::PLANAR) {
        __syncthreads();
        for (int e = threadIdx.x; e < PAIRS * MW; e += 64 * NW) Ulds[e] = gf16::planar::to_nibbles(Ulds[e]);
    }
}

// ---- unreduced shifted sums ---------------------------------------------------
// acc (UW words) ^= z^ell * src (an m-vector of MW words in LDS), no reduction.
__device__ __forceinline__ void add_shift(uint32_t (&acc)[UW], const uint32_t *src, int ell) {
    const int ws = ell >> 3, ns = 4 * (ell & 7);
#pragma unroll
    for (int q = 0; q < UW; ++q) {
        const int hi = q - ws, lo = q - ws - 1;
        uint32_t h = (hi >= 0 && hi < MW) ? src[hi] : 0u;
        uint32_t l = (lo >= 0 && lo < MW) ? src[lo] : 0u;
        acc[q] ^= ns ? __builtin_amdgcn_alignbit(h, l, 32 - ns) : h;
    }
}

// Reduce an unreduced sum modulo f(z) into MW words: the nibbles at degree
// m + t (t < pairs) fold back as z^t (F0 + F1 z + F2 z^2 + F3 z^3).
__device__ __forceinline__ void reduce_unreduced(const uint32_t (&acc)[UW], uint32_t (&out)[MW]) {
    constexpr int QS = M / 8, QN = 4 * (M % 8);
    uint32_t ov[OVW + 1];
#pragma unroll
    for (int k = 0; k <= OVW; ++k) {
        const int a = QS + k, b = QS + k + 1;
        uint32_t lo = a < UW ? acc[a] : 0u;
        uint32_t hi = b < UW ? acc[b] : 0u;
        ov[k] = QN ? __builtin_amdgcn_alignbit(hi, lo, QN) : lo;
    }
#pragma unroll
    for (int q = 0; q < MW; ++q) out[q] = acc[q];
    out[MW - 1] &= mvec::LAST_MASK;
#pragma unroll
    for (int k = 0; k <= OVW && k < MW; ++k) {
        gf16::Multiples mo(ov[k]);
        const uint32_t t0 = mo.times(F0), t1 = mo.times(F1), t2 = mo.times(F2), t3 = mo.times(F3);
        out[k] ^= t0 ^ (t1 << 4) ^ (t2 << 8) ^ (t3 << 12);
        if (k + 1 < MW) out[k + 1] ^= (t1 >> 28) ^ (t2 >> 24) ^ (t3 >> 20);
    }
}

// Nibble-packed m-vector (MW words) -> 4 bit-planes of NG words over rows.
__device__ __forceinline__ void to_planes(const uint32_t (&v)[MW], uint32_t (&pl)[4][NG]) {
#pragma unroll
    for (int d = 0; d < 4; ++d)
#pragma unroll
        for (int g = 0; g < NG; ++g) {
            uint32_t o = 0;
#pragma unroll
            for (int s = 0; s < 4; ++s) {
                const int wi = 4 * g + s;
                if (wi < MW) o |= nib_plane8(v[wi], d) << (8 * s);
            }
            pl[d][g] = o;
        }
}

// Horner order of the pair (i, j): l(i, j) in the A schedule (sign.cu:336-392).
__host__ __device__ constexpr int ell_of(int i, int j) {
    // rows i' < i contribute K - i' pairs each; within row i, j runs K-1 down
    return i * K - i * (i - 1) / 2 + (K - 1 - j);
}

// Unreduced sum_l z^l u_l over the pair sums U (nibble-packed, MW words per
// pair, pair p = pair_index(i, j) shifted by l = ell_of(i, j)). Lanes are
// (word q of the sum, group g of YPG consecutive pairs); the groups' partial
// sums meet in ypart (YG * UW words), whose first UW words hold the sum on
// return.
constexpr int YG = YG_MAX;
constexpr int YPG = (PAIRS + YG - 1) / YG;
template <class Sync>
__device__ __forceinline__ void unreduced_y(const uint32_t *U, uint32_t *ypart) {
    const int lane = lane_here() & 63, q = lane % UW, g = lane / UW;
    if (g < YG) {
        // first pair (i, j) of the group
        int p = g * YPG, i = 0;
        while (pair_index(i + 1, i + 1) <= p && i + 1 < K) ++i;
        int j = i + (p - pair_index(i, i));
        uint32_t acc = 0;
        for (int k = 0; k < YPG && p < PAIRS; ++k, ++p) {
            const int ell = ell_of(i, j), ws = ell >> 3, ns = 4 * (ell & 7);
            const uint32_t *src = U + p * MW;
            const int hi = q - ws, lo = q - ws - 1;
            const uint32_t h = (hi >= 0 && hi < MW) ? src[hi] : 0u;
            const uint32_t l = (lo >= 0 && lo < MW) ? src[lo] : 0u;
            acc ^= ns ? __builtin_amdgcn_alignbit(h, l, 32 - ns) : h;
            if (++j == K) j = ++i;
        }
        ypart[g * UW + q] = acc;
    }
    Sync::sync();
    if (lane < UW) {
        uint32_t a = ypart[lane];
#pragma unroll
        for (int gg = 1; gg < YG; ++gg) a ^= ypart[gg * UW + lane];
        ypart[lane] = a;
    }
    Sync::sync();
}

// y = t + sum_l E^l u_l computed by the wave; result in S.y (MW words).
template <class Sync = BlockSync>
__device__ __forceinline__ void wave_y(Shared &S, const uint32_t *t) {
    const int lane = lane_here() & 63;
    unreduced_y<Sync>(S.U, S.yu);
    if (lane == 0) {
        uint32_t acc[UW], out[MW];
#pragma unroll
        for (int q = 0; q < UW; ++q) acc[q] = S.yu[q];
        reduce_unreduced(acc, out);
#pragma unroll
        for (int q = 0; q < MW; ++q) S.y[q] = out[q] ^ t[q];
    }
    Sync::sync();
}

// GF(16) product of a per-lane nibble and a uniform nibble via the LDS table.
template <class LDS>
__device__ __forceinline__ uint32_t tmul(const LDS &S, uint32_t a, uint32_t b) {
    return S.mul[(a << 4) | b];
}

// LDS of the solve kernel: the bit planes of M, y and the elimination
// bookkeeping (the signing-loop kernel uses the same fields of Shared).
struct SolveShared {
    uint8_t vr[VR_BYTES];
    uint32_t Mb[KO * 4 * NG];
    uint32_t y[MW];
    uint32_t yplanes[4 * NG];
    int16_t prow[KO];
    uint8_t pinv[KO];
    uint8_t x[KO];
    uint8_t mul[256];
};

// Bit planes of the K*O vectors M_j[x] (MW words each, nibble-packed or
// planar (PL), vector v = j*O + x at src + v*MW, global or LDS):
// Mb[(v*4 + d)*NG + g] holds bit d of rows 32g .. 32g+31.
template <bool PL, class SRC>
__device__ __forceinline__ void planes_of_M(const SRC *src, uint32_t *Mb) {
    for (int it = lane_here(); it < KO * NG; it += 64) {
        const int v = it / NG, g = it % NG;
        uint32_t w[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) w[q] = (4 * g + q < MW) ? src[v * MW + 4 * g + q] : 0u;
#pragma unroll
        for (int d = 0; d < 4; ++d) {
            uint32_t o = 0;
#pragma unroll
            for (int q = 0; q < 4; ++q) o |= (PL ? (w[q] >> (8 * d)) & 0xFFu : nib_plane8(w[q], d)) << (8 * q);
            Mb[(v * 4 + d) * NG + g] = o;
        }
    }
}

// Words of an unreduced column (rows < m + pairs - 1), the largest word
// offset of a shift z^l (l < pairs), and the words of the rows >= m.
constexpr int NGX = (M + PAIRS - 1 + 31) / 32;
constexpr int WSX = (PAIRS - 1) / 32;
constexpr int NOV = (PAIRS - 1 + 31) / 32;

// Constant GF(16) multiple F * p of a bitsliced vector (4 planes).
template <uint32_t F>
__device__ __forceinline__ void bs_const_mul(const uint32_t (&p)[4], uint32_t (&o)[4]) {
    uint32_t q0 = p[0], q1 = p[1], q2 = p[2], q3 = p[3];
    o[0] = o[1] = o[2] = o[3] = 0;
#pragma unroll
    for (int b = 0; b < 4; ++b) {
        if ((F >> b) & 1) { o[0] ^= q0; o[1] ^= q1; o[2] ^= q2; o[3] ^= q3; }
        bs_mul_a(q0, q1, q2, q3);
    }
}

// Column c = (block b, x) of A (sign.cu:333-393), bitsliced over its m rows:
// sum over the pairs of block b of z^l M_j[x], the z^l being row shifts
// (per lane, since b is), accumulated over m + pairs - 1 rows and reduced
// once: row m + t folds into rows t .. t+3 as z^t (F0 + F1 z + F2 z^2 + F3 z^3).
template <class LDS>
__device__ __forceinline__ void build_columns(const LDS &S, uint32_t (&cols)[NCL][4][NG]) {
    const int lane = lane_here();
#pragma unroll
    for (int s = 0; s < NCL; ++s) {
        const int c = col_of(s, lane);
#pragma unroll
        for (int d = 0; d < 4; ++d)
#pragma unroll
            for (int g = 0; g < NG; ++g) cols[s][d][g] = 0;
        if (c < KO) {
            const int b = c / O, x = c % O;
            uint32_t acc[4][NGX];
#pragma unroll
            for (int d = 0; d < 4; ++d)
#pragma unroll
                for (int q = 0; q < NGX; ++q) acc[d][q] = 0;
#pragma unroll
            for (int j = 0; j < K; ++j) {
                const int ell = j >= b ? ell_of(b, j) : ell_of(j, b);
                const int ws = ell >> 5;
                const uint32_t ns = (uint32_t)ell & 31u;
                uint32_t e[WSX + 1];
#pragma unroll
                for (int t = 0; t <= WSX; ++t) e[t] = ws == t ? 0xFFFFFFFFu : 0u;
                const uint32_t *src = &S.Mb[(j * O + x) * 4 * NG];
#pragma unroll
                for (int d = 0; d < 4; ++d)
#pragma unroll
                    for (int i = 0; i < NG; ++i) {
                        const uint64_t sh = (uint64_t)src[d * NG + i] << ns;
                        const uint32_t lo = (uint32_t)sh, hi = (uint32_t)(sh >> 32);
#pragma unroll
                        for (int t = 0; t <= WSX; ++t) {
                            if (i + t < NGX) acc[d][i + t] = gf16::axm(acc[d][i + t], lo, e[t]);
                            if (i + t + 1 < NGX) acc[d][i + t + 1] = gf16::axm(acc[d][i + t + 1], hi, e[t]);
                        }
                    }
            }
            // rows >= m: ov row t = acc row m + t
            uint32_t ov[NOV][4];
#pragma unroll
            for (int k = 0; k < NOV; ++k)
#pragma unroll
                for (int d = 0; d < 4; ++d) {
                    const int q = M / 32 + k;
                    const uint32_t lo = q < NGX ? acc[d][q] : 0u, hi = q + 1 < NGX ? acc[d][q + 1] : 0u;
                    ov[k][d] = (M % 32) ? __builtin_amdgcn_alignbit(hi, lo, M % 32) : lo;
                }
            // fold: rows t + i += F_i * ov row t (t + i < m)
            auto fold = [&](auto Fc, int i) {
                constexpr uint32_t F = decltype(Fc)::value;
                if constexpr (F != 0) {
                    uint32_t prev[4] = {0, 0, 0, 0};
#pragma unroll
                    for (int k = 0; k <= NOV; ++k) {
                        uint32_t pr[4] = {0, 0, 0, 0};
                        if (k < NOV) bs_const_mul<F>(ov[k], pr);
#pragma unroll
                        for (int d = 0; d < 4; ++d) {
                            const uint32_t v = i ? __builtin_amdgcn_alignbit(pr[d], prev[d], 32 - i) : pr[d];
                            if (k < NGX) acc[d][k] ^= v;
                            prev[d] = pr[d];
                        }
                    }
                }
            };
            fold(std::integral_constant<uint32_t, F0>{}, 0);
            fold(std::integral_constant<uint32_t, F1>{}, 1);
            fold(std::integral_constant<uint32_t, F2>{}, 2);
            fold(std::integral_constant<uint32_t, F3>{}, 3);
#pragma unroll
            for (int d = 0; d < 4; ++d)
#pragma unroll
                for (int g = 0; g < NG; ++g)
                    cols[s][d][g] = (g == NG - 1 && (M % 32)) ? acc[d][g] & ((1u << (M % 32)) - 1u) : acc[d][g];
        } else if (c == KO) {
            uint32_t v[MW];
#pragma unroll
            for (int q = 0; q < MW; ++q) v[q] = S.y[q];
            to_planes(v, cols[s]);
        }
    }
}

// GF(16) product at compile time (x^4 + x + 1)
__host__ __device__ constexpr uint32_t gf_mul_c(uint32_t a, uint32_t b) {
    uint32_t r = 0;
    for (int i = 0; i < 4; ++i)
        if ((b >> i) & 1u) r ^= a << i;
    for (int i = 6; i >= 4; --i)
        if ((r >> i) & 1u) r ^= 0x13u << (i - 4);
    return r;
}
// a^b * p^-1 as a 64-bit table over p (nibble p; 0 -> 0), from the inverse
// table 0x834A5C2F67BDE910 (nibble a = a^-1)
__host__ __device__ constexpr uint64_t im_table(int b) {
    uint64_t t = 0;
    for (int p = 1; p < 16; ++p)
        t |= (uint64_t)gf_mul_c(1u << b, (uint32_t)(0x834A5C2F67BDE910ull >> (4 * p)) & 0xFu) << (4 * p);
    return t;
}
constexpr uint64_t IM_TABLE[4] = {im_table(0), im_table(1), im_table(2), im_table(3)};
static_assert(IM_TABLE[0] == 0x834A5C2F67BDE910ull, "inverse table");

// f(integral_constant<int, g>) for the runtime word index g < NG (uniform)
template <int G = 0, class F>
__device__ __forceinline__ void dispatch_word(int g, F &&f) {
    if constexpr (G < NG) {
        if (g == G) f(std::integral_constant<int, G>{});
        else dispatch_word<G + 1>(g, f);
    }
}

// o[d] = XOR over the set bits b of s of x[b][d], where x[b] holds the four
// bit planes of a^b v (uniform 64-bit words): the product s v for a uniform
// scalar s, as a switch over s so that each case is a fixed scalar xor
// sequence (branch-free masking would take four and-xor pairs per plane).
template <uint32_t SC>
__device__ __forceinline__ void scale_case(const uint64_t (&x)[4][4], uint64_t (&o)[4]) {
#pragma unroll
    for (int d = 0; d < 4; ++d) {
        uint64_t t = 0;
#pragma unroll
        for (int b = 0; b < 4; ++b)
            if ((SC >> b) & 1u) t ^= x[b][d];
        o[d] = t;
    }
}
__device__ __forceinline__ void scale_planes(uint32_t sc, const uint64_t (&x)[4][4], uint64_t (&o)[4]) {
    switch (sc) {
        case 1: scale_case<1>(x, o); break;
        case 2: scale_case<2>(x, o); break;
        case 3: scale_case<3>(x, o); break;
        case 4: scale_case<4>(x, o); break;
        case 5: scale_case<5>(x, o); break;
        case 6: scale_case<6>(x, o); break;
        case 7: scale_case<7>(x, o); break;
        case 8: scale_case<8>(x, o); break;
        case 9: scale_case<9>(x, o); break;
        case 10: scale_case<10>(x, o); break;
        case 11: scale_case<11>(x, o); break;
        case 12: scale_case<12>(x, o); break;
        case 13: scale_case<13>(x, o); break;
        case 14: scale_case<14>(x, o); break;
        case 15: scale_case<15>(x, o); break;
        default: scale_case<0>(x, o); break;
    }
}

// A uniform value as a VGPR operand (the compiler would otherwise read it
// from its SGPR).
__device__ __forceinline__ uint32_t vgpr_copy(uint32_t x) {
    uint32_t r;
    asm("v_mov_b32 %0, %1" : "=v"(r) : "s"(x));
    return r;
}
// Gauss-Jordan on the columns of [A | y - A r]; x (S.x) = r + z. Returns 1
// on full rank. Each pivot step: the pivot column f is broadcast (SGPRs),
// its pivot row pr is the lowest unused row with a nonzero entry, p = f[pr];
// every column c becomes c + (c[pr] / p) f', with f' = f except f'[pr] = 0,
// which clears column pc outside row pr and keeps row pr. The per-lane
// scalar q = c[pr] / p takes 8 ops; the multiples of f' are uniform.
// Pivot columns are the greedy ones in column order, as in the reference's
// echelon form (sign.cu:10-145), and z on them is y'[pr] / p: the solution
// with the free variables at zero is unique, so x equals the reference's.
template <class Sync = BlockSync, bool VS = false, class LDS>
__device__ int solve(LDS &S, uint32_t (&cols)[NCL][4][NG], bool force_fail = false) {
    const int lane = lane_here();
    const uint8_t *rv = S.vr + R_OFF;
    // y' = y - A r: every A column times its r nibble, XOR-reduced over the
    // lanes by DPP (an LDS atomic at one address per word would become a
    // loop over the active lanes)
    {
        uint32_t part[4][NG];
#pragma unroll
        for (int d = 0; d < 4; ++d)
#pragma unroll
            for (int g = 0; g < NG; ++g) part[d][g] = 0;
#pragma unroll
        for (int s = 0; s < NCL; ++s) {
            const int c = col_of(s, lane);
            const uint32_t rc = c < KO ? mvec::nib(rv, c) : 0u;
            uint32_t rm[4];
#pragma unroll
            for (int b = 0; b < 4; ++b) rm[b] = gf16::bitmask(rc, b);
#pragma unroll
            for (int g = 0; g < NG; ++g) {
                uint32_t p0 = cols[s][0][g], p1 = cols[s][1][g], p2 = cols[s][2][g], p3 = cols[s][3][g];
#pragma unroll
                for (int b = 0; b < 4; ++b) {
                    part[0][g] = gf16::axm(part[0][g], p0, rm[b]);
                    part[1][g] = gf16::axm(part[1][g], p1, rm[b]);
                    part[2][g] = gf16::axm(part[2][g], p2, rm[b]);
                    part[3][g] = gf16::axm(part[3][g], p3, rm[b]);
                    bs_mul_a(p0, p1, p2, p3);
                }
            }
        }
#pragma unroll
        for (int d = 0; d < 4; ++d)
#pragma unroll
            for (int g = 0; g < NG; ++g) {
                const uint32_t ar = wave_xor(part[d][g]);
                if (lane == Y_LANE) cols[Y_SET][d][g] ^= ar;
            }
    }
    constexpr int NH = (NG + 1) / 2;
    uint64_t used[NH];
#pragma unroll
    for (int h = 0; h < NH; ++h) used[h] = 0;
    int rank = 0;
    // pivots in column order: set NCL-1 (columns < SHIFT) first, then 0, 1, ...
    static_for<NCL>([&](auto Oc) {
        constexpr int ord = decltype(Oc)::value;
        constexpr int ps = ord == 0 ? NCL - 1 : ord - 1;  // set holding these pivot columns
        constexpr int nl = ps == NCL - 1 ? SHIFT : 64;
#pragma unroll 1
        for (int pl = 0; pl < nl; ++pl) {
            const int pc = col_of(ps, pl);
            if (pc >= KO) break;
            // the pivot column f to SGPRs, then its pivot row: the lowest
            // unused row with a nonzero entry (a scalar search, which
            // measured faster here than every lane searching its own column
            // on the vector unit)
            // (64-bit pairs of words: the scalar unit has 64-bit logic ops,
            // and uniform 64-bit values stay on it)
            uint64_t fw[4][NH];
#pragma unroll
            for (int d = 0; d < 4; ++d)
#pragma unroll
                for (int h = 0; h < NH; ++h) {
                    const uint32_t lo = __builtin_amdgcn_readlane(cols[ps][d][2 * h], pl);
                    const uint32_t hi = 2 * h + 1 < NG ? __builtin_amdgcn_readlane(cols[ps][d][2 * h + 1], pl) : 0u;
                    fw[d][h] = ((uint64_t)hi << 32) | lo;
                }
            int pr = -1;
            if constexpr (VS) {
                // every lane searches its own column of the set; the pivot
                // lane's answer is read back
#pragma unroll
                for (int g = NG - 1; g >= 0; --g) {
                    const uint32_t ug = (g & 1) ? (uint32_t)(used[g >> 1] >> 32) : (uint32_t)used[g >> 1];
                    const uint32_t nz = (cols[ps][0][g] | cols[ps][1][g] | cols[ps][2][g] | cols[ps][3][g]) & ~ug;
                    pr = nz ? 32 * g + (int)__builtin_ctz(nz) : pr;
                }
                pr = __builtin_amdgcn_readlane(pr, pl);
            } else {
#pragma unroll
                for (int h = NH - 1; h >= 0; --h) {
                    const uint64_t nz = (fw[0][h] | fw[1][h] | fw[2][h] | fw[3][h]) & ~used[h];
                    pr = nz ? 64 * h + (int)__builtin_ctzll(nz) : pr;
                }
                pr = __builtin_amdgcn_readfirstlane(pr);
            }
            if (pr >= 0) {
                const int pg = pr >> 5;
                const uint32_t pb = (uint32_t)pr & 31u;
                // everything uniform is computed once per pivot: the pivot
                // row's bit, f' = f without row pr and its multiples a^b f'
                uint64_t wm[NH];
#pragma unroll
                for (int h = 0; h < NH; ++h) wm[h] = h == (pr >> 6) ? 1ull << (pr & 63) : 0ull;
#pragma unroll
                for (int d = 0; d < 4; ++d)
#pragma unroll
                    for (int h = 0; h < NH; ++h) fw[d][h] &= ~wm[h];
                // bit masks of c[pr] for every column of the sets this step
                // reads (the pivot set and the sets it updates): one
                // sign-extending one-bit extract per plane from the pivot
                // row's word, chosen by a uniform branch on that word's
                // index (selects on it would cost two more ops per plane)
                uint32_t cmall[NCL][4];
                auto masks_at = [&](auto Gc) {
                    constexpr int g = decltype(Gc)::value;
                    static_for<NCL>([&](auto Sc) {
                        constexpr int s = decltype(Sc)::value;
                        if constexpr (s != ps && s != Y_SET && set_rank(s) < ord) return;
#pragma unroll
                        for (int d = 0; d < 4; ++d)
                            cmall[s][d] = (uint32_t)__builtin_amdgcn_sbfe((int)cols[s][d][g], pb, 1u);
                    });
                };
                static_for<NG>([&](auto Gc) {
                    if (pg == decltype(Gc)::value) masks_at(Gc);
                });
                // the pivot value p = f[pr], read from the pivot lane's masks
                uint32_t cmp[4];
#pragma unroll
                for (int d = 0; d < 4; ++d) cmp[d] = cmall[ps][d];
                const uint32_t pv = (cmp[0] & 1u) | (cmp[1] & 2u) | (cmp[2] & 4u) | (cmp[3] & 8u);
                const uint32_t p = __builtin_amdgcn_readlane(pv, pl);
                const uint32_t pinv = (uint32_t)(IM_TABLE[0] >> (4 * p)) & 0xFu;
                // G[d] = a^d f' / p on the scalar unit (G[0] = p^-1 f' by a
                // uniform switch over p^-1, then multiples by a), so that
                // c += (c[pr] / p) f' is sum_d bit_d(c[pr]) G[d]: the masks of
                // c[pr] are the update masks, with no per-lane division
                uint64_t A[4][4][NH];  // A[b] = a^b f'
#pragma unroll
                for (int h = 0; h < NH; ++h) {
#pragma unroll
                    for (int d = 0; d < 4; ++d) A[0][d][h] = fw[d][h];
#pragma unroll
                    for (int b = 1; b < 4; ++b) {
                        A[b][0][h] = A[b - 1][3][h];
                        A[b][1][h] = A[b - 1][0][h] ^ A[b - 1][3][h];
                        A[b][2][h] = A[b - 1][1][h];
                        A[b][3][h] = A[b - 1][2][h];
                    }
                }
                // G[0] = p^-1 f' = sum of A[b] over the set bits of p^-1: four
                // uniform branches, each a run of scalar xors
                uint64_t G[4][4][NH];
#pragma unroll
                for (int d = 0; d < 4; ++d)
#pragma unroll
                    for (int h = 0; h < NH; ++h) G[0][d][h] = 0;
#pragma unroll
                for (int b = 0; b < 4; ++b) {
                    if (__builtin_amdgcn_readfirstlane((pinv >> b) & 1u)) {
#pragma unroll
                        for (int d = 0; d < 4; ++d)
#pragma unroll
                            for (int h = 0; h < NH; ++h) G[0][d][h] ^= A[b][d][h];
                    }
                }
#pragma unroll
                for (int h = 0; h < NH; ++h)
#pragma unroll
                    for (int b = 1; b < 4; ++b) {
                        G[b][0][h] = G[b - 1][3][h];
                        G[b][1][h] = G[b - 1][0][h] ^ G[b - 1][3][h];
                        G[b][2][h] = G[b - 1][1][h];
                        G[b][3][h] = G[b - 1][2][h];
                    }
#pragma unroll
                for (int g = 0; g < NG; ++g) {
                    // word g of the multiples as VGPR operands (7 distinct of
                    // the 16): AND-XORs that read no SGPR can issue paired
                    // with another wave's VALU op
                    uint32_t vg[4][4];
#pragma unroll
                    for (int d = 0; d < 4; ++d)
#pragma unroll
                        for (int e = 0; e < 4; ++e) {
                            const uint64_t gw = G[d][e][g >> 1];
                            vg[d][e] = vgpr_copy((g & 1) ? (uint32_t)(gw >> 32) : (uint32_t)gw);
                        }
                    static_for<NCL>([&](auto Sc) {
                        constexpr int s = decltype(Sc)::value;
                        // sets whose columns all precede pc need no update
                        // (pivot columns are p e_pr, which the later steps
                        // leave alone; free columns do not enter x), unless
                        // they hold y'
                        if constexpr (s != Y_SET && set_rank(s) < ord) return;
#pragma unroll
                        for (int e = 0; e < 4; ++e) {
                            uint32_t c = cols[s][e][g];
#pragma unroll
                            for (int d = 0; d < 4; ++d) c = gf16::axm(c, vg[d][e], cmall[s][d]);
                            cols[s][e][g] = c;
                        }
                    });
                }
                if (lane == 0) S.pinv[pc] = (uint8_t)pinv;
#pragma unroll
                for (int h = 0; h < NH; ++h) used[h] |= wm[h];
                ++rank;
            }
            if (lane == 0) S.prow[pc] = (int16_t)pr;
        }
    });
    // x = r, plus y'[pivot row] / p on the pivot columns
    if (lane == Y_LANE) {
#pragma unroll
        for (int d = 0; d < 4; ++d)
#pragma unroll
            for (int g = 0; g < NG; ++g) S.yplanes[d * NG + g] = cols[Y_SET][d][g];
    }
    Sync::sync();
    // force_fail (test hook): report a singular system; x stays r, as when
    // the rank check fails
    const int ok = rank == M && !force_fail;
    for (int c = lane; c < KO; c += 64) {
        uint32_t xv = mvec::nib(rv, c);
        const int pr = S.prow[c];
        if (ok && pr >= 0) {
            uint32_t yn = 0;
#pragma unroll
            for (int d = 0; d < 4; ++d) yn |= ((S.yplanes[d * NG + (pr >> 5)] >> (pr & 31)) & 1u) << d;
            xv ^= tmul(S, yn, S.pinv[c]);
        }
        S.x[c] = (uint8_t)xv;
    }
    Sync::sync();
    return ok;
}

__device__ __forceinline__ void wave_scol(Shared &S) {
    for (int c = lane_here(); c < V; c += 64) {
        uint64_t v = 0;
        for (int j = 0; j < K; ++j) v |= (uint64_t)mvec::nib(S.vr + j * V_BYTES, c) << (4 * j);
        S.scol[c] = v;
    }
}

// s_i = (V_i + O x_i || x_i) and the salt (sign.cu:147-191, 472-477)
template <class LDS>
__device__ __forceinline__ void wave_write_sig(const LDS &S, const SignItem &it) {
    const int lane = lane_here();
    const uint8_t *Ob = it.esk + SK;
    for (int b = lane; b < SIG_S_BYTES; b += 64) {
        uint32_t out = 0;
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const int e = 2 * b + h, i = e / N, t = e % N;
            uint32_t val;
            if (t < V) {
                uint32_t a = 0;
#pragma unroll
                for (int kk = 0; kk < O; ++kk) a ^= tmul(S, mvec::nib(Ob, t * O + kk), S.x[i * O + kk]);
                val = mvec::nib(S.vr + i * V_BYTES, t) ^ a;
            } else {
                val = S.x[i * O + (t - V)];
            }
            out |= val << (4 * h);
        }
        it.sig[b] = (uint8_t)out;
    }
    for (int b = lane; b < SK; b += 64) it.sig[SIG_S_BYTES + b] = it.scratch[SCR_SALT + b];
}

// V || r of attempt c >= 1, re-derived by one wave (sign.cu:251-252):
// SHAKE256(digest || salt || seed || c) into S.vr.
template <class Sync>
__device__ __forceinline__ void derive_vr(Shared &S, const SignItem &it, int c) {
    const int lane = lane_here() & 63;
    for (int b = lane; b < DIGEST + SK; b += 64) S.hbuf[b] = it.scratch[SCR_DIGEST + b];
    for (int b = lane; b < SK; b += 64) S.hbuf[DIGEST + SK + b] = it.esk[b];
    if (lane == 0) S.hbuf[DIGEST + 2 * SK] = (uint8_t)c;
    Sync::sync();
    keccak::wave_shake256(S.vr, VR_BYTES, S.hbuf, DIGEST + 2 * SK + 1);
}

// Waves per SIMD of the whole-loop kernel below. MAYO_3's spills 182 VGPRs
// at 3 waves and 55 at 2, and signs 6% faster at 2 (0.90 -> 0.95 M/s over
// 32,768 pitched distinct keys). MAYO_1's spills 105 at 3 and none at 2, yet
// signs 9% slower at 2, so it keeps 3.
constexpr int SIGN_WAVE_OCC = M > 80 && M <= 108 ? 2 : QF_OCC;

// The whole signing loop in one wave (sign.cu:244-478), for the reference
// pool. Attempts below fail_first report a singular system (test hook).
// LF: the key row holds P2 where the expanded key has L, and M is formed
// without L (the L-free quadratic form above).
template <class Acc, class KL = RefKeys, bool LF = false>
__global__ void __launch_bounds__(64, SIGN_WAVE_OCC) k_sign_wave(Acc acc, size_t n, int fail_first = 0) {
    __shared__ Shared S;
    __shared__ uint32_t ocol[LF ? V * OWD : 1];
    __shared__ uint32_t Wc[LF ? QF_RT * SL * K * MW : 1];
    const int lane = threadIdx.x;
    SignItem it = acc(blockIdx.x);
    if constexpr (LF) load_ocol(it.esk + SK, ocol);
    for (int w = lane; w < MW; w += 64) S.t[w] = mvec::load_mw(it.scratch + SCR_T, w);
    for (int e = lane; e < 256; e += 64) S.mul[e] = (uint8_t)gf16::mul(e >> 4, e & 15);
    for (int b = lane; b < VR_BYTES; b += 64) S.vr[b] = it.scratch[SCR_VR + b];
    int ctr = 0, ok = 0;
    for (;; ++ctr) {
        __syncthreads();
        wave_scol(S);
        for (int e = lane; e < PAIRS * MW; e += 64) S.U[e] = 0;
        __syncthreads();
        quad_form<KL, O, false, 1, QF_RT, LF>(it.esk, S.scol, S.U, S.Mv, ocol, Wc);
        __syncthreads();
        wave_y(S, S.t);
        planes_of_M<KL::PLANAR>(S.Mv, S.Mb);  // Mb overlays U, which wave_y has consumed
        __syncthreads();
        uint32_t cols[NCL][4][NG];
        build_columns(S, cols);
        ok = solve(S, cols, ctr < fail_first);
        if (ok || ctr == MAX_ATTEMPTS - 1) break;
        derive_vr<BlockSync>(S, it, ctr + 1);
    }
    if (lane == 0) it.scratch[SCR_ATTEMPTS] = ok ? (uint8_t)(ctr + 1) : 0;
    __syncthreads();
    wave_write_sig(S, it);
}

// The whole signing stage in one persistent launch: each wave takes items
// k = blockIdx.x, + gridDim.x, ... (positions of the batch, or of the
// XCD-aware order) and runs attempt 0 and any retries of each item itself:
// quad form, M and y stay in LDS (no workspace hand-off), then the columns
// of A, the elimination and the signature (sign.cu:244-478). One wave per
// workgroup; the grid holds as many waves as the GPU keeps resident.
template <bool PL>
__device__ __noinline__ void retry_linear_algebra(Shared &S, int ctr, bool force_fail, int *state);

template <class Acc, class KL, int CODE>
__global__ void __launch_bounds__(64, CODE % 10) k_sign_fused(Acc acc, size_t n, int fail_first = 0,
                                                              int ordered = 0) {
    __shared__ Shared S;
    __shared__ int state;  // 0 next attempt, 1 solved, 2 attempts exhausted
    const int lane = threadIdx.x;
    for (int e = lane; e < 256; e += 64) S.mul[e] = (uint8_t)gf16::mul(e >> 4, e & 15);
    for (size_t k = blockIdx.x; k < n; k += gridDim.x) {
        SignItem it = acc(ordered ? acc.order(k) : k);
        __syncthreads();
        for (int w = lane; w < MW; w += 64) S.t[w] = mvec::load_mw(it.scratch + SCR_T, w);
        for (int b = lane; b < VR_BYTES; b += 64) S.vr[b] = it.scratch[SCR_VR + b];
        int ctr = 0;
        for (;; ++ctr) {
            __syncthreads();
            wave_scol(S);
            for (int e = lane; e < PAIRS * MW; e += 64) S.U[e] = 0;
            __syncthreads();
            quad_form<KL, O, false, 1, CODE / 10>(it.esk, S.scol, S.U, S.Mv);
            __syncthreads();
            // y, the columns of A and the elimination: a call of its own,
            // which keeps the quadratic form's registers out of it
            retry_linear_algebra<KL::PLANAR>(S, ctr, ctr < fail_first, &state);
            __syncthreads();
            if (state) break;
            derive_vr<BlockSync>(S, it, ctr + 1);
        }
        if (lane == 0) it.scratch[SCR_ATTEMPTS] = state == 1 ? (uint8_t)(ctr + 1) : 0;
        __syncthreads();
        wave_write_sig(S, it);
    }
}

// wave 0 of the retry kernel: y, the columns of A, the elimination (a call
// of its own, which keeps the register use of the kernel down)
template <bool PL>
__device__ __noinline__ void retry_linear_algebra(Shared &S, int ctr, bool force_fail, int *state) {
    wave_y<WaveSync>(S, S.t);
    planes_of_M<PL>(S.Mv, S.Mb);
    WaveSync::sync();
    uint32_t cols[NCL][4][NG];
    build_columns(S, cols);
    const int ok = solve<WaveSync>(S, cols, force_fail);
    if (threadIdx.x == 0) *state = ok ? 1 : (ctr == MAX_ATTEMPTS - 1 ? 2 : 0);
}

// Retry pass of the batch form: the items whose attempt 0 failed (about 1
// in 5,000), listed by the solve kernel, from attempt 1 on. One workgroup of
// RNW waves per item, so that one item's attempt takes a fraction of a
// single wave's time: the waves share the row chunks of the quadratic form,
// wave 0 re-derives V || r and runs the linear algebra. Grid: any size; a
// workgroup takes list entries blockIdx.x, + gridDim.x, ... of the failure
// list kept at item `list` (the first item of the launch part that made it).
// Eight waves, two per SIMD: the kernel needs at most 243 VGPRs (MAYO_5),
// within the 256 that two waves per SIMD leave each.
constexpr int RNW_MAX = 8;
constexpr int RNW = qf_chunks<O, false>() < RNW_MAX ? qf_chunks<O, false>() : RNW_MAX;
template <class Acc, class KL>
__global__ void __launch_bounds__(64 * RNW, 1) k_sign_retry(Acc acc, size_t n, int fail_first = 0,
                                                              size_t list = 0) {
    __shared__ Shared S;
    __shared__ int state;  // 0 next attempt, 1 solved, 2 attempts exhausted
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const uint32_t count = *(const uint32_t *)(acc.scratch_of(list) + SCR_LIST_COUNT);
    for (uint32_t k = blockIdx.x; k < count; k += gridDim.x) {
        const uint32_t item = *(const uint32_t *)(acc.scratch_of(list + k) + SCR_LIST);
        SignItem it = acc(item);
        if (wave == 0) {
            for (int w = lane; w < MW; w += 64) S.t[w] = mvec::load_mw(it.scratch + SCR_T, w);
            for (int e = lane; e < 256; e += 64) S.mul[e] = (uint8_t)gf16::mul(e >> 4, e & 15);
        }
        int ctr = 1;
        for (;; ++ctr) {
            if (wave == 0) {
                derive_vr<WaveSync>(S, it, ctr);
                WaveSync::sync();
                wave_scol(S);
            }
            for (int e = threadIdx.x; e < PAIRS * MW; e += 64 * RNW) S.U[e] = 0;
            __syncthreads();
            quad_form<KL, O, false, RNW>(it.esk, S.scol, S.U, S.Mv);
            __syncthreads();
            if (wave == 0) retry_linear_algebra<KL::PLANAR>(S, ctr, ctr < fail_first, &state);
            __syncthreads();
            if (state) break;
        }
        if (wave == 0) {
            if (lane == 0) it.scratch[SCR_ATTEMPTS] = state == 1 ? (uint8_t)(ctr + 1) : 0;
            wave_write_sig(S, it);
        }
        __syncthreads();
    }
}

// Batch form, attempt 0, part 1: M, U and y, handed over through the
// workspace. Register-heavy (the pair accumulators of the quadratic form).
// One item per workgroup, positions base + blockIdx.x (`end`, the part's
// end, is implied by the grid; the shared-key variant takes the same
// arguments).
template <class Acc, class KL = RefKeys, int CODE = qf_code(QF_RT, QF_OCC)>
__global__ void __launch_bounds__(64, CODE % 10) k_sign_prep(Acc acc, size_t n, int ordered = 0, size_t base = 0,
                                                             size_t end = ~(size_t)0) {
    __shared__ Shared S;
    const int lane = threadIdx.x;
    SignItem it = acc(item_of(acc, ordered, base));
    // the failure list of this launch part's solve kernel starts empty; it
    // lives in the part's own workspace slots (count in slot base, entries
    // in slots base, base + 1, ...), so parts never share one
    if (blockIdx.x == 0 && lane == 0) *(uint32_t *)(acc.scratch_of(base) + SCR_LIST_COUNT) = 0;
    for (int b = lane; b < VR_BYTES; b += 64) S.vr[b] = it.scratch[SCR_VR + b];
    for (int w = lane; w < MW; w += 64) S.t[w] = mvec::load_mw(it.scratch + SCR_T, w);
    __syncthreads();
    wave_scol(S);
    for (int e = lane; e < PAIRS * MW; e += 64) S.U[e] = 0;
    __syncthreads();
    quad_form<KL, O, false, 1, CODE / 10>(it.esk, S.scol, S.U, S.Mv);
    __syncthreads();
    wave_y(S, S.t);
    uint32_t *mv = (uint32_t *)(it.scratch + SCR_MV);
    for (int e = lane; e < K * O * MW; e += 64) mv[e] = S.Mv[e];
    uint32_t *yv = (uint32_t *)(it.scratch + SCR_Y);
    for (int e = lane; e < MW; e += 64) yv[e] = S.y[e];
}

// Batch form, attempt 0, part 2: columns of [A | y - A r], Gauss-Jordan,
// signature. Few registers, so many waves per SIMD hide the pivot chain.
// PL: the M vectors in the workspace are planar (prepared keys). CODE: a
// solve variant (SOLVE_OCCS): waves per SIMD, pivot search schedule, and
// signatures per workgroup (one wave each, positions base + blockIdx.x W +
// wave below end).
template <class Acc, bool PL, int CODE = SOLVE_OCC>
__global__ void __launch_bounds__(64 * solve_sigs(CODE), solve_occ(CODE))
    k_sign_solve(Acc acc, size_t n, int fail_first = 0, int ordered = 0, size_t base = 0, size_t end = ~(size_t)0) {
    constexpr int W = solve_sigs(CODE);
    using Sync = std::conditional_t<(W > 1), WaveSync, BlockSync>;
    __shared__ SolveShared Sw[W];
    const int wave = W > 1 ? (int)(threadIdx.x >> 6) : 0;
    SolveShared &S = Sw[wave];
    const int lane = threadIdx.x & 63;
    const size_t k = base + (size_t)blockIdx.x * W + wave;
    if (W > 1 && k >= end) return;  // whole waves only; no workgroup barrier follows
    const size_t item = ordered ? acc.order(k) : k;
    SignItem it = acc(item);
    for (int b = lane; b < VR_BYTES; b += 64) S.vr[b] = it.scratch[SCR_VR + b];
    planes_of_M<PL>((const uint32_t *)(it.scratch + SCR_MV), S.Mb);
    const uint32_t *yv = (const uint32_t *)(it.scratch + SCR_Y);
    for (int e = lane; e < MW; e += 64) S.y[e] = yv[e];
    for (int e = lane; e < 256; e += 64) S.mul[e] = (uint8_t)gf16::mul(e >> 4, e & 15);
    Sync::sync();
    uint32_t cols[NCL][4][NG];
    build_columns(S, cols);
    const int ok = solve<Sync, solve_schedule(CODE) == 1>(S, cols, fail_first > 0);
    if (!ok) {
        // pending: listed for the retry pass
        if (lane == 0) {
            it.scratch[SCR_ATTEMPTS] = ATTEMPT_PENDING;
            const uint32_t kk = atomicAdd((uint32_t *)(acc.scratch_of(base) + SCR_LIST_COUNT), 1u);
            *(uint32_t *)(acc.scratch_of(base + kk) + SCR_LIST) = (uint32_t)item;
        }
        return;
    }
    if (lane == 0) it.scratch[SCR_ATTEMPTS] = 1;
    wave_write_sig(S, it);
}

// ---- key-side products of P1 and O (expandsk.cu:37-80, keypair.cu:55-119) ----
// One wave per key. The entries of O are the scalars, uniform across the
// wave, so their bits are SGPR masks; lanes are (word w, row slot) and each
// lane accumulates the four bit planes of T_x[r] = sum_c P1'[r][c] O[c][x]
// for its rows, as in the quadratic form. SYM: P1' = P1 + P1^T, whose
// diagonal cancels (entry (r, c) is P1[r][c] above the diagonal, P1[c][r]
// below). Otherwise P1' is the upper triangle with its diagonal. For the
// rows of a chunk, columns left of the chunk read lower entries and columns
// right of it upper entries at lane-constant offsets; only the chunk's own
// columns need a per-lane choice. Entry words are fetched one column ahead.
constexpr int KRT = O <= 12 ? 2 : 1;        // rows per lane per chunk

// row_done(r, T) is called for each valid row of the lane with T[x] = T_x[r]
// (this lane's word w). Entry words are loaded PF columns ahead.
template <bool SYM, int PF, class RowDone>
__device__ __forceinline__ void p1_times_O(const uint8_t *P1, const uint32_t *ocol, RowDone &&row_done) {
    const int lane = threadIdx.x;
    const int w = lane % MW, slot = lane / MW;
    const bool active = slot < SL;
    const int nv = mvec::word_valid(w);
    constexpr int RPC = KRT * SL;
    constexpr int NCH = (V + RPC - 1) / RPC;
#pragma unroll 1
    for (int ch = 0; ch < NCH; ++ch) {
        const int r0 = ch * RPC, r1 = (r0 + RPC < V ? r0 + RPC : V) - 1;
        int rr[KRT];
        bool val[KRT];
        uint32_t up[KRT], lo[KRT];
        uint32_t Tb[KRT][O][4];
#pragma unroll
        for (int tt = 0; tt < KRT; ++tt) {
            rr[tt] = r0 + tt * SL + slot;
            val[tt] = active && rr[tt] < V;
            const int r = val[tt] ? rr[tt] : 0;
            up[tt] = (uint32_t)((tri_entry(r, r, V) - r) * MB + 4 * w);  // + c MB: entry (r, c)
            lo[tt] = (uint32_t)(r * MB + 4 * w);                         // + (tri(c,c) - c) MB: entry (c, r)
#pragma unroll
            for (int x = 0; x < O; ++x)
#pragma unroll
                for (int b = 0; b < 4; ++b) Tb[tt][x][b] = 0;
        }
        // address of column c for row slot tt (phase: 0 below, 1 mixed, 2 above)
        auto offset = [&](int c, int tt) -> uint32_t {
            const uint32_t lb = (uint32_t)((tri_entry(c, c, V) - c) * MB), ub = (uint32_t)(c * MB);
            if (c < r0) return lo[tt] + lb;
            if (c > r1) return up[tt] + ub;
            return (SYM && c < rr[tt]) ? lo[tt] + lb : up[tt] + ub;
        };
        // the loads of the next PF columns are in flight during each column
        wv::EntryLoad<RefKeys> ld[PF][KRT];
        const int cstart = SYM ? 0 : r0;
#pragma unroll
        for (int k = 0; k < PF; ++k)
            if (cstart + k < V) {
#pragma unroll
                for (int tt = 0; tt < KRT; ++tt) ld[k][tt].fetch(P1, offset(cstart + k, tt), nv);
            }
#pragma unroll 1
        for (int c0 = cstart; c0 < V; c0 += PF) {
#pragma unroll
            for (int k = 0; k < PF; ++k) {
                const int c = c0 + k;
                if (c >= V) break;
                uint32_t x[KRT];
#pragma unroll
                for (int tt = 0; tt < KRT; ++tt) {
                    x[tt] = ld[k][tt].word(nv);
                    // the chunk's own columns: the diagonal cancels (SYM); the
                    // upper product stops below the diagonal
                    if (c >= r0 && c <= r1) x[tt] = (SYM ? c == rr[tt] : c < rr[tt]) ? 0u : x[tt];
                }
                if (c + PF < V) {
#pragma unroll
                    for (int tt = 0; tt < KRT; ++tt) ld[k][tt].fetch(P1, offset(c + PF, tt), nv);
                }
                uint32_t oc[OWD];
#pragma unroll
                for (int q = 0; q < OWD; ++q) oc[q] = uni(ocol[c * OWD + q]);
#pragma unroll
                for (int xo = 0; xo < O; ++xo)
#pragma unroll
                    for (int b = 0; b < 4; ++b) {
                        const uint32_t m = (uint32_t)(-(int32_t)((oc[xo >> 3] >> (4 * (xo & 7) + b)) & 1u));
#pragma unroll
                        for (int tt = 0; tt < KRT; ++tt) Tb[tt][xo][b] = gf16::axm(Tb[tt][xo][b], x[tt], m);
                    }
            }
        }
#pragma unroll
        for (int tt = 0; tt < KRT; ++tt) {
            if (!val[tt]) continue;
            uint32_t T[O];
#pragma unroll
            for (int xo = 0; xo < O; ++xo)
                T[xo] = gf16::mul_a(gf16::mul_a(gf16::mul_a(Tb[tt][xo][3]) ^ Tb[tt][xo][2]) ^ Tb[tt][xo][1]) ^
                        Tb[tt][xo][0];
            row_done(rr[tt], T);
        }
    }
}

// Columns k_expand_L_wave loads ahead: 3 for MAYO_1 and 2 for MAYO_3 (65,536
// expansions: 16.3 -> 15.3 ms and 43.7 -> 41.4 ms); 1 for MAYO_2, whose
// kernel drops from 5 to 4 waves per SIMD with more (11.0 -> 13.4 ms), and
// for MAYO_5, which spills with more. The keypair kernel loads one ahead:
// more would cost it occupancy.
constexpr int EXPAND_PF = KRT == 1 || M > 108 ? 1 : (M <= 80 ? 3 : 2);

// expandsk: L = (P1 + P1^T) O + P2 in place over P2.
__global__ void __launch_bounds__(64, 3) k_expand_L_wave(uint8_t *pool, size_t pitch, size_t n) {
    __shared__ uint32_t ocol[V * OWD];
    uint8_t *esk = pool + (size_t)blockIdx.x * pitch;
    const uint8_t *P1 = esk + SK + O_BYTES;
    uint8_t *L = (uint8_t *)P1 + P1_BYTES;
    load_ocol(esk + SK, ocol);
    __syncthreads();
    const int w = threadIdx.x % MW;
    p1_times_O<true, EXPAND_PF>(P1, ocol, [&](int r, const uint32_t (&T)[O]) {
#pragma unroll
        for (int xo = 0; xo < O; ++xo) {
            uint8_t *e = L + (size_t)(r * O + xo) * MB;
            mvec::store_mw(e, w, T[xo] ^ mvec::load_mw(e, w));
        }
    });
}

// keypair: P3 = Upper(O^T (P1 O + P2)). O is read from the P3 area where the
// seed kernel parked it, then P3 overwrites it. Each row r of B = P1 O + P2
// adds O[r][x] B[r][y] + O[r][y] B[r][x] into entry {x, y} (x < y) and
// O[r][x] B[r][x] into {x, x}, summed over the wave in LDS.
__global__ void __launch_bounds__(64, 2) k_keypair_P3_wave(uint8_t *pool, size_t pitch, size_t n,
                                                           size_t pk_off, size_t p_off) {
    __shared__ uint32_t ocol[V * OWD];
    __shared__ uint32_t P3acc[P3_ENTRIES * MW];
    uint8_t *row = pool + (size_t)blockIdx.x * pitch;
    uint8_t *P3 = row + pk_off + 16;
    const uint8_t *P1 = row + p_off;
    const uint8_t *P2 = P1 + P1_BYTES;
    load_ocol(P3, ocol);
    for (int e = threadIdx.x; e < P3_ENTRIES * MW; e += 64) P3acc[e] = 0;
    __syncthreads();
    const int w = threadIdx.x % MW;
    p1_times_O<false, 1>(P1, ocol, [&](int r, const uint32_t (&T)[O]) {
        uint32_t om[O][4];  // bit masks of O[r][x]
#pragma unroll
        for (int x = 0; x < O; ++x) {
            const uint32_t s = (ocol[r * OWD + (x >> 3)] >> (4 * (x & 7))) & 0xFu;
#pragma unroll
            for (int b = 0; b < 4; ++b) om[x][b] = gf16::bitmask(s, b);
        }
#pragma unroll
        for (int y = 0; y < O; ++y) {
            const gf16::Multiples mb(T[y] ^ mvec::load_mw(P2 + (size_t)(r * O + y) * MB, w));  // B[r][y]
#pragma unroll
            for (int x = 0; x < O; ++x) {
                // O[r][x] B[r][y] lands in {min(x,y), max(x,y)}
                const uint32_t v = gf16::mac(0u, mb.x0, mb.x1, mb.x2, mb.x3, om[x][0], om[x][1], om[x][2], om[x][3]);
                const int e = x <= y ? tri_entry(x, y, O) : tri_entry(y, x, O);
                atomicXor(&P3acc[e * MW + w], v);
            }
        }
    });
    __syncthreads();
    for (int item = threadIdx.x; item < P3_ENTRIES * MW; item += 64) {
        const int e = item / MW, ww = item % MW;
        mvec::store_mw(P3 + (size_t)e * MB, ww, P3acc[item]);
    }
}

// Verification, one wave per item (verify.cu:9-211).
template <class Acc, class KL = RefKeys, int CODE = qf_code(QF_RT, QF_OCC)>
__global__ void __launch_bounds__(64, CODE % 10) k_verify_wave(Acc acc, size_t n, int ordered = 0) {
    __shared__ uint64_t scol[N];
    __shared__ uint32_t U[PAIRS * MW];
    __shared__ uint32_t yu[YG_MAX * UW];
    __shared__ uint8_t tb[MB + 4];
    __shared__ uint8_t hbuf[DIGEST + SK];
    const int lane = threadIdx.x;
    VerifyItem it = acc(item_of(acc, ordered));
    for (int c = lane; c < N; c += 64) {
        uint64_t v = 0;
        for (int j = 0; j < K; ++j) v |= (uint64_t)mvec::nib(it.sig, j * N + c) << (4 * j);
        scol[c] = v;
    }
    for (int e = lane; e < PAIRS * MW; e += 64) U[e] = 0;
    if (it.t) {
        for (int b = lane; b < MB; b += 64) tb[b] = it.t[b];
    } else {
        // digest = SHAKE256(msg), t = SHAKE256(digest || salt) (verify.cu:46-48)
        keccak::wave_shake256(hbuf, DIGEST, it.msg, it.mlen);
        for (int b = lane; b < SK; b += 64) hbuf[DIGEST + b] = it.sig[SIG_S_BYTES + b];
        __syncthreads();
        keccak::wave_shake256(tb, MB, hbuf, DIGEST + SK);
    }
    __syncthreads();
    quad_form<KL, 0, true, 1, CODE / 10>(it.epk, scol, U, nullptr);
    __syncthreads();
    unreduced_y<BlockSync>(U, yu);
    if (lane == 0) {
        uint32_t a[UW], y[MW];
#pragma unroll
        for (int q = 0; q < UW; ++q) a[q] = yu[q];
        reduce_unreduced(a, y);
        bool eq = true;
#pragma unroll
        for (int w = 0; w < MW; ++w) {
            const int nv = mvec::word_valid(w);
            uint32_t tw = 0;
            for (int b = 0; b < nv; ++b) tw |= (uint32_t)tb[4 * w + b] << (8 * b);
            eq &= (y[w] == tw);
        }
        it.result[0] = eq ? 0 : 1;
    }
}

}  // namespace wv
}  // namespace mayo
