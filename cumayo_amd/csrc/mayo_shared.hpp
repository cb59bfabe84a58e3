// Shared-key quadratic forms: a workgroup of W wavefronts computes the
// quadratic form of up to W items under one key, one wavefront per item, in
// two kernels: the table kernel (the column loop, below) and the fold.
//
// The quadratic form's column loop needs, for every entry P[r][c] of the key
// and every scalar S_j[c] of an item, the product S_j[c] P[r][c]. A product
// depends on the key entry and on a nibble only, so the workgroup builds, in
// LDS, the 16 multiples s P[r][c] (s = 0..15) of the entries of a few
// columns of the current row chunk once, and every wavefront of the group
// then takes each of its products as one LDS read at an address set by its
// scalar (uniform across the wave, so the read is conflict-free) and one
// XOR: 1.5 VALU instructions per entry word and scalar against 4 AND-XORs
// in the bitsliced form, with the table built once per W items. The loop is
// bound by LDS bandwidth (8-byte reads: about 120 bytes per clock and CU).
//
// Lanes of the column loop are (word pair g, row slot): a chunk holds one
// row per slot, and each lane accumulates T_j[r] = sum_c P[r][c] S_j[c] for
// its row and its two words, for every j, in 2K registers. Tables are
// double-buffered: while the wavefronts read the multiples of columns
// [c0, c0 + CG) the producing threads write those of the next CG columns,
// and one workgroup barrier per CG columns separates the two.
//
// At a chunk's end the lanes store their rows: for signing, the first NLR
// rows are the columns x of L, M_j[x] = T_j[x], stored to M; the rows of P
// go to the item's T region in the workspace. The fold kernel (one
// wavefront per item) then forms the pair sums, U_(a,j) += S_a[r] T_j[r],
// in a transposed layout: lanes own (j, word pair) items and keep U_(a,j)
// for all a in registers (2K per item), and rows are uniform across the
// wave. U_(a,j) + U_(j,a) is the pair sum of the bitsliced kernels
// (sign.cu:395-437, verify.cu:72-182).
//
// The items of a workgroup are W consecutive positions of the key order
// (launch_key_sort), so they share a key except at the boundaries between
// keys; a workgroup whose items hold several keys runs the table loop once
// per key, each wavefront reading in the pass for its own key.
#pragma once
#include "mayo_wave.hpp"

namespace mayo {
namespace wv {

// ---- key order ------------------------------------------------------------------
// Items sorted by key: a counting sort over SORT_BINS bins (key mod 4,096:
// exact for key ranges up to 4,096), sorted position k to acc.set_order(k),
// positions within a bin in any order. Three kernels over workgroups of
// KS_ITEMS items: per-workgroup histograms, one workgroup that turns them
// into each workgroup's first position per bin, and the scatter. `hist`
// holds SORT_BINS words per workgroup (key_sort_bytes(n), the end of the
// table-form workspaces, after the T regions).
constexpr int SORT_BINS = 4096;
constexpr int KS_ITEMS = 4096;  // items per workgroup, 4 per thread
constexpr int KS_U = KS_ITEMS / 1024;
inline unsigned key_sort_groups(size_t n) { return (unsigned)((n + KS_ITEMS - 1) / KS_ITEMS); }
inline size_t key_sort_bytes(size_t n) { return (size_t)key_sort_groups(n) * SORT_BINS * 4; }

template <class Acc>
__device__ __forceinline__ void ks_keys(const Acc &acc, size_t n, uint32_t (&kv)[KS_U]) {
#pragma unroll
    for (int u = 0; u < KS_U; ++u) {
        const size_t i = (size_t)blockIdx.x * KS_ITEMS + threadIdx.x + 1024 * u;
        kv[u] = i < n ? acc.key(i) % SORT_BINS : 0u;
    }
}

template <class Acc>
__global__ void __launch_bounds__(1024) k_key_hist(Acc acc, size_t n, uint32_t *hist) {
    __shared__ uint32_t h[SORT_BINS];
    const int t = threadIdx.x;
    for (int b = t; b < SORT_BINS; b += 1024) h[b] = 0;
    uint32_t kv[KS_U];
    ks_keys(acc, n, kv);
    __syncthreads();
#pragma unroll
    for (int u = 0; u < KS_U; ++u)
        if ((size_t)blockIdx.x * KS_ITEMS + t + 1024 * u < n) atomicAdd(&h[kv[u]], 1u);
    __syncthreads();
    for (int b = t; b < SORT_BINS; b += 1024) hist[(size_t)blockIdx.x * SORT_BINS + b] = h[b];
}

// hist[g][b] (counts) -> the first sorted position of workgroup g's items
// of bin b: bins in order, within a bin the workgroups in order
__global__ void __launch_bounds__(1024) k_key_scan(uint32_t *hist, unsigned groups) {
    constexpr int PER = SORT_BINS / 1024;  // thread t: bins PER t .. PER t + PER - 1
    constexpr int GB = 8;                  // workgroups' counts loaded together
    static_assert(PER == 4, "one 16-byte load per workgroup");
    __shared__ uint32_t wsum[16];
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    auto row = [&](unsigned g) { return (uint4 *)(hist + (size_t)g * SORT_BINS) + t; };
    uint32_t tot[PER] = {0u, 0u, 0u, 0u};
    for (unsigned g0 = 0; g0 < groups; g0 += GB) {
        uint4 c[GB];
#pragma unroll
        for (int gg = 0; gg < GB; ++gg) c[gg] = g0 + gg < groups ? *row(g0 + gg) : uint4{0u, 0u, 0u, 0u};
#pragma unroll
        for (int gg = 0; gg < GB; ++gg) {
            tot[0] += c[gg].x;
            tot[1] += c[gg].y;
            tot[2] += c[gg].z;
            tot[3] += c[gg].w;
        }
    }
    const uint32_t s = tot[0] + tot[1] + tot[2] + tot[3];
    uint32_t incl = s;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const uint32_t up = __shfl_up(incl, d, 64);
        if (lane >= d) incl += up;
    }
    if (lane == 63) wsum[w] = incl;
    __syncthreads();
    uint32_t run[PER];
    run[0] = incl - s;
    for (int k = 0; k < w; ++k) run[0] += wsum[k];
#pragma unroll
    for (int u = 1; u < PER; ++u) run[u] = run[u - 1] + tot[u - 1];
    for (unsigned g0 = 0; g0 < groups; g0 += GB) {
        uint4 c[GB];
#pragma unroll
        for (int gg = 0; gg < GB; ++gg) c[gg] = g0 + gg < groups ? *row(g0 + gg) : uint4{0u, 0u, 0u, 0u};
#pragma unroll
        for (int gg = 0; gg < GB; ++gg) {
            if (g0 + gg >= groups) break;
            *row(g0 + gg) = uint4{run[0], run[1], run[2], run[3]};
            run[0] += c[gg].x;
            run[1] += c[gg].y;
            run[2] += c[gg].z;
            run[3] += c[gg].w;
        }
    }
}

template <class Acc>
__global__ void __launch_bounds__(1024) k_key_scatter(Acc acc, size_t n, const uint32_t *hist) {
    __shared__ uint32_t pos[SORT_BINS];
    const int t = threadIdx.x;
    for (int b = t; b < SORT_BINS; b += 1024) pos[b] = hist[(size_t)blockIdx.x * SORT_BINS + b];
    uint32_t kv[KS_U];
    ks_keys(acc, n, kv);
    __syncthreads();
#pragma unroll
    for (int u = 0; u < KS_U; ++u) {
        const size_t i = (size_t)blockIdx.x * KS_ITEMS + t + 1024 * u;
        if (i < n) acc.set_order(atomicAdd(&pos[kv[u]], 1u), (uint32_t)i);
    }
}

// the three kernels on `stream`, with `hist` of key_sort_groups(n) x
// SORT_BINS words
template <class Acc>
inline void launch_key_sort(const Acc &acc, size_t n, uint32_t *hist, hipStream_t stream) {
    const unsigned g = key_sort_groups(n);
    hipLaunchKernelGGL(k_key_hist<Acc>, dim3(g), dim3(1024), 0, stream, acc, n, hist);
    hipLaunchKernelGGL(k_key_scan, dim3(1), dim3(1024), 0, stream, hist, g);
    hipLaunchKernelGGL(k_key_scatter<Acc>, dim3(g), dim3(1024), 0, stream, acc, n, (const uint32_t *)hist);
}

// ---- shared-key tables ----------------------------------------------------------
constexpr int SK_MWL = 2;                  // entry words per lane in the column loop
static_assert(MW % SK_MWL == 0, "m-vectors of whole word pairs");
constexpr int SK_NWP = MW / SK_MWL;        // word pairs per entry
constexpr int SK_RPC = 64 / SK_NWP;        // rows per chunk: one per lane slot
constexpr int SK_CSTR = SK_RPC * MW * 4;   // bytes of one column's entries in a table
constexpr int sk_pow2ge(int x) {
    int p = 1;
    while (p < x) p <<= 1;
    return p;
}
// bytes per multiple (plane s of a table holds s times each entry word)
template <int CG>
constexpr int sk_sstr() {
    return sk_pow2ge(CG * SK_CSTR);
}
// fold items (j, word pair) and items per lane
constexpr int SK_NFI = K * SK_NWP;
constexpr int SK_IPL = (SK_NFI + 63) / 64;

// Shared-key variant codes (in the prep and verification variant lists):
// 1000 + 10 W + CG, W items per workgroup, CG columns per table.
constexpr bool sk_code(int code) { return code >= 1000; }
constexpr int sk_w(int code) { return (code % 1000) / 10; }
constexpr int sk_cg(int code) { return code % 10; }

// LDS of one wavefront of a table kernel
struct SkWave {
    uint64_t scol[N];  // scol[c] nibble j = S_j[c]
};

template <int W, int CG>
struct SkGroup {
    uint8_t tab[2][16 * sk_sstr<CG>()] __attribute__((aligned(16)));
    SkWave wa[W];
    int key[W];  // key index of each wavefront's item, -1 for none
};

// The column loop of one key's quadratic form for the wavefronts with `mine`
// set, the tables built by the whole workgroup (every wavefront calls it; it
// ends with a workgroup barrier). Rows: with NLR > 0 (signing), the columns x
// of L (entries L[c][x], every c) and then the rows of P1; with FULL
// (verification), the rows of [[P1, P2], [0, P3]] over the columns c < n.
// Entries come from a prepared key row (aligned, planar words). The rows of
// L store M_j[x] to Mv (planar, (j O + x) MW + w), the P rows T_j[r] to Tg
// (planar, (r K + j) MW + w).
template <int NLR, bool FULL, int W, int CG>
__device__ __forceinline__ void table_rows_shared(const uint8_t *key, bool mine, SkGroup<W, CG> &G, SkWave &A,
                                                  uint32_t *Mv, uint32_t *Tg) {
    using KL = PrepKeys;
    constexpr int EB = KL::EB;
    constexpr int NP = FULL ? N : V;   // P rows and columns
    constexpr int NVR = NLR + NP;      // rows of the schedule
    constexpr int NCH = (NVR + SK_RPC - 1) / SK_RPC;
    constexpr int SSTR = sk_sstr<CG>();
    constexpr int NQ = SK_RPC * CG * MW;  // words of one table
    constexpr int NT = 64 * W;
    constexpr int QPT = (NQ + NT - 1) / NT;
    const int tid = threadIdx.x, lane = tid & 63;
    mine = uni((uint32_t)mine) != 0u;  // uniform per wavefront (SGPR control flow)
    // first column of chunk ch (its first row's start)
    auto cstart = [](int ch) {
        const int v0 = ch * SK_RPC;
        return v0 < NLR ? 0 : v0 - NLR;
    };
    auto next = [&](int &ch, int &c0) {
        c0 += CG;
        if (c0 >= NP) {
            ++ch;
            c0 = cstart(ch);
        }
    };
    // producing threads: word q of the table = (column cl, row, word w)
    int qcl[QPT], qrow[QPT], qw[QPT];
#pragma unroll
    for (int u = 0; u < QPT; ++u) {
        const int q = tid + u * NT;
        qcl[u] = q / (SK_RPC * MW);
        qrow[u] = (q / MW) % SK_RPC;
        qw[u] = q % MW;
    }
    // Entry of (row vr, column c) at byte base[c >= V] + c stride[c >= V]
    // for cmin <= c < NP (zero elsewhere): set per producing word when the
    // producer moves to a new chunk.
    int fb_ch = -1;
    int rb0[QPT], rb1[QPT], cmin[QPT];
    auto row_bases = [&](int ch) {
#pragma unroll
        for (int u = 0; u < QPT; ++u) {
            const int vr = ch * SK_RPC + qrow[u];
            rb0[u] = rb1[u] = 0;
            cmin[u] = NP;  // no entries
            if (tid + u * NT >= NQ || vr >= NVR) continue;
            if (vr < NLR) {
                // column x = vr of L: entry (c, x) at ESK_L + (c O + x) EB
                rb0[u] = (int)KL::ESK_L + vr * EB + 4 * qw[u];
                cmin[u] = 0;
            } else {
                const int r = vr - NLR;
                cmin[u] = r;
                if constexpr (FULL) {
                    if (r < V) {
                        rb0[u] = (int)KL::EPK_P1 + (tri_entry(r, 0, V)) * EB + 4 * qw[u];
                        rb1[u] = (int)KL::EPK_P2 + (r * O - V) * EB + 4 * qw[u];
                    } else {
                        rb1[u] = (int)KL::EPK_P3 + (tri_entry(r - V, 0, O) - V) * EB + 4 * qw[u];
                    }
                } else {
                    rb0[u] = (int)KL::ESK_P1 + tri_entry(r, 0, V) * EB + 4 * qw[u];
                }
            }
        }
    };
    // entry words of interval (ch, c0), zero outside the matrix
    auto fetch = [&](int ch, int c0, uint32_t (&x)[QPT]) {
        if (ch != fb_ch && ch < NCH) {
            row_bases(ch);
            fb_ch = ch;
        }
#pragma unroll
        for (int u = 0; u < QPT; ++u) {
            const int c = c0 + qcl[u];
            const bool lrow = NLR > 0 && ch * SK_RPC + qrow[u] < NLR;
            // P rows: consecutive columns are consecutive entries; L columns
            // are O entries apart
            const int off = c < V ? rb0[u] + c * (lrow ? O * EB : EB) : rb1[u] + c * EB;
            x[u] = ch < NCH && c >= cmin[u] && c < NP ? *(const uint32_t *)(key + off) : 0u;
        }
    };
    // the 15 nonzero multiples of each word into table `buf`
    auto produce = [&](int buf, const uint32_t (&x)[QPT]) {
#pragma unroll
        for (int u = 0; u < QPT; ++u) {
            const int q = tid + u * NT;
            if (q >= NQ) continue;
            using namespace gf16::planar;
            const uint32_t x1 = x[u], x2 = mul_a(x1), x4 = mul_a(x2), x8 = mul_a(x4);
            const uint32_t x3 = x1 ^ x2, x5 = x1 ^ x4, x6 = x2 ^ x4, x7 = x3 ^ x4;
            uint32_t *d = (uint32_t *)G.tab[buf] + q;
            constexpr int S4 = SSTR / 4;
            d[1 * S4] = x1;
            d[2 * S4] = x2;
            d[3 * S4] = x3;
            d[4 * S4] = x4;
            d[5 * S4] = x5;
            d[6 * S4] = x6;
            d[7 * S4] = x7;
            d[8 * S4] = x8;
            d[9 * S4] = x1 ^ x8;
            d[10 * S4] = x2 ^ x8;
            d[11 * S4] = x3 ^ x8;
            d[12 * S4] = x4 ^ x8;
            d[13 * S4] = x5 ^ x8;
            d[14 * S4] = x6 ^ x8;
            d[15 * S4] = x7 ^ x8;
        }
    };
    // column-loop lane: word pair g of row slot `slot`. The lanes past the
    // last slot (64 - SK_RPC SK_NWP of them) read the padding after the
    // table's last row, which keeps them off the banks of the real rows.
    const int g = lane % SK_NWP, slot = lane / SK_NWP;
    const bool active = slot < SK_RPC;
    static_assert((CG - 1) * SK_CSTR + SK_CSTR + 8 * (64 - SK_RPC * SK_NWP) <= SSTR, "idle lanes stay in the plane");
    const int vlane = slot * MW * 4 + g * 8;
    uint32_t T[K][SK_MWL];
#pragma unroll
    for (int j = 0; j < K; ++j) T[j][0] = T[j][1] = 0;
    // scalars of the interval's columns, read one interval ahead
    uint64_t scv[CG];
    auto scalars = [&](int c0) {
#pragma unroll
        for (int cl = 0; cl < CG; ++cl) scv[cl] = c0 + cl < NP ? A.scol[c0 + cl] : 0ull;
    };
    // the K table reads of column cl of an interval (issued together), and
    // their sums into T
    auto issue = [&](int buf, int cl, uint2 (&v)[K]) {
        const uint8_t *bb = G.tab[buf] + vlane + cl * SK_CSTR;
        const uint64_t sc = uni64(scv[cl]);
#pragma unroll
        for (int j = 0; j < K; ++j) {
            const uint32_t s = (uint32_t)(sc >> (4 * j)) & 15u;
            v[j] = *(const uint2 *)(bb + s * SSTR);
        }
#pragma unroll
        for (int j = 0; j < K; ++j) {
            __builtin_amdgcn_sched_group_barrier(0x0002, 1, 0);  // the address
            __builtin_amdgcn_sched_group_barrier(0x0100, 1, 0);  // its read
        }
    };
    auto absorb = [&](const uint2 (&v)[K]) {
#pragma unroll
        for (int j = 0; j < K; ++j) {
            T[j][0] ^= v[j].x;
            T[j][1] ^= v[j].y;
        }
    };
    // end of chunk ch: the L rows go to Mv, the P rows to Tg
    auto chunk_end = [&](int ch) {
        const int vr = ch * SK_RPC + slot;
        if (active && vr < NVR) {
            if (vr < NLR) {
#pragma unroll
                for (int j = 0; j < K; ++j)
                    *(uint2 *)&Mv[(j * O + vr) * MW + 2 * g] = uint2{T[j][0], T[j][1]};
            } else {
                uint32_t *row = Tg + (size_t)(vr - NLR) * K * MW + 2 * g;
#pragma unroll
                for (int j = 0; j < K; ++j) *(uint2 *)&row[j * MW] = uint2{T[j][0], T[j][1]};
            }
        }
#pragma unroll
        for (int j = 0; j < K; ++j) T[j][0] = T[j][1] = 0;
    };
    // the interval sequence: (chunk, first column), consumed in order; the
    // table of interval i + 1 is produced during interval i
    int ch = 0, c0 = cstart(0), pch = 0, pc0 = c0;
    // entry words of the next PFD intervals to produce, fetched PFD
    // intervals ahead (one interval lasts less than a global-memory round
    // trip under load); (fch, fc0) is the next interval to fetch
    constexpr int PFD = 3;
    uint32_t xr[PFD][QPT];
    int fch = pch, fc0 = pc0;
    auto advance = [&]() {
#pragma unroll
        for (int d = 0; d + 1 < PFD; ++d)
#pragma unroll
            for (int u = 0; u < QPT; ++u) xr[d][u] = xr[d + 1][u];
        fetch(fch, fc0, xr[PFD - 1]);
        next(fch, fc0);
        next(pch, pc0);
    };
#pragma unroll
    for (int d = 0; d < PFD; ++d) {
        fetch(fch, fc0, xr[d]);
        next(fch, fc0);
    }
    produce(0, xr[0]);
    advance();
    if (mine) scalars(c0);
    __syncthreads();
    int buf = 0;
#pragma unroll 1
    for (;;) {
        int nch = ch, nc0 = c0;
        next(nch, nc0);
        // column 0's table reads go out first; the next table's production
        // runs while they are in flight
        uint2 v[K];
        if (mine) issue(buf, 0, v);
        if (pch < NCH) {
            produce(buf ^ 1, xr[0]);
            advance();
        }
        if (mine) {
            absorb(v);
#pragma unroll
            for (int cl = 1; cl < CG; ++cl) {
                if (c0 + cl >= NP) break;
                issue(buf, cl, v);
                absorb(v);
            }
            if (nch < NCH) scalars(nc0);
            if (nch != ch) chunk_end(ch);
        }
        ch = nch;
        c0 = nc0;
        __syncthreads();
        if (ch >= NCH) break;
        buf ^= 1;
    }
}

// The workgroup's items and their keys; kernels then loop over the keys.
template <class Acc, int W, int CG>
__device__ __forceinline__ size_t sk_item(const Acc &acc, SkGroup<W, CG> &G, size_t base, size_t end, bool &has) {
    const int wave = (int)uni(threadIdx.x >> 6);
    const size_t k = base + (size_t)blockIdx.x * W + wave;
    has = k < end;
    const size_t item = has ? uni64(acc.order(k)) : 0;  // uniform: scalar registers
    if ((threadIdx.x & 63) == 0) G.key[wave] = has ? (int)acc.key(item) : -1;
    return item;
}

// Bit set w: wave w's item has key G.key of the first wave not yet done.
template <int W, int CG>
__device__ __forceinline__ int sk_next_key(const SkGroup<W, CG> &G, int done, int &kk) {
    kk = -1;
#pragma unroll
    for (int w = 0; w < W; ++w)
        if (kk < 0 && !((done >> w) & 1) && G.key[w] >= 0) kk = G.key[w];
    int sel = 0;
#pragma unroll
    for (int w = 0; w < W; ++w)
        if (kk >= 0 && G.key[w] == kk) sel |= 1 << w;
    return sel;
}

// Batch form over prepared keys, attempt 0, part 1 (as k_sign_prep, with
// k_sign_fold after it): M and the T rows of positions [base, end) of the
// key order, W per workgroup; the failure list of the part starts empty.
template <class Acc, int CODE>
__global__ void __launch_bounds__(64 * sk_w(CODE), 8)
    k_sign_table_shared(Acc acc, size_t n, int ordered, size_t base, size_t end) {
    constexpr int W = sk_w(CODE), CG = sk_cg(CODE);
    __shared__ SkGroup<W, CG> G;
    const int tid = threadIdx.x, lane = tid & 63, wave = (int)uni(threadIdx.x >> 6);
    if (blockIdx.x == 0 && tid == 0) *(uint32_t *)(acc.scratch_of(base) + SCR_LIST_COUNT) = 0;
    bool has;
    const size_t item = sk_item(acc, G, base, end, has);
    SignItem it = has ? acc(item) : SignItem{};
    SkWave &A = G.wa[wave];
    if (has) {
        const uint8_t *vr = it.scratch + SCR_VR;
        for (int c = lane; c < V; c += 64) {
            uint64_t v = 0;
            for (int j = 0; j < K; ++j) v |= (uint64_t)mvec::nib(vr + j * V_BYTES, c) << (4 * j);
            A.scol[c] = v;
        }
    }
    for (int q = tid; q < sk_sstr<CG>() / 4; q += 64 * W) {
        ((uint32_t *)G.tab[0])[q] = 0;
        ((uint32_t *)G.tab[1])[q] = 0;
    }
    __syncthreads();
    uint32_t *Mv = has ? (uint32_t *)(it.scratch + SCR_MV) : nullptr;
    uint32_t *Tg = has ? acc.trow(item) : nullptr;
    int done = 0;
#pragma unroll 1
    for (;;) {
        int kk;
        const int sel = sk_next_key(G, done, kk);
        if (kk < 0) break;
        table_rows_shared<O, false, W, CG>(acc.key_row(uni((uint32_t)kk)), (sel >> wave) & 1, G, A, Mv, Tg);
        done |= sel;
    }
}

// Verification over prepared keys, part 1 (k_verify_fold after it): the T
// rows of positions [0, n) of the key order, W per workgroup.
template <class Acc, int CODE>
__global__ void __launch_bounds__(64 * sk_w(CODE), 8) k_verify_table_shared(Acc acc, size_t n) {
    constexpr int W = sk_w(CODE), CG = sk_cg(CODE);
    __shared__ SkGroup<W, CG> G;
    const int tid = threadIdx.x, lane = tid & 63, wave = (int)uni(threadIdx.x >> 6);
    bool has;
    const size_t item = sk_item(acc, G, 0, n, has);
    VerifyItem it = has ? acc(item) : VerifyItem{};
    SkWave &A = G.wa[wave];
    if (has) {
        for (int c = lane; c < N; c += 64) {
            uint64_t v = 0;
            for (int j = 0; j < K; ++j) v |= (uint64_t)mvec::nib(it.sig, j * N + c) << (4 * j);
            A.scol[c] = v;
        }
    }
    for (int q = tid; q < sk_sstr<CG>() / 4; q += 64 * W) {
        ((uint32_t *)G.tab[0])[q] = 0;
        ((uint32_t *)G.tab[1])[q] = 0;
    }
    __syncthreads();
    uint32_t *Tg = has ? acc.trow(item) : nullptr;
    int done = 0;
#pragma unroll 1
    for (;;) {
        int kk;
        const int sel = sk_next_key(G, done, kk);
        if (kk < 0) break;
        table_rows_shared<0, true, W, CG>(acc.key_row(uni((uint32_t)kk)), (sel >> wave) & 1, G, A, nullptr, Tg);
        done |= sel;
    }
}

// ---- the fold -------------------------------------------------------------------
// U_(a,j) += S_a[r] T_j[r] over the NP rows of T (one wavefront per item), in
// the transposed layout: lanes own fold items (j, word pair gp) and keep
// U_(a,j) for every a in registers (2K per item); a row's scalars S_a[r] are
// uniform across the wave and select each product from the 16 multiples of
// the lane's T words, and the lanes' T words of a row are K MW consecutive
// words. Rows are read FOLD_AHEAD ahead of their use.
// U_(a,j) + U_(j,a) is the pair sum of the bitsliced kernels
// (sign.cu:395-437, verify.cu:72-182).
constexpr int FOLD_AHEAD = 8;
constexpr int FOLD_W = 4;  // wavefronts per workgroup

struct SkFold {
    uint64_t scol[N];
    uint32_t t[MW];             // t, read before the fold
    uint32_t U[PAIRS * MW];     // pair sums, nibble-packed
    uint32_t yu[YG_MAX * UW];   // unreduced y
};

// One row of the fold for one fold item: U[a][h] += S_a[r] x_h for the K
// scalars S_a[r] (nibble a of the uniform lo:hi) and the item's two words
// x_0, x_1. The 16 multiples s x_h are built in v0..v31 (multiple s of word
// h in v[2 s + h]; x_h itself arrives in v2 / v3), and each product is one
// XOR whose first operand is read relative to the scalar (GPR indexing mode,
// index 2 S_a[r]): 63 VALU and 40 SALU instructions per row against 4
// AND-XORs per product. Registers only; no memory access. mul_a (s -> a s in the planar
// layout, a^4 = a + 1): the bytes (p0, p1, p2, p3) -> (p3, p0 ^ p3, p1, p2),
// a rotation and the top byte XORed into byte 1.
#define SKF_MULA(SRC, DST)                                                                      \
    "v_alignbit_b32 " DST ", " SRC ", " SRC ", 24\n\t"                                        \
    "v_lshlrev_b32_sdwa v1, v0, " SRC " dst_sel:DWORD dst_unused:UNUSED_PAD src0_sel:DWORD "  \
    "src1_sel:BYTE_3\n\t"                                                                      \
    "v_xor_b32 " DST ", " DST ", v1\n\t"
#define SKF_MULTIPLES                                                                           \
    "v_mov_b32 v0, 8\n\t"                                                                       \
    SKF_MULA("v2", "v4") SKF_MULA("v4", "v8") SKF_MULA("v8", "v16")                             \
    SKF_MULA("v3", "v5") SKF_MULA("v5", "v9") SKF_MULA("v9", "v17")                             \
    "v_xor_b32 v6, v2, v4\n\tv_xor_b32 v7, v3, v5\n\t"                                          \
    "v_xor_b32 v10, v2, v8\n\tv_xor_b32 v11, v3, v9\n\t"                                        \
    "v_xor_b32 v12, v4, v8\n\tv_xor_b32 v13, v5, v9\n\t"                                        \
    "v_xor_b32 v14, v6, v8\n\tv_xor_b32 v15, v7, v9\n\t"                                        \
    "v_xor_b32 v18, v2, v16\n\tv_xor_b32 v19, v3, v17\n\t"                                      \
    "v_xor_b32 v20, v4, v16\n\tv_xor_b32 v21, v5, v17\n\t"                                      \
    "v_xor_b32 v22, v6, v16\n\tv_xor_b32 v23, v7, v17\n\t"                                      \
    "v_xor_b32 v24, v8, v16\n\tv_xor_b32 v25, v9, v17\n\t"                                      \
    "v_xor_b32 v26, v10, v16\n\tv_xor_b32 v27, v11, v17\n\t"                                    \
    "v_xor_b32 v28, v12, v16\n\tv_xor_b32 v29, v13, v17\n\t"                                    \
    "v_xor_b32 v30, v14, v16\n\tv_xor_b32 v31, v15, v17\n\t"                                    \
    "v_mov_b32 v0, 0\n\tv_mov_b32 v1, 0\n\t"
// product a: index 2 (nibble a of lo:hi), then U[a] ^= the two indexed
// words, indexing mode switched on and off around them (the form the
// compiler itself emits for a register-indexed read; the index is not
// switched while the mode is on)
#define SKF_A(A, SRC, BFE)                                                                      \
    "s_bfe_u32 %[t], %[" #SRC "], " #BFE "\n\t"                                                 \
    "s_lshl_b32 %[t], %[t], 1\n\t"                                                              \
    "s_set_gpr_idx_on %[t], gpr_idx(SRC0)\n\t"                                                  \
    "v_xor_b32 %[u" #A "a], v0, %[u" #A "a]\n\t"                                                \
    "v_xor_b32 %[u" #A "b], v1, %[u" #A "b]\n\t"                                                \
    "s_set_gpr_idx_off\n\t"
#define SKF_OP(A) [u##A##a] "+v"(U[A][0]), [u##A##b] "+v"(U[A][1])
#define SKF_A0_3 SKF_A(0, lo, 0x40000) SKF_A(1, lo, 0x40004) SKF_A(2, lo, 0x40008) SKF_A(3, lo, 0x4000c)
#define SKF_A4_7 SKF_A(4, lo, 0x40010) SKF_A(5, lo, 0x40014) SKF_A(6, lo, 0x40018) SKF_A(7, lo, 0x4001c)
#if MAYO_VERSION == 2
#define SKF_PRODUCTS SKF_A0_3
#define SKF_OPS SKF_OP(0), SKF_OP(1), SKF_OP(2), SKF_OP(3)
#elif MAYO_VERSION == 1
#define SKF_PRODUCTS SKF_A0_3 SKF_A4_7 SKF_A(8, hi, 0x40000) SKF_A(9, hi, 0x40004)
#define SKF_OPS SKF_OP(0), SKF_OP(1), SKF_OP(2), SKF_OP(3), SKF_OP(4), SKF_OP(5), SKF_OP(6), SKF_OP(7), \
                SKF_OP(8), SKF_OP(9)
#elif MAYO_VERSION == 3
#define SKF_PRODUCTS SKF_A0_3 SKF_A4_7 SKF_A(8, hi, 0x40000) SKF_A(9, hi, 0x40004) SKF_A(10, hi, 0x40008)
#define SKF_OPS SKF_OP(0), SKF_OP(1), SKF_OP(2), SKF_OP(3), SKF_OP(4), SKF_OP(5), SKF_OP(6), SKF_OP(7), \
                SKF_OP(8), SKF_OP(9), SKF_OP(10)
#else
#define SKF_PRODUCTS \
    SKF_A0_3 SKF_A4_7 SKF_A(8, hi, 0x40000) SKF_A(9, hi, 0x40004) SKF_A(10, hi, 0x40008) SKF_A(11, hi, 0x4000c)
#define SKF_OPS SKF_OP(0), SKF_OP(1), SKF_OP(2), SKF_OP(3), SKF_OP(4), SKF_OP(5), SKF_OP(6), SKF_OP(7), \
                SKF_OP(8), SKF_OP(9), SKF_OP(10), SKF_OP(11)
#endif
__device__ __forceinline__ void fold_row_asm(uint32_t (&U)[K][SK_MWL], uint2 x, uint32_t lo, uint32_t hi) {
    uint32_t x0 = x.x, x1 = x.y, t;
    asm volatile(SKF_MULTIPLES SKF_PRODUCTS
                 : [t] "=&s"(t), "+{v2}"(x0), "+{v3}"(x1), SKF_OPS
                 : [lo] "s"(lo), [hi] "s"(hi)
                 : "v0", "v1", "v4", "v5", "v6", "v7", "v8", "v9", "v10", "v11", "v12", "v13", "v14", "v15",
                   "v16", "v17", "v18", "v19", "v20", "v21", "v22", "v23", "v24", "v25", "v26", "v27", "v28",
                   "v29", "v30", "v31", "m0", "scc");
}

template <int NP>
__device__ __forceinline__ void fold_rows(const uint64_t *scol, const uint32_t *Tg, uint32_t (&U)[SK_IPL][K][SK_MWL]) {
    const int lane = lane_here();
    constexpr int RW = K * MW;  // words of a row
    constexpr int NR = (NP + FOLD_AHEAD - 1) / FOLD_AHEAD * FOLD_AHEAD;
#pragma unroll
    for (int i = 0; i < SK_IPL; ++i)
#pragma unroll
        for (int a = 0; a < K; ++a) U[i][a][0] = U[i][a][1] = 0;
    // word pair of fold item it = (j, gp) in row r: r K MW + j MW + 2 gp = r RW + 2 it
    int off[SK_IPL];
#pragma unroll
    for (int i = 0; i < SK_IPL; ++i) {
        const int it = lane + 64 * i;
        off[i] = it < SK_NFI ? 2 * it : 0;
    }
    auto load = [&](int r, uint2 (&t)[SK_IPL]) {
        const uint32_t *row = Tg + (size_t)(r < NP ? r : NP - 1) * RW;
#pragma unroll
        for (int i = 0; i < SK_IPL; ++i) t[i] = *(const uint2 *)(row + off[i]);
    };
    uint2 ring[FOLD_AHEAD][SK_IPL];
#pragma unroll
    for (int d = 0; d < FOLD_AHEAD; ++d) load(d, ring[d]);
#pragma unroll 1
    for (int r0 = 0; r0 < NR; r0 += FOLD_AHEAD) {
#pragma unroll
        for (int d = 0; d < FOLD_AHEAD; ++d) {
            const int r = r0 + d;
            // rows past the last read the last row again and add nothing
            const uint64_t sr = uni64(r < NP ? scol[r] : 0ull);
            const uint32_t sh[2] = {(uint32_t)sr, (uint32_t)(sr >> 32)};
#pragma unroll
            for (int i = 0; i < SK_IPL; ++i) fold_row_asm(U[i], ring[d][i], sh[0], sh[1]);
            load(r + FOLD_AHEAD, ring[d]);
        }
    }
}

// The pair sums of the fold items' ordered products into Ul (nibble words):
// U_(a,j) + U_(j,a) for a < j, U_(a,a). The planar words are summed first
// (a sum is the same XOR in either layout) and each pair sum is turned into
// nibble form once. Lanes of fold items past the last add nothing.
__device__ __forceinline__ void sk_pair_sums(uint32_t *Ul, const uint32_t (&U)[SK_IPL][K][SK_MWL]) {
    const int lane = lane_here();
    for (int e = lane; e < PAIRS * MW; e += 64) Ul[e] = 0;
    WaveSync::sync();
#pragma unroll
    for (int i = 0; i < SK_IPL; ++i) {
        const int it = lane + 64 * i;
        if (it >= SK_NFI) continue;
        const int j = it / SK_NWP, gp = it % SK_NWP;
#pragma unroll
        for (int a = 0; a < K; ++a) {
            const int p = a <= j ? pair_index(a, j) : pair_index(j, a);
#pragma unroll
            for (int h = 0; h < SK_MWL; ++h) atomicXor(&Ul[p * MW + 2 * gp + h], U[i][a][h]);
        }
    }
    WaveSync::sync();
    for (int e = lane; e < PAIRS * MW; e += 64) Ul[e] = gf16::planar::to_nibbles(Ul[e]);
    WaveSync::sync();
}

// Signing, attempt 0, part 1 after k_sign_table_shared: U and y of positions
// [base, end) of the key order, one wavefront per item.
template <class Acc>
__global__ void __launch_bounds__(64 * FOLD_W) k_sign_fold(Acc acc, size_t n, size_t base, size_t end) {
    __shared__ SkFold F[FOLD_W];
    const int lane = lane_here(), wave = (int)uni(threadIdx.x >> 6);
    const size_t k = base + (size_t)blockIdx.x * FOLD_W + wave;
    if (k >= end) return;  // whole wavefronts; no workgroup barrier follows
    const size_t item = uni64(acc.order(k));
    SkFold &A = F[wave];
    const SignItem it = acc(item);
    const uint8_t *vr = it.scratch + SCR_VR;
    for (int c = lane; c < V; c += 64) {
        uint64_t v = 0;
        for (int j = 0; j < K; ++j) v |= (uint64_t)mvec::nib(vr + j * V_BYTES, c) << (4 * j);
        A.scol[c] = v;
    }
    for (int q = lane; q < MW; q += 64) A.t[q] = mvec::load_mw(it.scratch + SCR_T, q);
    WaveSync::sync();
    uint32_t U[SK_IPL][K][SK_MWL];
    fold_rows<V>(A.scol, acc.trow(item), U);
    sk_pair_sums(A.U, U);
    unreduced_y<WaveSync>(A.U, A.yu);
    if (lane == 0) {
        uint32_t a[UW], out[MW];
#pragma unroll
        for (int q = 0; q < UW; ++q) a[q] = A.yu[q];
        reduce_unreduced(a, out);
        uint32_t *yv = (uint32_t *)(it.scratch + SCR_Y);
#pragma unroll
        for (int q = 0; q < MW; ++q) yv[q] = out[q] ^ A.t[q];
    }
}

// Verification after k_verify_table_shared: y of positions [0, n) of the key
// order against t (from the hashing pass), one wavefront per item.
template <class Acc>
__global__ void __launch_bounds__(64 * FOLD_W) k_verify_fold(Acc acc, size_t n) {
    __shared__ SkFold F[FOLD_W];
    const int lane = lane_here(), wave = (int)uni(threadIdx.x >> 6);
    const size_t k = (size_t)blockIdx.x * FOLD_W + wave;
    if (k >= n) return;
    const size_t item = uni64(acc.order(k));
    SkFold &A = F[wave];
    const VerifyItem it = acc(item);
    for (int c = lane; c < N; c += 64) {
        uint64_t v = 0;
        for (int j = 0; j < K; ++j) v |= (uint64_t)mvec::nib(it.sig, j * N + c) << (4 * j);
        A.scol[c] = v;
    }
    for (int w = lane; w < MW; w += 64) {
        const int nv = mvec::word_valid(w);
        uint32_t tw = 0;
        for (int b = 0; b < nv; ++b) tw |= (uint32_t)it.t[4 * w + b] << (8 * b);
        A.t[w] = tw;
    }
    WaveSync::sync();
    uint32_t U[SK_IPL][K][SK_MWL];
    fold_rows<N>(A.scol, acc.trow(item), U);
    sk_pair_sums(A.U, U);
    unreduced_y<WaveSync>(A.U, A.yu);
    if (lane == 0) {
        uint32_t a[UW], y[MW];
#pragma unroll
        for (int q = 0; q < UW; ++q) a[q] = A.yu[q];
        reduce_unreduced(a, y);
        bool eq = true;
#pragma unroll
        for (int w = 0; w < MW; ++w) eq &= (y[w] == A.t[w]);
        it.result[0] = eq ? 0 : 1;
    }
}

// Shared-key variants compiled into every library: 8 items per workgroup
// with two or one table columns (two tables of 16 KB or 8 KB per multiple
// set), and 4 items with two columns. The prep and verification variants the
// launch configuration accepts are the wave kernels' and these.
inline constexpr int SK_CODES[3] = {1082, 1081, 1042};
inline constexpr int PREP_VARIANTS[7] = {PREP_OCCS[0], PREP_OCCS[1], PREP_OCCS[2], PREP_OCCS[3],
                                         SK_CODES[0],  SK_CODES[1],  SK_CODES[2]};
inline constexpr int VERIFY_VARIANTS[7] = {VERIFY_OCCS[0], VERIFY_OCCS[1], VERIFY_OCCS[2], VERIFY_OCCS[3],
                                           SK_CODES[0],    SK_CODES[1],    SK_CODES[2]};

}  // namespace wv
}  // namespace mayo
