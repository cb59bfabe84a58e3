// Stage kernels of the MAYO engine for one parameter set.
//
// Work split (gfx950, wave64):
//  * hashing (SHAKE for seeds, digests, salts, targets, vinegar vectors):
//    one lane per item, so a wave hashes 64 items at once;
//  * AES-128-CTR expansion of seed_pk: one workgroup per key, one 16-byte
//    keystream block per lane-iteration, T-tables in LDS;
//  * GF(16) linear algebra: one 256-lane workgroup per item, m-vectors as
//    nibble-packed words, every scalar x vector product as 4 AND-XORs on the
//    shared multiples of the vector (gf16.hpp);
//  * elimination: rows packed 8 nibbles per word in LDS, all lanes sweep the
//    rows below the pivot, pivot search by LDS atomicMin.
// Kernels are templates over an item accessor so the reference's pitched
// pools (api.cu layouts) and the key-table batch form share the same code.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

#include "aes.hpp"
#include "gf16.hpp"
#include "keccak.hpp"
#include "mayo_params.hpp"
#include "mvec.hpp"

namespace mayo {

constexpr int WG = 256;
constexpr int RW = (KO + 7) / 8;  // words per packed row of A

// Scratch per signing item: digest || salt || t || (V || r) of attempt 0.
constexpr int SCR_DIGEST = 0;
constexpr int SCR_SALT = DIGEST;
constexpr int SCR_T = DIGEST + SK;
constexpr int SCR_VR = DIGEST + SK + MB;
constexpr int SCR_ATTEMPTS = SCR_VR + VR_BYTES;  // attempts used (0 = all failed)
constexpr int SIGN_SCRATCH = SCR_ATTEMPTS + 1;
// Batch form only: the list of items whose attempt 0 failed (entry k in
// item k's scratch, the count in item 0's), and the M products and y of
// attempt 0, handed from the product kernel to the solve kernel.
constexpr int SCR_LIST = (SIGN_SCRATCH + 3) / 4 * 4;
constexpr int SCR_LIST_COUNT = SCR_LIST + 4;
constexpr int SCR_ORDER = SCR_LIST_COUNT + 4;  // hashing order (entry k in item k's scratch)
constexpr int SCR_MV = (SCR_ORDER + 4 + 63) / 64 * 64;
constexpr int SCR_Y = SCR_MV + K * O * MW * 4;
constexpr int SIGN_SCRATCH_STRIDE = (SCR_Y + MW * 4 + 63) / 64 * 64;
constexpr uint8_t ATTEMPT_PENDING = 0xFF;   // rank failure left for the retry pass
// Scratch per verifying item: t, and the hashing order slot.
constexpr int VSCR_ORDER = (MB + 3) / 4 * 4;
constexpr int VERIFY_SCRATCH_STRIDE = (VSCR_ORDER + 4 + 15) / 16 * 16;
// Table-form workspaces end with one T region per item for the shared-key
// variants: T_j[r] of every row r the pair sums fold (the V rows of P1 for
// signing, the N rows of [[P1, P2], [0, P3]] for verification), K m-vectors
// of MW words per row, row-major (r K + j) MW + w, planar words.
constexpr size_t SIGN_T_BYTES = (size_t)V * K * MW * 4;
constexpr size_t VERIFY_T_BYTES = (size_t)N * K * MW * 4;
// Hashing order: items grouped by the SHAKE256 blocks (136 bytes) their
// message absorbs; ORDER_BINS - 1 and more share the last group.
constexpr int ORDER_BINS = 256;
__device__ __forceinline__ uint32_t order_bin(int32_t len) {
    const uint32_t b = (uint32_t)len / 136u;
    return b < ORDER_BINS - 1 ? b : ORDER_BINS - 1;
}

// ---------------------------------------------------------------------------
// Item accessors
// ---------------------------------------------------------------------------

struct SignItem {
    const uint8_t *esk;  // sk || O || P1 || L
    const uint8_t *msg;
    uint32_t mlen;
    uint8_t *sig;
    uint8_t *scratch;
};

// Reference pitched pool: every row holds esk, message, signature and the
// per-row scratch region at fixed offsets (api.cu:97-107, 125-137).
struct PitchedSign {
    uint8_t *pool;
    size_t pitch, msg_off, sig_off, scr_off;
    uint32_t mlen;
    __device__ SignItem operator()(size_t i) const {
        uint8_t *row = pool + i * pitch;
        return {row, row + msg_off, mlen, row + sig_off, row + scr_off};
    }
};

// Batch form: a table of expanded keys, per-item key index and ragged
// messages; signatures and scratch are dense arrays.
struct TableSign {
    const uint8_t *esk_table;
    size_t esk_stride;
    const int32_t *key_idx;
    const uint8_t *msgs;
    const int64_t *msg_off;
    const int32_t *msg_len;
    uint8_t *sigs;
    uint8_t *scratch;
    uint8_t *trows = nullptr;  // T regions (SIGN_T_BYTES per item), or none
    __device__ SignItem operator()(size_t i) const {
        return {esk_table + (size_t)key_idx[i] * esk_stride, msgs + msg_off[i],
                (uint32_t)msg_len[i], sigs + i * (size_t)SIG_BYTES,
                scratch + i * (size_t)SIGN_SCRATCH_STRIDE};
    }
    __device__ uint8_t *scratch_of(size_t i) const { return scratch + i * (size_t)SIGN_SCRATCH_STRIDE; }
    __device__ uint32_t bin(size_t i) const { return order_bin(msg_len[i]); }
    __device__ uint32_t key(size_t i) const { return (uint32_t)key_idx[i]; }
    __device__ const uint8_t *key_row(uint32_t k) const { return esk_table + (size_t)k * esk_stride; }
    __device__ uint32_t *trow(size_t i) const { return (uint32_t *)(trows + i * SIGN_T_BYTES); }
    __device__ void set_order(size_t k, uint32_t i) const { *(uint32_t *)(scratch_of(k) + SCR_ORDER) = i; }
    __device__ size_t order(size_t k) const { return *(const uint32_t *)(scratch_of(k) + SCR_ORDER); }
};

struct VerifyItem {
    const uint8_t *epk;  // P1 || P2 || P3
    const uint8_t *msg;
    uint32_t mlen;
    const uint8_t *sig;
    uint8_t *result;
    uint8_t *t;          // precomputed target, or nullptr to derive in-kernel
};

struct PitchedVerify {
    uint8_t *pool;
    size_t pitch, msg_off, sig_off, res_off;
    uint32_t mlen;
    __device__ VerifyItem operator()(size_t i) const {
        uint8_t *row = pool + i * pitch;
        return {row, row + msg_off, mlen, row + sig_off, row + res_off, nullptr};
    }
    __device__ size_t order(size_t k) const { return k; }  // pools run in row order
};

struct TableVerify {
    const uint8_t *epk_table;
    size_t epk_stride;
    const int32_t *key_idx;
    const uint8_t *msgs;
    const int64_t *msg_off;
    const int32_t *msg_len;
    const uint8_t *sigs;
    uint8_t *results;
    uint8_t *scratch;
    uint8_t *trows = nullptr;  // T regions (VERIFY_T_BYTES per item), or none
    __device__ VerifyItem operator()(size_t i) const {
        return {epk_table + (size_t)key_idx[i] * epk_stride, msgs + msg_off[i],
                (uint32_t)msg_len[i], sigs + i * (size_t)SIG_BYTES, results + i,
                scratch ? scratch + i * (size_t)VERIFY_SCRATCH_STRIDE : nullptr};
    }
    __device__ uint8_t *scratch_of(size_t i) const { return scratch + i * (size_t)VERIFY_SCRATCH_STRIDE; }
    __device__ uint32_t bin(size_t i) const { return order_bin(msg_len[i]); }
    __device__ uint32_t key(size_t i) const { return (uint32_t)key_idx[i]; }
    __device__ const uint8_t *key_row(uint32_t k) const { return epk_table + (size_t)k * epk_stride; }
    __device__ uint32_t *trow(size_t i) const { return (uint32_t *)(trows + i * VERIFY_T_BYTES); }
    __device__ void set_order(size_t k, uint32_t i) const { *(uint32_t *)(scratch_of(k) + VSCR_ORDER) = i; }
    __device__ size_t order(size_t k) const { return *(const uint32_t *)(scratch_of(k) + VSCR_ORDER); }
};

// ---------------------------------------------------------------------------
// Hashing: one lane per item
// ---------------------------------------------------------------------------

// keypair seed rule (keypair.cu:32-45) then S = SHAKE256(sk) = seed_pk || O
// (keypair.cu:47, expandsk.cu:29). When gen_sk is set the sk is generated
// from LE64(rand_index + 2 i + 1) and stored at sk; otherwise it is read there.
__global__ void __launch_bounds__(64) k_seed(uint8_t *pool, size_t pitch, size_t n, int gen_sk,
                                             uint64_t rand_index, size_t sk_off, size_t seed_off,
                                             size_t o_off) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint8_t *row = pool + i * pitch;
    uint8_t *sk = row + sk_off;
    if (gen_sk) {
        uint64_t ctr = rand_index + 2 * (uint64_t)i + 1;
        keccak::Shake<168> h;
        h.init();
        h.st[0] ^= ctr;  // LE64(ctr), the whole input
        h.finalize_static<8>();
        h.squeeze(sk, SK);
    }
    keccak::Shake<136> h;
    h.init();
    h.absorb_at<0, SK>(sk);
    h.finalize_static<SK>();
    // seed_pk (16 bytes) and O (O_BYTES) may go to different places
    uint8_t *seed = row + seed_off, *ob = row + o_off;
    for (int base = 0;; base += 136) {
#pragma unroll
        for (int w = 0; w < 17; ++w) {
#pragma unroll
            for (int b = 0; b < 8; ++b) {
                int off = base + 8 * w + b;
                uint8_t byte = (uint8_t)(h.st[w] >> (8 * b));
                if (off < 16) seed[off] = byte;
                else if (off < 16 + O_BYTES) ob[off - 16] = byte;
            }
        }
        if (base + 136 >= 16 + O_BYTES) break;
        keccak::f1600(h.st);
    }
}

// Exclusive prefix sum of the ORDER_BINS counters in pos (LDS), in place, by
// a 1024-lane workgroup: each of the first ORDER_BINS / 64 waves scans 64
// counters with lane shuffles, then adds the totals of the waves before it.
__device__ __forceinline__ void scan_bins(uint32_t *pos, uint32_t *wsum) {
    static_assert(ORDER_BINS % 64 == 0 && ORDER_BINS <= 1024, "bins");
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    uint32_t v = 0, incl = 0;
    if (t < ORDER_BINS) {
        v = pos[t];
        incl = v;
#pragma unroll
        for (int d = 1; d < 64; d <<= 1) {
            const uint32_t u = __shfl_up(incl, d, 64);
            if (lane >= d) incl += u;
        }
        if (lane == 63) wsum[w] = incl;
    }
    __syncthreads();
    if (t < ORDER_BINS) {
        uint32_t before = 0;
        for (int k = 0; k < w; ++k) before += wsum[k];
        pos[t] = before + incl - v;
    }
    __syncthreads();
}

// Hashing order for the lane-per-item hashing kernels below: items grouped
// by the SHAKE256 blocks their message absorbs, so that the lanes of a wave
// hash messages of similar length (a wave lasts as long as its longest
// message). Workgroup s sorts the segment [s ORDER_SEG, (s+1) ORDER_SEG) of
// the batch by counting sort into the same positions: a wave's 64 lanes only
// need neighbours of similar length, and the segments sort in parallel.
// Entry k goes to acc.set_order(k).
constexpr int ORDER_U = 4;                  // items per thread
constexpr size_t ORDER_SEG = 1024 * ORDER_U;
template <class Acc>
__global__ void __launch_bounds__(1024) k_length_order(Acc acc, size_t n) {
    __shared__ uint32_t pos[ORDER_BINS];
    __shared__ uint32_t wsum[ORDER_BINS / 64];
    const size_t base = (size_t)blockIdx.x * ORDER_SEG;
    const size_t end = base + ORDER_SEG < n ? base + ORDER_SEG : n;
    for (int b = threadIdx.x; b < ORDER_BINS; b += 1024) pos[b] = 0;
    __syncthreads();
    uint32_t bin[ORDER_U];
#pragma unroll
    for (int u = 0; u < ORDER_U; ++u) {
        const size_t i = base + threadIdx.x + 1024 * u;
        bin[u] = i < end ? acc.bin(i) : ORDER_BINS;
    }
#pragma unroll
    for (int u = 0; u < ORDER_U; ++u)
        if (bin[u] < ORDER_BINS) atomicAdd(&pos[bin[u]], 1u);
    __syncthreads();
    scan_bins(pos, wsum);
    uint32_t k[ORDER_U];
#pragma unroll
    for (int u = 0; u < ORDER_U; ++u) k[u] = bin[u] < ORDER_BINS ? atomicAdd(&pos[bin[u]], 1u) : 0u;
#pragma unroll
    for (int u = 0; u < ORDER_U; ++u)
        if (bin[u] < ORDER_BINS) acc.set_order(base + k[u], (uint32_t)(base + threadIdx.x + 1024 * u));
}

// XCD-aware order of the one-wave-per-item kernels (the slot of the hashing
// order is reused: the hashing pass has consumed it). The dispatcher deals
// workgroup b of a launch to XCD b mod XCDS, and each XCD has an L2 of its
// own. Items are sorted by key (counting sort over the top 8 bits of the
// key range) and the sorted list is cut into XCDS runs, run x going to the
// workgroups x, x + XCDS, x + 2 XCDS, ... in order: an XCD then works
// through its own few keys one after another instead of sharing every key
// with the other seven. Entry b goes to acc.set_order(b); the result is a
// permutation of [0, n).
constexpr int XCDS = 8;
template <class Acc>
__global__ void __launch_bounds__(1024) k_key_order(Acc acc, size_t n) {
    constexpr int U = 8;
    __shared__ uint32_t pos[ORDER_BINS];
    __shared__ uint32_t wsum[ORDER_BINS / 64];
    __shared__ uint32_t kmax;
    if (threadIdx.x == 0) kmax = 0;
    for (int b = threadIdx.x; b < ORDER_BINS; b += 1024) pos[b] = 0;
    __syncthreads();
    uint32_t km = 0;
    for (size_t i = threadIdx.x; i < n; i += 1024) km = max(km, acc.key(i));
    atomicMax(&kmax, km);
    __syncthreads();
    const int bits = kmax ? 32 - __builtin_clz(kmax) : 0;
    const int sh = bits > 8 ? bits - 8 : 0;  // key >> sh < ORDER_BINS
    for (size_t i0 = threadIdx.x; i0 < n; i0 += 1024 * U) {
        uint32_t bin[U];
#pragma unroll
        for (int u = 0; u < U; ++u) bin[u] = i0 + 1024 * u < n ? acc.key(i0 + 1024 * u) >> sh : ORDER_BINS;
#pragma unroll
        for (int u = 0; u < U; ++u)
            if (bin[u] < ORDER_BINS) atomicAdd(&pos[bin[u]], 1u);
    }
    __syncthreads();
    scan_bins(pos, wsum);
    // run x holds sorted positions [x q + min(x, r), ...) of q + (x < r) items
    const size_t q = n / XCDS, r = n % XCDS, big = r * (q + 1);
    for (size_t i0 = threadIdx.x; i0 < n; i0 += 1024 * U) {
        uint32_t bin[U], k[U];
#pragma unroll
        for (int u = 0; u < U; ++u) bin[u] = i0 + 1024 * u < n ? acc.key(i0 + 1024 * u) >> sh : ORDER_BINS;
#pragma unroll
        for (int u = 0; u < U; ++u) k[u] = bin[u] < ORDER_BINS ? atomicAdd(&pos[bin[u]], 1u) : 0u;
#pragma unroll
        for (int u = 0; u < U; ++u) {
            if (bin[u] >= ORDER_BINS) continue;
            size_t x, s;
            if (k[u] < big) {
                x = k[u] / (q + 1);
                s = k[u] - x * (q + 1);
            } else {
                x = r + (k[u] - big) / q;
                s = (k[u] - big) - (x - r) * q;
            }
            acc.set_order(XCDS * s + x, (uint32_t)(i0 + 1024 * u));
        }
    }
}

// digest, salt (R = 0), t and V || r for ctr = 0 (sign.cu:224-252). With
// ORDERED, thread k hashes item acc.order(k).
template <class Acc, bool ORDERED = false>
__global__ void __launch_bounds__(64) k_sign_hash(Acc acc, size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if constexpr (ORDERED) i = acc.order(i);
    SignItem it = acc(i);
    uint8_t *s = it.scratch;
    // the digest, salt and seed stay in registers between the four hashes
    // (each also goes to the workspace for the later kernels, but is never
    // read back from it here: a store-to-load round trip through memory per
    // hash, with one wave per SIMD, stalled the whole pass)
    static_assert(DIGEST % 8 == 0 && SK % 8 == 0, "whole words");
    constexpr int DW = DIGEST / 8, SW = SK / 8;
    uint64_t dig[DW], salt[SW], seed[SW];
#pragma unroll
    for (int k = 0; k < SW; ++k) seed[k] = keccak::Shake<136>::load64(it.esk + 8 * k);
    keccak::Shake<136> h;
    h.init();
    h.finalize_at(h.absorb_msg(it.msg, it.mlen));
    h.squeeze_words<DW>(dig);
    h.squeeze(s + SCR_DIGEST, DIGEST);
    // salt = SHAKE256(digest || R || seed_sk), R = 0 (absorbing zeros is a
    // no-op; no block ends inside R, so no permutation is skipped with it)
    static_assert((DIGEST + SK) <= 136, "R lies in the first block");
    h.init();
    h.absorb_words_at<0, DW>(dig);
    h.absorb_words_at<DIGEST + SK, SW>(seed);
    h.finalize_static<DIGEST + 2 * SK>();
    h.squeeze_words<SW>(salt);
    h.squeeze(s + SCR_SALT, SK);
    // t = SHAKE256(digest || salt)
    h.init();
    h.absorb_words_at<0, DW>(dig);
    h.absorb_words_at<DIGEST, SW>(salt);
    h.finalize_static<DIGEST + SK>();
    h.squeeze(s + SCR_T, MB);
    // V || r = SHAKE256(digest || salt || seed_sk || ctr), ctr = 0
    h.init();
    h.absorb_words_at<0, DW>(dig);
    h.absorb_words_at<DIGEST, SW>(salt);
    h.absorb_words_at<DIGEST + SK, SW>(seed);
    h.finalize_static<DIGEST + 2 * SK + 1>();
    h.squeeze(s + SCR_VR, VR_BYTES);
}

// digest and t for verification (verify.cu:37-48).
__device__ __forceinline__ void verify_target(const uint8_t *msg, uint32_t mlen,
                                              const uint8_t *sig, uint8_t *digest_tmp,
                                              uint8_t *t) {
    keccak::Shake<136> h;
    h.init();
    h.finalize_at(h.absorb_msg(msg, mlen));
    h.squeeze(digest_tmp, DIGEST);
    h.init();
    h.absorb_at<0, DIGEST>(digest_tmp);
    h.absorb_at<DIGEST, SK>(sig + SIG_S_BYTES);  // salt
    h.finalize_static<DIGEST + SK>();
    h.squeeze(t, MB);
}

template <class Acc, bool ORDERED = false>
__global__ void __launch_bounds__(64) k_verify_hash(Acc acc, size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if constexpr (ORDERED) i = acc.order(i);
    VerifyItem it = acc(i);
    uint8_t digest[DIGEST];
    verify_target(it.msg, it.mlen, it.sig, digest, it.t);
}

// ---------------------------------------------------------------------------
// AES-128-CTR expansion
// ---------------------------------------------------------------------------

__constant__ uint32_t c_aes_T0[256];
__constant__ uint32_t c_aes_S[256];

// out[0, P_BYTES) = keystream of key (aes128_ctr.cu:140-150 counter format)
// for n pool rows: row r's 16-byte key at key_off, its output at out_off.
// A workgroup stages the replicated round table once and takes groups of
// consecutive rows (grid-stride), KW waves per row (a power of two up to 16,
// chosen by the launch so that small batches still fill the GPU): the
// row's round keys are expanded into LDS, then read into SGPRs by the waves
// that encrypt that row's counter blocks, so they (and the first-round
// lookups of the key-only state words) stay out of the block loop. A row's
// key is read before any of that row's output is stored, so the two may
// overlap. The 64 KB table allows two workgroups per CU; at 1,024 threads
// each that is 8 waves per SIMD.
constexpr int AES_WG = 1024;
constexpr int AES_KEYS = AES_WG / 64;  // rows per group at one wave per row
template <int KW>
__global__ void __launch_bounds__(AES_WG, 8) k_aes_expand(uint8_t *pool, size_t pitch, size_t n,
                                                          size_t key_off, size_t out_off) {
    constexpr int kw = KW;
    __shared__ aes::Replicated tb;
    __shared__ uint32_t S[256];
    __shared__ uint32_t rk[AES_KEYS][44];
    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    constexpr int keys = AES_KEYS / kw;  // rows per group
    const int row = wave / kw, sub = wave % kw;
    for (int e = tid; e < 256 * aes::Replicated::COPIES; e += AES_WG) tb.T0[e] = c_aes_T0[e / aes::Replicated::COPIES];
    for (int e = tid; e < 256; e += AES_WG) S[e] = c_aes_S[e];
    __syncthreads();
    const size_t ngroups = (n + keys - 1) / keys;
    for (size_t g = blockIdx.x; g < ngroups; g += gridDim.x) {
        const size_t k0 = g * keys;
        const int nk = n - k0 < (size_t)keys ? (int)(n - k0) : keys;
        if (tid < nk) {
            const uint8_t *src = pool + (k0 + tid) * pitch + key_off;
            uint8_t key[16];
#pragma unroll
            for (int b = 0; b < 16; ++b) key[b] = src[b];
            aes::expand_key(key, rk[tid], S);
        }
        __syncthreads();
        if (row < nk) {
            uint32_t rks[44];
#pragma unroll
            for (int i = 0; i < 44; ++i) rks[i] = __builtin_amdgcn_readfirstlane(rk[row][i]);
            uint8_t *out = pool + (k0 + row) * pitch + out_off;
            const bool aligned = (((uintptr_t)out) & 3) == 0;
            static_assert(AES_BLOCKS <= 65536, "counter blocks below 2^16");
            const aes::CtrKey key(rks, c_aes_T0);
            for (int base = sub * 64; base < AES_BLOCKS; base += 64 * kw) {
                const aes::CtrStep step(key, c_aes_T0, (uint32_t)base);
                const int blk = base + lane;
                if (blk >= AES_BLOCKS) break;
                uint32_t o[4];
                aes::ctr_block_step(key, step, tb, (uint32_t)blk, o);
                uint8_t *p = out + 16 * (size_t)blk;
                int nb = P_BYTES - 16 * blk;
                if (nb >= 16 && aligned) {
                    uint32_t *q = (uint32_t *)p;
                    q[0] = __builtin_bswap32(o[0]); q[1] = __builtin_bswap32(o[1]);
                    q[2] = __builtin_bswap32(o[2]); q[3] = __builtin_bswap32(o[3]);
                } else {
                    if (nb > 16) nb = 16;
                    for (int b = 0; b < nb; ++b) p[b] = (uint8_t)(o[b >> 2] >> (24 - 8 * (b & 3)));
                }
            }
        }
        __syncthreads();  // rk is rewritten for the next group
    }
}
// Launch of k_aes_expand over n rows: waves per row kw = the power of two
// (1 .. 16) that brings n kw to about 8,192 waves (32 per CU), and at most
// 2,048 workgroups, each looping over its groups of rows.
inline void launch_aes_expand(uint8_t *pool, size_t pitch, size_t n, size_t key_off, size_t out_off,
                              hipStream_t stream) {
    int kw = 1;
    while (kw < AES_KEYS && n * (size_t)kw < 8192) kw *= 2;
    const size_t groups = (n + (size_t)(AES_KEYS / kw) - 1) / (size_t)(AES_KEYS / kw);
    const dim3 grid((unsigned)(groups < 2048 ? groups : 2048));
    switch (kw) {
        case 1: k_aes_expand<1><<<grid, AES_WG, 0, stream>>>(pool, pitch, n, key_off, out_off); break;
        case 2: k_aes_expand<2><<<grid, AES_WG, 0, stream>>>(pool, pitch, n, key_off, out_off); break;
        case 4: k_aes_expand<4><<<grid, AES_WG, 0, stream>>>(pool, pitch, n, key_off, out_off); break;
        case 8: k_aes_expand<8><<<grid, AES_WG, 0, stream>>>(pool, pitch, n, key_off, out_off); break;
        default: k_aes_expand<16><<<grid, AES_WG, 0, stream>>>(pool, pitch, n, key_off, out_off); break;
    }
}

// ---------------------------------------------------------------------------
// Quadratic-form evaluation U_ij = S_i^T P S_j + S_j^T P S_i
// ---------------------------------------------------------------------------

// S columns packed over j: scol[c] nibble j = S_j[c]. P is P1 (v x v upper)
// and, for verification, the P2 and P3 blocks of the full n x n form
// (sign.cu:396-437, verify.cu:72-182). Results XOR into Ulds[pair][w].
template <bool FULL>
__device__ __forceinline__ void quad_form(const uint8_t *P1, const uint8_t *P2, const uint8_t *P3,
                                          const uint64_t *scol, uint32_t *Ulds) {
    constexpr int NR = FULL ? N : V;
    constexpr int SLOTS = WG / MW;
    const int tid = threadIdx.x;
    if (tid >= SLOTS * MW) return;
    const int w = tid % MW, slot = tid / MW;
    const int nv = mvec::word_valid(w);
    uint32_t U[PAIRS];
#pragma unroll
    for (int p = 0; p < PAIRS; ++p) U[p] = 0;
    for (int r = slot; r < NR; r += SLOTS) {
        uint32_t T[K];
#pragma unroll
        for (int j = 0; j < K; ++j) T[j] = 0;
        if (r < V) {
            for (int c = r; c < V; ++c) {
                gf16::Multiples mp(mvec::load_word(P1 + (size_t)tri_entry(r, c, V) * MB + 4 * w, nv));
                const uint64_t sc = scol[c];
#pragma unroll
                for (int j = 0; j < K; ++j) T[j] ^= mp.times((uint32_t)(sc >> (4 * j)));
            }
            if (FULL) {
                for (int c = 0; c < O; ++c) {
                    gf16::Multiples mp(mvec::load_word(P2 + (size_t)(r * O + c) * MB + 4 * w, nv));
                    const uint64_t sc = scol[V + c];
#pragma unroll
                    for (int j = 0; j < K; ++j) T[j] ^= mp.times((uint32_t)(sc >> (4 * j)));
                }
            }
        } else {
            const int rr = r - V;
            for (int c = rr; c < O; ++c) {
                gf16::Multiples mp(mvec::load_word(P3 + (size_t)tri_entry(rr, c, O) * MB + 4 * w, nv));
                const uint64_t sc = scol[V + c];
#pragma unroll
                for (int j = 0; j < K; ++j) T[j] ^= mp.times((uint32_t)(sc >> (4 * j)));
            }
        }
        const uint64_t sr = scol[r];
#pragma unroll
        for (int j = 0; j < K; ++j) {
            gf16::Multiples mt(T[j]);
            // every ordered (a, j) adds S_a[r] T_j[r] to the pair {a, j}
#pragma unroll
            for (int a = 0; a < K; ++a) {
                const int p = a <= j ? pair_index(a, j) : pair_index(j, a);
                U[p] ^= mt.times((uint32_t)(sr >> (4 * a)));
            }
        }
    }
#pragma unroll
    for (int p = 0; p < PAIRS; ++p) atomicXor(&Ulds[p * MW + w], U[p]);
}

// y = sum E^l u_l by Horner over the pair order (sign.cu:440-455).
__device__ __forceinline__ void horner(const uint32_t *Ulds, uint32_t (&y)[MW]) {
#pragma unroll
    for (int w = 0; w < MW; ++w) y[w] = 0;
    for (int i = K - 1; i >= 0; --i)
        for (int j = i; j < K; ++j) {
            mvec::E_inplace(y);
            const uint32_t *u = Ulds + pair_index(i, j) * MW;
#pragma unroll
            for (int w = 0; w < MW; ++w) y[w] ^= u[w];
        }
}

// ---------------------------------------------------------------------------
// Signing: one workgroup per item
// ---------------------------------------------------------------------------

struct SignShared {
    uint8_t vr[VR_BYTES];
    uint64_t scol[V];
    uint32_t Mv[K * O * MW];   // M_i column x as an m-vector: Mv[(i*O + x)*MW + w]
    uint32_t U[PAIRS * MW];
    uint32_t A[M * RW];        // row-major, 8 nibbles per word
    uint32_t piv[4 * RW];      // multiples of the normalised pivot row
    uint32_t t[MW];
    uint8_t y[M];
    uint8_t f[M];
    uint8_t x[KO];
    int pivcol[M];
    int p, inv, ok;
};

__device__ __forceinline__ uint32_t a_nib(const SignShared &S, int row, int col) {
    return (S.A[row * RW + (col >> 3)] >> (4 * (col & 7))) & 0xFu;
}

// SampleSolution (sign.cu:10-145). x holds r on entry; returns 1 on full rank.
__device__ int sample_solution(SignShared &S) {
    const int tid = threadIdx.x;
    const uint8_t *rv = S.vr + R_OFF;
    for (int c = tid; c < KO; c += WG) S.x[c] = (uint8_t)mvec::nib(rv, c);
    // y <- y - A r (sign.cu:36-45)
    for (int a = tid; a < M; a += WG) {
        uint32_t acc = 0;
        for (int c = 0; c < KO; ++c) acc ^= gf16::mul(a_nib(S, a, c), mvec::nib(rv, c));
        S.y[a] ^= (uint8_t)acc;
    }
    __syncthreads();
    int piv_row = 0;
    for (int piv_col = 0; piv_col < KO && piv_row < M; ++piv_col) {
        if (tid == 0) S.p = M;
        __syncthreads();
        for (int a = piv_row + tid; a < M; a += WG)
            if (a_nib(S, a, piv_col)) atomicMin(&S.p, a);
        __syncthreads();
        const int p = S.p;
        if (p == M) continue;  // uniform
        if (p != piv_row) {
            for (int w = tid; w < RW; w += WG) {
                uint32_t tmp = S.A[piv_row * RW + w];
                S.A[piv_row * RW + w] = S.A[p * RW + w];
                S.A[p * RW + w] = tmp;
            }
            if (tid == 0) { uint8_t ty = S.y[piv_row]; S.y[piv_row] = S.y[p]; S.y[p] = ty; }
            __syncthreads();
        }
        if (tid == 0) {
            S.inv = (int)gf16::inv(a_nib(S, piv_row, piv_col));
            S.pivcol[piv_row] = piv_col;
        }
        __syncthreads();
        const uint32_t inv = (uint32_t)S.inv;
        for (int w = tid; w < RW; w += WG) {
            gf16::Multiples m0(S.A[piv_row * RW + w]);
            uint32_t v = m0.times(inv);
            S.A[piv_row * RW + w] = v;
            gf16::Multiples mv(v);
            S.piv[w] = mv.x0; S.piv[RW + w] = mv.x1; S.piv[2 * RW + w] = mv.x2; S.piv[3 * RW + w] = mv.x3;
        }
        if (tid == 0) S.y[piv_row] = (uint8_t)gf16::mul(S.y[piv_row], inv);
        __syncthreads();
        const uint32_t ypiv = S.y[piv_row];
        for (int a = piv_row + 1 + tid; a < M; a += WG) {
            uint32_t f = a_nib(S, a, piv_col);
            S.f[a] = (uint8_t)f;
            S.y[a] ^= (uint8_t)gf16::mul(f, ypiv);
        }
        __syncthreads();
        const int rows_below = M - piv_row - 1;
        for (int item = tid; item < rows_below * RW; item += WG) {
            const int a = piv_row + 1 + item / RW, w = item % RW;
            const uint32_t f = S.f[a];
            S.A[a * RW + w] ^= (S.piv[w] & gf16::bitmask(f, 0)) ^ (S.piv[RW + w] & gf16::bitmask(f, 1)) ^
                               (S.piv[2 * RW + w] & gf16::bitmask(f, 2)) ^
                               (S.piv[3 * RW + w] & gf16::bitmask(f, 3));
        }
        __syncthreads();
        ++piv_row;
    }
    if (piv_row < M) return 0;  // rank check (sign.cu:107-115)
    // back-substitution (sign.cu:117-142)
    for (int rr = M - 1; rr >= 0; --rr) {
        const int c = S.pivcol[rr];
        const uint32_t alpha = S.y[rr];
        __syncthreads();
        if (alpha) {
            if (tid == 0) S.x[c] ^= (uint8_t)alpha;
            for (int a = tid; a < M; a += WG) S.y[a] ^= (uint8_t)gf16::mul(alpha, a_nib(S, a, c));
        }
        __syncthreads();
    }
    return 1;
}

template <class Acc>
__global__ void __launch_bounds__(WG) k_sign_main(Acc acc, size_t n) {
    __shared__ SignShared S;
    const int tid = threadIdx.x;
    SignItem it = acc(blockIdx.x);
    const uint8_t *Ob = it.esk + SK;
    const uint8_t *P1 = Ob + O_BYTES;
    const uint8_t *L = P1 + P1_BYTES;
    for (int b = tid; b < VR_BYTES; b += WG) S.vr[b] = it.scratch[SCR_VR + b];
    for (int w = tid; w < MW; w += WG) S.t[w] = mvec::load_mw(it.scratch + SCR_T, w);
    for (int ctr = 0;; ++ctr) {
        __syncthreads();
        for (int c = tid; c < V; c += WG) {
            uint64_t v = 0;
            for (int j = 0; j < K; ++j) v |= (uint64_t)mvec::nib(S.vr + j * V_BYTES, c) << (4 * j);
            S.scol[c] = v;
        }
        for (int e = tid; e < K * O * MW; e += WG) S.Mv[e] = 0;
        for (int e = tid; e < PAIRS * MW; e += WG) S.U[e] = 0;
        for (int e = tid; e < M * RW; e += WG) S.A[e] = 0;
        __syncthreads();
        // M_i[.][x] = sum_r V_i[r] L[r][x] (sign.cu:256-330)
        {
            constexpr int ITEMS = O * MW;
            constexpr int RS = (WG / ITEMS) > 0 ? (WG / ITEMS) : 1;
            for (int item = tid; item < ITEMS * RS; item += WG) {
                const int q = item / ITEMS, rem = item % ITEMS, x = rem / MW, w = rem % MW;
                const int nv = mvec::word_valid(w);
                uint32_t a[K];
#pragma unroll
                for (int i = 0; i < K; ++i) a[i] = 0;
                for (int r = q; r < V; r += RS) {
                    gf16::Multiples ml(mvec::load_word(L + (size_t)(r * O + x) * MB + 4 * w, nv));
                    const uint64_t sc = S.scol[r];
#pragma unroll
                    for (int i = 0; i < K; ++i) a[i] ^= ml.times((uint32_t)(sc >> (4 * i)));
                }
#pragma unroll
                for (int i = 0; i < K; ++i) atomicXor(&S.Mv[(i * O + x) * MW + w], a[i]);
            }
        }
        quad_form<false>(P1, nullptr, nullptr, S.scol, S.U);
        __syncthreads();
        // A (sign.cu:333-393): column (b, x) = sum of E^l M_other[x]
        for (int item = tid; item < K * O; item += WG) {
            const int b = item / O, x = item % O;
            uint32_t col[MW];
#pragma unroll
            for (int w = 0; w < MW; ++w) col[w] = 0;
            int ell = 0;
            for (int i = 0; i < K; ++i)
                for (int j = K - 1; j >= i; --j, ++ell) {
                    if (i == b) mvec::add_E_pow(col, &S.Mv[(j * O + x) * MW], ell);
                    if (i != j && j == b) mvec::add_E_pow(col, &S.Mv[(i * O + x) * MW], ell);
                }
            const int cidx = b * O + x;
#pragma unroll
            for (int w = 0; w < MW; ++w)
#pragma unroll
                for (int e = 0; e < 8; ++e) {
                    const int a = 8 * w + e;
                    if (a < M) {
                        uint32_t v = (col[w] >> (4 * e)) & 0xFu;
                        if (v) atomicXor(&S.A[a * RW + (cidx >> 3)], v << (4 * (cidx & 7)));
                    }
                }
        }
        if (tid == 0) {
            uint32_t y[MW];
            horner(S.U, y);
#pragma unroll
            for (int w = 0; w < MW; ++w) y[w] ^= S.t[w];
#pragma unroll
            for (int w = 0; w < MW; ++w)
#pragma unroll
                for (int e = 0; e < 8; ++e)
                    if (8 * w + e < M) S.y[8 * w + e] = (uint8_t)((y[w] >> (4 * e)) & 0xF);
        }
        __syncthreads();
        const int ok = sample_solution(S);
        if (ok || ctr == MAX_ATTEMPTS - 1) {
            if (tid == 0) it.scratch[SCR_ATTEMPTS] = ok ? (uint8_t)(ctr + 1) : 0;
            break;
        }
        // rare path: next attempt's V || r (sign.cu:251-252)
        __syncthreads();
        if (tid == 0) {
            keccak::Shake<136> h;
            h.init();
            h.absorb(it.scratch + SCR_DIGEST, DIGEST + SK);
            h.absorb(it.esk, SK);
            h.absorb_byte((uint32_t)(ctr + 1));
            h.finalize();
            h.squeeze(S.vr, VR_BYTES);
        }
    }
    __syncthreads();
    // s_i = (V_i + O x_i || x_i) and the salt (sign.cu:147-191, 472-477)
    for (int b = tid; b < SIG_S_BYTES; b += WG) {
        uint32_t out = 0;
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const int e = 2 * b + h, i = e / N, t = e % N;
            uint32_t val;
            if (t < V) {
                uint32_t a = 0;
#pragma unroll
                for (int kk = 0; kk < O; ++kk) a ^= gf16::mul(mvec::nib(Ob, t * O + kk), S.x[i * O + kk]);
                val = mvec::nib(S.vr + i * V_BYTES, t) ^ a;
            } else {
                val = S.x[i * O + (t - V)];
            }
            out |= val << (4 * h);
        }
        it.sig[b] = (uint8_t)out;
    }
    for (int b = tid; b < SK; b += WG) it.sig[SIG_S_BYTES + b] = it.scratch[SCR_SALT + b];
}

// ---------------------------------------------------------------------------
// Verification: one workgroup per item
// ---------------------------------------------------------------------------

template <class Acc>
__global__ void __launch_bounds__(WG) k_verify_main(Acc acc, size_t n) {
    __shared__ uint64_t scol[N];
    __shared__ uint32_t U[PAIRS * MW];
    __shared__ uint8_t tb[MB + 4];
    __shared__ uint8_t digest[DIGEST];
    const int tid = threadIdx.x;
    VerifyItem it = acc(blockIdx.x);
    const uint8_t *P1 = it.epk, *P2 = P1 + P1_BYTES, *P3 = P2 + P2_BYTES;
    for (int c = tid; c < N; c += WG) {
        uint64_t v = 0;
        for (int j = 0; j < K; ++j) v |= (uint64_t)mvec::nib(it.sig, j * N + c) << (4 * j);
        scol[c] = v;
    }
    for (int e = tid; e < PAIRS * MW; e += WG) U[e] = 0;
    __syncthreads();
    if (it.t) {
        for (int b = tid; b < MB; b += WG) tb[b] = it.t[b];
    } else if (tid == WG - 1) {
        // the last lane is idle in quad_form when WG % MW != 0; otherwise
        // it simply hashes first
        verify_target(it.msg, it.mlen, it.sig, digest, tb);
    }
    quad_form<true>(P1, P2, P3, scol, U);
    __syncthreads();
    if (tid == 0) {
        uint32_t y[MW];
        horner(U, y);
        bool eq = true;
#pragma unroll
        for (int w = 0; w < MW; ++w) {
            const int nv = mvec::word_valid(w);
            uint32_t tw = 0;
            for (int b = 0; b < nv; ++b) tw |= (uint32_t)tb[4 * w + b] << (8 * b);
            eq &= (y[w] == tw);
        }
        it.result[0] = eq ? 0 : 1;  // verify.cu:209-211
    }
}

}  // namespace mayo
