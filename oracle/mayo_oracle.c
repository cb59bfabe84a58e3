/*
 * mayo_oracle.c -- CPU restatement of the cuMAYO algorithm (TEST INFRASTRUCTURE).
 *
 * This file is the checker for the HIP engine. Only tests/, the smoke check
 * and the CPU-baseline leg of the benchmark load it; the product never does.
 *
 * It restates, stage by stage, what the reference kernels compute. Cited
 * files are relative to the cuMAYO source tree:
 *   params.h          parameter sets
 *   code.h            nibble packing and m-vector indexing
 *   gf16.cuh/gf16.cu  GF(16) multiply / inverse (semantics only)
 *   poly.cuh          multiplication by z modulo f(z)
 *   fips202.cu/.cuh   SHAKE interface (this file has its own FIPS 202 Keccak)
 *   aes128_ctr.cu     AES-128-CTR counter format (own FIPS 197 AES here)
 *   keypair.cu, expandsk.cu, expandpk.cu, sign.cu, verify.cu
 *
 * Arithmetic style: m-vectors stay packed two nibbles per byte (equation 2i
 * in the low nibble of byte i, code.h:7-11), and every multiply-accumulate
 * acts on both nibbles of a byte, as the reference's byte lanes do
 * (gf16.cuh:30-41). The per-byte product table is computed from the field
 * definition at start-up, not copied.
 */
#include "mayo_oracle.h"

#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ------------------------------------------------------------------------ */
/* Parameter sets (params.h:10-96). Derived sizes are computed, not typed.   */
/* ------------------------------------------------------------------------ */

static mayo_params g_params[4];
static int g_params_ready = 0;

static void fill_params(mayo_params *p, int ver, int n, int m, int o, int k, int skb,
                        int digest, uint8_t c0, uint8_t c1, uint8_t c2, uint8_t c3) {
    p->version = ver;
    p->n = n; p->m = m; p->o = o; p->k = k; p->v = n - o;
    p->sk_bytes = skb;
    p->digest_bytes = digest;
    p->m_bytes = m / 2;
    p->o_bytes = (p->v * o + 1) / 2;
    p->p1_bytes = p->v * (p->v + 1) / 2 * p->m_bytes;
    p->p2_bytes = p->v * o * p->m_bytes;
    p->p3_bytes = o * (o + 1) / 2 * p->m_bytes;
    p->p_bytes = p->p1_bytes + p->p2_bytes;
    p->pk_bytes = 16 + p->p3_bytes;
    p->pairs = k * (k + 1) / 2;
    p->sig_bytes = k * n / 2 + skb;
    p->esk_bytes = skb + p->o_bytes + p->p_bytes;
    p->epk_bytes = p->p_bytes + p->p3_bytes;
    p->vr_bytes = k * n / 2;
    p->f_tail[0] = c0; p->f_tail[1] = c1; p->f_tail[2] = c2; p->f_tail[3] = c3;
}

static void init_params(void) {
    if (g_params_ready) return;
    /* params.h:13-29, 34-50, 55-71, 76-92 */
    fill_params(&g_params[0], 1, 86, 78, 8, 10, 24, 32, 0x8, 0x1, 0x1, 0x0);
    fill_params(&g_params[1], 2, 81, 64, 17, 4, 24, 32, 0x8, 0x0, 0x2, 0x8);
    fill_params(&g_params[2], 3, 118, 108, 10, 11, 32, 48, 0x8, 0x0, 0x1, 0x7);
    fill_params(&g_params[3], 5, 154, 142, 12, 12, 40, 64, 0x4, 0x0, 0x8, 0x1);
    g_params_ready = 1;
}

const mayo_params *mayo_oracle_params(int version) {
    init_params();
    switch (version) {
    case 1: return &g_params[0];
    case 2: return &g_params[1];
    case 3: return &g_params[2];
    case 5: return &g_params[3];
    default: return NULL;
    }
}

/* ------------------------------------------------------------------------ */
/* Keccak-f[1600] and SHAKE (FIPS 202). Round constants and rotation         */
/* offsets are generated from their definitions (rc LFSR, (x,y) walk).      */
/* ------------------------------------------------------------------------ */

static uint64_t g_rc[24];
static int g_rho[25];
static int g_keccak_ready = 0;

static int rc_bit(int t) {
    /* FIPS 202 Algorithm 5: LFSR over x^8 + x^6 + x^5 + x^4 + 1 */
    if (t % 255 == 0) return 1;
    uint8_t R = 0x01;
    for (int i = 1; i <= t % 255; i++) {
        uint8_t r8 = (R & 0x80) ? 1 : 0;
        R = (uint8_t)(R << 1);
        if (r8) R ^= 0x71; /* bits 0,4,5,6 */
    }
    return R & 1;
}

static void keccak_init(void) {
    if (g_keccak_ready) return;
    for (int ir = 0; ir < 24; ir++) {
        uint64_t rc = 0;
        for (int j = 0; j <= 6; j++)
            if (rc_bit(j + 7 * ir)) rc |= (uint64_t)1 << ((1 << j) - 1);
        g_rc[ir] = rc;
    }
    g_rho[0] = 0;
    int x = 1, y = 0;
    for (int t = 0; t < 24; t++) {
        g_rho[x + 5 * y] = ((t + 1) * (t + 2) / 2) % 64;
        int nx = y, ny = (2 * x + 3 * y) % 5;
        x = nx; y = ny;
    }
    g_keccak_ready = 1;
}

static inline uint64_t rotl64(uint64_t a, int s) {
    return s == 0 ? a : (a << s) | (a >> (64 - s));
}

void oracle_keccak_f1600(uint64_t A[25]) {
    keccak_init();
    for (int round = 0; round < 24; round++) {
        uint64_t C[5], D[5], B[25];
        for (int x = 0; x < 5; x++)
            C[x] = A[x] ^ A[x + 5] ^ A[x + 10] ^ A[x + 15] ^ A[x + 20];
        for (int x = 0; x < 5; x++)
            D[x] = C[(x + 4) % 5] ^ rotl64(C[(x + 1) % 5], 1);
        for (int i = 0; i < 25; i++) A[i] ^= D[i % 5];
        /* rho + pi: B[y, 2x+3y] = rot(A[x,y]) */
        for (int x = 0; x < 5; x++)
            for (int y = 0; y < 5; y++)
                B[y + 5 * ((2 * x + 3 * y) % 5)] = rotl64(A[x + 5 * y], g_rho[x + 5 * y]);
        for (int x = 0; x < 5; x++)
            for (int y = 0; y < 5; y++)
                A[x + 5 * y] = B[x + 5 * y] ^ (~B[(x + 1) % 5 + 5 * y] & B[(x + 2) % 5 + 5 * y]);
        A[0] ^= g_rc[round];
    }
}

static void shake(int rate, uint8_t *out, size_t outlen, const uint8_t *in, size_t inlen) {
    uint64_t st[25];
    memset(st, 0, sizeof st);
    uint8_t *sb = (uint8_t *)st; /* little-endian host assumed (x86-64) */
    while (inlen >= (size_t)rate) {
        for (int i = 0; i < rate; i++) sb[i] ^= in[i];
        oracle_keccak_f1600(st);
        in += rate; inlen -= rate;
    }
    for (size_t i = 0; i < inlen; i++) sb[i] ^= in[i];
    sb[inlen] ^= 0x1F;
    sb[rate - 1] ^= 0x80;
    for (;;) {
        oracle_keccak_f1600(st);
        size_t take = outlen < (size_t)rate ? outlen : (size_t)rate;
        memcpy(out, sb, take);
        out += take; outlen -= take;
        if (!outlen) break;
    }
}

void oracle_shake128(uint8_t *out, size_t outlen, const uint8_t *in, size_t inlen) {
    shake(168, out, outlen, in, inlen);
}

void oracle_shake256(uint8_t *out, size_t outlen, const uint8_t *in, size_t inlen) {
    shake(136, out, outlen, in, inlen);
}

/* ------------------------------------------------------------------------ */
/* AES-128 (FIPS 197), byte oriented; S-box derived from the GF(2^8)        */
/* inverse and the affine map. CTR keystream as aes128_ctr.cu:140-150:      */
/* block b encrypts 0^64 || BE64(b), output bytes in order, tail truncated. */
/* ------------------------------------------------------------------------ */

static uint8_t g_sbox[256];
static int g_aes_ready = 0;

static uint8_t gf256_mul(uint8_t a, uint8_t b) {
    uint8_t r = 0;
    while (b) {
        if (b & 1) r ^= a;
        uint8_t hi = a & 0x80;
        a = (uint8_t)(a << 1);
        if (hi) a ^= 0x1B;
        b >>= 1;
    }
    return r;
}

static void aes_init(void) {
    if (g_aes_ready) return;
    for (int a = 0; a < 256; a++) {
        uint8_t inv = 0;
        if (a) {
            for (int b = 1; b < 256; b++)
                if (gf256_mul((uint8_t)a, (uint8_t)b) == 1) { inv = (uint8_t)b; break; }
        }
        uint8_t s = inv, x = inv;
        for (int i = 0; i < 4; i++) {
            x = (uint8_t)((x << 1) | (x >> 7));
            s ^= x;
        }
        g_sbox[a] = s ^ 0x63;
    }
    g_aes_ready = 1;
}

static void aes128_key_schedule(const uint8_t key[16], uint8_t rk[176]) {
    aes_init();
    memcpy(rk, key, 16);
    uint8_t rcon = 1;
    for (int i = 4; i < 44; i++) {
        uint8_t t[4];
        memcpy(t, rk + 4 * (i - 1), 4);
        if (i % 4 == 0) {
            uint8_t t0 = t[0];
            t[0] = g_sbox[t[1]] ^ rcon;
            t[1] = g_sbox[t[2]];
            t[2] = g_sbox[t[3]];
            t[3] = g_sbox[t0];
            rcon = gf256_mul(rcon, 2);
        }
        for (int j = 0; j < 4; j++) rk[4 * i + j] = rk[4 * (i - 4) + j] ^ t[j];
    }
}

static void aes128_encrypt_rk(const uint8_t rk[176], const uint8_t in[16], uint8_t out[16]) {
    uint8_t s[16], t[16];
    for (int i = 0; i < 16; i++) s[i] = in[i] ^ rk[i];
    for (int round = 1; round <= 10; round++) {
        for (int i = 0; i < 16; i++) s[i] = g_sbox[s[i]];
        /* ShiftRows: byte (row r, col c) at index r + 4c takes column c + r */
        for (int c = 0; c < 4; c++)
            for (int r = 0; r < 4; r++) t[r + 4 * c] = s[r + 4 * ((c + r) % 4)];
        if (round != 10) {
            for (int c = 0; c < 4; c++) {
                uint8_t a0 = t[4 * c], a1 = t[4 * c + 1], a2 = t[4 * c + 2], a3 = t[4 * c + 3];
                s[4 * c + 0] = gf256_mul(a0, 2) ^ gf256_mul(a1, 3) ^ a2 ^ a3;
                s[4 * c + 1] = a0 ^ gf256_mul(a1, 2) ^ gf256_mul(a2, 3) ^ a3;
                s[4 * c + 2] = a0 ^ a1 ^ gf256_mul(a2, 2) ^ gf256_mul(a3, 3);
                s[4 * c + 3] = gf256_mul(a0, 3) ^ a1 ^ a2 ^ gf256_mul(a3, 2);
            }
        } else {
            memcpy(s, t, 16);
        }
        for (int i = 0; i < 16; i++) s[i] ^= rk[16 * round + i];
    }
    memcpy(out, s, 16);
}

void oracle_aes128_ecb_block(const uint8_t key[16], const uint8_t in[16], uint8_t out[16]) {
    uint8_t rk[176];
    aes128_key_schedule(key, rk);
    aes128_encrypt_rk(rk, in, out);
}

void oracle_aes128_ctr(const uint8_t key[16], uint8_t *out, size_t outlen) {
    uint8_t rk[176], ctr[16], ks[16];
    aes128_key_schedule(key, rk);
    uint64_t blk = 0;
    while (outlen) {
        memset(ctr, 0, 8);
        for (int i = 0; i < 8; i++) ctr[8 + i] = (uint8_t)(blk >> (56 - 8 * i));
        aes128_encrypt_rk(rk, ctr, ks);
        size_t take = outlen < 16 ? outlen : 16;
        memcpy(out, ks, take);
        out += take; outlen -= take; blk++;
    }
}

/* ------------------------------------------------------------------------ */
/* GF(16) arithmetic (gf16.cuh:21-41, inverse table gf16.cu:3-8).          */
/* ------------------------------------------------------------------------ */

uint8_t oracle_gf16_mul(uint8_t a, uint8_t b) {
    a &= 0xF; b &= 0xF;
    uint8_t r = 0;
    for (int i = 0; i < 4; i++) {
        if (b & 1) r ^= a;
        b >>= 1;
        uint8_t hi = a & 0x8;
        a = (uint8_t)((a << 1) & 0xF);
        if (hi) a ^= 0x3; /* x^4 = x + 1 */
    }
    return r;
}

uint8_t oracle_gf16_inv(uint8_t a) {
    a &= 0xF;
    if (!a) return 0; /* the reference table maps 0 to 0 */
    uint8_t r = 1;
    for (int i = 0; i < 14; i++) r = oracle_gf16_mul(r, a); /* a^14 = a^-1 */
    return r;
}

/* mul2[b][byte] = both nibbles of byte times b (gf16_mul2nibble). */
static uint8_t g_mul2[16][256];
static int g_gf_ready = 0;

static void gf_init(void) {
    if (g_gf_ready) return;
    for (int b = 0; b < 16; b++)
        for (int x = 0; x < 256; x++)
            g_mul2[b][x] = (uint8_t)(oracle_gf16_mul((uint8_t)(x & 0xF), (uint8_t)b) |
                                     (oracle_gf16_mul((uint8_t)(x >> 4), (uint8_t)b) << 4));
    g_gf_ready = 1;
}

static void oracle_init_all(void) {
    init_params();
    keccak_init();
    aes_init();
    gf_init();
}

/* acc[0..len) ^= b * src[0..len), both nibbles of every byte. */
static inline void mvec_mac(uint8_t *acc, const uint8_t *src, uint8_t b, int len) {
    if (!(b &= 0xF)) return;
    const uint8_t *tab = g_mul2[b];
    for (int i = 0; i < len; i++) acc[i] ^= tab[src[i]];
}

static inline void mvec_xor(uint8_t *acc, const uint8_t *src, int len) {
    for (int i = 0; i < len; i++) acc[i] ^= src[i];
}

/* Nibble e of a packed array, low nibble first (code.h:7-11). */
static inline uint8_t nib(const uint8_t *p, int e) {
    return (uint8_t)((p[e >> 1] >> ((e & 1) * 4)) & 0xF);
}

static inline void set_nib(uint8_t *p, int e, uint8_t val) {
    int sh = (e & 1) * 4;
    p[e >> 1] = (uint8_t)((p[e >> 1] & ~(0xF << sh)) | ((val & 0xF) << sh));
}

/* Upper-triangular entry (r, c), c >= r, of a dim x dim matrix of m-vectors
 * (code.h:15-21, idx_mattri with the entry counted from row r). */
static inline size_t tri_entry(int r, int c, int dim) {
    return (size_t)(c - r) + (size_t)dim * r - (size_t)r * (r - 1) / 2;
}

/* Row-major entry (r, c) of a rows x cols matrix of m-vectors (code.h:23-28). */
static inline size_t rect_entry(int r, int c, int cols) {
    return (size_t)r * cols + c;
}

/* Pair index of (i, j), i <= j (code.h:36-39). */
static inline int pair_index(int i, int j, int k) {
    return i * k - i * (i - 1) / 2 + (j - i);
}

/* ------------------------------------------------------------------------ */
/* E: y <- z*y mod f(z) (poly.cuh:11-69).                                   */
/* ------------------------------------------------------------------------ */

static void mvec_E(const mayo_params *p, uint8_t *y) {
    const int mb = p->m_bytes;
    uint8_t last = (uint8_t)(y[mb - 1] >> 4); /* coefficient of z^(m-1) */
    /* shift every nibble up by one position */
    for (int i = mb - 1; i >= 0; i--) {
        uint8_t prevhi = i > 0 ? (uint8_t)(y[i - 1] >> 4) : 0;
        y[i] = (uint8_t)(((y[i] & 0xF) << 4) | prevhi);
    }
    /* z^m = C0 + C1 z + C2 z^2 + C3 z^3 (characteristic 2) */
    y[0] ^= (uint8_t)(oracle_gf16_mul(last, p->f_tail[0]) | (oracle_gf16_mul(last, p->f_tail[1]) << 4));
    y[1] ^= (uint8_t)(oracle_gf16_mul(last, p->f_tail[2]) | (oracle_gf16_mul(last, p->f_tail[3]) << 4));
}

void oracle_mvec_E(int version, uint8_t *y_packed) {
    oracle_init_all();
    mvec_E(mayo_oracle_params(version), y_packed);
}

/* ------------------------------------------------------------------------ */
/* Key generation (keypair.cu:10-120).                                      */
/* ------------------------------------------------------------------------ */

void oracle_keypair_seed(int version, uint64_t rand_index, uint64_t item, uint8_t *sk) {
    oracle_init_all();
    const mayo_params *p = mayo_oracle_params(version);
    /* keypair.cu:32-45: seed = LE64(rand_index + 2*blockIdx + 1), sk = SHAKE128(seed) */
    uint64_t ctr = rand_index + 2 * item + 1;
    uint8_t buf[8];
    for (int i = 0; i < 8; i++) buf[i] = (uint8_t)(ctr >> (8 * i));
    oracle_shake128(sk, (size_t)p->sk_bytes, buf, 8);
}

/* S = SHAKE256(sk)[0 : 16 + O_bytes]: seed_pk || O (keypair.cu:47,
 * expandsk.cu:29). */
static void derive_S(const mayo_params *p, const uint8_t *sk, uint8_t *S) {
    oracle_shake256(S, (size_t)(16 + p->o_bytes), sk, (size_t)p->sk_bytes);
}

/* B[r][b] = P2[r][b] + sum_{c>=r} P1[r][c] * O[c][b] (keypair.cu:64-86). */
static void keypair_P3(const mayo_params *p, const uint8_t *O, const uint8_t *P1,
                       const uint8_t *P2, uint8_t *P3) {
    const int v = p->v, o = p->o, mb = p->m_bytes;
    uint8_t *B = (uint8_t *)calloc((size_t)v * o * mb, 1);
    for (int r = 0; r < v; r++) {
        for (int b = 0; b < o; b++) {
            uint8_t *acc = B + rect_entry(r, b, o) * mb;
            for (int c = r; c < v; c++)
                mvec_mac(acc, P1 + tri_entry(r, c, v) * mb, nib(O, c * o + b), mb);
            mvec_xor(acc, P2 + rect_entry(r, b, o) * mb, mb);
        }
    }
    /* P3[a][b] = sum_r O[r][a]*B[r][b] + (a != b) O[r][b]*B[r][a]
     * (keypair.cu:89-118) */
    memset(P3, 0, (size_t)p->p3_bytes);
    for (int a = 0; a < o; a++) {
        for (int b = a; b < o; b++) {
            uint8_t *acc = P3 + tri_entry(a, b, o) * mb;
            for (int r = 0; r < v; r++) {
                mvec_mac(acc, B + rect_entry(r, b, o) * mb, nib(O, r * o + a), mb);
                if (a != b) mvec_mac(acc, B + rect_entry(r, a, o) * mb, nib(O, r * o + b), mb);
            }
        }
    }
    free(B);
}

void oracle_keypair_from_sk(int version, const uint8_t *sk, uint8_t *pk) {
    oracle_init_all();
    const mayo_params *p = mayo_oracle_params(version);
    uint8_t *S = (uint8_t *)malloc((size_t)(16 + p->o_bytes));
    uint8_t *P = (uint8_t *)malloc((size_t)p->p_bytes);
    derive_S(p, sk, S);
    oracle_aes128_ctr(S, P, (size_t)p->p_bytes); /* keypair.cu:52 */
    memcpy(pk, S, 16);                           /* keypair.cu:49-50 */
    keypair_P3(p, S + 16, P, P + p->p1_bytes, pk + 16);
    free(S); free(P);
}

void oracle_keypair(int version, uint64_t rand_index, uint64_t item, uint8_t *pk, uint8_t *sk) {
    oracle_keypair_seed(version, rand_index, item, sk);
    oracle_keypair_from_sk(version, sk, pk);
}

/* ------------------------------------------------------------------------ */
/* Key expansion (expandsk.cu:9-81, expandpk.cu:9-24).                      */
/* ------------------------------------------------------------------------ */

void oracle_expandsk(int version, const uint8_t *sk, uint8_t *esk) {
    oracle_init_all();
    const mayo_params *p = mayo_oracle_params(version);
    const int v = p->v, o = p->o, mb = p->m_bytes;
    uint8_t *S = (uint8_t *)malloc((size_t)(16 + p->o_bytes));
    derive_S(p, sk, S);
    uint8_t *O = esk + p->sk_bytes;
    uint8_t *P1 = O + p->o_bytes;
    uint8_t *L = P1 + p->p1_bytes;
    memcpy(esk, sk, (size_t)p->sk_bytes);
    memcpy(O, S + 16, (size_t)p->o_bytes);
    oracle_aes128_ctr(S, P1, (size_t)p->p_bytes); /* P1 || P2, P2 becomes L */
    /* L[r][b] = P2[r][b] + sum_{c>r} P1[r][c] O[c][b] + sum_{c<r} P1[c][r] O[c][b]
     * (expandsk.cu:37-80) */
    for (int r = 0; r < v; r++) {
        for (int b = 0; b < o; b++) {
            uint8_t *acc = L + rect_entry(r, b, o) * mb;
            for (int c = r + 1; c < v; c++)
                mvec_mac(acc, P1 + tri_entry(r, c, v) * mb, nib(O, c * o + b), mb);
            for (int c = 0; c < r; c++)
                mvec_mac(acc, P1 + tri_entry(c, r, v) * mb, nib(O, c * o + b), mb);
        }
    }
    free(S);
}

void oracle_expandpk(int version, const uint8_t *cpk, uint8_t *epk) {
    oracle_init_all();
    const mayo_params *p = mayo_oracle_params(version);
    /* epk = AES-CTR(seed_pk) || P3 (expandpk.cu:17-23; api.cu:64-85) */
    oracle_aes128_ctr(cpk, epk, (size_t)p->p_bytes);
    memcpy(epk + p->p_bytes, cpk + 16, (size_t)p->p3_bytes);
}

/* ------------------------------------------------------------------------ */
/* Signing (sign.cu:10-478).                                                */
/* ------------------------------------------------------------------------ */

void oracle_sign_digest_salt_t(int version, const uint8_t *esk, const uint8_t *msg, size_t mlen,
                               uint8_t *digest, uint8_t *salt, uint8_t *t) {
    oracle_init_all();
    const mayo_params *p = mayo_oracle_params(version);
    const int D = p->digest_bytes, SK = p->sk_bytes;
    uint8_t buf[64 + 40 + 40 + 1];
    oracle_shake256(digest, (size_t)D, msg, mlen);      /* sign.cu:225 */
    memcpy(buf, digest, (size_t)D);
    memset(buf + D, 0, (size_t)SK);                     /* R = 0, sign.cu:230-231 */
    memcpy(buf + D + SK, esk, (size_t)SK);              /* seed_sk, sign.cu:233-234 */
    oracle_shake256(salt, (size_t)SK, buf, (size_t)(D + 2 * SK)); /* sign.cu:238 */
    memcpy(buf + D, salt, (size_t)SK);                  /* sign.cu:240-241 */
    oracle_shake256(t, (size_t)p->m_bytes, buf, (size_t)(D + SK)); /* sign.cu:243 */
}

void oracle_sign_vr(int version, const uint8_t *digest, const uint8_t *salt,
                    const uint8_t *seed_sk, int ctr, uint8_t *vr) {
    oracle_init_all();
    const mayo_params *p = mayo_oracle_params(version);
    const int D = p->digest_bytes, SK = p->sk_bytes;
    uint8_t buf[64 + 40 + 40 + 1];
    memcpy(buf, digest, (size_t)D);
    memcpy(buf + D, salt, (size_t)SK);
    memcpy(buf + D + SK, seed_sk, (size_t)SK);
    buf[D + 2 * SK] = (uint8_t)ctr;                     /* sign.cu:251-252 */
    oracle_shake256(vr, (size_t)p->vr_bytes, buf, (size_t)(D + 2 * SK + 1));
}

/* V_i occupies v/2 bytes at offset i*v/2 of vr; r follows at k*v/2
 * (sign.cu:250, with v == m). */
static inline const uint8_t *V_of(const mayo_params *p, const uint8_t *vr, int i) {
    return vr + (size_t)i * (p->v / 2);
}

/* M_i[a][col] = sum_r V_i[r] * L[r][col][a] (sign.cu:256-330). */
void oracle_sign_M(int version, const uint8_t *esk, const uint8_t *vr, uint8_t *M) {
    oracle_init_all();
    const mayo_params *p = mayo_oracle_params(version);
    const int v = p->v, o = p->o, m = p->m, k = p->k, mb = p->m_bytes;
    const uint8_t *L = esk + p->sk_bytes + p->o_bytes + p->p1_bytes;
    uint8_t *col = (uint8_t *)malloc((size_t)mb);
    for (int i = 0; i < k; i++) {
        const uint8_t *Vi = V_of(p, vr, i);
        for (int c = 0; c < o; c++) {
            memset(col, 0, (size_t)mb);
            for (int r = 0; r < v; r++)
                mvec_mac(col, L + rect_entry(r, c, o) * mb, nib(Vi, r), mb);
            for (int a = 0; a < m; a++)
                M[((size_t)i * m + a) * o + c] = nib(col, a);
        }
    }
    free(col);
}

/* A = sum over the (i, j) schedule of E^l M_j into block i and E^l M_i into
 * block j (sign.cu:333-393). E^l is applied as the reference does: rows
 * shift down by l, and the l overflowing rows fold back through C0..C3 into
 * rows 0..l+2 (valid because l + 3 < m for every parameter set). */
void oracle_sign_A(int version, const uint8_t *M, uint8_t *A) {
    oracle_init_all();
    const mayo_params *p = mayo_oracle_params(version);
    const int m = p->m, o = p->o, k = p->k, KO = k * o;
    memset(A, 0, (size_t)m * KO);
    int ell = 0;
    for (int i = 0; i < k; i++) {
        for (int j = k - 1; j >= i; j--) {
            for (int pass = 0; pass < (i == j ? 1 : 2); pass++) {
                const int dst_blk = pass == 0 ? i : j;
                const uint8_t *src = M + (size_t)(pass == 0 ? j : i) * m * o;
                for (int c = 0; c < o; c++) {
                    /* shifted part: row a + ell gets source row a */
                    for (int a = 0; a + ell < m; a++)
                        A[(size_t)(a + ell) * KO + dst_blk * o + c] ^= src[(size_t)a * o + c];
                    /* overflow part: source row m - ell + t, t < ell, is z^(m+t) */
                    for (int t = 0; t < ell; t++) {
                        uint8_t s = src[(size_t)(m - ell + t) * o + c];
                        for (int nn = 0; nn < 4; nn++)
                            A[(size_t)(t + nn) * KO + dst_blk * o + c] ^=
                                oracle_gf16_mul(s, p->f_tail[nn]);
                    }
                }
            }
            ell++;
        }
    }
}

/* u_ij = V_i^T P1 V_j + V_j^T P1 V_i (i != j), V_i^T P1 V_i (i == j), with
 * P1 upper triangular (sign.cu:396-437). */
void oracle_sign_U(int version, const uint8_t *esk, const uint8_t *vr, uint8_t *U) {
    oracle_init_all();
    const mayo_params *p = mayo_oracle_params(version);
    const int v = p->v, k = p->k, mb = p->m_bytes;
    const uint8_t *P1 = esk + p->sk_bytes + p->o_bytes;
    uint8_t *T = (uint8_t *)malloc((size_t)k * mb);
    memset(U, 0, (size_t)p->pairs * mb);
    for (int r = 0; r < v; r++) {
        memset(T, 0, (size_t)k * mb);
        for (int c = r; c < v; c++) {
            const uint8_t *e = P1 + tri_entry(r, c, v) * mb;
            for (int i = 0; i < k; i++) mvec_mac(T + (size_t)i * mb, e, nib(V_of(p, vr, i), c), mb);
        }
        for (int i = 0; i < k; i++) {
            uint8_t vir = nib(V_of(p, vr, i), r);
            for (int j = i; j < k; j++) {
                uint8_t *u = U + (size_t)pair_index(i, j, k) * mb;
                mvec_mac(u, T + (size_t)j * mb, vir, mb);
                if (i != j) mvec_mac(u, T + (size_t)i * mb, nib(V_of(p, vr, j), r), mb);
            }
        }
    }
    free(T);
}

/* Horner over the pair order: y = 0; for i desc, j asc: y = E y + u_ij
 * (sign.cu:440-455, verify.cu:185-200). */
static void horner(const mayo_params *p, const uint8_t *U, uint8_t *y) {
    const int k = p->k, mb = p->m_bytes;
    memset(y, 0, (size_t)mb);
    for (int i = k - 1; i >= 0; i--)
        for (int j = i; j < k; j++) {
            mvec_E(p, y);
            mvec_xor(y, U + (size_t)pair_index(i, j, k) * mb, mb);
        }
}

void oracle_sign_y(int version, const uint8_t *U, const uint8_t *t, uint8_t *y) {
    oracle_init_all();
    const mayo_params *p = mayo_oracle_params(version);
    horner(p, U, y);
    mvec_xor(y, t, p->m_bytes); /* sign.cu:457-459 */
}

/* SampleSolution (sign.cu:10-145) on an unpacked rows x cols system.
 * x holds r on entry. y <- y - A r; echelon form with row swaps and
 * normalised pivots, eliminating below; rank check on the last row;
 * back-substitution x[c] += y[row] and y -= y[row] * A[:, c]. */
int oracle_sample_solution(int rows, int cols, uint8_t *A, uint8_t *y, uint8_t *x) {
    oracle_init_all();
    for (int r = 0; r < rows; r++) {                     /* sign.cu:36-45 */
        uint8_t acc = 0;
        for (int c = 0; c < cols; c++) acc ^= oracle_gf16_mul(A[(size_t)r * cols + c], x[c]);
        y[r] ^= acc;
    }
    int piv_row = 0;
    for (int piv_col = 0; piv_col < cols && piv_row < rows; piv_col++) { /* :48-105 */
        int pr = -1;
        for (int rr = piv_row; rr < rows; rr++)
            if (A[(size_t)rr * cols + piv_col]) { pr = rr; break; }
        if (pr < 0) continue;
        if (pr != piv_row) {
            uint8_t ty = y[piv_row]; y[piv_row] = y[pr]; y[pr] = ty;
            for (int c = 0; c < cols; c++) {
                uint8_t ta = A[(size_t)piv_row * cols + c];
                A[(size_t)piv_row * cols + c] = A[(size_t)pr * cols + c];
                A[(size_t)pr * cols + c] = ta;
            }
        }
        uint8_t inv = oracle_gf16_inv(A[(size_t)piv_row * cols + piv_col]);
        y[piv_row] = oracle_gf16_mul(y[piv_row], inv);
        for (int c = 0; c < cols; c++)
            A[(size_t)piv_row * cols + c] = oracle_gf16_mul(A[(size_t)piv_row * cols + c], inv);
        for (int rr = piv_row + 1; rr < rows; rr++) {
            uint8_t f = A[(size_t)rr * cols + piv_col];
            if (!f) continue;
            y[rr] ^= oracle_gf16_mul(y[piv_row], f);
            for (int c = 0; c < cols; c++)
                A[(size_t)rr * cols + c] ^= oracle_gf16_mul(A[(size_t)piv_row * cols + c], f);
        }
        piv_row++;
    }
    int last_zero = 1;                                   /* :107-115 */
    for (int c = 0; c < cols; c++)
        if (A[(size_t)(rows - 1) * cols + c]) { last_zero = 0; break; }
    if (last_zero) return 0;
    for (int rr = rows - 1; rr >= 0; rr--) {             /* :117-142 */
        int c = -1;
        for (int cc = 0; cc < cols; cc++)
            if (A[(size_t)rr * cols + cc]) { c = cc; break; }
        if (c < 0) continue;
        uint8_t alpha = y[rr];
        if (!alpha) continue;
        x[c] ^= alpha;
        for (int r = 0; r < rows; r++) y[r] ^= oracle_gf16_mul(alpha, A[(size_t)r * cols + c]);
    }
    return 1;
}

int oracle_sign(int version, const uint8_t *esk, const uint8_t *msg, size_t mlen,
                uint8_t *sig, int fail_first) {
    oracle_init_all();
    const mayo_params *p = mayo_oracle_params(version);
    const int n = p->n, m = p->m, o = p->o, k = p->k, v = p->v, KO = k * o;
    uint8_t digest[64], salt[40], t[72], y[72];
    uint8_t *vr = (uint8_t *)malloc((size_t)p->vr_bytes);
    uint8_t *M = (uint8_t *)malloc((size_t)k * m * o);
    uint8_t *A = (uint8_t *)malloc((size_t)m * KO);
    uint8_t *U = (uint8_t *)malloc((size_t)p->pairs * p->m_bytes);
    uint8_t *yu = (uint8_t *)malloc((size_t)m);
    uint8_t *x = (uint8_t *)malloc((size_t)KO);
    const uint8_t *O = esk + p->sk_bytes;

    oracle_sign_digest_salt_t(version, esk, msg, mlen, digest, salt, t);
    int attempts = 256;
    for (int ctr = 0; ctr < 255; ctr++) {                /* sign.cu:244 */
        oracle_sign_vr(version, digest, salt, esk, ctr, vr);
        oracle_sign_M(version, esk, vr, M);
        oracle_sign_A(version, M, A);
        oracle_sign_U(version, esk, vr, U);
        oracle_sign_y(version, U, t, y);
        for (int a = 0; a < m; a++) yu[a] = nib(y, a);
        const uint8_t *rvec = vr + (size_t)k * (v / 2);
        for (int c = 0; c < KO; c++) x[c] = nib(rvec, c);
        int ok = oracle_sample_solution(m, KO, A, yu, x);
        if (ctr < fail_first) {
            /* test hook: report rank failure; x reverts to r as the
             * reference leaves it when the rank check fails (sign.cu:31-33,115) */
            ok = 0;
            for (int c = 0; c < KO; c++) x[c] = nib(rvec, c);
        }
        if (ok) { attempts = ctr + 1; break; }
    }
    /* s_i = (V_i + O x_i || x_i), packed, then the salt (sign.cu:147-191,
     * 472-477). */
    memset(sig, 0, (size_t)p->sig_bytes);
    for (int i = 0; i < k; i++) {
        const uint8_t *Vi = V_of(p, vr, i);
        for (int tt = 0; tt < n; tt++) {
            uint8_t val;
            if (tt < v) {
                uint8_t acc = 0;
                for (int kk = 0; kk < o; kk++)
                    acc ^= oracle_gf16_mul(nib(O, tt * o + kk), x[i * o + kk]);
                val = nib(Vi, tt) ^ acc;
            } else {
                val = x[i * o + (tt - v)];
            }
            set_nib(sig, i * n + tt, val);
        }
    }
    memcpy(sig + (size_t)k * n / 2, salt, (size_t)p->sk_bytes);
    free(vr); free(M); free(A); free(U); free(yu); free(x);
    return attempts;
}

/* ------------------------------------------------------------------------ */
/* Verification (verify.cu:9-211).                                          */
/* ------------------------------------------------------------------------ */

/* U_ij over P = [[P1, P2], [0, P3]] with s_i = (V_i || X_i), then Horner. */
void oracle_eval_public(int version, const uint8_t *epk, const uint8_t *s, uint8_t *y) {
    oracle_init_all();
    const mayo_params *p = mayo_oracle_params(version);
    const int n = p->n, o = p->o, k = p->k, v = p->v, mb = p->m_bytes;
    const uint8_t *P1 = epk, *P2 = epk + p->p1_bytes, *P3 = P2 + p->p2_bytes;
    uint8_t *U = (uint8_t *)calloc((size_t)p->pairs * mb, 1);
    uint8_t *T = (uint8_t *)malloc((size_t)k * mb);
#define SV(i, r) s[(size_t)(i) * n + (r)]
#define SX(i, r) s[(size_t)(i) * n + v + (r)]
    /* P1 block, verify.cu:78-111 */
    for (int r = 0; r < v; r++) {
        memset(T, 0, (size_t)k * mb);
        for (int c = r; c < v; c++)
            for (int i = 0; i < k; i++) mvec_mac(T + (size_t)i * mb, P1 + tri_entry(r, c, v) * mb, SV(i, c), mb);
        for (int i = 0; i < k; i++)
            for (int j = i; j < k; j++) {
                uint8_t *u = U + (size_t)pair_index(i, j, k) * mb;
                mvec_mac(u, T + (size_t)j * mb, SV(i, r), mb);
                if (i != j) mvec_mac(u, T + (size_t)i * mb, SV(j, r), mb);
            }
    }
    /* P2 block, verify.cu:113-146 */
    for (int r = 0; r < v; r++) {
        memset(T, 0, (size_t)k * mb);
        for (int c = 0; c < o; c++)
            for (int i = 0; i < k; i++) mvec_mac(T + (size_t)i * mb, P2 + rect_entry(r, c, o) * mb, SX(i, c), mb);
        for (int i = 0; i < k; i++)
            for (int j = i; j < k; j++) {
                uint8_t *u = U + (size_t)pair_index(i, j, k) * mb;
                mvec_mac(u, T + (size_t)j * mb, SV(i, r), mb);
                if (i != j) mvec_mac(u, T + (size_t)i * mb, SV(j, r), mb);
            }
    }
    /* P3 block, verify.cu:148-181 */
    for (int r = 0; r < o; r++) {
        memset(T, 0, (size_t)k * mb);
        for (int c = r; c < o; c++)
            for (int i = 0; i < k; i++) mvec_mac(T + (size_t)i * mb, P3 + tri_entry(r, c, o) * mb, SX(i, c), mb);
        for (int i = 0; i < k; i++)
            for (int j = i; j < k; j++) {
                uint8_t *u = U + (size_t)pair_index(i, j, k) * mb;
                mvec_mac(u, T + (size_t)j * mb, SX(i, r), mb);
                if (i != j) mvec_mac(u, T + (size_t)i * mb, SX(j, r), mb);
            }
    }
#undef SV
#undef SX
    horner(p, U, y);
    free(U); free(T);
}

int oracle_verify(int version, const uint8_t *epk, const uint8_t *msg, size_t mlen,
                  const uint8_t *sig) {
    oracle_init_all();
    const mayo_params *p = mayo_oracle_params(version);
    const int n = p->n, k = p->k, D = p->digest_bytes, SK = p->sk_bytes;
    uint8_t buf[64 + 40], t[72], y[72];
    uint8_t *s = (uint8_t *)malloc((size_t)k * n);
    oracle_shake256(buf, (size_t)D, msg, mlen);           /* verify.cu:46 */
    memcpy(buf + D, sig + (size_t)k * n / 2, (size_t)SK);  /* salt, :37-38 */
    oracle_shake256(t, (size_t)p->m_bytes, buf, (size_t)(D + SK)); /* :48 */
    for (int e = 0; e < k * n; e++) s[e] = nib(sig, e);    /* :62-69 */
    oracle_eval_public(version, epk, s, y);
    free(s);
    return memcmp(y, t, (size_t)p->m_bytes) == 0 ? 0 : 1;  /* :203-211 */
}

/* ------------------------------------------------------------------------ */
/* Batch helpers.                                                           */
/* ------------------------------------------------------------------------ */

void oracle_sign_batch(int version, const uint8_t *esk_table, const int32_t *key_idx,
                       const uint8_t *msgs, const int64_t *msg_off, const int32_t *msg_len,
                       uint8_t *sigs, int64_t n, int threads) {
    oracle_init_all();
    const mayo_params *p = mayo_oracle_params(version);
#ifdef _OPENMP
    if (threads > 0) omp_set_num_threads(threads);
#pragma omp parallel for schedule(dynamic, 1)
#endif
    for (int64_t i = 0; i < n; i++)
        oracle_sign(version, esk_table + (size_t)key_idx[i] * p->esk_bytes,
                    msgs + msg_off[i], (size_t)msg_len[i], sigs + (size_t)i * p->sig_bytes, 0);
}

void oracle_verify_batch(int version, const uint8_t *epk_table, const int32_t *key_idx,
                         const uint8_t *msgs, const int64_t *msg_off, const int32_t *msg_len,
                         const uint8_t *sigs, uint8_t *results, int64_t n, int threads) {
    oracle_init_all();
    const mayo_params *p = mayo_oracle_params(version);
#ifdef _OPENMP
    if (threads > 0) omp_set_num_threads(threads);
#pragma omp parallel for schedule(dynamic, 1)
#endif
    for (int64_t i = 0; i < n; i++)
        results[i] = (uint8_t)oracle_verify(version, epk_table + (size_t)key_idx[i] * p->epk_bytes,
                                            msgs + msg_off[i], (size_t)msg_len[i],
                                            sigs + (size_t)i * p->sig_bytes);
}
