/*
 * mayo_oracle.h -- CPU restatement of the cuMAYO algorithm (TEST INFRASTRUCTURE).
 *
 * This library is the correctness oracle for the HIP engine in cumayo_amd/.
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load it, and only as the checker or the reported CPU baseline; the product
 * path never links or calls it.
 *
 * Every function restates one stage of the reference (file:line cited at each
 * definition in mayo_oracle.c). All four parameter sets live in one build,
 * selected at run time by a descriptor (the reference selects one at compile
 * time, params.h:6-96).
 */
#ifndef MAYO_ORACLE_H
#define MAYO_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    int version;        /* 1, 2, 3 or 5 */
    int n, m, o, k, v;  /* v = n - o (equals m in every shipped set) */
    int sk_bytes;       /* compact sk = seed_sk; also the salt length */
    int pk_bytes;       /* 16-byte seed_pk || P3 */
    int p1_bytes, p2_bytes, p3_bytes, p_bytes;
    int o_bytes;        /* v*o nibbles */
    int digest_bytes;
    int pairs;          /* k(k+1)/2 */
    int m_bytes;        /* one m-vector = m/2 bytes */
    int sig_bytes;      /* k*n/2 + salt */
    int esk_bytes;      /* sk || O || P1 || L */
    int epk_bytes;      /* P1 || P2 || P3 */
    int vr_bytes;       /* V || r squeezed per attempt = k*n/2 */
    uint8_t f_tail[4];  /* C0..C3 of f(z) = z^m + C3 z^3 + C2 z^2 + C1 z + C0 */
} mayo_params;

/* Returns NULL for an unknown version. */
const mayo_params *mayo_oracle_params(int version);

/* FIPS 202 / FIPS 197 primitives. */
void oracle_shake128(uint8_t *out, size_t outlen, const uint8_t *in, size_t inlen);
void oracle_shake256(uint8_t *out, size_t outlen, const uint8_t *in, size_t inlen);
void oracle_keccak_f1600(uint64_t st[25]);
void oracle_aes128_ecb_block(const uint8_t key[16], const uint8_t in[16], uint8_t out[16]);
void oracle_aes128_ctr(const uint8_t key[16], uint8_t *out, size_t outlen);

/* GF(16) = GF(2)[x]/(x^4+x+1). */
uint8_t oracle_gf16_mul(uint8_t a, uint8_t b);
uint8_t oracle_gf16_inv(uint8_t a);

/* y <- z*y mod f(z) over a packed m-vector (poly.cuh:11-69). */
void oracle_mvec_E(int version, uint8_t *y_packed);

/* Stages. All buffers are host memory in the reference byte formats. */
void oracle_keypair_seed(int version, uint64_t rand_index, uint64_t item, uint8_t *sk);
void oracle_keypair(int version, uint64_t rand_index, uint64_t item, uint8_t *pk, uint8_t *sk);
void oracle_keypair_from_sk(int version, const uint8_t *sk, uint8_t *pk);
void oracle_expandsk(int version, const uint8_t *sk, uint8_t *esk);
void oracle_expandpk(int version, const uint8_t *cpk, uint8_t *epk);
/* Returns the number of attempts used (1..255); 256 means every attempt
 * failed and the last attempt's vector was written anyway (sign.cu:472-477).
 * fail_first forces the first fail_first attempts to report rank failure
 * (a test hook for the retry path; 0 in normal use). */
int oracle_sign(int version, const uint8_t *esk, const uint8_t *msg, size_t mlen,
                uint8_t *sig, int fail_first);
/* 0 = valid, 1 = invalid (verify.cu:209-211). */
int oracle_verify(int version, const uint8_t *epk, const uint8_t *msg, size_t mlen,
                  const uint8_t *sig);

/* Intermediate products exposed for the tests. */
void oracle_sign_digest_salt_t(int version, const uint8_t *esk, const uint8_t *msg, size_t mlen,
                               uint8_t *digest, uint8_t *salt, uint8_t *t);
void oracle_sign_vr(int version, const uint8_t *digest, const uint8_t *salt,
                    const uint8_t *seed_sk, int ctr, uint8_t *vr);
/* M_i = V_i * L, output [k][m][o] nibbles unpacked (one per byte). */
void oracle_sign_M(int version, const uint8_t *esk, const uint8_t *vr, uint8_t *M);
/* A, output [m][k*o] nibbles unpacked. */
void oracle_sign_A(int version, const uint8_t *M, uint8_t *A);
/* U per pair, output [pairs][m_bytes] packed. */
void oracle_sign_U(int version, const uint8_t *esk, const uint8_t *vr, uint8_t *U);
/* y = t + sum E^l u_l, packed m_bytes. */
void oracle_sign_y(int version, const uint8_t *U, const uint8_t *t, uint8_t *y);
/* Generic SampleSolution over an unpacked rows x cols system; x is in/out:
 * on entry it holds r, on exit r + z. Returns 1 if rank == rows. */
int oracle_sample_solution(int rows, int cols, uint8_t *A, uint8_t *y, uint8_t *x);
/* Evaluate the public map: U over pairs for s (k*n nibbles unpacked), then
 * Horner, giving the packed m-vector y (verify.cu:62-200). */
void oracle_eval_public(int version, const uint8_t *epk, const uint8_t *s_unpacked, uint8_t *y);

/* Batch helpers (OpenMP over items) used by tests and the CPU baseline. */
void oracle_sign_batch(int version, const uint8_t *esk_table, const int32_t *key_idx,
                       const uint8_t *msgs, const int64_t *msg_off, const int32_t *msg_len,
                       uint8_t *sigs, int64_t n, int threads);
void oracle_verify_batch(int version, const uint8_t *epk_table, const int32_t *key_idx,
                         const uint8_t *msgs, const int64_t *msg_off, const int32_t *msg_len,
                         const uint8_t *sigs, uint8_t *results, int64_t n, int threads);

#ifdef __cplusplus
}
#endif

#endif
