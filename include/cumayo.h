/*
 * cumayo.h -- C ABI of the MI355X MAYO engine (one shared library per
 * parameter set: libcumayo_mayo1.so, libcumayo_mayo2.so, libcumayo_mayo3.so,
 * libcumayo_mayo5.so, mirroring the reference's one build per MAYO_VERSION,
 * params.h:6-8 and tests/CMakeLists.txt:28-45).
 *
 * Part 1 is a drop-in for include/api.cuh of cuMAYO: the same 20 entry
 * points with the same names, argument order, pitched device pool layouts and
 * stream semantics (cudaStream_t becomes hipStream_t), plus
 * aes128_tables_init (aes128_ctr.cuh:16). Like the reference these calls
 * return void, never synchronise and only enqueue work on `stream`; unlike
 * the reference, the last launch or copy error is kept and can be read with
 * cumayo_last_error().
 *
 * Part 2 extends the boundary with the batched key-table form used for
 * shared resident keys and ragged messages, Part 3 with a
 * continuous-batching request server over it, and Part 4 with the launch
 * autotuner and HIP-graph captures of the batch pipelines; none of them has
 * a counterpart in the reference.
 */
#ifndef CUMAYO_H
#define CUMAYO_H

#include <stddef.h>
#include <stdint.h>
#include <hip/hip_runtime_api.h>

#ifdef __cplusplus
#define CUMAYO_DEFAULT(x) = x
extern "C" {
#else
#define CUMAYO_DEFAULT(x)
#endif

/* Most items (keys, signatures or verifications) one call of a batched entry
 * point takes: HIP launches at most 2^32 work-items per grid dimension, and
 * the workgroup-per-item kernels use 256 work-items per item. A larger batch
 * size is refused before any work is enqueued, with hipErrorInvalidValue
 * (returned, or recorded for cumayo_last_error by the void calls of Part 1);
 * split larger batches into several calls. */
#define CUMAYO_MAX_BATCH 16777215u

/* ---------------- Part 1: drop-in for api.cuh ---------------- */

/* api.cuh:9 -- generates batch_size key pairs into pitched pool rows:
 * sk at 0, pk at A128(sk), P1||P2 scratch after A128(pk) (api.cu:17-26).
 * The seed of row i is LE64(rand_index + 2 i + 1) (keypair.cu:32-40). */
void crypto_sign_keypair(uint8_t *d_keypair_mem_pool, size_t keypair_mem_pool_pitch,
                         size_t batch_size CUMAYO_DEFAULT(1), hipStream_t stream CUMAYO_DEFAULT(nullptr),
                         size_t rand_index CUMAYO_DEFAULT(0));
/* api.cuh:11 (api.cu:29-38) */
void crypto_sign_keypair_DtH(uint8_t *pk, uint8_t *sk, uint8_t *d_keypair_mem_pool, size_t pitch,
                             size_t batch_size, hipStream_t stream CUMAYO_DEFAULT(nullptr));
/* api.cuh:13-17 (api.cu:40-62): compact sk in, esk = sk||O||P1||L out */
void crypto_sign_expandsk_HtD(uint8_t *csk, uint8_t *d_expandsk_mem_pool, size_t pitch,
                              size_t batch_size, hipStream_t stream CUMAYO_DEFAULT(nullptr));
void crypto_sign_expandsk(uint8_t *d_expandsk_mem_pool, size_t pitch, size_t batch_size,
                          hipStream_t stream CUMAYO_DEFAULT(nullptr));
void crypto_sign_expandsk_DtH(uint8_t *esk, uint8_t *d_expandsk_mem_pool, size_t pitch,
                              size_t batch_size, hipStream_t stream CUMAYO_DEFAULT(nullptr));
/* api.cuh:19-23 (api.cu:64-85): compact pk placed at P_BYTES-16, epk = P1||P2||P3 */
void crypto_sign_expandpk_HtD(uint8_t *cpk, uint8_t *d_expandpk_mem_pool, size_t pitch,
                              size_t batch_size, hipStream_t stream CUMAYO_DEFAULT(nullptr));
void crypto_sign_expandpk(uint8_t *d_expandpk_mem_pool, size_t pitch, size_t batch_size,
                          hipStream_t stream CUMAYO_DEFAULT(nullptr));
void crypto_sign_expandpk_DtH(uint8_t *epk, uint8_t *d_expandpk_mem_pool, size_t pitch,
                              size_t batch_size, hipStream_t stream CUMAYO_DEFAULT(nullptr));
/* api.cuh:25-29 (api.cu:88-114): rows esk || msg || sig || scratch
 * (scratch of k*m*ceil(o/2) bytes, as the reference's d_M) */
void crypto_sign_sign_HtD(uint8_t *cesk, uint8_t *cMess, size_t mlen, uint8_t *d_sign_mem_pool,
                          size_t pitch, size_t batch_size, hipStream_t stream CUMAYO_DEFAULT(nullptr));
void crypto_sign_sign(uint8_t *d_sign_mem_pool, size_t Mlen, size_t pitch, size_t batch_size,
                      hipStream_t stream CUMAYO_DEFAULT(nullptr));
void crypto_sign_sign_DtH(uint8_t *sig, uint8_t *d_sig_mem_pool, size_t Mlen, size_t pitch,
                          size_t batch_size, hipStream_t stream CUMAYO_DEFAULT(nullptr));
/* api.cuh:31-35 (api.cu:116-144): compact sk in, expanded then signed;
 * rows A128(esk) || A128(msg) || A128(sig) || scratch. Where
 * cumayo_signwithesk_lfree() is 1 (MAYO_2) it signs without L: after the
 * call the esk region holds sk || O || P1 || P2 (the reference leaves
 * sk || O || P1 || L there); the signatures are the reference's. */
void crypto_sign_signWithesk_HtD(uint8_t *cesk, uint8_t *cMess, size_t mlen,
                                 uint8_t *d_sign_mem_pool, size_t pitch, size_t batch_size,
                                 hipStream_t stream CUMAYO_DEFAULT(nullptr));
void crypto_sign_signWithesk(uint8_t *d_sign_mem_pool, size_t Mlen, size_t pitch,
                             size_t batch_size, hipStream_t stream CUMAYO_DEFAULT(nullptr));
void crypto_sign_signWithesk_DtH(uint8_t *sig, uint8_t *d_sig_mem_pool, size_t Mlen, size_t pitch,
                                 size_t batch_size, hipStream_t stream CUMAYO_DEFAULT(nullptr));
/* api.cuh:37-41 (api.cu:146-178): rows A128(epk) || A128(msg) || A128(sig) || result;
 * host epk and sig rows are A128-padded, as in the reference */
void crypto_sign_verify_HtD(uint8_t *epk, uint8_t *Mess, uint8_t *sig, size_t mlen,
                            uint8_t *d_verify_mem_pool, size_t pitch, size_t batch_size,
                            hipStream_t stream CUMAYO_DEFAULT(nullptr));
void crypto_sign_verify(uint8_t *d_verify_mem_pool, size_t Mlen, size_t pitch, size_t batch_size,
                        hipStream_t stream CUMAYO_DEFAULT(nullptr));
void crypto_sign_verify_DtH(uint8_t *verify, size_t Mlen, uint8_t *d_verify_mem_pool, size_t pitch,
                            size_t batch_size, hipStream_t stream CUMAYO_DEFAULT(nullptr));
/* api.cuh:43-47 (api.cu:180-218): compact pk in, expanded then verified;
 * result byte 0 = valid, 1 = invalid (verify.cu:209-211) */
void crypto_sign_verifyWithepk_HtD(uint8_t *epk, uint8_t *Mess, uint8_t *sig, size_t mlen,
                                   uint8_t *d_verify_mem_pool, size_t pitch, size_t batch_size,
                                   hipStream_t stream CUMAYO_DEFAULT(nullptr));
void crypto_sign_verifyWithepk(uint8_t *d_verify_mem_pool, size_t Mlen, size_t pitch,
                               size_t batch_size, hipStream_t stream CUMAYO_DEFAULT(nullptr));
void crypto_sign_verifyWithepk_DtH(uint8_t *verify, size_t Mlen, uint8_t *d_verify_mem_pool,
                                   size_t pitch, size_t batch_size,
                                   hipStream_t stream CUMAYO_DEFAULT(nullptr));
/* aes128_ctr.cuh:16 (aes128_ctr.cu:69-72): uploads the AES tables to the
 * current device. The entry points above also do it on first use. */
void aes128_tables_init(void);

/* ---------------- Part 2: engine extensions ---------------- */

typedef struct {
    int version, n, m, o, k;
    int sk_bytes, pk_bytes, sig_bytes, esk_bytes, epk_bytes, digest_bytes;
    int p1_bytes, p2_bytes, p3_bytes, o_bytes;
    /* batch-form workspace: bytes per signing item, and the offset within
     * an item's slot of the byte that reports the attempts used (1..255,
     * 0 when every attempt failed) */
    int sign_ws_stride, sign_ws_attempts;
} cumayo_sizes;

/* Parameter set compiled into this library. */
void cumayo_get_sizes(cumayo_sizes *out);
/* Last HIP error recorded by any entry point of this library (0 = none);
 * reading clears it. */
int cumayo_last_error(void);
/* 1 when crypto_sign_signWithesk signs without L in this library (the
 * parameter sets where that is faster; the pool then holds P2 in L's place),
 * 0 when it expands and uses L as the reference does. */
int cumayo_signwithesk_lfree(void);
/* Test hook (no counterpart in the reference): with attempts > 0, the first
 * `attempts` attempts of every item of the drop-in signing kernels
 * (crypto_sign_sign, crypto_sign_signWithesk) report a singular system, as
 * a rank failure would; 0 turns it off. */
void cumayo_test_pitched_fail_first(int attempts);

/* Batched signing over a table of expanded secret keys (reference esk
 * format, rows esk_stride bytes apart): item i signs msg_len[i] bytes at
 * msgs + msg_off[i] with key key_idx[i] into sigs + i*sig_bytes. All
 * pointers are device pointers; workspace holds
 * cumayo_sign_workspace_size(n) bytes: n per-item slots, then one T region
 * per item that the shared-key launch variants write and read (the rows
 * T_j = P1 V_j, 31.2 KB per item for MAYO_1), then 16 KB per 4,096 items for
 * their key sort. Returns 0 or a hipError_t. The
 * per-item arrays are device data and are not checked: key_idx[i] must name
 * a row of the table, and msg_off[i] >= 0, msg_len[i] >= 0 must lie within
 * msgs (as for every table and prepared form below); other values make the
 * kernels read outside the caller's buffers. */
size_t cumayo_sign_workspace_size(size_t n);
int cumayo_sign_table(const uint8_t *esk_table, size_t esk_stride, const int32_t *key_idx,
                      const uint8_t *msgs, const int64_t *msg_off, const int32_t *msg_len,
                      uint8_t *sigs, size_t n, uint8_t *workspace, hipStream_t stream);

/* The same in two stages, selected by the bit mask `stages`: 1 = the
 * hashing pass (digest, salt, target, vinegar vector of attempt 0, one lane
 * per item), 2 = the linear-algebra pass (one wavefront per item: M, U and
 * y, then the solve, then a retry pass over the items whose first attempt
 * met a singular system). Stage 2 reads what stage 1 left in the workspace;
 * 3 runs both. Adding 4 runs stage 2 with the workgroup-per-item kernels
 * instead (same results; kept as the comparison form). Test hook: a nonzero
 * value f in bits 8..15 makes attempts 0 .. f-1 of every item report a
 * singular system (the retry path; f = 255 exhausts every attempt and
 * writes the last one's vector, attempts byte 0), as the oracle's
 * fail_first does; not with bit 4. */
int cumayo_sign_table_stages(const uint8_t *esk_table, size_t esk_stride, const int32_t *key_idx,
                             const uint8_t *msgs, const int64_t *msg_off, const int32_t *msg_len,
                             uint8_t *sigs, size_t n, uint8_t *workspace, hipStream_t stream,
                             int stages);

/* Batched verification over a table of expanded public keys (P1||P2||P3);
 * results[i] = 0 valid, 1 invalid. The workspace holds
 * cumayo_verify_workspace_size(n) bytes: n per-item slots, then one T region
 * per item for the shared-key variants (34.4 KB per item for MAYO_1) and
 * their key sort's 16 KB per 4,096 items. */
size_t cumayo_verify_workspace_size(size_t n);
int cumayo_verify_table(const uint8_t *epk_table, size_t epk_stride, const int32_t *key_idx,
                        const uint8_t *msgs, const int64_t *msg_off, const int32_t *msg_len,
                        const uint8_t *sigs, uint8_t *results, size_t n, uint8_t *workspace,
                        hipStream_t stream);

int cumayo_verify_table_stages(const uint8_t *epk_table, size_t epk_stride, const int32_t *key_idx,
                               const uint8_t *msgs, const int64_t *msg_off, const int32_t *msg_len,
                               const uint8_t *sigs, uint8_t *results, size_t n, uint8_t *workspace,
                               hipStream_t stream, int stages);

/* Prepared key tables: the reference's esk blocks (sk || O || P1 || L) and
 * epk blocks (P1 || P2 || P3) with every m-vector entry padded with zero
 * bytes to a multiple of 4 bytes and every block 16-byte aligned, rows
 * cumayo_prepared_esk_bytes() / cumayo_prepared_epk_bytes() apart (multiples
 * of 128), and each little-endian entry dword stored bit-plane-major (bit
 * 4e + d of the nibble form at bit 8d + e: byte d holds bit d of the 8
 * elements). Built once per key on the device from a reference-format table;
 * the kernels then read each entry word with one aligned load. An opaque
 * format: `prepared` must be 16-byte aligned; n_keys <= 65535 per call. */
size_t cumayo_prepared_esk_bytes(void);
size_t cumayo_prepared_epk_bytes(void);
int cumayo_prepare_esk(const uint8_t *esk_table, size_t esk_stride, size_t n_keys, uint8_t *prepared,
                       hipStream_t stream);
int cumayo_prepare_epk(const uint8_t *epk_table, size_t epk_stride, size_t n_keys, uint8_t *prepared,
                       hipStream_t stream);
/* cumayo_sign_table_stages / cumayo_verify_table_stages over prepared
 * tables (same results; stage bit 4 is not available here). */
int cumayo_sign_prepared(const uint8_t *prepared, const int32_t *key_idx, const uint8_t *msgs,
                         const int64_t *msg_off, const int32_t *msg_len, uint8_t *sigs, size_t n,
                         uint8_t *workspace, hipStream_t stream, int stages);
int cumayo_verify_prepared(const uint8_t *prepared, const int32_t *key_idx, const uint8_t *msgs,
                           const int64_t *msg_off, const int32_t *msg_len, const uint8_t *sigs,
                           uint8_t *results, size_t n, uint8_t *workspace, hipStream_t stream,
                           int stages);

/* cumayo_sign_prepared (stages 3) without a workspace argument and without
 * allocating: no counterpart in the reference. For callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc). The workspace is a file-scope device
 * array of CUMAYO_NOALLOC_MAX_SIGN items' workspace slots that each GPU holds
 * for the life of the process (CUMAYO_NOALLOC_MAX_SIGN x sign_ws_stride
 * bytes: 7.5 MiB for MAYO_1, 23 MiB for MAYO_5); its address is looked up once
 * per GPU (hipGetSymbolAddress) on the first call with that GPU current.
 * Batches above CUMAYO_NOALLOC_MAX_SIGN items return hipErrorInvalidValue
 * with nothing enqueued. Every call on one GPU shares that array: calls that
 * overlap on one GPU (other streams, other threads) race and must be
 * serialised by the caller. Same results as cumayo_sign_prepared. */
#define CUMAYO_NOALLOC_MAX_SIGN 2048u
int cumayo_sign_prepared_noalloc(const uint8_t *prepared, const int32_t *key_idx, const uint8_t *msgs,
                                 const int64_t *msg_off, const int32_t *msg_len, uint8_t *sigs, size_t n,
                                 hipStream_t stream);

/* NIST wire format (the crypto_sign / crypto_sign_open convention of the
 * MAYO submission: sm = signature || message), batched over prepared key
 * tables; no counterpart in the reference, which exposes detached signatures
 * only. Device pointers throughout; ragged offsets are caller-chosen
 * (sm_off[i] = msg_off[i] + i * sig_bytes packs them densely).
 * cumayo_crypto_sign writes sm_len = sig_bytes + msg_len[i] bytes at
 * sm + sm_off[i]. cumayo_crypto_sign_open splits each sm into the signature
 * and the message (copied to msgs_out + out_off[i]), verifies, and sets
 * results[i] = 0 valid / 1 invalid (also for sm shorter than a signature)
 * and mlen_out[i] = message length, or -1 when invalid (mlen_out may be
 * NULL). */
size_t cumayo_crypto_sign_workspace_size(size_t n);
int cumayo_crypto_sign(const uint8_t *prepared, const int32_t *key_idx, const uint8_t *msgs,
                       const int64_t *msg_off, const int32_t *msg_len, uint8_t *sm, const int64_t *sm_off,
                       size_t n, uint8_t *workspace, hipStream_t stream);
size_t cumayo_crypto_sign_open_workspace_size(size_t n);
int cumayo_crypto_sign_open(const uint8_t *prepared, const int32_t *key_idx, const uint8_t *sm,
                            const int64_t *sm_off, const int32_t *sm_len, uint8_t *msgs_out,
                            const int64_t *out_off, int32_t *mlen_out, uint8_t *results, size_t n,
                            uint8_t *workspace, hipStream_t stream);

/* ---------------- Part 3: continuous-batching server ---------------- */

/* Requests queue on the host; a worker thread admits them into free batch
 * slots (one HIP stream, pinned staging and device buffers each) as earlier
 * batches finish, so batches launched while others run overlap on the GPU.
 * Device memory is bounded by the prepared copies of the key tables the
 * server keeps (n_keys rows of cumayo_prepared_esk_bytes() and
 * cumayo_prepared_epk_bytes()) plus n_slots x (slot_msg_bytes + per-request
 * buffers for max_batch requests); a request waits until a slot with room
 * frees up. Results are identical to the batch entry points. No counterpart
 * in the reference. */
#define CUMAYO_OP_SIGN 0
#define CUMAYO_OP_VERIFY 1

typedef struct {
    int device;
    const uint8_t *esk_table; size_t esk_stride; /* device; NULL if no signing */
    const uint8_t *epk_table; size_t epk_stride; /* device; NULL if no verifying */
    int32_t n_keys;
    int32_t max_batch;      /* requests per slot (0 = 4096; at most CUMAYO_MAX_BATCH) */
    int32_t n_slots;        /* batches in flight */
    size_t slot_msg_bytes;  /* message bytes per slot */
    int64_t max_queue;      /* pending requests beyond which submit refuses (0 = no limit) */
    int32_t host_threads;   /* threads that copy requests into and results out of the
                               pinned staging (0 = 4; 1 = the worker alone) */
} cumayo_serve_config;

typedef struct {
    uint8_t op;             /* CUMAYO_OP_SIGN or CUMAYO_OP_VERIFY */
    int32_t key;            /* row of the key table */
    const uint8_t *msg;     /* host message */
    uint32_t len;
    const uint8_t *sig_in;  /* verify: host signature */
    uint8_t *sig_out;       /* sign: host output (sig_bytes) or NULL */
    uint8_t *result_out;    /* host byte or NULL: verify result (0 valid / 1 invalid);
                               sign: attempts used (0 = all 255 failed) */
    uint64_t tag;
    int64_t t_arrival_ns;   /* 0 = now; else the arrival time on cumayo_serve_clock_ns */
} cumayo_request;

typedef struct {
    uint64_t tag;
    int64_t latency_ns;     /* arrival to result on the host */
    uint8_t status;         /* verify: 0 valid / 1 invalid; sign: 0 ok, 1 all attempts failed;
                               2 = refused (key out of range, no key table for the op,
                               or a message longer than slot_msg_bytes);
                               3 = cancelled before it was launched;
                               4 = CUMAYO_STATUS_ERROR: a copy or launch of the request's
                               batch failed, or its work ended in a HIP error; no output
                               (signature, result byte) was written */
} cumayo_completion;

#define CUMAYO_STATUS_ERROR 4

typedef struct {
    uint64_t batches, requests, rejected, max_in_flight, overlapped_launches;
    size_t device_bytes;
    uint64_t cancelled;
    uint64_t failed;        /* requests completed with CUMAYO_STATUS_ERROR */
    int32_t last_error;     /* the last HIP error the server met (0 = none) */
} cumayo_serve_stats;

typedef struct {
    double offered_per_s, achieved_per_s;
    double p50_ms, p90_ms, p99_ms, mean_ms, max_ms;
    uint64_t submitted, completed, refused;
    double duration_s;
} cumayo_serve_report;

void *cumayo_serve_create(const cumayo_serve_config *cfg);
void cumayo_serve_destroy(void *server);
/* Returns the number of requests accepted (admission stops at max_queue). */
size_t cumayo_serve_submit(void *server, const cumayo_request *reqs, size_t n);
size_t cumayo_serve_poll(void *server, cumayo_completion *out, size_t max);
/* Withdraws a queued request (not yet in a batch): it completes with status
 * 3 and its buffers are no longer touched. Returns 1 if it was withdrawn, 0
 * if it had already been launched or completed. */
int cumayo_serve_cancel(void *server, uint64_t tag);
void cumayo_serve_get_stats(void *server, cumayo_serve_stats *out);
/* Test hook: the next n batches the server launches fail as if an enqueue
 * had returned an error; nothing is launched for them and their requests
 * complete with CUMAYO_STATUS_ERROR. */
void cumayo_serve_inject_failures(void *server, int n);
int64_t cumayo_serve_clock_ns(void);
/* Open-loop load: Poisson arrivals at rate_per_s for duration_s over a pool
 * of prepared requests (message i of the pool, key msg_key[i], and for
 * verification the signature sig_pool + i * sig_bytes); a fraction
 * verify_fraction of arrivals are verifications. Drains, then reports
 * latency percentiles (arrival to result); with lat_out, the first lat_cap
 * latencies (ns, completion order) are also written there. Returns 0 or an
 * error. */
int cumayo_serve_run_poisson(void *server, double rate_per_s, double duration_s,
                             double verify_fraction, uint64_t seed, const uint8_t *msg_pool,
                             const int64_t *msg_off, const int32_t *msg_len, const int32_t *msg_key,
                             const uint8_t *sig_pool, size_t pool_n, cumayo_serve_report *out,
                             int64_t *lat_out, size_t lat_cap);

/* ---------------- Part 4: launch autotuner and captured pipelines ---------------- */

/* Launch configuration of the batch entry points (process-wide for this
 * library; read at every launch and at graph capture). Over prepared key
 * tables the wave kernels come in compiled variants: the quadratic-form
 * kernels (sign_prep, verify) are coded 10 x rows per lane + waves per SIMD
 * they are built for (23: each lane owns two rows of a chunk, three waves per
 * SIMD; one row per lane halves the accumulators but shares each column's
 * scalar masks over fewer rows), the solve kernel by its waves per SIMD; the
 * one-wave-per-item kernels of every table form run their items either in
 * batch order (0) or in an XCD-aware order (1: items sorted by key and dealt
 * to the eight XCDs in contiguous runs, so each XCD's L2 holds few keys at a
 * time; one extra one-workgroup sort per launch, batches over 64 items
 * only). sign_prep and verify also take the shared-key variants, coded
 * 1000 + 10 W + CG: a workgroup of W wavefronts (one item each, items of one
 * key: the items are sorted by key first, exactly for key indices below
 * 4,096) builds the 16 GF(16) multiples of CG key columns at a time in LDS,
 * and each item's quadratic form reads its products from those tables (the
 * item order fields do not apply to them; with sign_chunks 0 the fused
 * kernel runs its default quadratic form). They pay
 * where many items share a key and the wave kernels run short of registers
 * (MAYO_5 signing). Every configuration gives the same bytes. */
typedef struct {
    int32_t sign_prep;    /* quadratic form + M + y kernel (10 x rows per lane + waves per SIMD,
                             or 1000 + 10 W + CG for the shared-key tables) */
    int32_t sign_solve;   /* columns of A + elimination kernel: 100 (signatures per
                             workgroup - 1) + 10 pivot-search schedule (0 scalar
                             unit, 1 vector unit) + waves per SIMD */
    int32_t sign_order;   /* 0 batch order, 1 XCD-aware key order */
    int32_t verify;       /* verification kernel (as sign_prep) */
    int32_t verify_order;
    int32_t sign_chunks;  /* 1: the signing linear-algebra pass as three kernels over the
                             batch (quadratic form + M + y, elimination, retries);
                             2: from 16,384 items on, that pass as two halves on two
                             streams (the second on a helper stream kept per calling thread
                             and caller stream, forked from and joined back to the caller's
                             stream), so that each half fills the other's kernel tails;
                             0: one persistent fused kernel, each wave signing items in a
                             grid-stride loop, retries included (sign_prep picks its
                             variant; sign_solve does not apply) */
} cumayo_launch_config;

#define CUMAYO_FIELD_SIGN_PREP 0
#define CUMAYO_FIELD_SIGN_SOLVE 1
#define CUMAYO_FIELD_VERIFY 2
#define CUMAYO_FIELD_ORDER 3
#define CUMAYO_FIELD_SIGN_CHUNKS 4

void cumayo_get_launch_config(cumayo_launch_config *out);
/* Returns hipErrorInvalidValue (nothing changed) for a value that is not a
 * compiled variant. */
int cumayo_set_launch_config(const cumayo_launch_config *cfg);
/* The values a field can take (entry 0 is the default); writes up to max of
 * them to out and returns their number, or -1 for an unknown field. */
int cumayo_launch_variants(int field, int32_t *out, int max);

typedef struct {
    cumayo_launch_config best;          /* the configuration now in force */
    double sign_ms, verify_ms;          /* linear-algebra pass, best configuration */
    double sign_default_ms, verify_default_ms;  /* the same with the defaults */
    int32_t configs_timed;
} cumayo_tune_report;

/* Launch autotuner: times the linear-algebra pass (stage 2) of
 * cumayo_sign_prepared and cumayo_verify_prepared on the caller's own batch
 * under every combination of compiled variants and item orders (HIP events
 * on `stream`, the least of `reps` runs after an untimed one), then makes the
 * fastest the configuration in force. The hashing pass runs once first.
 * Signing is tuned when prep_esk and sign_ws are given, verification when
 * prep_epk, verify_ws and results are; verification reads sigs. sigs,
 * results and the workspaces end up holding the results of the batch.
 * Synchronises `stream`. Returns 0 or a hipError_t. */
int cumayo_autotune(const uint8_t *prep_esk, const uint8_t *prep_epk, const int32_t *key_idx,
                    const uint8_t *msgs, const int64_t *msg_off, const int32_t *msg_len, uint8_t *sigs,
                    uint8_t *results, size_t n, uint8_t *sign_ws, uint8_t *verify_ws, hipStream_t stream,
                    int reps, cumayo_tune_report *out);

/* HIP-graph captures of cumayo_sign_prepared / cumayo_verify_prepared for
 * fixed buffers and batch size (the kernels of both stages, under the launch
 * configuration in force at capture; no memory set or copy nodes). Each
 * cumayo_graph_launch replays the whole pipeline on `stream` as one graph
 * launch, with the inputs' contents at that time. Create returns NULL on
 * invalid arguments or a capture error (cumayo_last_error). Replays of one
 * graph must not overlap each other or other work on the same buffers. */
typedef struct cumayo_graph cumayo_graph;
cumayo_graph *cumayo_sign_graph_create(const uint8_t *prepared, const int32_t *key_idx, const uint8_t *msgs,
                                       const int64_t *msg_off, const int32_t *msg_len, uint8_t *sigs, size_t n,
                                       uint8_t *workspace, int stages);
cumayo_graph *cumayo_verify_graph_create(const uint8_t *prepared, const int32_t *key_idx, const uint8_t *msgs,
                                         const int64_t *msg_off, const int32_t *msg_len, const uint8_t *sigs,
                                         uint8_t *results, size_t n, uint8_t *workspace, int stages);
int cumayo_graph_launch(cumayo_graph *g, hipStream_t stream);
void cumayo_graph_destroy(cumayo_graph *g);
/* One graph whose n child graphs are the pipelines `parts` (created by this
 * or another parameter set's library: the graph object is the same in each),
 * with no dependencies between them, so one cumayo_graph_launch runs them
 * all and they may overlap on the device: a batch that mixes parameter sets
 * as one graph launch. The children are copies; the parts stay valid and
 * are destroyed by their owner. NULL on invalid arguments or a HIP error
 * (cumayo_last_error). The parts' buffers must not be shared between
 * them. */
cumayo_graph *cumayo_graph_compose(cumayo_graph *const *parts, int n);

#ifdef __cplusplus
}
#endif

#endif
