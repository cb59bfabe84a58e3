// C ABI of the engine: the reference's pitched-pool entry points
// (api.cu:17-218) and the batched key-table extensions.
#include <hip/hip_runtime.h>

#include <atomic>
#include <cstring>
#include <mutex>
#include <vector>

#include "../../include/cumayo.h"
#include "mayo_kernels.hpp"
#include "mayo_wave.hpp"
#include "mayo_shared.hpp"

using namespace mayo;

namespace {

std::atomic<int> g_last_error{0};

void note(hipError_t e) {
    if (e != hipSuccess) g_last_error.store((int)e);
}
void note_launch() { note(hipGetLastError()); }

// Batch sizes above CUMAYO_MAX_BATCH are refused before anything is enqueued:
// a grid of n workgroups of 256 work-items must stay below 2^32 work-items.
bool batch_ok(size_t n) {
    if (n <= CUMAYO_MAX_BATCH) return true;
    note(hipErrorInvalidValue);
    return false;
}

constexpr size_t A128(size_t x) { return align128(x); }

// AES tables, computed from the field definition (FIPS 197 section 5.1.1):
// S-box = affine map of the GF(2^8) inverse; T0[x] = (2s, s, s, 3s).
uint8_t gf256_mul(uint8_t a, uint8_t b) {
    uint8_t r = 0;
    while (b) {
        if (b & 1) r ^= a;
        a = (uint8_t)((a << 1) ^ ((a & 0x80) ? 0x1B : 0));
        b >>= 1;
    }
    return r;
}

void build_aes_tables(uint32_t *T0, uint32_t *S) {
    for (int a = 0; a < 256; ++a) {
        uint8_t inv = 0;
        for (int b = 1; a && b < 256; ++b)
            if (gf256_mul((uint8_t)a, (uint8_t)b) == 1) { inv = (uint8_t)b; break; }
        uint8_t s = inv, x = inv;
        for (int i = 0; i < 4; ++i) { x = (uint8_t)((x << 1) | (x >> 7)); s ^= x; }
        s ^= 0x63;
        S[a] = s;
        T0[a] = ((uint32_t)gf256_mul(s, 2) << 24) | ((uint32_t)s << 16) | ((uint32_t)s << 8) |
                (uint32_t)gf256_mul(s, 3);
    }
}

std::mutex g_init_mu;
bool g_init_done[64] = {};

void ensure_tables() {
    int dev = 0;
    note(hipGetDevice(&dev));
    if (dev < 0 || dev >= 64) return;
    std::lock_guard<std::mutex> lk(g_init_mu);
    if (g_init_done[dev]) return;
    uint32_t T0[256], S[256];
    build_aes_tables(T0, S);
    note(hipMemcpyToSymbol(HIP_SYMBOL(c_aes_T0), T0, sizeof T0));
    note(hipMemcpyToSymbol(HIP_SYMBOL(c_aes_S), S, sizeof S));
    g_init_done[dev] = true;
}

inline unsigned blocks64(size_t n) { return (unsigned)((n + 63) / 64); }

// SIMDs of the current device (compute units x 4), looked up once per device
int device_simds() {
    static std::atomic<int> simds[64];
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 1024;
    int v = simds[dev].load();
    if (!v) {
        int cus = 0;
        if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0) cus = 256;
        v = 4 * cus;
        simds[dev].store(v);
    }
    return v;
}
inline unsigned order_grid(size_t n) { return (unsigned)((n + ORDER_SEG - 1) / ORDER_SEG); }

// Rows of `width` bytes from src (pitch spitch) to dst (pitch dpitch); one
// workgroup per group of rows, dwords when every address and pitch allows.
template <class T>
__global__ void __launch_bounds__(256) k_copy_rows(uint8_t *__restrict__ dst, size_t dpitch,
                                                   const uint8_t *__restrict__ src, size_t spitch,
                                                   size_t width, size_t rows, size_t rows_per_block) {
    const size_t r0 = (size_t)blockIdx.x * rows_per_block;
    const size_t r1 = r0 + rows_per_block < rows ? r0 + rows_per_block : rows;
    const size_t wn = width / sizeof(T);
    for (size_t r = r0; r < r1; ++r) {
        const T *s = (const T *)(src + r * spitch);
        T *d = (T *)(dst + r * dpitch);
        for (size_t b = threadIdx.x; b < wn; b += 256) d[b] = s[b];
    }
}

// Address at which the GPU reaches a host buffer: the buffer itself when it
// is pinned (page-locked and mapped at the same address), else null.
const void *device_view(const void *host) {
    hipPointerAttribute_t a;
    if (hipPointerGetAttributes(&a, host) != hipSuccess) {
        (void)hipGetLastError();  // pageable memory: not an error of the caller's
        return nullptr;
    }
    if (a.type != hipMemoryTypeHost || !a.devicePointer || a.devicePointer != a.hostPointer) return nullptr;
    return host;
}

// Pitched host <-> device copies of the entry points. hipMemcpy2DAsync moves
// one row per DMA transfer, which costs microseconds per row for the short
// rows of these pools; from pinned host memory a copy kernel reads or writes
// the host rows over PCIe directly instead. Pageable host memory goes
// through hipMemcpy2DAsync.
void copy2d(void *dst, size_t dpitch, const void *src, size_t spitch, size_t width, size_t rows,
            hipMemcpyKind kind, hipStream_t stream) {
    if (!rows || !width) return;
    const void *host = kind == hipMemcpyHostToDevice ? src : (const void *)dst;
    if (device_view(host)) {
        const size_t rpb = rows >= 4096 ? (rows + 4095) / 4096 : 1;
        const unsigned grid = (unsigned)((rows + rpb - 1) / rpb);
        const uintptr_t al = (uintptr_t)dst | (uintptr_t)src | dpitch | spitch | width;
        if ((al & 3) == 0)
            k_copy_rows<uint32_t><<<grid, 256, 0, stream>>>((uint8_t *)dst, dpitch, (const uint8_t *)src, spitch,
                                                             width, rows, rpb);
        else
            k_copy_rows<uint8_t><<<grid, 256, 0, stream>>>((uint8_t *)dst, dpitch, (const uint8_t *)src, spitch,
                                                            width, rows, rpb);
        note_launch();
        return;
    }
    note(hipMemcpy2DAsync(dst, dpitch, src, spitch, width, rows, kind, stream));
}

// crypto_sign_signWithesk signs without L (1) or expands L and signs with it
// as the reference does (0). Without L costs K (P1 entries + V O) MACs per
// signature against V^2 O for L, which pays for small K/O: over distinct
// keys (64-byte messages) MAYO_2 signs at 6.8 M/s without L against 4.8 M/s
// with it; MAYO_3 at 0.98 against 0.94 M/s; MAYO_1 (2.60 against 2.72 M/s)
// and MAYO_5 (0.29 against 0.35 M/s) are faster with L.
#ifndef MAYO_SIGNW_LFREE
#define MAYO_SIGNW_LFREE (MAYO_VERSION == 2)
#endif
constexpr int SIGNW_LFREE = MAYO_SIGNW_LFREE;

// Test hook: the first attempts of every item of the pitched signing kernels
// report a singular system (0 = off).
std::atomic<int> g_pitched_fail_first{0};

// Pool offsets of the sign layouts (api.cu:97-104, 125-132).
constexpr size_t SIGN_ESK_LEN = ESK_BYTES;
constexpr size_t SIGNW_MSG_OFF = A128(ESK_BYTES);
// Verify layouts (api.cu:147-151, 159-166).
constexpr size_t EPK_LEN = (size_t)P_BYTES + PK_BYTES - 16;
constexpr size_t EPK_PAD = A128(EPK_LEN);
constexpr size_t SIG_PAD = A128(SIG_BYTES);

// expandsk (expandsk.cu:37-81): sk -> sk || O || P1 || L in place. With
// with_L false the row keeps P2 where L would go (P1 || P2 is the keystream),
// for the L-free signing kernel.
void launch_expandsk(uint8_t *pool, size_t pitch, size_t n, hipStream_t stream, bool with_L = true) {
    if (!n || !batch_ok(n)) return;
    ensure_tables();
    const size_t p1_off = SK + O_BYTES;
    k_seed<<<blocks64(n), 64, 0, stream>>>(pool, pitch, n, 0, 0, 0, p1_off, SK);
    note_launch();
    launch_aes_expand(pool, pitch, n, p1_off, p1_off, stream);
    note_launch();
    if (!with_L) return;
    wv::k_expand_L_wave<<<(unsigned)n, 64, 0, stream>>>(pool, pitch, n);
    note_launch();
}

// Signing over a reference-layout pool; lfree: the rows hold P2 in L's place.
void launch_sign_pitched(uint8_t *pool, size_t pitch, size_t n, size_t msg_off, size_t sig_off,
                         size_t scr_off, size_t mlen, hipStream_t stream, bool lfree = false) {
    if (!n || !batch_ok(n)) return;
    PitchedSign acc{pool, pitch, msg_off, sig_off, scr_off, (uint32_t)mlen};
    k_sign_hash<PitchedSign><<<blocks64(n), 64, 0, stream>>>(acc, n);
    note_launch();
    const int ff = g_pitched_fail_first.load(std::memory_order_relaxed);
    if constexpr (wv::REF_WORD_ALIGNED) {
        if (((uintptr_t)pool | pitch) % 4 == 0) {
            if (lfree)
                wv::k_sign_wave<PitchedSign, wv::PitchedAlignedKeys, true><<<(unsigned)n, 64, 0, stream>>>(acc, n, ff);
            else
                wv::k_sign_wave<PitchedSign, wv::PitchedAlignedKeys><<<(unsigned)n, 64, 0, stream>>>(acc, n, ff);
            note_launch();
            return;
        }
    }
    if (lfree)
        wv::k_sign_wave<PitchedSign, wv::PitchedRefKeys, true><<<(unsigned)n, 64, 0, stream>>>(acc, n, ff);
    else
        wv::k_sign_wave<PitchedSign, wv::PitchedRefKeys><<<(unsigned)n, 64, 0, stream>>>(acc, n, ff);
    note_launch();
}

void launch_verify_pitched(uint8_t *pool, size_t pitch, size_t n, size_t mlen, hipStream_t stream) {
    if (!n || !batch_ok(n)) return;
    const size_t msg_off = EPK_PAD, sig_off = EPK_PAD + A128(mlen), res_off = sig_off + SIG_PAD;
    PitchedVerify acc{pool, pitch, msg_off, sig_off, res_off, (uint32_t)mlen};
    if constexpr (wv::REF_WORD_ALIGNED) {
        if (((uintptr_t)pool | pitch) % 4 == 0) {
            wv::k_verify_wave<PitchedVerify, wv::PitchedAlignedKeys><<<(unsigned)n, 64, 0, stream>>>(acc, n);
            note_launch();
            return;
        }
    }
    wv::k_verify_wave<PitchedVerify, wv::PitchedRefKeys><<<(unsigned)n, 64, 0, stream>>>(acc, n);
    note_launch();
}

// NIST wire format (sm = signature || message), one wavefront per item.
__global__ void __launch_bounds__(64) k_sm_pack(const uint8_t *__restrict__ sigs, const uint8_t *__restrict__ msgs,
                                                const int64_t *__restrict__ msg_off,
                                                const int32_t *__restrict__ msg_len, uint8_t *__restrict__ sm,
                                                const int64_t *__restrict__ sm_off) {
    const size_t i = blockIdx.x;
    uint8_t *out = sm + sm_off[i];
    const uint8_t *sg = sigs + i * (size_t)SIG_BYTES, *m = msgs + msg_off[i];
    for (int b = threadIdx.x; b < SIG_BYTES; b += 64) out[b] = sg[b];
    for (int b = threadIdx.x; b < msg_len[i]; b += 64) out[SIG_BYTES + b] = m[b];
}

__global__ void __launch_bounds__(64) k_sm_split(const uint8_t *__restrict__ sm, const int64_t *__restrict__ sm_off,
                                                 const int32_t *__restrict__ sm_len, uint8_t *__restrict__ sigs,
                                                 uint8_t *__restrict__ msgs, const int64_t *__restrict__ out_off,
                                                 int32_t *__restrict__ len) {
    const size_t i = blockIdx.x;
    const int32_t total = sm_len[i];
    const int32_t ml = total >= SIG_BYTES ? total - SIG_BYTES : 0;
    const uint8_t *in = sm + sm_off[i];
    uint8_t *sg = sigs + i * (size_t)SIG_BYTES, *m = msgs + out_off[i];
    for (int b = threadIdx.x; b < SIG_BYTES; b += 64) sg[b] = total >= SIG_BYTES ? in[b] : 0;
    for (int b = threadIdx.x; b < ml; b += 64) m[b] = in[SIG_BYTES + b];
    if (threadIdx.x == 0) len[i] = ml;
}

__global__ void k_sm_finish(const int32_t *__restrict__ sm_len, const int32_t *__restrict__ len,
                            uint8_t *__restrict__ results, int32_t *__restrict__ mlen_out, size_t n) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (sm_len[i] < SIG_BYTES) results[i] = 1;
    if (mlen_out) mlen_out[i] = results[i] == 0 ? len[i] : -1;
}

}  // namespace

extern "C" {

void aes128_tables_init(void) { ensure_tables(); }

void crypto_sign_keypair(uint8_t *pool, size_t pitch, size_t batch_size, hipStream_t stream,
                         size_t rand_index) {
    if (!batch_size || !batch_ok(batch_size)) return;
    ensure_tables();
    const size_t pk_off = A128(SK), p_off = pk_off + A128(PK_BYTES);
    // seed_pk lands in pk; O is parked in the P3 area until P3 replaces it
    k_seed<<<blocks64(batch_size), 64, 0, stream>>>(pool, pitch, batch_size, 1, rand_index, 0,
                                                    pk_off, pk_off + 16);
    note_launch();
    launch_aes_expand(pool, pitch, batch_size, pk_off, p_off, stream);
    note_launch();
    wv::k_keypair_P3_wave<<<(unsigned)batch_size, 64, 0, stream>>>(pool, pitch, batch_size, pk_off, p_off);
    note_launch();
}

void crypto_sign_keypair_DtH(uint8_t *pk, uint8_t *sk, uint8_t *pool, size_t pitch,
                             size_t batch_size, hipStream_t stream) {
    copy2d(pk, PK_BYTES, pool + A128(SK), pitch, PK_BYTES, batch_size, hipMemcpyDeviceToHost, stream);
    copy2d(sk, SK, pool, pitch, SK, batch_size, hipMemcpyDeviceToHost, stream);
}

void crypto_sign_expandsk_HtD(uint8_t *csk, uint8_t *pool, size_t pitch, size_t batch_size,
                              hipStream_t stream) {
    copy2d(pool, pitch, csk, SK, SK, batch_size, hipMemcpyHostToDevice, stream);
}

void crypto_sign_expandsk(uint8_t *pool, size_t pitch, size_t batch_size, hipStream_t stream) {
    launch_expandsk(pool, pitch, batch_size, stream);
}

void crypto_sign_expandsk_DtH(uint8_t *esk, uint8_t *pool, size_t pitch, size_t batch_size,
                              hipStream_t stream) {
    copy2d(esk, ESK_BYTES, pool, pitch, ESK_BYTES, batch_size, hipMemcpyDeviceToHost, stream);
}

void crypto_sign_expandpk_HtD(uint8_t *cpk, uint8_t *pool, size_t pitch, size_t batch_size,
                              hipStream_t stream) {
    copy2d(pool + P_BYTES - 16, pitch, cpk, PK_BYTES, PK_BYTES, batch_size, hipMemcpyHostToDevice,
           stream);
}

void crypto_sign_expandpk(uint8_t *pool, size_t pitch, size_t batch_size, hipStream_t stream) {
    if (!batch_size || !batch_ok(batch_size)) return;
    ensure_tables();
    launch_aes_expand(pool, pitch, batch_size, P_BYTES - 16, 0, stream);
    note_launch();
}

void crypto_sign_expandpk_DtH(uint8_t *epk, uint8_t *pool, size_t pitch, size_t batch_size,
                              hipStream_t stream) {
    copy2d(epk, EPK_LEN, pool, pitch, EPK_LEN, batch_size, hipMemcpyDeviceToHost, stream);
}

void crypto_sign_sign_HtD(uint8_t *cesk, uint8_t *cMess, size_t mlen, uint8_t *pool, size_t pitch,
                          size_t batch_size, hipStream_t stream) {
    copy2d(pool, pitch, cesk, SIGN_ESK_LEN, SIGN_ESK_LEN, batch_size, hipMemcpyHostToDevice, stream);
    copy2d(pool + SIGN_ESK_LEN, pitch, cMess, mlen, mlen, batch_size, hipMemcpyHostToDevice, stream);
}

void crypto_sign_sign(uint8_t *pool, size_t Mlen, size_t pitch, size_t batch_size,
                      hipStream_t stream) {
    const size_t msg_off = SIGN_ESK_LEN, sig_off = msg_off + Mlen, scr_off = sig_off + SIG_BYTES;
    launch_sign_pitched(pool, pitch, batch_size, msg_off, sig_off, scr_off, Mlen, stream);
}

void crypto_sign_sign_DtH(uint8_t *sig, uint8_t *pool, size_t Mlen, size_t pitch, size_t batch_size,
                          hipStream_t stream) {
    copy2d(sig, SIG_BYTES, pool + SIGN_ESK_LEN + Mlen, pitch, SIG_BYTES, batch_size,
           hipMemcpyDeviceToHost, stream);
}

void crypto_sign_signWithesk_HtD(uint8_t *csk, uint8_t *cMess, size_t mlen, uint8_t *pool,
                                 size_t pitch, size_t batch_size, hipStream_t stream) {
    copy2d(pool, pitch, csk, SK, SK, batch_size, hipMemcpyHostToDevice, stream);
    copy2d(pool + SIGNW_MSG_OFF, pitch, cMess, mlen, mlen, batch_size, hipMemcpyHostToDevice, stream);
}

void crypto_sign_signWithesk(uint8_t *pool, size_t Mlen, size_t pitch, size_t batch_size,
                             hipStream_t stream) {
    // the key is expanded to sk || O || P1 || P2 and signed without L
    launch_expandsk(pool, pitch, batch_size, stream, SIGNW_LFREE == 0);
    const size_t msg_off = SIGNW_MSG_OFF, sig_off = msg_off + A128(Mlen), scr_off = sig_off + SIG_PAD;
    launch_sign_pitched(pool, pitch, batch_size, msg_off, sig_off, scr_off, Mlen, stream, SIGNW_LFREE != 0);
}

void crypto_sign_signWithesk_DtH(uint8_t *sig, uint8_t *pool, size_t Mlen, size_t pitch,
                                 size_t batch_size, hipStream_t stream) {
    copy2d(sig, SIG_BYTES, pool + SIGNW_MSG_OFF + A128(Mlen), pitch, SIG_BYTES, batch_size,
           hipMemcpyDeviceToHost, stream);
}

void crypto_sign_verify_HtD(uint8_t *epk, uint8_t *Mess, uint8_t *sig, size_t mlen, uint8_t *pool,
                            size_t pitch, size_t batch_size, hipStream_t stream) {
    copy2d(pool, pitch, epk, EPK_PAD, EPK_PAD, batch_size, hipMemcpyHostToDevice, stream);
    copy2d(pool + EPK_PAD, pitch, Mess, mlen, mlen, batch_size, hipMemcpyHostToDevice, stream);
    copy2d(pool + EPK_PAD + A128(mlen), pitch, sig, SIG_PAD, SIG_PAD, batch_size,
           hipMemcpyHostToDevice, stream);
}

void crypto_sign_verify(uint8_t *pool, size_t Mlen, size_t pitch, size_t batch_size,
                        hipStream_t stream) {
    launch_verify_pitched(pool, pitch, batch_size, Mlen, stream);
}

void crypto_sign_verify_DtH(uint8_t *verify, size_t Mlen, uint8_t *pool, size_t pitch,
                            size_t batch_size, hipStream_t stream) {
    copy2d(verify, 1, pool + EPK_PAD + A128(Mlen) + SIG_PAD, pitch, 1, batch_size,
           hipMemcpyDeviceToHost, stream);
}

void crypto_sign_verifyWithepk_HtD(uint8_t *cpk, uint8_t *Mess, uint8_t *sig, size_t mlen,
                                   uint8_t *pool, size_t pitch, size_t batch_size,
                                   hipStream_t stream) {
    copy2d(pool + P_BYTES - 16, pitch, cpk, PK_BYTES, PK_BYTES, batch_size, hipMemcpyHostToDevice,
           stream);
    copy2d(pool + EPK_PAD, pitch, Mess, mlen, mlen, batch_size, hipMemcpyHostToDevice, stream);
    copy2d(pool + EPK_PAD + A128(mlen), pitch, sig, SIG_BYTES, SIG_BYTES, batch_size,
           hipMemcpyHostToDevice, stream);
}

void crypto_sign_verifyWithepk(uint8_t *pool, size_t Mlen, size_t pitch, size_t batch_size,
                               hipStream_t stream) {
    if (!batch_size || !batch_ok(batch_size)) return;
    ensure_tables();
    launch_aes_expand(pool, pitch, batch_size, P_BYTES - 16, 0, stream);
    note_launch();
    launch_verify_pitched(pool, pitch, batch_size, Mlen, stream);
}

void crypto_sign_verifyWithepk_DtH(uint8_t *verify, size_t Mlen, uint8_t *pool, size_t pitch,
                                   size_t batch_size, hipStream_t stream) {
    crypto_sign_verify_DtH(verify, Mlen, pool, pitch, batch_size, stream);
}

void cumayo_get_sizes(cumayo_sizes *out) {
    out->version = MAYO_VERSION;
    out->n = N; out->m = M; out->o = O; out->k = K;
    out->sk_bytes = SK; out->pk_bytes = PK_BYTES; out->sig_bytes = SIG_BYTES;
    out->esk_bytes = ESK_BYTES; out->epk_bytes = EPK_BYTES; out->digest_bytes = DIGEST;
    out->p1_bytes = P1_BYTES; out->p2_bytes = P2_BYTES; out->p3_bytes = P3_BYTES;
    out->o_bytes = O_BYTES;
    out->sign_ws_stride = SIGN_SCRATCH_STRIDE;
    out->sign_ws_attempts = SCR_ATTEMPTS;
}

int cumayo_last_error(void) { return g_last_error.exchange(0); }

int cumayo_signwithesk_lfree(void) { return SIGNW_LFREE != 0; }

void cumayo_test_pitched_fail_first(int attempts) {
    g_pitched_fail_first.store(attempts > 0 ? (attempts < MAX_ATTEMPTS ? attempts : MAX_ATTEMPTS) : 0);
}

size_t cumayo_sign_workspace_size(size_t n) {
    return n * ((size_t)SIGN_SCRATCH_STRIDE + SIGN_T_BYTES) + wv::key_sort_bytes(n);
}

}  // extern "C"

namespace {

// Launch configuration of the batch form (process-wide for this library).
// Over prepared keys, the compiled variant of each wave kernel (quadratic
// form: 10 x rows per lane + waves per SIMD; solve: waves per SIMD) and the
// split of signing's linear-algebra pass into one or two parts; for every
// table form, the item order of the one-wave-per-item kernels (0 batch
// order, 1 XCD-aware key order).
struct LaunchCfg {
    int prep, solve, sign_order, verify, verify_order, sign_chunks;
};
std::atomic<int> g_prep{wv::PREP_OCCS[0]}, g_solve{wv::SOLVE_OCCS[0]}, g_sign_order{0},
    g_verify{wv::VERIFY_OCCS[0]}, g_verify_order{0}, g_sign_chunks{1};

LaunchCfg current_cfg() {
    return {g_prep.load(),   g_solve.load(),        g_sign_order.load(),
            g_verify.load(), g_verify_order.load(), g_sign_chunks.load()};
}

// Stage 2 of signing in two parts needs at least this many items (each
// part then still holds several rounds of waves per SIMD).
constexpr size_t CHUNK_MIN = 16384;

// Helper streams of two-part passes, per calling thread, keyed by (device,
// caller stream): passes issued from one thread on different streams (the
// server's slots, say) get different helpers and stay independent. Each
// thread keeps at most HELPERS of them, dropping the oldest (a helper is
// only ever joined through events, so destroying it with work still queued
// is safe: HIP releases it once that work is done), and destroys the rest
// when it exits. Helpers run at the default priority.
class HelperStreams {
  public:
    ~HelperStreams() {
        for (auto &e : v_) (void)hipStreamDestroy(e.helper);
    }
    hipStream_t get(hipStream_t caller) {
        int dev = 0;
        if (hipGetDevice(&dev) != hipSuccess) return nullptr;
        for (auto &e : v_)
            if (e.dev == dev && e.caller == caller) return e.helper;
        hipStream_t h = nullptr;
        if (hipStreamCreateWithFlags(&h, hipStreamNonBlocking) != hipSuccess) return nullptr;
        if (v_.size() >= HELPERS) {
            (void)hipStreamDestroy(v_.front().helper);
            v_.erase(v_.begin());
        }
        v_.push_back({dev, caller, h});
        return h;
    }

  private:
    static constexpr size_t HELPERS = 16;
    struct Entry {
        int dev;
        hipStream_t caller, helper;
    };
    std::vector<Entry> v_;
};

hipStream_t aux_stream(hipStream_t caller) {
    thread_local HelperStreams helpers;
    return helpers.get(caller);
}

// Stage 2 of signing: attempt 0 in two kernels tuned for their register
// needs (prep_k, solve_k; with the shared-key tables prep_k writes T rows
// and fold_k folds them into y between the two), then the retry pass over the
// items whose attempt met a singular system (about 1 in 4,000). Over the
// whole batch, or, with
// parts = 2, over two halves, the second on the helper stream after a fork
// from `stream` and joined back at the end: the kernels of one half fill the
// other half's tails, and each half's retry pass (over the half's own
// failure list) runs beside the other half's kernels.
using FoldK = void (*)(TableSign, size_t, size_t, size_t);

template <class KL, class PK, class SK>
hipError_t launch_sign_stage2(PK prep_k, int prep_w, FoldK fold_k, SK solve_k, int solve_w, const TableSign &acc,
                              size_t n, int ordered, int fail_first, int parts, hipStream_t stream) {
    const unsigned retry_threads = 64 * wv::RNW;
    // prep, (fold,) solve and retry over the positions [base, base + m) on s
    auto part = [&](size_t base, size_t m, hipStream_t s) {
        hipLaunchKernelGGL(prep_k, dim3((unsigned)((m + prep_w - 1) / prep_w)), dim3(64 * prep_w), 0, s, acc, n,
                           ordered, base, base + m);
        if (fold_k)
            hipLaunchKernelGGL(fold_k, dim3((unsigned)((m + wv::FOLD_W - 1) / wv::FOLD_W)), dim3(64 * wv::FOLD_W),
                               0, s, acc, n, base, base + m);
        hipLaunchKernelGGL(solve_k, dim3((unsigned)((m + solve_w - 1) / solve_w)), dim3(64 * solve_w), 0, s, acc, n,
                           fail_first, ordered, base, base + m);
        hipLaunchKernelGGL((wv::k_sign_retry<TableSign, KL>), dim3((unsigned)(m < 256 ? m : 256)), dim3(retry_threads), 0, s,
                           acc, n, fail_first, base);
    };
    hipStream_t aux = parts > 1 ? aux_stream(stream) : nullptr;
    hipEvent_t fork = nullptr, join = nullptr;
    if (parts > 1 && aux && hipEventCreateWithFlags(&fork, hipEventDisableTiming) == hipSuccess &&
        hipEventCreateWithFlags(&join, hipEventDisableTiming) == hipSuccess) {
        hipError_t e = hipEventRecord(fork, stream);
        if (e == hipSuccess) e = hipStreamWaitEvent(aux, fork, 0);
        const size_t h = n / 2;
        if (e == hipSuccess) {
            part(0, h, stream);
            part(h, n - h, aux);
            e = hipEventRecord(join, aux);
            if (e == hipSuccess) e = hipStreamWaitEvent(stream, join, 0);
        }
        (void)hipEventDestroy(fork);
        (void)hipEventDestroy(join);
        if (e != hipSuccess) return e;
    } else {
        if (fork) (void)hipEventDestroy(fork);
        part(0, n, stream);
    }
    return hipGetLastError();
}

template <size_t N>
bool listed(const int (&arr)[N], int v) {
    for (size_t i = 0; i < N; ++i)
        if (arr[i] == v) return true;
    return false;
}

// f(integral_constant<int, ARR[i]>) for the entry ARR[i] == want; false when
// no compiled variant matches
template <const auto &ARR, size_t I = 0, class F>
bool with_occ(int want, F &&f) {
    if constexpr (I < sizeof(ARR) / sizeof(ARR[0])) {
        if (ARR[I] == want) {
            f(std::integral_constant<int, ARR[I]>{});
            return true;
        }
        return with_occ<ARR, I + 1>(want, f);
    }
    return false;
}

template <class KL>
int sign_table_impl(const TableSign &acc, size_t n, hipStream_t stream, int stages, const LaunchCfg &cfg) {
    hipError_t e = hipSuccess;
    if (stages & 1) {
        // lanes of a hashing wave take messages of similar length
        if (n > 64) {
            k_length_order<TableSign><<<order_grid(n), 1024, 0, stream>>>(acc, n);
            k_sign_hash<TableSign, true><<<blocks64(n), 64, 0, stream>>>(acc, n);
        } else {
            k_sign_hash<TableSign><<<blocks64(n), 64, 0, stream>>>(acc, n);
        }
        e = hipGetLastError();
    }
    if (e == hipSuccess && (stages & 2)) {
        if (stages & 4) {
            k_sign_main<TableSign><<<(unsigned)n, WG, 0, stream>>>(acc, n);
        } else {
            const int fail_first = (stages >> 8) & 0xFF;
            // shared-key tables (prepared keys, split launches, a 16-byte
            // aligned T region in the workspace): items in key order, several
            // per workgroup
            const bool shared = std::is_same<KL, wv::PrepKeys>::value && cfg.sign_chunks != 0 &&
                                wv::sk_code(cfg.prep) && acc.trows && !((uintptr_t)acc.trows & 15);
            const int ordered = shared || (cfg.sign_order && n > 64);
            const int parts = cfg.sign_chunks > 1 && n >= CHUNK_MIN ? 2 : 1;
            if (shared)
                wv::launch_key_sort(acc, n, (uint32_t *)(acc.trows + n * SIGN_T_BYTES), stream);
            else if (ordered)
                k_key_order<TableSign><<<1, 1024, 0, stream>>>(acc, n);
            if (cfg.sign_chunks == 0) {
                // one persistent fused launch (prepared or reference tables)
                bool found = false;
                const bool own = std::is_same<KL, wv::PrepKeys>::value && !wv::sk_code(cfg.prep);
                // (the shared-key codes stand for PREP_OCCS[0] here)
                with_occ<wv::PREP_OCCS>(own ? cfg.prep : wv::PREP_OCCS[0], [&](auto P) {
                    constexpr int code = decltype(P)::value;
                    found = true;
                    const size_t resident = (size_t)device_simds() * (code % 10);
                    const unsigned grid = (unsigned)(n < resident ? n : resident);
                    hipLaunchKernelGGL((wv::k_sign_fused<TableSign, KL, code>), dim3(grid), dim3(64), 0, stream,
                                       acc, n, fail_first, ordered);
                });
                if (!found) {
                    note(hipErrorInvalidValue);
                    return (int)hipErrorInvalidValue;
                }
            } else if constexpr (std::is_same<KL, wv::PrepKeys>::value) {
                // prepared keys: the variants the launch autotuner chooses among
                bool found = false;
                if (shared) {
                    with_occ<wv::SK_CODES>(cfg.prep, [&](auto P) {
                        with_occ<wv::SOLVE_OCCS>(cfg.solve, [&](auto Q) {
                            found = true;
                            e = launch_sign_stage2<KL>(&wv::k_sign_table_shared<TableSign, decltype(P)::value>,
                                                       wv::sk_w(decltype(P)::value), &wv::k_sign_fold<TableSign>,
                                                       &wv::k_sign_solve<TableSign, KL::PLANAR, decltype(Q)::value>,
                                                       wv::solve_sigs(decltype(Q)::value), acc, n, ordered,
                                                       fail_first, parts, stream);
                        });
                    });
                } else {
                    // a shared-key code without a T region runs the wave kernel
                    with_occ<wv::PREP_OCCS>(wv::sk_code(cfg.prep) ? wv::PREP_OCCS[0] : cfg.prep, [&](auto P) {
                        with_occ<wv::SOLVE_OCCS>(cfg.solve, [&](auto Q) {
                            found = true;
                            e = launch_sign_stage2<KL>(&wv::k_sign_prep<TableSign, KL, decltype(P)::value>, 1,
                                                       nullptr, &wv::k_sign_solve<TableSign, KL::PLANAR, decltype(Q)::value>,
                                                       wv::solve_sigs(decltype(Q)::value), acc, n, ordered,
                                                       fail_first, parts, stream);
                        });
                    });
                }
                if (!found) {
                    note(hipErrorInvalidValue);
                    return (int)hipErrorInvalidValue;
                }
            } else {
                e = launch_sign_stage2<KL>(&wv::k_sign_prep<TableSign, KL>, 1, nullptr,
                                           &wv::k_sign_solve<TableSign, KL::PLANAR>, 1, acc, n, ordered, fail_first,
                                           parts, stream);
            }
        }
        if (e == hipSuccess) e = hipGetLastError();
    }
    note(e);
    return (int)e;
}

template <class KL>
int verify_table_impl(const TableVerify &acc, size_t n, hipStream_t stream, int stages, const LaunchCfg &cfg) {
    hipError_t e = hipSuccess;
    if (stages & 1) {
        if (n > 64) {
            k_length_order<TableVerify><<<order_grid(n), 1024, 0, stream>>>(acc, n);
            k_verify_hash<TableVerify, true><<<blocks64(n), 64, 0, stream>>>(acc, n);
        } else {
            k_verify_hash<TableVerify><<<blocks64(n), 64, 0, stream>>>(acc, n);
        }
        e = hipGetLastError();
    }
    if (e == hipSuccess && (stages & 2)) {
        if (stages & 4) {
            k_verify_main<TableVerify><<<(unsigned)n, WG, 0, stream>>>(acc, n);
        } else {
            const bool shared = std::is_same<KL, wv::PrepKeys>::value && wv::sk_code(cfg.verify) && acc.trows &&
                                !((uintptr_t)acc.trows & 15);
            const int ordered = !shared && cfg.verify_order && n > 64;
            if (shared)
                wv::launch_key_sort(acc, n, (uint32_t *)(acc.trows + n * VERIFY_T_BYTES), stream);
            else if (ordered)
                k_key_order<TableVerify><<<1, 1024, 0, stream>>>(acc, n);
            if constexpr (std::is_same<KL, wv::PrepKeys>::value) {
                const bool found = shared ? with_occ<wv::SK_CODES>(cfg.verify, [&](auto O) {
                    constexpr int W = wv::sk_w(decltype(O)::value);
                    wv::k_verify_table_shared<TableVerify, decltype(O)::value>
                        <<<(unsigned)((n + W - 1) / W), 64 * W, 0, stream>>>(acc, n);
                    wv::k_verify_fold<TableVerify>
                        <<<(unsigned)((n + wv::FOLD_W - 1) / wv::FOLD_W), 64 * wv::FOLD_W, 0, stream>>>(acc, n);
                }) : with_occ<wv::VERIFY_OCCS>(wv::sk_code(cfg.verify) ? wv::VERIFY_OCCS[0] : cfg.verify,
                                               [&](auto O) {
                    wv::k_verify_wave<TableVerify, KL, decltype(O)::value>
                        <<<(unsigned)n, 64, 0, stream>>>(acc, n, ordered);
                });
                if (!found) {
                    note(hipErrorInvalidValue);
                    return (int)hipErrorInvalidValue;
                }
            } else {
                wv::k_verify_wave<TableVerify, KL><<<(unsigned)n, 64, 0, stream>>>(acc, n, ordered);
            }
        }
        e = hipGetLastError();
    }
    note(e);
    return (int)e;
}

int prepare_keys(const uint8_t *src, size_t src_stride, size_t n_keys, uint8_t *dst, uint32_t dst_bytes,
                 uint32_t hdr, wv::PrepRegion r0, wv::PrepRegion r1, wv::PrepRegion r2,
                 hipStream_t stream) {
    if (!n_keys) return 0;
    if (!src || !dst || ((uintptr_t)dst & 15) || n_keys > 65535) return (int)hipErrorInvalidValue;
    const dim3 grid((dst_bytes / 4 + 255) / 256, (unsigned)n_keys);
    wv::k_prepare_keys<<<grid, 256, 0, stream>>>(src, src_stride, dst, dst_bytes, dst_bytes, hdr, r0, r1,
                                                   r2);
    const hipError_t e = hipGetLastError();
    note(e);
    return (int)e;
}

}  // namespace

extern "C" {

int cumayo_sign_table_stages(const uint8_t *esk_table, size_t esk_stride, const int32_t *key_idx,
                             const uint8_t *msgs, const int64_t *msg_off, const int32_t *msg_len,
                             uint8_t *sigs, size_t n, uint8_t *workspace, hipStream_t stream,
                             int stages) {
    if (!n) return 0;
    if (!batch_ok(n)) return (int)hipErrorInvalidValue;
    if (!esk_table || !key_idx || !msgs || !msg_off || !msg_len || !sigs || !workspace)
        return (int)hipErrorInvalidValue;
    TableSign acc{esk_table, esk_stride, key_idx, msgs, msg_off, msg_len, sigs, workspace};
    return sign_table_impl<wv::RefKeys>(acc, n, stream, stages, current_cfg());
}

size_t cumayo_prepared_esk_bytes(void) { return wv::PrepKeys::ESK_BYTES; }
size_t cumayo_prepared_epk_bytes(void) { return wv::PrepKeys::EPK_BYTES; }

int cumayo_prepare_esk(const uint8_t *esk_table, size_t esk_stride, size_t n_keys, uint8_t *prepared,
                       hipStream_t stream) {
    using P = wv::PrepKeys;
    return prepare_keys(esk_table, esk_stride, n_keys, prepared, P::ESK_BYTES, SK + O_BYTES,
                        {SK + O_BYTES, P::ESK_P1, P1_ENTRIES},
                        {SK + O_BYTES + P1_BYTES, P::ESK_L, P2_ENTRIES}, {0, 0, 0}, stream);
}

int cumayo_prepare_epk(const uint8_t *epk_table, size_t epk_stride, size_t n_keys, uint8_t *prepared,
                       hipStream_t stream) {
    using P = wv::PrepKeys;
    return prepare_keys(epk_table, epk_stride, n_keys, prepared, P::EPK_BYTES, 0, {0, P::EPK_P1, P1_ENTRIES},
                        {P1_BYTES, P::EPK_P2, P2_ENTRIES}, {P1_BYTES + P2_BYTES, P::EPK_P3, P3_ENTRIES},
                        stream);
}

}  // extern "C"

namespace {
// cumayo_sign_prepared over per-item slots at `workspace` and T regions at
// `trows` (none: the wave kernels run whatever the launch configuration)
int sign_prepared(const uint8_t *prepared, const int32_t *key_idx, const uint8_t *msgs, const int64_t *msg_off,
                  const int32_t *msg_len, uint8_t *sigs, size_t n, uint8_t *workspace, uint8_t *trows,
                  hipStream_t stream, int stages) {
    if (!n) return 0;
    if (!batch_ok(n)) return (int)hipErrorInvalidValue;
    if (!prepared || ((uintptr_t)prepared & 15) || !key_idx || !msgs || !msg_off || !msg_len || !sigs ||
        !workspace || (stages & 4))
        return (int)hipErrorInvalidValue;
    TableSign acc{prepared, wv::PrepKeys::ESK_BYTES, key_idx, msgs, msg_off, msg_len, sigs, workspace, trows};
    return sign_table_impl<wv::PrepKeys>(acc, n, stream, stages, current_cfg());
}
}  // namespace

extern "C" {

int cumayo_sign_prepared(const uint8_t *prepared, const int32_t *key_idx, const uint8_t *msgs,
                         const int64_t *msg_off, const int32_t *msg_len, uint8_t *sigs, size_t n,
                         uint8_t *workspace, hipStream_t stream, int stages) {
    uint8_t *trows = workspace ? workspace + n * (size_t)SIGN_SCRATCH_STRIDE : nullptr;
    return sign_prepared(prepared, key_idx, msgs, msg_off, msg_len, sigs, n, workspace, trows, stream, stages);
}

// cumayo_sign_prepared with its workspace in `scratch`, a file-scope device
// array of CUMAYO_NOALLOC_MAX_SIGN items' workspace slots: no allocation,
// for callers that cannot allocate. Every GPU holds its own copy of the
// array (until the process ends) at an address that hipGetSymbolAddress
// returns with that GPU current; the name `scratch` in host code is only a
// host-side stand-in, so host code uses scratch_ptr() and never the name.
// Nothing in the workspace needs clearing between calls (the kernels reset
// the failure-list count themselves and write every slot before reading
// it); calls that overlap on one GPU share the array and race.
static __device__ __align__(256) uint8_t scratch[(size_t)CUMAYO_NOALLOC_MAX_SIGN * SIGN_SCRATCH_STRIDE];

namespace {
std::mutex g_scratch_mu;
uint8_t *g_scratch_ptr[64] = {};

// Device address of `scratch` on the current GPU, looked up once per GPU;
// null (error noted) if the lookup fails.
uint8_t *scratch_ptr() {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess || dev < 0 || dev >= 64) {
        note(e != hipSuccess ? e : hipErrorInvalidDevice);
        return nullptr;
    }
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            note(e != hipSuccess ? e : hipErrorInvalidSymbol);
            return nullptr;
        }
        g_scratch_ptr[dev] = static_cast<uint8_t *>(p);
    }
    return g_scratch_ptr[dev];
}
}  // namespace

int cumayo_sign_prepared_noalloc(const uint8_t *prepared, const int32_t *key_idx, const uint8_t *msgs,
                                 const int64_t *msg_off, const int32_t *msg_len, uint8_t *sigs, size_t n,
                                 hipStream_t stream) {
    if (!n) return 0;
    if (n > CUMAYO_NOALLOC_MAX_SIGN) {
        note(hipErrorInvalidValue);
        return (int)hipErrorInvalidValue;
    }
    uint8_t *ws = scratch_ptr();
    if (!ws) return (int)hipErrorInvalidSymbol;
    // one launch part (n is far below the two-part threshold): no helper
    // stream or event is created; no T regions, so the wave kernels run
    static_assert(CUMAYO_NOALLOC_MAX_SIGN < CHUNK_MIN, "the no-allocation path runs stage 2 in one part");
    return sign_prepared(prepared, key_idx, msgs, msg_off, msg_len, sigs, n, ws, nullptr, stream, 3);
}

int cumayo_verify_prepared(const uint8_t *prepared, const int32_t *key_idx, const uint8_t *msgs,
                           const int64_t *msg_off, const int32_t *msg_len, const uint8_t *sigs,
                           uint8_t *results, size_t n, uint8_t *workspace, hipStream_t stream,
                           int stages) {
    if (!n) return 0;
    if (!batch_ok(n)) return (int)hipErrorInvalidValue;
    if (!prepared || ((uintptr_t)prepared & 15) || !key_idx || !msgs || !msg_off || !msg_len || !sigs ||
        !results || !workspace || (stages & 4))
        return (int)hipErrorInvalidValue;
    TableVerify acc{prepared, wv::PrepKeys::EPK_BYTES, key_idx, msgs, msg_off, msg_len, sigs, results,
                    workspace, workspace + n * (size_t)VERIFY_SCRATCH_STRIDE};
    return verify_table_impl<wv::PrepKeys>(acc, n, stream, stages, current_cfg());
}

int cumayo_sign_table(const uint8_t *esk_table, size_t esk_stride, const int32_t *key_idx,
                      const uint8_t *msgs, const int64_t *msg_off, const int32_t *msg_len,
                      uint8_t *sigs, size_t n, uint8_t *workspace, hipStream_t stream) {
    return cumayo_sign_table_stages(esk_table, esk_stride, key_idx, msgs, msg_off, msg_len, sigs, n,
                                    workspace, stream, 3);
}

size_t cumayo_verify_workspace_size(size_t n) {
    return n * ((size_t)VERIFY_SCRATCH_STRIDE + VERIFY_T_BYTES) + wv::key_sort_bytes(n);
}

int cumayo_verify_table_stages(const uint8_t *epk_table, size_t epk_stride, const int32_t *key_idx,
                               const uint8_t *msgs, const int64_t *msg_off, const int32_t *msg_len,
                               const uint8_t *sigs, uint8_t *results, size_t n, uint8_t *workspace,
                               hipStream_t stream, int stages) {
    if (!n) return 0;
    if (!batch_ok(n)) return (int)hipErrorInvalidValue;
    if (!epk_table || !key_idx || !msgs || !msg_off || !msg_len || !sigs || !results || !workspace)
        return (int)hipErrorInvalidValue;
    TableVerify acc{epk_table, epk_stride, key_idx, msgs, msg_off, msg_len, sigs, results, workspace};
    return verify_table_impl<wv::RefKeys>(acc, n, stream, stages, current_cfg());
}

int cumayo_verify_table(const uint8_t *epk_table, size_t epk_stride, const int32_t *key_idx,
                        const uint8_t *msgs, const int64_t *msg_off, const int32_t *msg_len,
                        const uint8_t *sigs, uint8_t *results, size_t n, uint8_t *workspace,
                        hipStream_t stream) {
    return cumayo_verify_table_stages(epk_table, epk_stride, key_idx, msgs, msg_off, msg_len, sigs,
                                      results, n, workspace, stream, 3);
}

size_t cumayo_crypto_sign_workspace_size(size_t n) {
    return cumayo_sign_workspace_size(n) + n * (size_t)SIG_BYTES;
}

int cumayo_crypto_sign(const uint8_t *prepared, const int32_t *key_idx, const uint8_t *msgs,
                       const int64_t *msg_off, const int32_t *msg_len, uint8_t *sm, const int64_t *sm_off,
                       size_t n, uint8_t *workspace, hipStream_t stream) {
    if (!n) return 0;
    if (!batch_ok(n)) return (int)hipErrorInvalidValue;
    if (!sm || !sm_off) return (int)hipErrorInvalidValue;
    uint8_t *sigs = workspace + cumayo_sign_workspace_size(n);
    int e = cumayo_sign_prepared(prepared, key_idx, msgs, msg_off, msg_len, sigs, n, workspace, stream, 3);
    if (e) return e;
    k_sm_pack<<<(unsigned)n, 64, 0, stream>>>(sigs, msgs, msg_off, msg_len, sm, sm_off);
    const hipError_t le = hipGetLastError();
    note(le);
    return (int)le;
}

size_t cumayo_crypto_sign_open_workspace_size(size_t n) {
    return cumayo_verify_workspace_size(n) + n * (sizeof(int32_t) + (size_t)SIG_BYTES);
}

int cumayo_crypto_sign_open(const uint8_t *prepared, const int32_t *key_idx, const uint8_t *sm,
                            const int64_t *sm_off, const int32_t *sm_len, uint8_t *msgs_out,
                            const int64_t *out_off, int32_t *mlen_out, uint8_t *results, size_t n,
                            uint8_t *workspace, hipStream_t stream) {
    if (!n) return 0;
    if (!batch_ok(n)) return (int)hipErrorInvalidValue;
    if (!prepared || ((uintptr_t)prepared & 15) || !key_idx || !sm || !sm_off || !sm_len || !msgs_out ||
        !out_off || !results || !workspace)
        return (int)hipErrorInvalidValue;
    int32_t *len = (int32_t *)(workspace + cumayo_verify_workspace_size(n));
    uint8_t *sigs = (uint8_t *)(len + n);
    k_sm_split<<<(unsigned)n, 64, 0, stream>>>(sm, sm_off, sm_len, sigs, msgs_out, out_off, len);
    hipError_t le = hipGetLastError();
    if (le != hipSuccess) { note(le); return (int)le; }
    int e = cumayo_verify_prepared(prepared, key_idx, msgs_out, out_off, len, sigs, results, n, workspace, stream, 3);
    if (e) return e;
    k_sm_finish<<<blocks64(n), 64, 0, stream>>>(sm_len, len, results, mlen_out, n);
    le = hipGetLastError();
    note(le);
    return (int)le;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Launch configuration, launch autotuner and captured pipelines (C ABI)
// ---------------------------------------------------------------------------
struct cumayo_graph {
    hipGraph_t graph;
    hipGraphExec_t exec;
};

namespace {

// Records launch(s) on a private stream in capture mode into an
// instantiated graph; null (and the error noted) on failure.
template <class F>
cumayo_graph *capture(F &&launch) {
    hipStream_t s = nullptr;
    hipError_t e = hipStreamCreateWithFlags(&s, hipStreamNonBlocking);
    if (e != hipSuccess) {
        note(e);
        return nullptr;
    }
    // a two-part signing pass forks onto the helper stream of (this thread,
    // s): create it before capture starts
    (void)aux_stream(s);
    hipGraph_t g = nullptr;
    e = hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal);
    if (e == hipSuccess) {
        const int le = launch(s);
        const hipError_t ee = hipStreamEndCapture(s, &g);
        e = le ? (hipError_t)le : ee;
    }
    (void)hipStreamDestroy(s);
    hipGraphExec_t x = nullptr;
    if (e == hipSuccess && g) e = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
    if (e != hipSuccess || !g) {
        if (g) (void)hipGraphDestroy(g);
        note(e != hipSuccess ? e : hipErrorUnknown);
        return nullptr;
    }
    return new cumayo_graph{g, x};
}

// Least duration in ms of launch() over reps timed runs after one untimed
// run, between two events on `stream`; *err gets the first error.
template <class F>
double time_min_ms(hipStream_t stream, hipEvent_t a, hipEvent_t b, int reps, F &&launch, int *err) {
    double best = 1e30;
    for (int r = -1; r < reps && !*err; ++r) {
        hipError_t he = hipEventRecord(a, stream);
        const int le = he == hipSuccess ? launch() : (int)he;
        he = hipEventRecord(b, stream);
        if (he == hipSuccess) he = hipEventSynchronize(b);
        float ms = 0;
        if (he == hipSuccess) he = hipEventElapsedTime(&ms, a, b);
        if (le || he != hipSuccess) {
            *err = le ? le : (int)he;
            note((hipError_t)*err);
            break;
        }
        if (r >= 0 && ms < best) best = ms;
    }
    return best;
}

}  // namespace

extern "C" {

void cumayo_get_launch_config(cumayo_launch_config *out) {
    if (!out) return;
    const LaunchCfg c = current_cfg();
    out->sign_prep = c.prep;
    out->sign_solve = c.solve;
    out->sign_order = c.sign_order;
    out->verify = c.verify;
    out->verify_order = c.verify_order;
    out->sign_chunks = c.sign_chunks;
}

int cumayo_set_launch_config(const cumayo_launch_config *cfg) {
    if (!cfg || !listed(wv::PREP_VARIANTS, cfg->sign_prep) || !listed(wv::SOLVE_OCCS, cfg->sign_solve) ||
        !listed(wv::VERIFY_VARIANTS, cfg->verify) || (unsigned)cfg->sign_order > 1u ||
        (unsigned)cfg->verify_order > 1u || cfg->sign_chunks < 0 || cfg->sign_chunks > 2)
        return (int)hipErrorInvalidValue;
    g_prep.store(cfg->sign_prep);
    g_solve.store(cfg->sign_solve);
    g_sign_order.store(cfg->sign_order);
    g_verify.store(cfg->verify);
    g_verify_order.store(cfg->verify_order);
    g_sign_chunks.store(cfg->sign_chunks);
    return 0;
}

int cumayo_launch_variants(int field, int32_t *out, int max) {
    const int *v = nullptr;
    int cnt = 0;
    static const int orders[2] = {0, 1};
    static const int chunks[3] = {1, 2, 0};
    auto use = [&](const auto &arr) { v = arr, cnt = (int)(sizeof(arr) / sizeof(arr[0])); };
    switch (field) {
        case CUMAYO_FIELD_SIGN_PREP: use(wv::PREP_VARIANTS); break;
        case CUMAYO_FIELD_SIGN_SOLVE: use(wv::SOLVE_OCCS); break;
        case CUMAYO_FIELD_VERIFY: use(wv::VERIFY_VARIANTS); break;
        case CUMAYO_FIELD_ORDER: use(orders); break;
        case CUMAYO_FIELD_SIGN_CHUNKS: use(chunks); break;
        default: return -1;
    }
    for (int i = 0; i < cnt && i < max && out; ++i) out[i] = v[i];
    return cnt;
}

int cumayo_autotune(const uint8_t *prep_esk, const uint8_t *prep_epk, const int32_t *key_idx,
                    const uint8_t *msgs, const int64_t *msg_off, const int32_t *msg_len, uint8_t *sigs,
                    uint8_t *results, size_t n, uint8_t *sign_ws, uint8_t *verify_ws, hipStream_t stream,
                    int reps, cumayo_tune_report *out) {
    const bool do_sign = prep_esk && sign_ws, do_verify = prep_epk && verify_ws && results;
    if (!n || n > CUMAYO_MAX_BATCH || !key_idx || !msgs || !msg_off || !msg_len || !sigs || !out ||
        (!do_sign && !do_verify) ||
        ((uintptr_t)prep_esk & 15) || ((uintptr_t)prep_epk & 15))
        return (int)hipErrorInvalidValue;
    if (reps < 1) reps = 1;
    std::memset(out, 0, sizeof *out);
    hipEvent_t a = nullptr, b = nullptr;
    hipError_t he = hipEventCreate(&a);
    if (he == hipSuccess) he = hipEventCreate(&b);
    if (he != hipSuccess) {
        if (a) (void)hipEventDestroy(a);
        note(he);
        return (int)he;
    }
    LaunchCfg cfg = current_cfg();
    int err = 0, timed = 0;
    if (do_sign) {
        TableSign acc{prep_esk, wv::PrepKeys::ESK_BYTES, key_idx, msgs, msg_off, msg_len, sigs, sign_ws,
                      sign_ws + n * (size_t)SIGN_SCRATCH_STRIDE};
        err = sign_table_impl<wv::PrepKeys>(acc, n, stream, 1, cfg);  // hashing pass, once
        double best = 1e30;
        LaunchCfg bc = cfg;
        // the split into two parts only applies from CHUNK_MIN items on;
        // chunks 0 is the persistent fused kernel
        const int max_chunks = n >= CHUNK_MIN ? 2 : 1;
        for (int chunks = 0; chunks <= max_chunks; ++chunks)
            for (int order = 0; order < 2; ++order)
                for (int p : wv::PREP_VARIANTS)
                    for (int so : wv::SOLVE_OCCS) {
                        if (err) break;
                        if (chunks == 0 && so != wv::SOLVE_OCCS[0]) continue;  // the fused kernel has no solve variant
                        // the shared-key variants run in their own key order
                        // and are not a fused-kernel option
                        if (wv::sk_code(p) && (order == 1 || chunks == 0)) continue;
                        LaunchCfg c = cfg;
                        c.prep = p, c.solve = so, c.sign_order = order, c.sign_chunks = chunks;
                        const double t = time_min_ms(
                            stream, a, b, reps,
                            [&] { return sign_table_impl<wv::PrepKeys>(acc, n, stream, 2, c); }, &err);
                        ++timed;
                        if (p == wv::PREP_OCCS[0] && so == wv::SOLVE_OCCS[0] && order == 0 && chunks == 1)
                            out->sign_default_ms = t;
                        if (t < best) best = t, bc = c;
                    }
        cfg.prep = bc.prep, cfg.solve = bc.solve, cfg.sign_order = bc.sign_order, cfg.sign_chunks = bc.sign_chunks;
        out->sign_ms = best;
    }
    if (do_verify && !err) {
        TableVerify acc{prep_epk, wv::PrepKeys::EPK_BYTES, key_idx, msgs, msg_off, msg_len, sigs, results,
                        verify_ws, verify_ws + n * (size_t)VERIFY_SCRATCH_STRIDE};
        err = verify_table_impl<wv::PrepKeys>(acc, n, stream, 1, cfg);
        double best = 1e30;
        LaunchCfg bc = cfg;
        for (int order = 0; order < 2; ++order)
            for (int vo : wv::VERIFY_VARIANTS) {
                if (err) break;
                if (wv::sk_code(vo) && order == 1) continue;  // own key order
                LaunchCfg c = cfg;
                c.verify = vo, c.verify_order = order;
                const double t = time_min_ms(
                    stream, a, b, reps, [&] { return verify_table_impl<wv::PrepKeys>(acc, n, stream, 2, c); },
                    &err);
                ++timed;
                if (vo == wv::VERIFY_OCCS[0] && order == 0) out->verify_default_ms = t;
                if (t < best) best = t, bc = c;
            }
        cfg.verify = bc.verify, cfg.verify_order = bc.verify_order;
        out->verify_ms = best;
    }
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    if (!err) {
        g_prep.store(cfg.prep);
        g_solve.store(cfg.solve);
        g_sign_order.store(cfg.sign_order);
        g_verify.store(cfg.verify);
        g_verify_order.store(cfg.verify_order);
        g_sign_chunks.store(cfg.sign_chunks);
    }
    cumayo_get_launch_config(&out->best);
    out->configs_timed = timed;
    return err;
}

cumayo_graph *cumayo_sign_graph_create(const uint8_t *prepared, const int32_t *key_idx, const uint8_t *msgs,
                                       const int64_t *msg_off, const int32_t *msg_len, uint8_t *sigs, size_t n,
                                       uint8_t *workspace, int stages) {
    if (!n) {
        note(hipErrorInvalidValue);
        return nullptr;
    }
    return capture([&](hipStream_t s) {
        return cumayo_sign_prepared(prepared, key_idx, msgs, msg_off, msg_len, sigs, n, workspace, s, stages);
    });
}

cumayo_graph *cumayo_verify_graph_create(const uint8_t *prepared, const int32_t *key_idx, const uint8_t *msgs,
                                         const int64_t *msg_off, const int32_t *msg_len, const uint8_t *sigs,
                                         uint8_t *results, size_t n, uint8_t *workspace, int stages) {
    if (!n) {
        note(hipErrorInvalidValue);
        return nullptr;
    }
    return capture([&](hipStream_t s) {
        return cumayo_verify_prepared(prepared, key_idx, msgs, msg_off, msg_len, sigs, results, n, workspace, s,
                                      stages);
    });
}

int cumayo_graph_launch(cumayo_graph *g, hipStream_t stream) {
    if (!g) return (int)hipErrorInvalidValue;
    const hipError_t e = hipGraphLaunch(g->exec, stream);
    note(e);
    return (int)e;
}

cumayo_graph *cumayo_graph_compose(cumayo_graph *const *parts, int n) {
    if (!parts || n <= 0) {
        note(hipErrorInvalidValue);
        return nullptr;
    }
    for (int i = 0; i < n; ++i)
        if (!parts[i] || !parts[i]->graph) {
            note(hipErrorInvalidValue);
            return nullptr;
        }
    hipGraph_t g = nullptr;
    hipError_t e = hipGraphCreate(&g, 0);
    for (int i = 0; i < n && e == hipSuccess; ++i) {
        hipGraphNode_t node = nullptr;
        e = hipGraphAddChildGraphNode(&node, g, nullptr, 0, parts[i]->graph);
    }
    hipGraphExec_t x = nullptr;
    if (e == hipSuccess) e = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
    if (e != hipSuccess) {
        if (g) (void)hipGraphDestroy(g);
        note(e);
        return nullptr;
    }
    return new cumayo_graph{g, x};
}

void cumayo_graph_destroy(cumayo_graph *g) {
    if (!g) return;
    (void)hipGraphExecDestroy(g->exec);
    (void)hipGraphDestroy(g->graph);
    delete g;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Continuous-batching server (C ABI)
// ---------------------------------------------------------------------------
#include "serve.hpp"

extern "C" {

void *cumayo_serve_create(const cumayo_serve_config *cfg) {
    if (!cfg) return nullptr;
    auto *s = new serve::Server(*cfg);
    if (s->error()) {
        note((hipError_t)s->error());
        delete s;
        return nullptr;
    }
    return s;
}

void cumayo_serve_destroy(void *server) { delete static_cast<serve::Server *>(server); }

size_t cumayo_serve_submit(void *server, const cumayo_request *reqs, size_t n) {
    return server ? static_cast<serve::Server *>(server)->submit(reqs, n) : 0;
}

int cumayo_serve_cancel(void *server, uint64_t tag) {
    return server ? static_cast<serve::Server *>(server)->cancel(tag) : 0;
}

size_t cumayo_serve_poll(void *server, cumayo_completion *out, size_t max) {
    return server ? static_cast<serve::Server *>(server)->poll(out, max) : 0;
}

void cumayo_serve_get_stats(void *server, cumayo_serve_stats *out) {
    if (server && out) static_cast<serve::Server *>(server)->stats(out);
}

void cumayo_serve_inject_failures(void *server, int n) {
    if (server) static_cast<serve::Server *>(server)->inject_failures(n);
}

int64_t cumayo_serve_clock_ns(void) { return serve::now_ns(); }

int cumayo_serve_run_poisson(void *server, double rate_per_s, double duration_s,
                             double verify_fraction, uint64_t seed, const uint8_t *msg_pool,
                             const int64_t *msg_off, const int32_t *msg_len, const int32_t *msg_key,
                             const uint8_t *sig_pool, size_t pool_n, cumayo_serve_report *out,
                             int64_t *lat_out, size_t lat_cap) {
    if (!server || !out) return (int)hipErrorInvalidValue;
    return serve::run_poisson(*static_cast<serve::Server *>(server), rate_per_s, duration_s,
                              verify_fraction, seed, msg_pool, msg_off, msg_len, msg_key, sig_pool,
                              pool_n, out, lat_out, lat_cap);
}

}  // extern "C"
