// The reference's benchmark flow through the drop-in C ABI (include/cumayo.h):
// key generation, signing with compact secret keys (crypto_sign_signWithesk:
// upload, per-item key expansion, signing, download) and verification with
// compact public keys (crypto_sign_verifyWithepk), each over a batch split
// across non-blocking streams and timed on the wall clock around a device
// synchronisation -- the method of the reference's tests/bench_stream.cu
// (batch 20,000 over 10 streams, 32-byte messages, 10 timed iterations).
//
// Usage: bench_flow_mayoN [batch] [streams] [iters] [mlen]
// Prints one JSON line; exits non-zero if a signature fails to verify.
#include <hip/hip_runtime.h>

#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "cumayo.h"

namespace {

size_t a128(size_t x) { return (x + 127) / 128 * 128; }

void check(hipError_t e, const char *what) {
    if (e != hipSuccess) {
        std::fprintf(stderr, "%s: %s\n", what, hipGetErrorString(e));
        std::exit(2);
    }
}

struct Split {
    size_t base, count;
};

// The batch in n_streams parts, the last one taking the remainder.
std::vector<Split> split(size_t batch, int n_streams) {
    std::vector<Split> out;
    const size_t sub = batch / (size_t)n_streams;
    for (int i = 0; i < n_streams; ++i) {
        const size_t cnt = sub + (i == n_streams - 1 ? batch % (size_t)n_streams : 0);
        if (cnt) out.push_back({(size_t)i * sub, cnt});
    }
    return out;
}

template <class F>
double ops_per_s(size_t batch, int iters, F &&round) {
    round();  // warm-up
    check(hipDeviceSynchronize(), "warm-up");
    double best = 0, total = 0;
    for (int it = 0; it < iters; ++it) {
        const auto t0 = std::chrono::steady_clock::now();
        round();
        check(hipDeviceSynchronize(), "round");
        const double s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        total += s;
        best = best > 0 && best < s ? best : s;
    }
    (void)best;
    return (double)batch * iters / total;
}

}  // namespace

int main(int argc, char **argv) {
    const size_t batch = argc > 1 ? std::strtoull(argv[1], nullptr, 10) : 20000;
    const int n_streams = argc > 2 ? std::atoi(argv[2]) : 10;
    const int iters = argc > 3 ? std::atoi(argv[3]) : 10;
    const size_t mlen = argc > 4 ? std::strtoull(argv[4], nullptr, 10) : 32;
    if (!batch || n_streams < 1 || iters < 1) return 2;

    cumayo_sizes z;
    cumayo_get_sizes(&z);
    aes128_tables_init();
    std::vector<hipStream_t> streams((size_t)n_streams);
    for (auto &s : streams) check(hipStreamCreateWithFlags(&s, hipStreamNonBlocking), "stream");
    const auto parts = split(batch, n_streams);

    // key generation: rows sk || pk || P1 P2 scratch
    uint8_t *h_pk, *h_sk;
    check(hipHostMalloc((void **)&h_pk, batch * z.pk_bytes, hipHostMallocDefault), "host pk");
    check(hipHostMalloc((void **)&h_sk, batch * z.sk_bytes, hipHostMallocDefault), "host sk");
    const size_t p_bytes = (size_t)z.p1_bytes + z.p2_bytes;
    uint8_t *kpool;
    size_t kpitch;
    check(hipMallocPitch((void **)&kpool, &kpitch, a128(z.sk_bytes) + a128(z.pk_bytes) + a128(p_bytes), batch),
          "keypair pool");
    const double keypair_s = ops_per_s(batch, iters, [&] {
        for (size_t i = 0; i < parts.size(); ++i) {
            uint8_t *pool = kpool + parts[i].base * kpitch;
            crypto_sign_keypair(pool, kpitch, parts[i].count, streams[i], 0);
            crypto_sign_keypair_DtH(h_pk + parts[i].base * z.pk_bytes, h_sk + parts[i].base * z.sk_bytes, pool,
                                    kpitch, parts[i].count, streams[i]);
        }
    });
    check(hipFree(kpool), "free");

    // signing with compact keys: rows A128(esk) || A128(msg) || A128(sig) || A128(scratch)
    uint8_t *h_msg, *h_sig;
    check(hipHostMalloc((void **)&h_msg, batch * mlen + 1, hipHostMallocDefault), "host msg");
    check(hipHostMalloc((void **)&h_sig, batch * z.sig_bytes, hipHostMallocDefault), "host sig");
    uint64_t st = 0x9E3779B97F4A7C15ull;
    for (size_t i = 0; i < batch * mlen; ++i) {
        st ^= st << 13; st ^= st >> 7; st ^= st << 17;
        h_msg[i] = (uint8_t)st;
    }
    const size_t scratch = (size_t)z.k * z.m * ((z.o + 1) / 2);
    uint8_t *spool;
    size_t spitch;
    check(hipMallocPitch((void **)&spool, &spitch,
                         a128(z.esk_bytes) + a128(mlen) + a128(z.sig_bytes) + a128(scratch), batch),
          "sign pool");
    const double sign_s = ops_per_s(batch, iters, [&] {
        for (size_t i = 0; i < parts.size(); ++i) {
            uint8_t *pool = spool + parts[i].base * spitch;
            crypto_sign_signWithesk_HtD(h_sk + parts[i].base * z.sk_bytes, h_msg + parts[i].base * mlen, mlen,
                                        pool, spitch, parts[i].count, streams[i]);
            crypto_sign_signWithesk(pool, mlen, spitch, parts[i].count, streams[i]);
            crypto_sign_signWithesk_DtH(h_sig + parts[i].base * z.sig_bytes, pool, mlen, spitch, parts[i].count,
                                        streams[i]);
        }
    });
    check(hipFree(spool), "free");

    // verification with compact keys: rows A128(epk) || A128(msg) || A128(sig) || result
    uint8_t *h_res;
    check(hipHostMalloc((void **)&h_res, batch, hipHostMallocDefault), "host res");
    std::memset(h_res, 0xEE, batch);
    const size_t epk_len = p_bytes + z.pk_bytes - 16;
    uint8_t *vpool;
    size_t vpitch;
    check(hipMallocPitch((void **)&vpool, &vpitch, a128(epk_len) + a128(mlen) + a128(z.sig_bytes) + 1, batch),
          "verify pool");
    const double verify_s = ops_per_s(batch, iters, [&] {
        for (size_t i = 0; i < parts.size(); ++i) {
            uint8_t *pool = vpool + parts[i].base * vpitch;
            crypto_sign_verifyWithepk_HtD(h_pk + parts[i].base * z.pk_bytes, h_msg + parts[i].base * mlen,
                                          h_sig + parts[i].base * z.sig_bytes, mlen, pool, vpitch,
                                          parts[i].count, streams[i]);
            crypto_sign_verifyWithepk(pool, mlen, vpitch, parts[i].count, streams[i]);
            crypto_sign_verifyWithepk_DtH(h_res + parts[i].base, mlen, pool, vpitch, parts[i].count, streams[i]);
        }
    });
    check(hipFree(vpool), "free");

    size_t invalid = 0;
    for (size_t i = 0; i < batch; ++i) invalid += h_res[i] != 0;
    const int err = cumayo_last_error();
    std::printf("{\"param_set\": \"MAYO_%d\", \"batch\": %zu, \"streams\": %d, \"iters\": %d, \"mlen\": %zu, "
                "\"keypair_per_s\": %.1f, \"sign_per_s\": %.1f, \"verify_per_s\": %.1f, "
                "\"invalid\": %zu, \"hip_error\": %d}\n",
                z.version, batch, n_streams, iters, mlen, keypair_s, sign_s, verify_s, invalid, err);
    for (auto &s : streams) (void)hipStreamDestroy(s);
    (void)hipHostFree(h_pk); (void)hipHostFree(h_sk); (void)hipHostFree(h_msg);
    (void)hipHostFree(h_sig); (void)hipHostFree(h_res);
    return (invalid || err) ? 1 : 0;
}
