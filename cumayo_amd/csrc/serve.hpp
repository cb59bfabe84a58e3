// Continuous-batching request server over the batched sign / verify kernels
// (host C++; included by the C ABI translation unit).
//
// The server keeps its own prepared copies of the key tables it is given
// (the padded kernel layout), resident for its lifetime.
// Requests queue on the host. A worker thread owns n_slots batch slots, each
// with its own HIP stream, pinned staging and device buffers. Whenever a slot
// is free and requests are waiting it admits as many as fit (max_batch
// requests, slot_msg_bytes of messages), copies them in, launches the sign
// and verify kernels and the copies back, and records an event; finished
// slots are harvested as their events complete. Requests that arrive while
// earlier batches run are therefore admitted into the next free slot and
// their work overlaps on the GPU with the batches already in flight.
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstdint>
#include <cstring>
#include <deque>
#include <functional>
#include <mutex>
#include <random>
#include <thread>
#include <vector>

#include "../../include/cumayo.h"
#include "mayo_kernels.hpp"

namespace serve {

// Attempt counters of a signing batch, gathered from the per-item workspace
// slots into a dense byte array (one contiguous copy back instead of one
// strided row per request).
__global__ void k_gather_attempts(const uint8_t *__restrict__ ws, uint8_t *__restrict__ out, size_t n) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = ws[i * mayo::SIGN_SCRATCH_STRIDE + mayo::SCR_ATTEMPTS];
}

// out[0, n) = v. Verification results start as 1 (invalid), so a result
// byte that no verification kernel wrote reads as a rejection.
__global__ void k_fill_bytes(uint8_t *__restrict__ out, uint8_t v, size_t n) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = v;
}

// Fork-join pool for the host-side copies of a batch: run(n, chunk, fn)
// calls fn(begin, end) over [0, n) in chunks taken by the pool's threads
// and the caller, and returns when all are done. One job at a time.
class HostPool {
  public:
    explicit HostPool(int threads) {
        for (int i = 0; i < threads; ++i) th_.emplace_back([this] { loop(); });
    }
    ~HostPool() {
        {
            std::lock_guard<std::mutex> lk(mu_);
            stop_ = true;
        }
        cv_.notify_all();
        for (auto &t : th_) t.join();
    }
    template <class F>
    void run(size_t n, size_t chunk, F &&fn) {
        if (th_.empty() || n <= chunk) {
            fn(size_t(0), n);
            return;
        }
        const std::function<void(size_t, size_t)> job = fn;
        {
            std::lock_guard<std::mutex> lk(mu_);
            job_ = &job;
            n_ = n;
            chunk_ = chunk;
            next_.store(0);
            busy_ = (int)th_.size();
            ++gen_;
        }
        cv_.notify_all();
        work();
        std::unique_lock<std::mutex> lk(mu_);
        done_cv_.wait(lk, [this] { return busy_ == 0; });
        job_ = nullptr;
    }

  private:
    void work() {
        size_t b;
        while ((b = next_.fetch_add(chunk_)) < n_) (*job_)(b, std::min(b + chunk_, n_));
    }
    void loop() {
        uint64_t seen = 0;
        for (;;) {
            {
                std::unique_lock<std::mutex> lk(mu_);
                cv_.wait(lk, [&] { return stop_ || gen_ != seen; });
                if (stop_) return;
                seen = gen_;
            }
            // run() waits for every thread before it returns, so each
            // thread takes part in each job once
            work();
            std::lock_guard<std::mutex> lk(mu_);
            if (--busy_ == 0) done_cv_.notify_one();
        }
    }
    std::vector<std::thread> th_;
    std::mutex mu_;
    std::condition_variable cv_, done_cv_;
    const std::function<void(size_t, size_t)> *job_ = nullptr;
    size_t n_ = 0, chunk_ = 1;
    std::atomic<size_t> next_{0};
    int busy_ = 0;
    uint64_t gen_ = 0;
    bool stop_ = false;
};

// Stable counting sort of requests by the SHAKE256 blocks their message
// absorbs (136 bytes each; 255 and more share the last bin).
inline void order_by_blocks(std::vector<cumayo_request> &v, std::vector<cumayo_request> &tmp) {
    constexpr size_t BINS = 256;
    if (v.size() < 2) return;
    size_t pos[BINS] = {};
    for (const auto &r : v) ++pos[std::min<size_t>(r.len / 136, BINS - 1)];
    size_t sum = 0;
    for (size_t b = 0; b < BINS; ++b) {
        const size_t c = pos[b];
        pos[b] = sum;
        sum += c;
    }
    tmp.resize(v.size());
    for (const auto &r : v) tmp[pos[std::min<size_t>(r.len / 136, BINS - 1)]++] = r;
    v.swap(tmp);
}

inline int64_t now_ns() {
    return std::chrono::duration_cast<std::chrono::nanoseconds>(
               std::chrono::steady_clock::now().time_since_epoch())
        .count();
}

struct Slot {
    hipStream_t stream = nullptr;
    hipEvent_t done = nullptr;
    bool busy = false;
    bool failed = false;  // a copy, launch or the batch's event reported an error
    uint8_t *h_msgs = nullptr, *h_sig = nullptr, *h_res = nullptr, *h_att = nullptr;
    int64_t *h_off = nullptr;
    int32_t *h_len = nullptr, *h_key = nullptr;
    uint8_t *d_msgs = nullptr, *d_sig = nullptr, *d_res = nullptr, *d_att = nullptr, *d_wss = nullptr,
            *d_wsv = nullptr;
    int64_t *d_off = nullptr;
    int32_t *d_len = nullptr, *d_key = nullptr;
    size_t n_sign = 0, n_verify = 0;
    std::vector<cumayo_request> reqs;  // sign requests first, then verifications
    std::vector<cumayo_request> vreqs; // verifications while admitting
    std::vector<cumayo_request> tmp;
    std::vector<cumayo_completion> comps;  // completions of the batch
};

class Server {
  public:
    explicit Server(const cumayo_serve_config &cfg)
        : cfg_(cfg), pool_(cfg.host_threads > 0 ? cfg.host_threads - 1 : 3) {
        if (cfg_.max_batch <= 0) cfg_.max_batch = 4096;
        if ((uint32_t)cfg_.max_batch > CUMAYO_MAX_BATCH) cfg_.max_batch = (int32_t)CUMAYO_MAX_BATCH;
        if (cfg_.n_slots <= 0) cfg_.n_slots = 3;
        if (!cfg_.slot_msg_bytes) cfg_.slot_msg_bytes = (size_t)cfg_.max_batch * 1024;
        ok(hipSetDevice(cfg_.device));
        // resident keys in the kernel layout, built once on the device (at
        // most PREP_KEYS keys per call of the preparation entry points)
        constexpr size_t PREP_KEYS = 65535;
        const size_t nk = cfg_.n_keys > 0 ? (size_t)cfg_.n_keys : 0;
        if (cfg_.esk_table && nk) {
            const size_t rb = cumayo_prepared_esk_bytes();
            dev_alloc(&pesk_, nk * rb);
            for (size_t k0 = 0; pesk_ && k0 < nk; k0 += PREP_KEYS)
                ok((hipError_t)cumayo_prepare_esk(cfg_.esk_table + k0 * cfg_.esk_stride, cfg_.esk_stride,
                                                  std::min(PREP_KEYS, nk - k0), pesk_ + k0 * rb, nullptr));
        }
        if (cfg_.epk_table && nk) {
            const size_t rb = cumayo_prepared_epk_bytes();
            dev_alloc(&pepk_, nk * rb);
            for (size_t k0 = 0; pepk_ && k0 < nk; k0 += PREP_KEYS)
                ok((hipError_t)cumayo_prepare_epk(cfg_.epk_table + k0 * cfg_.epk_stride, cfg_.epk_stride,
                                                  std::min(PREP_KEYS, nk - k0), pepk_ + k0 * rb, nullptr));
        }
        ok(hipDeviceSynchronize());
        const size_t mb = (size_t)cfg_.max_batch;
        slots_.resize((size_t)cfg_.n_slots);
        for (auto &s : slots_) {
            ok(hipStreamCreateWithFlags(&s.stream, hipStreamNonBlocking));
            ok(hipEventCreateWithFlags(&s.done, hipEventDisableTiming));
            ok(hipHostMalloc((void **)&s.h_msgs, cfg_.slot_msg_bytes + 8, hipHostMallocDefault));
            ok(hipHostMalloc((void **)&s.h_off, mb * sizeof(int64_t), hipHostMallocDefault));
            ok(hipHostMalloc((void **)&s.h_len, mb * sizeof(int32_t), hipHostMallocDefault));
            ok(hipHostMalloc((void **)&s.h_key, mb * sizeof(int32_t), hipHostMallocDefault));
            ok(hipHostMalloc((void **)&s.h_sig, mb * mayo::SIG_BYTES, hipHostMallocDefault));
            ok(hipHostMalloc((void **)&s.h_res, mb, hipHostMallocDefault));
            ok(hipHostMalloc((void **)&s.h_att, mb, hipHostMallocDefault));
            dev_alloc(&s.d_msgs, cfg_.slot_msg_bytes + 8);
            dev_alloc((uint8_t **)&s.d_off, mb * sizeof(int64_t));
            dev_alloc((uint8_t **)&s.d_len, mb * sizeof(int32_t));
            dev_alloc((uint8_t **)&s.d_key, mb * sizeof(int32_t));
            dev_alloc(&s.d_sig, mb * mayo::SIG_BYTES);
            dev_alloc(&s.d_res, mb);
            dev_alloc(&s.d_att, mb);
            dev_alloc(&s.d_wss, cumayo_sign_workspace_size(mb));
            dev_alloc(&s.d_wsv, cumayo_verify_workspace_size(mb));
            s.reqs.reserve(mb);
            s.vreqs.reserve(mb);
            s.tmp.reserve(mb);
            s.comps.resize(mb);
        }
        thread_ = std::thread([this] { worker(); });
    }

    ~Server() {
        {
            std::lock_guard<std::mutex> lk(mu_);
            stop_ = true;
        }
        cv_.notify_all();
        if (thread_.joinable()) thread_.join();
        (void)hipSetDevice(cfg_.device);
        for (auto &s : slots_) {
            if (s.stream) (void)hipStreamSynchronize(s.stream);
            (void)hipHostFree(s.h_msgs); (void)hipHostFree(s.h_off); (void)hipHostFree(s.h_len);
            (void)hipHostFree(s.h_key); (void)hipHostFree(s.h_sig); (void)hipHostFree(s.h_res);
            (void)hipHostFree(s.h_att);
            (void)hipFree(s.d_msgs); (void)hipFree(s.d_off); (void)hipFree(s.d_len);
            (void)hipFree(s.d_key); (void)hipFree(s.d_sig); (void)hipFree(s.d_res); (void)hipFree(s.d_att);
            (void)hipFree(s.d_wss); (void)hipFree(s.d_wsv);
            if (s.done) (void)hipEventDestroy(s.done);
            if (s.stream) (void)hipStreamDestroy(s.stream);
        }
        (void)hipFree(pesk_);
        (void)hipFree(pepk_);
    }

    size_t submit(const cumayo_request *reqs, size_t n) {
        const int64_t t = now_ns();
        size_t accepted = 0;
        {
            std::lock_guard<std::mutex> lk(mu_);
            for (size_t i = 0; i < n; ++i) {
                if (cfg_.max_queue > 0 && (int64_t)pending_.size() >= cfg_.max_queue) {
                    rejected_ += n - i;
                    break;
                }
                cumayo_request r = reqs[i];
                if (!r.t_arrival_ns) r.t_arrival_ns = t;
                // a request that can never fit a slot, or names no key, is
                // answered at once with status 2
                if (r.len > cfg_.slot_msg_bytes || r.key < 0 || r.key >= cfg_.n_keys ||
                    (r.op == CUMAYO_OP_SIGN && !cfg_.esk_table) ||
                    (r.op == CUMAYO_OP_VERIFY && !cfg_.epk_table)) {
                    completed_.push_back({r.tag, now_ns() - r.t_arrival_ns, 2});
                } else {
                    pending_.push_back(r);
                }
                ++accepted;
            }
        }
        cv_.notify_one();
        return accepted;
    }

    // Withdraws a request that is still queued; it completes with status 3.
    int cancel(uint64_t tag) {
        std::lock_guard<std::mutex> lk(mu_);
        for (auto it = pending_.begin(); it != pending_.end(); ++it)
            if (it->tag == tag) {
                completed_.push_back({tag, now_ns() - it->t_arrival_ns, 3});
                pending_.erase(it);
                ++cancelled_;
                return 1;
            }
        return 0;
    }

    size_t poll(cumayo_completion *out, size_t max) {
        std::lock_guard<std::mutex> lk(mu_);
        const size_t k = std::min(max, completed_.size());
        std::copy(completed_.begin(), completed_.begin() + (long)k, out);
        completed_.erase(completed_.begin(), completed_.begin() + (long)k);
        return k;
    }

    void stats(cumayo_serve_stats *out) {
        std::lock_guard<std::mutex> lk(mu_);
        out->batches = batches_;
        out->requests = requests_done_;
        out->rejected = rejected_;
        out->max_in_flight = max_in_flight_;
        out->overlapped_launches = overlapped_;
        out->device_bytes = device_bytes_;
        out->cancelled = cancelled_;
        out->failed = failed_requests_;
        out->last_error = err_.load();
    }

    // Test hook: the next n batches fail as if an enqueue had returned an
    // error (nothing is launched for them).
    void inject_failures(int n) { inject_.store(n); }

    int error() const { return err_.load(); }
    size_t in_system() {
        std::lock_guard<std::mutex> lk(mu_);
        return pending_.size() + in_flight_requests_;
    }

  private:
    void ok(hipError_t e) {
        if (e != hipSuccess) err_.store((int)e);
    }
    void fail(Slot &s, hipError_t e) {
        s.failed = true;
        err_.store((int)e);
    }
    void dev_alloc(uint8_t **p, size_t bytes) {
        ok(hipMalloc((void **)p, bytes ? bytes : 1));
        device_bytes_ += bytes;
    }

    void worker() {
        (void)hipSetDevice(cfg_.device);
        for (;;) {
            bool progressed = false;
            for (auto &s : slots_) {
                if (!s.busy) continue;
                const hipError_t q = hipEventQuery(s.done);
                if (q == hipErrorNotReady) continue;
                if (q != hipSuccess) fail(s, q);  // the batch's work ended in an error
                finish(s);
                progressed = true;
            }
            for (auto &s : slots_)
                if (!s.busy) {
                    if (!launch(s)) break;
                    progressed = true;
                }
            if (progressed) continue;
            std::unique_lock<std::mutex> lk(mu_);
            if (stop_ && pending_.empty() && in_flight_requests_ == 0) return;
            bool any_busy = in_flight_requests_ > 0;
            if (!any_busy && pending_.empty() && !stop_) {
                cv_.wait_for(lk, std::chrono::microseconds(200));
            } else {
                lk.unlock();
                std::this_thread::yield();
            }
        }
    }

    // Admit waiting requests into slot s and launch its batch.
    bool launch(Slot &s) {
        s.reqs.clear();
        s.vreqs.clear();
        {
            std::lock_guard<std::mutex> lk(mu_);
            size_t bytes = 0, taken = 0;
            while (!pending_.empty() && taken < (size_t)cfg_.max_batch) {
                const cumayo_request &r = pending_.front();
                if (bytes + r.len > cfg_.slot_msg_bytes) break;
                bytes += r.len;
                (r.op == CUMAYO_OP_SIGN ? s.reqs : s.vreqs).push_back(r);
                pending_.pop_front();
                ++taken;
            }
            if (!taken) return false;
            in_flight_requests_ += taken;
        }
        // sign requests first, then verifications, each grouped by message
        // length (the hashing pass gives each lane one message, so a wave
        // takes as long as its longest one)
        order_by_blocks(s.reqs, s.tmp);
        order_by_blocks(s.vreqs, s.tmp);
        s.n_sign = s.reqs.size();
        s.n_verify = s.vreqs.size();
        s.reqs.insert(s.reqs.end(), s.vreqs.begin(), s.vreqs.end());
        const size_t n = s.reqs.size();
        int64_t off = 0;
        for (size_t i = 0; i < n; ++i) {
            const cumayo_request &r = s.reqs[i];
            s.h_off[i] = off;
            s.h_len[i] = (int32_t)r.len;
            s.h_key[i] = r.key;
            off += r.len;
        }
        // the copies into the pinned staging, spread over the host pool
        pool_.run(n, 256, [&](size_t b, size_t e) {
            for (size_t i = b; i < e; ++i) {
                const cumayo_request &r = s.reqs[i];
                if (r.len) std::memcpy(s.h_msgs + s.h_off[i], r.msg, r.len);
                if (i >= s.n_sign) std::memcpy(s.h_sig + i * mayo::SIG_BYTES, r.sig_in, mayo::SIG_BYTES);
            }
        });
        const hipStream_t st = s.stream;
        // every enqueue is checked: an error fails the whole batch (its
        // requests complete with STATUS_ERROR and no output is read back)
        auto chk = [&](int e) {
            if (e) fail(s, (hipError_t)e);
        };
        if (inject_.load() > 0 && inject_.fetch_sub(1) > 0) fail(s, hipErrorLaunchFailure);  // test hook
        // result staging starts failed-closed: "all attempts failed" for
        // signing, "invalid" for verification
        if (s.n_sign) std::memset(s.h_att, 0, s.n_sign);
        if (s.n_verify) std::memset(s.h_res, 1, s.n_verify);
        if (!s.failed) chk(hipMemcpyAsync(s.d_msgs, s.h_msgs, (size_t)off + 1, hipMemcpyHostToDevice, st));
        if (!s.failed) chk(hipMemcpyAsync(s.d_off, s.h_off, n * sizeof(int64_t), hipMemcpyHostToDevice, st));
        if (!s.failed) chk(hipMemcpyAsync(s.d_len, s.h_len, n * sizeof(int32_t), hipMemcpyHostToDevice, st));
        if (!s.failed) chk(hipMemcpyAsync(s.d_key, s.h_key, n * sizeof(int32_t), hipMemcpyHostToDevice, st));
        if (s.n_verify && !s.failed)
            chk(hipMemcpyAsync(s.d_sig + s.n_sign * mayo::SIG_BYTES, s.h_sig + s.n_sign * mayo::SIG_BYTES,
                               s.n_verify * mayo::SIG_BYTES, hipMemcpyHostToDevice, st));
        if (s.n_sign && !s.failed) {
            chk(cumayo_sign_prepared(pesk_, s.d_key, s.d_msgs, s.d_off, s.d_len, s.d_sig, s.n_sign, s.d_wss, st,
                                     3));
            if (!s.failed)
                chk(hipMemcpyAsync(s.h_sig, s.d_sig, s.n_sign * mayo::SIG_BYTES, hipMemcpyDeviceToHost, st));
            if (!s.failed) {
                k_gather_attempts<<<(unsigned)((s.n_sign + 255) / 256), 256, 0, st>>>(s.d_wss, s.d_att, s.n_sign);
                chk(hipGetLastError());
            }
            if (!s.failed) chk(hipMemcpyAsync(s.h_att, s.d_att, s.n_sign, hipMemcpyDeviceToHost, st));
        }
        if (s.n_verify && !s.failed) {
            k_fill_bytes<<<(unsigned)((s.n_verify + 255) / 256), 256, 0, st>>>(s.d_res, 1, s.n_verify);
            chk(hipGetLastError());
            if (!s.failed)
                chk(cumayo_verify_prepared(pepk_, s.d_key + s.n_sign, s.d_msgs, s.d_off + s.n_sign,
                                           s.d_len + s.n_sign, s.d_sig + s.n_sign * mayo::SIG_BYTES, s.d_res,
                                           s.n_verify, s.d_wsv, st, 3));
            if (!s.failed) chk(hipMemcpyAsync(s.h_res, s.d_res, s.n_verify, hipMemcpyDeviceToHost, st));
        }
        // the event completes the slot even when nothing was enqueued; if it
        // cannot be recorded the slot is finished at once, as failed
        if (hipEventRecord(s.done, st) != hipSuccess) {
            fail(s, hipErrorInvalidValue);
            s.busy = true;
            finish(s);
            return true;
        }
        s.busy = true;
        size_t busy = 0;
        for (auto &o : slots_) busy += o.busy ? 1 : 0;
        std::lock_guard<std::mutex> lk(mu_);
        ++batches_;
        if (busy > 1) ++overlapped_;
        max_in_flight_ = std::max<uint64_t>(max_in_flight_, busy);
        return true;
    }

    void finish(Slot &s) {
        const int64_t t = now_ns();
        const size_t n = s.reqs.size();
        const bool failed = s.failed;
        pool_.run(n, 256, [&](size_t b, size_t e) {
            for (size_t i = b; i < e; ++i) {
                const cumayo_request &r = s.reqs[i];
                if (failed) {
                    // nothing of this batch is trusted: no output is written
                    s.comps[i] = {r.tag, t - r.t_arrival_ns, CUMAYO_STATUS_ERROR};
                    continue;
                }
                uint8_t status;
                if (i < s.n_sign) {
                    if (r.sig_out) std::memcpy(r.sig_out, s.h_sig + i * mayo::SIG_BYTES, mayo::SIG_BYTES);
                    status = s.h_att[i] == 0 ? 1 : 0;
                    if (r.result_out) *r.result_out = s.h_att[i];
                } else {
                    status = s.h_res[i - s.n_sign];
                    if (r.result_out) *r.result_out = status;
                }
                s.comps[i] = {r.tag, t - r.t_arrival_ns, status};
            }
        });
        s.busy = false;
        s.failed = false;
        std::lock_guard<std::mutex> lk(mu_);
        if (failed) failed_requests_ += n;
        completed_.insert(completed_.end(), s.comps.begin(), s.comps.begin() + (long)n);
        requests_done_ += s.reqs.size();
        in_flight_requests_ -= s.reqs.size();
    }

    cumayo_serve_config cfg_;
    HostPool pool_;
    uint8_t *pesk_ = nullptr, *pepk_ = nullptr;  // prepared key tables
    std::vector<Slot> slots_;
    std::mutex mu_;
    std::condition_variable cv_;
    std::deque<cumayo_request> pending_;
    std::vector<cumayo_completion> completed_;
    std::thread thread_;
    bool stop_ = false;
    std::atomic<int> err_{0};
    std::atomic<int> inject_{0};
    uint64_t batches_ = 0, requests_done_ = 0, rejected_ = 0, max_in_flight_ = 0, overlapped_ = 0,
             cancelled_ = 0, failed_requests_ = 0;
    size_t in_flight_requests_ = 0, device_bytes_ = 0;
};

// Open-loop Poisson load generator (in the calling thread).
inline int run_poisson(Server &srv, double rate, double duration, double verify_fraction,
                       uint64_t seed, const uint8_t *msg_pool, const int64_t *msg_off,
                       const int32_t *msg_len, const int32_t *msg_key, const uint8_t *sig_pool,
                       size_t pool_n, cumayo_serve_report *out, int64_t *lat_out, size_t lat_cap) {
    if (rate <= 0 || duration <= 0 || !pool_n) return (int)hipErrorInvalidValue;
    std::mt19937_64 rng(seed);
    std::exponential_distribution<double> gap(rate);
    std::uniform_real_distribution<double> uni(0.0, 1.0);
    std::vector<int64_t> lat;
    lat.reserve((size_t)(rate * duration * 1.1) + 16);
    std::vector<cumayo_request> batch;
    std::vector<cumayo_completion> comp(1 << 16);
    // results land in a ring of output rows (row i % pool_n), so the copy of
    // every signature and result byte to its caller is part of the measurement
    std::vector<uint8_t> sig_out(pool_n * (size_t)mayo::SIG_BYTES), res_out(pool_n);
    uint64_t submitted = 0, refused = 0, k = 0;
    const int64_t t0 = now_ns(), t_end = t0 + (int64_t)(duration * 1e9);
    double next = (double)t0 + gap(rng) * 1e9;
    auto drain = [&] {
        size_t m;
        while ((m = srv.poll(comp.data(), comp.size())) > 0)
            for (size_t i = 0; i < m; ++i) lat.push_back(comp[i].latency_ns);
    };
    for (;;) {
        const int64_t t = now_ns();
        if (t >= t_end) break;
        batch.clear();
        while (next <= (double)t && next < (double)t_end) {
            const size_t i = (size_t)(k % pool_n);
            cumayo_request r{};
            r.op = (uni(rng) < verify_fraction) ? CUMAYO_OP_VERIFY : CUMAYO_OP_SIGN;
            r.key = msg_key[i];
            r.msg = msg_pool + msg_off[i];
            r.len = (uint32_t)msg_len[i];
            r.sig_in = sig_pool ? sig_pool + i * (size_t)mayo::SIG_BYTES : nullptr;
            if (r.op == CUMAYO_OP_VERIFY && !r.sig_in) r.op = CUMAYO_OP_SIGN;
            r.sig_out = r.op == CUMAYO_OP_SIGN ? sig_out.data() + i * (size_t)mayo::SIG_BYTES : nullptr;
            r.result_out = res_out.data() + i;
            r.tag = k;
            r.t_arrival_ns = (int64_t)next;
            batch.push_back(r);
            ++k;
            next += gap(rng) * 1e9;
        }
        if (!batch.empty()) {
            const size_t acc = srv.submit(batch.data(), batch.size());
            submitted += acc;
            refused += batch.size() - acc;
        }
        drain();
        // sleep through long gaps between arrivals rather than contend for the queue lock
        const double wait = next - (double)now_ns();
        if (wait > 20e3) std::this_thread::sleep_for(std::chrono::nanoseconds((int64_t)std::min(wait, 100e3)));
    }
    const int64_t t_gen_end = now_ns();
    while (srv.in_system() > 0) {
        drain();
        std::this_thread::sleep_for(std::chrono::microseconds(50));
        if (srv.error()) break;
    }
    drain();
    const int64_t t_done = now_ns();
    if (lat_out) std::copy(lat.begin(), lat.begin() + (long)std::min(lat.size(), lat_cap), lat_out);
    std::sort(lat.begin(), lat.end());
    auto pct = [&](double q) {
        if (lat.empty()) return 0.0;
        size_t idx = (size_t)std::min<double>((double)lat.size() - 1, q * (double)(lat.size() - 1));
        return lat[idx] / 1e6;
    };
    double sum = 0;
    for (auto v : lat) sum += (double)v;
    out->offered_per_s = (double)k / ((t_gen_end - t0) / 1e9);
    out->achieved_per_s = (double)lat.size() / ((t_done - t0) / 1e9);
    out->p50_ms = pct(0.50);
    out->p90_ms = pct(0.90);
    out->p99_ms = pct(0.99);
    out->mean_ms = lat.empty() ? 0 : sum / lat.size() / 1e6;
    out->max_ms = lat.empty() ? 0 : lat.back() / 1e6;
    out->submitted = submitted;
    out->completed = lat.size();
    out->refused = refused;
    out->duration_s = (t_done - t0) / 1e9;
    return srv.error();
}

}  // namespace serve
