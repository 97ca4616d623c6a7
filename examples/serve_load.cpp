// Load generator for the request loop (turtle_server_*).  Open loop: requests
// arrive as a Poisson process of rate --rate per second for --seconds;
// sizes are log-uniform in [--lo, --hi] keys, the key type is uniform over
// {uint8, int16, int32, int64}, keys are uniform over the type's range, and
// half of the requests carry an int32 index (iota).  Every result is
// checked: keys non-decreasing, key sum equal to the input's, and at 64
// sampled positions keys_out[i] == input[idx_out[i]].  Prints one JSON line
// with p50/p99 latency (submission to completion), the aggregate keys/s and
// the five slowest requests with their queueing and batch figures.
// Closed loop (--closed C): C clients each submit a request of the same
// distribution and wait for it, again and again, for --seconds; the line
// then reports the throughput the server sustains at that concurrency.
// Built with plain g++ against libturtle_hip.so.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "turtle_c.h"

namespace {

uint64_t mix64(uint64_t z)
{
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// splitmix64 draw i (0-based) of seed s, as in turtlesort_amd/workloads.py
uint64_t draw(uint64_t seed, uint64_t i) { return mix64(seed + (i + 1) * 0x9E3779B97F4A7C15ull); }

double unit(uint64_t seed, uint64_t i) { return (double)(draw(seed, i) >> 11) * (1.0 / 9007199254740992.0); }

double now_s()
{
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

const int TYPES[4] = {TURTLE_UINT8, TURTLE_INT16, TURTLE_INT32, TURTLE_INT64};

struct Req {
    uint64_t ticket = 0;
    int dtype = 0, kb = 0;
    int64_t n = 0, src_off = 0;
    void *keys = nullptr;
    int32_t *idx = nullptr;
    uint64_t sum = 0;   // sum of the input keys (as unsigned, wrapping)
    double lat = 0.0;
    int status = 0;
    int idx_flag = 0;   // the request carried an index
    turtle_request_info info{};
};

template <typename T>
bool check_keys(const Req &r, const T *src)
{
    const T *k = (const T *)r.keys;
    uint64_t s = 0;
    for (int64_t i = 0; i < r.n; i++) {
        if (i && k[i] < k[i - 1]) return false;
        s += (uint64_t)k[i];
    }
    if (s != r.sum) return false;
    if (r.idx) {
        for (int j = 0; j < 64; j++) {
            const int64_t i = (int64_t)(draw(r.ticket, (uint64_t)j) % (uint64_t)r.n);
            const int32_t x = r.idx[i];
            if (x < 0 || x >= r.n || k[i] != src[r.src_off + x]) return false;
        }
    }
    return true;
}

}  // namespace

int main(int argc, char **argv)
{
    double rate = 2000.0, seconds = 60.0;
    int64_t lo = 1024, hi = 4194304, slot_keys = 8 << 20;
    double pool_gb = 64.0;
    int slots = 2, waiters = 6, closed = 0;
    uint64_t seed = 0x5EED0007ull;
    for (int a = 1; a + 1 < argc; a += 2) {
        const std::string k = argv[a];
        const char *v = argv[a + 1];
        if (k == "--rate") rate = atof(v);
        else if (k == "--seconds") seconds = atof(v);
        else if (k == "--lo") lo = atoll(v);
        else if (k == "--hi") hi = atoll(v);
        else if (k == "--pool-gb") pool_gb = atof(v);
        else if (k == "--slot-keys") slot_keys = atoll(v);
        else if (k == "--slots") slots = atoi(v);
        else if (k == "--waiters") waiters = atoi(v);
        else if (k == "--closed") closed = atoi(v);
        else if (k == "--seed") seed = strtoull(v, nullptr, 0);
        else {
            fprintf(stderr, "unknown option %s\n", k.c_str());
            return 2;
        }
    }
    // one source array per key type; a request copies a slice of it
    const int64_t SRC = hi + 65536;
    std::vector<std::vector<uint8_t>> src(4);
    for (int t = 0; t < 4; t++) {
        const int kb = turtle_dtype_size(TYPES[t]);
        src[t].resize((size_t)SRC * kb);
        for (int64_t i = 0; i < SRC; i++) {
            const uint64_t x = draw(seed + 101 + t, (uint64_t)i);
            memcpy(&src[t][(size_t)i * kb], &x, kb);   // low bytes: uniform over the type's range
        }
    }
    // prefix sums of every source array (wrapping, signed types sign-extended),
    // so a slice's key sum costs two lookups
    std::vector<std::vector<uint64_t>> pre(4);
    for (int t = 0; t < 4; t++) {
        const int kb = turtle_dtype_size(TYPES[t]);
        pre[t].assign((size_t)SRC + 1, 0);
        for (int64_t i = 0; i < SRC; i++) {
            uint64_t x = 0;
            memcpy(&x, &src[t][(size_t)i * kb], kb);
            if (TYPES[t] != TURTLE_UINT8 && kb < 8 && ((x >> (8 * kb - 1)) & 1)) x |= ~0ull << (8 * kb);
            pre[t][i + 1] = pre[t][i] + x;
        }
    }
    turtle_server_config cfg{};
    cfg.pool_bytes = (int64_t)(pool_gb * (double)(1ull << 30));
    cfg.slot_keys = slot_keys;
    cfg.slots_per_dtype = slots;
    for (int t = 0; t < 4; t++) cfg.dtype_mask |= 1 << TYPES[t];
    turtle_server_handle sv = nullptr;
    int rc = turtle_server_create(&cfg, &sv);
    if (rc) {
        fprintf(stderr, "turtle_server_create: %s\n", turtle_status_string(rc));
        return 1;
    }
    auto type_of = [](int dtype) {
        return dtype == TURTLE_UINT8 ? 0 : dtype == TURTLE_INT16 ? 1 : dtype == TURTLE_INT32 ? 2 : 3;
    };
    auto check = [&](const Req &r) {
        const int t = type_of(r.dtype);
        switch (r.dtype) {
        case TURTLE_UINT8: return check_keys<uint8_t>(r, (const uint8_t *)src[t].data());
        case TURTLE_INT16: return check_keys<int16_t>(r, (const int16_t *)src[t].data());
        case TURTLE_INT32: return check_keys<int32_t>(r, (const int32_t *)src[t].data());
        default: return check_keys<int64_t>(r, (const int64_t *)src[t].data());
        }
    };
    const double la0 = std::log((double)lo), lb0 = std::log((double)hi);
    // request i of the stream: size log-uniform in [lo, hi], type uniform,
    // index on half of them, a random slice of the type's source array
    auto make = [&](uint64_t i, Req &r) {
        const int64_t n = (int64_t)std::floor(std::exp(la0 + unit(seed, 4 * i + 1) * (lb0 - la0)));
        r.n = std::min(std::max(n, lo), hi);
        const uint64_t coin = draw(seed, 4 * i + 2);
        r.dtype = TYPES[coin & 3];
        r.kb = turtle_dtype_size(r.dtype);
        r.src_off = (int64_t)(draw(seed, 4 * i + 3) % (uint64_t)(SRC - r.n + 1));
        const int t = type_of(r.dtype);
        r.keys = malloc((size_t)r.n * r.kb);
        memcpy(r.keys, &src[t][(size_t)r.src_off * r.kb], (size_t)r.n * r.kb);
        r.sum = pre[t][r.src_off + r.n] - pre[t][r.src_off];
        r.idx = nullptr;
        if ((coin >> 2) & 1) {
            r.idx = (int32_t *)malloc((size_t)r.n * 4);
            for (int64_t j = 0; j < r.n; j++) r.idx[j] = (int32_t)j;
        }
    };
    if (closed > 0) {
        // closed loop: `closed` clients, each submitting and waiting in turn
        std::atomic<uint64_t> next{0};
        std::atomic<int> bad_c{0};
        std::vector<std::vector<double>> lats(closed);
        std::vector<uint64_t> keys_done(closed, 0);
        const double t0c = now_s(), t_stop = t0c + seconds;
        std::vector<std::thread> cl;
        for (int c = 0; c < closed; c++) {
            cl.emplace_back([&, c] {
                while (now_s() < t_stop) {
                    Req r;
                    r.ticket = next.fetch_add(1);
                    make(r.ticket, r);
                    uint64_t tk = 0;
                    int e = turtle_server_submit(sv, r.dtype, r.keys, r.idx, r.n, &tk);
                    if (!e) e = turtle_server_wait_info(sv, tk, &r.info);
                    if (e || !check(r)) bad_c++;
                    else {
                        lats[c].push_back(r.info.latency_seconds);
                        keys_done[c] += (uint64_t)r.n;
                    }
                    free(r.keys);
                    free(r.idx);
                }
            });
        }
        for (auto &t : cl) t.join();
        const double el = now_s() - t0c;
        turtle_server_stats st{};
        turtle_server_get_stats(sv, &st);
        turtle_server_destroy(sv);
        std::vector<double> l;
        uint64_t kd = 0;
        for (int c = 0; c < closed; c++) {
            l.insert(l.end(), lats[c].begin(), lats[c].end());
            kd += keys_done[c];
        }
        std::sort(l.begin(), l.end());
        const size_t N = l.size();
        auto pct = [&](double p) { return N ? l[std::min(N - 1, (size_t)std::floor(p * (N - 1) + 0.5))] * 1e3 : 0.0; };
        printf("{\"mode\": \"closed\", \"clients\": %d, \"requests\": %zu, \"keys\": %llu, \"elapsed_s\": %.3f, "
               "\"Gkeys_s\": %.5f, \"requests_s\": %.1f, \"p50_ms\": %.4f, \"p99_ms\": %.4f, \"max_ms\": %.4f, "
               "\"errors\": %d, \"batches\": %llu, \"max_in_flight_batches\": %u, \"max_bytes_in_flight\": %llu}\n",
               closed, N, (unsigned long long)kd, el, kd / el / 1e9, N / el, pct(0.5), pct(0.99),
               N ? l.back() * 1e3 : 0.0, bad_c.load(), (unsigned long long)st.batches, st.max_in_flight_batches,
               (unsigned long long)st.max_bytes_in_flight);
        return bad_c.load() ? 1 : 0;
    }
    // the schedule: arrival times, sizes, types, index coins, source offsets
    std::vector<double> arrive;
    for (uint64_t i = 0;; i++) {
        const double gap = -std::log(1.0 - unit(seed, 4 * i)) / rate;
        const double t = (arrive.empty() ? 0.0 : arrive.back()) + gap;
        if (t >= seconds) break;
        arrive.push_back(t);
    }
    const size_t N = arrive.size();
    std::vector<Req> reqs(N);
    const double la = std::log((double)lo), lb = std::log((double)hi);
    for (size_t i = 0; i < N; i++) {
        Req &r = reqs[i];
        int64_t n = (int64_t)std::floor(std::exp(la + unit(seed, 4 * i + 1) * (lb - la)));
        r.n = std::min(std::max(n, lo), hi);
        const uint64_t coin = draw(seed, 4 * i + 2);
        r.dtype = TYPES[coin & 3];
        r.kb = turtle_dtype_size(r.dtype);
        r.idx = nullptr;
        r.src_off = (int64_t)(draw(seed, 4 * i + 3) % (uint64_t)(SRC - r.n + 1));
        r.status = (coin >> 2) & 1;   // index coin, replaced by the status later
    }
    std::atomic<size_t> prepared{0}, submitted{0};
    std::atomic<int> bad{0};
    std::vector<double> lat(N, 0.0);
    const double t0 = now_s() + 0.05;
    double lag_max = 0.0;
    auto type_slot = [](int dtype) {
        return dtype == TURTLE_UINT8 ? 0 : dtype == TURTLE_INT16 ? 1 : dtype == TURTLE_INT32 ? 2 : 3;
    };
    // the request arrays are made ahead of their arrival by one thread (at
    // most 512 requests ahead) and submitted at their arrival time by another
    std::thread preparer([&] {
        for (size_t i = 0; i < N; i++) {
            while (i >= submitted.load(std::memory_order_acquire) + 512) std::this_thread::sleep_for(std::chrono::microseconds(200));
            Req &r = reqs[i];
            const int t = type_slot(r.dtype);
            const bool want_idx = r.status != 0;
            r.status = 0;
            r.keys = malloc((size_t)r.n * r.kb);
            memcpy(r.keys, &src[t][(size_t)r.src_off * r.kb], (size_t)r.n * r.kb);
            r.sum = pre[t][r.src_off + r.n] - pre[t][r.src_off];
            r.idx_flag = want_idx ? 1 : 0;
            if (want_idx) {
                r.idx = (int32_t *)malloc((size_t)r.n * 4);
                for (int64_t j = 0; j < r.n; j++) r.idx[j] = (int32_t)j;
            }
            prepared.store(i + 1, std::memory_order_release);
        }
    });
    std::thread submitter([&] {
        for (size_t i = 0; i < N; i++) {
            Req &r = reqs[i];
            while (prepared.load(std::memory_order_acquire) <= i) std::this_thread::sleep_for(std::chrono::microseconds(50));
            const double due = t0 + arrive[i];
            for (double w = now_s(); w < due; w = now_s())
                std::this_thread::sleep_for(std::chrono::duration<double>(std::min(due - w, 0.001)));
            const double lag = now_s() - due;
            if (lag > lag_max) lag_max = lag;
            const int e = turtle_server_submit(sv, r.dtype, r.keys, r.idx, r.n, &r.ticket);
            if (e) {
                r.status = e;
                bad++;
            }
            submitted.store(i + 1, std::memory_order_release);
        }
    });
    std::vector<std::thread> ws;
    for (int w = 0; w < waiters; w++) {
        ws.emplace_back([&, w] {
            for (size_t i = (size_t)w; i < N; i += (size_t)waiters) {
                while (submitted.load(std::memory_order_acquire) <= i) std::this_thread::sleep_for(std::chrono::microseconds(200));
                Req &r = reqs[i];
                if (r.status == 0) {
                    const int e = turtle_server_wait_info(sv, r.ticket, &r.info);
                    r.lat = r.info.latency_seconds;
                    lat[i] = r.lat;
                    bool ok = e == 0;
                    const int t = type_slot(r.dtype);
                    if (ok) {
                        switch (r.dtype) {
                        case TURTLE_UINT8: ok = check_keys<uint8_t>(r, (const uint8_t *)src[t].data()); break;
                        case TURTLE_INT16: ok = check_keys<int16_t>(r, (const int16_t *)src[t].data()); break;
                        case TURTLE_INT32: ok = check_keys<int32_t>(r, (const int32_t *)src[t].data()); break;
                        default: ok = check_keys<int64_t>(r, (const int64_t *)src[t].data()); break;
                        }
                    }
                    if (!ok) bad++;
                }
                free(r.keys);
                free(r.idx);
                r.keys = nullptr;
                r.idx = nullptr;
            }
        });
    }
    preparer.join();
    submitter.join();
    for (auto &t : ws) t.join();
    const double t_end = now_s();
    turtle_server_stats st{};
    turtle_server_get_stats(sv, &st);
    turtle_server_destroy(sv);
    std::vector<double> l(lat);
    std::sort(l.begin(), l.end());
    auto pct = [&](double p) { return N ? l[std::min(N - 1, (size_t)std::floor(p * (N - 1) + 0.5))] * 1e3 : 0.0; };
    double mean = 0.0;
    uint64_t keys = 0;
    int64_t by_type[4] = {0, 0, 0, 0};
    for (size_t i = 0; i < N; i++) {
        mean += lat[i];
        keys += (uint64_t)reqs[i].n;
        by_type[type_slot(reqs[i].dtype)]++;
    }
    mean = N ? mean / N : 0.0;
    const double elapsed = t_end - t0;
    // the five slowest requests, with where their time went
    std::vector<size_t> order(N);
    for (size_t i = 0; i < N; i++) order[i] = i;
    std::sort(order.begin(), order.end(), [&](size_t a, size_t b) { return lat[a] > lat[b]; });
    std::string slow = "[";
    for (size_t j = 0; j < std::min<size_t>(5, N); j++) {
        const Req &r = reqs[order[j]];
        char buf[320];
        snprintf(buf, sizeof buf,
                 "%s{\"arrival_s\": %.4f, \"n\": %lld, \"key_bytes\": %d, \"index\": %d, \"latency_ms\": %.3f, "
                 "\"queued_ms\": %.3f, \"batch_ms\": %.3f, \"batch_keys\": %lld, \"batch_requests\": %d}",
                 j ? ", " : "", arrive[order[j]], (long long)r.n, r.kb, r.idx_flag, r.info.latency_seconds * 1e3,
                 r.info.queued_seconds * 1e3, r.info.batch_seconds * 1e3, (long long)r.info.batch_keys,
                 r.info.batch_requests);
        slow += buf;
    }
    slow += "]";
    printf("{\"requests\": %zu, \"keys\": %llu, \"seconds_offered\": %.3f, \"elapsed_s\": %.3f, "
           "\"offered_Gkeys_s\": %.5f, \"Gkeys_s\": %.5f, \"p50_ms\": %.4f, \"p99_ms\": %.4f, \"mean_ms\": %.4f, "
           "\"max_ms\": %.4f, \"submit_lag_max_ms\": %.3f, \"errors\": %d, \"batches\": %llu, "
           "\"max_in_flight_batches\": %u, \"slots\": %u, \"slot_keys\": %llu, \"pool_bytes\": %llu, "
           "\"busy_seconds\": %.3f, \"requests_by_type\": {\"uint8\": %lld, \"int16\": %lld, \"int32\": %lld, "
           "\"int64\": %lld}, \"rate\": %.1f, \"max_bytes_in_flight\": %llu, \"slowest\": %s}\n",
           N, (unsigned long long)keys, seconds, elapsed, keys / seconds / 1e9, keys / elapsed / 1e9, pct(0.50),
           pct(0.99), mean * 1e3, N ? l.back() * 1e3 : 0.0, lag_max * 1e3, bad.load(),
           (unsigned long long)st.batches, st.max_in_flight_batches, st.slots, (unsigned long long)st.slot_keys,
           (unsigned long long)st.pool_bytes, st.busy_seconds, (long long)by_type[0], (long long)by_type[1],
           (long long)by_type[2], (long long)by_type[3], rate, (unsigned long long)st.max_bytes_in_flight, slow.c_str());
    return bad.load() ? 1 : 0;
}
