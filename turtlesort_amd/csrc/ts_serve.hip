// ts_serve.hip -- request loop with continuous batching (no reference
// counterpart: the reference supports one in-flight sort per TurtleSort
// object, turtle-sort.cuh:53-61, with no admission or batching).
//
// The device memory budget is carved into batch slots when the server is
// created: per served key type, `slots_per_dtype` sorter handles (each with
// its own HIP stream and key/index buffers for `slot_keys` keys) plus a
// pinned host buffer of the same capacity.  Every slot has a worker thread.
// Each request is charged its device bytes (request_bytes) against the
// in-flight budget while its batch runs.  A worker that is free takes, in
// arrival order, every pending request of its key type that still fits the
// slot and the budget, copies them back to back into its
// pinned buffer, sorts them as the segments of one engine run
// (turtle_segmented_sort_device), copies the result back into each
// request's arrays and signals completion.  Requests arriving while other
// batches are on the GPU wait only until a slot of their type is free, and
// batches of different slots run on different streams, so their kernels
// overlap on the device.  Each request is one segment of its batch and the
// engine is deterministic per segment, so a result does not depend on what
// else was in the batch.
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../../include/turtle_c.h"

namespace {

double now_s()
{
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

int key_bytes(int dtype) { return turtle_dtype_size(dtype); }

// device bytes of a request while its batch runs: input and output keys and
// indices of the slot, the engine's ping-pong keys and indices, per-key bins
int64_t request_bytes(int dtype, int64_t n) { return n * (4 * (int64_t)key_bytes(dtype) + 17); }

struct Request {
    uint64_t ticket = 0;
    int dtype = 0;
    void *keys = nullptr;
    int32_t *idx = nullptr;
    int64_t n = 0;
    double t_submit = 0.0, t_start = 0.0, t_done = 0.0;
    int64_t batch_keys = 0;
    int32_t batch_requests = 0;
    int status = TURTLE_OK;
    bool done = false;
};

struct Slot {
    int dtype = 0, kb = 0;
    turtle_handle h = nullptr;
    hipStream_t stream = nullptr;
    void *dkin = nullptr, *dkout = nullptr;
    int32_t *diin = nullptr, *diout = nullptr;
    void *pin_keys = nullptr;
    int32_t *pin_idx = nullptr;
    std::thread worker;
};

}  // namespace

struct turtle_server {
    turtle_server_config cfg{};
    std::vector<std::unique_ptr<Slot>> slots;
    std::mutex mu;
    std::condition_variable cv_work;   // new requests / shutdown
    std::condition_variable cv_done;   // a request finished
    std::deque<std::shared_ptr<Request>> pending[TURTLE_FLOAT64 + 1];
    std::unordered_map<uint64_t, std::shared_ptr<Request>> live;
    uint64_t next_ticket = 1;
    bool stop = false;
    int device = 0;
    turtle_server_stats st{};
    int64_t max_keys[TURTLE_FLOAT64 + 1] = {0};
    int64_t inflight = 0;           // bytes of the requests being sorted
};

namespace {

void slot_free(Slot *s)
{
    if (s->h) turtle_destroy(s->h);
    void *d[] = {s->dkin, s->dkout, s->diin, s->diout};
    for (void *p : d)
        if (p) (void)hipFree(p);
    if (s->pin_keys) (void)hipHostFree(s->pin_keys);
    if (s->pin_idx) (void)hipHostFree(s->pin_idx);
    if (s->stream) (void)hipStreamDestroy(s->stream);
}

// sorts one batch on the slot; returns a turtle status
int run_batch(turtle_server *sv, Slot *s, std::vector<std::shared_ptr<Request>> &batch)
{
    const size_t kb = (size_t)s->kb;
    std::vector<int64_t> off(batch.size() + 1, 0);
    bool any_idx = false;
    for (size_t i = 0; i < batch.size(); i++) {
        off[i + 1] = off[i] + batch[i]->n;
        any_idx |= batch[i]->idx != nullptr;
    }
    const int64_t total = off.back();
    for (size_t i = 0; i < batch.size(); i++) {
        const Request &r = *batch[i];
        memcpy((char *)s->pin_keys + off[i] * kb, r.keys, (size_t)r.n * kb);
        if (r.idx) memcpy(s->pin_idx + off[i], r.idx, (size_t)r.n * 4);
    }
    if (hipSetDevice(sv->device) != hipSuccess) return TURTLE_ERR_DEVICE;
    if (hipMemcpyAsync(s->dkin, s->pin_keys, (size_t)total * kb, hipMemcpyHostToDevice, s->stream) != hipSuccess)
        return TURTLE_ERR_DEVICE;
    if (any_idx &&
        hipMemcpyAsync(s->diin, s->pin_idx, (size_t)total * 4, hipMemcpyHostToDevice, s->stream) != hipSuccess)
        return TURTLE_ERR_DEVICE;
    int rc = turtle_segmented_sort_device(s->h, s->dkin, any_idx ? s->diin : nullptr, s->dkout,
                                          any_idx ? s->diout : nullptr, off.data(), (int64_t)batch.size(), s->stream);
    // turtle_synchronize adds levels, or grows the tables and sorts the batch
    // again from its (untouched) input, when the first levels were not enough
    if (rc == TURTLE_OK) rc = turtle_synchronize(s->h, nullptr);
    if (rc != TURTLE_OK) return rc;
    if (hipMemcpyAsync(s->pin_keys, s->dkout, (size_t)total * kb, hipMemcpyDeviceToHost, s->stream) != hipSuccess)
        return TURTLE_ERR_DEVICE;
    if (any_idx &&
        hipMemcpyAsync(s->pin_idx, s->diout, (size_t)total * 4, hipMemcpyDeviceToHost, s->stream) != hipSuccess)
        return TURTLE_ERR_DEVICE;
    if (hipStreamSynchronize(s->stream) != hipSuccess) return TURTLE_ERR_DEVICE;
    for (size_t i = 0; i < batch.size(); i++) {
        Request &r = *batch[i];
        memcpy(r.keys, (const char *)s->pin_keys + off[i] * kb, (size_t)r.n * kb);
        if (r.idx) {
            // the engine carries the index values; a request's indices are
            // its own (they were copied in unchanged)
            memcpy(r.idx, s->pin_idx + off[i], (size_t)r.n * 4);
        }
    }
    return TURTLE_OK;
}

void worker_loop(turtle_server *sv, Slot *s)
{
    (void)hipSetDevice(sv->device);
    for (;;) {
        std::vector<std::shared_ptr<Request>> batch;
        int64_t batch_bytes = 0, batch_keys = 0;
        double t_start = 0.0;
        {
            std::unique_lock<std::mutex> lk(sv->mu);
            auto &q = sv->pending[s->dtype];
            const int64_t budget = sv->cfg.inflight_bytes;
            // the first pending request must fit the bytes left in flight
            sv->cv_work.wait(lk, [&] {
                return (sv->stop && q.empty()) ||
                       (!q.empty() && sv->inflight + request_bytes(s->dtype, q.front()->n) <= budget);
            });
            if (sv->stop && q.empty()) return;
            // every pending request of this type that still fits the slot
            // and the budget, in order
            int64_t room = sv->cfg.slot_keys;
            while (!q.empty() && q.front()->n <= room &&
                   sv->inflight + batch_bytes + request_bytes(s->dtype, q.front()->n) <= budget) {
                room -= q.front()->n;
                batch_bytes += request_bytes(s->dtype, q.front()->n);
                batch_keys += q.front()->n;
                batch.push_back(q.front());
                q.pop_front();
            }
            sv->inflight += batch_bytes;
            sv->st.bytes_in_flight = (uint64_t)sv->inflight;
            if (sv->st.bytes_in_flight > sv->st.max_bytes_in_flight) sv->st.max_bytes_in_flight = sv->st.bytes_in_flight;
            t_start = now_s();
            for (auto &r : batch) {
                r->t_start = t_start;
                r->batch_keys = batch_keys;
                r->batch_requests = (int32_t)batch.size();
            }
            sv->st.batches++;
            sv->st.in_flight_batches++;
            if (sv->st.in_flight_batches > sv->st.max_in_flight_batches)
                sv->st.max_in_flight_batches = sv->st.in_flight_batches;
        }
        const int rc = run_batch(sv, s, batch);
        const double t1 = now_s();
        {
            std::lock_guard<std::mutex> lk(sv->mu);
            sv->st.in_flight_batches--;
            sv->st.busy_seconds += t1 - t_start;
            sv->inflight -= batch_bytes;
            sv->st.bytes_in_flight = (uint64_t)sv->inflight;
            for (auto &r : batch) {
                r->status = rc;
                r->done = true;
                r->t_done = t1;
                sv->st.completed++;
                sv->st.keys += (uint64_t)r->n;
                if (rc != TURTLE_OK) sv->st.failed++;
            }
        }
        sv->cv_done.notify_all();
        sv->cv_work.notify_all();   // bytes were freed: other workers may admit now
    }
}

}  // namespace

extern "C" {

int turtle_server_create(const turtle_server_config *cfg, turtle_server_handle *out)
{
    if (!out) return TURTLE_ERR_INVALID;
    *out = nullptr;
    turtle_server_config c{};
    if (cfg) c = *cfg;
    if (c.pool_bytes <= 0) c.pool_bytes = (int64_t)64 << 30;
    if (c.slots_per_dtype <= 0) c.slots_per_dtype = 2;
    if (c.dtype_mask == 0) c.dtype_mask = 0x7fe;   // all ten key types
    if (c.slots_per_dtype > 16 || (c.dtype_mask & ~0x7fe)) return TURTLE_ERR_INVALID;
    // device bytes per key of a slot: 4 key buffers of the handle, 4 index
    // buffers, and the server's in/out key and index buffers
    int64_t per_key_sum = 0;
    for (int d = TURTLE_INT8; d <= TURTLE_FLOAT64; d++)
        if (c.dtype_mask & (1 << d)) per_key_sum += (int64_t)c.slots_per_dtype * (6 * key_bytes(d) + 6 * 4);
    if (c.slot_keys <= 0) {
        c.slot_keys = c.pool_bytes / per_key_sum;
        if (c.slot_keys > ((int64_t)32 << 20)) c.slot_keys = (int64_t)32 << 20;   // 32M keys per batch
    }
    if (c.slot_keys < 1 || c.slot_keys >= ((int64_t)1 << 31)) return TURTLE_ERR_INVALID;
    if (c.slot_keys * per_key_sum > c.pool_bytes) return TURTLE_ERR_BUDGET;
    if (c.inflight_bytes <= 0) c.inflight_bytes = c.pool_bytes;
    turtle_server *sv = new turtle_server();
    sv->cfg = c;
    (void)hipGetDevice(&sv->device);
    int rc = TURTLE_OK;
    for (int d = TURTLE_INT8; d <= TURTLE_FLOAT64 && rc == TURTLE_OK; d++) {
        if (!(c.dtype_mask & (1 << d))) continue;
        sv->max_keys[d] = c.slot_keys;
        for (int k = 0; k < c.slots_per_dtype && rc == TURTLE_OK; k++) {
            std::unique_ptr<Slot> s(new Slot());
            s->dtype = d;
            s->kb = key_bytes(d);
            const size_t kbytes = (size_t)c.slot_keys * s->kb, ibytes = (size_t)c.slot_keys * 4;
            rc = turtle_create(d, c.slot_keys, 1, 0, &s->h);
            if (rc == TURTLE_OK &&
                (hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking) != hipSuccess ||
                 hipMalloc(&s->dkin, kbytes) != hipSuccess || hipMalloc(&s->dkout, kbytes) != hipSuccess ||
                 hipMalloc((void **)&s->diin, ibytes) != hipSuccess ||
                 hipMalloc((void **)&s->diout, ibytes) != hipSuccess ||
                 hipHostMalloc(&s->pin_keys, kbytes, hipHostMallocDefault) != hipSuccess ||
                 hipHostMalloc((void **)&s->pin_idx, ibytes, hipHostMallocDefault) != hipSuccess))
                rc = TURTLE_ERR_OOM;
            if (rc != TURTLE_OK) {
                slot_free(s.get());
                break;
            }
            // one warm-up run per engine form (with and without an index,
            // at a small and a large size, which take different tunings):
            // the first launch of a kernel pays its loading, which would
            // otherwise land in the first requests' latency
            if (hipMemsetAsync(s->dkin, 0, kbytes, s->stream) != hipSuccess ||
                hipMemsetAsync(s->diin, 0, ibytes, s->stream) != hipSuccess) {
                slot_free(s.get());
                rc = TURTLE_ERR_DEVICE;
                break;
            }
            const int64_t big = c.slot_keys < ((int64_t)4 << 20) ? c.slot_keys : ((int64_t)4 << 20);
            for (int64_t wn : {(int64_t)65536, big}) {
                if (wn > c.slot_keys) continue;
                const int64_t off[2] = {0, wn};
                for (int with_idx = 0; with_idx < 2 && rc == TURTLE_OK; with_idx++) {
                    rc = turtle_segmented_sort_device(s->h, s->dkin, with_idx ? s->diin : nullptr, s->dkout,
                                                      with_idx ? s->diout : nullptr, off, 1, s->stream);
                    if (rc == TURTLE_OK) rc = turtle_synchronize(s->h, nullptr);
                }
            }
            if (rc != TURTLE_OK) {
                slot_free(s.get());
                break;
            }
            sv->slots.push_back(std::move(s));
        }
    }
    if (rc != TURTLE_OK) {
        for (auto &s : sv->slots) slot_free(s.get());
        delete sv;
        return rc;
    }
    sv->st.pool_bytes = (uint64_t)c.pool_bytes;
    sv->st.inflight_bytes = (uint64_t)c.inflight_bytes;
    sv->st.slot_keys = (uint64_t)c.slot_keys;
    sv->st.slots = (uint32_t)sv->slots.size();
    for (auto &s : sv->slots) s->worker = std::thread(worker_loop, sv, s.get());
    *out = sv;
    return TURTLE_OK;
}

int turtle_server_submit(turtle_server_handle sv, int dtype, void *keys, int32_t *idx, int64_t n, uint64_t *ticket)
{
    if (!sv || !ticket) return TURTLE_ERR_INVALID;
    *ticket = 0;
    if (dtype < TURTLE_INT8 || dtype > TURTLE_FLOAT64) return TURTLE_ERR_DTYPE;
    if (n < 0 || (n > 0 && !keys)) return TURTLE_ERR_INVALID;
    std::shared_ptr<Request> r(new Request());
    r->dtype = dtype;
    r->keys = keys;
    r->idx = idx;
    r->n = n;
    r->t_submit = now_s();
    {
        std::lock_guard<std::mutex> lk(sv->mu);
        if (sv->max_keys[dtype] == 0) return TURTLE_ERR_DTYPE;   // type not served
        if (n > sv->max_keys[dtype] || request_bytes(dtype, n) > sv->cfg.inflight_bytes) {
            sv->st.rejected++;
            return TURTLE_ERR_BUDGET;
        }
        r->ticket = sv->next_ticket++;
        sv->st.submitted++;
        sv->live[r->ticket] = r;
        if (n == 0) {
            r->done = true;
            r->t_done = r->t_submit;
            sv->st.completed++;
        } else {
            sv->pending[dtype].push_back(r);
        }
    }
    sv->cv_work.notify_all();
    *ticket = r->ticket;
    return TURTLE_OK;
}

int turtle_server_wait_info(turtle_server_handle sv, uint64_t ticket, turtle_request_info *info)
{
    if (!sv) return TURTLE_ERR_INVALID;
    std::unique_lock<std::mutex> lk(sv->mu);
    auto it = sv->live.find(ticket);
    if (it == sv->live.end()) return TURTLE_ERR_STATE;
    std::shared_ptr<Request> r = it->second;
    sv->cv_done.wait(lk, [&] { return r->done; });
    sv->live.erase(ticket);
    if (info) {
        info->latency_seconds = r->t_done - r->t_submit;
        info->queued_seconds = r->n ? r->t_start - r->t_submit : 0.0;
        info->batch_seconds = r->n ? r->t_done - r->t_start : 0.0;
        info->batch_keys = r->batch_keys;
        info->batch_requests = r->batch_requests;
        info->status = r->status;
    }
    return r->status;
}

int turtle_server_wait(turtle_server_handle sv, uint64_t ticket, double *latency_seconds)
{
    turtle_request_info info{};
    const int rc = turtle_server_wait_info(sv, ticket, &info);
    if (latency_seconds && rc != TURTLE_ERR_STATE && rc != TURTLE_ERR_INVALID) *latency_seconds = info.latency_seconds;
    return rc;
}

int turtle_server_test(turtle_server_handle sv, uint64_t ticket, int *done)
{
    if (!sv || !done) return TURTLE_ERR_INVALID;
    std::lock_guard<std::mutex> lk(sv->mu);
    auto it = sv->live.find(ticket);
    if (it == sv->live.end()) return TURTLE_ERR_STATE;
    *done = it->second->done ? 1 : 0;
    return TURTLE_OK;
}

int turtle_server_get_stats(turtle_server_handle sv, turtle_server_stats *out)
{
    if (!sv || !out) return TURTLE_ERR_INVALID;
    std::lock_guard<std::mutex> lk(sv->mu);
    *out = sv->st;
    return TURTLE_OK;
}

int turtle_server_destroy(turtle_server_handle sv)
{
    if (!sv) return TURTLE_OK;
    {
        std::lock_guard<std::mutex> lk(sv->mu);
        sv->stop = true;
    }
    sv->cv_work.notify_all();
    for (auto &s : sv->slots)
        if (s->worker.joinable()) s->worker.join();
    for (auto &s : sv->slots) slot_free(s.get());
    delete sv;
    return TURTLE_OK;
}

}  // extern "C"
