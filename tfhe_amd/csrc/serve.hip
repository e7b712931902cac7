// Serving loop: continuous batching of gate-circuit requests from many
// clients, each with its own cloud key resident in HBM.
//
// A request is a gate netlist (the shape of the reference's random circuits,
// test/test-long-run.cpp:61-185, and adders, test/test-addition-boot.cpp:30-111)
// with host input ciphertexts; the caller receives its output ciphertexts.
// One scheduler thread owns every HIP call of the server:
//   - admission: requests wait in arrival order until their wires fit in the
//     device wire arena (the HBM budget given at creation); admitted requests
//     get a contiguous range of ciphertext slots and their inputs are copied in;
//   - batching: for each client, the next dependency level of every admitted
//     request of that client (as far as the batch limit allows) goes into one
//     launch, so requests that arrive while others run join the next launch;
//   - overlap: clients are spread over several streams and each client keeps up
//     to `depth` launches in flight; launches of one client share a stream, so a
//     request's levels execute in order without host round trips;
//   - completion: when a request's last launch has finished, its output wires
//     are gathered and copied to the caller's buffer and its slots are freed.
// Every gate is computed exactly as in a stand-alone batch (the kernels give
// each bootstrap the same result whatever the batch composition), so a
// request's outputs are bit-identical alone and in any mix.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "engine.h"

namespace tfhe_amd {
namespace {

using Clock = std::chrono::steady_clock;

// out[r] = pool[rows[r]] for r < nrows, one workgroup per row of w int32.
__global__ void __launch_bounds__(256) gather_rows_kernel(const int32_t* __restrict__ pool,
                                                          const int32_t* __restrict__ rows, int nrows, int w,
                                                          int32_t* __restrict__ out) {
    const int r = blockIdx.x;
    if (r >= nrows) return;
    const int32_t* src = pool + (size_t)rows[r] * w;
    int32_t* dst = out + (size_t)r * w;
    for (int i = threadIdx.x; i < w; i += blockDim.x) dst[i] = src[i];
}

struct Request {
    int64_t id = 0;
    int client = 0;
    int32_t W = 0, G = 0, n_in = 0, n_out = 0;
    std::vector<int32_t> gates;      // [5][G] op, in0, in1, in2, out (local slots), sorted by level
    std::vector<int32_t> level_end;  // end (exclusive) of each level in the sorted order
    std::vector<int32_t> inputs;     // [n_in][w]; local slots 0..n_in-1
    std::vector<int32_t> out_rows;   // local slot of each output
    int32_t* out_dst = nullptr;      // caller's [n_out][w]
    size_t slot0 = 0, slots = 0;     // arena range: W wires, n_out staged outputs, row list
    int level = 0;
    int32_t pos = 0;                 // next gate to launch (sorted order)
    int inflight = 0;
    bool done = false;
    int status = 0;
    std::string error;
    Clock::time_point t_submit, t_done;
};

// A launch's completion event, shared by every client whose gates it carries.
struct EvBox {
    hipEvent_t e = nullptr;
    ~EvBox() {
        if (e) (void)hipEventDestroy(e);
    }
};

struct Launch {
    std::shared_ptr<EvBox> ev;
    std::vector<Request*> reqs;
    int ring = 0;  // the client's index-ring slot; -1 for a packed launch
};

struct Client {
    tfhe_amd_device_key* dk = nullptr;
    int stream = 0;
    std::vector<Request*> active;  // admission order
    std::deque<Launch> inflight;
    int32_t* h_idx = nullptr;      // pinned [depth][5*max_batch]
    int32_t* d_idx = nullptr;      // device [depth][5*max_batch]
    int next_ring = 0;
    int last_stream = 0;           // stream of the newest in-flight launch (-1: the packing stream)
};

// Gates of one client's requests in one launch: request r's sorted gates [from, to).
struct Part {
    Request* r;
    int32_t from, to;
};

// The next level (or the rest of it) of each admitted request of the client, in
// admission order, up to `limit` gates. Changes nothing.
int gather_parts(const Client& C, int limit, std::vector<Part>& parts) {
    int count = 0;
    for (Request* r : C.active) {
        if (r->pos >= r->G) continue;
        const int32_t end = r->level_end[r->level];
        const int take = std::min(end - r->pos, limit - count);
        if (take <= 0) break;
        parts.push_back({r, r->pos, r->pos + take});
        count += take;
    }
    return count;
}

// Index rows of the parts into h[5][stride] (op, in0, in1, in2, out) from column
// `at` on, wires as absolute arena slots.
void write_parts(const std::vector<Part>& parts, int32_t* h, int stride, int at) {
    int k = at;
    for (const Part& p : parts) {
        const Request* r = p.r;
        const int32_t G = r->G;
        const int32_t s0 = (int32_t)r->slot0;
        for (int32_t g = p.from; g < p.to; g++, k++) {
            h[k] = r->gates[g];
            h[stride + k] = s0 + r->gates[G + g];
            h[2 * stride + k] = s0 + r->gates[2 * G + g];
            h[3 * stride + k] = s0 + r->gates[3 * G + g];
            h[4 * stride + k] = s0 + r->gates[4 * G + g];
        }
    }
}

// Marks the parts launched under L.
void advance_parts(const std::vector<Part>& parts, Launch& L) {
    for (const Part& p : parts) {
        Request* r = p.r;
        r->pos = p.to;
        if (r->pos == r->level_end[r->level]) r->level++;
        r->inflight++;
        L.reqs.push_back(r);
    }
}

int record_event(hipStream_t s, std::shared_ptr<EvBox>& ev) {
    ev = std::make_shared<EvBox>();
    if (hipEventCreateWithFlags(&ev->e, hipEventDisableTiming) != hipSuccess || hipEventRecord(ev->e, s) != hipSuccess)
        return set_error("server: event record failed");
    return 0;
}

// A segment table holds at most this many clients per packed launch.
constexpr int PACK_SEGMENTS = 32;

// First-fit allocator over the slots of the wire arena.
struct SlotAllocator {
    std::map<size_t, size_t> free_;  // start -> length
    size_t total = 0, used = 0, peak = 0;
    void init(size_t n) {
        total = n;
        free_.clear();
        if (n) free_[0] = n;
    }
    bool alloc(size_t n, size_t* at) {
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second < n) continue;
            *at = it->first;
            const size_t rest = it->second - n, start = it->first + n;
            free_.erase(it);
            if (rest) free_[start] = rest;
            used += n;
            peak = std::max(peak, used);
            return true;
        }
        return false;
    }
    void release(size_t at, size_t n) {
        used -= n;
        auto it = free_.emplace(at, n).first;
        auto nx = std::next(it);
        if (nx != free_.end() && it->first + it->second == nx->first) {
            it->second += nx->second;
            free_.erase(nx);
        }
        if (it != free_.begin()) {
            auto pv = std::prev(it);
            if (pv->first + pv->second == it->first) {
                pv->second += it->second;
                free_.erase(it);
            }
        }
    }
};

}  // namespace
}  // namespace tfhe_amd

using namespace tfhe_amd;

struct tfhe_amd_server {
    int device = 0;
    int n_streams = 4, depth = 2, max_batch = 4096;
    size_t budget = 0;
    int w = 0;  // ciphertext width n+1 (fixed by the first client)
    tfhe_amd_device_key* first_dk = nullptr;
    std::vector<hipStream_t> streams;
    std::vector<char*> scratch;
    std::vector<size_t> scratch_cap;
    hipStream_t io = nullptr;
    int32_t* arena = nullptr;
    SlotAllocator slots;
    int32_t* h_bounce = nullptr;
    size_t bounce_cap = 0;
    std::vector<std::unique_ptr<Client>> clients;
    size_t key_bytes = 0;

    std::mutex mu;  // guards everything below
    int pack_min = 0;  // packing threshold in gates (0: each launch carries one client)
    std::condition_variable cv_work, cv_done;
    std::deque<Request*> incoming;
    std::unordered_map<int64_t, std::unique_ptr<Request>> all;
    int64_t next_id = 0;
    bool stop = false;
    std::string fatal;
    int64_t n_submitted = 0, n_completed = 0, n_failed = 0, n_gates = 0, n_launches = 0, max_launch = 0, n_packed = 0;
    int32_t n_queued = 0, n_active = 0;
    size_t wire_in_use = 0, wire_peak = 0;
    std::thread worker;

    // --- scheduler-thread state (touched only by the worker) ---
    std::deque<Request*> pending;
    std::vector<Client*> worker_clients;  // snapshot of `clients` for the current pass
    // packed launches: their own stream, index ring and workspace
    hipStream_t pack_stream = nullptr;
    int32_t* h_pack = nullptr;  // pinned [depth][5*max_batch]
    int32_t* d_pack = nullptr;  // device [depth][5*max_batch]
    int pack_ring = 0;
    std::deque<std::shared_ptr<EvBox>> pack_inflight;
    char* pack_scr = nullptr;
    size_t pack_scr_cap = 0;

    void run();
    int admit_pending(bool* progress);
    int launch_client(int c, bool packing, bool* progress);
    int launch_packed(int min_gates, bool* progress);
    int poll_client(int c, bool* progress);
    int finish(Request* r);
    int ensure_bounce(size_t bytes);
    void fail_all(const std::string& msg);
};

int tfhe_amd_server::ensure_bounce(size_t bytes) {
    if (bounce_cap >= bytes) return 0;
    if (h_bounce) (void)hipHostFree(h_bounce);
    h_bounce = nullptr;
    bounce_cap = 0;
    const size_t cap = bytes + bytes / 2 + (1 << 16);
    if (hipHostMalloc((void**)&h_bounce, cap, hipHostMallocDefault) != hipSuccess)
        return set_error("server: pinned staging allocation failed");
    bounce_cap = cap;
    return 0;
}

int tfhe_amd_server::admit_pending(bool* progress) {
    while (!pending.empty()) {
        Request* r = pending.front();
        size_t at = 0;
        if (!slots.alloc(r->slots, &at)) return 0;  // arrival order: wait for space
        r->slot0 = at;
        // inputs to slots [slot0, slot0 + n_in); output row list after the staged outputs
        const size_t in_words = (size_t)r->n_in * w, row_words = (size_t)r->n_out;
        if (int rc = ensure_bounce(sizeof(int32_t) * (in_words + row_words))) return rc;
        if (in_words) memcpy(h_bounce, r->inputs.data(), sizeof(int32_t) * in_words);
        for (int j = 0; j < r->n_out; j++) h_bounce[in_words + j] = (int32_t)(r->slot0 + r->out_rows[j]);
        int32_t* base = arena + r->slot0 * w;
        int32_t* rows_dst = arena + (r->slot0 + r->W + r->n_out) * w;
        if ((in_words && hipMemcpyAsync(base, h_bounce, sizeof(int32_t) * in_words, hipMemcpyHostToDevice, io) != hipSuccess) ||
            (row_words && hipMemcpyAsync(rows_dst, h_bounce + in_words, sizeof(int32_t) * row_words,
                                         hipMemcpyHostToDevice, io) != hipSuccess) ||
            hipStreamSynchronize(io) != hipSuccess)
            return set_error("server: input upload failed");
        std::vector<int32_t>().swap(r->inputs);
        pending.pop_front();
        worker_clients[r->client]->active.push_back(r);
        {
            std::lock_guard<std::mutex> g(mu);
            n_queued--;
            n_active++;
            wire_in_use = slots.used * (size_t)w * sizeof(int32_t);
            wire_peak = std::max(wire_peak, wire_in_use);
        }
        *progress = true;
    }
    return 0;
}

int tfhe_amd_server::launch_client(int c, bool packing, bool* progress) {
    Client& C = *worker_clients[c];
    if ((int)C.inflight.size() >= depth) return 0;
    // while packing, a client whose newest launch is a packed one waits for the next packed launch
    if (packing && !C.inflight.empty() && C.last_stream < 0) return 0;
    std::vector<Part> parts;
    const int count = gather_parts(C, max_batch, parts);
    if (!count) return 0;
    const int ring = C.next_ring;
    int32_t* h = C.h_idx + (size_t)ring * 5 * max_batch;
    write_parts(parts, h, count, 0);
    const int si = C.stream;
    hipStream_t s = streams[si];
    const size_t need = engine_batch_scratch_bytes(C.dk, max_batch);
    if (scratch_cap[si] < need) {
        if (hipStreamSynchronize(s) != hipSuccess) return set_error("server: stream synchronize failed");
        (void)hipFree(scratch[si]);
        scratch[si] = nullptr;
        scratch_cap[si] = 0;
        if (hipMalloc(&scratch[si], need) != hipSuccess) return set_error("server: scratch allocation failed");
        scratch_cap[si] = need;
    }
    // the client's previous level ran in a packed launch: order after it on the device
    if (!C.inflight.empty() && C.last_stream != si && hipStreamWaitEvent(s, C.inflight.back().ev->e, 0) != hipSuccess)
        return set_error("server: stream wait failed");
    int32_t* d = C.d_idx + (size_t)ring * 5 * max_batch;
    if (hipMemcpyAsync(d, h, sizeof(int32_t) * 5 * (size_t)count, hipMemcpyHostToDevice, s) != hipSuccess)
        return set_error("server: index upload failed");
    // other launches still on the device (any client's or a packed one): the kernel choice
    // then leaves out the latency kernel, which takes a whole CU per bootstrap
    size_t others = pack_inflight.size();
    for (const Client* o : worker_clients) others += o->inflight.size();
    {
        std::lock_guard<std::mutex> g(engine_mutex());
        if (int rc = engine_launch_batch_locked(C.dk, count, d, d + count, d + 2 * count, d + 3 * count,
                                                d + 4 * count, arena, scratch[si], s, true, others > 0))
            return rc;
    }
    Launch L;
    L.ring = ring;
    if (int rc = record_event(s, L.ev)) return rc;
    advance_parts(parts, L);
    C.inflight.push_back(std::move(L));
    C.last_stream = si;
    C.next_ring = (ring + 1) % depth;
    {
        std::lock_guard<std::mutex> g(mu);
        n_launches++;
        n_gates += count;
        max_launch = std::max<int64_t>(max_launch, count);
    }
    *progress = true;
    return 0;
}

// One launch carrying the next level of several clients (each with its own key,
// as a segment of the batched blind-rotation kernel), made when at least two
// clients have gates ready and together at least `min_gates`. Packed launches
// run in order on their own stream, at most `depth` in flight.
int tfhe_amd_server::launch_packed(int min_gates, bool* progress) {
    while (!pack_inflight.empty()) {
        const hipError_t e = hipEventQuery(pack_inflight.front()->e);
        if (e == hipErrorNotReady) break;
        if (e != hipSuccess) return set_error("server: packed launch failed: %s", hipGetErrorString(e));
        pack_inflight.pop_front();
    }
    if ((int)pack_inflight.size() >= depth) return 0;
    std::vector<int> cand;
    std::vector<std::vector<Part>> cparts;
    int total = 0;
    const tfhe_amd_device_key* ref = nullptr;
    for (size_t c = 0; c < worker_clients.size() && (int)cand.size() < PACK_SEGMENTS && total < max_batch; c++) {
        const Client& C = *worker_clients[c];
        if ((int)C.inflight.size() >= depth || !engine_can_segment(C.dk)) continue;
        if (ref && !engine_same_params(ref, C.dk)) continue;
        std::vector<Part> parts;
        const int cnt = gather_parts(C, max_batch - total, parts);
        if (!cnt) continue;
        if (!ref) ref = C.dk;
        cand.push_back((int)c);
        cparts.push_back(std::move(parts));
        total += cnt;
    }
    if (cand.size() < 2 || total < min_gates) return 0;
    if (!pack_stream) {
        const size_t idx_bytes = sizeof(int32_t) * 5 * (size_t)max_batch * depth;
        if (hipStreamCreateWithFlags(&pack_stream, hipStreamNonBlocking) != hipSuccess ||
            hipHostMalloc((void**)&h_pack, idx_bytes, hipHostMallocDefault) != hipSuccess ||
            hipMalloc(&d_pack, idx_bytes) != hipSuccess)
            return set_error("server: packing stream or index buffers failed");
    }
    const int nseg = (int)cand.size();
    int32_t* h = h_pack + (size_t)pack_ring * 5 * max_batch;
    std::vector<tfhe_amd_device_key*> dks(nseg);
    std::vector<int32_t> counts(nseg);
    for (int k = 0, at = 0; k < nseg; k++) {
        int cnt = 0;
        for (const Part& p : cparts[k]) cnt += p.to - p.from;
        write_parts(cparts[k], h, total, at);
        dks[k] = worker_clients[cand[k]]->dk;
        counts[k] = cnt;
        at += cnt;
    }
    const size_t need = engine_segments_scratch_bytes(dks.data(), counts.data(), nseg);
    if (pack_scr_cap < need) {
        if (hipStreamSynchronize(pack_stream) != hipSuccess) return set_error("server: stream synchronize failed");
        (void)hipFree(pack_scr);
        pack_scr = nullptr;
        pack_scr_cap = 0;
        const size_t cap = need + need / 4;
        if (hipMalloc(&pack_scr, cap) != hipSuccess) return set_error("server: packing workspace allocation failed");
        pack_scr_cap = cap;
    }
    // each client's previous level may still run on its own stream
    for (int k = 0; k < nseg; k++) {
        const Client& C = *worker_clients[cand[k]];
        if (!C.inflight.empty() && C.last_stream >= 0 &&
            hipStreamWaitEvent(pack_stream, C.inflight.back().ev->e, 0) != hipSuccess)
            return set_error("server: stream wait failed");
    }
    int32_t* d = d_pack + (size_t)pack_ring * 5 * max_batch;
    if (hipMemcpyAsync(d, h, sizeof(int32_t) * 5 * (size_t)total, hipMemcpyHostToDevice, pack_stream) != hipSuccess)
        return set_error("server: index upload failed");
    {
        std::lock_guard<std::mutex> g(engine_mutex());
        if (int rc = engine_launch_segments_locked(dks.data(), counts.data(), nseg, d, arena, pack_scr, pack_stream))
            return rc;
    }
    std::shared_ptr<EvBox> ev;
    if (int rc = record_event(pack_stream, ev)) return rc;
    pack_inflight.push_back(ev);
    pack_ring = (pack_ring + 1) % depth;
    for (int k = 0; k < nseg; k++) {
        Client& C = *worker_clients[cand[k]];
        Launch L;
        L.ring = -1;
        L.ev = ev;
        advance_parts(cparts[k], L);
        C.inflight.push_back(std::move(L));
        C.last_stream = -1;
    }
    {
        std::lock_guard<std::mutex> g(mu);
        n_launches++;
        n_packed++;
        n_gates += total;
        max_launch = std::max<int64_t>(max_launch, total);
    }
    *progress = true;
    return 0;
}

int tfhe_amd_server::finish(Request* r) {
    int32_t* staged = arena + (r->slot0 + r->W) * w;
    const int32_t* rows = arena + (r->slot0 + r->W + r->n_out) * w;
    const size_t bytes = sizeof(int32_t) * (size_t)r->n_out * w;
    if (r->n_out) {
        if (int rc = ensure_bounce(bytes)) return rc;
        hipLaunchKernelGGL(gather_rows_kernel, dim3(r->n_out), dim3(256), 0, io, arena, rows, r->n_out, w, staged);
        if (hipGetLastError() != hipSuccess ||
            hipMemcpyAsync(h_bounce, staged, bytes, hipMemcpyDeviceToHost, io) != hipSuccess ||
            hipStreamSynchronize(io) != hipSuccess)
            return set_error("server: output download failed");
        memcpy(r->out_dst, h_bounce, bytes);
    }
    slots.release(r->slot0, r->slots);
    auto& act = worker_clients[r->client]->active;
    act.erase(std::find(act.begin(), act.end(), r));
    std::lock_guard<std::mutex> g(mu);
    r->t_done = Clock::now();
    r->done = true;
    n_completed++;
    n_active--;
    wire_in_use = slots.used * (size_t)w * sizeof(int32_t);
    cv_done.notify_all();
    return 0;
}

int tfhe_amd_server::poll_client(int c, bool* progress) {
    Client& C = *worker_clients[c];
    while (!C.inflight.empty()) {
        Launch& L = C.inflight.front();
        const hipError_t e = hipEventQuery(L.ev->e);
        if (e == hipErrorNotReady) return 0;
        if (e != hipSuccess) return set_error("server: launch failed: %s", hipGetErrorString(e));
        std::vector<Request*> reqs = std::move(L.reqs);
        C.inflight.pop_front();
        for (Request* r : reqs)
            if (--r->inflight == 0 && r->pos == r->G)
                if (int rc = finish(r)) return rc;
        *progress = true;
    }
    return 0;
}

void tfhe_amd_server::fail_all(const std::string& msg) {
    std::lock_guard<std::mutex> g(mu);
    fatal = msg;
    for (auto& kv : all)
        if (!kv.second->done) {
            kv.second->done = true;
            kv.second->status = -1;
            kv.second->error = msg;
            kv.second->t_done = Clock::now();
            n_failed++;
        }
    incoming.clear();
    n_queued = 0;
    n_active = 0;
    cv_done.notify_all();
}

void tfhe_amd_server::run() {
    if (hipSetDevice(device) != hipSuccess) {
        fail_all("server: hipSetDevice failed");
        return;
    }
    for (;;) {
        size_t ncl;
        int pack_now;
        {
            // clients are appended under the lock; the worker only touches the
            // first ncl of them in this pass (their lists are worker-owned)
            std::lock_guard<std::mutex> g(mu);
            while (!incoming.empty()) {
                pending.push_back(incoming.front());
                incoming.pop_front();
            }
            ncl = clients.size();
            pack_now = pack_min;
            bool busy = !pending.empty();
            for (size_t c = 0; c < ncl; c++) busy = busy || !clients[c]->active.empty() || !clients[c]->inflight.empty();
            if (stop && !busy) break;
            worker_clients.assign(ncl, nullptr);
            for (size_t c = 0; c < ncl; c++) worker_clients[c] = clients[c].get();
        }
        bool progress = false;
        int rc = 0;
        for (size_t c = 0; c < ncl && !rc; c++) rc = poll_client((int)c, &progress);
        if (!rc) rc = admit_pending(&progress);
        if (!rc && pack_now > 0) rc = launch_packed(pack_now, &progress);
        for (size_t c = 0; c < ncl && !rc; c++) rc = launch_client((int)c, pack_now > 0, &progress);
        if (rc) {
            fail_all(tfhe_amd_last_error());
            // drain what is queued on the device before the buffers go away
            for (auto s : streams) (void)hipStreamSynchronize(s);
            if (pack_stream) (void)hipStreamSynchronize(pack_stream);
            for (Client* cl : worker_clients) {
                cl->inflight.clear();
                cl->active.clear();
            }
            pack_inflight.clear();
            pending.clear();
            std::unique_lock<std::mutex> lk(mu);
            cv_work.wait(lk, [&] { return stop; });
            return;
        }
        if (!progress) {
            std::unique_lock<std::mutex> lk(mu);
            cv_work.wait_for(lk, std::chrono::microseconds(100), [&] { return stop || !incoming.empty(); });
        }
    }
}

// ------------------------------------------------------------------ C-ABI --

TFHE_AMD_EXPORT tfhe_amd_server* tfhe_amd_server_create(int32_t n_streams, int32_t depth, int32_t max_batch,
                                                        size_t wire_budget_bytes) {
    if (n_streams <= 0) n_streams = 8;
    if (depth <= 0) depth = 2;
    if (max_batch <= 0) max_batch = 4096;
    if (n_streams > 16 || depth > 8 || max_batch > (1 << 20) || wire_budget_bytes < (1 << 20)) {
        set_error("server_create: n_streams <= 16, depth <= 8, max_batch <= 2^20, budget >= 1 MiB");
        return nullptr;
    }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) {
        set_error("server_create: no GPU device visible");
        return nullptr;
    }
    auto* S = new tfhe_amd_server();
    S->n_streams = n_streams;
    S->depth = depth;
    S->max_batch = max_batch;
    S->budget = wire_budget_bytes;
    bool ok = hipGetDevice(&S->device) == hipSuccess;
    S->streams.assign(n_streams, nullptr);
    S->scratch.assign(n_streams, nullptr);
    S->scratch_cap.assign(n_streams, 0);
    for (int i = 0; ok && i < n_streams; i++) ok = hipStreamCreateWithFlags(&S->streams[i], hipStreamNonBlocking) == hipSuccess;
    ok = ok && hipStreamCreateWithFlags(&S->io, hipStreamNonBlocking) == hipSuccess;
    ok = ok && hipMalloc(&S->arena, wire_budget_bytes) == hipSuccess;
    if (!ok) {
        set_error("server_create: stream or arena allocation failed");
        for (auto s : S->streams)
            if (s) (void)hipStreamDestroy(s);
        if (S->io) (void)hipStreamDestroy(S->io);
        (void)hipFree(S->arena);
        delete S;
        return nullptr;
    }
    S->worker = std::thread([S] { S->run(); });
    return S;
}

TFHE_AMD_EXPORT int32_t tfhe_amd_server_add_client(tfhe_amd_server* S, const TFheGateBootstrappingCloudKeySet* ck) {
    if (!S || !ck) return set_error("server_add_client: bad arguments");
    // pinned: releasing or deleting the client's key set later leaves this device
    // copy allocated until the server is destroyed
    tfhe_amd_device_key* dk = engine_key_pinned(ck->bkFFT);
    if (!dk) return -1;
    std::lock_guard<std::mutex> g(S->mu);
    if (S->first_dk && !engine_same_params(S->first_dk, dk)) {
        engine_key_unpin(dk);
        return set_error("server_add_client: all clients of a server use one parameter set");
    }
    auto cl = std::make_unique<Client>();
    cl->dk = dk;
    cl->stream = (int)S->clients.size() % S->n_streams;
    const size_t idx_bytes = sizeof(int32_t) * 5 * (size_t)S->max_batch * S->depth;
    if (hipHostMalloc((void**)&cl->h_idx, idx_bytes, hipHostMallocDefault) != hipSuccess ||
        hipMalloc(&cl->d_idx, idx_bytes) != hipSuccess) {
        if (cl->h_idx) (void)hipHostFree(cl->h_idx);
        engine_key_unpin(dk);
        return set_error("server_add_client: staging allocation failed");
    }
    if (!S->first_dk) {
        S->first_dk = dk;
        S->w = engine_param_n(dk) + 1;
        S->slots.init(S->budget / (sizeof(int32_t) * (size_t)S->w));
    }
    S->key_bytes += engine_key_bytes(dk);
    S->clients.push_back(std::move(cl));
    return (int32_t)S->clients.size() - 1;
}

TFHE_AMD_EXPORT int64_t tfhe_amd_server_submit(tfhe_amd_server* S, int32_t client, int32_t n_wires, int32_t n_gates,
                                               const int32_t* op, const int32_t* in0, const int32_t* in1,
                                               const int32_t* in2, const int32_t* out, int32_t n_inputs,
                                               const int32_t* input_wires, const int32_t* input_cts,
                                               int32_t n_outputs, const int32_t* output_wires, int32_t* output_cts) {
    if (!S) return set_error("server_submit: null server");
    int w;
    size_t total_slots;
    {
        std::lock_guard<std::mutex> g(S->mu);
        if (client < 0 || client >= (int32_t)S->clients.size()) return set_error("server_submit: unknown client %d", client);
        if (!S->fatal.empty()) return set_error("server_submit: server failed: %s", S->fatal.c_str());
        if (S->stop) return set_error("server_submit: server is stopping");
        w = S->w;
        total_slots = S->slots.total;
    }
    if (n_wires <= 0 || n_gates <= 0 || !op || !in0 || !in1 || !out || n_inputs < 0 || n_outputs < 0 ||
        (n_inputs && (!input_wires || !input_cts)) || (n_outputs && (!output_wires || !output_cts)))
        return set_error("server_submit: bad arguments");
    auto r = std::make_unique<Request>();
    std::vector<int32_t> order, level_count;
    if (int rc = engine_levelize(n_wires, n_gates, op, in0, in1, in2, out, 0, order, level_count)) return rc;
    // local slots: the inputs first (in the order given), then every other wire
    std::vector<int32_t> local(n_wires, -1);
    std::vector<char> driven(n_wires, 0);
    for (int g = 0; g < n_gates; g++) driven[out[g]] = 1;
    for (int j = 0; j < n_inputs; j++) {
        const int32_t x = input_wires[j];
        if (x < 0 || x >= n_wires) return set_error("server_submit: input wire %d out of range", x);
        if (local[x] >= 0) return set_error("server_submit: input wire %d listed twice", x);
        if (driven[x]) return set_error("server_submit: input wire %d is driven by a gate", x);
        local[x] = j;
    }
    int32_t next = n_inputs;
    for (int x = 0; x < n_wires; x++)
        if (local[x] < 0) local[x] = next++;
    for (int g = 0; g < n_gates; g++) {
        const int nin = op[g] == TFHE_AMD_MUX ? 3 : 2;
        const int32_t ins[3] = {in0[g], in1[g], in2 ? in2[g] : 0};
        for (int j = 0; j < nin; j++)
            if (!driven[ins[j]] && local[ins[j]] >= n_inputs)
                return set_error("server_submit: gate %d reads wire %d, which is neither an input nor driven", g,
                                 ins[j]);
    }
    for (int j = 0; j < n_outputs; j++)
        if (output_wires[j] < 0 || output_wires[j] >= n_wires)
            return set_error("server_submit: output wire %d out of range", output_wires[j]);
    r->client = client;
    r->W = n_wires;
    r->G = n_gates;
    r->n_in = n_inputs;
    r->n_out = n_outputs;
    r->gates.resize(5 * (size_t)n_gates);
    for (int k = 0; k < n_gates; k++) {
        const int g = order[k];
        r->gates[k] = op[g];
        r->gates[n_gates + k] = local[in0[g]];
        r->gates[2 * (size_t)n_gates + k] = local[in1[g]];
        r->gates[3 * (size_t)n_gates + k] = (op[g] == TFHE_AMD_MUX && in2) ? local[in2[g]] : local[in0[g]];
        r->gates[4 * (size_t)n_gates + k] = local[out[g]];
    }
    int32_t acc = 0;
    for (int32_t c : level_count) r->level_end.push_back(acc += c);
    r->inputs.assign(input_cts, input_cts + (size_t)n_inputs * w);
    for (int j = 0; j < n_outputs; j++) r->out_rows.push_back(local[output_wires[j]]);
    r->out_dst = output_cts;
    r->slots = (size_t)n_wires + n_outputs + ((size_t)n_outputs + w - 1) / w;
    if (r->slots > total_slots)
        return set_error("server_submit: request needs %zu ciphertext slots, the wire budget holds %zu", r->slots,
                         total_slots);
    std::lock_guard<std::mutex> g(S->mu);
    r->id = S->next_id++;
    r->t_submit = Clock::now();
    const int64_t id = r->id;
    S->incoming.push_back(r.get());
    S->all.emplace(id, std::move(r));
    S->n_submitted++;
    S->n_queued++;
    S->cv_work.notify_one();
    return id;
}

TFHE_AMD_EXPORT int32_t tfhe_amd_server_wait(tfhe_amd_server* S, int64_t request, int32_t timeout_ms,
                                             double* latency_ms) {
    if (!S) return set_error("server_wait: null server");
    std::unique_lock<std::mutex> lk(S->mu);
    auto it = S->all.find(request);
    if (it == S->all.end()) return set_error("server_wait: unknown request %lld", (long long)request);
    Request* r = it->second.get();
    auto ready = [&] { return r->done; };
    if (timeout_ms < 0) S->cv_done.wait(lk, ready);
    else if (!S->cv_done.wait_for(lk, std::chrono::milliseconds(timeout_ms), ready)) return 1;
    if (latency_ms) *latency_ms = std::chrono::duration<double, std::milli>(r->t_done - r->t_submit).count();
    const int status = r->status;
    const std::string err = r->error;
    S->all.erase(it);
    if (status) return set_error("server_wait: request failed: %s", err.c_str());
    return 0;
}

TFHE_AMD_EXPORT int32_t tfhe_amd_server_get_stats(tfhe_amd_server* S, tfhe_amd_server_stats* st) {
    if (!S || !st) return set_error("server_get_stats: bad arguments");
    std::lock_guard<std::mutex> g(S->mu);
    st->submitted = S->n_submitted;
    st->completed = S->n_completed;
    st->failed = S->n_failed;
    st->gates = S->n_gates;
    st->launches = S->n_launches;
    st->max_launch_gates = S->max_launch;
    st->queued = S->n_queued;
    st->active = S->n_active;
    st->clients = (int32_t)S->clients.size();
    st->n_streams = S->n_streams;
    st->wire_bytes_budget = S->slots.total * (size_t)std::max(S->w, 1) * sizeof(int32_t);
    st->wire_bytes_in_use = S->wire_in_use;
    st->wire_bytes_peak = S->wire_peak;
    st->key_bytes = S->key_bytes;
    return 0;
}

TFHE_AMD_EXPORT int32_t tfhe_amd_server_set_packing(tfhe_amd_server* S, int32_t min_gates) {
    if (!S || min_gates < 0) return set_error("server_set_packing: bad arguments");
    std::lock_guard<std::mutex> g(S->mu);
    S->pack_min = min_gates;
    S->cv_work.notify_all();
    return 0;
}

TFHE_AMD_EXPORT int64_t tfhe_amd_server_packed_launches(tfhe_amd_server* S) {
    if (!S) return set_error("server_packed_launches: bad arguments");
    std::lock_guard<std::mutex> g(S->mu);
    return S->n_packed;
}

TFHE_AMD_EXPORT void tfhe_amd_server_destroy(tfhe_amd_server* S) {
    if (!S) return;
    {
        std::lock_guard<std::mutex> g(S->mu);
        S->stop = true;
        S->cv_work.notify_all();
    }
    S->worker.join();  // finishes every admitted and queued request first
    {
        std::lock_guard<std::mutex> g(S->mu);
        for (auto& kv : S->all)
            if (!kv.second->done) {
                kv.second->done = true;
                kv.second->status = -1;
                kv.second->error = "server destroyed";
            }
        S->cv_done.notify_all();
    }
    for (auto s : S->streams) (void)hipStreamSynchronize(s);
    if (S->pack_stream) (void)hipStreamSynchronize(S->pack_stream);
    for (auto& cl : S->clients) {
        (void)hipHostFree(cl->h_idx);
        (void)hipFree(cl->d_idx);
        engine_key_unpin(cl->dk);
    }
    for (char* p : S->scratch) (void)hipFree(p);
    if (S->pack_stream) {
        S->pack_inflight.clear();
        (void)hipHostFree(S->h_pack);
        (void)hipFree(S->d_pack);
        (void)hipFree(S->pack_scr);
        (void)hipStreamDestroy(S->pack_stream);
    }
    for (auto s : S->streams) (void)hipStreamDestroy(s);
    (void)hipStreamDestroy(S->io);
    (void)hipFree(S->arena);
    if (S->h_bounce) (void)hipHostFree(S->h_bounce);
    delete S;
}
