// Level scheduler for gate netlists (the callers of the gate API that the
// reference's circuits are: test/test-addition-boot.cpp:30-111,
// test/test-long-run.cpp:61-185). A netlist is a set of bootstrapped gates
// over wires; wires are slots of a device ciphertext pool. Gates are
// levelised (as soon as possible, as late as possible, or balanced to the
// widths the GPU runs at the same cost), each level is one batch launch
// (blind rotation + keyswitch), and the whole sequence of launches is
// captured once into a hipGraph and replayed.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <mutex>
#include <queue>
#include <vector>

#include "engine.h"

struct tfhe_amd_netlist {
    const LweBootstrappingKeyFFT* bf = nullptr;
    int32_t n_wires = 0, n_gates = 0;
    std::vector<int32_t> level_count;    // gates per level
    std::vector<size_t> level_offset;    // offset (in gates) of each level in the sorted order
    int32_t* d_idx = nullptr;            // [5][n_gates]: op, in0, in1, in2, out (sorted by level)
    char* d_scr = nullptr;               // scratch for the widest level
    size_t scr_bytes = 0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    int32_t* graph_pool = nullptr;
    void* graph_stream = nullptr;
    uint64_t graph_key = 0;  // upload id of the device key the graph's kernels read
};

namespace tfhe_amd {
namespace {

// Base width (bootstraps) of the balanced schedule: schedule values of 64 and
// above give it directly; 2 takes the SIMD count of the current GPU (four per
// CU), the number of bootstraps the one-wavefront batched kernel runs at one
// wavefront per SIMD, i.e. at its shortest latency; 1024 without a GPU.
int balanced_base(int schedule) {
    if (schedule >= 64) return schedule;
    int dev = 0, cus = 0;
    if (hipGetDevice(&dev) == hipSuccess &&
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess && cus > 0)
        return 4 * cus;
    (void)hipGetLastError();
    return 1024;
}

int build_levels(tfhe_amd_netlist* nl, const int32_t* op, const int32_t* in0, const int32_t* in1,
                 const int32_t* in2, const int32_t* out, int schedule, std::vector<int32_t>& order,
                 std::vector<int32_t>& level) {
    if (schedule < 0 || (schedule > 2 && schedule < 64)) return set_error("netlist: unknown schedule %d", schedule);
    const int G = nl->n_gates, W = nl->n_wires;
    std::vector<int32_t> producer(W, -1);
    for (int g = 0; g < G; g++) {
        if (op[g] < TFHE_AMD_NAND || op[g] > TFHE_AMD_MUX) return set_error("netlist gate %d: bad gate code %d", g, op[g]);
        if (op[g] == TFHE_AMD_MUX && !in2) return set_error("netlist gate %d: MUX needs a third input (in2 is NULL)", g);
        const int nin = op[g] == TFHE_AMD_MUX ? 3 : 2;
        const int32_t ins[3] = {in0[g], in1[g], in2 ? in2[g] : 0};
        for (int j = 0; j < nin; j++)
            if (ins[j] < 0 || ins[j] >= W) return set_error("netlist gate %d: input wire %d out of range", g, ins[j]);
        if (out[g] < 0 || out[g] >= W) return set_error("netlist gate %d: output wire out of range", g);
        if (producer[out[g]] >= 0) return set_error("netlist: wire %d driven by two gates", out[g]);
        producer[out[g]] = g;
    }
    // longest path from the inputs (as soon as possible), Kahn's algorithm
    std::vector<std::vector<int32_t>> users(G);
    std::vector<int32_t> pending(G, 0);
    for (int g = 0; g < G; g++) {
        const int nin = op[g] == TFHE_AMD_MUX ? 3 : 2;
        const int32_t ins[3] = {in0[g], in1[g], in2 ? in2[g] : 0};
        for (int j = 0; j < nin; j++) {
            const int p = producer[ins[j]];
            if (p >= 0) {
                users[p].push_back(g);
                pending[g]++;
            }
        }
    }
    level.assign(G, 1);
    std::vector<int32_t> queue;
    for (int g = 0; g < G; g++)
        if (!pending[g]) queue.push_back(g);
    size_t head = 0;
    while (head < queue.size()) {
        const int g = queue[head++];
        for (int u : users[g]) {
            level[u] = std::max(level[u], level[g] + 1);
            if (--pending[u] == 0) queue.push_back(u);
        }
    }
    if ((int)queue.size() != G) return set_error("netlist has a cycle");
    int depth = 0;
    for (int g = 0; g < G; g++) depth = std::max(depth, level[g]);
    if (schedule == 1) {
        // as late as possible: each gate just before its earliest user; gates
        // nobody reads (circuit outputs) keep their earliest level
        for (int q = G - 1; q >= 0; q--) {
            const int g = queue[q];
            if (users[g].empty()) continue;
            int lv = depth;
            for (int u : users[g]) lv = std::min(lv, level[u] - 1);
            level[g] = lv;
        }
    } else if (schedule >= 2) {
        // balanced: the same depth, each level filled to the width the GPU runs at the
        // cost of the gates that must run there. A launch of up to `base` bootstraps
        // takes one bootstrap's latency (one wavefront per SIMD), up to 2 base about
        // 1.7 times that, and so on; so a level takes the gates whose deadline it is
        // (the as-late-as-possible level; gates nobody reads may wait until the last
        // level), sets its width to the smallest base * 2^k that holds them, and fills
        // it with ready gates in order of deadline. A MUX gate counts as two
        // bootstraps. The deadlines keep the depth: every gate's inputs have earlier
        // deadlines, so a gate is ready at its deadline at the latest.
        const int base = balanced_base(schedule);
        std::vector<int32_t> deadline(G);
        for (int q = G - 1; q >= 0; q--) {
            const int g = queue[q];
            int lv = depth;
            for (int u : users[g]) lv = std::min(lv, deadline[u] - 1);
            deadline[g] = lv;
        }
        auto weight = [&](int g) { return op[g] == TFHE_AMD_MUX ? 2 : 1; };
        for (int g = 0; g < G; g++) pending[g] = 0;
        for (int g = 0; g < G; g++)
            for (int u : users[g]) pending[u]++;
        // ready gates ordered by deadline (earliest first), ties by gate index
        using Item = std::pair<int32_t, int32_t>;
        std::priority_queue<Item, std::vector<Item>, std::greater<Item>> ready;
        for (int g = 0; g < G; g++)
            if (!pending[g]) ready.push({deadline[g], g});
        std::vector<int32_t> taken;
        for (int lv = 1; lv <= depth; lv++) {
            taken.clear();
            long forced = 0;
            while (!ready.empty() && ready.top().first <= lv) {
                const int g = ready.top().second;
                ready.pop();
                forced += weight(g);
                taken.push_back(g);
            }
            long width = base;
            while (width < forced) width *= 2;
            long used = forced;
            while (!ready.empty() && used + weight(ready.top().second) <= width) {
                const int g = ready.top().second;
                ready.pop();
                used += weight(g);
                taken.push_back(g);
            }
            for (int g : taken) level[g] = lv;
            // users become ready for the next level once all their inputs are placed
            for (int g : taken)
                for (int u : users[g])
                    if (--pending[u] == 0) ready.push({deadline[u], u});
        }
        if (!ready.empty()) return set_error("netlist: balanced schedule left gates unplaced");
    }
    order.resize(G);
    for (int g = 0; g < G; g++) order[g] = g;
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return level[a] < level[b]; });
    nl->level_count.assign(depth, 0);
    for (int g = 0; g < G; g++) nl->level_count[level[g] - 1]++;
    nl->level_offset.assign(depth, 0);
    for (int l = 1; l < depth; l++) nl->level_offset[l] = nl->level_offset[l - 1] + nl->level_count[l - 1];
    return 0;
}

int launch_all(tfhe_amd_netlist* nl, tfhe_amd_device_key* dk, int32_t* d_pool, void* stream, bool capturing) {
    const int G = nl->n_gates;
    for (size_t l = 0; l < nl->level_count.size(); l++) {
        const int cnt = nl->level_count[l];
        if (!cnt) continue;
        const int32_t* base = nl->d_idx + nl->level_offset[l];
        if (int rc = engine_launch_batch_locked(dk, cnt, base, base + G, base + 2 * G, base + 3 * G, base + 4 * G,
                                                d_pool, nl->d_scr, stream, capturing))
            return rc;
    }
    return 0;
}

void drop_graph(tfhe_amd_netlist* nl) {
    if (nl->exec) (void)hipGraphExecDestroy(nl->exec);
    if (nl->graph) (void)hipGraphDestroy(nl->graph);
    nl->exec = nullptr;
    nl->graph = nullptr;
    nl->graph_pool = nullptr;
    nl->graph_key = 0;
}

}  // namespace

int engine_levelize(int32_t n_wires, int32_t n_gates, const int32_t* op, const int32_t* in0, const int32_t* in1,
                    const int32_t* in2, const int32_t* out, int schedule, std::vector<int32_t>& order,
                    std::vector<int32_t>& level_count) {
    tfhe_amd_netlist tmp;
    tmp.n_wires = n_wires;
    tmp.n_gates = n_gates;
    std::vector<int32_t> level;
    if (int rc = build_levels(&tmp, op, in0, in1, in2, out, schedule, order, level)) return rc;
    level_count = tmp.level_count;
    return 0;
}

}  // namespace tfhe_amd

using namespace tfhe_amd;

TFHE_AMD_EXPORT tfhe_amd_netlist* tfhe_amd_netlist_create(const TFheGateBootstrappingCloudKeySet* ck, int32_t n_wires,
                                                          int32_t n_gates, const int32_t* op, const int32_t* in0,
                                                          const int32_t* in1, const int32_t* in2, const int32_t* out,
                                                          int32_t schedule) {
    if (!ck || n_wires <= 0 || n_gates <= 0 || !op || !in0 || !in1 || !out) {
        set_error("netlist_create: bad arguments");
        return nullptr;
    }
    auto* nl = new tfhe_amd_netlist();
    nl->bf = ck->bkFFT;
    nl->n_wires = n_wires;
    nl->n_gates = n_gates;
    std::vector<int32_t> order, level;
    if (build_levels(nl, op, in0, in1, in2, out, schedule, order, level)) {
        delete nl;
        return nullptr;
    }
    tfhe_amd_device_key* dk = engine_key(ck->bkFFT);
    if (!dk) {
        delete nl;
        return nullptr;
    }
    std::vector<int32_t> idx(5 * (size_t)n_gates);
    for (int k = 0; k < n_gates; k++) {
        const int g = order[k];
        idx[k] = op[g];
        idx[n_gates + k] = in0[g];
        idx[2 * (size_t)n_gates + k] = in1[g];
        idx[3 * (size_t)n_gates + k] = (op[g] == TFHE_AMD_MUX && in2) ? in2[g] : in0[g];
        idx[4 * (size_t)n_gates + k] = out[g];
    }
    int widest = 0;
    for (int c : nl->level_count) widest = std::max(widest, c);
    nl->scr_bytes = engine_batch_scratch_bytes(dk, widest);
    if (hipMalloc(&nl->d_idx, idx.size() * sizeof(int32_t)) != hipSuccess ||
        hipMalloc(&nl->d_scr, nl->scr_bytes) != hipSuccess ||
        hipMemcpy(nl->d_idx, idx.data(), idx.size() * sizeof(int32_t), hipMemcpyHostToDevice) != hipSuccess) {
        set_error("netlist_create: device allocation failed");
        (void)hipFree(nl->d_idx);
        (void)hipFree(nl->d_scr);
        delete nl;
        return nullptr;
    }
    return nl;
}

TFHE_AMD_EXPORT int32_t tfhe_amd_netlist_levelize(int32_t n_wires, int32_t n_gates, const int32_t* op,
                                                  const int32_t* in0, const int32_t* in1, const int32_t* in2,
                                                  const int32_t* out, int32_t schedule, int32_t* level_of_gate) {
    if (n_wires <= 0 || n_gates <= 0 || !op || !in0 || !in1 || !out) return set_error("levelize: bad arguments");
    tfhe_amd_netlist tmp;
    tmp.n_wires = n_wires;
    tmp.n_gates = n_gates;
    std::vector<int32_t> order, level;
    if (int rc = build_levels(&tmp, op, in0, in1, in2, out, schedule, order, level)) return rc;
    if (level_of_gate) memcpy(level_of_gate, level.data(), sizeof(int32_t) * n_gates);
    return (int32_t)tmp.level_count.size();
}

TFHE_AMD_EXPORT int32_t tfhe_amd_netlist_levels(const tfhe_amd_netlist* nl, int32_t* widths) {
    if (!nl) return set_error("null netlist");
    if (widths)
        for (size_t l = 0; l < nl->level_count.size(); l++) widths[l] = nl->level_count[l];
    return (int32_t)nl->level_count.size();
}

TFHE_AMD_EXPORT int32_t tfhe_amd_netlist_run(tfhe_amd_netlist* nl, int32_t* d_pool, void* stream, int32_t use_graph) {
    if (!nl || !d_pool) return set_error("netlist_run: bad arguments");
    tfhe_amd_device_key* dk = engine_key(nl->bf);
    if (!dk) return -1;
    std::lock_guard<std::mutex> g(engine_mutex());
    void* s = stream ? stream : engine_stream();
    if (!s) return -1;
    if (!use_graph) return launch_all(nl, dk, d_pool, s, false);
    // the captured kernels hold the device key's addresses: a key released and
    // uploaded again since the capture (a new upload id) needs a new graph
    if (!nl->exec || nl->graph_pool != d_pool || nl->graph_stream != s || nl->graph_key != engine_key_gen(dk)) {
        drop_graph(nl);
        hipStream_t hs = (hipStream_t)s;
        if (hipStreamBeginCapture(hs, hipStreamCaptureModeThreadLocal) != hipSuccess)
            return set_error("netlist_run: stream capture failed to start");
        const int rc = launch_all(nl, dk, d_pool, s, true);
        hipGraph_t graph = nullptr;
        const hipError_t e = hipStreamEndCapture(hs, &graph);
        if (rc) {
            if (graph) (void)hipGraphDestroy(graph);
            return rc;
        }
        if (e != hipSuccess) return set_error("netlist_run: stream capture failed: %s", hipGetErrorString(e));
        nl->graph = graph;
        if (hipGraphInstantiate(&nl->exec, graph, nullptr, nullptr, 0) != hipSuccess) {
            drop_graph(nl);
            return set_error("netlist_run: graph instantiation failed");
        }
        nl->graph_pool = d_pool;
        nl->graph_stream = s;
        nl->graph_key = engine_key_gen(dk);
    }
    if (hipGraphLaunch(nl->exec, (hipStream_t)s) != hipSuccess) return set_error("netlist_run: graph launch failed");
    return 0;
}

TFHE_AMD_EXPORT void tfhe_amd_netlist_destroy(tfhe_amd_netlist* nl) {
    if (!nl) return;
    {
        std::lock_guard<std::mutex> g(engine_mutex());
        (void)hipDeviceSynchronize();
        drop_graph(nl);
    }
    (void)hipFree(nl->d_idx);
    (void)hipFree(nl->d_scr);
    delete nl;
}
