// Step-loop graphs for one membrane.
//
// A step with Philox draws is some 15-20 launches whose arguments change
// from step to step only through the step number and the rotation of the
// particle buffers (the cell sort gathers into the spare arrays, the move
// writes the spare positions; the rotation repeats every 6 steps).  Small
// membranes spend more time issuing those launches than running them, so a
// stretch of n steps is recorded once into a HIP graph (stream capture of
// the ordinary launch sequence) and replayed:
//
//   * every launch of a recording takes its step as an offset from d_step
//     (StepRef), and the host writes the replay's first step into d_step with
//     a one-thread kernel before each replay, so a replay draws exactly the
//     numbers the same steps would draw launch by launch;
//   * a recording is keyed by the host state it starts from (buffer rotation,
//     slot_of / work-list validity), its length, and whether it images and
//     accumulates observables; a replay leaves the host state the recording
//     left (recorded at capture time);
//   * any allocation or parameter change the recorded arguments depend on
//     bumps graph_gen (graph_touch) and drops every recording.  A recording
//     is made only after an eager step of the same kind ran without such a
//     change, so first-use allocations never happen inside a capture.
//
// Steps that write per-step host-indexed rows (the census ring) run launch
// by launch; the faithful pair cap is decided on the device inside step
// loops (Cells::cap_dev), so it needs no read-back.  If a capture fails, the
// membrane runs eagerly from then on.
//
// Measured on MI355X with ROCm 7.2 (4k to 1M particles, imaging on): a
// replayed step takes 20-35 us longer than the same launches issued one by
// one on the stream (whose kernels already run back to back with no gaps),
// so recording is off unless EDM_TUNE_GRAPHS asks for it.
#include <hip/hip_runtime.h>

#include <cstring>

#include "edm_internal.h"

namespace edm {

__global__ void k_step_set(uint32_t* d_step, uint32_t v) {
    if (blockIdx.x == 0 && threadIdx.x == 0) *d_step = v;
}

static constexpr int kGraphMin = 4;      // shorter runs go launch by launch
static constexpr int kGraphMax = 256;    // longest recording, in steps
static constexpr size_t kGraphKeep = 8;  // recordings kept per membrane

static StepHost host_state(const edm_membrane* m) {
    StepHost h;
    h.a = m->a;
    h.slot_of_valid = m->slot_of_valid;
    h.wl_valid = m->wl_valid;
    h.wl_built = m->wl_built;
    return h;
}

static void set_host_state(edm_membrane* m, const StepHost& h) {
    m->a = h.a;
    m->slot_of_valid = h.slot_of_valid;
    m->wl_valid = h.wl_valid;
    m->wl_built = h.wl_built;
}

static bool same_host(const StepHost& u, const StepHost& v) {
    return std::memcmp(&u.a, &v.a, sizeof(DevArrays)) == 0 &&
           u.slot_of_valid == v.slot_of_valid && u.wl_valid == v.wl_valid &&
           u.wl_built == v.wl_built;
}

static int eager_step(edm_membrane* m, bool image) {
    m->need_pair_offsets = loop_pair_offsets(m);
    m->pair_sync = false;
    if (int rc = dyn_step_begin(m)) return rc;
    if (int rc = dyn_step_finish(m, nullptr, false)) return rc;
    if (image)
        if (int rc = ph_fused_image(m, nullptr, nullptr)) return rc;
    if (m->rec_flags & EDM_REC_AUTO)
        if (int rc = rec_step(m)) return rc;
    return 0;
}

static void graph_drop(StepGraph& g) {
    if (g.exec) (void)hipGraphExecDestroy(g.exec);
    g.exec = nullptr;
}

void graphs_free(edm_membrane* m) {
    for (StepGraph& g : m->graphs) graph_drop(g);
    m->graphs.clear();
}

void graph_touch(edm_membrane* m) {
    m->graph_gen += 1;
    if (!m->graphs.empty()) {
        // a replay in flight still holds its executable
        if (m->stream) (void)hipStreamSynchronize(m->stream);
        graphs_free(m);
    }
}

bool graphs_eligible(const edm_membrane* m) {
    return m->use_graphs && !m->graph_failed && !(m->rec_flags & EDM_REC_AUTO) && !m->begun &&
           !m->time_force;
}

static int replay(edm_membrane* m, const StepGraph& g) {
    hipLaunchKernelGGL(k_step_set, dim3(1), dim3(64), 0, m->stream, m->d_step, m->step);
    EDM_HIP(hipGetLastError());
    EDM_HIP(hipGraphLaunch(g.exec, m->stream));
    set_host_state(m, g.h1);
    m->step += (uint32_t)g.n;
    m->graph_replays += 1;
    return 0;
}

// Records n steps starting from the current host state and replays them
// once.  Returns 0 with *done = false when the capture failed (the host
// state is restored and nothing ran); the caller then steps eagerly.
static int record(edm_membrane* m, int n, bool image, bool* done) {
    *done = false;
    if (!m->d_step) {
        EDM_HIP(hipMalloc((void**)&m->d_step, 64));
        graph_touch(m);
        return 0;  // the next call records (after an eager step of this generation)
    }
    const StepHost h0 = host_state(m);
    const uint32_t step0 = m->step;
    const uint64_t gen0 = m->graph_gen;
    const hipError_t eb = hipStreamBeginCapture(m->stream, hipStreamCaptureModeThreadLocal);
    if (eb != hipSuccess) {
        (void)hipGetLastError();
        m->graph_failed = true;
        m->graph_status = 1000 + (int)eb;
        return 0;
    }
    m->graph_rec = true;
    m->graph_step0 = step0;
    int rc = 0;
    for (int k = 0; k < n && rc == 0; ++k) rc = eager_step(m, image);
    m->graph_rec = false;
    hipGraph_t graph = nullptr;
    const hipError_t ec = hipStreamEndCapture(m->stream, &graph);
    const StepHost h1 = host_state(m);
    set_host_state(m, h0);
    m->step = step0;
    hipGraphExec_t exec = nullptr;
    // why a recording failed, kept for edm_get_tuning(EDM_TUNE_GRAPH_STATUS)
    int status = 0;
    if (rc) status = 2000 + rc;
    else if (ec != hipSuccess || !graph) status = 3000 + (int)ec;
    else if (m->graph_gen != gen0) status = 5000;
    if (status == 0) {
        const hipError_t ei = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
        if (ei != hipSuccess) status = 4000 + (int)ei;
    }
    if (graph) (void)hipGraphDestroy(graph);
    if (status) {
        (void)hipGetLastError();
        if (exec) (void)hipGraphExecDestroy(exec);
        m->graph_failed = true;
        m->graph_status = status;
        return 0;
    }
    if (m->graphs.size() >= kGraphKeep) {
        EDM_HIP(hipStreamSynchronize(m->stream));
        graph_drop(m->graphs.front());
        m->graphs.erase(m->graphs.begin());
    }
    StepGraph g;
    g.h0 = h0;
    g.h1 = h1;
    g.n = n;
    g.image = image;
    g.stats = m->st_acc != nullptr;
    g.exec = exec;
    m->graphs.push_back(g);
    if (int rc2 = replay(m, m->graphs.back())) return rc2;
    *done = true;
    return 0;
}

int steps_run(edm_membrane* m, int n, bool image) {
    const int slot = image ? 1 : 0;
    int done = 0;
    while (done < n) {
        const int left = n - done;
        if (!graphs_eligible(m) || left < kGraphMin || m->warm_gen[slot] != m->graph_gen) {
            const uint64_t gen = m->graph_gen;
            if (int rc = eager_step(m, image)) return rc;
            // a step that allocated nothing: the next stretch may be recorded
            if (m->graph_gen == gen) m->warm_gen[slot] = gen;
            done += 1;
            continue;
        }
        const int len = left < kGraphMax ? left : kGraphMax;
        const StepHost h = host_state(m);
        const bool stats = m->st_acc != nullptr;
        const StepGraph* hit = nullptr;
        for (const StepGraph& g : m->graphs)
            if (g.n == len && g.image == image && g.stats == stats && same_host(g.h0, h)) hit = &g;
        if (hit) {
            if (int rc = replay(m, *hit)) return rc;
            done += len;
            continue;
        }
        bool ran = false;
        if (int rc = record(m, len, image, &ran)) return rc;
        if (ran) done += len;
    }
    return 0;
}

}  // namespace edm
