// Runtime services of the backend: error strings, the pooled device allocator,
// per-stream scratch, hipGraph capture helpers.
//
// Reference counterparts: GPUAllocator / BlockManager (libs/Einsums/GPUMemory,
// libs/Einsums/BlockManager/src/BlockManager.cpp:18-90) and the per-thread streams of
// libs/Einsums/GPUStreams/src/GPUStreams.hip:55-200. Here leaves never synchronise;
// scratch is per stream so concurrent streams never share it. Freed pool blocks are reused
// only once the work that may read them is known to be finished (pool_book.hpp).
#include <map>
#include <mutex>
#include <string>
#include <unordered_map>
#include <vector>

#include "ein_common.hpp"
#include "pool_book.hpp"

namespace ein {

namespace {
thread_local std::string g_last_error;
std::mutex g_ws_mutex;
// slot 0 serves the kernels of one call; slot 1 holds a call's packed operands while a
// nested call uses slot 0
std::map<std::pair<hipStream_t, int>, WorkspaceSlot> g_ws;
// buffers superseded while a captured graph may still point at them; kept to process end
std::vector<void *> g_ws_retired;

std::mutex g_pool_mutex;
PoolBook<hipEvent_t> g_pool;

bool stream_is_capturing(hipStream_t s) {
    hipStreamCaptureStatus st = hipStreamCaptureStatusNone;
    if (hipStreamIsCapturing(s, &st) != hipSuccess) return false;
    return st == hipStreamCaptureStatusActive;
}

bool event_done(hipEvent_t e) {
    hipError_t const r = hipEventQuery(e);
    if (r == hipSuccess) return true;
    (void)hipGetLastError(); // hipErrorNotReady is not an error of the caller's next call
    return false;
}

void event_destroy(hipEvent_t e) { (void)hipEventDestroy(e); }
} // namespace

void set_error(std::string const &msg) { g_last_error = msg; }

int fail(int code, std::string const &msg) {
    set_error(msg);
    return code;
}

int workspace(hipStream_t stream, size_t bytes, void **out) { return workspace_slot(stream, 0, bytes, out); }

bool stream_capturing(hipStream_t stream) { return stream_is_capturing(stream); }

bool workspace_fits(hipStream_t stream, int slot, size_t bytes) {
    std::lock_guard<std::mutex> lock(g_ws_mutex);
    auto it = g_ws.find({stream, slot});
    return it != g_ws.end() && it->second.bytes >= bytes;
}

int workspace_slot(hipStream_t stream, int slot, size_t bytes, void **out) {
    std::lock_guard<std::mutex> lock(g_ws_mutex);
    WorkspaceSlot &w          = g_ws[{stream, slot}];
    bool const capturing      = stream_is_capturing(stream);
    WorkspaceDecision const d = workspace_decide(w, bytes);
    if (!d.grow) {
        if (capturing) w.captured = true;
        *out = w.ptr;
        return EIN_SUCCESS;
    }
    if (capturing) {
        return fail(EIN_ERROR_WORKSPACE, "workspace of " + std::to_string(bytes) +
                                             " bytes needed during graph capture; call ein_workspace_reserve first");
    }
    if (w.ptr != nullptr) {
        // queued kernels on this stream may still use the old buffer
        EIN_HIP_CHECK(hipStreamSynchronize(stream));
        if (d.retire_old) g_ws_retired.push_back(w.ptr);
        else EIN_HIP_CHECK(hipFree(w.ptr));
        w.ptr      = nullptr;
        w.bytes    = 0;
        w.captured = false;
    }
    void *p = nullptr;
    if (hipMalloc(&p, d.new_bytes) != hipSuccess) {
        (void)hipGetLastError();
        return fail(EIN_ERROR_ALLOC, "workspace hipMalloc of " + std::to_string(d.new_bytes) + " bytes failed");
    }
    w.ptr   = p;
    w.bytes = d.new_bytes;
    *out    = p;
    return EIN_SUCCESS;
}

int num_cus() {
    static int cus = 0;
    if (cus == 0) {
        int dev = 0;
        hipGetDevice(&dev);
        hipDeviceProp_t prop;
        if (hipGetDeviceProperties(&prop, dev) == hipSuccess) {
            cus = prop.multiProcessorCount;
        } else {
            cus = 256;
        }
    }
    return cus;
}

int resident_blocks(void const *kernel, int threads) {
    static std::mutex mu;
    static std::unordered_map<void const *, int> cache;
    std::lock_guard<std::mutex> lock(mu);
    auto it = cache.find(kernel);
    if (it != cache.end()) return it->second;
    int blocks = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, kernel, threads, 0) != hipSuccess || blocks < 1) blocks = 1;
    cache.emplace(kernel, blocks);
    return blocks;
}

} // namespace ein

extern "C" {

const char *ein_version(void) { return "einsums_amd 0.1.0 (gfx950)"; }

const char *ein_last_error_string(void) { return ein::g_last_error.c_str(); }

int ein_init(void) {
    int n = 0;
    EIN_HIP_CHECK(hipGetDeviceCount(&n));
    if (n < 1) return ein::fail(EIN_ERROR_HIP, "no HIP device");
    (void)ein::num_cus();
    return EIN_SUCCESS;
}

int ein_malloc(void **ptr, size_t bytes) {
    if (ptr == nullptr) return -1;
    size_t const sz = ein::pool_round_size(bytes);
    for (int attempt = 0; attempt < 2; ++attempt) {
        bool needs_sync = false;
        {
            std::lock_guard<std::mutex> lock(ein::g_pool_mutex);
            void *p = ein::g_pool.take(sz, ein::event_done, ein::event_destroy, &needs_sync);
            if (p != nullptr) {
                *ptr = p;
                return EIN_SUCCESS;
            }
        }
        // a block freed without a stream fits: make it safe by waiting for the device once
        if (!needs_sync || attempt == 1 || hipDeviceSynchronize() != hipSuccess) break;
        std::lock_guard<std::mutex> lock(ein::g_pool_mutex);
        ein::g_pool.synchronized();
    }
    void *p = nullptr;
    if (hipMalloc(&p, sz) != hipSuccess) {
        (void)hipGetLastError();
        // return every cached block to the device (after the device is idle) and retry once
        if (hipDeviceSynchronize() == hipSuccess) {
            std::lock_guard<std::mutex> lock(ein::g_pool_mutex);
            ein::g_pool.synchronized();
            ein::g_pool.drain(ein::event_done, [](void *q, auto const &b) {
                if (b.has_event) ein::event_destroy(b.event);
                (void)hipFree(q);
            });
        }
        (void)hipGetLastError();
        if (hipMalloc(&p, sz) != hipSuccess) {
            (void)hipGetLastError();
            return ein::fail(EIN_ERROR_ALLOC, "ein_malloc of " + std::to_string(bytes) + " bytes failed");
        }
    }
    std::lock_guard<std::mutex> lock(ein::g_pool_mutex);
    ein::g_pool.add_live(p, sz);
    *ptr = p;
    return EIN_SUCCESS;
}

int ein_free(void *ptr) {
    if (ptr == nullptr) return EIN_SUCCESS;
    std::lock_guard<std::mutex> lock(ein::g_pool_mutex);
    if (!ein::g_pool.release(ptr, nullptr, false))
        return ein::fail(EIN_ERROR_VALUE, "ein_free: pointer not from ein_malloc");
    return EIN_SUCCESS;
}

int ein_free_async(void *ptr, ein_stream_t stream) {
    if (ptr == nullptr) return EIN_SUCCESS;
    hipStream_t const s = ein::as_stream(stream);
    {
        std::lock_guard<std::mutex> lock(ein::g_pool_mutex);
        if (!ein::g_pool.is_live(ptr)) return ein::fail(EIN_ERROR_VALUE, "ein_free_async: pointer not from ein_malloc");
    }
    hipEvent_t ev = nullptr;
    bool const use_event =
        !ein::stream_is_capturing(s) && hipEventCreateWithFlags(&ev, hipEventDisableTiming) == hipSuccess &&
        hipEventRecord(ev, s) == hipSuccess;
    if (!use_event) {
        (void)hipGetLastError();
        if (ev != nullptr) (void)hipEventDestroy(ev);
        ev = nullptr;
    }
    std::lock_guard<std::mutex> lock(ein::g_pool_mutex);
    (void)ein::g_pool.release(ptr, ev, use_event);
    return EIN_SUCCESS;
}

int ein_workspace_reserve(ein_stream_t stream, size_t bytes) {
    void *p = nullptr;
    return ein::workspace(ein::as_stream(stream), bytes, &p);
}

size_t ein_workspace_size(ein_stream_t stream) {
    std::lock_guard<std::mutex> lock(ein::g_ws_mutex);
    auto it = ein::g_ws.find({ein::as_stream(stream), 0});
    return it == ein::g_ws.end() ? 0 : it->second.bytes;
}

int ein_stream_synchronize(ein_stream_t stream) {
    EIN_HIP_CHECK(hipStreamSynchronize(ein::as_stream(stream)));
    return EIN_SUCCESS;
}

int ein_graph_begin(ein_stream_t stream) {
    if (stream == nullptr) return ein::fail(EIN_ERROR_VALUE, "graph capture needs a non-null stream");
    EIN_HIP_CHECK(hipStreamBeginCapture(ein::as_stream(stream), hipStreamCaptureModeThreadLocal));
    return EIN_SUCCESS;
}

int ein_graph_end(ein_stream_t stream, void **graph_exec) {
    if (graph_exec == nullptr) return -2;
    hipGraph_t graph = nullptr;
    EIN_HIP_CHECK(hipStreamEndCapture(ein::as_stream(stream), &graph));
    hipGraphExec_t exec = nullptr;
    hipError_t e        = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (e != hipSuccess) return ein::fail(EIN_ERROR_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
    *graph_exec = exec;
    return EIN_SUCCESS;
}

int ein_graph_launch(void *graph_exec, ein_stream_t stream) {
    EIN_HIP_CHECK(hipGraphLaunch(reinterpret_cast<hipGraphExec_t>(graph_exec), ein::as_stream(stream)));
    return EIN_SUCCESS;
}

int ein_graph_destroy(void *graph_exec) {
    EIN_HIP_CHECK(hipGraphExecDestroy(reinterpret_cast<hipGraphExec_t>(graph_exec)));
    return EIN_SUCCESS;
}

} // extern "C"
