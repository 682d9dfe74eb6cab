// Runtime services of the backend: error strings, the pooled device allocator,
// per-stream scratch, hipGraph capture helpers.
//
// Reference counterparts: GPUAllocator / BlockManager (libs/Einsums/GPUMemory,
// libs/Einsums/BlockManager/src/BlockManager.cpp:18-90) and the per-thread streams of
// libs/Einsums/GPUStreams/src/GPUStreams.hip:55-200. Here leaves never synchronise;
// scratch is per stream so concurrent streams never share it.
#include <map>
#include <mutex>
#include <string>
#include <unordered_map>
#include <vector>

#include "ein_common.hpp"

namespace ein {

namespace {
thread_local std::string g_last_error;
std::mutex g_ws_mutex;
struct Workspace {
    void *ptr    = nullptr;
    size_t bytes = 0;
};
// slot 0 serves the kernels of one call; slot 1 holds a call's packed operands while a
// nested call uses slot 0
std::map<std::pair<hipStream_t, int>, Workspace> g_ws;

// Size-class pool: freed blocks are kept for reuse (a stream-ordered arena would need
// per-stream bookkeeping; callers of ein_malloc own their tensors for long periods, so
// a simple pool keyed on rounded size is enough and keeps hipMalloc off hot paths).
std::mutex g_pool_mutex;
std::multimap<size_t, void *> g_pool_free;
std::unordered_map<void *, size_t> g_pool_live;

size_t round_size(size_t bytes) {
    if (bytes < 256) return 256;
    // round to 1/8 of the next power of two to bound internal fragmentation
    size_t p = 1;
    while (p < bytes) p <<= 1;
    size_t step = p >= 8192 ? p / 8 : 256;
    return (bytes + step - 1) / step * step;
}

bool stream_is_capturing(hipStream_t s) {
    hipStreamCaptureStatus st = hipStreamCaptureStatusNone;
    if (hipStreamIsCapturing(s, &st) != hipSuccess) return false;
    return st == hipStreamCaptureStatusActive;
}
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
    Workspace &w = g_ws[{stream, slot}];
    if (w.bytes >= bytes) {
        *out = w.ptr;
        return EIN_SUCCESS;
    }
    if (stream_is_capturing(stream)) {
        return fail(EIN_ERROR_WORKSPACE, "workspace of " + std::to_string(bytes) +
                                             " bytes needed during graph capture; call ein_workspace_reserve first");
    }
    if (w.ptr != nullptr) {
        // the old buffer may still be in use by queued kernels on this stream
        EIN_HIP_CHECK(hipStreamSynchronize(stream));
        EIN_HIP_CHECK(hipFree(w.ptr));
        w.ptr   = nullptr;
        w.bytes = 0;
    }
    size_t want = bytes < (size_t{1} << 20) ? (size_t{1} << 20) : bytes;
    want        = (want + 4095) / 4096 * 4096;
    void *p     = nullptr;
    if (hipMalloc(&p, want) != hipSuccess) {
        (void)hipGetLastError();
        return fail(EIN_ERROR_ALLOC, "workspace hipMalloc of " + std::to_string(want) + " bytes failed");
    }
    w.ptr   = p;
    w.bytes = want;
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
    size_t sz = ein::round_size(bytes);
    {
        std::lock_guard<std::mutex> lock(ein::g_pool_mutex);
        auto it = ein::g_pool_free.lower_bound(sz);
        if (it != ein::g_pool_free.end() && it->first <= sz + sz / 4) {
            *ptr = it->second;
            ein::g_pool_live[it->second] = it->first;
            ein::g_pool_free.erase(it);
            return EIN_SUCCESS;
        }
    }
    void *p = nullptr;
    if (hipMalloc(&p, sz) != hipSuccess) {
        (void)hipGetLastError();
        // release cached blocks and retry once
        {
            std::lock_guard<std::mutex> lock(ein::g_pool_mutex);
            for (auto &kv : ein::g_pool_free) (void)hipFree(kv.second);
            ein::g_pool_free.clear();
        }
        if (hipMalloc(&p, sz) != hipSuccess) {
            (void)hipGetLastError();
            return ein::fail(EIN_ERROR_ALLOC, "ein_malloc of " + std::to_string(bytes) + " bytes failed");
        }
    }
    std::lock_guard<std::mutex> lock(ein::g_pool_mutex);
    ein::g_pool_live[p] = sz;
    *ptr                = p;
    return EIN_SUCCESS;
}

int ein_free(void *ptr) {
    if (ptr == nullptr) return EIN_SUCCESS;
    std::lock_guard<std::mutex> lock(ein::g_pool_mutex);
    auto it = ein::g_pool_live.find(ptr);
    if (it == ein::g_pool_live.end()) return ein::fail(EIN_ERROR_VALUE, "ein_free: pointer not from ein_malloc");
    ein::g_pool_free.emplace(it->second, ptr);
    ein::g_pool_live.erase(it);
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
