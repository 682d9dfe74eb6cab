// Bookkeeping of the pooled device allocator and of the per-stream workspaces, kept free of
// device calls so that the host tests can drive it with simulated events.
//
// The reference allocates per DeviceTensor with hipMalloc and frees with hipFree, which
// waits for the device (GPUAllocator.hpp:53-180, BlockManager.cpp:66). Here freed blocks go
// back to a size-class pool without waiting, so each freed block carries the condition under
// which it may be handed out again:
//   * freed with a stream (ein_free_async): an event recorded on that stream at the free;
//     the block is reusable once the event has completed;
//   * freed without a stream (ein_free): the work that may still read it is unknown, so the
//     block is reusable only after the device has been synchronised once after the free
//     (a synchronisation "epoch" that the pool advances).
#pragma once

#include <cstddef>
#include <cstdint>
#include <map>
#include <unordered_map>
#include <utility>
#include <vector>

namespace ein {

// Size classes: 256 B minimum, otherwise rounded up to 1/8 of the next power of two (internal
// fragmentation below 12.5 %).
inline size_t pool_round_size(size_t bytes) {
    if (bytes < 256) return 256;
    size_t p = 1;
    while (p < bytes) p <<= 1;
    size_t const step = p >= 8192 ? p / 8 : 256;
    return (bytes + step - 1) / step * step;
}

template <class Event>
class PoolBook {
public:
    struct Block {
        void *ptr;
        size_t size;
        Event event;
        bool has_event;
        uint64_t epoch; // pool epoch at the free (blocks without an event)
    };

    void add_live(void *p, size_t size) { live_[p] = size; }
    bool is_live(void *p) const { return live_.count(p) != 0; }
    size_t live_count() const { return live_.size(); }
    size_t free_count() const { return free_.size(); }
    uint64_t epoch() const { return epoch_; }

    // Moves a live block to the free list with its reuse condition; false if the pool does not
    // own `p` (nothing changes then).
    bool release(void *p, Event ev, bool has_event) {
        auto it = live_.find(p);
        if (it == live_.end()) return false;
        free_.emplace(it->second, Block{p, it->second, ev, has_event, epoch_});
        live_.erase(it);
        return true;
    }

    // A free block of class `size` (at most 25 % larger) whose condition holds, now live; or
    // nullptr. `done(ev)` tells whether an event has completed; the event of a taken block is
    // passed to `retire(ev)` (to be destroyed). *needs_sync is set when a block of the right
    // size waits only for a device synchronisation.
    template <class Done, class Retire>
    void *take(size_t size, Done &&done, Retire &&retire, bool *needs_sync) {
        if (needs_sync) *needs_sync = false;
        for (auto it = free_.lower_bound(size); it != free_.end() && it->first <= size + size / 4; ++it) {
            Block const &b = it->second;
            bool ok;
            if (b.has_event) {
                ok = done(b.event);
            } else {
                ok = b.epoch < epoch_;
                if (!ok && needs_sync) *needs_sync = true;
            }
            if (!ok) continue;
            void *p = b.ptr;
            if (b.has_event) retire(b.event);
            live_[p] = it->first;
            free_.erase(it);
            return p;
        }
        return nullptr;
    }

    // The device has been synchronised: every block freed before this call is reusable.
    void synchronized() { ++epoch_; }

    // Removes every free block whose condition holds and hands it to `release_fn(ptr, block)`
    // (which returns the memory to the device); the others stay.
    template <class Done, class Fn>
    size_t drain(Done &&done, Fn &&release_fn) {
        size_t n = 0;
        for (auto it = free_.begin(); it != free_.end();) {
            Block const &b = it->second;
            bool const ok  = b.has_event ? done(b.event) : b.epoch < epoch_;
            if (ok) {
                release_fn(b.ptr, b);
                it = free_.erase(it);
                ++n;
            } else {
                ++it;
            }
        }
        return n;
    }

private:
    std::multimap<size_t, Block> free_;
    std::unordered_map<void *, size_t> live_;
    uint64_t epoch_ = 0;
};

// Per-stream scratch buffer bookkeeping. A buffer handed out while its stream was being
// captured into a graph may be baked into that graph, so when the slot grows it is retired
// (kept until the process ends) instead of being freed; other superseded buffers are freed
// once the stream has drained.
struct WorkspaceSlot {
    void *ptr      = nullptr;
    size_t bytes   = 0;
    bool captured  = false; // handed out during a stream capture
};

// Decision for a request of `bytes` on a slot: use the current buffer, or grow it (then
// `free_old` / `retire_old` say what becomes of the old one). Growth is at least 2x so that a
// sequence of growing requests retires at most as much memory as the final buffer holds.
struct WorkspaceDecision {
    bool grow       = false;
    bool free_old   = false;
    bool retire_old = false;
    size_t new_bytes = 0;
};

inline WorkspaceDecision workspace_decide(WorkspaceSlot const &w, size_t bytes) {
    WorkspaceDecision d;
    if (w.bytes >= bytes && w.ptr != nullptr) return d;
    d.grow = true;
    size_t want = bytes < (size_t{1} << 20) ? (size_t{1} << 20) : bytes;
    if (w.ptr != nullptr) {
        if (want < 2 * w.bytes) want = 2 * w.bytes;
        d.retire_old = w.captured;
        d.free_old   = !w.captured;
    }
    d.new_bytes = (want + 4095) / 4096 * 4096;
    return d;
}

} // namespace ein
