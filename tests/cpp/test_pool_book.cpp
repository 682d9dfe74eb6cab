// Host check of the device pool's and workspaces' bookkeeping (pool_book.hpp) with simulated
// events: a freed block is never handed out while the work that may read it can still run.
#include <cstdio>
#include <set>
#include <vector>

#include "pool_book.hpp"

using namespace ein;

static int failures = 0;
#define CHECK(cond)                                                                                                    \
    do {                                                                                                               \
        if (!(cond)) {                                                                                                 \
            std::printf("FAIL line %d: %s\n", __LINE__, #cond);                                                        \
            ++failures;                                                                                                \
        }                                                                                                              \
    } while (0)

int main() {
    std::set<int> done;  // completed simulated events
    std::vector<int> retired;
    auto is_done = [&](int ev) { return done.count(ev) != 0; };
    auto retire  = [&](int ev) { retired.push_back(ev); };
    char mem[16];
    void *a = mem, *b = mem + 1, *c = mem + 2, *d = mem + 3;

    CHECK(pool_round_size(1) == 256 && pool_round_size(256) == 256 && pool_round_size(257) == 512);
    CHECK(pool_round_size(8193) == 16384 / 8 * 5);  // 1/8 of 16384 steps
    CHECK(pool_round_size(1 << 20) == (1 << 20));

    PoolBook<int> pool;
    pool.add_live(a, 4096);
    pool.add_live(b, 4096);
    pool.add_live(c, 4096);
    pool.add_live(d, 1 << 20);
    CHECK(!pool.release(mem + 9, 0, false));  // not owned: nothing changes
    CHECK(pool.live_count() == 4);

    // freed on a stream: reusable only once the event completed
    CHECK(pool.release(a, 7, true));
    bool sync = true;
    CHECK(pool.take(4096, is_done, retire, &sync) == nullptr && !sync);
    done.insert(7);
    CHECK(pool.take(4096, is_done, retire, &sync) == a && retired.size() == 1 && retired[0] == 7);
    CHECK(pool.take(4096, is_done, retire, &sync) == nullptr);  // a is live again

    // freed without a stream: reusable only after a device synchronisation
    CHECK(pool.release(b, 0, false));
    CHECK(pool.take(4096, is_done, retire, &sync) == nullptr && sync);
    pool.synchronized();
    CHECK(pool.take(4096, is_done, retire, &sync) == b);
    // a block freed after the synchronisation waits for the next one
    CHECK(pool.release(c, 0, false));
    CHECK(pool.take(4096, is_done, retire, &sync) == nullptr && sync);
    pool.synchronized();
    CHECK(pool.take(4096, is_done, retire, &sync) == c);

    // size classes: a 1 MiB block does not serve a 4 KiB request, but serves 1 MiB - 100 KiB
    CHECK(pool.release(d, 9, true));
    done.insert(9);
    CHECK(pool.take(4096, is_done, retire, &sync) == nullptr);
    CHECK(pool.take((1 << 20) - 100 * 1024, is_done, retire, &sync) == d);

    // drain: only the blocks whose condition holds go back to the device
    CHECK(pool.release(a, 11, true));  // event pending
    CHECK(pool.release(b, 0, false));  // needs a sync
    CHECK(pool.release(c, 12, true));
    done.insert(12);
    std::vector<void *> drained;
    CHECK(pool.drain(is_done, [&](void *p, auto const &) { drained.push_back(p); }) == 1 && drained[0] == c);
    pool.synchronized();
    done.insert(11);
    CHECK(pool.drain(is_done, [&](void *p, auto const &) { drained.push_back(p); }) == 2 && pool.free_count() == 0);

    // workspace growth
    WorkspaceSlot w;
    WorkspaceDecision g = workspace_decide(w, 10);
    CHECK(g.grow && !g.free_old && !g.retire_old && g.new_bytes == (size_t{1} << 20));
    w.ptr   = a;
    w.bytes = g.new_bytes;
    CHECK(!workspace_decide(w, 1000).grow && !workspace_decide(w, w.bytes).grow);
    g = workspace_decide(w, w.bytes + 1);  // at least doubles, frees the old buffer
    CHECK(g.grow && g.free_old && !g.retire_old && g.new_bytes == 2 * w.bytes);
    w.captured = true;  // handed out during a capture: retired, never freed
    g          = workspace_decide(w, 5 * w.bytes);
    CHECK(g.grow && !g.free_old && g.retire_old && g.new_bytes == 5 * w.bytes);

    std::printf("test_pool_book: %d failures\n", failures);
    return failures == 0 ? 0 : 1;
}
