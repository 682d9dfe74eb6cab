// Host check of the hybrid data-parallel + stream-K schedule used by the GEMM-like
// kernels: for many tile / slab / slot counts, every (tile, slab) unit is owned by exactly
// one workgroup, data-parallel workgroups own whole tiles, the partial-tile piece slots
// written by the stream-K workgroups never collide and stay inside the slots the launcher
// allocates, and the fix-up pass (one workgroup per stream-K tile) finds exactly the
// pieces of every cut tile. Mirrors the loops of gemm_dense_kernel / gemm_streamk_fixup.
// The split-K schedule (plan_split_k) is checked the same way: every unit owned once, one
// piece per workgroup in slot c * tiles + t, the fix-up's chunk loop finds exactly those.
#include <cstdio>
#include <map>
#include <set>
#include <vector>

#include "gemm_core.hpp"

using namespace ein;

static int check(uint64_t tiles, uint64_t slabs, uint64_t slots, bool split) {
    StreamK sk = split ? plan_split_k(tiles, slabs, slots) : plan_stream_k(tiles, slabs, slots);
    if (split && sk.ksplit == 1) { std::printf("split-K with one chunk\n"); return 1; }
    uint64_t const S = sk.slabs;
    std::vector<int> owner(tiles * slabs, 0);
    std::map<uint64_t, std::set<uint64_t>> pieces_of_tile; // tile -> slots written
    std::set<uint64_t> used_slots;
    if (sk.dp_tiles > sk.tiles || sk.grid < sk.dp_tiles) { std::printf("bad dp split\n"); return 1; }
    if (sk.units == 0 && sk.grid != sk.dp_tiles) { std::printf("stream-K workgroups without units\n"); return 1; }
    for (uint32_t b = 0; b < sk.grid; ++b) {
        uint64_t u0, u1, wlog;
        sk_range_of(sk, b, u0, u1, wlog);
        if (b < sk.dp_tiles && (u0 % S != 0 || u1 - u0 != S)) { std::printf("dp workgroup without a whole tile\n"); return 1; }
        for (uint64_t u = u0; u < u1;) {
            uint64_t t = u / S, s0 = u - t * S, s1 = (S - s0 < u1 - u) ? S : s0 + (u1 - u);
            for (uint64_t s = s0; s < s1; ++s) owner[t * S + s]++;
            u += s1 - s0;
            if (!(s0 == 0 && s1 == S)) {
                if (sk.units == 0 || t < sk.dp_tiles) { std::printf("partial tile outside the stream-K part\n"); return 1; }
                uint64_t slot = sk_slot(sk, wlog, t);
                if (slot >= sk_piece_slots(sk) || !used_slots.insert(slot).second) {
                    std::printf("slot %llu collides/out of range\n", (unsigned long long)slot);
                    return 1;
                }
                pieces_of_tile[t].insert(slot);
            }
        }
    }
    for (size_t i = 0; i < owner.size(); ++i)
        if (owner[i] != 1) { std::printf("unit %zu owned %d times\n", i, owner[i]); return 1; }
    if (sk.ksplit) {
        if (sk.dp_tiles != 0 || sk.grid != tiles * sk.ksplit) { std::printf("split-K grid\n"); return 1; }
        for (uint64_t t = 0; t < tiles; ++t) {
            std::set<uint64_t> found; // the fix-up's loop
            for (uint64_t c = 0; c < sk.ksplit; ++c) found.insert(c * sk.tiles + t);
            if (found != pieces_of_tile[t]) { std::printf("split-K pieces differ for tile %llu\n", (unsigned long long)t); return 1; }
        }
    } else if (sk.units) {
        for (uint64_t t = sk.dp_tiles; t < tiles; ++t) {
            uint64_t g0, g1;
            sk_owners(sk, t, g0, g1);
            std::set<uint64_t> found;
            if (g0 != g1)
                for (uint64_t g = g0; g <= g1; ++g) found.insert(sk_slot(sk, g, t));
            if (found != pieces_of_tile[t]) { std::printf("fixup pieces differ for tile %llu\n", (unsigned long long)t); return 1; }
        }
    }
    return 0;
}

int main() {
    int bad = 0, n = 0;
    uint64_t tile_counts[] = {1, 2, 3, 7, 25, 81, 100, 324, 511, 512, 513, 648, 1000, 1296, 1500, 5184, 9999, 20000};
    uint64_t slab_counts[] = {1, 2, 3, 16, 36, 144, 576, 1000};
    uint64_t slot_counts[] = {1, 2, 64, 512, 600};
    int nsplit = 0;
    for (auto t : tile_counts)
        for (auto s : slab_counts)
            for (auto g : slot_counts)
                for (bool split : {false, true}) {
                    ++n;
                    if (split && plan_split_k(t, s, g).ksplit) ++nsplit;
                    if (check(t, s, g, split)) {
                        std::printf("  failed: tiles=%llu slabs=%llu slots=%llu split=%d\n", (unsigned long long)t,
                                    (unsigned long long)s, (unsigned long long)g, int(split));
                        ++bad;
                    }
                }
    // the Wmnij case: 81 tiles of 64 x 64, 576 slabs, 512 slots -> 6 chunks of 96 slabs
    StreamK const w = plan_split_k(81, 576, 512);
    if (w.ksplit != 6 || w.kchunk != 96 || w.grid != 486) { std::printf("Wmnij split-K plan\n"); ++bad; }
    std::printf("%d schedules checked (%d split-K), %d failures\n", n, nsplit, bad);
    return bad ? 1 : 0;
}
