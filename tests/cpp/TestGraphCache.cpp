// CPU unit test of the keyed graph cache (openmm_rigidbody_plugin_amd/csrc/
// rbhip_graph_cache.h) with fake graph handles: a graph is replayed only
// while the live key equals the one it was captured under, and it is
// destroyed only after the stream it was launched on has been waited for.
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "rbhip_graph_cache.h"

using rbhip::GraphKey;

namespace {

int failures = 0;
void check(bool ok, const std::string& what) {
    if (!ok) {
        std::printf("FAILED: %s\n", what.c_str());
        failures++;
    }
}

// graph handles are integers; the log records waits and destroys in order
struct FakeOps {
    std::vector<std::string> log;
    void wait(const void* stream) { log.push_back("wait " + std::to_string((long)(size_t)stream)); }
    void destroy(int exec, int graph) { log.push_back("destroy " + std::to_string(exec) + "/" + std::to_string(graph)); }
};

GraphKey key(long stream, long tiles, int block, int kind = 3) {
    GraphKey k;
    k.stream = reinterpret_cast<const void*>((size_t)stream);
    k.systems = reinterpret_cast<const void*>((size_t)1000);
    k.partTiles = reinterpret_cast<const void*>((size_t)tiles);
    k.forceTiles = reinterpret_cast<const void*>((size_t)(tiles + 1));
    k.numPart = 10;
    k.numForce = 20;
    k.block = block;
    k.kind = kind;
    return k;
}

}  // namespace

int main() {
    rbhip::KeyedGraph<int, int> g;
    FakeOps ops;
    int captures = 0;
    auto capture = [&](int& exec, int& graph) {
        captures++;
        exec = 100 + captures;
        graph = 200 + captures;
    };

    // first use captures; the same key replays without capturing again
    int e1 = g.get(key(7, 500, 256), ops, capture);
    check(e1 == 101 && captures == 1, "first get captures");
    int e2 = g.get(key(7, 500, 256), ops, capture);
    check(e2 == 101 && captures == 1 && ops.log.empty(), "same key replays the cached graph");

    // any change of what the kernels were captured with recaptures, and the
    // old graph goes only after its own stream was waited for
    int e3 = g.get(key(7, 600, 256), ops, capture);
    check(e3 == 102 && captures == 2, "new tile table recaptures");
    check(ops.log.size() == 2 && ops.log[0] == "wait 7" && ops.log[1] == "destroy 101/201",
          "stale graph destroyed after its stream drained");
    check(g.staleDrops() == 1, "stale drop counted");

    ops.log.clear();
    int e4 = g.get(key(9, 600, 256), ops, capture);
    check(e4 == 103 && ops.log.size() == 2 && ops.log[0] == "wait 7", "new stream: old stream drained first");
    ops.log.clear();
    int e5 = g.get(key(9, 600, 512), ops, capture);
    check(e5 == 104 && ops.log[0] == "wait 9", "new block size recaptures");
    GraphKey k = key(9, 600, 512);
    k.numSites = 3;
    ops.log.clear();
    int e6 = g.get(k, ops, capture);
    check(e6 == 105 && ops.log.size() == 2, "site kernels joining the sequence recapture");
    k.layout = 1;
    int e7 = g.get(k, ops, capture);
    check(e7 == 106, "a new atom layout recaptures");

    // an explicit drop also waits first, and leaves nothing to replay
    ops.log.clear();
    g.drop(ops);
    check(!g.valid() && ops.log.size() == 2 && ops.log[0] == "wait 9" && ops.log[1] == "destroy 106/206",
          "drop waits for the stream, then destroys");
    ops.log.clear();
    g.drop(ops);
    check(ops.log.empty(), "dropping an empty cache does nothing");

    // a capture that fails leaves the cache empty, so nothing stale replays
    int before = captures;
    try {
        g.get(key(7, 500, 256), ops, [&](int&, int&) { throw 1; });
        check(false, "failing capture throws");
    } catch (int) {
    }
    check(!g.valid() && captures == before, "failed capture leaves no graph");
    int e8 = g.get(key(7, 500, 256), ops, capture);
    check(e8 == 107, "capture after a failure");

    if (failures == 0) std::printf("Done\n");
    return failures == 0 ? 0 : 1;
}
