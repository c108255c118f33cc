// Host-side bookkeeping of captured launch sequences (no HIP dependency, so
// the logic is unit-tested on the CPU).
//
// A captured step sequence is valid only for what its kernels were captured
// with: the launch stream, the device descriptor table and the tile tables
// they read, the tile counts, the launch shape and the atom layout.  A
// KeyedGraph remembers that key with its executable graph and
//   - replays the graph only while the live key equals it (any change the
//     owner forgot to announce recaptures instead of replaying a graph that
//     reads freed or replaced tables);
//   - drops a graph only after waiting for the stream it was launched on, so
//     no executable graph is destroyed while a launch of it is in flight.
// Ops provides wait(stream) and destroy(exec, graph) for the handle types.
#pragma once

namespace rbhip {

struct GraphKey {
    const void* stream = nullptr;
    const void* systems = nullptr;
    const void* partTiles = nullptr;
    const void* forceTiles = nullptr;
    const void* clusterTiles = nullptr;
    int numPart = 0, numForce = 0, numClusterTiles = 0, numSites = 0;
    int block = 0, maxBlocks = 0, layout = 0;
    int kind = 0;   // which launch sequence
    bool operator==(const GraphKey& o) const {
        return stream == o.stream && systems == o.systems && partTiles == o.partTiles &&
               forceTiles == o.forceTiles && clusterTiles == o.clusterTiles && numPart == o.numPart &&
               numForce == o.numForce && numClusterTiles == o.numClusterTiles && numSites == o.numSites &&
               block == o.block && maxBlocks == o.maxBlocks && layout == o.layout && kind == o.kind;
    }
    bool operator!=(const GraphKey& o) const { return !(*this == o); }
};

template <class Exec, class Graph> class KeyedGraph {
public:
    bool valid() const { return exec_ != Exec{}; }
    const GraphKey& key() const { return key_; }
    // number of graphs dropped because the live key no longer matched
    long long staleDrops() const { return stale_; }

    // the graph to launch for the live key: the cached one if it was
    // captured under that key, else a new capture() (which returns the
    // executable graph and its graph; the old ones are dropped first)
    template <class Ops, class Capture> Exec get(const GraphKey& live, Ops& ops, Capture&& capture) {
        if (valid() && key_ == live) return exec_;
        if (valid()) stale_++;
        drop(ops);
        capture(exec_, graph_);
        key_ = live;
        return exec_;
    }

    // destroys the graph after its stream has drained
    template <class Ops> void drop(Ops& ops) {
        if (valid() || graph_ != Graph{}) {
            ops.wait(key_.stream);
            ops.destroy(exec_, graph_);
        }
        exec_ = Exec{};
        graph_ = Graph{};
        key_ = GraphKey{};
    }

private:
    Exec exec_{};
    Graph graph_{};
    GraphKey key_{};
    long long stale_ = 0;
};

}  // namespace rbhip
