// Host side of the rbhip C ABI (include/rbhip.h): contexts, batches, tiling,
// graph capture, timing and the tile autotuner.
//
// Integer bookkeeping follows the reference bit for bit:
//   cleanBodyIndices            openmmapi/src/RigidBodySystem.cpp:28-49
//   RigidBodySystem::initialize openmmapi/src/RigidBodySystem.cpp:55-114
// and is what rbhip_get_body_index / rbhip_get_atom_index return.  The
// device works in its own internal order: free atoms first, then the bodies
// grouped by atom count (ascending), each body's atoms contiguous, bodies of
// one size in reference order.  Tiles of one size class never mix sizes, so
// every lane of a wave runs the same atom loop.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <functional>
#include <map>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "rbhip.h"
#include "rbhip_bookkeeping.h"
#include "rbhip_graph_cache.h"
#include "rbhip_shapes.h"
#include "rbhip_internal.h"
#include "rbhip_kernels.h"

using namespace rbhip;

namespace {

thread_local std::string g_err;

struct RbError : std::runtime_error {
    int code;
    RbError(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

#define HIPCHK(expr)                                                                                       \
    do {                                                                                                   \
        hipError_t e_ = (expr);                                                                            \
        if (e_ != hipSuccess) throw RbError(RBHIP_E_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

// (messages are built only on failure: the checks run in per-atom loops)
void require(bool ok, const char* msg, int code = RBHIP_E_ARG) {
    if (!ok) throw RbError(code, msg);
}
void require(bool ok, const std::string& msg, int code = RBHIP_E_ARG) {
    if (!ok) throw RbError(code, msg);
}

template <class F> int guard(F&& f) {
    try {
        f();
        return RBHIP_OK;
    } catch (const RbError& e) {
        g_err = e.what();
        return e.code;
    } catch (const std::bad_alloc&) {
        g_err = "host allocation failed";
        return RBHIP_E_NOMEM;
    } catch (const std::exception& e) {
        g_err = e.what();
        return RBHIP_E_ARG;
    }
}

// cleanBodyIndices (RigidBodySystem.cpp:28-49): distinct positive labels are
// renumbered 1..B in ascending label order, labels <= 0 become 0
std::vector<int> clean_body_indices(const int* in, int n) {
    // (label, atom) pairs of the body atoms sorted by label with a stable
    // 4 x 8-bit radix sort, then ranked in one pass: O(n), no per-atom search
    std::vector<unsigned> key, atom;
    key.reserve(n);
    atom.reserve(n);
    for (int i = 0; i < n; i++)
        if (in[i] > 0) {
            key.push_back((unsigned)in[i]);
            atom.push_back((unsigned)i);
        }
    const size_t m = key.size();
    std::vector<unsigned> key2(m), atom2(m);
    for (int shift = 0; shift < 32 && m > 0; shift += 8) {
        size_t count[257] = {0};
        for (size_t i = 0; i < m; i++) count[((key[i] >> shift) & 255u) + 1]++;
        if (count[((key[0] >> shift) & 255u) + 1] == m) continue;  // one bucket: the pass keeps the order
        for (int b = 0; b < 256; b++) count[b + 1] += count[b];
        for (size_t i = 0; i < m; i++) {
            size_t d = count[(key[i] >> shift) & 255u]++;
            key2[d] = key[i];
            atom2[d] = atom[i];
        }
        key.swap(key2);
        atom.swap(atom2);
    }
    std::vector<int> out(n, 0);
    int rank = 0;
    for (size_t i = 0; i < m; i++) {
        if (i == 0 || key[i] != key[i - 1]) rank++;
        out[atom[i]] = rank;
    }
    return out;
}

// Device memory of one owner.  With a stream set, blocks are allocated and
// freed in that stream's order (hipMallocAsync / hipFreeAsync from the
// device's memory pool, which keeps freed memory for reuse), so creating or
// destroying a context never waits for work queued on other streams; the
// owner's device work runs on that stream or is ordered after it by events.
// Without a stream the blocks use hipMalloc / hipFree.
struct DeviceArena {
    std::vector<void*> blocks;
    size_t bytes = 0;
    hipStream_t stream = nullptr;
    void* alloc(size_t n) {
        if (n == 0) n = 256;
        void* p = nullptr;
        hipError_t e = stream ? hipMallocAsync(&p, n, stream) : hipMalloc(&p, n);
        if (e != hipSuccess) throw RbError(RBHIP_E_NOMEM, std::string("device allocation failed: ") + hipGetErrorString(e));
        blocks.push_back(p);
        bytes += n;
        return p;
    }
    template <class T> T* allocT(size_t count) { return reinterpret_cast<T*>(alloc(count * sizeof(T))); }
    void release() {
        for (void* p : blocks) (void)(stream ? hipFreeAsync(p, stream) : hipFree(p));
        blocks.clear();
        bytes = 0;
    }
};

// Streams come from a per-device pool created once per process (creating a
// HIP stream costs milliseconds, which dominated the set-up of small
// systems); a context or batch takes the next pool stream in turn.  Pool
// streams are non-blocking and live until the process ends.
hipStream_t pooled_stream(int device) {
    static std::mutex mtx;
    static std::map<int, std::vector<hipStream_t>> pool;
    static std::map<int, unsigned> next;
    constexpr int kPoolSize = 4;
    std::lock_guard<std::mutex> lock(mtx);
    std::vector<hipStream_t>& v = pool[device];
    if (v.empty()) {
        for (int i = 0; i < kPoolSize; i++) {
            hipStream_t st = nullptr;
            HIPCHK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
            v.push_back(st);
        }
        // the stream-ordered allocations keep freed memory in the device's
        // pool for the next context instead of returning it.  Memory freed on
        // one stream is reused on another only once the free has completed:
        // the pool inserts no waits between streams, so setting up a context
        // never waits behind the steps queued before another context's free.
        hipMemPool_t mp = nullptr;
        if (hipDeviceGetDefaultMemPool(&mp, device) == hipSuccess && mp) {
            uint64_t keep = UINT64_MAX;
            (void)hipMemPoolSetAttribute(mp, hipMemPoolAttrReleaseThreshold, &keep);
            int no = 0;
            (void)hipMemPoolSetAttribute(mp, hipMemPoolReuseAllowInternalDependencies, &no);
        }
    }
    return v[next[device]++ % v.size()];
}

int pow2_at_least(int v) {
    int p = 1;
    while (p < v) p <<= 1;
    return p;
}

// HIP handles for KeyedGraph: a graph is destroyed only after the stream it
// was launched on has drained
struct HipGraphOps {
    void wait(const void* stream) { (void)hipStreamSynchronize(static_cast<hipStream_t>(const_cast<void*>(stream))); }
    void destroy(hipGraphExec_t exec, hipGraph_t graph) {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
    }
};
using GraphCache = KeyedGraph<hipGraphExec_t, hipGraph_t>;

// the graph of the launches `launches` makes on key.stream, captured once
// per key
template <class F> hipGraphExec_t ensure_graph(GraphCache& g, const GraphKey& key, F&& launches) {
    HipGraphOps ops;
    hipStream_t stream = static_cast<hipStream_t>(const_cast<void*>(key.stream));
    return g.get(key, ops, [&](hipGraphExec_t& exec, hipGraph_t& graph) {
        HIPCHK(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
        try {
            launches();
        } catch (...) {
            hipGraph_t tmp = nullptr;
            (void)hipStreamEndCapture(stream, &tmp);
            if (tmp) (void)hipGraphDestroy(tmp);
            throw;
        }
        HIPCHK(hipStreamEndCapture(stream, &graph));
        hipError_t e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
        if (e != hipSuccess) {
            (void)hipGraphDestroy(graph);
            graph = nullptr;
            exec = nullptr;
            throw RbError(RBHIP_E_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
        }
    });
}

// The repeated launch sequences of a call, as HIP graphs: the two-pass step
// (part1, forces, part2), one fused step boundary (fused pass, forces), and
// kFusedChunk of those in one graph; the *Last graphs are the same steps with
// the potential energy, which the last step of a call takes.  A graph launch
// costs the host tens of microseconds, as much as a small system's whole
// step, so long calls replay the chunk.
struct StepGraphs {
    static constexpr int kFusedChunk = 16;
    enum Kind { TWO_PASS = 1, FUSED = 2, CHUNK = 3, TWO_PASS_LAST = 4, FUSED_LAST = 5 };
    GraphCache twoPass, fused, chunk, twoPassLast, fusedLast;
    // drops every graph (each after its stream has drained)
    void reset() {
        HipGraphOps ops;
        twoPass.drop(ops);
        fused.drop(ops);
        chunk.drop(ops);
        twoPassLast.drop(ops);
        fusedLast.drop(ops);
    }
    bool any() const {
        return twoPass.valid() || fused.valid() || chunk.valid() || twoPassLast.valid() || fusedLast.valid();
    }
};

// RigidBodyIntegrator::step(steps) on a context or a batch (`o` provides the
// launches, its stream, graphs and switches).  With fusion the steps'
// interior boundaries run as one fused pass each: part1, forces,
// (steps - 1) x [fused, forces], part2.  With `refine` the last step is
// bracketed by the refined-energy kernels.  Only the last force evaluation
// takes the potential energy (a step asks for forces alone); the graphs hold
// the other steps, and the last one is launched on its own.
template <class O> void run_steps(O& o, int steps, bool refine) {
    if (steps <= 0) return;
    const int tail = refine ? 1 : 0;
    if (o.useFusion && steps >= 2) {
        int interior = steps - 1 - tail;
        o.launchPart1();
        o.launchForces(interior == 0 && !tail);
        if (interior > 0) interior--;                 // the last one below, with the energy
        if (o.useGraph) {
            if (interior >= StepGraphs::kFusedChunk) {
                hipGraphExec_t g = ensure_graph(o.graphs.chunk, o.graphKey(StepGraphs::CHUNK), [&] {
                    for (int i = 0; i < StepGraphs::kFusedChunk; i++) {
                        o.launchFused();
                        o.launchForces(false);
                    }
                });
                for (; interior >= StepGraphs::kFusedChunk; interior -= StepGraphs::kFusedChunk)
                    HIPCHK(hipGraphLaunch(g, o.stream));
            }
            if (interior > 0) {
                hipGraphExec_t g = ensure_graph(o.graphs.fused, o.graphKey(StepGraphs::FUSED), [&] {
                    o.launchFused();
                    o.launchForces(false);
                });
                for (; interior > 0; interior--) HIPCHK(hipGraphLaunch(g, o.stream));
            }
        } else {
            for (; interior > 0; interior--) {
                o.launchFused();
                o.launchForces(false);
            }
        }
        if (steps - 1 - tail > 0) {
            if (o.useGraph && !tail) {
                HIPCHK(hipGraphLaunch(ensure_graph(o.graphs.fusedLast, o.graphKey(StepGraphs::FUSED_LAST), [&] {
                                          o.launchFused();
                                          o.launchForces(true);
                                      }),
                                      o.stream));
            } else {
                o.launchFused();
                o.launchForces(!tail);
            }
        }
        if (tail) {
            o.launchRefineBegin(false);
            o.launchFused();
            o.launchForces(true);
        }
        o.launchPart2();
    } else {
        const int plain = steps - tail;
        if (o.useGraph && plain > 1) {
            hipGraphExec_t g = ensure_graph(o.graphs.twoPass, o.graphKey(StepGraphs::TWO_PASS), [&] {
                o.launchPart1();
                o.launchForces(false);
                o.launchPart2();
            });
            for (int s = 0; s < plain - 1; s++) HIPCHK(hipGraphLaunch(g, o.stream));
        } else {
            for (int s = 0; s < plain - 1; s++) {
                o.launchPart1();
                o.launchForces(false);
                o.launchPart2();
            }
        }
        if (plain > 0) {
            if (o.useGraph && !tail) {
                HIPCHK(hipGraphLaunch(ensure_graph(o.graphs.twoPassLast, o.graphKey(StepGraphs::TWO_PASS_LAST), [&] {
                                          o.launchPart1();
                                          o.launchForces(true);
                                          o.launchPart2();
                                      }),
                                      o.stream));
            } else {
                o.launchPart1();
                o.launchForces(!tail);
                o.launchPart2();
            }
        }
        if (tail) {
            o.launchRefineBegin(true);
            o.launchPart1();
            o.launchForces(true);
            o.launchPart2();
        }
    }
    if (tail) o.launchRefineEnd();
}

}  // namespace

namespace {
struct EventSet {
    std::vector<hipEvent_t> ev;
    explicit EventSet(size_t n) : ev(n, nullptr) {
        for (auto& e : ev) HIPCHK(hipEventCreate(&e));
    }
    ~EventSet() {
        for (auto& e : ev)
            if (e) (void)hipEventDestroy(e);
    }
};

// times `steps` repetitions of the phases (each a launch sequence) on
// `stream`: avgMs[i] = the mean duration of the step kernel of phase i (its
// first k_part1 / k_part2 / k_fused / k_forces launch), from events attached
// to that kernel's dispatch, so launch overheads and gaps are not counted;
// totalMs = the whole run, gaps and any other kernels included
void time_phases(hipStream_t stream, int steps, const std::vector<std::function<void()>>& phases, double* avgMs,
                 double* totalMs) {
    const size_t np = phases.size();
    EventSet starts(np * (size_t)steps), stops(np * (size_t)steps), span(2);
    HIPCHK(hipEventRecord(span.ev[0], stream));
    for (int s = 0; s < steps; s++)
        for (size_t k = 0; k < np; k++) {
            const size_t e = np * s + k;
            set_step_kernel_events(starts.ev[e], stops.ev[e]);
            try {
                phases[k]();
            } catch (...) {
                (void)clear_step_kernel_events();
                throw;
            }
            if (clear_step_kernel_events()) {   // no step kernel in this phase
                HIPCHK(hipEventRecord(starts.ev[e], stream));
                HIPCHK(hipEventRecord(stops.ev[e], stream));
            }
        }
    HIPCHK(hipEventRecord(span.ev[1], stream));
    HIPCHK(hipEventSynchronize(span.ev[1]));
    std::vector<double> acc(np, 0.0);
    for (int s = 0; s < steps; s++)
        for (size_t k = 0; k < np; k++) {
            float ms = 0.f;
            HIPCHK(hipEventElapsedTime(&ms, starts.ev[np * s + k], stops.ev[np * s + k]));
            acc[k] += ms;
        }
    for (size_t k = 0; k < np; k++) avgMs[k] = acc[k] / steps;
    if (totalMs) {
        float ms = 0.f;
        HIPCHK(hipEventElapsedTime(&ms, span.ev[0], span.ev[1]));
        *totalMs = ms;
    }
}

template <class P1, class F, class P2>
void time_steps(hipStream_t stream, int steps, P1 p1, F f, P2 p2, double* avgMs, double* totalMs) {
    time_phases(stream, steps, {p1, f, p2}, avgMs, totalMs);
}
}  // namespace

// device bytes rbhip_create allocates for a system (the arena; tiles and
// bonds come on top), kept in step with rbhip_create by a check there
static size_t context_arena_bytes(size_t na, int numBodies, int numFree, int numBodyAtoms, int numClasses, size_t R,
                                  int clusterSlots = 0) {
    auto a = [](size_t n) { return n == 0 ? (size_t)256 : n; };
    const size_t nb = std::max(numBodies, 1);
    const size_t bodyStride = (nb + 63) / 64 * 64, atomStride = (na + 63) / 64 * 64;
    const size_t dStride = (std::max<size_t>(numBodyAtoms, 1) + 63) / 64 * 64;
    const bool mixed = R == 4;
    const size_t maxKe = ((size_t)numFree + 63) / 64 + 1 + numBodies + numClasses + 1 + clusterSlots;
    const size_t maxPe = (na + 63) / 64 + 1;
    size_t b = 0;
    b += a(4 * R * nb) + a(3 * R * nb) + a(3 * R * nb) + a(4 * R * nb);           // q, L, pcm, iI
    b += a(3 * sizeof(double) * bodyStride) + a(3 * R * dStride);                  // rcm, d
    b += a(3 * R * atomStride) + (mixed ? a(3 * sizeof(float) * atomStride) : 0);  // pos, posLo
    b += a(3 * R * atomStride) + a(3 * R * atomStride);                            // vel, force
    b += a(R * std::max(numFree, 1)) + a(sizeof(double) * na);                    // freeInvM, mass
    b += a(3 * sizeof(double) * na) * 2;                                           // stageR, stageV
    b += a(R * na) + a(3 * R * atomStride);                                       // tether k, r0
    b += a(4 * sizeof(double) * maxKe) + a(4 * sizeof(double) * maxPe);            // partials
    b += a(sizeof(DevSys)) + a(4 * sizeof(int)) + a(sizeof(int) * nb) * 2;
    b += a(sizeof(double) * std::max(3 * na, 36 * nb)) + a(4 * sizeof(double));  // staging, reduction
    b += a(sizeof(int) * na) * 3;                      // body index, atom index, internal order
    const size_t nc = std::max(numClasses, 1);
    b += a(sizeof(int) * nb) + a(sizeof(int) * 8 * nc) + a(sizeof(int) * nc);   // dOff, shape rows, matches
    return b;
}

// upper bound of the tile tables (64-wide workgroups, one lane per body)
static size_t tile_bytes_bound(size_t na, int numBodies, int numFree, int numClasses) {
    size_t part = ((size_t)numFree + 63) / 64 + ((size_t)numBodies + 63) / 64 + numClasses + 1;
    size_t force = (na + 63) / 64 + 1;
    return (2 * part + force) * sizeof(Tile);
}

// ---------------------------------------------------------------------------
struct rbhip_context {
    int device = 0;
    int precision = RBHIP_PRECISION_MIXED;
    int rotationMode = 10;
    double dt = 0.0;
    double kickDt = 0.0;      // step size of the half-kick folded into the body state
    hipStream_t ownStream = nullptr;
    hipStream_t stream = nullptr;

    int numAtoms = 0, numFree = 0, numBodies = 0, numBodyAtoms = 0, numDOF = 0;
    // distance constraints between free atoms: the atoms of each cluster are
    // contiguous and follow the numFree - numClustered unconstrained ones
    int numConstraints = 0, numClusters = 0, numClustered = 0;
    // massless particles, after the body atoms; virtual sites among them
    int numMassless = 0, numSites = 0, numSiteParents = 0;
    DeviceArena siteArena;
    std::vector<int> clusterN;                   // atoms per cluster (small classes first)
    std::vector<int> clusterK;                   // constraints per cluster
    // built on the device: the reference's cleanBodyIndices result and
    // atomIndex, and the internal order (internal atom -> original atom)
    int* dRefBodyIndex = nullptr;
    int* dRefAtomIndex = nullptr;
    int* dAtomOrig = nullptr;
    std::vector<int> refBodyN, refBodyLoc;       // per reference body
    std::vector<int> internalBody;               // internal body -> reference body
    std::vector<int> bodyN, bodyAtom0;           // per internal body
    // host copies of the device arrays above, read back when first needed
    std::vector<int> hostBodyIndex, hostAtomIndex, hostAtomOrig;
    const std::vector<int>& hostCopy(std::vector<int>& copy, const int* src, size_t count) {
        if (copy.size() != count) {
            std::vector<int> v(count);
            if (count > 0) {
                setDevice();
                HIPCHK(hipMemcpyAsync(v.data(), src, count * sizeof(int), hipMemcpyDeviceToHost, stream));
                HIPCHK(hipStreamSynchronize(stream));
            }
            copy.swap(v);
        }
        return copy;
    }
    const std::vector<int>& bodyIndex() { return hostCopy(hostBodyIndex, dRefBodyIndex, numAtoms); }
    const std::vector<int>& atomIndex() { return hostCopy(hostAtomIndex, dRefAtomIndex, numFree + numBodyAtoms); }
    const std::vector<int>& atomOrig() { return hostCopy(hostAtomOrig, dAtomOrig, numAtoms); }
    // shapes: how many shapes of body-fixed coordinates the class's bodies
    // share (0: each body reads its own rows)
    struct Class { int N, firstBody, count, firstAtom, shapes = 0; };
    std::vector<Class> classes;
    std::map<int, int> lanes;                    // atoms per body -> lanes per body
    int block = 256, maxBlocks = 0;
    // bodies of one shape share body-fixed coordinate rows (mixed precision)
    static constexpr int kMaxShapes = 8, kShapeSample = 32, kShareMinBodies = 64;
    bool shareFramesOn = true;                   // rbhip_set_shared_frames
    int* dShapeRows = nullptr;                   // candidate rows, kMaxShapes per class
    int* dShapeMatched = nullptr;                // bodies that matched, per class
    int shapesVersion = 0;                       // bumped when a class's shapes change

    DeviceArena arena;
    DeviceArena tileArena;
    DeviceArena clusterArena;
    DevSys h{};
    DevSys* dSys = nullptr;
    Tile* dPartTiles = nullptr;
    Tile* dForceTiles = nullptr;
    Tile* dSetupTiles = nullptr;
    int nPartTiles = 0, nForceTiles = 0, nSetupTiles = 0;
    // cluster tiles, 64 clusters each: KE slots after the part tiles, and a
    // copy whose slots follow the set-up tiles
    Tile* dClusterTiles = nullptr;
    Tile* dClusterSetupTiles = nullptr;
    int nClusterTiles = 0;
    int clusterClassTiles[4] = {0, 0, 0, 0};      // tiles of 2, 3, 4 atoms, then larger
    static constexpr int kClusterBlock = 64;
    int maxKeSlots = 0, maxPeSlots = 0;
    int keSlots = 0;                             // slots written by the last KE producer
    int* dSlotCounts = nullptr;                  // [0] ke, [1] pe, [2] refined ke, [3] refined pe
    int* dBodyN = nullptr;
    int* dBodyAtom0 = nullptr;
    double* dTmp = nullptr;
    size_t tmpCount = 0;
    double* dRed = nullptr;
    double hRed[4] = {0.0, 0.0, 0.0, 0.0};
    hipEvent_t handoff = nullptr;                // orders a new stream after the old one

    // refined energies (setComputeRefinedEnergies): buffers in their own
    // arena, allocated when enabled; the accumulators are valid after a
    // step and reset by every state change, as the reference's upload zeroes them
    bool refined = false;
    bool refinedValid = false;
    int refKeSlots = 0;
    DeviceArena refArena;
    void* refBuf[4] = {nullptr, nullptr, nullptr, nullptr};

    bool built = false;
    bool stageVFresh = false;
    bool positionsDerived = false;               // body-atom positions hold fp32 r_cm + A(q)^T d only
    // the stored body momenta carry the next step's first half-kick, made
    // with the forces in the force array (after the set-up and part 2; not
    // between part 1 and part 2)
    bool kickFolded = false;
    // atom state adopted from a caller in OpenMM's layout (h.omm*): two slot
    // maps, the live one and the next (a new atom order is moved to it)
    bool adopted = false;
    int* dSlot[2] = {nullptr, nullptr};
    int liveSlot = 0;
    DeviceArena adoptArena;
    DeviceArena corrArena;                        // freeCorr, at the first constrained part 1
    void* dFreeCorr = nullptr;
    // the free atoms' KE partials predate the caller's velocity constraints
    // (set by a part 2 that completes a constrained part 1)
    bool freeKeStale = false;
    bool useGraph = true;
    bool useFusion = true;
    StepGraphs graphs;
    double time = 0.0;
    long long stepCount = 0;

    // bonds in original numbering
    std::vector<int> bondA, bondB;
    std::vector<double> bondLen, bondK;

    size_t realSize() const { return precision == RBHIP_PRECISION_MIXED ? 4 : 8; }

    void setDevice() const { HIPCHK(hipSetDevice(device)); }

    int sysVersion = 0;                          // bumped on every change of h
    void uploadSys() {
        HIPCHK(hipMemcpyAsync(dSys, &h, sizeof(DevSys), hipMemcpyHostToDevice, stream));
        sysVersion++;
    }

    // device copy of everything a step writes, so tuning can run real steps
    // and put the state back
    struct Snapshot {
        DeviceArena arena;
        std::vector<void*> live, copy;
        std::vector<size_t> bytes;
        int keSlots = 0, refKeSlots = 0;
        bool positionsDerived = false, refinedValid = false, freeKeStale = false;
        ~Snapshot() { arena.release(); }
    };
    void takeSnapshot(Snapshot& s) {
        s.arena.stream = stream;
        const size_t R = realSize();
        const size_t nb = std::max(numBodies, 1), as = (size_t)h.atomStride;
        std::vector<std::pair<void*, size_t>> spans = {
            {h.q, 4 * R * nb}, {h.L, 3 * R * nb}, {h.pcm, 3 * R * nb},
            {h.rcm, 3 * sizeof(double) * (size_t)h.bodyStride}, {h.pos, 3 * R * as}, {h.vel, 3 * R * as},
            {h.force, 3 * R * as}, {h.kePart, 4 * sizeof(double) * (size_t)maxKeSlots},
            {h.pePart, 4 * sizeof(double) * (size_t)maxPeSlots}};
        if (h.posLo) spans.push_back({h.posLo, 3 * sizeof(float) * as});
        if (h.refKePart) spans.push_back({h.refKePart, 4 * sizeof(double) * (size_t)maxKeSlots});
        if (adopted) {
            const size_t ps = (size_t)h.ommStride;
            spans.push_back({h.ommPosq, 4 * R * ps});
            if (h.ommPosqCorr) spans.push_back({h.ommPosqCorr, 4 * sizeof(float) * ps});
            spans.push_back({h.ommVelm, 4 * sizeof(double) * ps});
            spans.push_back({h.ommForce, 3 * sizeof(long long) * ps});
        }
        for (auto& sp : spans) {
            void* c = s.arena.alloc(sp.second);
            HIPCHK(hipMemcpyAsync(c, sp.first, sp.second, hipMemcpyDeviceToDevice, stream));
            s.live.push_back(sp.first);
            s.copy.push_back(c);
            s.bytes.push_back(sp.second);
        }
        s.keSlots = keSlots;
        s.refKeSlots = refKeSlots;
        s.positionsDerived = positionsDerived;
        s.refinedValid = refinedValid;
        s.freeKeStale = freeKeStale;
    }
    // enqueues the copies back on `on` (the caller orders and synchronises)
    void restoreSnapshot(const Snapshot& s, hipStream_t on) {
        for (size_t i = 0; i < s.live.size(); i++)
            HIPCHK(hipMemcpyAsync(s.live[i], s.copy[i], s.bytes[i], hipMemcpyDeviceToDevice, on));
        keSlots = s.keSlots;
        refKeSlots = s.refKeSlots;
        positionsDerived = s.positionsDerived;
        refinedValid = s.refinedValid;
        freeKeStale = s.freeKeStale;
    }

    std::vector<Tile> partTiles(int blk, const std::map<int, int>& lanesByN) const {
        std::vector<Tile> t;
        const int freeTile = FREE_PER_THREAD * blk;
        const int numUnconstrained = numFree - numClustered;
        for (int f = 0; f < numUnconstrained; f += freeTile)
            t.push_back(Tile{0, TILE_FREE, f, std::min(freeTile, numUnconstrained - f), 0, 0, 1, 0});
        for (const Class& c : classes) {
            auto it = lanesByN.find(c.N);
            int L = it == lanesByN.end() ? 1 : it->second;
            int per = blk / L;
            for (int b = 0; b < c.count; b += per)
                t.push_back(Tile{0, TILE_BODIES, c.firstBody + b, std::min(per, c.count - b), c.N,
                                 c.firstAtom + b * c.N, L, 0, c.shapes > 0 ? 1 : 0});
        }
        for (size_t i = 0; i < t.size(); i++) t[i].slot = (int)i;
        return launchOrder(t);
    }

    // launch order of part tiles: body tiles by atom count, largest first,
    // with the free-atom tiles spread evenly among them.  A one-tile
    // workgroup of a large body class lives longest (9 us for 20-atom
    // clusters, 6 for 3-site bodies, 3.6 for 256 free atoms on MI355X), so
    // the pass ends with short tiles filling the wave slots instead of long
    // ones draining them; the memory-only free tiles keep sharing the CUs
    // with the rotation-heavy body tiles all through the pass (all of them
    // last makes the end bandwidth-bound).  Each tile keeps its slot, so the
    // partial sums are reduced in the same order whatever the launch order.
    static std::vector<Tile> launchOrder(const std::vector<Tile>& t) {
        std::vector<Tile> freeT, bodyT;
        for (const Tile& x : t) (x.kind == TILE_FREE ? freeT : bodyT).push_back(x);
        std::stable_sort(bodyT.begin(), bodyT.end(), [](const Tile& a, const Tile& b) { return a.N > b.N; });
        const size_t nb = bodyT.size(), nf = freeT.size(), n = nb + nf;
        std::vector<Tile> out;
        out.reserve(n);
        size_t ib = 0, jf = 0;
        for (size_t k = 0; k < n; k++) {
            // free tile j goes to position (j + 1/2) n / nf
            const bool takeFree = jf < nf && (ib == nb || (double)k + 0.5 >= ((double)jf + 0.5) * (double)n / (double)nf);
            out.push_back(takeFree ? freeT[jf++] : bodyT[ib++]);
        }
        return out;
    }

    std::vector<Tile> forceTiles(int blk) const {
        std::vector<Tile> t;
        const int atomTile = FORCE_PER_THREAD * blk;
        for (int a = 0; a < numAtoms; a += atomTile)
            t.push_back(Tile{0, TILE_ATOMS, a, std::min(atomTile, numAtoms - a), 0, 0, 1, 0});
        for (size_t i = 0; i < t.size(); i++) t[i].slot = (int)i;
        return t;
    }

    // every tile must address only this system's arrays and slots
    void validateTiles(const std::vector<Tile>& tiles, int blk, int maxSlots) const {
        for (const Tile& t : tiles) {
            require(t.slot >= 0 && t.slot < maxSlots, "tile slot out of range", RBHIP_E_STATE);
            require(t.count > 0, "empty tile", RBHIP_E_STATE);
            if (t.kind == TILE_FREE) {
                require(t.count <= FREE_PER_THREAD * blk && t.first >= 0 && t.first + t.count <= numFree - numClustered,
                        "free tile out of range",
                        RBHIP_E_STATE);
            } else if (t.kind == TILE_ATOMS) {
                require(t.count <= FORCE_PER_THREAD * blk && t.first >= 0 && t.first + t.count <= numAtoms,
                        "atom tile out of range",
                        RBHIP_E_STATE);
            } else {
                require(t.lanes >= 1 && t.lanes <= 64 && (t.lanes & (t.lanes - 1)) == 0 && blk % t.lanes == 0,
                        "bad lanes per body", RBHIP_E_STATE);
                require(t.count * t.lanes <= blk, "body tile larger than the workgroup", RBHIP_E_STATE);
                require(t.first >= 0 && t.first + t.count <= numBodies, "body tile out of range", RBHIP_E_STATE);
                require(t.atomFirst >= numFree && t.atomFirst + t.count * t.N <= numAtoms,
                        "body tile atoms out of range", RBHIP_E_STATE);
                for (int b = 0; b < t.count; b += std::max(1, t.count - 1))
                    require(bodyN[t.first + b] == t.N && bodyAtom0[t.first + b] == t.atomFirst + b * t.N,
                            "body tile does not match the body layout", RBHIP_E_STATE);
            }
        }
    }

    // tiles of up to kClusterBlock clusters (one lane each) of one size class
    // (N = 2, 3 or 4 atoms with at most N (N - 1) / 2 constraints, held in
    // registers, or N = 0 for the others), KE slots from slot0
    std::vector<Tile> clusterTiles(int slot0) const {
        std::vector<Tile> t;
        auto cls = [&](int c) {
            int n = clusterN[c];
            return n <= 4 && clusterK[c] <= n * (n - 1) / 2 ? n : 0;
        };
        for (int c = 0; c < numClusters;) {
            int e = c;
            while (e < numClusters && e - c < kClusterBlock && cls(e) == cls(c)) e++;
            Tile x{0, TILE_CLUSTERS, c, e - c, cls(c), 0, 1, slot0 + (int)t.size()};
            require(x.slot < maxKeSlots, "cluster tile slot out of range", RBHIP_E_STATE);
            t.push_back(x);
            c = e;
        }
        return t;
    }

    // this context's validated tiles for workgroup size `blk` and its lanes per
    // body, kept for reuse (a batch concatenates them on every change of its
    // member list)
    struct TileCache {
        int block = -1, shapesVersion = -1;
        std::map<int, int> lanes;
        std::vector<Tile> part, force;
    };
    TileCache tileCache;
    const TileCache& tilesFor(int blk) {
        if (tileCache.block != blk || tileCache.lanes != lanes || tileCache.shapesVersion != shapesVersion) {
            tileCache.part = partTiles(blk, lanes);
            tileCache.force = forceTiles(blk);
            validateTiles(tileCache.part, blk, maxKeSlots);
            validateTiles(tileCache.force, blk, maxPeSlots);
            tileCache.block = blk;
            tileCache.shapesVersion = shapesVersion;
            tileCache.lanes = lanes;
        }
        return tileCache;
    }

    // largest workgroup the kernels are compiled for: 512, or 256 for the FP64
    // exact rotation; the FP64 body-state build always runs at <= 256
    int maxBlock() const { return rotationMode == 0 ? 256 : 512; }
    int setupBlock() const { return std::min(block, 256); }

    void rebuildTiles() {
        require(block >= 64 && block <= maxBlock() && block % 64 == 0,
                "block size must be a multiple of 64 in [64, " + std::to_string(maxBlock()) + "]");
        std::vector<Tile> pt = partTiles(block, lanes);
        std::vector<Tile> ft = forceTiles(block);
        std::map<int, int> ones;
        std::vector<Tile> st = partTiles(setupBlock(), ones);
        validateTiles(pt, block, maxKeSlots);
        validateTiles(ft, block, maxPeSlots);
        validateTiles(st, setupBlock(), maxKeSlots);
        // the old tables may still be read by queued work
        HIPCHK(hipStreamSynchronize(stream));
        tileArena.release();
        dPartTiles = tileArena.allocT<Tile>(pt.size() + 1);
        dForceTiles = tileArena.allocT<Tile>(ft.size() + 1);
        dSetupTiles = tileArena.allocT<Tile>(st.size() + 1);
        std::vector<Tile> ct = clusterTiles((int)pt.size()), cst = clusterTiles((int)st.size());
        dClusterTiles = tileArena.allocT<Tile>(ct.size() + 1);
        dClusterSetupTiles = tileArena.allocT<Tile>(cst.size() + 1);
        if (!ct.empty()) {
            HIPCHK(hipMemcpyAsync(dClusterTiles, ct.data(), ct.size() * sizeof(Tile), hipMemcpyHostToDevice, stream));
            HIPCHK(hipMemcpyAsync(dClusterSetupTiles, cst.data(), cst.size() * sizeof(Tile), hipMemcpyHostToDevice,
                                  stream));
        }
        if (!pt.empty())
            HIPCHK(hipMemcpyAsync(dPartTiles, pt.data(), pt.size() * sizeof(Tile), hipMemcpyHostToDevice, stream));
        if (!ft.empty())
            HIPCHK(hipMemcpyAsync(dForceTiles, ft.data(), ft.size() * sizeof(Tile), hipMemcpyHostToDevice, stream));
        if (!st.empty())
            HIPCHK(hipMemcpyAsync(dSetupTiles, st.data(), st.size() * sizeof(Tile), hipMemcpyHostToDevice, stream));
        HIPCHK(hipStreamSynchronize(stream));
        nPartTiles = (int)pt.size();
        nForceTiles = (int)ft.size();
        nSetupTiles = (int)st.size();
        nClusterTiles = (int)ct.size();
        // classes in order 2, 3, 4, larger (N = 0), as the clusters are sorted
        int prev = 0;
        for (int i = 0; i < 4; i++) clusterClassTiles[i] = 0;
        for (const Tile& x : ct) {
            int cls = x.N == 0 ? 3 : x.N - 2;
            require(cls >= prev && cls >= 0 && cls < 4, "cluster tiles out of class order", RBHIP_E_STATE);
            clusterClassTiles[cls]++;
            prev = cls;
        }
        graphs.reset();
    }

    void setSlotCounts(int ke, int pe) {
        int v[2] = {ke, pe};
        HIPCHK(hipMemcpyAsync(dSlotCounts, v, sizeof(v), hipMemcpyHostToDevice, stream));
        HIPCHK(hipStreamSynchronize(stream));
    }

    int layout() const { return adopted ? 1 : 0; }
    // everything a captured step sequence of this context reads
    GraphKey graphKey(int kind) const {
        GraphKey k;
        k.stream = stream;
        k.systems = dSys;
        k.partTiles = dPartTiles;
        k.forceTiles = dForceTiles;
        k.clusterTiles = dClusterTiles;
        k.numPart = nPartTiles;
        k.numForce = nForceTiles;
        k.numClusterTiles = nClusterTiles;
        k.numSites = numSites;
        k.block = block;
        k.maxBlocks = maxBlocks;
        k.layout = layout();
        k.kind = kind;
        return k;
    }
    // (energy: the potential-energy partials too, for the last evaluation of a call)
    void launchForces(bool energy = true) {
        HIPCHK(launch_forces(precision, dSys, dForceTiles, nForceTiles, block, maxBlocks, stream, layout(), energy));
        if (numSites > 0) HIPCHK(launch_site_forces(precision, dSys, numSites, numSiteParents, stream));
    }
    void launchClusters(int phase) {
        if (nClusterTiles > 0)
            HIPCHK(launch_clusters(precision, dSys, dClusterTiles, clusterClassTiles, phase, kClusterBlock, stream));
    }
    void launchSites() {
        if (numSites > 0) HIPCHK(launch_site_positions(precision, dSys, numSites, stream));
    }
    void launchPart1() {
        HIPCHK(launch_part1(precision, rotationMode, dSys, dPartTiles, nPartTiles, block, maxBlocks, stream, layout()));
        launchClusters(CLUSTER_PART1);
        launchSites();
    }
    void launchPart2() {
        HIPCHK(launch_part2(precision, dSys, dPartTiles, nPartTiles, block, maxBlocks, stream, layout()));
        launchClusters(CLUSTER_PART2);
    }
    void launchFused() {
        HIPCHK(launch_fused(precision, rotationMode, dSys, dPartTiles, nPartTiles, block, maxBlocks, stream, layout()));
        launchClusters(CLUSTER_FUSED);
        launchSites();
    }
    int partKeSlots() const { return nPartTiles + nClusterTiles; }
    void launchRefineBegin(bool folded) {
        HIPCHK(launch_refine_begin(precision, rotationMode, dSys, dPartTiles, nPartTiles, folded ? 1 : 0, block,
                                   maxBlocks, stream, layout()));
    }
    void launchRefineEnd() {
        HIPCHK(launch_refine_end(precision, rotationMode, dSys, dPartTiles, nPartTiles, block, maxBlocks, stream,
                                 layout()));
    }

    // 4 doubles reduced from the KE (which = 0), PE (1), refined KE (2) or
    // refined PE (3) partials
    void reduce(int which, double* out4) {
        int counts[4] = {keSlots, nForceTiles, refKeSlots, nPartTiles};
        HIPCHK(hipMemcpyAsync(dSlotCounts, counts, sizeof(counts), hipMemcpyHostToDevice, stream));
        HIPCHK(launch_reduce(dSys, dSlotCounts + which, 1, which, dRed, stream));
        HIPCHK(hipMemcpyAsync(hRed, dRed, 4 * sizeof(double), hipMemcpyDeviceToHost, stream));
        HIPCHK(hipStreamSynchronize(stream));
        for (int c = 0; c < 4; c++) out4[c] = hRed[c];
    }

    void runSetup(int geometry) {
        h.useStageV = stageVFresh ? 1 : 0;
        uploadSys();
        HIPCHK(launch_setup(precision, dSys, dSetupTiles, nSetupTiles, geometry, setupBlock(), stream));
        if (geometry) shareFrames();
        if (nClusterTiles > 0)
            HIPCHK(launch_clusters(precision, dSys, dClusterSetupTiles, clusterClassTiles, CLUSTER_KE, kClusterBlock,
                                   stream));
        keSlots = nSetupTiles + nClusterTiles;
        freeKeStale = false;
        kickDt = dt;
        built = true;
        kickFolded = true;
        refinedValid = false;
        double r[4];
        reduce(0, r);
        // RigidBodySystem.cpp:130: each constraint removes a degree of freedom
        if (geometry) numDOF = numFree - numConstraints + (int)std::lround(r[2]);
    }

    // Bodies of one shape (mixed precision).  Rigid bodies of one molecule
    // type have the same body-fixed coordinates d up to the frame's axis
    // signs and rounding far below fp32's, so most bodies of such a class
    // can read a few shared rows (cached) instead of 12 B per atom of their
    // own in every step pass.  For each class of at least kShareMinBodies
    // bodies, its first kShapeSample bodies give the candidate shapes
    // (pick_shapes, rbhip_shapes.h); k_match_frames then gives every body of
    // the class that lies within tol of a candidate that candidate's values
    // and rows, so a body moves by at most four fp32 ulps of its size (1e-8
    // nm for water).  A class whose bodies matched for at least half of them
    // reads the shared rows (Tile::sharedD).  Runs after every build of the
    // body frames.
    void shareFrames() {
        std::vector<int> before;
        for (Class& c : classes) {
            before.push_back(c.shapes);
            c.shapes = 0;
        }
        if (shareFramesOn && precision == RBHIP_PRECISION_MIXED && numBodies > 0) {
            const size_t nc = classes.size();
            std::vector<int> rows(kMaxShapes * nc, 0), ncand(nc, 0);
            std::vector<double> tol(nc, 0.0);
            std::vector<float> sample;
            for (size_t ci = 0; ci < nc; ci++) {
                const Class& c = classes[ci];
                if (c.count < kShareMinBodies) continue;
                const int ns = std::min(c.count, kShapeSample), row0 = c.firstAtom - numFree, len = ns * c.N;
                sample.assign(3 * (size_t)len, 0.0f);
                for (int comp = 0; comp < 3; comp++)
                    HIPCHK(hipMemcpyAsync(sample.data() + (size_t)comp * len,
                                          reinterpret_cast<const float*>(h.d) + (size_t)comp * h.dStride + row0,
                                          sizeof(float) * len, hipMemcpyDeviceToHost, stream));
                HIPCHK(hipStreamSynchronize(stream));
                std::vector<int> cands;
                if (!pick_shapes(sample.data(), ns, c.N, kMaxShapes, cands, tol[ci])) continue;
                ncand[ci] = (int)cands.size();
                for (size_t i = 0; i < cands.size(); i++) rows[kMaxShapes * ci + i] = row0 + cands[i] * c.N;
            }
            bool any = false;
            for (int n : ncand) any = any || n > 0;
            if (any) {
                std::vector<int> matched(nc, 0);
                HIPCHK(hipMemcpyAsync(dShapeRows, rows.data(), sizeof(int) * rows.size(), hipMemcpyHostToDevice, stream));
                HIPCHK(hipMemcpyAsync(dShapeMatched, matched.data(), sizeof(int) * nc, hipMemcpyHostToDevice, stream));
                for (size_t ci = 0; ci < nc; ci++) {
                    const Class& c = classes[ci];
                    if (ncand[ci] > 0)
                        HIPCHK(launch_match_frames(precision, dSys, c.firstBody, c.count, c.N, c.firstAtom - numFree,
                                                   dShapeRows + kMaxShapes * ci, ncand[ci], tol[ci],
                                                   dShapeMatched + ci, stream));
                }
                HIPCHK(hipMemcpyAsync(matched.data(), dShapeMatched, sizeof(int) * nc, hipMemcpyDeviceToHost, stream));
                HIPCHK(hipStreamSynchronize(stream));
                for (size_t ci = 0; ci < nc; ci++)
                    if (ncand[ci] > 0 && 2 * matched[ci] >= classes[ci].count) classes[ci].shapes = ncand[ci];
            }
        }
        bool changed = false;
        for (size_t ci = 0; ci < classes.size(); ci++) changed = changed || before[ci] != classes[ci].shapes;
        if (changed) {
            shapesVersion++;
            rebuildTiles();
        }
    }

    // original-order FP64 xyz rows (host) -> the staging buffer dTmp, still
    // in original order: the kernels that read it gather through dAtomOrig
    void uploadRows(const double* xyz, size_t rowLength) {
        HIPCHK(hipMemcpyAsync(dTmp, xyz, rowLength * (size_t)numAtoms * sizeof(double), hipMemcpyHostToDevice,
                              stream));
    }

    // the state in original order: the kernels scatter through dAtomOrig
    void download(int what, double* xyz) {
        HIPCHK(launch_stage_out(precision, dSys, dTmp, numAtoms, what, stream, layout(), dAtomOrig));
        if (what == 0 && positionsDerived && precision == RBHIP_PRECISION_MIXED && !adopted)
            HIPCHK(launch_body_atoms_out(precision, dSys, dBodyN, dBodyAtom0, numBodies, dTmp, stream, dAtomOrig));
        HIPCHK(hipMemcpyAsync(xyz, dTmp, 3 * (size_t)numAtoms * sizeof(double), hipMemcpyDeviceToHost, stream));
        HIPCHK(hipStreamSynchronize(stream));
    }

    // Forces replaced outside a step: momenta that carry the next step's
    // half-kick (made with the forces the force array held) get it again
    // with the new ones.  beginForceChange keeps a copy of the old forces
    // (in a buffer allocated at the first such change) when the kick is
    // folded; endForceChange redoes the kick from the difference.
    DeviceArena prevArena;
    void* dForcePrev = nullptr;
    bool beginForceChange() {
        const bool refold = built && kickFolded && numBodies > 0;
        if (!refold) return false;
        fixKick();
        adoptedForcesIn();
        const size_t bytes = 3 * realSize() * (size_t)h.atomStride;
        if (!dForcePrev) {
            prevArena.stream = stream;
            dForcePrev = prevArena.alloc(bytes);
        }
        HIPCHK(hipMemcpyAsync(dForcePrev, h.force, bytes, hipMemcpyDeviceToDevice, stream));
        return true;
    }
    void endForceChange(bool refold) {
        if (!refold) return;
        adoptedForcesIn();
        HIPCHK(launch_refold(precision, dSys, dBodyN, dBodyAtom0, numBodies, dForcePrev, stream));
    }
    // evaluates the library's force field outside a step
    void refreshForces() {
        const bool refold = beginForceChange();
        launchForces();
        endForceChange(refold);
    }

    // the library's force array takes the adopted forces (the body-state
    // build, the re-kick and the body-state download read it there)
    void adoptedForcesIn() {
        if (adopted) HIPCHK(launch_adopted_forces_in(precision, dSys, numAtoms, stream));
    }
    // FP64 copies of the adopted positions and velocities, as asked, for
    // the body-state build, and the adopted forces in the library's force
    // array
    void adoptedStateIn(bool positions, bool velocities) {
        if (!adopted) return;
        if (positions) HIPCHK(launch_stage_out(precision, dSys, h.stageR, numAtoms, 0, stream, 1));
        if (velocities) {
            HIPCHK(launch_stage_out(precision, dSys, h.stageV, numAtoms, 1, stream, 1));
            stageVFresh = true;
        }
        adoptedForcesIn();
    }

    void fixKick() {
        if (built && kickDt != dt) {
            if (kickFolded) {
                adoptedForcesIn();
                HIPCHK(launch_rekick(precision, dSys, dBodyN, dBodyAtom0, numBodies, dt - kickDt, stream));
            }
            kickDt = dt;
        }
    }

    // RigidBodyIntegrator::step(steps) (run_steps)
    void executeSteps(int steps) {
        require(built, "positions must be set before stepping", RBHIP_E_STATE);
        require(steps >= 0, "negative step count");
        fixKick();
        if (steps == 0) return;
        run_steps(*this, steps, refined);
        if (refined) {
            refKeSlots = nPartTiles;
            refinedValid = true;
        }
        keSlots = partKeSlots();
        freeKeStale = false;
        stageVFresh = false;
        positionsDerived = true;
        kickFolded = true;
        time += dt * steps;
        stepCount += steps;
    }

    // turns the refined-energy accumulation on (allocating its buffers once)
    // or off; either way the accumulators start empty.  Kernels see the
    // buffers only while it is on.
    void setRefined(bool on) {
        if (on && numConstraints > 0)
            throw RbError(RBHIP_E_UNSUPPORTED, "refined energies are not supported with constraints");
        if (on && numSites > 0)
            throw RbError(RBHIP_E_UNSUPPORTED, "refined energies are not supported with virtual sites");
        if (on && dFreeCorr)
            throw RbError(RBHIP_E_UNSUPPORTED, "refined energies are not supported with the caller's constraints");
        if (on && !refBuf[0]) {
            const size_t R = realSize();
            refArena.stream = stream;
            refBuf[0] = refArena.alloc(8 * R * (size_t)std::max(numBodies, 1));
            refBuf[1] = refArena.alloc(3 * R * (size_t)std::max(numFree, 1));
            refBuf[2] = refArena.alloc(4 * sizeof(double) * (size_t)maxKeSlots);
            refBuf[3] = refArena.alloc(4 * sizeof(double) * (size_t)maxKeSlots);
        }
        refined = on;
        refinedValid = false;
        refKeSlots = 0;
        h.refBody = on ? refBuf[0] : nullptr;
        h.refFree = on ? refBuf[1] : nullptr;
        h.refKePart = on ? static_cast<double*>(refBuf[2]) : nullptr;
        h.refPePart = on ? static_cast<double*>(refBuf[3]) : nullptr;
        uploadSys();   // (captured graphs hold only plain launches and stay valid)
    }

    ~rbhip_context() {
        graphs.reset();
        adoptArena.release();
        corrArena.release();
        prevArena.release();
        tileArena.release();
        refArena.release();
        clusterArena.release();
        siteArena.release();
        arena.release();
        if (handoff) (void)hipEventDestroy(handoff);
    }
};

// ---------------------------------------------------------------------------
struct rbhip_batch {
    int device = 0;
    hipStream_t stream = nullptr;
    int block = 256, maxBlocks = 0;
    bool useGraph = true;
    std::vector<rbhip_context*> members;
    DeviceArena arena;
    DevSys* dSystems = nullptr;
    Tile* dPartTiles = nullptr;
    Tile* dForceTiles = nullptr;
    int nPart = 0, nForce = 0;
    std::vector<int> memberPartTiles, memberForceTiles;
    std::vector<int> memberVersions;             // sysVersion of each member's DevSys copy
    int* dKeSlots = nullptr;
    double* dRed = nullptr;
    int precision = 0, rotationMode = 0;
    bool useFusion = true;
    StepGraphs graphs;

    void rebuild() {
        graphs.reset();
        arena.release();
        nPart = nForce = 0;
        memberPartTiles.assign(members.size(), 0);
        memberForceTiles.assign(members.size(), 0);
        if (members.empty()) return;
        std::vector<DevSys> sys;
        std::vector<Tile> pt, ft;
        memberVersions.assign(members.size(), 0);
        size_t np = 0, nf = 0;
        for (rbhip_context* c : members) {
            const auto& tc = c->tilesFor(block);
            np += tc.part.size();
            nf += tc.force.size();
        }
        pt.reserve(np);
        ft.reserve(nf);
        for (size_t m = 0; m < members.size(); m++) {
            rbhip_context* c = members[m];
            sys.push_back(c->h);
            memberVersions[m] = c->sysVersion;
            const auto& tc = c->tilesFor(block);
            for (Tile t : tc.part) { t.sys = (int)m; pt.push_back(t); }
            for (Tile t : tc.force) { t.sys = (int)m; ft.push_back(t); }
            memberPartTiles[m] = (int)tc.part.size();
            memberForceTiles[m] = (int)tc.force.size();
        }
        // one launch order over every member's tiles (each keeps its
        // member's slot)
        pt = rbhip_context::launchOrder(pt);
        dSystems = arena.allocT<DevSys>(sys.size());
        dPartTiles = arena.allocT<Tile>(pt.size() + 1);
        dForceTiles = arena.allocT<Tile>(ft.size() + 1);
        dKeSlots = arena.allocT<int>(members.size());
        dRed = arena.allocT<double>(4 * members.size());
        HIPCHK(hipMemcpyAsync(dSystems, sys.data(), sys.size() * sizeof(DevSys), hipMemcpyHostToDevice, stream));
        if (!pt.empty())
            HIPCHK(hipMemcpyAsync(dPartTiles, pt.data(), pt.size() * sizeof(Tile), hipMemcpyHostToDevice, stream));
        if (!ft.empty())
            HIPCHK(hipMemcpyAsync(dForceTiles, ft.data(), ft.size() * sizeof(Tile), hipMemcpyHostToDevice, stream));
        HIPCHK(hipStreamSynchronize(stream));
        nPart = (int)pt.size();
        nForce = (int)ft.size();
    }

    // re-sends the descriptor of every member whose DevSys changed since it
    // was copied (step size, force field); the table's address is unchanged,
    // so captured graphs stay valid
    void refreshSystems() {
        bool copied = false;
        for (size_t m = 0; m < members.size(); m++) {
            rbhip_context* c = members[m];
            if (c->sysVersion == memberVersions[m]) continue;
            HIPCHK(hipMemcpyAsync(dSystems + m, &c->h, sizeof(DevSys), hipMemcpyHostToDevice, stream));
            memberVersions[m] = c->sysVersion;
            copied = true;
        }
        if (copied) HIPCHK(hipStreamSynchronize(stream));
    }

    GraphKey graphKey(int kind) const {
        GraphKey k;
        k.stream = stream;
        k.systems = dSystems;
        k.partTiles = dPartTiles;
        k.forceTiles = dForceTiles;
        k.numPart = nPart;
        k.numForce = nForce;
        k.block = block;
        k.maxBlocks = maxBlocks;
        k.kind = kind;
        return k;
    }
    // (members hold their atoms in the library's own arrays)
    void launchPart1() {
        HIPCHK(launch_part1(precision, rotationMode, dSystems, dPartTiles, nPart, block, maxBlocks, stream, 0));
    }
    void launchForces(bool energy = true) {
        HIPCHK(launch_forces(precision, dSystems, dForceTiles, nForce, block, maxBlocks, stream, 0, energy));
    }
    void launchPart2() {
        HIPCHK(launch_part2(precision, dSystems, dPartTiles, nPart, block, maxBlocks, stream, 0));
    }
    void launchFused() {
        HIPCHK(launch_fused(precision, rotationMode, dSystems, dPartTiles, nPart, block, maxBlocks, stream, 0));
    }
    void launchRefineBegin(bool folded) {
        HIPCHK(launch_refine_begin(precision, rotationMode, dSystems, dPartTiles, nPart, folded ? 1 : 0, block,
                                   maxBlocks, stream, 0));
    }
    void launchRefineEnd() {
        HIPCHK(launch_refine_end(precision, rotationMode, dSystems, dPartTiles, nPart, block, maxBlocks, stream, 0));
    }
    bool anyRefined() const {
        for (const rbhip_context* c : members)
            if (c->refined) return true;
        return false;
    }

    void executeSteps(int steps) { run_steps(*this, steps, anyRefined()); }

    void finishSteps(int steps) {
        for (size_t m = 0; m < members.size(); m++) {
            rbhip_context* c = members[m];
            if (steps > 0 && c->refined) {
                c->refKeSlots = memberPartTiles[m];
                c->refinedValid = true;
            }
            c->keSlots = memberPartTiles[m];
            c->freeKeStale = false;
            c->stageVFresh = false;
            if (steps > 0) {
                c->positionsDerived = true;
                c->kickFolded = true;
            }
            c->time += c->dt * steps;
            c->stepCount += steps;
        }
    }

    // the batch stream waits for work queued on member streams, and members
    // wait for the batch, so member calls and batch steps stay ordered
    void joinMembers() {
        for (rbhip_context* c : members) {
            HIPCHK(hipEventRecord(ordering, c->stream));
            HIPCHK(hipStreamWaitEvent(stream, ordering, 0));
        }
    }
    void releaseMembers() {
        HIPCHK(hipEventRecord(ordering, stream));
        for (rbhip_context* c : members) HIPCHK(hipStreamWaitEvent(c->stream, ordering, 0));
    }

    hipEvent_t ordering = nullptr;

    ~rbhip_batch() {
        graphs.reset();
        arena.release();
        if (ordering) (void)hipEventDestroy(ordering);
    }
};

// ---------------------------------------------------------------------------
extern "C" {

const char* rbhip_last_error(void) { return g_err.c_str(); }

int rbhip_version(void) { return RBHIP_VERSION; }

int rbhip_device_count(int* count) {
    return guard([&] {
        require(count != nullptr, "null output");
        int n = 0;
        if (hipGetDeviceCount(&n) != hipSuccess) n = 0;
        *count = n;
    });
}

int rbhip_device_synchronize(void) {
    return guard([&] { HIPCHK(hipDeviceSynchronize()); });
}

// the checks of rbhip_create on the device bookkeeping's status flags
static void bookkeeping_status(const int* small) {
    if (small[0] & 2) throw RbError(RBHIP_E_ARG, "negative or invalid mass");
    if (small[0] & 1) throw RbError(RBHIP_E_ARG, "massless particles cannot belong to rigid bodies");
}

static int create_context(int numAtoms, const double* masses, const int* bodyIndices, int numConstraints,
                          const int* constraintAtoms, const double* constraintDistances, double stepSize,
                          int rotationMode, int precision, int device, rbhip_context** out) {
    return guard([&] {
        require(out != nullptr, "null output handle");
        *out = nullptr;
        require(numAtoms > 0 && masses != nullptr && bodyIndices != nullptr, "empty system");
        require(rotationMode >= 0, "Rotation mode cannot be negative");
        require(precision == RBHIP_PRECISION_MIXED || precision == RBHIP_PRECISION_DOUBLE, "unknown precision");
        require(stepSize > 0.0, "step size must be positive");
        for (int i = 0; i < numAtoms; i++)
            if (!(masses[i] >= 0.0))
                throw RbError(RBHIP_E_ARG, "negative or invalid mass (particle " + std::to_string(i) + ")");
        int ndev = 0;
        HIPCHK(hipGetDeviceCount(&ndev));
        require(device >= 0 && device < ndev, "no such device");

        std::unique_ptr<rbhip_context> c(new rbhip_context());
        rbhip_context& C = *c;
        C.device = device;
        C.precision = precision;
        C.rotationMode = rotationMode;
        C.dt = C.kickDt = stepSize;
        C.numAtoms = numAtoms;
        C.setDevice();

        // massless particles (virtual sites among them) are neither free nor
        // body atoms (RigidBodySystem.cpp:66-76); they must be labelled free,
        // because the reference's body-atom loop (:99-104) does not skip them
        std::vector<int> massless;
        for (int i = 0; i < numAtoms; i++)
            if (masses[i] == 0.0) {
                if (bodyIndices[i] > 0)
                    throw RbError(RBHIP_E_ARG, "massless particles cannot belong to rigid bodies (particle " +
                                                   std::to_string(i) + ")");
                massless.push_back(i);
            }

        // ---- reference bookkeeping on the device (RigidBodySystem.cpp:28-105):
        // cleanBodyIndices, atomIndex and each body's first atom, kept on the
        // device; the host reads back one int per body
        C.ownStream = pooled_stream(device);
        C.stream = C.ownStream;
        C.arena.stream = C.tileArena.stream = C.stream;
        hipStream_t st = C.stream;
        const size_t na = (size_t)numAtoms;
        C.dRefBodyIndex = C.arena.allocT<int>(na);
        C.dRefAtomIndex = C.arena.allocT<int>(na);
        C.dAtomOrig = C.arena.allocT<int>(na);
        DeviceArena setupArena;   // inputs and scratch of the set-up, freed at its end
        setupArena.stream = st;
        struct ReleaseAtExit {
            DeviceArena& a;
            ~ReleaseAtExit() { a.release(); }
        } releaseSetup{setupArena};
        int* dLabels = setupArena.allocT<int>(na);
        double* dMassIn = setupArena.allocT<double>(na);
        int* dRefStart = setupArena.allocT<int>(na + 1);
        size_t bkBytes = 0;
        HIPCHK(bookkeeping_scratch_bytes(numAtoms, st, &bkBytes));
        void* bkScratch = setupArena.alloc(bkBytes);
        HIPCHK(hipMemcpyAsync(dLabels, bodyIndices, na * sizeof(int), hipMemcpyHostToDevice, st));
        HIPCHK(hipMemcpyAsync(dMassIn, masses, na * sizeof(double), hipMemcpyHostToDevice, st));
        int small[5] = {0, 0, 0, 0, 0};
        HIPCHK(launch_bookkeeping(numAtoms, dLabels, dMassIn, C.dRefBodyIndex, C.dRefAtomIndex, dRefStart, bkScratch,
                                  bkBytes, st, small));
        bookkeeping_status(small);
        C.numFree = small[1];
        C.numBodies = small[2];
        C.numMassless = small[3];
        C.numBodyAtoms = small[4];
        require(C.numMassless == (int)massless.size(), "internal: massless particle count", RBHIP_E_STATE);
        std::vector<int> refStart(C.numBodies + 1);
        HIPCHK(hipMemcpyAsync(refStart.data(), dRefStart, refStart.size() * sizeof(int), hipMemcpyDeviceToHost, st));
        HIPCHK(hipStreamSynchronize(st));
        C.refBodyN.assign(C.numBodies, 0);
        C.refBodyLoc.assign(C.numBodies, 0);
        int maxN = 0;
        for (int b = 0; b < C.numBodies; b++) {
            C.refBodyLoc[b] = refStart[b];
            C.refBodyN[b] = refStart[b + 1] - refStart[b];
            if (C.refBodyN[b] < 2)
                throw RbError(RBHIP_E_UNSUPPORTED, "rigid body " + std::to_string(b + 1) +
                                                       " has a single atom; single-atom bodies are not supported");
            maxN = std::max(maxN, C.refBodyN[b]);
        }

        // ---- constraints between free atoms (RigidBodySystem.cpp:107-112):
        // connected clusters, in the order of their first constraint, go
        // after the unconstrained free atoms, each cluster's atoms together
        std::vector<int> freeOrder;   // (with constraints only; else the reference's free atoms)
        std::vector<std::vector<int>> clusterAtoms, clusterCons;
        require(numConstraints >= 0 && (numConstraints == 0 || (constraintAtoms && constraintDistances)),
                "bad constraint arrays");
        if (numConstraints > 0) {
            freeOrder.reserve(C.numFree);
            for (int i = 0; i < numAtoms; i++)
                if (bodyIndices[i] <= 0 && masses[i] != 0.0) freeOrder.push_back(i);
            require((int)freeOrder.size() == C.numFree, "internal: free atom count", RBHIP_E_STATE);
            std::vector<int> root(numAtoms);
            for (int i = 0; i < numAtoms; i++) root[i] = i;
            auto find = [&](int x) {
                while (root[x] != x) x = root[x] = root[root[x]];
                return x;
            };
            for (int k = 0; k < numConstraints; k++) {
                int a = constraintAtoms[2 * k], b = constraintAtoms[2 * k + 1];
                double d = constraintDistances[k];
                require(a >= 0 && a < numAtoms && b >= 0 && b < numAtoms && a != b,
                        "constraint " + std::to_string(k) + ": bad atom indices");
                require(d > 0.0 && std::isfinite(d), "constraint " + std::to_string(k) + ": bad distance");
                if (bodyIndices[a] > 0 || bodyIndices[b] > 0)
                    throw RbError(RBHIP_E_ARG, "Constraints involving rigid-body atoms are not allowed");
                if (masses[a] == 0.0 || masses[b] == 0.0)
                    throw RbError(RBHIP_E_ARG, "A constraint cannot involve a massless particle");
                root[find(a)] = find(b);
            }
            std::vector<int> id(numAtoms, -1);
            for (int k = 0; k < numConstraints; k++) {
                int r = find(constraintAtoms[2 * k]);
                if (id[r] < 0) {
                    id[r] = (int)clusterCons.size();
                    clusterCons.emplace_back();
                    clusterAtoms.emplace_back();
                }
                clusterCons[id[r]].push_back(k);
            }
            std::vector<int> unconstrained;
            for (int i : freeOrder) {
                int r = find(i);
                if (id[r] < 0) unconstrained.push_back(i);
                else clusterAtoms[id[r]].push_back(i);
            }
            freeOrder = unconstrained;
            // clusters of n = 2-4 atoms with at most n (n - 1) / 2
            // constraints first, by size, then the others: each tile holds
            // one class
            auto key = [&](int c2) {
                size_t n = clusterAtoms[c2].size();
                return n <= 4 && clusterCons[c2].size() <= n * (n - 1) / 2 ? n : 100 + n;
            };
            std::vector<int> bySize(clusterAtoms.size());
            for (size_t c2 = 0; c2 < bySize.size(); c2++) bySize[c2] = (int)c2;
            std::stable_sort(bySize.begin(), bySize.end(), [&](int x, int y) { return key(x) < key(y); });
            std::vector<std::vector<int>> sortedAtoms, sortedCons;
            for (int c2 : bySize) {
                sortedAtoms.push_back(std::move(clusterAtoms[c2]));
                sortedCons.push_back(std::move(clusterCons[c2]));
            }
            clusterAtoms.swap(sortedAtoms);
            clusterCons.swap(sortedCons);
            for (const auto& cc : clusterCons) C.clusterK.push_back((int)cc.size());
            for (const auto& ca : clusterAtoms) {
                if ((int)ca.size() > CLUSTER_MAX_ATOMS)
                    throw RbError(RBHIP_E_UNSUPPORTED, "a cluster of constraints joins " + std::to_string(ca.size()) +
                                                           " atoms; at most " + std::to_string(CLUSTER_MAX_ATOMS) +
                                                           " are supported");
                freeOrder.insert(freeOrder.end(), ca.begin(), ca.end());
                C.numClustered += (int)ca.size();
                C.clusterN.push_back((int)ca.size());
            }
            C.numConstraints = numConstraints;
            C.numClusters = (int)clusterAtoms.size();
        }

        // ---- internal order: bodies grouped by size class (a stable
        // counting sort by atom count), their atoms placed on the device ----
        std::vector<int> order(C.numBodies);
        {
            std::vector<int> next(maxN + 2, 0);
            for (int b = 0; b < C.numBodies; b++) next[C.refBodyN[b] + 1]++;
            for (int v = 0; v <= maxN; v++) next[v + 1] += next[v];
            for (int b = 0; b < C.numBodies; b++) order[next[C.refBodyN[b]]++] = b;
        }
        C.internalBody = order;
        C.bodyN.resize(C.numBodies);
        C.bodyAtom0.resize(C.numBodies);
        int a = C.numFree;
        for (int ib = 0; ib < C.numBodies; ib++) {
            int N = C.refBodyN[order[ib]];
            C.bodyN[ib] = N;
            C.bodyAtom0[ib] = a;
            if (C.classes.empty() || C.classes.back().N != N)
                C.classes.push_back(rbhip_context::Class{N, ib, 0, a});
            C.classes.back().count++;
            a += N;
        }

        // lanes per body: one lane per body up to 7 atoms, then a power of
        // two that leaves each lane 4-7 atoms (shorter serial loops)
        for (const auto& cl : C.classes) {
            int L = 1;
            while (L < 16 && cl.N >= 8 * L) L <<= 1;
            if (L > 1) C.lanes[cl.N] = L;
        }

        // ---- device allocation ----
        if (rotationMode == 0) C.block = 64;   // the FP64 exact rotation is register heavy
        const size_t R = C.realSize();
        const size_t nb = std::max(C.numBodies, 1);
        DevSys& h = C.h;
        h.numFree = C.numFree;
        h.numBodies = C.numBodies;
        h.numBodyAtoms = C.numBodyAtoms;
        h.numAtoms = numAtoms;
        h.rotationMode = rotationMode;
        h.bodyStride = (int)((nb + 63) / 64 * 64);
        h.atomStride = (int)((na + 63) / 64 * 64);
        h.dStride = (int)((std::max<size_t>(C.numBodyAtoms, 1) + 63) / 64 * 64);
        const size_t as = h.atomStride;
        h.dt = stepSize;
        h.tetherScale = 0.0;
        h.q = C.arena.alloc(4 * R * nb);
        h.L = C.arena.alloc(3 * R * nb);
        h.pcm = C.arena.alloc(3 * R * nb);
        h.iI = C.arena.alloc(4 * R * nb);
        h.rcm = C.arena.allocT<double>(3 * (size_t)h.bodyStride);
        h.d = C.arena.alloc(3 * R * (size_t)h.dStride);
        h.dOff = C.arena.allocT<int>(nb);
        h.pos = C.arena.alloc(3 * R * as);
        h.posLo = precision == RBHIP_PRECISION_MIXED ? C.arena.allocT<float>(3 * as) : nullptr;
        h.vel = C.arena.alloc(3 * R * as);
        h.force = C.arena.alloc(3 * R * as);
        h.freeInvM = C.arena.alloc(R * std::max(C.numFree, 1));
        h.mass = C.arena.allocT<double>(na);
        h.stageR = C.arena.allocT<double>(3 * na);
        h.stageV = C.arena.allocT<double>(3 * na);
        h.tetherK = C.arena.alloc(R * na);
        h.r0 = C.arena.alloc(3 * R * as);
        h.bondStart = nullptr;
        h.bondOther = nullptr;
        h.bondLen = nullptr;
        h.bondK = nullptr;
        h.refBody = h.refFree = nullptr;
        h.refKePart = h.refPePart = nullptr;
        size_t nFreeTilesMax = (C.numFree + 63) / 64 + 1;
        // one partial tile per size class (2, 3, 4, larger) at most
        const int clusterSlots = C.numClusters > 0 ? (C.numClusters + 63) / 64 + 4 : 0;
        C.maxKeSlots = (int)(nFreeTilesMax + C.numBodies + C.classes.size() + 1) + clusterSlots;
        C.maxPeSlots = (int)((na + 63) / 64 + 1);
        h.kePart = C.arena.allocT<double>(4 * (size_t)C.maxKeSlots);
        h.pePart = C.arena.allocT<double>(4 * (size_t)C.maxPeSlots);
        C.dSys = C.arena.allocT<DevSys>(1);
        C.dSlotCounts = C.arena.allocT<int>(4);
        C.dBodyN = C.arena.allocT<int>(nb);
        C.dBodyAtom0 = C.arena.allocT<int>(nb);
        C.tmpCount = std::max(3 * na, 36 * nb);
        C.dTmp = C.arena.allocT<double>(C.tmpCount);
        C.dRed = C.arena.allocT<double>(4);
        C.dShapeRows = C.arena.allocT<int>((size_t)rbhip_context::kMaxShapes * std::max<size_t>(C.classes.size(), 1));
        C.dShapeMatched = C.arena.allocT<int>(std::max<size_t>(C.classes.size(), 1));
        require(C.arena.bytes == context_arena_bytes(na, C.numBodies, C.numFree, C.numBodyAtoms,
                                                     (int)C.classes.size(), R, clusterSlots),
                "internal: device size estimate out of step with the allocation", RBHIP_E_STATE);

        // zero state, constants (queued on the context's stream; the host
        // buffers live until the synchronisation below)
        HIPCHK(hipMemsetAsync(h.pos, 0, 3 * R * as, st));
        if (h.posLo) HIPCHK(hipMemsetAsync(h.posLo, 0, 3 * sizeof(float) * as, st));
        HIPCHK(hipMemsetAsync(h.vel, 0, 3 * R * as, st));
        HIPCHK(hipMemsetAsync(h.force, 0, 3 * R * as, st));
        HIPCHK(hipMemsetAsync(h.d, 0, 3 * R * (size_t)h.dStride, st));
        HIPCHK(hipMemsetAsync(h.tetherK, 0, R * na, st));
        HIPCHK(hipMemsetAsync(h.r0, 0, 3 * R * as, st));
        HIPCHK(hipMemsetAsync(h.kePart, 0, 4 * sizeof(double) * C.maxKeSlots, st));
        HIPCHK(hipMemsetAsync(h.pePart, 0, 4 * sizeof(double) * C.maxPeSlots, st));
        if (C.numBodies > 0) {
            HIPCHK(hipMemcpyAsync(C.dBodyN, C.bodyN.data(), C.numBodies * sizeof(int), hipMemcpyHostToDevice, st));
            HIPCHK(hipMemcpyAsync(C.dBodyAtom0, C.bodyAtom0.data(), C.numBodies * sizeof(int), hipMemcpyHostToDevice,
                                  st));
        }
        // the internal atom order, then the masses and free-atom inverse
        // masses gathered through it
        int* dOrder = setupArena.allocT<int>(C.numBodies);
        int* dFreeOrder = freeOrder.empty() ? nullptr : setupArena.allocT<int>(freeOrder.size());
        int* dMassless = setupArena.allocT<int>(massless.size());
        if (C.numBodies > 0)
            HIPCHK(hipMemcpyAsync(dOrder, order.data(), order.size() * sizeof(int), hipMemcpyHostToDevice, st));
        if (dFreeOrder)
            HIPCHK(hipMemcpyAsync(dFreeOrder, freeOrder.data(), freeOrder.size() * sizeof(int),
                                  hipMemcpyHostToDevice, st));
        if (!massless.empty())
            HIPCHK(hipMemcpyAsync(dMassless, massless.data(), massless.size() * sizeof(int), hipMemcpyHostToDevice,
                                  st));
        HIPCHK(launch_internal_order(C.numFree, C.numBodies, C.numMassless, C.numFree + C.numBodyAtoms,
                                     C.dRefAtomIndex, dRefStart, dFreeOrder, dOrder, C.dBodyN, C.dBodyAtom0,
                                     dMassless, C.dAtomOrig, st));
        HIPCHK(launch_masses_in(precision, numAtoms, C.numFree, C.dAtomOrig, dMassIn, h.mass, h.freeInvM, st));
        if (C.numClusters > 0) {
            // cluster tables in internal atom numbering (the free atoms'
            // internal order is freeOrder)
            std::vector<int> internalOf(numAtoms, -1);
            for (int k = 0; k < C.numFree; k++) internalOf[freeOrder[k]] = k;
            std::vector<int> cs(1, C.numFree - C.numClustered), ks(1, 0), ca, cb;
            std::vector<double> cd;
            for (int c2 = 0; c2 < C.numClusters; c2++) {
                cs.push_back(cs.back() + (int)clusterAtoms[c2].size());
                for (int k : clusterCons[c2]) {
                    ca.push_back(internalOf[constraintAtoms[2 * k]]);
                    cb.push_back(internalOf[constraintAtoms[2 * k + 1]]);
                    cd.push_back(constraintDistances[k]);
                }
                ks.push_back((int)ca.size());
            }
            C.clusterArena.stream = st;
            h.numClusters = C.numClusters;
            h.clusterStart = C.clusterArena.allocT<int>(cs.size());
            h.consStart = C.clusterArena.allocT<int>(ks.size());
            h.consA = C.clusterArena.allocT<int>(ca.size());
            h.consB = C.clusterArena.allocT<int>(cb.size());
            h.consDist = C.clusterArena.allocT<double>(cd.size());
            h.constraintFail = C.clusterArena.allocT<unsigned>(1);
            h.constraintTol = 1e-5;
            HIPCHK(hipMemcpyAsync(h.clusterStart, cs.data(), cs.size() * sizeof(int), hipMemcpyHostToDevice, st));
            HIPCHK(hipMemcpyAsync(h.consStart, ks.data(), ks.size() * sizeof(int), hipMemcpyHostToDevice, st));
            HIPCHK(hipMemcpyAsync(h.consA, ca.data(), ca.size() * sizeof(int), hipMemcpyHostToDevice, st));
            HIPCHK(hipMemcpyAsync(h.consB, cb.data(), cb.size() * sizeof(int), hipMemcpyHostToDevice, st));
            HIPCHK(hipMemcpyAsync(h.consDist, cd.data(), cd.size() * sizeof(double), hipMemcpyHostToDevice, st));
            HIPCHK(hipMemsetAsync(h.constraintFail, 0, sizeof(unsigned), st));
        }
        C.uploadSys();
        HIPCHK(hipStreamSynchronize(st));
        C.rebuildTiles();
        *out = c.release();
    });
}

int rbhip_create(int numAtoms, const double* masses, const int* bodyIndices, double stepSize, int rotationMode,
                 int precision, int device, rbhip_context** out) {
    return create_context(numAtoms, masses, bodyIndices, 0, nullptr, nullptr, stepSize, rotationMode, precision,
                          device, out);
}

int rbhip_create_constrained(int numAtoms, const double* masses, const int* bodyIndices, int numConstraints,
                             const int* constraintAtoms, const double* constraintDistances, double stepSize,
                             int rotationMode, int precision, int device, rbhip_context** out) {
    return create_context(numAtoms, masses, bodyIndices, numConstraints, constraintAtoms, constraintDistances,
                          stepSize, rotationMode, precision, device, out);
}

int rbhip_set_virtual_sites(rbhip_context* ctx, int count, const int* types, const int* siteAtoms,
                            const int* parents, const double* weights) {
    return guard([&] {
        require(ctx && count >= 0 && (count == 0 || (types && siteAtoms && parents && weights)), "bad arguments");
        rbhip_context& C = *ctx;
        require(C.numSites == 0, "virtual sites are already set", RBHIP_E_STATE);
        require(!C.refined, "refined energies are not supported with virtual sites", RBHIP_E_UNSUPPORTED);
        if (count == 0) return;
        std::vector<int> internalOf(C.numAtoms);
        const std::vector<int>& atomOrig = C.atomOrig();
        for (int k = 0; k < C.numAtoms; k++) internalOf[atomOrig[k]] = k;
        const int firstMassless = C.numAtoms - C.numMassless;
        std::vector<char> isSite(C.numAtoms, 0);
        std::vector<int> sa(count), st(count), sp(3 * (size_t)count, 0);
        std::vector<double> sw(12 * (size_t)count, 0.0);
        for (int v = 0; v < count; v++) {
            const int t = types[v], site = siteAtoms[v];
            require(t == RBHIP_SITE_TWO_AVERAGE || t == RBHIP_SITE_THREE_AVERAGE || t == RBHIP_SITE_OUT_OF_PLANE ||
                        t == RBHIP_SITE_LOCAL_COORDINATES,
                    "unknown virtual site type");
            require(site >= 0 && site < C.numAtoms && internalOf[site] >= firstMassless,
                    "a virtual site must be a massless particle");
            require(!isSite[site], "particle " + std::to_string(site) + " is given two virtual sites");
            isSite[site] = 1;
            const int np = t == RBHIP_SITE_TWO_AVERAGE ? 2 : 3;
            for (int k = 0; k < np; k++) {
                int pa = parents[3 * v + k];
                require(pa >= 0 && pa < C.numAtoms && pa != site && internalOf[pa] < firstMassless,
                        "a virtual site's parents must be massive particles");
                sp[3 * v + k] = internalOf[pa];
            }
            const int nw = t == RBHIP_SITE_LOCAL_COORDINATES ? 12 : np;
            for (int k = 0; k < nw; k++) sw[12 * (size_t)v + k] = weights[12 * (size_t)v + k];
            sa[v] = internalOf[site];
            st[v] = t;
        }
        // each parent's (site, role) entries in site order
        std::map<int, std::vector<std::pair<int, int>>> byParent;
        for (int v = 0; v < count; v++)
            for (int k = 0; k < (st[v] == RBHIP_SITE_TWO_AVERAGE ? 2 : 3); k++) byParent[sp[3 * v + k]].push_back({v, k});
        std::vector<int> pa, ps(1, 0), pv, pr;
        for (auto& kv : byParent) {
            pa.push_back(kv.first);
            for (auto& e : kv.second) {
                pv.push_back(e.first);
                pr.push_back(e.second);
            }
            ps.push_back((int)pv.size());
        }
        C.setDevice();
        hipStream_t s = C.stream;
        C.siteArena.stream = s;
        DevSys& h = C.h;
        h.siteAtom = C.siteArena.allocT<int>(sa.size());
        h.siteType = C.siteArena.allocT<int>(st.size());
        h.siteParent = C.siteArena.allocT<int>(sp.size());
        h.siteW = C.siteArena.allocT<double>(sw.size());
        h.parentAtom = C.siteArena.allocT<int>(pa.size());
        h.parentStart = C.siteArena.allocT<int>(ps.size());
        h.parentSite = C.siteArena.allocT<int>(pv.size());
        h.parentRole = C.siteArena.allocT<int>(pr.size());
        HIPCHK(hipMemcpyAsync(h.siteAtom, sa.data(), sa.size() * sizeof(int), hipMemcpyHostToDevice, s));
        HIPCHK(hipMemcpyAsync(h.siteType, st.data(), st.size() * sizeof(int), hipMemcpyHostToDevice, s));
        HIPCHK(hipMemcpyAsync(h.siteParent, sp.data(), sp.size() * sizeof(int), hipMemcpyHostToDevice, s));
        HIPCHK(hipMemcpyAsync(h.siteW, sw.data(), sw.size() * sizeof(double), hipMemcpyHostToDevice, s));
        HIPCHK(hipMemcpyAsync(h.parentAtom, pa.data(), pa.size() * sizeof(int), hipMemcpyHostToDevice, s));
        HIPCHK(hipMemcpyAsync(h.parentStart, ps.data(), ps.size() * sizeof(int), hipMemcpyHostToDevice, s));
        HIPCHK(hipMemcpyAsync(h.parentSite, pv.data(), pv.size() * sizeof(int), hipMemcpyHostToDevice, s));
        HIPCHK(hipMemcpyAsync(h.parentRole, pr.data(), pr.size() * sizeof(int), hipMemcpyHostToDevice, s));
        h.numSites = count;
        h.numSiteParents = (int)pa.size();
        C.numSites = count;
        C.numSiteParents = (int)pa.size();
        C.uploadSys();
        C.graphs.reset();   // the step sequence gains the site kernels
        HIPCHK(hipStreamSynchronize(s));
    });
}

int rbhip_set_constraint_tolerance(rbhip_context* ctx, double tolerance) {
    return guard([&] {
        require(ctx, "null context");
        require(tolerance > 0.0 && std::isfinite(tolerance), "constraint tolerance must be positive");
        ctx->setDevice();
        ctx->h.constraintTol = tolerance;
        ctx->uploadSys();
    });
}

int rbhip_constraint_failures(rbhip_context* ctx, long long* count) {
    return guard([&] {
        require(ctx && count, "null argument");
        *count = 0;
        if (ctx->numClusters == 0) return;
        ctx->setDevice();
        unsigned n = 0;
        HIPCHK(hipMemcpyAsync(&n, ctx->h.constraintFail, sizeof(unsigned), hipMemcpyDeviceToHost, ctx->stream));
        HIPCHK(hipStreamSynchronize(ctx->stream));
        *count = n;
    });
}

int rbhip_destroy(rbhip_context* ctx) {
    return guard([&] {
        if (!ctx) return;
        (void)hipSetDevice(ctx->device);
        // device memory is freed in the context's stream order; only
        // instantiated graphs wait for the stream before they go
        if (ctx->graphs.any()) (void)hipStreamSynchronize(ctx->stream);
        delete ctx;
    });
}

int rbhip_set_stream(rbhip_context* ctx, rbhip_stream stream) {
    return guard([&] {
        require(ctx, "null context");
        hipStream_t next = stream ? reinterpret_cast<hipStream_t>(stream) : ctx->ownStream;
        if (next == ctx->stream) return;
        HIPCHK(hipSetDevice(ctx->device));
        if (ctx->graphs.any()) {
            // captured graphs may still be running on the old stream
            HIPCHK(hipStreamSynchronize(ctx->stream));
            ctx->graphs.reset();
        } else {
            // the new stream starts after the work queued on the old one
            if (!ctx->handoff) HIPCHK(hipEventCreateWithFlags(&ctx->handoff, hipEventDisableTiming));
            HIPCHK(hipEventRecord(ctx->handoff, ctx->stream));
            HIPCHK(hipStreamWaitEvent(next, ctx->handoff, 0));
        }
        ctx->stream = next;
        ctx->arena.stream = ctx->tileArena.stream = ctx->refArena.stream = ctx->clusterArena.stream =
            ctx->siteArena.stream = ctx->adoptArena.stream = ctx->prevArena.stream = ctx->corrArena.stream = ctx->stream;
    });
}

int rbhip_device_bytes(rbhip_context* ctx, size_t* bytes) {
    return guard([&] {
        require(ctx && bytes, "null argument");
        *bytes = ctx->arena.bytes + ctx->tileArena.bytes + ctx->refArena.bytes + ctx->clusterArena.bytes +
                 ctx->siteArena.bytes + ctx->adoptArena.bytes + ctx->prevArena.bytes;
    });
}

int rbhip_estimate_bytes(int numAtoms, const int* bodyIndices, int precision, size_t* bytes, int* numBodiesOut) {
    return guard([&] {
        require(numAtoms > 0 && bodyIndices && bytes, "bad arguments");
        require(precision == RBHIP_PRECISION_MIXED || precision == RBHIP_PRECISION_DOUBLE, "unknown precision");
        // the body and free-atom counts and the distinct body sizes that
        // clean_body_indices gives; labels spanning at most 4n + 1024 values
        // are counted directly, wider ones through the renumbering
        int lo = INT_MAX, hi = 0, numFree = 0;
        for (int i = 0; i < numAtoms; i++) {
            int v = bodyIndices[i];
            if (v <= 0) {
                numFree++;
            } else {
                lo = std::min(lo, v);
                hi = std::max(hi, v);
            }
        }
        std::vector<int> n;
        if (numFree < numAtoms && (long long)hi - lo <= 4LL * numAtoms + 1024) {
            n.assign((size_t)(hi - lo) + 1, 0);
            for (int i = 0; i < numAtoms; i++)
                if (bodyIndices[i] > 0) n[bodyIndices[i] - lo]++;
        } else if (numFree < numAtoms) {
            std::vector<int> bi = clean_body_indices(bodyIndices, numAtoms);
            n.assign(*std::max_element(bi.begin(), bi.end()), 0);
            for (int v : bi)
                if (v > 0) n[v - 1]++;
        }
        int numBodies = 0, classes = 0;
        std::vector<char> seen((size_t)numAtoms + 1, 0);
        for (int c : n) {
            if (c == 0) continue;
            numBodies++;
            if (!seen[c]) {
                seen[c] = 1;
                classes++;
            }
        }
        size_t R = precision == RBHIP_PRECISION_MIXED ? 4 : 8;
        *bytes = context_arena_bytes(numAtoms, numBodies, numFree, numAtoms - numFree, classes, R) +
                 tile_bytes_bound(numAtoms, numBodies, numFree, classes);
        if (numBodiesOut) *numBodiesOut = numBodies;
    });
}

int rbhip_body_bookkeeping(int numAtoms, const int* bodyIndices, const double* masses, int* bodyIndexOut,
                           int* atomIndexOut, int* bodyAtomStartOut, int* counts4, int device, rbhip_stream stream) {
    return guard([&] {
        require(numAtoms > 0 && bodyIndices && masses && bodyIndexOut && atomIndexOut && bodyAtomStartOut && counts4,
                "bad arguments");
        int ndev = 0;
        HIPCHK(hipGetDeviceCount(&ndev));
        require(device >= 0 && device < ndev, "no such device");
        HIPCHK(hipSetDevice(device));
        hipStream_t st = stream ? reinterpret_cast<hipStream_t>(stream) : pooled_stream(device);
        size_t bytes = 0;
        HIPCHK(bookkeeping_scratch_bytes(numAtoms, st, &bytes));
        DeviceArena arena;
        arena.stream = st;
        void* scratch = arena.alloc(bytes);
        int small[5] = {0, 0, 0, 0, 0};
        hipError_t e = launch_bookkeeping(numAtoms, bodyIndices, masses, bodyIndexOut, atomIndexOut, bodyAtomStartOut,
                                          scratch, bytes, st, small);
        arena.release();
        HIPCHK(e);
        bookkeeping_status(small);
        for (int k = 0; k < 4; k++) counts4[k] = small[k + 1];
    });
}

// ---------------------------------------------------------------------------
// rbhip_body_bookkeeping_noalloc: the bookkeeping on a scratch buffer that is
// a file-scope device array (one copy per GPU, held until the process ends)
// instead of a stream-ordered allocation.  In host code `scratch` names a
// host-side stand-in, not device memory: the device address of each GPU's
// copy comes from hipGetSymbolAddress, taken once per device with that
// device current, and only that address is handed to kernels.  Calls on one
// device share the array, so they are serialised here (each call waits for
// its own stream before it returns); every region is written before it is
// read, and the status words are cleared by the first kernel of each call.
// ---------------------------------------------------------------------------
constexpr size_t kStaticScratchBytes = size_t(40) << 20;
static __device__ __align__(256) unsigned char scratch[kStaticScratchBytes];

struct StaticScratch {
    std::mutex mtx;                 // held for a whole call on its device
    void* ptr = nullptr;            // the device address of this GPU's copy
};

// (the device must be current)
static StaticScratch& static_scratch(int device) {
    static std::mutex mapMtx;
    static std::map<int, std::unique_ptr<StaticScratch>> perDevice;
    std::lock_guard<std::mutex> lock(mapMtx);
    std::unique_ptr<StaticScratch>& s = perDevice[device];
    if (!s) {
        std::unique_ptr<StaticScratch> fresh(new StaticScratch());
        void* ptr = nullptr;
        HIPCHK(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
        require(ptr != nullptr, "no device address for the static scratch", RBHIP_E_HIP);
        fresh->ptr = ptr;
        s = std::move(fresh);
    }
    return *s;
}

int rbhip_body_bookkeeping_noalloc(int numAtoms, const int* bodyIndices, const double* masses, int* bodyIndexOut,
                                   int* atomIndexOut, int* bodyAtomStartOut, int* counts4, int device,
                                   rbhip_stream stream) {
    return guard([&] {
        require(numAtoms > 0 && bodyIndices && masses && bodyIndexOut && atomIndexOut && bodyAtomStartOut && counts4,
                "bad arguments");
        int ndev = 0;
        HIPCHK(hipGetDeviceCount(&ndev));
        require(device >= 0 && device < ndev, "no such device");
        HIPCHK(hipSetDevice(device));
        hipStream_t st = reinterpret_cast<hipStream_t>(stream);   // NULL: the device's null stream
        size_t bytes = 0;
        HIPCHK(bookkeeping_scratch_bytes(numAtoms, st, &bytes));
        if (bytes > kStaticScratchBytes)
            throw RbError(RBHIP_E_ARG, std::to_string(numAtoms) + " atoms need " + std::to_string(bytes) +
                                           " bytes of scratch; the static scratch holds " +
                                           std::to_string(kStaticScratchBytes) +
                                           " (up to about 1.2 million atoms): use rbhip_body_bookkeeping");
        StaticScratch& sc = static_scratch(device);
        std::lock_guard<std::mutex> lock(sc.mtx);
        int small[5] = {0, 0, 0, 0, 0};
        HIPCHK(launch_bookkeeping(numAtoms, bodyIndices, masses, bodyIndexOut, atomIndexOut, bodyAtomStartOut,
                                  sc.ptr, kStaticScratchBytes, st, small));
        bookkeeping_status(small);
        for (int k = 0; k < 4; k++) counts4[k] = small[k + 1];
    });
}

int rbhip_get_counts(rbhip_context* ctx, int* out5) {
    return guard([&] {
        require(ctx && out5, "null argument");
        // RigidBodySystem.cpp:64-79: virtual sites are not actual atoms, and
        // numBodyAtoms = numActualAtoms - numFree counts massless non-site
        // particles too
        const int actual = ctx->numAtoms - ctx->numSites;
        out5[0] = ctx->numFree;
        out5[1] = ctx->numBodies;
        out5[2] = actual;
        out5[3] = actual - ctx->numFree;
        out5[4] = ctx->numDOF;
    });
}

int rbhip_get_atom_index(rbhip_context* ctx, int* out) {
    return guard([&] {
        require(ctx && out, "null argument");
        ctx->setDevice();
        // free and body atoms, then zeros up to numActualAtoms
        const std::vector<int>& ai = ctx->atomIndex();
        const size_t n = (size_t)(ctx->numAtoms - ctx->numSites);
        std::copy(ai.begin(), ai.begin() + std::min(n, ai.size()), out);
        std::fill(out + std::min(n, ai.size()), out + n, 0);
    });
}

int rbhip_get_body_index(rbhip_context* ctx, int* out) {
    return guard([&] {
        require(ctx && out, "null argument");
        ctx->setDevice();
        const std::vector<int>& bi = ctx->bodyIndex();
        std::copy(bi.begin(), bi.end(), out);
    });
}

int rbhip_set_tether(rbhip_context* ctx, const double* k, const double* r0) {
    return guard([&] {
        require(ctx && k && r0, "null argument");
        rbhip_context& C = *ctx;
        C.setDevice();
        size_t na = C.numAtoms;
        // anchors, then spring constants, at the state precision, each
        // through the staging buffer (the second upload waits in stream
        // order for the first gather)
        C.uploadRows(r0, 3);
        HIPCHK(launch_tether_in(C.precision, C.dSys, C.dTmp, C.numAtoms, 0, C.dAtomOrig, C.stream));
        C.uploadRows(k, 1);
        HIPCHK(launch_tether_in(C.precision, C.dSys, C.dTmp, C.numAtoms, 1, C.dAtomOrig, C.stream));
        C.h.tetherScale = 1.0;
        C.h.tetherUniform = 1;
        C.h.tetherK0 = k[0];
        for (size_t i = 1; i < na; i++)
            if (k[i] != k[0]) C.h.tetherUniform = 0;
        C.uploadSys();
        HIPCHK(hipStreamSynchronize(C.stream));
    });
}

int rbhip_add_bonds(rbhip_context* ctx, int count, const int* atomPairs, const double* length, const double* k) {
    return guard([&] {
        require(ctx && count >= 0 && (count == 0 || (atomPairs && length && k)), "bad bond arguments");
        rbhip_context& C = *ctx;
        C.setDevice();
        for (int b = 0; b < count; b++) {
            int x = atomPairs[2 * b], y = atomPairs[2 * b + 1];
            require(x >= 0 && y >= 0 && x < C.numAtoms && y < C.numAtoms && x != y, "bond atom out of range");
            C.bondA.push_back(x);
            C.bondB.push_back(y);
            C.bondLen.push_back(length[b]);
            C.bondK.push_back(k[b]);
        }
        // CSR per internal atom: each bond listed for both of its atoms
        std::vector<int> origToInt(C.numAtoms);
        const std::vector<int>& atomOrig = C.atomOrig();
        for (int i = 0; i < C.numAtoms; i++) origToInt[atomOrig[i]] = i;
        std::vector<int> start(C.numAtoms + 1, 0);
        for (size_t b = 0; b < C.bondA.size(); b++) {
            start[origToInt[C.bondA[b]] + 1]++;
            start[origToInt[C.bondB[b]] + 1]++;
        }
        for (int i = 0; i < C.numAtoms; i++) start[i + 1] += start[i];
        size_t ne = start[C.numAtoms];
        std::vector<int> other(ne), pos(start.begin(), start.end() - 1);
        std::vector<double> len(ne), kk(ne);
        for (size_t b = 0; b < C.bondA.size(); b++) {
            int x = origToInt[C.bondA[b]], y = origToInt[C.bondB[b]];
            other[pos[x]] = y; len[pos[x]] = C.bondLen[b]; kk[pos[x]++] = C.bondK[b];
            other[pos[y]] = x; len[pos[y]] = C.bondLen[b]; kk[pos[y]++] = C.bondK[b];
        }
        C.h.bondStart = C.arena.allocT<int>(start.size());
        C.h.bondOther = C.arena.allocT<int>(std::max<size_t>(ne, 1));
        C.h.bondLen = C.arena.allocT<double>(std::max<size_t>(ne, 1));
        C.h.bondK = C.arena.allocT<double>(std::max<size_t>(ne, 1));
        HIPCHK(hipMemcpyAsync(C.h.bondStart, start.data(), start.size() * sizeof(int), hipMemcpyHostToDevice, C.stream));
        if (ne) {
            HIPCHK(hipMemcpyAsync(C.h.bondOther, other.data(), ne * sizeof(int), hipMemcpyHostToDevice, C.stream));
            HIPCHK(hipMemcpyAsync(C.h.bondLen, len.data(), ne * sizeof(double), hipMemcpyHostToDevice, C.stream));
            HIPCHK(hipMemcpyAsync(C.h.bondK, kk.data(), ne * sizeof(double), hipMemcpyHostToDevice, C.stream));
        }
        C.uploadSys();
        HIPCHK(hipStreamSynchronize(C.stream));
    });
}

int rbhip_compute_forces(rbhip_context* ctx, double* energy) {
    return guard([&] {
        require(ctx, "null context");
        ctx->setDevice();
        ctx->refreshForces();
        if (energy) {
            double r[4];
            ctx->reduce(1, r);
            *energy = r[0];
        }
    });
}

int rbhip_set_forces(rbhip_context* ctx, const double* xyz) {
    return guard([&] {
        require(ctx && xyz, "null argument");
        rbhip_context& C = *ctx;
        require(!C.adopted, "the forces of this context live in its adopted force array", RBHIP_E_STATE);
        C.setDevice();
        C.uploadRows(xyz, 3);
        const bool refold = C.beginForceChange();
        HIPCHK(launch_stage_in(C.precision, C.dSys, C.dTmp, C.numAtoms, 2, C.stream, 0, C.dAtomOrig));
        // forces on virtual sites move to their parents, as a force
        // evaluation leaves them
        if (C.numSites > 0) HIPCHK(launch_site_forces(C.precision, C.dSys, C.numSites, C.numSiteParents, C.stream));
        C.endForceChange(refold);
        HIPCHK(hipStreamSynchronize(C.stream));
    });
}

// internal atom -> slot in arrays of OpenMM order `order` (order[s] = the
// original atom held at slot s; a permutation of 0 .. length - 1)
static std::vector<int> slots_for_order(rbhip_context& C, const int* order, int length) {
    require(order != nullptr && length >= C.numAtoms, "atom order shorter than the system");
    std::vector<int> inv(length, -1);
    for (int s = 0; s < length; s++) {
        int o = order[s];
        require(o >= 0 && o < length && inv[o] < 0, "atom order is not a permutation");
        inv[o] = s;
    }
    std::vector<int> slot(C.numAtoms);
    const std::vector<int>& atomOrig = C.atomOrig();
    for (int i = 0; i < C.numAtoms; i++) slot[i] = inv[atomOrig[i]];
    return slot;
}

int rbhip_adopt_arrays(rbhip_context* ctx, void* posq, void* posqCorrection, void* velm, long long* force,
                       int paddedNumAtoms, const int* atomOrder, int orderLength) {
    return guard([&] {
        require(ctx, "null context");
        rbhip_context& C = *ctx;
        require(!C.adopted, "the context already holds adopted arrays", RBHIP_E_STATE);
        require(posq && velm && force, "null array");
        require((C.precision == RBHIP_PRECISION_MIXED) == (posqCorrection != nullptr),
                "posqCorrection is required in mixed precision and absent in double precision");
        require(paddedNumAtoms >= C.numAtoms && orderLength <= paddedNumAtoms, "arrays shorter than the system");
        require(C.numConstraints == 0, "adopted arrays are not supported with constraints", RBHIP_E_UNSUPPORTED);
        require(C.numSites == 0, "adopted arrays are not supported with virtual sites", RBHIP_E_UNSUPPORTED);
        for (void* p : {posq, posqCorrection, velm, (void*)force})
            require(((uintptr_t)p & 15) == 0, "adopted arrays must be 16-byte aligned");
        std::vector<int> slot = slots_for_order(C, atomOrder, orderLength);
        C.setDevice();
        HIPCHK(hipStreamSynchronize(C.stream));
        C.adoptArena.stream = C.stream;
        C.dSlot[0] = C.adoptArena.allocT<int>(C.numAtoms);
        C.dSlot[1] = C.adoptArena.allocT<int>(C.numAtoms);
        C.liveSlot = 0;
        HIPCHK(hipMemcpyAsync(C.dSlot[0], slot.data(), slot.size() * sizeof(int), hipMemcpyHostToDevice, C.stream));
        C.h.ommSlot = C.dSlot[0];
        C.h.ommPosq = posq;
        C.h.ommPosqCorr = static_cast<float*>(posqCorrection);
        C.h.ommVelm = static_cast<double*>(velm);
        C.h.ommForce = force;
        C.h.ommStride = paddedNumAtoms;
        C.adopted = true;
        C.built = false;                 // the state is read from the arrays at the next state change
        C.positionsDerived = false;
        C.graphs.reset();
        C.uploadSys();
        HIPCHK(hipStreamSynchronize(C.stream));
    });
}

int rbhip_set_atom_order(rbhip_context* ctx, const int* atomOrder, int orderLength) {
    return guard([&] {
        require(ctx, "null context");
        rbhip_context& C = *ctx;
        require(C.adopted, "the context holds no adopted arrays", RBHIP_E_STATE);
        require(orderLength <= C.h.ommStride, "atom order longer than the arrays");
        std::vector<int> slot = slots_for_order(C, atomOrder, orderLength);
        C.setDevice();
        const int next = 1 - C.liveSlot;
        HIPCHK(hipMemcpyAsync(C.dSlot[next], slot.data(), slot.size() * sizeof(int), hipMemcpyHostToDevice, C.stream));
        // the forces stay with their atoms (ReorderListener,
        // CudaRigidBodyKernels.cpp:69-113): the arrays' owner permutes
        // positions and velocities, the force array is moved here
        HIPCHK(launch_adopted_move_forces(C.dSys, reinterpret_cast<long long*>(C.dTmp), C.dSlot[next], C.numAtoms,
                                          C.stream));
        C.liveSlot = next;
        C.h.ommSlot = C.dSlot[next];
        C.uploadSys();
        HIPCHK(hipStreamSynchronize(C.stream));
    });
}

int rbhip_state_changed(rbhip_context* ctx, int what) {
    return guard([&] {
        require(ctx, "null context");
        rbhip_context& C = *ctx;
        require(C.adopted, "rbhip_state_changed reads adopted arrays; use rbhip_set_positions / rbhip_set_velocities",
                RBHIP_E_STATE);
        require(what == RBHIP_STATE_POSITIONS || what == RBHIP_STATE_VELOCITIES, "unknown state change");
        require(what == RBHIP_STATE_POSITIONS || C.built, "positions must be set before velocities", RBHIP_E_STATE);
        C.setDevice();
        C.adoptedStateIn(what == RBHIP_STATE_POSITIONS, true);
        const int geometry = what == RBHIP_STATE_POSITIONS ? 1 : 0;
        C.h.useStageR = geometry;
        C.runSetup(geometry);
        C.h.useStageR = 0;
        C.h.corrPending = 0;      // (a pending constrained step is abandoned)
        C.uploadSys();
        HIPCHK(hipStreamSynchronize(C.stream));
    });
}

int rbhip_set_positions(rbhip_context* ctx, const double* xyz) {
    return guard([&] {
        require(ctx && xyz, "null argument");
        rbhip_context& C = *ctx;
        C.setDevice();
        C.uploadRows(xyz, 3);
        HIPCHK(launch_stage_in(C.precision, C.dSys, C.dTmp, C.numAtoms, 0, C.stream, C.layout(), C.dAtomOrig,
                               C.h.stageR));
        // stateChanged(Positions): forces, then geometry + dynamics (RigidBodyIntegrator.cpp:65-68)
        C.h.useStageR = 1;
        C.uploadSys();
        C.launchForces();
        // (velocities set since the last step are staged in FP64 already)
        C.adoptedStateIn(false, !C.stageVFresh);
        C.runSetup(1);
        C.positionsDerived = false;
        C.h.useStageR = 0;
        C.uploadSys();
        HIPCHK(hipStreamSynchronize(C.stream));
    });
}

int rbhip_set_velocities(rbhip_context* ctx, const double* xyz) {
    return guard([&] {
        require(ctx && xyz, "null argument");
        rbhip_context& C = *ctx;
        C.setDevice();
        C.uploadRows(xyz, 3);
        HIPCHK(launch_stage_in(C.precision, C.dSys, C.dTmp, C.numAtoms, 1, C.stream, C.layout(), C.dAtomOrig,
                               C.h.stageV));
        C.stageVFresh = true;
        // stateChanged(Velocities): dynamics rebuild on the live body geometry
        if (C.built) {
            C.adoptedForcesIn();
            C.runSetup(0);
        }
        HIPCHK(hipStreamSynchronize(C.stream));
    });
}

int rbhip_get_positions(rbhip_context* ctx, double* xyz) {
    return guard([&] { require(ctx && xyz, "null argument"); ctx->setDevice(); ctx->download(0, xyz); });
}

int rbhip_get_velocities(rbhip_context* ctx, double* xyz) {
    return guard([&] { require(ctx && xyz, "null argument"); ctx->setDevice(); ctx->download(1, xyz); });
}

int rbhip_get_forces(rbhip_context* ctx, double* xyz) {
    return guard([&] { require(ctx && xyz, "null argument"); ctx->setDevice(); ctx->download(2, xyz); });
}

int rbhip_set_step_size(rbhip_context* ctx, double dt) {
    return guard([&] {
        require(ctx && dt > 0.0, "step size must be positive");
        rbhip_context& C = *ctx;
        C.setDevice();
        C.dt = dt;
        C.h.dt = dt;
        C.uploadSys();
        C.fixKick();
        HIPCHK(hipStreamSynchronize(C.stream));
    });
}

int rbhip_get_time(rbhip_context* ctx, double* time, long long* stepCount) {
    return guard([&] {
        require(ctx, "null context");
        if (time) *time = ctx->time;
        if (stepCount) *stepCount = ctx->stepCount;
    });
}

// nothing steps between rbhip_part1_constrained and the rbhip_part2 that adds
// its correction to the velocities
static void require_step_complete(const rbhip_context& C) {
    require(!C.h.corrPending, "a constrained part 1 must be completed by rbhip_part2 first", RBHIP_E_STATE);
}

int rbhip_part1(rbhip_context* ctx) {
    return guard([&] {
        require(ctx, "null context");
        require(ctx->built, "positions must be set before stepping", RBHIP_E_STATE);
        require_step_complete(*ctx);
        ctx->setDevice();
        ctx->fixKick();
        if (ctx->refined) ctx->launchRefineBegin(true);
        ctx->launchPart1();
        ctx->positionsDerived = true;
        ctx->kickFolded = false;
    });
}

int rbhip_free_delta(rbhip_context* ctx, double* posDelta) {
    return guard([&] {
        require(ctx && posDelta, "null argument");
        rbhip_context& C = *ctx;
        require(C.adopted, "free-atom displacements are for adopted arrays", RBHIP_E_STATE);
        require(!C.refined, "refined energies are not supported with the caller's constraints", RBHIP_E_UNSUPPORTED);
        require(C.built, "positions must be set before stepping", RBHIP_E_STATE);
        require_step_complete(C);
        require(((uintptr_t)posDelta & 15) == 0, "posDelta must be 16-byte aligned");
        C.setDevice();
        C.fixKick();
        HIPCHK(launch_free_delta(C.precision, C.dSys, C.numFree, posDelta, C.stream));
    });
}

int rbhip_part1_constrained(rbhip_context* ctx, const double* posDelta) {
    return guard([&] {
        require(ctx && posDelta, "null argument");
        rbhip_context& C = *ctx;
        require(C.adopted, "constrained displacements are for adopted arrays", RBHIP_E_STATE);
        require(!C.refined, "refined energies are not supported with the caller's constraints", RBHIP_E_UNSUPPORTED);
        require(C.built, "positions must be set before stepping", RBHIP_E_STATE);
        require_step_complete(C);
        require(((uintptr_t)posDelta & 15) == 0, "posDelta must be 16-byte aligned");
        C.setDevice();
        C.fixKick();
        if (!C.dFreeCorr) {
            C.corrArena.stream = C.stream;
            C.dFreeCorr = C.corrArena.alloc(3 * C.realSize() * (size_t)std::max(C.numFree, 1));
        }
        C.h.posDelta = const_cast<double*>(posDelta);
        C.h.freeCorr = C.dFreeCorr;
        C.uploadSys();
        try {
            C.launchPart1();
        } catch (...) {
            // the caller's buffer is not kept past this call
            C.h.posDelta = nullptr;
            C.uploadSys();
            throw;
        }
        // part 1 read posDelta; the correction waits for part 2
        C.h.posDelta = nullptr;
        C.h.corrPending = 1;
        C.uploadSys();
        C.positionsDerived = true;
        C.kickFolded = false;
    });
}

int rbhip_part2(rbhip_context* ctx) {
    return guard([&] {
        require(ctx, "null context");
        require(ctx->built, "positions must be set before stepping", RBHIP_E_STATE);
        ctx->setDevice();
        ctx->launchPart2();
        ctx->keSlots = ctx->partKeSlots();
        ctx->freeKeStale = ctx->h.corrPending != 0;
        if (ctx->h.corrPending) {
            ctx->h.corrPending = 0;
            ctx->uploadSys();
        }
        ctx->stageVFresh = false;
        ctx->kickFolded = true;
        // the step is complete (ReferenceRigidBodyKernels.cpp:106-107)
        ctx->time += ctx->dt;
        ctx->stepCount++;
        if (ctx->refined) {
            ctx->launchRefineEnd();
            ctx->refKeSlots = ctx->nPartTiles;
            ctx->refinedValid = true;
        }
    });
}

int rbhip_execute(rbhip_context* ctx, int steps) {
    return guard([&] {
        require(ctx, "null context");
        require_step_complete(*ctx);
        ctx->setDevice();
        ctx->executeSteps(steps);
    });
}

int rbhip_use_graph(rbhip_context* ctx, int enable) {
    return guard([&] {
        require(ctx, "null context");
        ctx->useGraph = enable != 0;
        if (!ctx->useGraph) ctx->graphs.reset();
    });
}

int rbhip_use_fusion(rbhip_context* ctx, int enable) {
    return guard([&] {
        require(ctx, "null context");
        ctx->useFusion = enable != 0;
    });
}

int rbhip_kinetic_energies(rbhip_context* ctx, double* out2) {
    return guard([&] {
        require(ctx && out2, "null argument");
        ctx->setDevice();
        double r[4] = {0, 0, 0, 0};
        // after a caller's velocity constraints, the free atoms' partials again
        if (ctx->freeKeStale) {
            HIPCHK(launch_free_ke(ctx->precision, ctx->dSys, ctx->dPartTiles, ctx->nPartTiles, ctx->block, ctx->stream));
        }
        if (ctx->keSlots > 0) ctx->reduce(0, r);
        out2[0] = 0.5 * r[0];
        out2[1] = 0.5 * r[1];
    });
}

int rbhip_set_refined_energies(rbhip_context* ctx, int enable) {
    return guard([&] {
        require(ctx, "null context");
        ctx->setDevice();
        ctx->setRefined(enable != 0);
        HIPCHK(hipStreamSynchronize(ctx->stream));
    });
}

int rbhip_refined_kinetic_energies(rbhip_context* ctx, double* out2) {
    return guard([&] {
        require(ctx && out2, "null argument");
        rbhip_context& C = *ctx;
        C.setDevice();
        double r[4] = {0, 0, 0, 0};
        if (!C.refined) {
            if (C.keSlots > 0) C.reduce(0, r);
            out2[0] = 0.5 * r[0];
            out2[1] = 0.5 * r[1];
            return;
        }
        if (C.refinedValid && C.refKeSlots > 0) C.reduce(2, r);
        const double f = 1.0 / (6.0 * C.dt);
        out2[0] = r[0] * f;
        out2[1] = r[1] * f;
    });
}

int rbhip_potential_energy_refinement(rbhip_context* ctx, double* energy) {
    return guard([&] {
        require(ctx && energy, "null argument");
        rbhip_context& C = *ctx;
        *energy = 0.0;
        if (!C.refined || !C.built || C.nPartTiles == 0) return;
        C.setDevice();
        HIPCHK(launch_refine_pe(C.precision, C.dSys, C.dPartTiles, C.nPartTiles, C.block, C.maxBlocks, C.stream,
                                C.layout()));
        double r[4];
        C.reduce(3, r);
        *energy = -r[0] * C.dt * C.dt / 24.0;
    });
}

int rbhip_potential_energy(rbhip_context* ctx, double* energy) {
    return guard([&] {
        require(ctx && energy, "null argument");
        ctx->setDevice();
        double r[4];
        ctx->reduce(1, r);
        *energy = r[0];
    });
}

int rbhip_get_body_states(rbhip_context* ctx, double* out) {
    return guard([&] {
        require(ctx && out, "null argument");
        rbhip_context& C = *ctx;
        require(C.built, "body state not built", RBHIP_E_STATE);
        if (C.numBodies == 0) return;
        C.setDevice();
        C.fixKick();
        C.adoptedForcesIn();
        HIPCHK(launch_body_states(C.precision, C.dSys, C.dBodyN, C.dBodyAtom0, C.numBodies, C.dTmp, C.stream));
        std::vector<double> tmp(36 * (size_t)C.numBodies);
        HIPCHK(hipMemcpyAsync(tmp.data(), C.dTmp, tmp.size() * sizeof(double), hipMemcpyDeviceToHost, C.stream));
        HIPCHK(hipStreamSynchronize(C.stream));
        for (int ib = 0; ib < C.numBodies; ib++) {
            int r = C.internalBody[ib];
            std::memcpy(out + 36 * (size_t)r, tmp.data() + 36 * (size_t)ib, 36 * sizeof(double));
            out[36 * (size_t)r + 32] = C.refBodyLoc[r];
        }
    });
}

int rbhip_get_body_fixed_positions(rbhip_context* ctx, double* xyz) {
    return guard([&] {
        require(ctx && xyz, "null argument");
        rbhip_context& C = *ctx;
        if (C.numBodyAtoms == 0) return;
        C.setDevice();
        const size_t ds = C.h.dStride, nba = C.numBodyAtoms;
        std::vector<double> soa(3 * ds), d(3 * nba);
        HIPCHK(hipStreamSynchronize(C.stream));
        if (C.precision == RBHIP_PRECISION_MIXED) {
            std::vector<float> f(3 * ds);
            HIPCHK(hipMemcpyAsync(f.data(), C.h.d, 3 * ds * 4, hipMemcpyDeviceToHost, C.stream));
            HIPCHK(hipStreamSynchronize(C.stream));
            for (size_t i = 0; i < 3 * ds; i++) soa[i] = f[i];
        } else {
            HIPCHK(hipMemcpyAsync(soa.data(), C.h.d, 3 * ds * 8, hipMemcpyDeviceToHost, C.stream));
            HIPCHK(hipStreamSynchronize(C.stream));
        }
        for (size_t i = 0; i < nba; i++)
            for (int c = 0; c < 3; c++) d[3 * i + c] = soa[c * ds + i];
        for (int ib = 0; ib < C.numBodies; ib++) {
            int r = C.internalBody[ib];
            size_t src = 3 * (size_t)(C.bodyAtom0[ib] - C.numFree), dst = 3 * (size_t)C.refBodyLoc[r];
            std::memcpy(xyz + dst, d.data() + src, 3 * C.bodyN[ib] * sizeof(double));
        }
    });
}

int rbhip_set_launch(rbhip_context* ctx, int blockSize, int maxBlocks) {
    return guard([&] {
        require(ctx, "null context");
        require(blockSize >= 64 && blockSize <= ctx->maxBlock() && blockSize % 64 == 0,
                "block size must be a multiple of 64 in [64, " + std::to_string(ctx->maxBlock()) + "]");
        require(maxBlocks >= 0, "negative grid cap");
        for (auto& kv : ctx->lanes) require(blockSize % kv.second == 0, "block size must be a multiple of the lanes per body");
        ctx->setDevice();
        HIPCHK(hipStreamSynchronize(ctx->stream));
        int oldBlock = ctx->block, oldMax = ctx->maxBlocks;
        ctx->block = blockSize;
        ctx->maxBlocks = maxBlocks;
        try {
            ctx->rebuildTiles();
        } catch (...) {
            ctx->block = oldBlock;
            ctx->maxBlocks = oldMax;
            ctx->rebuildTiles();
            throw;
        }
    });
}

int rbhip_set_lanes(rbhip_context* ctx, int atomsPerBody, int lanes) {
    return guard([&] {
        require(ctx, "null context");
        require(lanes >= 1 && lanes <= 64 && (lanes & (lanes - 1)) == 0, "lanes per body must be a power of two in [1, 64]");
        require(ctx->block % lanes == 0, "lanes per body must divide the block size");
        ctx->setDevice();
        HIPCHK(hipStreamSynchronize(ctx->stream));
        ctx->lanes[atomsPerBody] = lanes;
        ctx->rebuildTiles();
    });
}

int rbhip_set_shared_frames(rbhip_context* ctx, int enable) {
    return guard([&] {
        require(ctx, "null context");
        ctx->shareFramesOn = enable != 0;
    });
}

int rbhip_describe(rbhip_context* ctx, char* buf, int len) {
    return guard([&] {
        require(ctx && buf && len > 0, "null argument");
        std::string s = "block=" + std::to_string(ctx->block) + " maxBlocks=" + std::to_string(ctx->maxBlocks) +
                        " partTiles=" + std::to_string(ctx->nPartTiles) + " classes=[";
        for (auto& c : ctx->classes) {
            auto it = ctx->lanes.find(c.N);
            s += "N" + std::to_string(c.N) + "x" + std::to_string(c.count) + ":L" +
                 std::to_string(it == ctx->lanes.end() ? 1 : it->second) + " ";
        }
        s += "] free=" + std::to_string(ctx->numFree);
        std::string shapes;
        for (auto& c : ctx->classes)
            if (c.shapes > 0) shapes += " N" + std::to_string(c.N) + ":" + std::to_string(c.shapes);
        if (!shapes.empty()) s += " sharedShapes=[" + shapes.substr(1) + "]";
        std::snprintf(buf, len, "%s", s.c_str());
    });
}


int rbhip_time_kernels(rbhip_context* ctx, int steps, double* avgMs, double* totalMs) {
    return guard([&] {
        require(ctx && avgMs && steps > 0, "bad arguments");
        rbhip_context& C = *ctx;
        require(C.built, "positions must be set before stepping", RBHIP_E_STATE);
        require_step_complete(C);
        C.setDevice();
        C.fixKick();
        time_steps(C.stream, steps, [&] { C.launchPart1(); }, [&] { C.launchForces(false); }, [&] { C.launchPart2(); },
                   avgMs, totalMs);
        // the fused path of step(steps + 1): part1, forces, steps x [fused, forces], part2
        C.launchPart1();
        C.launchForces();
        double fusedMs[2];
        time_phases(C.stream, steps, {[&] { C.launchFused(); }, [&] { C.launchForces(false); }}, fusedMs, nullptr);
        C.launchForces(true);   // (the same forces again, with the energy)
        C.launchPart2();
        avgMs[RBHIP_KERNEL_FUSED] = fusedMs[0];
        // the state advanced 2 steps + 1 as rbhip_execute leaves it (the
        // refined-energy accumulators do not describe it)
        C.keSlots = C.partKeSlots();
        C.freeKeStale = false;
        C.stageVFresh = false;
        C.positionsDerived = true;
        C.kickFolded = true;
        C.refinedValid = false;
        C.refKeSlots = 0;
        C.time += C.dt * (2 * steps + 1);
        C.stepCount += 2 * steps + 1;
    });
}

int rbhip_autotune(rbhip_context* ctx, int reps) {
    return guard([&] {
        require(ctx && reps > 0, "bad arguments");
        rbhip_context& C = *ctx;
        require(C.built, "positions must be set before tuning", RBHIP_E_STATE);
        require_step_complete(C);
        C.setDevice();
        C.fixKick();
        HIPCHK(hipStreamSynchronize(C.stream));
        rbhip_context::Snapshot snap;
        C.takeSnapshot(snap);
        auto restore = [&] {
            C.restoreSnapshot(snap, C.stream);
            HIPCHK(hipStreamSynchronize(C.stream));
        };
        // the time of one step as rbhip_execute runs it: fused pass + forces
        // with fusion, part1 + forces + part2 without
        auto measure = [&]() -> double {
            double avg[3], total;
            double t;
            if (C.useFusion) {
                C.launchPart1(); C.launchForces();
                C.launchFused(); C.launchForces();
                time_phases(C.stream, reps, {[&] { C.launchFused(); }, [&] { C.launchForces(false); }}, avg, &total);
                C.launchPart2();
                t = avg[0] + avg[1];
            } else {
                C.launchPart1(); C.launchForces(); C.launchPart2();
                time_steps(C.stream, reps, [&] { C.launchPart1(); }, [&] { C.launchForces(false); },
                           [&] { C.launchPart2(); }, avg, &total);
                t = avg[0] + avg[1] + avg[2];
            }
            restore();
            return t;
        };
        try {
            const int blocks[] = {64, 128, 256, 512};
            int bestBlock = C.block;
            double best = 1e300;
            std::map<int, int> savedLanes = C.lanes;
            for (int b : blocks) {
                bool ok = b <= C.maxBlock();
                for (auto& kv : C.lanes) ok = ok && b % kv.second == 0;
                if (!ok) continue;
                C.block = b;
                C.rebuildTiles();
                double t = measure();
                if (t < best) { best = t; bestBlock = b; }
            }
            C.block = bestBlock;
            C.rebuildTiles();
            for (auto& cl : C.classes) {
                int bestL = savedLanes.count(cl.N) ? savedLanes[cl.N] : 1;
                double bestT = 1e300;
                int maxL = std::min(16, pow2_at_least(cl.N));
                for (int L = 1; L <= maxL; L <<= 1) {
                    if (C.block % L) continue;
                    C.lanes[cl.N] = L;
                    C.rebuildTiles();
                    double t = measure();
                    if (t < bestT) { bestT = t; bestL = L; }
                }
                C.lanes[cl.N] = bestL;
            }
            C.rebuildTiles();
        } catch (...) {
            restore();
            throw;
        }
    });
}

int rbhip_copy_calibration(size_t bytes, int reps, double* avgMs) {
    return guard([&] {
        require(bytes >= 16 && reps > 0 && avgMs, "bad arguments");
        DeviceArena a;
        void* src = a.alloc(bytes);
        void* dst = a.alloc(bytes);
        HIPCHK(hipMemset(src, 1, bytes));
        hipStream_t s;
        HIPCHK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
        EventSet es(2 * (size_t)reps);
        double acc = 0.0;
        try {
            for (int r = 0; r < reps; r++) {
                HIPCHK(hipEventRecord(es.ev[2 * r], s));
                HIPCHK(launch_copy(src, dst, bytes, s));
                HIPCHK(hipEventRecord(es.ev[2 * r + 1], s));
            }
            HIPCHK(hipStreamSynchronize(s));
            for (int r = 0; r < reps; r++) {
                float ms;
                HIPCHK(hipEventElapsedTime(&ms, es.ev[2 * r], es.ev[2 * r + 1]));
                acc += ms;
            }
        } catch (...) {
            (void)hipStreamDestroy(s);
            a.release();
            throw;
        }
        (void)hipStreamDestroy(s);
        a.release();
        *avgMs = acc / reps;
    });
}

// ---------------------------------------------------------------------------
int rbhip_batch_create(int device, rbhip_batch** out) {
    return guard([&] {
        require(out, "null output handle");
        int ndev = 0;
        HIPCHK(hipGetDeviceCount(&ndev));
        require(device >= 0 && device < ndev, "no such device");
        HIPCHK(hipSetDevice(device));
        std::unique_ptr<rbhip_batch> b(new rbhip_batch());
        b->device = device;
        b->stream = pooled_stream(device);
        b->arena.stream = b->stream;
        HIPCHK(hipEventCreateWithFlags(&b->ordering, hipEventDisableTiming));
        *out = b.release();
    });
}

int rbhip_batch_destroy(rbhip_batch* batch) {
    return guard([&] {
        if (!batch) return;
        (void)hipSetDevice(batch->device);
        (void)hipStreamSynchronize(batch->stream);
        delete batch;
    });
}

int rbhip_batch_set_members(rbhip_batch* batch, int count, rbhip_context* const* members) {
    return guard([&] {
        require(batch && count >= 0 && (count == 0 || members), "bad arguments");
        HIPCHK(hipSetDevice(batch->device));
        HIPCHK(hipStreamSynchronize(batch->stream));
        std::vector<rbhip_context*> m(members, members + count);
        for (size_t i = 0; i < m.size(); i++) {
            require(m[i] != nullptr, "null member");
            require(m[i]->device == batch->device, "member on another device");
            require(m[i]->built, "member body state not built", RBHIP_E_STATE);
            require(m[i]->precision == m[0]->precision, "members must share the precision");
            require(m[i]->numConstraints == 0, "members with constraints cannot be batched", RBHIP_E_UNSUPPORTED);
            require(m[i]->numSites == 0, "members with virtual sites cannot be batched", RBHIP_E_UNSUPPORTED);
            require(!m[i]->adopted, "members with adopted atom arrays cannot be batched", RBHIP_E_UNSUPPORTED);
            require((m[i]->rotationMode == 0) == (m[0]->rotationMode == 0), "members must share the rotation kind");
            for (auto& kv : m[i]->lanes)
                require(batch->block % kv.second == 0, "member lanes per body must divide the batch block size");
            if (m[i]->kickDt != m[i]->dt) {
                HIPCHK(hipStreamSynchronize(m[i]->stream));
                m[i]->fixKick();
                HIPCHK(hipStreamSynchronize(m[i]->stream));
            }
        }
        if (!m.empty()) {
            batch->precision = m[0]->precision;
            batch->rotationMode = m[0]->rotationMode;
            if (batch->rotationMode == 0 && batch->block > 256) batch->block = 256;
        }
        batch->members = m;
        batch->rebuild();
    });
}

int rbhip_batch_set_stream(rbhip_batch* batch, rbhip_stream stream) {
    return guard([&] {
        require(batch, "null batch");
        HIPCHK(hipSetDevice(batch->device));
        HIPCHK(hipStreamSynchronize(batch->stream));
        batch->stream = stream ? reinterpret_cast<hipStream_t>(stream) : pooled_stream(batch->device);
        batch->arena.stream = batch->stream;
        batch->graphs.reset();
    });
}

int rbhip_batch_set_launch(rbhip_batch* batch, int blockSize, int maxBlocks) {
    return guard([&] {
        require(batch, "null batch");
        int cap = batch->rotationMode == 0 && !batch->members.empty() ? 256 : 512;
        require(blockSize >= 64 && blockSize <= cap && blockSize % 64 == 0,
                "block size must be a multiple of 64 in [64, " + std::to_string(cap) + "]");
        require(maxBlocks >= 0, "negative grid cap");
        for (rbhip_context* c : batch->members)
            for (auto& kv : c->lanes)
                require(blockSize % kv.second == 0, "member lanes per body must divide the batch block size");
        HIPCHK(hipSetDevice(batch->device));
        HIPCHK(hipStreamSynchronize(batch->stream));
        batch->block = blockSize;
        batch->maxBlocks = maxBlocks;
        batch->rebuild();
    });
}

int rbhip_batch_use_graph(rbhip_batch* batch, int enable) {
    return guard([&] {
        require(batch, "null batch");
        batch->useGraph = enable != 0;
        if (!batch->useGraph) batch->graphs.reset();
    });
}

int rbhip_batch_use_fusion(rbhip_batch* batch, int enable) {
    return guard([&] {
        require(batch, "null batch");
        batch->useFusion = enable != 0;
    });
}

int rbhip_batch_execute(rbhip_batch* batch, int steps) {
    return guard([&] {
        require(batch && steps >= 0, "bad arguments");
        if (batch->members.empty() || steps == 0) return;
        HIPCHK(hipSetDevice(batch->device));
        for (rbhip_context* c : batch->members)
            require(c->kickDt == c->dt, "member step size changed since it joined the batch", RBHIP_E_STATE);
        batch->joinMembers();
        batch->refreshSystems();
        batch->executeSteps(steps);
        batch->releaseMembers();
        batch->finishSteps(steps);
    });
}

int rbhip_batch_kinetic_energies(rbhip_batch* batch, double* out) {
    return guard([&] {
        require(batch && out, "null argument");
        size_t n = batch->members.size();
        if (n == 0) return;
        HIPCHK(hipSetDevice(batch->device));
        batch->joinMembers();
        std::vector<int> slots(n);
        for (size_t m = 0; m < n; m++) slots[m] = batch->members[m]->keSlots;
        HIPCHK(hipMemcpyAsync(batch->dKeSlots, slots.data(), n * sizeof(int), hipMemcpyHostToDevice, batch->stream));
        HIPCHK(launch_reduce(batch->dSystems, batch->dKeSlots, (int)n, 0, batch->dRed, batch->stream));
        std::vector<double> r(4 * n);
        HIPCHK(hipMemcpyAsync(r.data(), batch->dRed, r.size() * sizeof(double), hipMemcpyDeviceToHost, batch->stream));
        HIPCHK(hipStreamSynchronize(batch->stream));
        for (size_t m = 0; m < n; m++) {
            out[2 * m] = 0.5 * r[4 * m];
            out[2 * m + 1] = 0.5 * r[4 * m + 1];
        }
    });
}

int rbhip_batch_time_kernels(rbhip_batch* batch, int steps, double* avgMs, double* totalMs) {
    return guard([&] {
        require(batch && avgMs && steps > 0, "bad arguments");
        require(!batch->members.empty(), "empty batch", RBHIP_E_STATE);
        HIPCHK(hipSetDevice(batch->device));
        rbhip_batch& B = *batch;
        for (rbhip_context* c : B.members)
            require(c->kickDt == c->dt, "member step size changed since it joined the batch", RBHIP_E_STATE);
        B.joinMembers();
        B.refreshSystems();
        time_steps(B.stream, steps, [&] { B.launchPart1(); }, [&] { B.launchForces(false); }, [&] { B.launchPart2(); },
                   avgMs, totalMs);
        B.launchPart1();
        B.launchForces();
        double fusedMs[2];
        time_phases(B.stream, steps, {[&] { B.launchFused(); }, [&] { B.launchForces(false); }}, fusedMs, nullptr);
        B.launchForces(true);   // (the same forces again, with the energy)
        B.launchPart2();
        avgMs[RBHIP_KERNEL_FUSED] = fusedMs[0];
        B.releaseMembers();
        B.finishSteps(2 * steps + 1);
    });
}

int rbhip_batch_autotune(rbhip_batch* batch, int reps) {
    return guard([&] {
        require(batch && reps > 0, "bad arguments");
        rbhip_batch& B = *batch;
        require(!B.members.empty(), "empty batch", RBHIP_E_STATE);
        HIPCHK(hipSetDevice(B.device));
        for (rbhip_context* c : B.members) {
            require(c->kickDt == c->dt, "member step size changed since it joined the batch", RBHIP_E_STATE);
            HIPCHK(hipStreamSynchronize(c->stream));
        }
        std::vector<std::unique_ptr<rbhip_context::Snapshot>> snaps;
        for (rbhip_context* c : B.members) {
            snaps.emplace_back(new rbhip_context::Snapshot());
            c->takeSnapshot(*snaps.back());
            HIPCHK(hipStreamSynchronize(c->stream));
        }
        B.joinMembers();
        B.refreshSystems();
        auto restore = [&] {
            for (size_t m = 0; m < B.members.size(); m++) B.members[m]->restoreSnapshot(*snaps[m], B.stream);
            HIPCHK(hipStreamSynchronize(B.stream));
        };
        // time of one step as rbhip_batch_execute runs it
        auto measure = [&]() -> double {
            double avg[3], total, t;
            if (B.useFusion) {
                B.launchPart1(); B.launchForces();
                B.launchFused(); B.launchForces();
                time_phases(B.stream, reps, {[&] { B.launchFused(); }, [&] { B.launchForces(false); }}, avg, &total);
                B.launchPart2();
                t = avg[0] + avg[1];
            } else {
                B.launchPart1(); B.launchForces(); B.launchPart2();
                time_steps(B.stream, reps, [&] { B.launchPart1(); }, [&] { B.launchForces(false); },
                           [&] { B.launchPart2(); }, avg, &total);
                t = avg[0] + avg[1] + avg[2];
            }
            restore();
            return t;
        };
        const int cap = B.rotationMode == 0 ? 256 : 512;
        auto lanesFit = [&](int blk) {
            for (rbhip_context* c : B.members)
                for (auto& kv : c->lanes)
                    if (blk % kv.second || c->block % kv.second) return false;
            return true;
        };
        try {
            // workgroup size and grid cap (0 = one workgroup per tile; a cap
            // makes each workgroup loop over tiles, a persistent grid)
            int bestBlock = B.block, bestGrid = B.maxBlocks;
            double best = 1e300;
            for (int blk : {64, 128, 256, 512}) {
                if (blk > cap || !lanesFit(blk)) continue;
                for (int grid : {0, 2048, 4096, 8192, 16384, 32768}) {
                    B.block = blk;
                    B.maxBlocks = grid;
                    B.rebuild();
                    double t = measure();
                    if (t < best) { best = t; bestBlock = blk; bestGrid = grid; }
                }
            }
            B.block = bestBlock;
            B.maxBlocks = bestGrid;
            // lanes per body for each body size present in the batch
            std::map<int, int> sizes;
            for (rbhip_context* c : B.members)
                for (auto& cl : c->classes) sizes[cl.N] = c->lanes.count(cl.N) ? c->lanes[cl.N] : 1;
            for (auto& kv : sizes) {
                const int N = kv.first;
                int bestL = kv.second;
                double bestT = 1e300;
                for (int L = 1; L <= std::min(16, pow2_at_least(N)); L <<= 1) {
                    bool ok = B.block % L == 0;
                    for (rbhip_context* c : B.members) ok = ok && c->block % L == 0;
                    if (!ok) continue;
                    for (rbhip_context* c : B.members)
                        if (c->lanes.count(N) || std::any_of(c->classes.begin(), c->classes.end(),
                                                             [&](const rbhip_context::Class& cl) { return cl.N == N; }))
                            c->lanes[N] = L;
                    B.rebuild();
                    double t = measure();
                    if (t < bestT) { bestT = t; bestL = L; }
                }
                for (rbhip_context* c : B.members)
                    if (c->lanes.count(N)) c->lanes[N] = bestL;
            }
            for (rbhip_context* c : B.members) c->rebuildTiles();
            B.rebuild();
        } catch (...) {
            restore();
            B.releaseMembers();
            throw;
        }
        B.releaseMembers();
    });
}

int rbhip_batch_describe(rbhip_batch* batch, char* buf, int len) {
    return guard([&] {
        require(batch && buf && len > 0, "bad arguments");
        std::map<int, int> lanes;
        long long bodies = 0, free = 0;
        for (rbhip_context* c : batch->members) {
            for (auto& kv : c->lanes) lanes[kv.first] = kv.second;
            bodies += c->numBodies;
            free += c->numFree;
        }
        std::string s = "members=" + std::to_string(batch->members.size()) + " bodies=" + std::to_string(bodies) +
                        " free=" + std::to_string(free) + " block=" + std::to_string(batch->block) +
                        " maxBlocks=" + std::to_string(batch->maxBlocks) + " partTiles=" + std::to_string(batch->nPart) +
                        " forceTiles=" + std::to_string(batch->nForce) + " lanes=[";
        for (auto& kv : lanes) s += "N" + std::to_string(kv.first) + ":L" + std::to_string(kv.second) + " ";
        s += "]";
        std::snprintf(buf, (size_t)len, "%s", s.c_str());
    });
}

}  // extern "C"
