// Device-side bookkeeping of the rigid-body system on labels and masses in
// device memory (the integer contract of the reference, bit for bit):
//   cleanBodyIndices            openmmapi/src/RigidBodySystem.cpp:28-49
//   RigidBodySystem::initialize openmmapi/src/RigidBodySystem.cpp:55-105
//                               (atomIndex, and N and loc of each body)
// Every atom gets a 32-bit key: 0 for a free atom (label <= 0), its label
// for a body atom, 0xFFFFFFFF for a massless particle (neither free nor a
// body atom).  A stable radix sort of (key, atom) pairs puts the free atoms
// first in ascending order, then each body's atoms in ascending order, bodies
// by ascending label: exactly the reference's atomIndex.  Heads of equal-key
// runs, scanned, give the body numbers 1..B (cleanBodyIndices' ranks), and
// the heads' positions give each body's first atom (loc).  The sort and the
// scan are rocPRIM's device algorithms.
#include <cstring>

#include <hip/hip_runtime.h>
#include <rocprim/rocprim.hpp>

#include "rbhip_bookkeeping.h"

namespace rbhip {

namespace {

constexpr unsigned kFreeKey = 0u;
constexpr unsigned kMasslessKey = 0xFFFFFFFFu;
constexpr size_t kAlign = 256;

size_t aligned(size_t b) { return (b + kAlign - 1) / kAlign * kAlign; }

// status flags and counts, zeroed by the first kernel of every call:
// small[0] flags (1: a massless particle has a body label, 2: a mass is
// negative or not a number), small[1..4] numFree, numBodies, numMassless,
// numBodyAtoms
__global__ void k_bk_clear(int* small) {
    if (threadIdx.x < 8) small[threadIdx.x] = 0;
}

__global__ void k_bk_keys(int n, const int* __restrict__ labels, const double* __restrict__ masses,
                          unsigned* __restrict__ keys, int* __restrict__ vals, int* __restrict__ small) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int l = labels[i];
    const double m = masses[i];
    unsigned k;
    if (m == 0.0) {
        k = kMasslessKey;
        if (l > 0) atomicOr(small, 1);
    } else if (!(m > 0.0)) {
        k = kMasslessKey;
        atomicOr(small, 2);
    } else {
        k = l > 0 ? (unsigned)l : kFreeKey;
    }
    keys[i] = k;
    vals[i] = i;
}

// head[j] = 1 where a body's run of sorted keys starts
__global__ void k_bk_heads(int n, const unsigned* __restrict__ keys, int* __restrict__ head) {
    int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    const unsigned k = keys[j];
    const bool body = k != kFreeKey && k != kMasslessKey;
    head[j] = body && (j == 0 || keys[j - 1] != k) ? 1 : 0;
}

// first index j of the sorted keys with keys[j] >= v
__device__ int lower_bound(const unsigned* keys, int n, unsigned v) {
    int lo = 0, hi = n;
    while (lo < hi) {
        int mid = lo + (hi - lo) / 2;
        if (keys[mid] < v) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

__global__ void k_bk_counts(int n, const unsigned* __restrict__ keys, const int* __restrict__ rank,
                            int* __restrict__ small) {
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    const int numFree = lower_bound(keys, n, 1u);
    const int bodyEnd = lower_bound(keys, n, kMasslessKey);
    small[1] = numFree;
    small[2] = bodyEnd > numFree ? rank[bodyEnd - 1] : 0;
    small[3] = n - bodyEnd;
    small[4] = bodyEnd - numFree;
}

__global__ void k_bk_out(int n, const int* __restrict__ vals, const int* __restrict__ head,
                         const int* __restrict__ rank, const int* __restrict__ small, int* __restrict__ bodyIndex,
                         int* __restrict__ atomIndex, int* __restrict__ bodyAtomStart) {
    int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    const int numFree = small[1], numBodies = small[2], numBodyAtoms = small[4];
    const int bodyEnd = numFree + numBodyAtoms;
    const int atom = vals[j];
    const bool body = j >= numFree && j < bodyEnd;
    bodyIndex[atom] = body ? rank[j] : 0;
    if (j < bodyEnd) atomIndex[j] = atom;
    if (body && head[j]) bodyAtomStart[rank[j] - 1] = j - numFree;
    if (j == 0) bodyAtomStart[numBodies] = numBodyAtoms;
}

struct Layout {
    size_t keys, keysSorted, vals, valsSorted, head, rank, small, temp, total, tempBytes;
};

// scratch regions for n atoms (256-byte aligned); the rocPRIM temporary
// storage is shared by the sort and the scan, which run one after the other
hipError_t layout_for(int n, hipStream_t stream, Layout& L) {
    size_t sortBytes = 0, scanBytes = 0;
    hipError_t e = rocprim::radix_sort_pairs(nullptr, sortBytes, (unsigned*)nullptr, (unsigned*)nullptr,
                                             (int*)nullptr, (int*)nullptr, (size_t)n, 0, 32, stream);
    if (e != hipSuccess) return e;
    e = rocprim::inclusive_scan(nullptr, scanBytes, (int*)nullptr, (int*)nullptr, (size_t)n, rocprim::plus<int>(),
                                stream);
    if (e != hipSuccess) return e;
    const size_t a = aligned(sizeof(int) * (size_t)n);
    size_t off = 0;
    L.keys = off; off += a;
    L.keysSorted = off; off += a;
    L.vals = off; off += a;
    L.valsSorted = off; off += a;
    L.head = off; off += a;
    L.rank = off; off += a;
    L.small = off; off += kAlign;
    L.temp = off;
    L.tempBytes = aligned(sortBytes > scanBytes ? sortBytes : scanBytes);
    off += L.tempBytes;
    L.total = off;
    return hipSuccess;
}

}  // namespace

hipError_t bookkeeping_scratch_bytes(int numAtoms, hipStream_t stream, size_t* bytes) {
    Layout L;
    hipError_t e = layout_for(numAtoms, stream, L);
    *bytes = L.total;
    return e;
}

hipError_t launch_bookkeeping(int n, const int* labels, const double* masses, int* bodyIndex, int* atomIndex,
                              int* bodyAtomStart, void* scratch, size_t scratchBytes, hipStream_t stream,
                              int* smallOut) {
    Layout L;
    hipError_t e = layout_for(n, stream, L);
    if (e != hipSuccess) return e;
    if (L.total > scratchBytes) return hipErrorInvalidValue;
    char* s = static_cast<char*>(scratch);
    unsigned* keys = reinterpret_cast<unsigned*>(s + L.keys);
    unsigned* keysSorted = reinterpret_cast<unsigned*>(s + L.keysSorted);
    int* vals = reinterpret_cast<int*>(s + L.vals);
    int* valsSorted = reinterpret_cast<int*>(s + L.valsSorted);
    int* head = reinterpret_cast<int*>(s + L.head);
    int* rank = reinterpret_cast<int*>(s + L.rank);
    int* small = reinterpret_cast<int*>(s + L.small);
    void* temp = s + L.temp;
    const int grid = (n + 255) / 256;
    k_bk_clear<<<1, 64, 0, stream>>>(small);
    k_bk_keys<<<grid, 256, 0, stream>>>(n, labels, masses, keys, vals, small);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    size_t tb = L.tempBytes;
    e = rocprim::radix_sort_pairs(temp, tb, keys, keysSorted, vals, valsSorted, (size_t)n, 0, 32, stream);
    if (e != hipSuccess) return e;
    k_bk_heads<<<grid, 256, 0, stream>>>(n, keysSorted, head);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    tb = L.tempBytes;
    e = rocprim::inclusive_scan(temp, tb, head, rank, (size_t)n, rocprim::plus<int>(), stream);
    if (e != hipSuccess) return e;
    k_bk_counts<<<1, 64, 0, stream>>>(n, keysSorted, rank, small);
    k_bk_out<<<grid, 256, 0, stream>>>(n, valsSorted, head, rank, small, bodyIndex, atomIndex, bodyAtomStart);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    e = hipMemcpyAsync(smallOut, small, 5 * sizeof(int), hipMemcpyDeviceToHost, stream);
    if (e != hipSuccess) return e;
    return hipStreamSynchronize(stream);
}

namespace {

__global__ void k_order_free(int numFree, int numMassless, int massStart, const int* __restrict__ atomIndex,
                             const int* __restrict__ freeOrder, const int* __restrict__ massless,
                             int* __restrict__ atomOrig) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < numFree) atomOrig[i] = freeOrder ? freeOrder[i] : atomIndex[i];
    else if (i - numFree < numMassless) atomOrig[massStart + i - numFree] = massless[i - numFree];
}

__global__ void k_order_bodies(int numBodies, int numFree, const int* __restrict__ atomIndex,
                               const int* __restrict__ bodyAtomStart, const int* __restrict__ order,
                               const int* __restrict__ bodyN, const int* __restrict__ bodyAtom0,
                               int* __restrict__ atomOrig) {
    int ib = blockIdx.x * blockDim.x + threadIdx.x;
    if (ib >= numBodies) return;
    const int* src = atomIndex + numFree + bodyAtomStart[order[ib]];
    int* dst = atomOrig + bodyAtom0[ib];
    for (int j = 0, N = bodyN[ib]; j < N; j++) dst[j] = src[j];
}

template <class Real>
__global__ void k_masses_in(int n, int numFree, const int* __restrict__ atomOrig, const double* __restrict__ raw,
                            double* __restrict__ mass, Real* __restrict__ freeInvM) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double m = raw[atomOrig[i]];
    mass[i] = m;
    if (i < numFree) freeInvM[i] = (Real)(1.0 / m);
}

}  // namespace

hipError_t launch_internal_order(int numFree, int numBodies, int numMassless, int massStart, const int* atomIndex,
                                 const int* bodyAtomStart, const int* freeOrder, const int* order, const int* bodyN,
                                 const int* bodyAtom0, const int* massless, int* atomOrig, hipStream_t stream) {
    const int nf = numFree + numMassless;
    if (nf > 0) k_order_free<<<(nf + 255) / 256, 256, 0, stream>>>(numFree, numMassless, massStart, atomIndex,
                                                                     freeOrder, massless, atomOrig);
    if (numBodies > 0)
        k_order_bodies<<<(numBodies + 255) / 256, 256, 0, stream>>>(numBodies, numFree, atomIndex, bodyAtomStart,
                                                                     order, bodyN, bodyAtom0, atomOrig);
    return hipGetLastError();
}

hipError_t launch_masses_in(int precision, int n, int numFree, const int* atomOrig, const double* raw, double* mass,
                            void* freeInvM, hipStream_t stream) {
    const int grid = (n + 255) / 256;
    if (precision == 0)
        k_masses_in<float><<<grid, 256, 0, stream>>>(n, numFree, atomOrig, raw, mass, static_cast<float*>(freeInvM));
    else
        k_masses_in<double><<<grid, 256, 0, stream>>>(n, numFree, atomOrig, raw, mass, static_cast<double*>(freeInvM));
    return hipGetLastError();
}

}  // namespace rbhip
