// Launching the step's kernels on one environment or on an ensemble.
//
// Every kernel of the step is a body: a struct K with
//   static __device__ void run(VB vb, Args...)
// where vb is the block's tile index in the launch and the number of tiles.
// The same body runs in two ways:
//   * k_solo<K>: one launch of one environment, one block per tile
//     (vb = {blockIdx.x, gridDim.x});
//   * k_batch<K>: one launch for every member of an ensemble. The tiles of
//     all members form a work list (member m owns the work indices
//     [start[m], start[m] + nb[m]), its range padded to a multiple of 8 so
//     that a tile lands on the same XCD as in the member's own launch), each
//     member's arguments sit in a device table, and a persistent grid strides
//     over the list. A tile computes exactly what it computes in the
//     member's own launch (same tile, same arguments, the member's own
//     reduction partials in its own fixed order), so a member's results do
//     not depend on the other members.
// klaunch issues the plain launch, or, while an ensemble step is being
// recorded, appends the launch and its argument bytes to the recorder.
#pragma once
#include <hip/hip_runtime.h>
#include <cstring>
#include <type_traits>
#include <utility>
#include "skt_internal.h"

namespace skt {

// The arguments of one launch as one trivially copyable record.
template <class... A>
struct Pack;
template <>
struct Pack<> {};
template <class H, class... T>
struct Pack<H, T...> {
    H h;
    Pack<T...> t;
};
template <size_t I, class H, class... T>
__host__ __device__ __forceinline__ const auto& pget(const Pack<H, T...>& p) {
    if constexpr (I == 0) return p.h;
    else return pget<I - 1>(p.t);
}
inline void pfill(Pack<>&) {}
template <class H, class... T>
inline void pfill(Pack<H, T...>& p, const H& h, const T&... t) {
    p.h = h;
    pfill(p.t, t...);
}

// Argument I of a body's argument list (for idle predicates).
template <size_t I, class H, class... T>
__host__ __device__ __forceinline__ const auto& nthArg(const H& h, const T&... t) {
    if constexpr (I == 0) return h;
    else return nthArg<I - 1>(t...);
}

// A body may declare  template <class... T> static bool idle(const T&...)
// over its run arguments (without vb): true when every tile of the launch
// except tile 0 returns without effect (a finished solve, an advection pass
// past the member's substep count). The ensemble launches then skip a
// member's other tiles without running them.
template <class K, class P, class = void>
struct HasIdle : std::false_type {};
template <class K, class... A>
struct HasIdle<K, Pack<A...>, std::void_t<decltype(K::idle(std::declval<const A&>()...))>> : std::true_type {};
template <class K, class... A, size_t... I>
__device__ __forceinline__ bool idlePacked(const Pack<A...>& p, std::index_sequence<I...>) {
    return K::idle(pget<I>(p)...);
}
template <class K, class... A>
__device__ __forceinline__ bool memberIdle(const Pack<A...>& p, unsigned local) {
    if constexpr (HasIdle<K, Pack<A...>>::value) return local != 0 && idlePacked<K>(p, std::index_sequence_for<A...>{});
    else return false;
}

template <class K, class... A>
__global__ __launch_bounds__(256) void k_solo(A... a) {
    K::run(VB{blockIdx.x, gridDim.x}, a...);
}

template <class K, class... A, size_t... I>
__device__ __forceinline__ void runPacked(VB vb, const Pack<A...>& p, std::index_sequence<I...>) {
    K::run(vb, pget<I>(p)...);
}

// Persistent ensemble launch: block b takes work items b, b + gridDim.x, ...
// The member of an item is found by a binary search of the (uniform) start
// table. Every body returns uniformly across its block before any barrier
// it contains (or contains none after a per-thread return), and the barrier
// at the end of each item keeps a block's next tile from overwriting LDS a
// wave of the previous one still reads.
// An idle member's remaining items are skipped in one move: to the block's
// first work index past the member's range (no tile of it ran, so no
// barrier is needed).
template <class K, class... A>
__global__ __launch_bounds__(256) void k_batch(const Pack<A...>* __restrict__ args, WorkMap wm) {
    for (unsigned w = blockIdx.x; w < wm.total;) {
        int lo = 0, hi = wm.M - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (wm.start[mid] <= w) lo = mid;
            else hi = mid - 1;
        }
        // the member index is uniform over the block: say so, so that its
        // argument record is read with scalar loads
        lo = __builtin_amdgcn_readfirstlane(lo);
        const unsigned local = w - wm.start[lo];
        const unsigned nb = wm.nb[lo];
        if (local < nb) {
            if (memberIdle<K>(args[lo], local)) {
                const unsigned end = wm.start[lo] + nb, g = gridDim.x;
                w += (end - w + g - 1) / g * g;
                continue;
            }
            runPacked<K>(VB{local, nb}, args[lo], std::index_sequence_for<A...>{});
        }
        __syncthreads();
        w += gridDim.x;
    }
}

// The same with one work item per block (a grid of wm.total blocks).
template <class K, class... A>
__global__ __launch_bounds__(256) void k_batch1(const Pack<A...>* __restrict__ args, WorkMap wm) {
    const unsigned w = blockIdx.x;
    int lo = 0, hi = wm.M - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (wm.start[mid] <= w) lo = mid;
        else hi = mid - 1;
    }
    lo = __builtin_amdgcn_readfirstlane(lo);
    const unsigned local = w - wm.start[lo];
    const unsigned nb = wm.nb[lo];
    if (local < nb && !memberIdle<K>(args[lo], local))
        runPacked<K>(VB{local, nb}, args[lo], std::index_sequence_for<A...>{});
}

template <class K, class... A>
hipError_t batchLaunch(const void* dargs, const WorkMap& wm, unsigned grid, dim3 block, uint32_t shm,
                       hipStream_t s) {
    if (grid >= wm.total)
        hipLaunchKernelGGL((k_batch1<K, A...>), dim3(wm.total), block, shm, s, static_cast<const Pack<A...>*>(dargs),
                           wm);
    else
        hipLaunchKernelGGL((k_batch<K, A...>), dim3(grid), block, shm, s, static_cast<const Pack<A...>*>(dargs), wm);
    return hipGetLastError();
}
template <class K, class... A>
int batchOccupancy(int threads, uint32_t shm) {
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_batch<K, A...>, threads, shm) != hipSuccess) return 0;
    return nb;
}

// The name of body K, from the compiler's signature of this function
// ("... [K = skt::KPcgA<false, true>]").
template <class K>
const char* bodyName() {
    return __PRETTY_FUNCTION__;
}

// Launch of body K with 'grid' tiles (its x, y and z counts multiplied: the
// bodies take a one-dimensional tile index) on the env's stream. When a
// timed launch has armed kevA/kevB, the events are attached to this dispatch
// (hipExtLaunchKernelGGL records the kernel's own start and end) and
// disarmed; marker events between kernels cost several microseconds of idle
// device time per launch. Every body is compiled for blocks of at most 256
// threads (__launch_bounds__(256)); a larger block is refused here.
template <class K, class... A>
hipError_t klaunch(EnvState& e, dim3 grid, dim3 block, uint32_t shm, A... args) {
    const unsigned long long nb = (unsigned long long)grid.x * grid.y * grid.z;
    if (!blockFits(block.x, block.y, block.z) || nb > 0x7FFFFFFFull) return hipErrorInvalidConfiguration;
    if (e.rec) {
        RecLaunch L;
        L.fn = &batchLaunch<K, A...>;
        L.occ = &batchOccupancy<K, A...>;
        L.name = bodyName<K>();
        L.nb = (unsigned)nb;
        L.block = block;
        L.shm = shm;
        Pack<A...> p;
        pfill(p, args...);
        static_assert(alignof(Pack<A...>) <= 256, "argument records are placed at 256-byte boundaries");
        L.args.resize(sizeof(p));
        std::memcpy(L.args.data(), &p, sizeof(p));
        e.rec->launches.push_back(std::move(L));
        return hipSuccess;
    }
    if (nb == 0) return hipSuccess;
    if (e.kevA && e.kevB) {
        hipEvent_t a = e.kevA, b = e.kevB;
        e.kevA = e.kevB = nullptr;
        hipExtLaunchKernelGGL((k_solo<K, A...>), dim3((unsigned)nb), block, shm, e.stream, a, b, 0u, args...);
    } else {
        hipLaunchKernelGGL((k_solo<K, A...>), dim3((unsigned)nb), block, shm, e.stream, args...);
    }
    return hipGetLastError();
}

}  // namespace skt
