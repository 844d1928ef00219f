// batch.hpp -- launches of one kernel over the tiles of many domains.
//
// An ensemble of members (one library context each) runs the same sequence
// of kernels per time step, each member with its own grid, domain and
// fields.  A plan (nemo_hip_plan_begin / end) traces one step of every
// member without launching anything, then turns position p of the members'
// launch sequences into ONE launch over the concatenated work list of
// (member, workgroup) pairs: workgroup b of the batched launch finds its
// member m from the members' first workgroups and runs the kernel body with
// m's arguments and m's own workgroup coordinates.  Kernels written as a
// body (a struct S with a static device function run(const VB &, args...))
// batch this way; any other launch of the step stays one launch per member
// in the plan.
//
// Each member's range of workgroups starts at a multiple of 8, so the
// XCD-aware tile order of a member (xcd_tile_v) sees the same workgroup
// numbering modulo 8 as in a launch of its own.
#pragma once
#include <cstring>
#include <functional>
#include <memory>
#include <string>
#include <type_traits>
#include <utility>
#include <vector>

// virtual workgroup coordinates of a kernel body: the hardware ones in a
// launch of its own, the member's own in a batched launch
struct VB {
    unsigned x, y, z, gx, gy, gz;
};
__device__ __forceinline__ VB vb_hw() { return VB{blockIdx.x, blockIdx.y, blockIdx.z, gridDim.x, gridDim.y, gridDim.z}; }

// xcd_tile on virtual coordinates
__device__ __forceinline__ void xcd_tile_v(const VB &v, unsigned &bx, unsigned &by)
{
    const unsigned nx = v.gx, np = v.gx * v.gy;
    const unsigned b = v.x + nx * v.y;
    const unsigned l = (np & 7u) ? b : (b & 7u) * (np >> 3) + (b >> 3);
    bx = l % nx;
    by = l / nx;
}
#define THREAD_IJ_V(v, i0, j0)                                          \
    unsigned bx_, by_;                                                  \
    xcd_tile_v(v, bx_, by_);                                            \
    const int i = (int)(bx_ * blockDim.x + threadIdx.x) + (i0);         \
    const int j = (int)(by_ * blockDim.y + threadIdx.y) + (j0);

// a plain aggregate holding a kernel's arguments (trivially copyable when
// they are, unlike std::tuple), with get<I>
template <class... A> struct Pack {};
template <class H, class... R> struct Pack<H, R...> {
    H h;
    Pack<R...> t;
};
template <size_t I, class H, class... R>
__host__ __device__ __forceinline__ const auto &pget(const Pack<H, R...> &p)
{
    if constexpr (I == 0) return p.h;
    else return pget<I - 1>(p.t);
}
template <class H, class... R>
inline Pack<H, R...> make_pack(const H &h, const R &...r)
{
    if constexpr (sizeof...(R) == 0) return Pack<H>{h, Pack<>{}};
    else return Pack<H, R...>{h, make_pack(r...)};
}

// one member of a batched launch: its kernel arguments and its own grid
template <class... A> struct BatchEnt {
    Pack<A...> a;
    unsigned gx, gy, gz;
};

template <class S, class... A, size_t... I>
__device__ __forceinline__ void batch_call(const VB &v, const Pack<A...> &p, std::index_sequence<I...>)
{
    S::run(v, pget<I>(p)...);
}

// start[0..nm]: the first workgroup of each member (start[nm] = total)
template <class S, class... A>
__global__ __launch_bounds__(S::NTHR, S::WAVES) void k_batch(const unsigned *__restrict__ start, int nm,
                                                        const BatchEnt<A...> *__restrict__ ent)
{
    const unsigned b = blockIdx.x;
    int lo = 0, hi = nm;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (start[mid] <= b) lo = mid;
        else hi = mid;
    }
    const BatchEnt<A...> &e = ent[lo];
    const unsigned l = b - start[lo];
    if (l >= e.gx * e.gy * e.gz) return;   // the padding of the member's range
    const VB v{l % e.gx, (l / e.gx) % e.gy, l / (e.gx * e.gy), e.gx, e.gy, e.gz};
    batch_call<S>(v, e.a, std::index_sequence_for<A...>{});
}

namespace nh {
bool planning();   // a plan is being traced: launches are recorded, not issued

using BatchFn = hipError_t (*)(unsigned nblk, dim3 block, hipStream_t s, hipEvent_t e0, hipEvent_t e1,
                               const unsigned *start, int nm, const void *ent);

// one traced launch of one member
struct PlanRec {
    std::string routine, kname;
    dim3 grid, block;
    BatchFn bfn = nullptr;             // batchable: the launcher of a batch of such entries
    std::vector<char> ent;             // ... and this member's BatchEnt
    std::function<hipError_t(hipStream_t)> single;   // otherwise: this member's launch alone
};
int plan_add(PlanRec &&r);

template <class S, class... A>
hipError_t batch_launch(unsigned nblk, dim3 block, hipStream_t s, hipEvent_t e0, hipEvent_t e1, const unsigned *start,
                        int nm, const void *ent)
{
    const BatchEnt<A...> *e = (const BatchEnt<A...> *)ent;
    if (e0)
        hipExtLaunchKernelGGL((k_batch<S, A...>), dim3(nblk), block, 0, s, e0, e1, 0, start, nm, e);
    else
        hipLaunchKernelGGL((k_batch<S, A...>), dim3(nblk), block, 0, s, start, nm, e);
    return hipGetLastError();
}

template <class S, class... A>
int plan_batch(const char *name, const char *kname, dim3 grid, dim3 block, const A &...a)
{
    using E = BatchEnt<std::decay_t<A>...>;
    static_assert(std::is_trivially_copyable<E>::value, "batched kernel arguments must be trivially copyable");
    E e{make_pack<std::decay_t<A>...>(a...), grid.x, grid.y, grid.z};
    PlanRec r;
    r.routine = name;
    r.kname = kname;
    r.grid = grid;
    r.block = block;
    r.bfn = &batch_launch<S, std::decay_t<A>...>;
    r.ent.resize(sizeof(E));
    std::memcpy(r.ent.data(), &e, sizeof(E));
    return plan_add(std::move(r));
}
}  // namespace nh

#define NH_UNPAREN(...) __VA_ARGS__
// LAUNCH of a kernel whose body is the struct S (given in parentheses):
// traced as a batchable launch while a plan is traced, launched otherwise
#define LAUNCHB(name, kern, S, g, ...)                                                        \
    do {                                                                                      \
        if (nh::planning()) {                                                                 \
            const int rc_ = nh::plan_batch<NH_UNPAREN S>(name, #kern, (g).grid, (g).block, __VA_ARGS__); \
            if (rc_) return rc_;                                                              \
            break;                                                                            \
        }                                                                                     \
        LAUNCH(name, kern, g, __VA_ARGS__);                                                   \
    } while (0)
