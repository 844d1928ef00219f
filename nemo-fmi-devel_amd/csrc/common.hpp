// common.hpp -- shared device/host definitions of the gfx950 NEMO hot-path library.
//
// Layout: every field keeps NEMO's Fortran layout (i fastest, then j, then k,
// then tracer), addressed with 1-based (i,j,k) exactly as the reference loop
// nests are written, so each kernel body reads like the loop it replaces.
//
// Masks are not stored as 3D arrays.  Under full-step z-coordinates every
// 3D mask of DOM/dommsk.F90 is, per column, the indicator of one interval of
// levels (tmask: [k_top,k_bot]; umask: the intersection of the two adjacent
// t-columns; ...), after the serial lbc_lnk of the ring.  The library keeps
// one (k0,k1) pair of int32 per column and point type, so a mask costs 0 HBM
// bytes per cell; the dom_msk entry point expands them to the reference's
// REAL 0/1 arrays for the bit-exact check.
//
// Vertical scale factors are addressed through (base, stride_k, stride_ij):
// a horizontally uniform z-coordinate registers a jpk-long table (stride_ij 0),
// anything else the full 3D array (stride_ij 1).
#pragma once

#include <hip/hip_runtime.h>
#include <hip/hip_ext.h>
#include <stdint.h>
#include <stddef.h>
#include <type_traits>

#include "../../include/nemo_hip.h"

#define NH_MAXSTR 1024

// ---------------------------------------------------------------------------
// per-column level interval: mask(k) = (k0 <= k && k <= k1), 1-based
// ---------------------------------------------------------------------------
struct KRange {
    const int32_t *k0, *k1;
};

template <typename T>
struct Metrics {
    // 2D (jpi,jpj)
    const T *e1t, *e2t, *e1u, *e2u, *e1v, *e2v;
    const T *e1e2t, *r1_e1e2t, *e1e2u, *r1_e1e2u, *e1e2v, *r1_e1e2v, *r1_e1e2f;
    const T *e2_e1u, *e1_e2v, *r1_e1u, *r1_e2u, *r1_e1v, *r1_e2v;
    const T *ff_f, *ff_t;
    const T *di_e2v_2e1e2f, *dj_e1u_2e1e2f, *di_e2u_2, *dj_e1v_2;   // DYN/dynvor.F90:919-939
    // vertical scale factors
    const T *e3t, *e3u, *e3v, *e3w, *e3f, *e3uw, *e3vw;
};

struct DevDomain {
    int jpi, jpj, jpk;
    int iper, jper;                 // l_Iperio, l_Jperio
    size_t n2, n3;
    size_t e3_sk, e3_sij;           // strides of the e3 arrays
    // options
    int ln_traldf_lap, ln_traldf_msc, ln_zdfddm;
    int n_dynadv, nn_dynkeg, nvor_scheme, nn_een_e3f;
    int ln_drgimp, ln_dynspg_ts, ln_bt_fw, ln_bt_av, nn_bt_flt, nn_baro;
    int neuler, nit000, ln_traqsr;
    double rdt, rn_atfp, rau0, grav, rn_scal_load;
    // masks as level intervals, DOM/dommsk.F90
    KRange kT, kU, kV, kF;
    // the same intervals packed as bytes, one 32-bit load per column for two
    // point types: ktu = T0 | T1 << 8 | U0 << 16 | U1 << 24, kvf = V0 | V1 << 8 | F0 << 16 | F1 << 24
    const uint32_t *ktu, *kvf;
    const double *ssmask, *ssumask, *ssvmask;
    const float *ssmask_f, *ssumask_f, *ssvmask_f;
    // k-indices, DOM/domzgr.F90:292-314
    const int32_t *mikt, *mbkt, *miku, *mikv, *mbku, *mbkv;
    Metrics<double> md;
    Metrics<float> mf;
};

template <typename T> __host__ __device__ inline const Metrics<T> &met(const DevDomain &d);
template <> __host__ __device__ inline const Metrics<double> &met<double>(const DevDomain &d) { return d.md; }
template <> __host__ __device__ inline const Metrics<float> &met<float>(const DevDomain &d) { return d.mf; }

// ---------------------------------------------------------------------------
// indexing (1-based, Fortran order)
// ---------------------------------------------------------------------------
#define IJ(i, j) ((size_t)((j) - 1) * (size_t)d.jpi + (size_t)((i) - 1))
#define IJK(i, j, k) (((size_t)((k) - 1) * (size_t)d.jpj + (size_t)((j) - 1)) * (size_t)d.jpi + (size_t)((i) - 1))
#define A2(p, i, j) (p)[IJ(i, j)]
// 3D element access with a 32-bit byte offset: with a wave-uniform base the
// load takes the saddr + voffset form (no 64-bit address per lane).  The
// domain registration refuses arrays of 4 GiB or more.
template <typename T>
__device__ __forceinline__ T &at32(T *p, uint32_t e)
{
    return *reinterpret_cast<T *>(reinterpret_cast<char *>(const_cast<typename std::remove_const<T>::type *>(p)) +
                                  (uint32_t)(e * (uint32_t)sizeof(T)));
}
#define IJK32(i, j, k) ((((uint32_t)((k) - 1) * (uint32_t)d.jpj + (uint32_t)((j) - 1)) * (uint32_t)d.jpi) + (uint32_t)((i) - 1))
#define A3(p, i, j, k) at32((p), IJK32(i, j, k))
#define E3(p, ij, k) (p)[(size_t)((k) - 1) * d.e3_sk + (size_t)(ij) * d.e3_sij]

// Plane-relative access for k-marched kernels: the level base pointer is
// uniform across the wave (kept in SGPRs), the lane's in-plane offset is a
// 32-bit byte offset, so loads use the saddr + 32-bit voffset form and the
// 64-bit per-lane address arithmetic disappears.  Valid while a plane is
// below 4 GiB (jpi*jpj*sizeof(T) < 2^32).
template <typename T>
__device__ __forceinline__ const T *plane_ptr(const T *base, size_t plane, int k)
{
    return base + (size_t)(k - 1) * plane;
}
template <typename T>
__device__ __forceinline__ T *plane_ptr(T *base, size_t plane, int k)
{
    return base + (size_t)(k - 1) * plane;
}
template <typename T>
__device__ __forceinline__ T ldp(const T *pl, uint32_t ij)
{
    return *(const T *)((const char *)pl + (uint32_t)(ij * (uint32_t)sizeof(T)));
}
template <typename T>
__device__ __forceinline__ void stp(T *pl, uint32_t ij, T v)
{
    *(T *)((char *)pl + (uint32_t)(ij * (uint32_t)sizeof(T))) = v;
}
#define IJ32(i, j) ((uint32_t)((j) - 1) * (uint32_t)d.jpi + (uint32_t)((i) - 1))

template <typename T>
__device__ __forceinline__ T kmask(const KRange &r, size_t ij, int k)
{
    return (k >= r.k0[ij] && k <= r.k1[ij]) ? T(1) : T(0);
}
// w-point masks: wmask(1) = tmask(1), wmask(k) = tmask(k) * tmask(k-1), DOM/dommsk.F90:182-190
template <typename T>
__device__ __forceinline__ T wkmask(const KRange &r, size_t ij, int k)
{
    int a = r.k0[ij], b = r.k1[ij];
    bool in = (k >= a && k <= b) && (k == 1 || (k - 1 >= a && k - 1 <= b));
    return in ? T(1) : T(0);
}

template <typename T, int SH>
__device__ __forceinline__ T pmask(const uint32_t *p, uint32_t ij, int k)
{
    const uint32_t w = p[ij] >> SH;
    const int a = (int)(w & 255u), b = (int)((w >> 8) & 255u);
    return (k >= a && k <= b) ? T(1) : T(0);
}
template <typename T, int SH>
__device__ __forceinline__ T wpmask(const uint32_t *p, uint32_t ij, int k)
{
    const uint32_t w = p[ij] >> SH;
    const int a = (int)(w & 255u), b = (int)((w >> 8) & 255u);
    return (k >= a && k <= b && (k == 1 || k - 1 >= a)) ? T(1) : T(0);
}

// surface masks ssmask / ssumask / ssvmask (1 where the column has a wet
// level, DOM/dommsk.F90) from the same packed intervals: the values the
// registration uploads as 2D arrays, read as 4 bytes per column
template <int SH>
__device__ __forceinline__ double psurf(const uint32_t *p, uint32_t ij)
{
    const uint32_t w = p[ij] >> SH;
    return ((w >> 8) & 255u) >= (w & 255u) ? 1.0 : 0.0;
}
#define SSMASK(ij) psurf<0>(d.ktu, (uint32_t)(ij))
#define SSUMASK(ij) psurf<16>(d.ktu, (uint32_t)(ij))
#define SSVMASK(ij) psurf<0>(d.kvf, (uint32_t)(ij))

#define TMASK(T_, i, j, k) pmask<T_, 0>(d.ktu, IJ32(i, j), k)
#define UMASK(T_, i, j, k) pmask<T_, 16>(d.ktu, IJ32(i, j), k)
#define VMASK(T_, i, j, k) pmask<T_, 0>(d.kvf, IJ32(i, j), k)
#define WMASK(T_, i, j, k) wpmask<T_, 0>(d.ktu, IJ32(i, j), k)
#define WUMASK(T_, i, j, k) wpmask<T_, 16>(d.ktu, IJ32(i, j), k)
#define WVMASK(T_, i, j, k) wpmask<T_, 0>(d.kvf, IJ32(i, j), k)

// A column Thomas solve may keep the FP64 pivot of every ZDF_SEG-th level
// only (a checkpoint in scratch): the back substitution then recomputes the
// pivots of each segment in registers from the checkpoint above it with the
// operations of the forward sweep, so they are the same values, and the
// per-level pivot store and reload (16 bytes per cell) disappear.
#define ZDF_SEG 4

// NEMO's SIGN with key_nosignedzero semantics, lib_fortran.F90:336-351
template <typename T>
__device__ __forceinline__ T nsign(T a, T b) { return b >= T(0) ? fabs(a) : -fabs(a); }
template <typename T>
__device__ __forceinline__ T fmax2(T a, T b) { return a > b ? a : b; }
template <typename T>
__device__ __forceinline__ T fmin2(T a, T b) { return a < b ? a : b; }

// ---------------------------------------------------------------------------
// serial lbc_lnk as an index map, LBC/lbc_lnk_generic.h90:80-103.
// The value a field holds at (i,j) after lbc_lnk is either the closed
// boundary value (returns false) or the value computed at (si,sj) before the
// call.  The east-west fill runs first and the north-south fill then copies
// whole rows, so a corner takes the E-W rule of its source row.
// ---------------------------------------------------------------------------
__device__ __forceinline__ bool lbc_src(const DevDomain &d, char nat, int i, int j, int &si, int &sj)
{
    si = i;
    sj = j;
    if (j == 1 || j == d.jpj) {
        if (d.jper) sj = (j == 1) ? d.jpj - 1 : 2;
        else if (!(nat == 'F' && j == 1)) return false;
    }
    if (i == 1 || i == d.jpi) {
        if (d.iper) si = (i == 1) ? d.jpi - 1 : 2;
        else if (!(nat == 'F' && i == 1)) return false;
    }
    return true;
}

// branch-free form for stencils: when the point has no source (closed ring)
// the source is moved to the interior point (2,2), so the caller can load its
// stencil unconditionally (every load stays in bounds) and select zero at the
// end; the compiler can then issue all loads of a stencil together
__device__ __forceinline__ bool lbc_src_safe(const DevDomain &d, char nat, int i, int j, int &si, int &sj)
{
    const bool ok = lbc_src(d, nat, i, j, si, sj);
    si = ok ? si : 2;
    sj = ok ? sj : 2;
    return ok;
}

__device__ __forceinline__ bool is_interior(const DevDomain &d, int i, int j)
{
    return i >= 2 && i <= d.jpi - 1 && j >= 2 && j <= d.jpj - 1;
}

// Every kernel runs 256-thread blocks (one wave per SIMD).  Declaring that
// bound lifts the register cap the compiler applies for 1024-thread blocks
// (128 VGPRs per lane), which made the larger stencils spill to scratch.
#define NH_BLOCK 256
#define KERNEL __global__ __launch_bounds__(NH_BLOCK)
// same, asking the compiler to keep at least W waves per SIMD resident
// (caps registers at 512/W per lane) for kernels whose stencils are large
#define KERNEL_W(W) __global__ __launch_bounds__(NH_BLOCK, W)

// 2D launch geometry over (i,j) with an explicit index origin
struct Grid2 {
    dim3 grid, block;
};
inline Grid2 grid2(int ni, int nj, int bx = 64, int by = NH_BLOCK / 64)
{
    Grid2 g;
    if (bx * by != NH_BLOCK) by = NH_BLOCK / bx;
    g.block = dim3(bx, by, 1);
    g.grid = dim3((unsigned)((ni + bx - 1) / bx), (unsigned)((nj + by - 1) / by), 1);
    return g;
}
// XCD-aware tile order.  The dispatcher deals the workgroups of a launch
// round-robin to the 8 XCDs, each with its own L2, so with the plain order the
// i-neighbour tiles of a tile sit in other L2s.  Within each level plane of
// the grid, workgroup b takes tile l = (b mod 8) * (tiles/8) + b div 8
// instead: XCD e then works through the contiguous band e of tile rows, where
// the i- and j-neighbours of a tile (and, for per-cell kernels, the band one
// level up) are cached in its own L2.  Planes whose tile count is not a
// multiple of 8 keep the plain order.
__device__ __forceinline__ void xcd_tile(unsigned &bx, unsigned &by)
{
    const unsigned nx = gridDim.x, np = gridDim.x * gridDim.y;
    const unsigned b = blockIdx.x + nx * blockIdx.y;
    const unsigned l = (np & 7u) ? b : (b & 7u) * (np >> 3) + (b >> 3);
    bx = l % nx;
    by = l / nx;
}
#define THREAD_IJ(i0, j0) \
    unsigned bx_, by_; \
    xcd_tile(bx_, by_); \
    const int i = (int)(bx_ * blockDim.x + threadIdx.x) + (i0); \
    const int j = (int)(by_ * blockDim.y + threadIdx.y) + (j0);

// ---------------------------------------------------------------------------
// host side: error state, stream, timing, launch wrapper
// ---------------------------------------------------------------------------
namespace nh {
int set_error(int code, const char *fmt, ...);
hipStream_t stream();
const DevDomain &dom();
bool capturing();              // a graph is being recorded on the library stream
int *spg_rot_u();              // dyn_spg_ts level buffers of the current context
int *spg_rot_s();
unsigned *spg_barrier();       // dyn_spg_ts's grid barrier (2 words) and its status word after them
void spg_loop_launched();      // a cooperative phase 2 was queued: check its status at the next sync
bool have_domain();
// tuning switches (nemo_hip_tune_set): kernel variants and tile shapes
enum Tune { T_FCT_FUSED, T_SPG_PERSISTENT, T_ZDF_REG, T_FCT_VLIM, T_NTUNE };
int tune(Tune which);
// per-routine timing: when this launch is timed, returns true with the two
// events its dispatch packet records at kernel start and end
bool timing_pair(const char *name, const char *kernel, hipEvent_t *start, hipEvent_t *stop);
// library-owned scratch buffers (HBM), sized at domain registration
enum Scratch {
    S_ZWI, S_ZWX, S_ZWY, S_ZWZ, S_ZBETUP, S_ZBETDO, S_ZTW, S_ZWT, S_ZWT2, S_ZUN, S_ZVN, S_ZWN, S_NSCRATCH3
};
void *scratch3(int which);     // n3 doubles each, slot s+1 directly after slot s
#define NH_NSCRATCH2 40
double *scratch2(int which);   // n2 doubles each (dyn_spg_ts work arrays), NH_NSCRATCH2 of them
#define NH_S2_SLP 34           // slots 34..37: ldf_slp mixed-layer slopes (dyn_spg_ts uses 0..33)
}  // namespace nh

#define HIPCHK(x)                                                                                  \
    do {                                                                                           \
        hipError_t e_ = (x);                                                                       \
        if (e_ != hipSuccess)                                                                      \
            return nh::set_error(4, "%s:%d: %s: %s", __FILE__, __LINE__, #x, hipGetErrorString(e_)); \
    } while (0)

#include "batch.hpp"

// launch on the library stream (timed when timing is on); while a plan is
// traced, the launch is recorded for this member instead
#define LAUNCH(name, kern, g, ...)                                                                    \
    do {                                                                                              \
        if (nh::planning()) {                                                                         \
            const Grid2 gp_ = (g);                                                                    \
            nh::PlanRec r_;                                                                           \
            r_.routine = name;                                                                        \
            r_.kname = #kern;                                                                         \
            r_.grid = gp_.grid;                                                                       \
            r_.block = gp_.block;                                                                     \
            r_.single = [=](hipStream_t s_) {                                                         \
                hipLaunchKernelGGL(kern, gp_.grid, gp_.block, 0, s_, __VA_ARGS__);                    \
                return hipGetLastError();                                                             \
            };                                                                                        \
            const int rc_ = nh::plan_add(std::move(r_));                                              \
            if (rc_) return rc_;                                                                      \
            break;                                                                                    \
        }                                                                                             \
        hipEvent_t ev0_, ev1_;                                                                        \
        if (nh::timing_pair(name, #kern, &ev0_, &ev1_))                                               \
            hipExtLaunchKernelGGL(kern, (g).grid, (g).block, 0, nh::stream(), ev0_, ev1_, 0, __VA_ARGS__); \
        else                                                                                          \
            hipLaunchKernelGGL(kern, (g).grid, (g).block, 0, nh::stream(), __VA_ARGS__);              \
        HIPCHK(hipGetLastError());                                                                    \
    } while (0)
