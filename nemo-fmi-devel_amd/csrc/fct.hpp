// fct.hpp -- device helpers shared by the FCT kernels (tra.hip, fct_fused.hip):
// the transports of tra_adv formed where used, the vertical upstream flux,
// the pivot table of interp_4th_cpt and the face limiter of nonosc.
#pragma once
#include "common.hpp"

// deepest column the data-independent pivot table of interp_4th_cpt covers
#define STEP_MAXK 160

template <typename T>
struct PivotTab {
    T p[STEP_MAXK + 2];
};

// (k is uniform across a wave: the level tests select, they do not branch,
// so the loads of a stencil can be issued together)
template <typename T>
__device__ __forceinline__ T trp_u(const DevDomain &d, const T *un, int p, int q, int k)
{
    const Metrics<T> &m = met<T>(d);
    const size_t pq = IJ(p, q);
    const int kc = k < d.jpk ? k : d.jpk;
    const T v = m.e2u[pq] * E3(m.e3u, pq, kc) * A3(un, p, q, kc);
    return k >= d.jpk ? T(0) : v;
}
template <typename T>
__device__ __forceinline__ T trp_v(const DevDomain &d, const T *vn, int p, int q, int k)
{
    const Metrics<T> &m = met<T>(d);
    const size_t pq = IJ(p, q);
    const int kc = k < d.jpk ? k : d.jpk;
    const T v = m.e1v[pq] * E3(m.e3v, pq, kc) * A3(vn, p, q, kc);
    return k >= d.jpk ? T(0) : v;
}
template <typename T>
__device__ __forceinline__ T trp_w(const DevDomain &d, const T *wn, int p, int q, int k)
{
    const int kc = k < d.jpk ? k : d.jpk;
    const T v = met<T>(d).e1e2t[IJ(p, q)] * A3(wn, p, q, kc);
    return k >= d.jpk ? T(0) : v;
}
// interp_4th_cpt pivot of row k in a column whose last wet level is ikb
template <typename T>
__device__ __forceinline__ T cpt_piv(const PivotTab<T> &pt, int k, int ikb)
{
    return (k >= 3 && k < ikb) ? pt.p[k] : T(1);
}

// The pivots of interp_4th_cpt's wet rows do not depend on the data: rows
// 3.. of a wet column follow p(k) = 4 - 1/p(k-1) from p(3) = 4 (the
// forward elimination of traadv_fct.F90:577-595 with wmask 1), every other
// row has pivot 1.  The table is evaluated at compile time (IEEE
// arithmetic, the operations of the elimination in its order) into constant
// memory, which the kernels read with a level index.
template <typename T>
constexpr PivotTab<T> cpt_table()
{
    PivotTab<T> pt{};
    for (int k = 0; k < STEP_MAXK + 2; k++) pt.p[k] = T(1);
    T zws_prev = T(0), zwt_prev = T(1);
    for (int k = 3; k <= STEP_MAXK; k++) {
        const T wm = T(1);
        const T zwd = T(3) * wm + T(1);
        const T zwt_k = zwd - wm * zws_prev / zwt_prev;
        pt.p[k] = zwt_k;
        zws_prev = wm;
        zwt_prev = zwt_k;
    }
    return pt;
}
static __constant__ PivotTab<float> g_cpt_f = cpt_table<float>();
static __constant__ PivotTab<double> g_cpt_d = cpt_table<double>();
template <typename T>
__device__ __forceinline__ const PivotTab<T> &cpt_tab()
{
    if constexpr (sizeof(T) == 4) return g_cpt_f;
    else return g_cpt_d;
}

// transports: TRP = 0 reads the arrays of tra_adv (pun, pvn, pwn, zero at
// jpk); TRP = 1 forms them from un, vn, wn where used (traadv.F90:100-124)
template <int TRP, typename T>
__device__ __forceinline__ T tr_u(const DevDomain &d, const T *pu, int p, int q, int k)
{
    if (TRP) return trp_u(d, pu, p, q, k);
    return A3(pu, p, q, k);
}
template <int TRP, typename T>
__device__ __forceinline__ T tr_v(const DevDomain &d, const T *pv, int p, int q, int k)
{
    if (TRP) return trp_v(d, pv, p, q, k);
    return A3(pv, p, q, k);
}
template <int TRP, typename T>
__device__ __forceinline__ T tr_w(const DevDomain &d, const T *pw, int p, int q, int k)
{
    if (TRP) return trp_w(d, pw, p, q, k);
    return A3(pw, p, q, k);
}
template <int TRP, typename T>
__device__ __forceinline__ T tp_upz(const DevDomain &d, const T *pw, const T *tb, int i, int j, int k)
{
    // k == 1: linear free surface flux; k >= jpk: zero (selects on the uniform k)
    const int kc = k < d.jpk ? k : d.jpk, km1 = kc > 1 ? kc - 1 : 1;
    const T w = tr_w<TRP>(d, pw, i, j, kc);
    const T tbk = A3(tb, i, j, kc);
    const T zfp = w + fabs(w), zfm = w - fabs(w);
    const T gen = T(0.5) * (zfp * tbk + zfm * A3(tb, i, j, km1)) * WMASK(T, i, j, kc);
    const T sfc = w * tbk;
    return k == 1 ? sfc : (k >= d.jpk ? T(0) : gen);
}


// limited fluxes of nonosc (traadv_fct.F90:404-422), FP64
template <typename T>
__device__ __forceinline__ double lim_face(double flux, double bdo_here, double bup_next, double bup_here,
                                           double bdo_next)
{
    const double za = fmin2(fmin2(1.0, bdo_here), bup_next);
    const double zb = fmin2(fmin2(1.0, bup_here), bdo_next);
    const double zc = (0.5 + nsign(0.5, flux));
    return flux * (zc * za + (1.0 - zc) * zb);
}

// tra_adv_fct of nt (1 or 2) tracers on a closed domain in one k-marched
// kernel per tile and tracer (fct_fused.hip)
// the compact scheme's forward sweep of nt (1 or 2) tracers into pzo
// (fct_fused.hip)
template <typename T>
int launch_fct_cpt_fwd(const DevDomain &d, int nt, const T *ptn, T *pzo, const char *nm);

// the vertical fluxes, first guess and limiter of nt (1 or 2) tracers on a
// closed domain in one bottom-up march (fct_vlim.hip); ps: the horizontal
// upstream sums of k_fct_h (VERT = 0), pzo: the forward sweep (order 4) or
// tn (order 2), paa / pbb: zwx / zwy after their lbc_lnk
template <int TRP, typename T>
int launch_fct_vlim(const DevDomain &d, double p2dt, int nt, int kn_fct_v, int ta0,
                    const T *pw, const T *ptb, const T *ps, const T *pzo, const T *paa, const T *pbb, T *pta,
                    const char *nm);

template <int TRP, typename T>
int launch_fct_fused(const DevDomain &d, double p2dt, int nt, const T *pu, const T *pv,
                     const T *pw, const T *ptb, const T *ptn, T *pta, int kn_fct_h, int kn_fct_v, int ta0, T *pzo,
                     const char *nm);
