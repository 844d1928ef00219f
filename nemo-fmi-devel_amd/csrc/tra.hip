// tra.hip -- tracer operators of the ocean time step on gfx950.
//
//   tra_adv_trp     TRA/traadv.F90:100-124      effective transport
//   tra_adv_fct     TRA/traadv_fct.F90:54-327   FCT advection (2nd/4th order)
//   nonosc          TRA/traadv_fct.F90:330-428  Zalesak limiter
//   interp_4th_cpt  TRA/traadv_fct.F90:517-616  compact T->w interpolation
//   tra_ldf_lap     TRA/traldf_lap_blp.F90:47-158
//   tra_ldf_iso     TRA/traldf_iso.F90:50-382
//   tra_zdf_imp     TRA/trazdf.F90:109-268 (+ the salinity reset of tra_zdf)
//
// Every kernel is an (i,j)-tiled, k-marched stencil: one thread owns a
// column, consecutive threads own consecutive i, so every level access is a
// coalesced row segment.  Neighbour columns come through L1/L2; values that
// NEMO exchanges with lbc_lnk between loop nests are either produced by a ring
// kernel (k_lbc) or recomputed at their lbc source point (lbc_src).
//
// Arithmetic: each expression is evaluated in the reference's order (Fortran
// left-to-right, parentheses honoured) and the library is compiled with
// -ffp-contract=off, so the FP64 form reproduces the reference bit for bit.
// The template parameter T is the storage (and stencil) type; the Zalesak
// betas/limiter and the tridiagonal eliminations run in FP64 in both forms.
#include "common.hpp"
#include "kernels.hpp"
#include "fct.hpp"

// levels per thread of k_fct_beta: a short k-march reuses the vertical
// neighbours the thread has loaded, where one level per thread re-fetches
// them through the fabric.  It pays for the FP64 betas; for FP32 storage one
// level per thread is faster (measured at 1024x1024x75), and so it is for the
// flux kernel k_fct_h, whose k-march needs too many registers
template <typename T> struct FctKB { static constexpr int beta = sizeof(T) == 8 ? 8 : 1; };
template <int TRP, typename T>
static int launch_fct_pairs(const DevDomain &d, double p2dt, const T *pu, const T *pv, const T *pw, const T *ptb,
                            const T *ptn, T *pta, int kjpt, int kn_fct_h, int kn_fct_v, int ta0);
template <typename T>
static int launch_fct_fluxes1(const DevDomain &d, double p2dt, const T *pu, const T *pv, const T *pw, const T *ptb,
                              const T *ptn, T *pta, int kn_h, int kn_v, T *zwi, T *zwx, T *zwy, T *zwz,
                              const char *nm);

// ---------------------------------------------------------------------------
// effective transport, TRA/traadv.F90:100-124
// ---------------------------------------------------------------------------
template <typename T>
struct S_tra_adv_trp {
    static constexpr int NTHR = NH_BLOCK, WAVES = 1;
    __device__ __forceinline__ static void run(const VB &vb_, const DevDomain &d, const T * __restrict__ un, const T * __restrict__ vn, const T * __restrict__ wn, T * __restrict__ zun, T * __restrict__ zvn, T * __restrict__ zwn)
    {
        THREAD_IJ_V(vb_, 1, 1);
        if (i > d.jpi || j > d.jpj) return;
        const Metrics<T> &m = met<T>(d);
        const size_t ij = IJ(i, j);
        const T e2u = m.e2u[ij], e1v = m.e1v[ij], e1e2t = m.e1e2t[ij];
        for (int k = 1; k <= d.jpk - 1; k++) {
            const size_t p = IJK(i, j, k);
            zun[p] = e2u * E3(m.e3u, ij, k) * un[p];
            zvn[p] = e1v * E3(m.e3v, ij, k) * vn[p];
            zwn[p] = e1e2t * wn[p];
        }
        const size_t p = IJK(i, j, d.jpk);
        zun[p] = T(0);
        zvn[p] = T(0);
        zwn[p] = T(0);
    }
};
template <typename T>
KERNEL void k_tra_adv_trp(DevDomain d, const T * __restrict__ un, const T * __restrict__ vn, const T * __restrict__ wn, T * __restrict__ zun, T * __restrict__ zvn, T * __restrict__ zwn)
{
    S_tra_adv_trp<T>::run(vb_hw(), d, un, vn, wn, zun, zvn, zwn);
}

template <typename T>
int launch_tra_adv_trp(const T *un, const T *vn, const T *wn, T *zun, T *zvn, T *zwn)
{
    const DevDomain &d = nh::dom();
    Grid2 g = grid2(d.jpi, d.jpj);
    LAUNCHB("tra_adv_trp", k_tra_adv_trp<T>, (S_tra_adv_trp<T>), g, d, un, vn, wn, zun, zvn, zwn);
    return 0;
}
template int launch_tra_adv_trp<double>(const double *, const double *, const double *, double *, double *, double *);
template int launch_tra_adv_trp<float>(const float *, const float *, const float *, float *, float *, float *);

// ---------------------------------------------------------------------------
// FCT: upstream fluxes, recomputed wherever a loop nest needs them
// (traadv_fct.F90:123-156)
// ---------------------------------------------------------------------------
template <typename T>
__device__ __forceinline__ T up_x(const DevDomain &d, const T *pun, const T *tb, int i, int j, int k)
{
    const T u = A3(pun, i, j, k);
    const T zfp = u + fabs(u), zfm = u - fabs(u);
    return T(0.5) * (zfp * A3(tb, i, j, k) + zfm * A3(tb, i + 1, j, k));
}
template <typename T>
__device__ __forceinline__ T up_y(const DevDomain &d, const T *pvn, const T *tb, int i, int j, int k)
{
    const T v = A3(pvn, i, j, k);
    const T zfp = v + fabs(v), zfm = v - fabs(v);
    return T(0.5) * (zfp * A3(tb, i, j, k) + zfm * A3(tb, i, j + 1, k));
}
// w-flux at level k: k == 1 linear-free-surface flux, k == jpk zero
template <typename T>
__device__ __forceinline__ T up_z(const DevDomain &d, const T *pwn, const T *tb, int i, int j, int k)
{
    if (k == 1) return A3(pwn, i, j, 1) * A3(tb, i, j, 1);
    if (k >= d.jpk) return T(0);
    const T w = A3(pwn, i, j, k);
    const T zfp = w + fabs(w), zfm = w - fabs(w);
    return T(0.5) * (zfp * A3(tb, i, j, k) + zfm * A3(tb, i, j, k - 1)) * WMASK(T, i, j, k);
}

// (1) low-order trend into pta and first guess zwi, traadv_fct.F90:158-170
template <typename T>
KERNEL void k_fct_low(DevDomain d, double p2dt, const T * __restrict__ pun, const T * __restrict__ pvn, const T * __restrict__ pwn, const T * __restrict__ tb, T * __restrict__ ta,
                          T * __restrict__ zwi)
{
    THREAD_IJ(2, 2);
    if (i > d.jpi - 1 || j > d.jpj - 1) return;
    const Metrics<T> &m = met<T>(d);
    const size_t ij = IJ(i, j);
    const T r1_e1e2t = m.r1_e1e2t[ij];
    const T zp2dt = (T)p2dt;
    T zwz_k = up_z(d, pwn, tb, i, j, 1);
    for (int k = 1; k <= d.jpk - 1; k++) {
        const T zwz_kp1 = up_z(d, pwn, tb, i, j, k + 1);
        const T ztra = -((up_x(d, pun, tb, i, j, k) - up_x(d, pun, tb, i - 1, j, k) + up_y(d, pvn, tb, i, j, k) -
                          up_y(d, pvn, tb, i, j - 1, k) + zwz_k - zwz_kp1) *
                         r1_e1e2t);
        const T e3t = E3(m.e3t, ij, k);
        const T tm = TMASK(T, i, j, k);
        const size_t p = IJK(i, j, k);
        ta[p] = ta[p] + ztra / e3t * tm;
        zwi[p] = (e3t * tb[p] + zp2dt * ztra) / e3t * tm;
        zwz_k = zwz_kp1;
    }
    zwi[IJK(i, j, d.jpk)] = T(0);
}

// (2) anti-diffusive fluxes, high order minus upstream, traadv_fct.F90:180-278.
// Threads cover 1..jpim1 x 1..jpjm1 (the horizontal flux loops); the
// vertical flux, and for kn_fct_v == 4 the compact interpolation of
// interp_4th_cpt, is done by the interior columns.
template <typename T>
__device__ __forceinline__ T fct_zltu(const DevDomain &d, const T *tn, int i, int j, int k, bool xdir)
{
    // zltu / zltv after lbc_lnk('T'), traadv_fct.F90:195-209
    int si, sj;
    const bool ok = lbc_src_safe(d, 'T', i, j, si, sj);
    const T r1_6 = T(1) / T(6);
    T r;
    if (xdir) {
        const T a = (A3(tn, si + 1, sj, k) - A3(tn, si, sj, k)) * UMASK(T, si, sj, k);
        const T b = (A3(tn, si, sj, k) - A3(tn, si - 1, sj, k)) * UMASK(T, si - 1, sj, k);
        r = (a + b) * r1_6;
    } else {
        const T a = (A3(tn, si, sj + 1, k) - A3(tn, si, sj, k)) * VMASK(T, si, sj, k);
        const T b = (A3(tn, si, sj, k) - A3(tn, si, sj - 1, k)) * VMASK(T, si, sj - 1, k);
        r = (a + b) * r1_6;
    }
    return ok ? r : T(0);
}

template <typename T>
KERNEL void k_fct_ad(DevDomain d, int kn_fct_h, int kn_fct_v, const T * __restrict__ pun, const T * __restrict__ pvn, const T * __restrict__ pwn,
                         const T * __restrict__ tb, const T * __restrict__ tn, T * __restrict__ zwx, T * __restrict__ zwy, T * __restrict__ zwz, double * __restrict__ zwt)
{
    THREAD_IJ(1, 1);
    if (i > d.jpi - 1 || j > d.jpj - 1) return;
    for (int k = 1; k <= d.jpk - 1; k++) {
        const size_t p = IJK(i, j, k);
        const T hu = up_x(d, pun, tb, i, j, k), hv = up_y(d, pvn, tb, i, j, k);
        const T c2u = A3(tn, i, j, k) + A3(tn, i + 1, j, k);
        const T c2v = A3(tn, i, j, k) + A3(tn, i, j + 1, k);
        if (kn_fct_h == 2) {
            zwx[p] = T(0.5) * pun[p] * c2u - hu;
            zwy[p] = T(0.5) * pvn[p] * c2v - hv;
        } else {
            zwx[p] = T(0.5) * pun[p] * (c2u + fct_zltu(d, tn, i, j, k, true) - fct_zltu(d, tn, i + 1, j, k, true)) - hu;
            zwy[p] = T(0.5) * pvn[p] * (c2v + fct_zltu(d, tn, i, j, k, false) - fct_zltu(d, tn, i, j + 1, k, false)) - hv;
        }
    }
    if (!is_interior(d, i, j)) return;
    const int jpk = d.jpk;
    zwz[IJK(i, j, 1)] = T(0);      // linear free surface: high order = upstream at the surface
    zwz[IJK(i, j, jpk)] = T(0);
    if (kn_fct_v == 2) {
        for (int k = 2; k <= jpk - 1; k++) {
            const size_t p = IJK(i, j, k);
            zwz[p] = (pwn[p] * T(0.5) * (tn[p] + A3(tn, i, j, k - 1)) - up_z(d, pwn, tb, i, j, k)) * WMASK(T, i, j, k);
        }
        return;
    }
    // 4th order compact: interp_4th_cpt on this column (traadv_fct.F90:535-614);
    // forward sweep keeps Tk in zwt and Zk in zwz, back substitution overwrites zwz.
    const int ikb = A2(d.mbkt, i, j);       // ikt = mikt + 1 = 2 without cavities
    T zws_prev = T(0), zwt_prev = T(1), zout_prev = T(0);
    for (int k = 2; k <= jpk - 1; k++) {
        T zwd, zwi, zws, zwrm;
        if (k == 2 || k == ikb) {
            zwd = T(1); zwi = T(0); zws = T(0);
            zwrm = T(0.5) * (A3(tn, i, j, k - 1) + A3(tn, i, j, k));
        } else {
            const T wm = WMASK(T, i, j, k);
            zwd = T(3) * wm + T(1); zwi = wm; zws = wm;
            zwrm = T(3) * wm * (A3(tn, i, j, k) + A3(tn, i, j, k - 1));
        }
        T zwt_k, zout;
        if (k == 2) {
            zwt_k = zwd;
            zout = zwrm;
        } else {
            zwt_k = zwd - zwi * zws_prev / zwt_prev;
            zout = zwrm - zwi / zwt_prev * zout_prev;
        }
        zwt[IJK(i, j, k)] = (double)zwt_k;
        zwz[IJK(i, j, k)] = zout;
        zws_prev = zws;
        zwt_prev = zwt_k;
        zout_prev = zout;
    }
    T znext = zout_prev / zwt_prev;          // X(jpkm1)
    {
        const size_t p = IJK(i, j, jpk - 1);
        zwz[p] = (pwn[p] * znext - up_z(d, pwn, tb, i, j, jpk - 1)) * WMASK(T, i, j, jpk - 1);
    }
    for (int k = jpk - 2; k >= 2; k--) {
        const T zws = (k == 2 || k == ikb) ? T(0) : WMASK(T, i, j, k);
        const size_t p = IJK(i, j, k);
        const T x = (zwz[p] - zws * znext) / (T)zwt[p];
        zwz[p] = (pwn[p] * x - up_z(d, pwn, tb, i, j, k)) * WMASK(T, i, j, k);
        znext = x;
    }
}

// (3) Zalesak betas, traadv_fct.F90:354-399 (FP64 arithmetic)
template <typename T>
__device__ __forceinline__ double zbup(const DevDomain &d, const T *pbef, const T *paft, int i, int j, int k)
{
    const double zbig = 1.e+40;
    const double tm = TMASK(double, i, j, k);
    const size_t p = IJK(i, j, k);
    const double a = (double)pbef[p] * tm - zbig * (1.0 - tm);
    const double b = (double)paft[p] * tm - zbig * (1.0 - tm);
    return a > b ? a : b;
}
template <typename T>
__device__ __forceinline__ double zbdo(const DevDomain &d, const T *pbef, const T *paft, int i, int j, int k)
{
    const double zbig = 1.e+40;
    const double tm = TMASK(double, i, j, k);
    const size_t p = IJK(i, j, k);
    const double a = (double)pbef[p] * tm + zbig * (1.0 - tm);
    const double b = (double)paft[p] * tm + zbig * (1.0 - tm);
    return a < b ? a : b;
}
template <typename T, int KB>
struct S_fct_beta {
    static constexpr int NTHR = NH_BLOCK, WAVES = 1;
    __device__ __forceinline__ static void run(const VB &vb_, const DevDomain &d, double p2dt, const T * __restrict__ pbef, const T * __restrict__ paft, const T * __restrict__ paa, const T * __restrict__ pbb, const T * __restrict__ pcc, T * __restrict__ zbetup, T * __restrict__ zbetdo)
    {
        // one thread per interior cell column and block of KB levels
        THREAD_IJ_V(vb_, 2, 2);
        if (i > d.jpi - 1 || j > d.jpj - 1) return;
        const int jpk = d.jpk;
        const int kb0 = (int)vb_.z * KB + 1;
        const int kb1 = kb0 + KB - 1 < jpk - 1 ? kb0 + KB - 1 : jpk - 1;
        const double zrtrn = 1.e-15;
        const size_t ij = IJ(i, j);
        const double e1e2t = d.md.e1e2t[ij];
        // the bounds of the column itself at the levels above, at and below k
        // (above level 1 is level 1 itself, as ikm1 = MAX(1, jk-1))
        double up_m = zbup(d, pbef, paft, i, j, kb0 > 1 ? kb0 - 1 : 1);
        double do_m = zbdo(d, pbef, paft, i, j, kb0 > 1 ? kb0 - 1 : 1);
        double up_c = zbup(d, pbef, paft, i, j, kb0), do_c = zbdo(d, pbef, paft, i, j, kb0);
        double c_k = A3(pcc, i, j, kb0);
    #pragma unroll 1
        for (int k = kb0; k <= kb1; k++) {
            const double up_p = zbup(d, pbef, paft, i, j, k + 1), do_p = zbdo(d, pbef, paft, i, j, k + 1);
            double zup = up_c;
            zup = fmax2(zup, zbup(d, pbef, paft, i - 1, j, k));
            zup = fmax2(zup, zbup(d, pbef, paft, i + 1, j, k));
            zup = fmax2(zup, zbup(d, pbef, paft, i, j - 1, k));
            zup = fmax2(zup, zbup(d, pbef, paft, i, j + 1, k));
            zup = fmax2(zup, up_m);
            zup = fmax2(zup, up_p);
            double zdo = do_c;
            zdo = fmin2(zdo, zbdo(d, pbef, paft, i - 1, j, k));
            zdo = fmin2(zdo, zbdo(d, pbef, paft, i + 1, j, k));
            zdo = fmin2(zdo, zbdo(d, pbef, paft, i, j - 1, k));
            zdo = fmin2(zdo, zbdo(d, pbef, paft, i, j + 1, k));
            zdo = fmin2(zdo, do_m);
            zdo = fmin2(zdo, do_p);
            const double a_im1 = A3(paa, i - 1, j, k), a_i = A3(paa, i, j, k);
            const double b_jm1 = A3(pbb, i, j - 1, k), b_j = A3(pbb, i, j, k);
            const double c_kp1 = A3(pcc, i, j, k + 1);
            const double zpos = fmax2(0.0, a_im1) - fmin2(0.0, a_i) + fmax2(0.0, b_jm1) - fmin2(0.0, b_j) +
                                fmax2(0.0, c_kp1) - fmin2(0.0, c_k);
            const double zneg = fmax2(0.0, a_i) - fmin2(0.0, a_im1) + fmax2(0.0, b_j) - fmin2(0.0, b_jm1) +
                                fmax2(0.0, c_k) - fmin2(0.0, c_kp1);
            const double zbt = e1e2t * E3(d.md.e3t, ij, k) / p2dt;
            const double paft_k = A3(paft, i, j, k);
            const double bu = (zup - paft_k) / (zpos + zrtrn) * zbt;
            const double bd = (paft_k - zdo) / (zneg + zrtrn) * zbt;
            if (sizeof(T) == 4) {
                // FP32 storage: the land sentinels make |beta| ~ 1e64; any value
                // below -1 or above 1 limits identically, so store it finite
                A3(zbetup, i, j, k) = (T)fmin2(fmax2(bu, -1.0e30), 1.0e30);
                A3(zbetdo, i, j, k) = (T)fmin2(fmax2(bd, -1.0e30), 1.0e30);
            } else {
                A3(zbetup, i, j, k) = (T)bu;
                A3(zbetdo, i, j, k) = (T)bd;
            }
            if (k == jpk - 1) {
                A3(zbetup, i, j, jpk) = T(0);
                A3(zbetdo, i, j, jpk) = T(0);
            }
            up_m = up_c;
            up_c = up_p;
            do_m = do_c;
            do_c = do_p;
            c_k = c_kp1;
        }
    }
};
template <typename T, int KB>
KERNEL void k_fct_beta(DevDomain d, double p2dt, const T * __restrict__ pbef, const T * __restrict__ paft, const T * __restrict__ paa, const T * __restrict__ pbb,
                           const T * __restrict__ pcc, T * __restrict__ zbetup, T * __restrict__ zbetdo)
{
    S_fct_beta<T, KB>::run(vb_hw(), d, p2dt, pbef, paft, paa, pbb, pcc, zbetup, zbetdo);
}

// paa at (p,q) after nonosc's limiting and its lbc_lnk('U')
template <typename T>
__device__ __forceinline__ double lim_x(const DevDomain &d, const T *paa, const T *bup, const T *bdo, int p, int q, int k)
{
    int sp, sq;
    const bool ok = lbc_src_safe(d, 'U', p, q, sp, sq);
    const double r = lim_face<T>(A3(paa, sp, sq, k), A3(bdo, sp, sq, k), A3(bup, sp + 1, sq, k), A3(bup, sp, sq, k),
                                 A3(bdo, sp + 1, sq, k));
    return ok ? r : 0.0;
}
template <typename T>
__device__ __forceinline__ double lim_y(const DevDomain &d, const T *pbb, const T *bup, const T *bdo, int p, int q, int k)
{
    int sp, sq;
    const bool ok = lbc_src_safe(d, 'V', p, q, sp, sq);
    const double r = lim_face<T>(A3(pbb, sp, sq, k), A3(bdo, sp, sq, k), A3(bup, sp, sq + 1, k), A3(bup, sp, sq, k),
                                 A3(bdo, sp, sq + 1, k));
    return ok ? r : 0.0;
}
// pcc at level kk (2..jpk) after limiting: set by the k = kk-1 iteration
template <typename T>
__device__ __forceinline__ double lim_z(const DevDomain &d, const T *pcc, const T *bup, const T *bdo, int i, int j, int kk)
{
    if (kk == 1) return A3(pcc, i, j, 1);
    const int k = kk - 1;
    return lim_face<T>(A3(pcc, i, j, kk), A3(bdo, i, j, k + 1), A3(bup, i, j, k), A3(bup, i, j, k + 1),
                       A3(bdo, i, j, k));
}

// (4) limited fluxes and final trend, traadv_fct.F90:404-426 + 288-297
template <typename T>
struct S_fct_final {
    static constexpr int NTHR = NH_BLOCK, WAVES = 1;
    __device__ __forceinline__ static void run(const VB &vb_, const DevDomain &d, const T * __restrict__ paa, const T * __restrict__ pbb, const T * __restrict__ pcc, const T * __restrict__ bup, const T * __restrict__ bdo, T * __restrict__ ta)
    {
        // one thread per interior column, k-marched: the limited vertical flux of
        // the lower face is carried to the next level
        THREAD_IJ_V(vb_, 2, 2);
        if (i > d.jpi - 1 || j > d.jpj - 1) return;
        const Metrics<T> &m = met<T>(d);
        const size_t ij = IJ(i, j);
        const T r1_e1e2t = m.r1_e1e2t[ij];
        double lz_k = lim_z(d, pcc, bup, bdo, i, j, 1);
        for (int k = 1; k <= d.jpk - 1; k++) {
            const double lz_kp1 = lim_z(d, pcc, bup, bdo, i, j, k + 1);
            const T lx_i = (T)lim_x(d, paa, bup, bdo, i, j, k), lx_im1 = (T)lim_x(d, paa, bup, bdo, i - 1, j, k);
            const T ly_j = (T)lim_y(d, pbb, bup, bdo, i, j, k), ly_jm1 = (T)lim_y(d, pbb, bup, bdo, i, j - 1, k);
            const size_t p = IJK(i, j, k);
            ta[p] = ta[p] - (lx_i - lx_im1 + ly_j - ly_jm1 + (T)lz_k - (T)lz_kp1) * r1_e1e2t / E3(m.e3t, ij, k);
            lz_k = lz_kp1;
        }
    }
};
template <typename T>
KERNEL void k_fct_final(DevDomain d, const T * __restrict__ paa, const T * __restrict__ pbb, const T * __restrict__ pcc, const T * __restrict__ bup, const T * __restrict__ bdo, T * __restrict__ ta)
{
    S_fct_final<T>::run(vb_hw(), d, paa, pbb, pcc, bup, bdo, ta);
}

// (3)+(4) fused for closed domains: the Zalesak betas and the limited
// divergence in one k-march per 64xTY tile.  The bounds zbup/zbdo (FP64) of
// the tile and a 2-point halo, for three levels, and the betas of the tile
// and a 1-point halo, for two levels, live in LDS, so the betas never go
// through HBM.  Every global load is issued one level ahead into registers
// and consumed after the next barrier, so the loads of the march overlap
// its LDS work.  Each plane is fetched once per tile: the antidiffusive
// fluxes the limited divergence needs are passed on through LDS by the
// threads that loaded them for the betas, the vertical flux of level k+1
// stays in a register as the beta's flux of level k, and (FP32 storage) the
// first guess of the betas comes from the plane fetched for the bounds.
// Every value is the one k_fct_beta and k_fct_final compute (same
// expressions, same rounding to the storage type), and the ring values
// lbc_lnk gives a closed domain (zero betas and limited fluxes) are selected
// where they apply.
// k_fct_lim tiles: LIM_TX x TY cells per workgroup of LIM_TX x WY threads,
// so each thread owns TY/WY cells of its i column (rows ty, ty+WY, ...);
// taller tiles spread the halo of the bounds (2 cells) and of the betas (1
// cell) over more cells and do more work between two barriers of the march
#define LIM_TX 64
template <typename T> struct FctLimTY {
    // 64x8 tiles for FP32, 64x4 for FP64, 256 threads each (512 threads per
    // 64x8 FP32 tile, two workgroups per CU, measured the same)
    static constexpr int value = sizeof(T) == 4 ? 8 : 4;
    static constexpr int nth = 256;
};
template <typename T, int TY, int NTH = NH_BLOCK>
struct S_fct_lim {
    static constexpr int NTHR = NTH, WAVES = NTH == 512 ? 4 : 3;
    __device__ __forceinline__ static void run(const VB &vb_, const DevDomain &d, double p2dt, const T *__restrict__ pbef_, const T *__restrict__ paft_, const T *__restrict__ paa_, const T *__restrict__ pbb_, const T *__restrict__ pcc_, T *__restrict__ ta_)
    {
        // bounds: FP64 with the reference's +-1e40 land sentinels; with FP32
        // storage they are exact in FP32 (max/min of stored values) and the
        // sentinels become +-inf, which give the same betas after the +-1e30
        // clamp of the FP32 form
        using ZB = typename std::conditional<sizeof(T) == 4, float, double>::type;
        const ZB zsent = sizeof(T) == 4 ? (ZB)INFINITY : (ZB)1.e+40;
        constexpr int ZX = LIM_TX + 4, ZY = TY + 4, BX = LIM_TX + 2, BY = TY + 2;
        constexpr int NZ = (ZX * ZY + NTH - 1) / NTH, NB = (BX * BY + NTH - 1) / NTH;
        constexpr int WY = NTH / LIM_TX, NR = TY / WY;   // workgroup rows, cells per thread
        static_assert(TY % WY == 0, "tile height must be a multiple of the workgroup height");
        // the first guess through LDS for FP32 storage only: with FP64 the
        // extra plane would cost a workgroup per CU
        constexpr bool AFT_LDS = sizeof(T) == 4;
        __shared__ ZB zup_s[3][ZY][ZX], zdo_s[3][ZY][ZX];
        __shared__ T bup_s[2][BY][BX], bdo_s[2][BY][BX];
        __shared__ T aft_s[AFT_LDS ? 3 : 1][AFT_LDS ? BY : 1][AFT_LDS ? BX : 1];
        // paa, pbb of level k and pcc of level k+1 over the beta region, for
        // the limited divergence of level k
        __shared__ T faa_s[BY][BX], fbb_s[BY][BX], fcc_s[BY][BX];
        // vb_.z selects the tracer of the pair
        const size_t o = (size_t)vb_.z * d.n3;
        const T *pbef = pbef_ + o, *paft = paft_ + o, *paa = paa_ + o, *pbb = pbb_ + o, *pcc = pcc_ + o;
        T *ta = ta_ + o;
        unsigned bx, by;
        xcd_tile_v(vb_, bx, by);
        const int tx = (int)threadIdx.x, ty = (int)threadIdx.y, tid = ty * LIM_TX + tx;
        const int i0 = 2 + (int)bx * LIM_TX, j0 = 2 + (int)by * TY;
        const int i = i0 + tx;
        const int jpi = d.jpi, jpj = d.jpj, jpk = d.jpk;
        const double zrtrn = 1.e-15;
        const Metrics<T> &m = met<T>(d);

        // ---- this thread's points of the bounds region (tile + 2)
        // Every global load of the march is unconditional, from an address
        // clamped into the domain (values of points outside it are loaded and
        // ignored), so the compiler can count the loads in flight and wait for a
        // prefetched level without waiting for the next one.
        int zpx[NZ], zpy[NZ];
        bool zin[NZ], zon[NZ];
        T ztb[NZ], zzw[NZ];
        uint32_t zkt[NZ], zof[NZ];   // packed tmask interval (loaded once), in-plane offset
    #pragma unroll
        for (int r = 0; r < NZ; r++) {
            const int q = tid + r * NTH;
            zon[r] = q < ZX * ZY;
            zpx[r] = q % ZX;
            zpy[r] = q / ZX;
            const int pi = i0 - 2 + zpx[r], pj = j0 - 2 + zpy[r];
            zin[r] = zon[r] && pi >= 1 && pi <= jpi && pj >= 1 && pj <= jpj;
            zof[r] = IJ32(min(max(pi, 1), jpi), min(max(pj, 1), jpj));
            zkt[r] = zin[r] ? d.ktu[zof[r]] : 0u;
        }
        // tmask(kk) of zb point r from its interval (TMASK without a load)
        auto zwet = [&](int r, int kk) { return kk >= (int)(zkt[r] & 255u) && kk <= (int)((zkt[r] >> 8) & 255u); };
        auto zb_fetch = [&](int kk) {
            const int kc = min(kk, jpk);
            const T *pb = plane_ptr(pbef, d.n2, kc), *pa = plane_ptr(paft, d.n2, kc);
    #pragma unroll
            for (int r = 0; r < NZ; r++) {
                ztb[r] = ldp(pb, zof[r]);
                zzw[r] = ldp(pa, zof[r]);
            }
        };
        // zbup / zbdo of traadv_fct.F90:361-364 from the fetched values: with
        // tmask 0 or 1, b*tmask -+ zbig*(1-tmask) is b or -+zbig exactly
        auto zb_store = [&](int kk) {
            const int sl = kk % 3;
    #pragma unroll
            for (int r = 0; r < NZ; r++) {
                if (!zon[r]) continue;
                ZB u = 0, w = 0;
                if (zin[r] && kk <= jpk) {
                    if constexpr (sizeof(T) == 4) {
                        const bool wet = zwet(r, kk);
                        const ZB a = (ZB)ztb[r], b = (ZB)zzw[r];
                        u = wet ? (a > b ? a : b) : -zsent;
                        w = wet ? (a < b ? a : b) : zsent;
                    } else {
                        const double tm = zwet(r, kk) ? 1.0 : 0.0;
                        const double au = (double)ztb[r] * tm - zsent * (1.0 - tm);
                        const double bu = (double)zzw[r] * tm - zsent * (1.0 - tm);
                        u = au > bu ? au : bu;
                        const double ad = (double)ztb[r] * tm + zsent * (1.0 - tm);
                        const double bd = (double)zzw[r] * tm + zsent * (1.0 - tm);
                        w = ad < bd ? ad : bd;
                    }
                }
                zup_s[sl][zpy[r]][zpx[r]] = u;
                zdo_s[sl][zpy[r]][zpx[r]] = w;
                if constexpr (AFT_LDS) {
                    const int bx_ = zpx[r] - 1, by_ = zpy[r] - 1;
                    if (bx_ >= 0 && bx_ < BX && by_ >= 0 && by_ < BY) aft_s[sl][by_][bx_] = zzw[r];
                }
            }
        };

        // ---- this thread's points of the beta region (tile + 1)
        int bpx[NB], bpy[NB];
        bool bon[NB], bin[NB];
        uint32_t bof[NB];
        T fa0[NB], fa1[NB], fb0[NB], fb1[NB], fc0[NB], fc1[NB], fzw[NB];
        T fa1_prev[NB], fb1_prev[NB];   // paa, pbb of the level before
        double fe1e2t[NB], fe3t[NB];
    #pragma unroll
        for (int r = 0; r < NB; r++) {
            const int q = tid + r * NTH;
            bon[r] = q < BX * BY;
            bpx[r] = q % BX;
            bpy[r] = q / BX;
            const int pi = i0 - 1 + bpx[r], pj = j0 - 1 + bpy[r];
            bin[r] = bon[r] && pi >= 2 && pi <= jpi - 1 && pj >= 2 && pj <= jpj - 1;
            bof[r] = IJ32(min(max(pi, 2), jpi - 1), min(max(pj, 2), jpj - 1));
            fe1e2t[r] = bin[r] ? d.md.e1e2t[bof[r]] : 0.0;
            fa1_prev[r] = fb1_prev[r] = T(0);
        }
        // first: level kk is the first, so pcc(kk) is fetched too; after
        // it, pcc(kk) is the pcc(kk+1) fetched for level kk-1 (levels up to
        // jpk - 1 are fetched unclamped, in order)
        auto beta_fetch = [&](int kk, bool first) {
            const int kc = min(kk, jpk - 1);
            const T *pa = plane_ptr(paa, d.n2, kc), *pb = plane_ptr(pbb, d.n2, kc), *pc = plane_ptr(pcc, d.n2, kc);
            const T *pc1 = plane_ptr(pcc, d.n2, kc + 1), *pz = plane_ptr(paft, d.n2, kc);
            const uint32_t w = (uint32_t)jpi;
    #pragma unroll
            for (int r = 0; r < NB; r++) {
                const uint32_t q = bof[r];
                fa0[r] = ldp(pa, q - 1u);
                fa1[r] = ldp(pa, q);
                fb0[r] = ldp(pb, q - w);
                fb1[r] = ldp(pb, q);
                fc0[r] = first ? ldp(pc, q) : fc1[r];
                fc1[r] = ldp(pc1, q);
                if constexpr (!AFT_LDS) fzw[r] = ldp(pz, q);
                fe3t[r] = E3(d.md.e3t, q, kc);
            }
        };
        // nonosc betas, lines 366-399, then the closed-ring lbc_lnk (zero off
        // the interior and at level jpk)
        // also hands paa, pbb of level kk-1 and pcc of level kk to the
        // limited divergence of level kk-1 (faa_s, fbb_s, fcc_s)
        auto beta_make = [&](int kk) {
            const int sb = kk & 1;
            const int sa = (kk > 1 ? kk - 1 : 1) % 3, sc = kk % 3, sn = (kk + 1) % 3;
    #pragma unroll
            for (int r = 0; r < NB; r++) {
                if (!bon[r]) continue;
                faa_s[bpy[r]][bpx[r]] = fa1_prev[r];
                fbb_s[bpy[r]][bpx[r]] = fb1_prev[r];
                fcc_s[bpy[r]][bpx[r]] = fc0[r];
                fa1_prev[r] = fa1[r];
                fb1_prev[r] = fb1[r];
                T bu_t = T(0), bd_t = T(0);
                if (bin[r] && kk <= jpk - 1) {
                    const int zx = bpx[r] + 1, zy = bpy[r] + 1;
                    ZB zup = zup_s[sc][zy][zx];
                    zup = fmax2(zup, zup_s[sc][zy][zx - 1]);
                    zup = fmax2(zup, zup_s[sc][zy][zx + 1]);
                    zup = fmax2(zup, zup_s[sc][zy - 1][zx]);
                    zup = fmax2(zup, zup_s[sc][zy + 1][zx]);
                    zup = fmax2(zup, zup_s[sa][zy][zx]);
                    zup = fmax2(zup, zup_s[sn][zy][zx]);
                    ZB zdo = zdo_s[sc][zy][zx];
                    zdo = fmin2(zdo, zdo_s[sc][zy][zx - 1]);
                    zdo = fmin2(zdo, zdo_s[sc][zy][zx + 1]);
                    zdo = fmin2(zdo, zdo_s[sc][zy - 1][zx]);
                    zdo = fmin2(zdo, zdo_s[sc][zy + 1][zx]);
                    zdo = fmin2(zdo, zdo_s[sa][zy][zx]);
                    zdo = fmin2(zdo, zdo_s[sn][zy][zx]);
                    const double a_im1 = fa0[r], a_i = fa1[r], b_jm1 = fb0[r], b_j = fb1[r];
                    const double c_k = fc0[r], c_kp1 = fc1[r];
                    const double zpos = fmax2(0.0, a_im1) - fmin2(0.0, a_i) + fmax2(0.0, b_jm1) - fmin2(0.0, b_j) +
                                        fmax2(0.0, c_kp1) - fmin2(0.0, c_k);
                    const double zneg = fmax2(0.0, a_i) - fmin2(0.0, a_im1) + fmax2(0.0, b_j) - fmin2(0.0, b_jm1) +
                                        fmax2(0.0, c_k) - fmin2(0.0, c_kp1);
                    const double zbt = fe1e2t[r] * fe3t[r] / p2dt;
                    double paft_k;
                    if constexpr (AFT_LDS) paft_k = aft_s[sc][bpy[r]][bpx[r]];
                    else paft_k = fzw[r];
                    const double bu = ((double)zup - paft_k) / (zpos + zrtrn) * zbt;
                    const double bd = (paft_k - (double)zdo) / (zneg + zrtrn) * zbt;
                    if (sizeof(T) == 4) {
                        bu_t = (T)fmin2(fmax2(bu, -1.0e30), 1.0e30);
                        bd_t = (T)fmin2(fmax2(bd, -1.0e30), 1.0e30);
                    } else {
                        bu_t = (T)bu;
                        bd_t = (T)bd;
                    }
                }
                bup_s[sb][bpy[r]][bpx[r]] = bu_t;
                bdo_s[sb][bpy[r]][bpx[r]] = bd_t;
            }
        };

        // ---- this thread's cells: the limited divergence (lines 404-422, 288-297)
        int jr[NR];
        bool own[NR];
        uint32_t ijr[NR], fof[NR];
        T gta[NR], ge3t[NR], r1_e1e2t[NR];
        double lz_k[NR];
    #pragma unroll
        for (int r = 0; r < NR; r++) {
            jr[r] = j0 + ty + WY * r;
            own[r] = (i <= jpi - 1 && jr[r] <= jpj - 1);
            ijr[r] = own[r] ? IJ32(i, jr[r]) : 0u;
            fof[r] = IJ32(min(i, jpi - 1), min(jr[r], jpj - 1));
            r1_e1e2t[r] = own[r] ? m.r1_e1e2t[ijr[r]] : T(0);
            lz_k[r] = own[r] ? (double)A3(pcc, i, jr[r], 1) : 0.0;
            gta[r] = ge3t[r] = T(1);
        }
        auto fin_fetch = [&](int k) {
            const int kc = min(k, jpk - 1);
            const T *pt = plane_ptr((const T *)ta, d.n2, kc);
    #pragma unroll
            for (int r = 0; r < NR; r++) {
                const uint32_t q = fof[r];
                gta[r] = ldp(pt, q);
                ge3t[r] = E3(m.e3t, q, kc);
            }
        };
        const int cx = tx + 1;
        const bool okxm = i - 1 >= 2;
        auto fin_do = [&](int k) {
            const int s0 = k & 1, s1 = (k + 1) & 1;
    #pragma unroll
            for (int r = 0; r < NR; r++) {
                if (!own[r]) continue;
                const int j = jr[r], cy = ty + WY * r + 1;
                const bool okym = j - 1 >= 2;
                const double bu_c = bup_s[s0][cy][cx], bd_c = bdo_s[s0][cy][cx];
                const double lx_i =
                    lim_face<T>(faa_s[cy][cx], bd_c, bup_s[s0][cy][cx + 1], bu_c, bdo_s[s0][cy][cx + 1]);
                const double lx_im1 =
                    okxm ? lim_face<T>(faa_s[cy][cx - 1], bdo_s[s0][cy][cx - 1], bu_c, bup_s[s0][cy][cx - 1], bd_c)
                         : 0.0;
                const double ly_j =
                    lim_face<T>(fbb_s[cy][cx], bd_c, bup_s[s0][cy + 1][cx], bu_c, bdo_s[s0][cy + 1][cx]);
                const double ly_jm1 =
                    okym ? lim_face<T>(fbb_s[cy - 1][cx], bdo_s[s0][cy - 1][cx], bu_c, bup_s[s0][cy - 1][cx], bd_c)
                         : 0.0;
                const double lz_kp1 = lim_face<T>(fcc_s[cy][cx], bdo_s[s1][cy][cx], bu_c, bup_s[s1][cy][cx], bd_c);
                A3(ta, i, j, k) = gta[r] - ((T)lx_i - (T)lx_im1 + (T)ly_j - (T)ly_jm1 + (T)lz_k[r] - (T)lz_kp1) *
                                               r1_e1e2t[r] / ge3t[r];
                lz_k[r] = lz_kp1;
            }
        };

        // ---- prologue: bounds of levels 1-3, betas of level 1, loads of the
        // first iteration in flight
        zb_fetch(1);
        zb_store(1);
        zb_fetch(2);
        zb_store(2);
        zb_fetch(3);
        zb_store(3);
        beta_fetch(1, true);
        __syncthreads();
        beta_make(1);
        beta_fetch(2, false);
        fin_fetch(1);
        zb_fetch(4);
        __syncthreads();
        // ---- march: at the top of iteration k the LDS holds the bounds of
        // levels k..k+2 and the betas of level k; the registers hold the loads
        // for the betas of level k+1, the cell of level k and the bounds of
        // level k+3
        for (int k = 1; k <= jpk - 1; k++) {
            beta_make(k + 1);
            beta_fetch(k + 2, false);
            __syncthreads();
            fin_do(k);
            fin_fetch(k + 1);
            zb_store(k + 3);
            zb_fetch(k + 4);
            __syncthreads();
        }
    }
};
template <typename T, int TY, int NTH = NH_BLOCK>
__global__ __launch_bounds__(NTH, (S_fct_lim<T, TY, NTH>::WAVES)) void k_fct_lim(DevDomain d, double p2dt, const T *__restrict__ pbef_, const T *__restrict__ paft_,
                      const T *__restrict__ paa_, const T *__restrict__ pbb_, const T *__restrict__ pcc_,
                      T *__restrict__ ta_)
{
    S_fct_lim<T, TY, NTH>::run(vb_hw(), d, p2dt, pbef_, paft_, paa_, pbb_, pcc_, ta_);
}

template <typename T>
int launch_tra_adv_fct(double p2dt, const T *pun, const T *pvn, const T *pwn, const T *ptb, const T *ptn, T *pta,
                       int kjpt, int kn_fct_h, int kn_fct_v)
{
    const DevDomain &d = nh::dom();
    if (d.jpk <= STEP_MAXK)
        return launch_fct_pairs<0, T>(d, p2dt, pun, pvn, pwn, ptb, ptn, pta, kjpt, kn_fct_h, kn_fct_v, 0);
    T *zwi = (T *)nh::scratch3(nh::S_ZWI), *zwx = (T *)nh::scratch3(nh::S_ZWX), *zwy = (T *)nh::scratch3(nh::S_ZWY);
    T *zwz = (T *)nh::scratch3(nh::S_ZWZ), *zbu = (T *)nh::scratch3(nh::S_ZBETUP), *zbd = (T *)nh::scratch3(nh::S_ZBETDO);
    double *zwt = (double *)nh::scratch3(nh::S_ZTW);
    Grid2 gi = grid2(d.jpi - 2, d.jpj - 2), gh = grid2(d.jpi - 1, d.jpj - 1);
    Grid2 gk = gi;
    gk.grid.z = (unsigned)((d.jpk - 1 + FctKB<T>::beta - 1) / FctKB<T>::beta);
    int st;
    for (int jn = 0; jn < kjpt; jn++) {
        const T *tb = ptb + (size_t)jn * d.n3, *tn = ptn + (size_t)jn * d.n3;
        T *ta = pta + (size_t)jn * d.n3;
        LAUNCH("tra_adv_fct", (k_fct_low<T>), gi, d, p2dt, pun, pvn, pwn, tb, ta, zwi);
        if ((st = launch_lbc<T>(zwi, d.jpk, 'T', T(0)))) return st;
        LAUNCH("tra_adv_fct", (k_fct_ad<T>), gh, d, kn_fct_h, kn_fct_v, pun, pvn, pwn, tb, tn, zwx, zwy, zwz, zwt);
        if ((st = launch_lbc<T>(zwx, d.jpk, 'U', T(0)))) return st;
        if ((st = launch_lbc<T>(zwy, d.jpk, 'V', T(0)))) return st;
        if ((st = launch_lbc<T>(zwz, d.jpk, 'W', T(0)))) return st;
        LAUNCHB("tra_adv_fct", (k_fct_beta<T, FctKB<T>::beta>), (S_fct_beta<T, FctKB<T>::beta>), gk, d, p2dt, tb, zwi, zwx, zwy, zwz, zbu, zbd);
        if ((st = launch_lbc<T>(zbu, d.jpk, 'T', T(0)))) return st;
        if ((st = launch_lbc<T>(zbd, d.jpk, 'T', T(0)))) return st;
        LAUNCHB("tra_adv_fct", (k_fct_final<T>), (S_fct_final<T>), gi, d, zwx, zwy, zwz, zbu, zbd, ta);
    }
    return 0;
}
template int launch_tra_adv_fct<double>(double, const double *, const double *, const double *, const double *,
                                        const double *, double *, int, int, int);
// ---------------------------------------------------------------------------
// the FCT limiter frozen (the tangent linear of the tracer step): the
// factors nonosc multiplies the antidiffusive fluxes by (traadv_fct.F90:
// 404-426), and the FCT trend with given factors.  One tracer, FP64.
// ---------------------------------------------------------------------------
__device__ __forceinline__ double lim_factor(double flux, double bdo_here, double bup_next, double bup_here,
                                             double bdo_next)
{
    const double za = fmin2(fmin2(1.0, bdo_here), bup_next);
    const double zb = fmin2(fmin2(1.0, bup_here), bdo_next);
    const double zc = (0.5 + nsign(0.5, flux));
    return zc * za + (1.0 - zc) * zb;
}

KERNEL void k_fct_coef(DevDomain d, const double *__restrict__ paa, const double *__restrict__ pbb,
                       const double *__restrict__ pcc, const double *__restrict__ bup, const double *__restrict__ bdo,
                       double *__restrict__ cx, double *__restrict__ cy, double *__restrict__ cz)
{
    THREAD_IJ(2, 2);
    if (i > d.jpi - 1 || j > d.jpj - 1) return;
    for (int k = 1; k <= d.jpk - 1; k++) {
        const size_t p = IJK(i, j, k);
        cx[p] = lim_factor(paa[p], bdo[p], A3(bup, i + 1, j, k), bup[p], A3(bdo, i + 1, j, k));
        cy[p] = lim_factor(pbb[p], bdo[p], A3(bup, i, j + 1, k), bup[p], A3(bdo, i, j + 1, k));
        cz[IJK(i, j, k + 1)] = lim_factor(A3(pcc, i, j, k + 1), A3(bdo, i, j, k + 1), bup[p], A3(bup, i, j, k + 1),
                                          bdo[p]);
    }
}

// k_fct_final with the factors given: each limited flux is the flux times
// its factor, at the flux's lbc source, as nonosc and its lbc_lnk give it
KERNEL void k_fct_final_frozen(DevDomain d, const double *__restrict__ paa, const double *__restrict__ pbb,
                               const double *__restrict__ pcc, const double *__restrict__ cx,
                               const double *__restrict__ cy, const double *__restrict__ cz, double *__restrict__ ta)
{
    THREAD_IJ(2, 2);
    if (i > d.jpi - 1 || j > d.jpj - 1) return;
    const Metrics<double> &m = d.md;
    const size_t ij = IJ(i, j);
    const double r1_e1e2t = m.r1_e1e2t[ij];
    auto lx = [&](int p, int q, int k) {
        int sp, sq;
        const bool ok = lbc_src_safe(d, 'U', p, q, sp, sq);
        return ok ? A3(paa, sp, sq, k) * A3(cx, sp, sq, k) : 0.0;
    };
    auto ly = [&](int p, int q, int k) {
        int sp, sq;
        const bool ok = lbc_src_safe(d, 'V', p, q, sp, sq);
        return ok ? A3(pbb, sp, sq, k) * A3(cy, sp, sq, k) : 0.0;
    };
    double lz_k = A3(pcc, i, j, 1);
    for (int k = 1; k <= d.jpk - 1; k++) {
        const double lz_kp1 = A3(pcc, i, j, k + 1) * A3(cz, i, j, k + 1);
        const size_t p = IJK(i, j, k);
        ta[p] = ta[p] - (lx(i, j, k) - lx(i - 1, j, k) + ly(i, j, k) - ly(i, j - 1, k) + lz_k - lz_kp1) * r1_e1e2t /
                            E3(m.e3t, ij, k);
        lz_k = lz_kp1;
    }
}

// mode 0: the factors of the state (tb, tn) into cx, cy, cz; mode 1: pta +=
// the FCT trend with the factors cx, cy, cz
int launch_fct_frozen(int mode, double p2dt, const double *pun, const double *pvn, const double *pwn,
                      const double *tb, const double *tn, int kn_h, int kn_v, double *cx, double *cy, double *cz,
                      double *pta)
{
    const DevDomain &d = nh::dom();
    double *zwi = (double *)nh::scratch3(nh::S_ZWI), *zwx = (double *)nh::scratch3(nh::S_ZWX);
    double *zwy = (double *)nh::scratch3(nh::S_ZWY), *zwz = (double *)nh::scratch3(nh::S_ZWZ);
    double *zbu = (double *)nh::scratch3(nh::S_ZBETUP), *zbd = (double *)nh::scratch3(nh::S_ZBETDO);
    double *zwt = (double *)nh::scratch3(nh::S_ZTW), *dummy = (double *)nh::scratch3(nh::S_ZUN);
    Grid2 gi = grid2(d.jpi - 2, d.jpj - 2), gh = grid2(d.jpi - 1, d.jpj - 1);
    Grid2 gk = gi;
    gk.grid.z = (unsigned)((d.jpk - 1 + FctKB<double>::beta - 1) / FctKB<double>::beta);
    int st;
    double *ta = mode == 0 ? dummy : pta;
    if (d.jpk <= STEP_MAXK) {
        // the flux kernels of the tracer-pair chain, for one tracer
        if ((st = launch_fct_fluxes1<double>(d, p2dt, pun, pvn, pwn, tb, tn, ta, kn_h, kn_v, zwi, zwx, zwy, zwz,
                                             "tra_adv_fct")))
            return st;
    } else {
        LAUNCH("tra_adv_fct", (k_fct_low<double>), gi, d, p2dt, pun, pvn, pwn, tb, ta, zwi);
        if ((st = launch_lbc<double>(zwi, d.jpk, 'T', 0.0))) return st;
        LAUNCH("tra_adv_fct", (k_fct_ad<double>), gh, d, kn_h, kn_v, pun, pvn, pwn, tb, tn, zwx, zwy, zwz, zwt);
        if ((st = launch_lbc<double>(zwx, d.jpk, 'U', 0.0))) return st;
        if ((st = launch_lbc<double>(zwy, d.jpk, 'V', 0.0))) return st;
        if ((st = launch_lbc<double>(zwz, d.jpk, 'W', 0.0))) return st;
    }
    if (mode == 1) {
        LAUNCH("tra_adv_fct", k_fct_final_frozen, gi, d, zwx, zwy, zwz, cx, cy, cz, pta);
        return 0;
    }
    LAUNCHB("tra_adv_fct", (k_fct_beta<double, FctKB<double>::beta>), (S_fct_beta<double, FctKB<double>::beta>), gk, d, p2dt, tb, zwi, zwx, zwy, zwz, zbu, zbd);
    if ((st = launch_lbc<double>(zbu, d.jpk, 'T', 0.0))) return st;
    if ((st = launch_lbc<double>(zbd, d.jpk, 'T', 0.0))) return st;
    LAUNCH("tra_adv_fct", k_fct_coef, gi, d, zwx, zwy, zwz, zbu, zbd, cx, cy, cz);
    return 0;
}

template int launch_tra_adv_fct<float>(double, const float *, const float *, const float *, const float *,
                                       const float *, float *, int, int, int);

// ---------------------------------------------------------------------------
// interp_4th_cpt alone (traadv_fct.F90:517-616), interior columns,
// levels 2..jpkm1 of pt_out
// ---------------------------------------------------------------------------
KERNEL void k_interp_4th_cpt(DevDomain d, const double * __restrict__ tin, double * __restrict__ tout, double * __restrict__ zwt)
{
    THREAD_IJ(2, 2);
    if (i > d.jpi - 1 || j > d.jpj - 1) return;
    const int jpk = d.jpk, ikb = A2(d.mbkt, i, j);
    double zws_prev = 0.0, zwt_prev = 1.0, zout_prev = 0.0;
    for (int k = 2; k <= jpk - 1; k++) {
        double zwd, zwi, zws, zwrm;
        if (k == 2 || k == ikb) {
            zwd = 1.0; zwi = 0.0; zws = 0.0;
            zwrm = 0.5 * (A3(tin, i, j, k - 1) + A3(tin, i, j, k));
        } else {
            const double wm = WMASK(double, i, j, k);
            zwd = 3.0 * wm + 1.0; zwi = wm; zws = wm;
            zwrm = 3.0 * wm * (A3(tin, i, j, k) + A3(tin, i, j, k - 1));
        }
        double zwt_k, zout;
        if (k == 2) { zwt_k = zwd; zout = zwrm; }
        else { zwt_k = zwd - zwi * zws_prev / zwt_prev; zout = zwrm - zwi / zwt_prev * zout_prev; }
        A3(zwt, i, j, k) = zwt_k;
        A3(tout, i, j, k) = zout;
        zws_prev = zws; zwt_prev = zwt_k; zout_prev = zout;
    }
    double znext = zout_prev / zwt_prev;
    A3(tout, i, j, jpk - 1) = znext;
    for (int k = jpk - 2; k >= 2; k--) {
        const double zws = (k == 2 || k == ikb) ? 0.0 : WMASK(double, i, j, k);
        const double x = (A3(tout, i, j, k) - zws * znext) / A3(zwt, i, j, k);
        A3(tout, i, j, k) = x;
        znext = x;
    }
}

int launch_interp_4th_cpt(const double *tin, double *tout)
{
    const DevDomain &d = nh::dom();
    Grid2 g = grid2(d.jpi - 2, d.jpj - 2);
    LAUNCH("interp_4th_cpt", k_interp_4th_cpt, g, d, tin, tout, (double *)nh::scratch3(nh::S_ZTW));
    return 0;
}

// ---------------------------------------------------------------------------
// tra_ldf_lap, TRA/traldf_lap_blp.F90:99-145 (ln_zps = F); all tracers
// of the call are updated by the same column thread
// ---------------------------------------------------------------------------
template <typename T>
struct S_tra_ldf_lap {
    static constexpr int NTHR = NH_BLOCK, WAVES = 1;
    __device__ __forceinline__ static void run(const VB &vb_, const DevDomain &d, T zsign, const T * __restrict__ pahu, const T * __restrict__ pahv, const T * __restrict__ ptb, T * __restrict__ pta, int kjpt)
    {
        THREAD_IJ_V(vb_, 2, 2);
        if (i > d.jpi - 1 || j > d.jpj - 1) return;
        const Metrics<T> &m = met<T>(d);
        const size_t ij = IJ(i, j), ijm = IJ(i - 1, j), ijs = IJ(i, j - 1);
        const T e2_e1u = m.e2_e1u[ij], e2_e1u_w = m.e2_e1u[ijm], e1_e2v = m.e1_e2v[ij], e1_e2v_s = m.e1_e2v[ijs];
        const T e1e2t = m.e1e2t[ij];
        for (int k = 1; k <= d.jpk - 1; k++) {
            const T zaheeu = zsign * A3(pahu, i, j, k) * e2_e1u * E3(m.e3u, ij, k);
            const T zaheeu_w = zsign * A3(pahu, i - 1, j, k) * e2_e1u_w * E3(m.e3u, ijm, k);
            const T zaheev = zsign * A3(pahv, i, j, k) * e1_e2v * E3(m.e3v, ij, k);
            const T zaheev_s = zsign * A3(pahv, i, j - 1, k) * e1_e2v_s * E3(m.e3v, ijs, k);
            const T e3t = E3(m.e3t, ij, k);
            for (int jn = 0; jn < kjpt; jn++) {
                const T *tb = ptb + (size_t)jn * d.n3;
                T *ta = pta + (size_t)jn * d.n3;
                const T tc = A3(tb, i, j, k);
                const T ztu = zaheeu * (A3(tb, i + 1, j, k) - tc);
                const T ztu_w = zaheeu_w * (tc - A3(tb, i - 1, j, k));
                const T ztv = zaheev * (A3(tb, i, j + 1, k) - tc);
                const T ztv_s = zaheev_s * (tc - A3(tb, i, j - 1, k));
                const size_t p = IJK(i, j, k);
                ta[p] = ta[p] + (ztu - ztu_w + ztv - ztv_s) / (e1e2t * e3t);
            }
        }
    }
};
template <typename T>
KERNEL void k_tra_ldf_lap(DevDomain d, T zsign, const T * __restrict__ pahu, const T * __restrict__ pahv, const T * __restrict__ ptb, T * __restrict__ pta, int kjpt)
{
    S_tra_ldf_lap<T>::run(vb_hw(), d, zsign, pahu, pahv, ptb, pta, kjpt);
}

template <typename T>
int launch_tra_ldf_lap(const T *pahu, const T *pahv, const T *ptb, T *pta, int kjpt, int kpass)
{
    const DevDomain &d = nh::dom();
    Grid2 g = grid2(d.jpi - 2, d.jpj - 2);
    const T zsign = (kpass == 1) ? T(1) : T(-1);
    LAUNCHB("tra_ldf_lap", k_tra_ldf_lap<T>, (S_tra_ldf_lap<T>), g, d, zsign, pahu, pahv, ptb, pta, kjpt);
    return 0;
}
template int launch_tra_ldf_lap<double>(const double *, const double *, const double *, double *, int, int);
template int launch_tra_ldf_lap<float>(const float *, const float *, const float *, float *, int, int);

// ---------------------------------------------------------------------------
// tra_ldf_iso, TRA/traldf_iso.F90:144-366
// ---------------------------------------------------------------------------
// pass-1 coefficients ah_wslp2 and akz, lines 146-202 (threads 1..jpim1 x 1..jpjm1)
template <typename T>
struct S_ldf_iso_coef {
    static constexpr int NTHR = NH_BLOCK, WAVES = 1;
    __device__ __forceinline__ static void run(const VB &vb_, const DevDomain &d, double z2dt, const T * __restrict__ pahu, const T * __restrict__ pahv, const T * __restrict__ wslpi, const T * __restrict__ wslpj, T * __restrict__ ah_wslp2, T * __restrict__ akz)
    {
        // one thread per (i,j) and level k = vb_.z + 2 (no recurrence in k)
        THREAD_IJ_V(vb_, 1, 1);
        const int k = (int)vb_.z + 2;
        if (i > d.jpi - 1 || j > d.jpj - 1) return;
        const Metrics<T> &m = met<T>(d);
        const bool inner = is_interior(d, i, j);
        const size_t ij = IJ(i, j);
        const T z1_2dt = T(1) / (T)z2dt;
        {
            const size_t p = IJK(i, j, k);
            if (inner) {
                const T wm = WMASK(T, i, j, k);
                const T zmsku = wm / fmax2(UMASK(T, i, j, k - 1) + UMASK(T, i - 1, j, k) + UMASK(T, i - 1, j, k - 1) +
                                               UMASK(T, i, j, k), T(1));
                const T zmskv = wm / fmax2(VMASK(T, i, j, k - 1) + VMASK(T, i, j - 1, k) + VMASK(T, i, j - 1, k - 1) +
                                               VMASK(T, i, j, k), T(1));
                const T zahu_w = (A3(pahu, i, j, k - 1) + A3(pahu, i - 1, j, k) + A3(pahu, i - 1, j, k - 1) +
                                  A3(pahu, i, j, k)) * zmsku;
                const T zahv_w = (A3(pahv, i, j, k - 1) + A3(pahv, i, j - 1, k) + A3(pahv, i, j - 1, k - 1) +
                                  A3(pahv, i, j, k)) * zmskv;
                ah_wslp2[p] = zahu_w * wslpi[p] * wslpi[p] + zahv_w * wslpj[p] * wslpj[p];
            }
            if (d.ln_traldf_msc) {
                T zakz = akz[p];
                if (inner) {
                    const T e1u = m.e1u[ij], e1u_w = m.e1u[IJ(i - 1, j)], e2v = m.e2v[ij], e2v_s = m.e2v[IJ(i, j - 1)];
                    zakz = T(0.25) * ((A3(pahu, i, j, k) + A3(pahu, i, j, k - 1)) / (e1u * e1u) +
                                      (A3(pahu, i - 1, j, k) + A3(pahu, i - 1, j, k - 1)) / (e1u_w * e1u_w) +
                                      (A3(pahv, i, j, k) + A3(pahv, i, j, k - 1)) / (e2v * e2v) +
                                      (A3(pahv, i, j - 1, k) + A3(pahv, i, j - 1, k - 1)) / (e2v_s * e2v_s));
                }
                const T e3w = E3(m.e3w, ij, k);
                const T ah = ah_wslp2[p];
                if (!d.ln_traldf_lap) {
                    akz[p] = T(16) * ah * (zakz + ah / (e3w * e3w));
                } else {
                    const T ze3w_2 = e3w * e3w;
                    const T zcoef0 = (T)z2dt * (zakz + ah / ze3w_2);
                    akz[p] = fmax2(zcoef0 - T(0.5), T(0)) * ze3w_2 * z1_2dt;
                }
            }
        }
    }
};
template <typename T>
KERNEL void k_ldf_iso_coef(DevDomain d, double z2dt, const T * __restrict__ pahu, const T * __restrict__ pahv, const T * __restrict__ wslpi,
                               const T * __restrict__ wslpj, T * __restrict__ ah_wslp2, T * __restrict__ akz)
{
    S_ldf_iso_coef<T>::run(vb_hw(), d, z2dt, pahu, pahv, wslpi, wslpj, ah_wslp2, akz);
}

// Per column the kernel marches down k with a register window: level k+1 of
// tb is loaded once for the five columns (C, W, E, S, N) the stencil touches,
// and every intermediate of the reference loop nests -- the masked gradients
// zdit/zdjt (lines 218-225), the vertical gradients zdkt/zdk1t (250-254), the
// w-point flux ztfw (298-357) -- is formed once and carried to the next level.
// The values, and the order in which each expression combines them, are those
// of the reference, so the FP64 form is bit-identical to it.  NT tracers share
// one pass over the coefficient arrays.
//
// MODE 0: ln_traldf_lap (line 323); 1: bilaplacian first pass (kpass 1, line
// 335); 2: bilaplacian second pass (kpass 2, line 341, reads ptbb).
struct IsoCols {
    uint32_t c, w, e, s, n;
    int t0[5], t1[5];            // t-intervals of C W E S N (w-masks)
    int u0[2], u1[2], v0[2], v1[2];   // u-intervals of C W, v-intervals of C S
};
__device__ __forceinline__ bool kin(int k, int a, int b) { return k >= a && k <= b; }
template <typename T>
__device__ __forceinline__ T wmk(const IsoCols &c, int q, int k)
{
    return (kin(k, c.t0[q], c.t1[q]) && (k == 1 || kin(k - 1, c.t0[q], c.t1[q]))) ? T(1) : T(0);
}
template <typename T>
__device__ __forceinline__ T umk(const IsoCols &c, int q, int k) { return kin(k, c.u0[q], c.u1[q]) ? T(1) : T(0); }
template <typename T>
__device__ __forceinline__ T vmk(const IsoCols &c, int q, int k) { return kin(k, c.v0[q], c.v1[q]) ? T(1) : T(0); }

enum { QC = 0, QW = 1, QE = 2, QS = 3, QN = 4 };

// NT tracers per thread; TSPLIT > 1 spreads TSPLIT tracers over
// neighbouring workgroups instead (the tracer is the lowest digit of the
// XCD-ordered tile index, so the workgroups of one tile run on one XCD at
// the same time and share the coefficient loads in its L2)
template <typename T, int MODE, int NT, int TSPLIT>
struct S_tra_ldf_iso {
    static constexpr int NTHR = NH_BLOCK, WAVES = 1;
    __device__ __forceinline__ static void run(const VB &vb_, const DevDomain &d, T zsign, const T * __restrict__ pahu, const T * __restrict__ pahv, const T * __restrict__ uslp, const T * __restrict__ vslp, const T * __restrict__ wslpi, const T * __restrict__ wslpj, const T * __restrict__ ah_wslp2, const T * __restrict__ akz, const T * __restrict__ ptb, const T * __restrict__ ptbb, T * __restrict__ pta)
    {
        unsigned bx_, by_;
        xcd_tile_v(vb_, bx_, by_);
        if (TSPLIT > 1) {
            const size_t toff = (size_t)(bx_ % TSPLIT) * d.n3;
            ptb += toff;
            ptbb += toff;
            pta += toff;
            bx_ /= TSPLIT;
        }
        const int i = (int)(bx_ * blockDim.x + threadIdx.x) + 2;
        const int j = (int)(by_ * blockDim.y + threadIdx.y) + 2;
        if (i > d.jpi - 1 || j > d.jpj - 1) return;
        const Metrics<T> &m = met<T>(d);
        const size_t n3 = d.n3, plane = d.n2;
        const int jpk = d.jpk;
        IsoCols c;
        c.c = IJ32(i, j); c.w = IJ32(i - 1, j); c.e = IJ32(i + 1, j); c.s = IJ32(i, j - 1); c.n = IJ32(i, j + 1);
        const uint32_t col[5] = {c.c, c.w, c.e, c.s, c.n};
    #pragma unroll
        for (int q = 0; q < 5; q++) { c.t0[q] = d.kT.k0[col[q]]; c.t1[q] = d.kT.k1[col[q]]; }
        c.u0[0] = d.kU.k0[c.c]; c.u1[0] = d.kU.k1[c.c]; c.u0[1] = d.kU.k0[c.w]; c.u1[1] = d.kU.k1[c.w];
        c.v0[0] = d.kV.k0[c.c]; c.v1[0] = d.kV.k1[c.c]; c.v0[1] = d.kV.k0[c.s]; c.v1[1] = d.kV.k1[c.s];

        const T r1_e1e2t = m.r1_e1e2t[c.c], e1e2t = m.e1e2t[c.c], e2t = m.e2t[c.c], e1t = m.e1t[c.c];
        const T e2_e1u_c = m.e2_e1u[c.c], e2_e1u_w = m.e2_e1u[c.w], e2u_c = m.e2u[c.c], e2u_w = m.e2u[c.w];
        const T e1_e2v_c = m.e1_e2v[c.c], e1_e2v_s = m.e1_e2v[c.s], e1v_c = m.e1v[c.c], e1v_s = m.e1v[c.s];

        // window state, level k ("now") -- per tracer
        T tb[NT][5], zdit_c[NT], zdit_w[NT], zdjt_c[NT], zdjt_s[NT], dkp[NT][5], tbb_k[NT], ztfw_k[NT];
        T ahu_c = ldp(pahu, c.c), ahu_w = ldp(pahu, c.w), ahv_c = ldp(pahv, c.c), ahv_s = ldp(pahv, c.s);
    #pragma unroll
        for (int t = 0; t < NT; t++) {
            const T *tbt = ptb + (size_t)t * n3;
    #pragma unroll
            for (int q = 0; q < 5; q++) { tb[t][q] = ldp(tbt, col[q]); dkp[t][q] = T(0); }
            zdit_c[t] = (tb[t][QE] - tb[t][QC]) * umk<T>(c, 0, 1);
            zdit_w[t] = (tb[t][QC] - tb[t][QW]) * umk<T>(c, 1, 1);
            zdjt_c[t] = (tb[t][QN] - tb[t][QC]) * vmk<T>(c, 0, 1);
            zdjt_s[t] = (tb[t][QC] - tb[t][QS]) * vmk<T>(c, 1, 1);
            tbb_k[t] = (MODE == 2) ? ldp(ptbb + (size_t)t * n3, c.c) : T(0);
            ztfw_k[t] = T(0);
        }
        // Every load of iteration k is issued one iteration ahead (Lv), without
        // conditions, so the march waits for a level's data while the next
        // level's loads are in flight.  All of them stay inside the arrays:
        // iteration k reads levels k and k+1 <= jpk.
        struct Lv {
            T ahu_c1, ahu_w1, ahv_c1, ahv_s1, us_c, us_w, vs_c, vs_s;
            T e3t, e3u_c, e3u_w, e3v_c, e3v_s, e3w1, wsi, wsj, ahw, akw;
            T tb1[NT][5], tbb1[NT], ta[NT];
        };
        auto fetch = [&](int k, Lv &v) {
            const T *ahu1 = plane_ptr(pahu, plane, k + 1), *ahv1 = plane_ptr(pahv, plane, k + 1);
            v.ahu_c1 = ldp(ahu1, c.c); v.ahu_w1 = ldp(ahu1, c.w);
            v.ahv_c1 = ldp(ahv1, c.c); v.ahv_s1 = ldp(ahv1, c.s);
            const T *usk = plane_ptr(uslp, plane, k), *vsk = plane_ptr(vslp, plane, k);
            v.us_c = ldp(usk, c.c); v.us_w = ldp(usk, c.w); v.vs_c = ldp(vsk, c.c); v.vs_s = ldp(vsk, c.s);
            v.e3t = E3(m.e3t, c.c, k); v.e3u_c = E3(m.e3u, c.c, k); v.e3u_w = E3(m.e3u, c.w, k);
            v.e3v_c = E3(m.e3v, c.c, k); v.e3v_s = E3(m.e3v, c.s, k);
            v.e3w1 = E3(m.e3w, c.c, k + 1);
            v.wsi = ldp(plane_ptr(wslpi, plane, k + 1), c.c);
            v.wsj = ldp(plane_ptr(wslpj, plane, k + 1), c.c);
            v.ahw = ldp(plane_ptr(ah_wslp2, plane, k + 1), c.c);
            v.akw = (MODE != 1) ? ldp(plane_ptr(akz, plane, k + 1), c.c) : T(0);
    #pragma unroll
            for (int t = 0; t < NT; t++) {
                const T *tb1p = plane_ptr(ptb + (size_t)t * n3, plane, k + 1);
    #pragma unroll
                for (int q = 0; q < 5; q++) v.tb1[t][q] = ldp(tb1p, col[q]);
                v.tbb1[t] = (MODE == 2) ? ldp(plane_ptr(ptbb + (size_t)t * n3, plane, k + 1), c.c) : T(0);
                v.ta[t] = ldp((const T *)plane_ptr(pta + (size_t)t * n3, plane, k), c.c);
            }
        };
        Lv cur, nxt;
        fetch(1, cur);
        for (int k = 1; k <= jpk - 1; k++) {
            fetch(k + 1 <= jpk - 1 ? k + 1 : jpk - 1, nxt);
            // level k and k+1 planes (k+1 = jpk only feeds terms the reference
            // multiplies by a zero mask or skips)
            const T ahu_c1 = cur.ahu_c1, ahu_w1 = cur.ahu_w1, ahv_c1 = cur.ahv_c1, ahv_s1 = cur.ahv_s1;
            const T us_c = cur.us_c, us_w = cur.us_w, vs_c = cur.vs_c, vs_s = cur.vs_s;
            const T e3t = cur.e3t, e3u_c = cur.e3u_c, e3u_w = cur.e3u_w, e3v_c = cur.e3v_c, e3v_s = cur.e3v_s;
            const bool wlev = (k + 1 <= jpk - 1);       // ztfw(k+1) is a computed level
            const T wsi = wlev ? cur.wsi : T(0), wsj = wlev ? cur.wsj : T(0), ahw = wlev ? cur.ahw : T(0);
            const T akw = (wlev && MODE != 1) ? cur.akw : T(0), e3w1 = wlev ? cur.e3w1 : T(1);
            // masks at k and k+1
            const T wm_c = wmk<T>(c, QC, k), wm_w = wmk<T>(c, QW, k), wm_e = wmk<T>(c, QE, k), wm_s = wmk<T>(c, QS, k),
                    wm_n = wmk<T>(c, QN, k);
            const T wm_c1 = wmk<T>(c, QC, k + 1), wm_w1 = wmk<T>(c, QW, k + 1), wm_e1 = wmk<T>(c, QE, k + 1),
                    wm_s1 = wmk<T>(c, QS, k + 1), wm_n1 = wmk<T>(c, QN, k + 1);
            const T um_c = umk<T>(c, 0, k), um_w = umk<T>(c, 1, k), vm_c = vmk<T>(c, 0, k), vm_s = vmk<T>(c, 1, k);
            const T um_c1 = umk<T>(c, 0, k + 1), um_w1 = umk<T>(c, 1, k + 1), vm_c1 = vmk<T>(c, 0, k + 1),
                    vm_s1 = vmk<T>(c, 1, k + 1);
            // zftu / zftv mask weights and coefficients, lines 257-276
            const T zmsku_c = T(1) / fmax2(wm_e + wm_c1 + wm_e1 + wm_c, T(1));
            const T zmsku_w = T(1) / fmax2(wm_c + wm_w1 + wm_c1 + wm_w, T(1));
            const T zmskv_c = T(1) / fmax2(wm_n + wm_c1 + wm_n1 + wm_c, T(1));
            const T zmskv_s = T(1) / fmax2(wm_c + wm_s1 + wm_c1 + wm_s, T(1));
            const T zabe1_c = ahu_c * e2_e1u_c * e3u_c, zabe1_w = ahu_w * e2_e1u_w * e3u_w;
            const T zcof1_c = -(ahu_c * e2u_c * us_c * zmsku_c), zcof1_w = -(ahu_w * e2u_w * us_w * zmsku_w);
            const T zabe2_c = ahv_c * e1_e2v_c * e3v_c, zabe2_s = ahv_s * e1_e2v_s * e3v_s;
            const T zcof2_c = -(ahv_c * e1v_c * vs_c * zmskv_c), zcof2_s = -(ahv_s * e1v_s * vs_s * zmskv_s);
            // ztfw(k+1) weights, lines 298-321
            T zcoef3 = T(0), zcoef4 = T(0);
            if (wlev) {
                const T zmsku = wm_c1 / fmax2(um_c + um_w1 + um_w + um_c1, T(1));
                const T zmskv = wm_c1 / fmax2(vm_c + vm_s1 + vm_s + vm_c1, T(1));
                const T zahu_w = (ahu_c + ahu_w1 + ahu_w + ahu_c1) * zmsku;
                const T zahv_w = (ahv_c + ahv_s1 + ahv_s + ahv_c1) * zmskv;
                zcoef3 = -(zahu_w * e2t * zmsku * wsi);
                zcoef4 = -(zahv_w * e1t * zmskv * wsj);
            }
    #pragma unroll
            for (int t = 0; t < NT; t++) {
                T tb1[5], dkn[5];
    #pragma unroll
                for (int q = 0; q < 5; q++) tb1[q] = cur.tb1[t][q];
                // DK(q,k+1) = (tb(k) - tb(k+1)) * wmask(k+1) = zdk1t(q,k) = zdkt(q,k+1)
                dkn[QC] = (tb[t][QC] - tb1[QC]) * wm_c1;
                dkn[QW] = (tb[t][QW] - tb1[QW]) * wm_w1;
                dkn[QE] = (tb[t][QE] - tb1[QE]) * wm_e1;
                dkn[QS] = (tb[t][QS] - tb1[QS]) * wm_s1;
                dkn[QN] = (tb[t][QN] - tb1[QN]) * wm_n1;
                T zdkt[5];
    #pragma unroll
                for (int q = 0; q < 5; q++) zdkt[q] = (k == 1) ? dkn[q] : dkp[t][q];
                const T zftu_c = (zabe1_c * zdit_c[t] + zcof1_c * (zdkt[QE] + dkn[QC] + dkn[QE] + zdkt[QC])) * um_c;
                const T zftu_w = (zabe1_w * zdit_w[t] + zcof1_w * (zdkt[QC] + dkn[QW] + dkn[QC] + zdkt[QW])) * um_w;
                const T zftv_c = (zabe2_c * zdjt_c[t] + zcof2_c * (zdkt[QN] + dkn[QC] + dkn[QN] + zdkt[QC])) * vm_c;
                const T zftv_s = (zabe2_s * zdjt_s[t] + zcof2_s * (zdkt[QC] + dkn[QS] + dkn[QC] + zdkt[QS])) * vm_s;
                // gradients at k+1
                const T zdit_c1 = (tb1[QE] - tb1[QC]) * um_c1, zdit_w1 = (tb1[QC] - tb1[QW]) * um_w1;
                const T zdjt_c1 = (tb1[QN] - tb1[QC]) * vm_c1, zdjt_s1 = (tb1[QC] - tb1[QS]) * vm_s1;
                T ztfw1 = T(0), tbb1 = T(0);
                if (MODE == 2) tbb1 = cur.tbb1[t];
                if (wlev) {
                    ztfw1 = zcoef3 * (zdit_c[t] + zdit_w1 + zdit_w[t] + zdit_c1) +
                            zcoef4 * (zdjt_c[t] + zdjt_s1 + zdjt_s[t] + zdjt_c1);
                    const T dt = tb[t][QC] - tb1[QC];
                    if (MODE == 0) ztfw1 = ztfw1 + e1e2t / e3w1 * wm_c1 * (ahw - akw) * dt;
                    else if (MODE == 1) ztfw1 = ztfw1 + ahw * e1e2t * dt / e3w1 * wm_c1;
                    else ztfw1 = ztfw1 + e1e2t / e3w1 * wm_c1 * (ahw * dt + akw * (tbb_k[t] - tbb1));
                }
                const T hdiv = zftu_c - zftu_w + zftv_c - zftv_s;
                T *tak = plane_ptr(pta + (size_t)t * n3, plane, k);
                T a = cur.ta[t];
                a = a + zsign * hdiv * r1_e1e2t / e3t;
                a = a + zsign * (ztfw_k[t] - ztfw1) * r1_e1e2t / e3t;
                stp(tak, c.c, a);
                // roll the window
    #pragma unroll
                for (int q = 0; q < 5; q++) { tb[t][q] = tb1[q]; dkp[t][q] = dkn[q]; }
                zdit_c[t] = zdit_c1; zdit_w[t] = zdit_w1; zdjt_c[t] = zdjt_c1; zdjt_s[t] = zdjt_s1;
                ztfw_k[t] = ztfw1;
                tbb_k[t] = tbb1;
            }
            ahu_c = ahu_c1; ahu_w = ahu_w1; ahv_c = ahv_c1; ahv_s = ahv_s1;
            cur = nxt;
        }
    }
};
template <typename T, int MODE, int NT, int TSPLIT>
KERNEL void k_tra_ldf_iso(DevDomain d, T zsign, const T * __restrict__ pahu, const T * __restrict__ pahv,
                          const T * __restrict__ uslp, const T * __restrict__ vslp, const T * __restrict__ wslpi,
                          const T * __restrict__ wslpj, const T * __restrict__ ah_wslp2, const T * __restrict__ akz,
                          const T * __restrict__ ptb, const T * __restrict__ ptbb, T * __restrict__ pta)
{
    S_tra_ldf_iso<T, MODE, NT, TSPLIT>::run(vb_hw(), d, zsign, pahu, pahv, uslp, vslp, wslpi, wslpj, ah_wslp2, akz, ptb, ptbb, pta);
}

template <typename T, int MODE>
static int launch_iso_mode(const DevDomain &d, T zsign, const T *pahu, const T *pahv, const T *uslp, const T *vslp,
                           const T *wslpi, const T *wslpj, const T *ah_wslp2, const T *akz, const T *ptb,
                           const T *ptbb, T *pta, int kjpt)
{
    Grid2 g = grid2(d.jpi - 2, d.jpj - 2);
    int jn = 0;
    // tracer pairs: one tracer per thread (a two-tracer register window needs
    // 170-255 VGPRs and runs two waves per SIMD or fewer), the pair in one
    // launch so that its two workgroups of a tile share the coefficients
    Grid2 g2 = g;
    g2.grid.x *= 2;
    for (; jn + 2 <= kjpt; jn += 2) {
        const size_t o = (size_t)jn * d.n3;
        LAUNCHB("tra_ldf_iso", (k_tra_ldf_iso<T, MODE, 1, 2>), (S_tra_ldf_iso<T, MODE, 1, 2>), g2, d, zsign, pahu, pahv, uslp, vslp, wslpi, wslpj,
               ah_wslp2, akz, ptb + o, ptbb + o, pta + o);
    }
    for (; jn < kjpt; jn++) {
        const size_t o = (size_t)jn * d.n3;
        LAUNCHB("tra_ldf_iso", (k_tra_ldf_iso<T, MODE, 1, 1>), (S_tra_ldf_iso<T, MODE, 1, 1>), g, d, zsign, pahu, pahv, uslp, vslp, wslpi, wslpj,
               ah_wslp2, akz, ptb + o, ptbb + o, pta + o);
    }
    return 0;
}

// whole-array fill and copy (kernels rather than memset/memcpy calls, so
// the calls can be recorded into a graph)
template <typename T>
struct S_fill3 {
    static constexpr int NTHR = NH_BLOCK, WAVES = 1;
    __device__ __forceinline__ static void run(const VB &vb_, T *__restrict__ p, size_t n, T v)
    {
        const size_t q = (size_t)vb_.x * NH_BLOCK + threadIdx.x;
        if (q < n) p[q] = v;
    }
};
template <typename T>
KERNEL void k_fill3(T *__restrict__ p, size_t n, T v)
{
    S_fill3<T>::run(vb_hw(), p, n, v);
}
template <typename T>
struct S_copy3 {
    static constexpr int NTHR = NH_BLOCK, WAVES = 1;
    __device__ __forceinline__ static void run(const VB &vb_, T *__restrict__ dst, const T *__restrict__ src, size_t n)
    {
        const size_t q = (size_t)vb_.x * NH_BLOCK + threadIdx.x;
        if (q < n) dst[q] = src[q];
    }
};
template <typename T>
KERNEL void k_copy3(T *__restrict__ dst, const T *__restrict__ src, size_t n)
{
    S_copy3<T>::run(vb_hw(), dst, src, n);
}

template <typename T>
int launch_tra_ldf_iso(int first, double z2dt, const T *pahu, const T *pahv, const T *ptb, const T *ptbb, T *pta,
                       int kjpt, int kpass, const T *uslp, const T *vslp, const T *wslpi, const T *wslpj, T *ah_wslp2,
                       T *akz)
{
    const DevDomain &d = nh::dom();
    if (kpass == 1) {
        const unsigned nb = (unsigned)((d.n3 + NH_BLOCK - 1) / NH_BLOCK);
        Grid2 g1;
        g1.block = dim3(NH_BLOCK, 1, 1);
        g1.grid = dim3(nb, 1, 1);
        if (first) {
            LAUNCHB("tra_ldf_iso", k_fill3<T>, (S_fill3<T>), g1, akz, d.n3, T(0));
            LAUNCHB("tra_ldf_iso", k_fill3<T>, (S_fill3<T>), g1, ah_wslp2, d.n3, T(0));
        }
        Grid2 gh = grid2(d.jpi - 1, d.jpj - 1);
        gh.grid.z = (unsigned)(d.jpk - 2);
        LAUNCHB("tra_ldf_iso", k_ldf_iso_coef<T>, (S_ldf_iso_coef<T>), gh, d, z2dt, pahu, pahv, wslpi, wslpj, ah_wslp2, akz);
        if (!d.ln_traldf_msc) LAUNCHB("tra_ldf_iso", k_copy3<T>, (S_copy3<T>), g1, akz, (const T *)ah_wslp2, d.n3);
    }
    const T zsign = (kpass == 1) ? T(1) : T(-1);
    if (d.ln_traldf_lap)
        return launch_iso_mode<T, 0>(d, zsign, pahu, pahv, uslp, vslp, wslpi, wslpj, ah_wslp2, akz, ptb, ptbb, pta, kjpt);
    if (kpass == 1)
        return launch_iso_mode<T, 1>(d, zsign, pahu, pahv, uslp, vslp, wslpi, wslpj, ah_wslp2, akz, ptb, ptbb, pta, kjpt);
    return launch_iso_mode<T, 2>(d, zsign, pahu, pahv, uslp, vslp, wslpi, wslpj, ah_wslp2, akz, ptb, ptbb, pta, kjpt);
}
template int launch_tra_ldf_iso<double>(int, double, const double *, const double *, const double *, const double *,
                                        double *, int, int, const double *, const double *, const double *,
                                        const double *, double *, double *);
template int launch_tra_ldf_iso<float>(int, double, const float *, const float *, const float *, const float *,
                                       float *, int, int, const float *, const float *, const float *, const float *,
                                       float *, float *);

// ---------------------------------------------------------------------------
// tra_zdf_imp (+ the salinity reset of tra_zdf), TRA/trazdf.F90.
// One thread per interior column: the forward sweep builds the matrix on the
// fly (FP64), keeps the pivot of every ZDF_SEG-th level in a scratch column
// and substitutes every tracer of the group; the backward sweep recomputes
// the upper diagonal and, segment by segment from those checkpoints, the
// other pivots, with the operations of the forward sweep.
// ---------------------------------------------------------------------------
// NT tracers share the matrix.  The forward sweep is unrolled by four with
// the pointers declared non-aliasing, so the loads of four levels are issued
// together ahead of the serial FP64 recurrence; the backward sweep loads a
// segment's operands together.
template <typename T, int NT>
struct S_tra_zdf {
    static constexpr int NTHR = NH_BLOCK, WAVES = 1;
    __device__ __forceinline__ static void run(const VB &vb_, const DevDomain &d, double p2dt, const T * __restrict__ ptb, T * __restrict__ pta, int jn0, const T * __restrict__ src, int l_ldfslp, const T * __restrict__ add, double * __restrict__ zwt, int clamp_jn)
    {
        THREAD_IJ_V(vb_, 2, 2);
        if (i > d.jpi - 1 || j > d.jpj - 1) return;
        const uint32_t ij = IJ32(i, j);
        const size_t plane = d.n2, n3 = d.n3;
        const Metrics<T> &m = met<T>(d);
        const int jpk = d.jpk;
        const T *tb[NT];
        T *ta[NT];
    #pragma unroll
        for (int n = 0; n < NT; n++) {
            tb[n] = ptb + (size_t)(jn0 + n) * n3;
            ta[n] = pta + (size_t)(jn0 + n) * n3;
        }
        // zwt(k) of trazdf.F90:153-176: avt/avs, 0 at the surface, plus the
        // iso-neutral contribution on 2..jpkm1.  Both loads are unconditional
        // (levels 1 <= k <= jpk, the second clamped to jpkm1), so the march can
        // keep several levels of loads in flight; the conditions only select.
        const T *addp = l_ldfslp ? add : src;
        auto coef = [&](int k) -> double {
            const double sv = (double)ldp(plane_ptr(src, plane, k), ij);
            const double av = (double)ldp(plane_ptr(addp, plane, k <= jpk - 1 ? k : jpk - 1), ij);
            double c = sv;
            if (l_ldfslp && k <= jpk - 1) c = c + av;
            return k <= 1 ? 0.0 : c;
        };
        // only checkpoint pivots (every ZDF_SEG-th level) are stored; the back
        // substitution recomputes the others segment by segment
        double prev[NT];
        double zws_prev, zwt_prev, zc_k;
        // k = 1 (zwi = 0)
        {
            const double zc_kp1 = coef(2);
            const double e3t = (double)E3(m.e3t, ij, 1);
            const double zwi = -(p2dt * 0.0 / (double)E3(m.e3w, ij, 1));
            const double zws = -(p2dt * zc_kp1 / (double)E3(m.e3w, ij, 2));
            const double zwd = e3t - zwi - zws;
    #pragma unroll
            for (int n = 0; n < NT; n++) {
                const double v = e3t * (double)ldp(tb[n], ij) + p2dt * e3t * (double)ldp((const T *)ta[n], ij);
                prev[n] = v;
                stp(ta[n], ij, (T)v);
            }
            zws_prev = zws;
            zwt_prev = zwd;
            zc_k = zc_kp1;
        }
    #pragma unroll 4
        for (int k = 2; k <= jpk - 1; k++) {
            const double zc_kp1 = coef(k + 1);
            const double e3t = (double)E3(m.e3t, ij, k);
            const double zwi = -(p2dt * zc_k / (double)E3(m.e3w, ij, k));
            const double zws = -(p2dt * zc_kp1 / (double)E3(m.e3w, ij, k + 1));
            const double zwd = e3t - zwi - zws;
            const double zwt_k = zwd - zwi * zws_prev / zwt_prev;
            if (k % ZDF_SEG == 0 && k <= jpk - 2) stp(plane_ptr(zwt, plane, k), ij, zwt_k);
    #pragma unroll
            for (int n = 0; n < NT; n++) {
                const double zrhs = e3t * (double)ldp(plane_ptr(tb[n], plane, k), ij) +
                                    p2dt * e3t * (double)ldp((const T *)plane_ptr(ta[n], plane, k), ij);
                const double v = zrhs - zwi / zwt_prev * prev[n];
                prev[n] = v;
                stp(plane_ptr(ta[n], plane, k), ij, (T)v);
            }
            zws_prev = zws;
            zwt_prev = zwt_k;
            zc_k = zc_kp1;
        }
        // back substitution, the result is the after tracer
        {
            const double tm = TMASK(double, i, j, jpk - 1);
    #pragma unroll
            for (int n = 0; n < NT; n++) {
                const double v = prev[n] / zwt_prev * tm;
                prev[n] = v;
                T out = (T)v;
                if (jn0 + n == clamp_jn && out < T(0)) out = T(0.1);
                stp(plane_ptr(ta[n], plane, jpk - 1), ij, out);
            }
        }
        const int t0 = d.kT.k0[ij], t1 = d.kT.k1[ij];
        const int kl = jpk - 2;
        // each segment's operands are loaded together, from levels within the
        // arrays (clamped to kl; the values of levels past kl are not used)
        auto ld_w = [&](int k, T &sv_, T &av_, T &e3w_) {   // w-level k <= jpk
            sv_ = ldp(plane_ptr(src, plane, k), ij);
            av_ = ldp(plane_ptr(addp, plane, k <= jpk - 1 ? k : jpk - 1), ij);
            e3w_ = E3(m.e3w, ij, k);
        };
        auto coef_v = [&](int k, T sv_, T av_) -> double {
            double c = (double)sv_;
            if (l_ldfslp && k <= jpk - 1) c = c + (double)av_;
            return k <= 1 ? 0.0 : c;
        };
        for (int s = (kl + ZDF_SEG - 1) / ZDF_SEG - 1; s >= 0; s--) {
            const int k0 = s * ZDF_SEG + 1;
            T e3t_[ZDF_SEG], sv_[ZDF_SEG], av_[ZDF_SEG], e3w_[ZDF_SEG], ta_[NT][ZDF_SEG];   // level k, w-level k + 1
    #pragma unroll
            for (int q = 0; q < ZDF_SEG; q++) {
                const int k = k0 + q <= kl ? k0 + q : kl;
                e3t_[q] = E3(m.e3t, ij, k);
                ld_w(k + 1, sv_[q], av_[q], e3w_[q]);
    #pragma unroll
                for (int n = 0; n < NT; n++) ta_[n][q] = ldp((const T *)plane_ptr(ta[n], plane, k), ij);
            }
            T svp, avp, e3wp;   // w-level k0
            ld_w(k0, svp, avp, e3wp);
            double piv[ZDF_SEG], zsv[ZDF_SEG];
            double zc = coef_v(k0, svp, avp);
            double zwt_p = 1.0, zws_p = 0.0;
            if (k0 > 1) {
                zwt_p = ldp((const double *)plane_ptr(zwt, plane, k0 - 1), ij);
                zws_p = -(p2dt * zc / (double)e3wp);
            }
    #pragma unroll
            for (int q = 0; q < ZDF_SEG; q++) {
                const int k = k0 + q;
                const double zc_kp1 = coef_v(k + 1, sv_[q], av_[q]);
                const double e3t = (double)e3t_[q];
                const double zwi = -(p2dt * zc / (double)(q == 0 ? e3wp : e3w_[q > 0 ? q - 1 : 0]));
                const double zws = -(p2dt * zc_kp1 / (double)e3w_[q]);
                const double zwd = e3t - zwi - zws;
                piv[q] = k == 1 ? zwd : zwd - zwi * zws_p / zwt_p;
                zsv[q] = zws;
                zws_p = zws;
                zwt_p = piv[q];
                zc = zc_kp1;
            }
    #pragma unroll
            for (int q = ZDF_SEG - 1; q >= 0; q--) {
                const int k = k0 + q;
                if (k <= kl) {
                    const double tm = (k >= t0 && k <= t1) ? 1.0 : 0.0;
    #pragma unroll
                    for (int n = 0; n < NT; n++) {
                        const double v = ((double)ta_[n][q] - zsv[q] * prev[n]) / piv[q] * tm;
                        prev[n] = v;
                        T out = (T)v;
                        if (jn0 + n == clamp_jn && out < T(0)) out = T(0.1);
                        stp(plane_ptr(ta[n], plane, k), ij, out);
                    }
                }
            }
        }
    #pragma unroll
        for (int n = 0; n < NT; n++) {
            if (jn0 + n == clamp_jn) {
                T *sa = plane_ptr(ta[n], plane, jpk);
                if (ldp((const T *)sa, ij) < T(0)) stp(sa, ij, T(0.1));
            }
        }
    }
};
template <typename T, int NT>
KERNEL void k_tra_zdf(DevDomain d, double p2dt, const T * __restrict__ ptb, T * __restrict__ pta, int jn0,
                      const T * __restrict__ src, int l_ldfslp, const T * __restrict__ add,
                      double * __restrict__ zwt, int clamp_jn)
{
    S_tra_zdf<T, NT>::run(vb_hw(), d, p2dt, ptb, pta, jn0, src, l_ldfslp, add, zwt, clamp_jn);
}

// FP32 storage with at most KMAX - 1 solved levels: the same solve with the
// forward-substituted values (rounded to storage, as k_tra_zdf stores them)
// kept on chip instead of written to pta and read back (the first tracer's
// in registers, the second's in LDS, one column per thread), and the FP64
// pivots kept at every REG_SEG-th level, the others recomputed segment by
// segment in the back substitution from a second read of the diffusivity.
// Every loop over levels is unrolled to KMAX, so the register array is
// indexed by constants; the level count only guards.  Results equal
// k_tra_zdf<float>'s bit for bit.
#define REG_SEG 8
template <int NT, int KMAX>
struct S_tra_zdf_reg {
    static constexpr int NTHR = NH_BLOCK, WAVES = 2;
    __device__ __forceinline__ static void run(const VB &vb_, const DevDomain &d, double p2dt, const float *__restrict__ ptb, float *__restrict__ pta, int jn0, const float *__restrict__ src, int l_ldfslp, const float *__restrict__ add, int clamp_jn)
    {
        THREAD_IJ_V(vb_, 2, 2);
        if (i > d.jpi - 1 || j > d.jpj - 1) return;
        const uint32_t ij = IJ32(i, j);
        const size_t plane = d.n2, n3 = d.n3;
        const Metrics<float> &m = met<float>(d);
        const int jpk = d.jpk;
        const float *tb[NT];
        float *ta[NT];
    #pragma unroll
        for (int n = 0; n < NT; n++) {
            tb[n] = ptb + (size_t)(jn0 + n) * n3;
            ta[n] = pta + (size_t)(jn0 + n) * n3;
        }
        const float *addp = l_ldfslp ? add : src;
        auto coef = [&](int k) -> double {
            const double sv = (double)ldp(plane_ptr(src, plane, k), ij);
            const double av = (double)ldp(plane_ptr(addp, plane, k <= jpk - 1 ? k : jpk - 1), ij);
            double c = sv;
            if (l_ldfslp && k <= jpk - 1) c = c + av;
            return k <= 1 ? 0.0 : c;
        };
        constexpr int NSEG = (KMAX + REG_SEG - 1) / REG_SEG;
        // level k of tracer n at index k - 1: tracer 0 in registers, tracer 1 in LDS
        float fw0[KMAX];
        __shared__ float s_fw1[NT > 1 ? KMAX : 1][NH_BLOCK];
        const int lt = (int)(threadIdx.y * blockDim.x + threadIdx.x);
        auto fw_put = [&](int n, int x, float v) {
            if (n == 0) fw0[x] = v;
            else s_fw1[x][lt] = v;
        };
        auto fw_get = [&](int n, int x) -> float { return n == 0 ? fw0[x] : s_fw1[x][lt]; };
        double ck[NSEG];               // the pivot of level s * REG_SEG at index s
        double prev[NT];
        double zws_prev, zwt_prev, zc_k;
        {
            const double zc_kp1 = coef(2);
            const double e3t = (double)E3(m.e3t, ij, 1);
            const double zwi = -(p2dt * 0.0 / (double)E3(m.e3w, ij, 1));
            const double zws = -(p2dt * zc_kp1 / (double)E3(m.e3w, ij, 2));
            const double zwd = e3t - zwi - zws;
    #pragma unroll
            for (int n = 0; n < NT; n++) {
                const double v = e3t * (double)ldp(tb[n], ij) + p2dt * e3t * (double)ldp((const float *)ta[n], ij);
                prev[n] = v;
                fw_put(n, 0, (float)v);
            }
            zws_prev = zws;
            zwt_prev = zwd;
            zc_k = zc_kp1;
        }
    #pragma unroll
        for (int k = 2; k < KMAX; k++) {
            if (k <= jpk - 1) {
                const double zc_kp1 = coef(k + 1);
                const double e3t = (double)E3(m.e3t, ij, k);
                const double zwi = -(p2dt * zc_k / (double)E3(m.e3w, ij, k));
                const double zws = -(p2dt * zc_kp1 / (double)E3(m.e3w, ij, k + 1));
                const double zwd = e3t - zwi - zws;
                const double zwt_k = zwd - zwi * zws_prev / zwt_prev;
                if (k % REG_SEG == 0) ck[k / REG_SEG] = zwt_k;
    #pragma unroll
                for (int n = 0; n < NT; n++) {
                    const double zrhs = e3t * (double)ldp(plane_ptr(tb[n], plane, k), ij) +
                                        p2dt * e3t * (double)ldp((const float *)plane_ptr(ta[n], plane, k), ij);
                    const double v = zrhs - zwi / zwt_prev * prev[n];
                    prev[n] = v;
                    fw_put(n, k - 1, (float)v);
                }
                zws_prev = zws;
                zwt_prev = zwt_k;
                zc_k = zc_kp1;
            }
        }
        {
            const double tm = TMASK(double, i, j, jpk - 1);
    #pragma unroll
            for (int n = 0; n < NT; n++) {
                const double v = prev[n] / zwt_prev * tm;
                prev[n] = v;
                float out = (float)v;
                if (jn0 + n == clamp_jn && out < 0.f) out = 0.1f;
                stp(plane_ptr(ta[n], plane, jpk - 1), ij, out);
            }
        }
        const int t0 = d.kT.k0[ij], t1 = d.kT.k1[ij];
        const int kl = jpk - 2;
    #pragma unroll
        for (int s = NSEG - 1; s >= 0; s--) {
            const int k0 = s * REG_SEG + 1;
            if (k0 <= kl) {
                double piv[REG_SEG], zsv[REG_SEG];
                double zc = coef(k0);
                double zwt_p = 1.0, zws_p = 0.0;
                if (s > 0) {
                    zwt_p = ck[s];
                    zws_p = -(p2dt * zc / (double)E3(m.e3w, ij, k0));
                }
    #pragma unroll
                for (int q = 0; q < REG_SEG; q++) {
                    const int k = k0 + q;
                    if (k < KMAX && k <= kl) {
                        const double zc_kp1 = coef(k + 1);
                        const double e3t = (double)E3(m.e3t, ij, k);
                        const double zwi = -(p2dt * zc / (double)E3(m.e3w, ij, k));
                        const double zws = -(p2dt * zc_kp1 / (double)E3(m.e3w, ij, k + 1));
                        const double zwd = e3t - zwi - zws;
                        piv[q] = k == 1 ? zwd : zwd - zwi * zws_p / zwt_p;
                        zsv[q] = zws;
                        zws_p = zws;
                        zwt_p = piv[q];
                        zc = zc_kp1;
                    }
                }
    #pragma unroll
                for (int q = REG_SEG - 1; q >= 0; q--) {
                    const int k = k0 + q;
                    if (k < KMAX && k <= kl) {
                        const double tm = (k >= t0 && k <= t1) ? 1.0 : 0.0;
    #pragma unroll
                        for (int n = 0; n < NT; n++) {
                            const double v = ((double)fw_get(n, k - 1) - zsv[q] * prev[n]) / piv[q] * tm;
                            prev[n] = v;
                            float out = (float)v;
                            if (jn0 + n == clamp_jn && out < 0.f) out = 0.1f;
                            stp(plane_ptr(ta[n], plane, k), ij, out);
                        }
                    }
                }
            }
        }
    #pragma unroll
        for (int n = 0; n < NT; n++) {
            if (jn0 + n == clamp_jn) {
                float *sa = plane_ptr(ta[n], plane, jpk);
                if (ldp((const float *)sa, ij) < 0.f) stp(sa, ij, 0.1f);
            }
        }
    }
};
template <int NT, int KMAX>
__global__ __launch_bounds__(NH_BLOCK, 2) void k_tra_zdf_reg(DevDomain d, double p2dt, const float *__restrict__ ptb,
                                                             float *__restrict__ pta, int jn0,
                                                             const float *__restrict__ src, int l_ldfslp,
                                                             const float *__restrict__ add, int clamp_jn)
{
    S_tra_zdf_reg<NT, KMAX>::run(vb_hw(), d, p2dt, ptb, pta, jn0, src, l_ldfslp, add, clamp_jn);
}

// the levels k_tra_zdf_reg holds on chip (jpk <= ZDF_REG_KMAX): 75 of the
// full-size configurations; 76 x 256 FP32 values of LDS per workgroup
#define ZDF_REG_KMAX 76

// WHERE( tsa(:,:,:,jp_sal) < 0 ) tsa = 0.1 on the columns the solve does not touch
template <typename T>
struct S_clamp_ring {
    static constexpr int NTHR = NH_BLOCK, WAVES = 1;
    __device__ __forceinline__ static void run(const VB &vb_, const DevDomain &d, T * __restrict__ sa)
    {
        THREAD_IJ_V(vb_, 1, 1);
        if (i > d.jpi || j > d.jpj || is_interior(d, i, j)) return;
        for (int k = 1; k <= d.jpk; k++)
            if (A3(sa, i, j, k) < T(0)) A3(sa, i, j, k) = T(0.1);
    }
};
template <typename T>
KERNEL void k_clamp_ring(DevDomain d, T * __restrict__ sa)
{
    S_clamp_ring<T>::run(vb_hw(), d, sa);
}

// one solve of NT tracers sharing the matrix: the register form where it
// applies (FP32 storage, jpk <= ZDF_REG_KMAX, switch zdf_reg), else k_tra_zdf
template <typename T, int NT>
static int zdf_group(const char *name, const DevDomain &d, Grid2 g, double p2dt, const T *ptb, T *pta, int jn,
                     const T *src, int l_ldfslp, const T *add, double *zwt, int clamp_jn)
{
    if constexpr (sizeof(T) == 4) {
        if (d.jpk <= ZDF_REG_KMAX && nh::tune(nh::T_ZDF_REG)) {
            LAUNCHB(name, (k_tra_zdf_reg<NT, ZDF_REG_KMAX>), (S_tra_zdf_reg<NT, ZDF_REG_KMAX>), g, d, p2dt, ptb, pta, jn, src, l_ldfslp, add, clamp_jn);
            return 0;
        }
    }
    LAUNCHB(name, (k_tra_zdf<T, NT>), (S_tra_zdf<T, NT>), g, d, p2dt, ptb, pta, jn, src, l_ldfslp, add, zwt, clamp_jn);
    return 0;
}

template <typename T>
int launch_tra_zdf_imp(double p2dt, int is_tra, const T *ptb, T *pta, int kjpt, const T *avt, const T *avs,
                       int l_ldfslp, const T *add, int clamp_sal)
{
    const DevDomain &d = nh::dom();
    Grid2 g = grid2(d.jpi - 2, d.jpj - 2);
    double *zwt = (double *)nh::scratch3(nh::S_ZWT);
    const int clamp_jn = clamp_sal ? 1 : -1;
    // matrix groups: TRA with ln_zdfddm: T on avt, S on avs; otherwise one
    // matrix (avt for TRA, avs for TRC) shared by every tracer
    int st;
    if (is_tra && d.ln_zdfddm) {
        for (int jn = 0; jn < kjpt; jn++)
            if ((st = zdf_group<T, 1>("tra_zdf", d, g, p2dt, ptb, pta, jn, jn == 0 ? avt : avs, l_ldfslp, add, zwt,
                                      clamp_jn)))
                return st;
    } else {
        const T *src = is_tra ? avt : avs;
        int jn = 0;
        for (; jn + 2 <= kjpt; jn += 2)
            if ((st = zdf_group<T, 2>("tra_zdf", d, g, p2dt, ptb, pta, jn, src, l_ldfslp, add, zwt, clamp_jn)))
                return st;
        if (jn < kjpt && (st = zdf_group<T, 1>("tra_zdf", d, g, p2dt, ptb, pta, jn, src, l_ldfslp, add, zwt, clamp_jn)))
            return st;
    }
    if (clamp_sal) {
        Grid2 ga = grid2(d.jpi, d.jpj);
        LAUNCHB("tra_zdf", k_clamp_ring<T>, (S_clamp_ring<T>), ga, d, pta + d.n3);
    }
    return 0;
}
template int launch_tra_zdf_imp<double>(double, int, const double *, double *, int, const double *, const double *,
                                        int, const double *, int);
template int launch_tra_zdf_imp<float>(double, int, const float *, float *, int, const float *, const float *, int,
                                       const float *, int);

// ===========================================================================
// FCT kernels shared by tra_adv_fct and the fused tracer step, and the fused
// tracer step itself: tsa = 0; tra_adv (transports + FCT); tra_ldf_iso (one
// pass, laplacian); tra_zdf -- the order of step.F90:234-257.  Both tracers
// of a pair go through each kernel together:
//   k_fct_h      per cell: horizontal upstream and antidiffusive fluxes,
//                low-order trend, first guess zwi;
//   k_fct_v      per column: vertical upstream flux, compact interpolation
//                and vertical antidiffusive flux zwz;
//   k_lbc        lbc_lnk of zwi, zwx, zwy, zwz;
//   k_fct_beta   per cell: the Zalesak betas (nonosc), then k_lbc;
//   k_fct_final  per column: limited antidiffusive divergence;
// then k_ldf_iso_coef / k_tra_ldf_iso and k_tra_zdf.  In the fused step the
// transports of tra_adv are formed where used and the trend starts from zero
// without a fill pass; every value the separate routines store is rounded to
// the storage type at the same point, so the fused step equals the sequence
// of separate calls bit for bit in either storage form.  The pivots of
// interp_4th_cpt do not depend on the data: rows 3..mbkt-1 follow
// p(k) = 4 - 1/p(k-1) from p(3) = 4, every other row has pivot 1, so they
// come from a table evaluated at compile time (fct.hpp).
// ===========================================================================

// per cell (one thread per (i,j,k), k = blockIdx.z + 1): the horizontal
// fluxes of the (i+1/2, j) and (i, j+1/2) faces (upstream and antidiffusive,
// lines 123-145, 180-221) and, on the interior, the low-order trend and the
// first guess zwi (158-170)
// NT: the number of tracers of the launch, a template constant so that the
// loads of both tracers of a pair are in flight together.  VERT = 0 leaves
// the vertical fluxes to k_fct_vlim: the kernel then stores, in zwi's place,
// the horizontal part s = hu - hum + hv - hvm of the upstream-flux sum (the
// leading terms of the sum ztra is formed from, in the same order) and
// leaves ta alone.
template <int TRP, typename T, int NH4, int TA0, int NT, int VERT = 1>
struct S_fct_h {
    static constexpr int NTHR = NH_BLOCK, WAVES = 1;
    __device__ __forceinline__ static void run(const VB &vb_, const DevDomain &d, double p2dt, const T *__restrict__ pu, const T *__restrict__ pv, const T *__restrict__ pw, const T *__restrict__ ptb, const T *__restrict__ ptn, T *__restrict__ pta, T *__restrict__ pzwi, T *__restrict__ pzwx, T *__restrict__ pzwy, int kjpt)
    {
        (void)kjpt;
        THREAD_IJ_V(vb_, 1, 1);
        const int k = (int)vb_.z + 1;
        if (i > d.jpi - 1 || j > d.jpj - 1) return;     // the flux loops: 1..jpim1 x 1..jpjm1
        const bool inner = (i >= 2 && j >= 2);
        const Metrics<T> &m = met<T>(d);
        const size_t n3 = d.n3;
        const T zp2dt = (T)p2dt;
        const T u = tr_u<TRP>(d, pu, i, j, k), v = tr_v<TRP>(d, pv, i, j, k);
        T um = T(0), vm = T(0), e3t = T(1), tm = T(0), r1_e1e2t = T(0);
        if (inner) {
            um = tr_u<TRP>(d, pu, i - 1, j, k);
            vm = tr_v<TRP>(d, pv, i, j - 1, k);
            const size_t ij = IJ(i, j);
            e3t = E3(m.e3t, ij, k);
            tm = TMASK(T, i, j, k);
            r1_e1e2t = m.r1_e1e2t[ij];
        }
    #pragma unroll
        for (int t = 0; t < NT; t++) {
            const T *tb = ptb + t * n3, *tn = ptn + t * n3;
            const T tbc = A3(tb, i, j, k), tbe = A3(tb, i + 1, j, k), tbn = A3(tb, i, j + 1, k);
            const T hu = T(0.5) * ((u + fabs(u)) * tbc + (u - fabs(u)) * tbe);
            const T hv = T(0.5) * ((v + fabs(v)) * tbc + (v - fabs(v)) * tbn);
            const T c2u = A3(tn, i, j, k) + A3(tn, i + 1, j, k);
            const T c2v = A3(tn, i, j, k) + A3(tn, i, j + 1, k);
            T *zwx = pzwx + t * n3, *zwy = pzwy + t * n3;
            if (NH4) {
                A3(zwx, i, j, k) = T(0.5) * u * (c2u + fct_zltu(d, tn, i, j, k, true) - fct_zltu(d, tn, i + 1, j, k, true)) - hu;
                A3(zwy, i, j, k) = T(0.5) * v * (c2v + fct_zltu(d, tn, i, j, k, false) - fct_zltu(d, tn, i, j + 1, k, false)) - hv;
            } else {
                A3(zwx, i, j, k) = T(0.5) * u * c2u - hu;
                A3(zwy, i, j, k) = T(0.5) * v * c2v - hv;
            }
            if (!inner) continue;
            const T hum = T(0.5) * ((um + fabs(um)) * A3(tb, i - 1, j, k) + (um - fabs(um)) * tbc);
            const T hvm = T(0.5) * ((vm + fabs(vm)) * A3(tb, i, j - 1, k) + (vm - fabs(vm)) * tbc);
            if (!VERT) {
                A3(pzwi + t * n3, i, j, k) = hu - hum + hv - hvm;
                continue;
            }
            const T ztra = -((hu - hum + hv - hvm + tp_upz<TRP>(d, pw, tb, i, j, k) - tp_upz<TRP>(d, pw, tb, i, j, k + 1)) *
                             r1_e1e2t);
            T *ta = pta + t * n3, *zwi = pzwi + t * n3;
            const T a = TA0 ? T(0) : A3(ta, i, j, k);
            A3(ta, i, j, k) = a + ztra / e3t * tm;
            A3(zwi, i, j, k) = (e3t * tbc + zp2dt * ztra) / e3t * tm;
            if (k == d.jpk - 1) A3(zwi, i, j, d.jpk) = T(0);
        }
    }
};
template <int TRP, typename T, int NH4, int TA0, int NT>
KERNEL void k_fct_h(DevDomain d, double p2dt, const T *__restrict__ pu, const T *__restrict__ pv,
                    const T *__restrict__ pw, const T *__restrict__ ptb, const T *__restrict__ ptn, T *__restrict__ pta,
                    T *__restrict__ pzwi, T *__restrict__ pzwx, T *__restrict__ pzwy, int kjpt)
{
    S_fct_h<TRP, T, NH4, TA0, NT, 1>::run(vb_hw(), d, p2dt, pu, pv, pw, ptb, ptn, pta, pzwi, pzwx, pzwy, kjpt);
}
// the horizontal half alone (VERT = 0)
template <int TRP, typename T, int NH4, int NT>
KERNEL void k_fct_hh(DevDomain d, double p2dt, const T *__restrict__ pu, const T *__restrict__ pv,
                     const T *__restrict__ pw, const T *__restrict__ ptb, const T *__restrict__ ptn, T *__restrict__ pta,
                     T *__restrict__ pzwi, T *__restrict__ pzwx, T *__restrict__ pzwy, int kjpt)
{
    S_fct_h<TRP, T, NH4, 1, NT, 0>::run(vb_hw(), d, p2dt, pu, pv, pw, ptb, ptn, pta, pzwi, pzwx, pzwy, kjpt);
}

// per interior column: the vertical antidiffusive flux zwz (lines 255-278),
// the compact scheme's sweeps (577-614) with the table pivots
template <int TRP, typename T, int NV4, int NT>
struct S_fct_v {
    static constexpr int NTHR = NH_BLOCK, WAVES = 1;
    __device__ __forceinline__ static void run(const VB &vb_, const DevDomain &d, const T *__restrict__ pw, const T *__restrict__ ptb, const T *__restrict__ ptn, T *__restrict__ pzwz, int kjpt)
    {
        (void)kjpt;
        THREAD_IJ_V(vb_, 2, 2);
        if (i > d.jpi - 1 || j > d.jpj - 1) return;
        const size_t n3 = d.n3;
        const int jpk = d.jpk;
        const int ikb = d.mbkt[IJ(i, j)];
    #pragma unroll
        for (int t = 0; t < NT; t++) {
            const T *tb = ptb + t * n3, *tn = ptn + t * n3;
            T *zwz = pzwz + t * n3;
            A3(zwz, i, j, 1) = T(0);
            A3(zwz, i, j, jpk) = T(0);
            if (!NV4) {
    #pragma unroll 4
                for (int k = 2; k <= jpk - 1; k++)
                    A3(zwz, i, j, k) = (tr_w<TRP>(d, pw, i, j, k) * T(0.5) * (A3(tn, i, j, k) + A3(tn, i, j, k - 1)) -
                                        tp_upz<TRP>(d, pw, tb, i, j, k)) *
                                       WMASK(T, i, j, k);
                continue;
            }
            T zout_prev = T(0), tn_km1 = A3(tn, i, j, 1);
    #pragma unroll 4
            for (int k = 2; k <= jpk - 1; k++) {
                const T tnk = A3(tn, i, j, k);
                T zwrm, zout, zwi_c = T(0);
                if (k == 2 || k == ikb) {
                    zwrm = T(0.5) * (tn_km1 + tnk);
                } else {
                    const T wm = WMASK(T, i, j, k);
                    zwi_c = wm;
                    zwrm = T(3) * wm * (tnk + tn_km1);
                }
                if (k == 2) zout = zwrm;
                else zout = zwrm - zwi_c / cpt_piv(cpt_tab<T>(), k - 1, ikb) * zout_prev;
                A3(zwz, i, j, k) = zout;
                zout_prev = zout;
                tn_km1 = tnk;
            }
            T xnext = zout_prev / cpt_piv(cpt_tab<T>(), jpk - 1, ikb);
            A3(zwz, i, j, jpk - 1) = (tr_w<TRP>(d, pw, i, j, jpk - 1) * xnext - tp_upz<TRP>(d, pw, tb, i, j, jpk - 1)) *
                                     WMASK(T, i, j, jpk - 1);
    #pragma unroll 4
            for (int k = jpk - 2; k >= 2; k--) {
                const T zws = (k == 2 || k == ikb) ? T(0) : WMASK(T, i, j, k);
                const T x = (A3(zwz, i, j, k) - zws * xnext) / cpt_piv(cpt_tab<T>(), k, ikb);
                A3(zwz, i, j, k) = (tr_w<TRP>(d, pw, i, j, k) * x - tp_upz<TRP>(d, pw, tb, i, j, k)) * WMASK(T, i, j, k);
                xnext = x;
            }
        }
    }
};
template <int TRP, typename T, int NV4, int NT>
KERNEL void k_fct_v(DevDomain d, const T *__restrict__ pw, const T *__restrict__ ptb,
                    const T *__restrict__ ptn, T *__restrict__ pzwz, int kjpt)
{
    S_fct_v<TRP, T, NV4, NT>::run(vb_hw(), d, pw, ptb, ptn, pzwz, kjpt);
}

#define FCTV_REG_KMAX 76
// k_fct_v for FP32 storage and the order-4 scheme with jpk <= KMAX, with
// the vertical half of k_fct_h as well.  The compact scheme's forward-swept
// values stay in registers (every level loop unrolled to KMAX, so the array
// is indexed by constants), so zwz is written once.  After k_fct_hh (k_fct_h
// with VERT = 0) has left the
// horizontal upstream sums s in zwi's place, the back substitution, which
// climbs the column with the vertical upstream flux of each w-level at hand,
// also forms the low-order trend ztra = -(s + upz(k) - upz(k+1)) r1, adds
// ztra / e3t to pta and stores the first guess zwi over s -- what k_fct_h
// with VERT = 1 stores, with the same operations in the same order.  The
// vertical planes of tb and w are then read once, by this column kernel,
// instead of from three levels by every cell of k_fct_h.
template <int TRP, int KMAX, int TA0>
struct S_fct_vz_reg {
    static constexpr int NTHR = NH_BLOCK, WAVES = 2;
    __device__ __forceinline__ static void run(const VB &vb_, const DevDomain &d, double p2dt, const float *__restrict__ pw, const float *__restrict__ ptb, const float *__restrict__ ptn, float *__restrict__ pta, float *__restrict__ pzwi, float *__restrict__ pzwz)
    {
        THREAD_IJ_V(vb_, 2, 2);
        if (i > d.jpi - 1 || j > d.jpj - 1) return;
        const size_t n3 = d.n3;
        const int jpk = d.jpk;
        const size_t ij = IJ(i, j);
        const int ikb = d.mbkt[ij];
        const Metrics<float> &m = met<float>(d);
        const float r1_e1e2t = m.r1_e1e2t[ij];
        const float zp2dt = (float)p2dt;
        const size_t t = vb_.z;
        const float *tb = ptb + t * n3, *tn = ptn + t * n3;
        float *zwz = pzwz + t * n3, *zwi = pzwi + t * n3, *ta = pta + t * n3;
        A3(zwz, i, j, 1) = 0.0f;
        A3(zwz, i, j, jpk) = 0.0f;
        A3(zwi, i, j, jpk) = 0.0f;
        float zo[KMAX];
        // every level's loads are issued two levels ahead, from addresses
        // within the arrays, so they are in flight across the level guards
        float zout_prev = 0.0f, tn_km1 = A3(tn, i, j, 1), tn_n1 = A3(tn, i, j, 2), tn_n2 = A3(tn, i, j, jpk > 2 ? 3 : jpk);
    #pragma unroll
        for (int k = 2; k < KMAX; k++) {
            if (k <= jpk - 1) {
                const float tnk = tn_n1;
                tn_n1 = tn_n2;
                tn_n2 = A3(tn, i, j, k + 2 <= jpk ? k + 2 : jpk);
                float zwrm, zout, zwi_c = 0.0f;
                if (k == 2 || k == ikb) {
                    zwrm = 0.5f * (tn_km1 + tnk);
                } else {
                    const float wm = WMASK(float, i, j, k);
                    zwi_c = wm;
                    zwrm = 3.0f * wm * (tnk + tn_km1);
                }
                if (k == 2) zout = zwrm;
                else zout = zwrm - zwi_c / cpt_piv(cpt_tab<float>(), k - 1, ikb) * zout_prev;
                zo[k] = zout;
                zout_prev = zout;
                tn_km1 = tnk;
            }
        }
        // level k's operands: wn (or pwn), tb(k), the stored s, ta, e3t (and
        // tb(k-1), level k-1's); the climb starts at level jpk - 1
        struct Lv { float w, tb, s, a, e3; };
        const float *wr = pw;
        const auto lv = [&](int k) {
            Lv l;
            l.w = A3(wr, i, j, k);
            l.tb = A3(tb, i, j, k);
            l.s = A3(zwi, i, j, k);
            l.a = TA0 ? 0.0f : A3(ta, i, j, k);
            l.e3 = E3(m.e3t, ij, k);
            return l;
        };
        const int kf = jpk - 1;
        Lv cur = lv(kf), nx = lv(kf > 1 ? kf - 1 : 1);
        const float e1e2t = TRP ? m.e1e2t[ij] : 1.0f;
        float xnext = 0.0f, upz_kp1 = 0.0f;   // upz(jpk) = 0
    #pragma unroll
        for (int k = KMAX - 1; k >= 1; k--) {
            if (k <= jpk - 1) {
                Lv nn = nx;
                if (k >= 3) nn = lv(k - 2);
                const float w_c = cur.w, tb_c = cur.tb, tbm_c = nx.tb, s_c = cur.s, a_c = cur.a, e3_c = cur.e3;
                // tp_upz and tr_w of level k, the same operations
                const float w = TRP ? e1e2t * w_c : w_c;
                const float zfp = w + fabsf(w), zfm = w - fabsf(w);
                const float upz = (k == 1) ? w * tb_c : 0.5f * (zfp * tb_c + zfm * tbm_c) * WMASK(float, i, j, k);
                if (k >= 2) {   // zwz, as k_fct_v
                    const float wm = WMASK(float, i, j, k);
                    float x;
                    if (k == jpk - 1) {
                        x = zo[k] / cpt_piv(cpt_tab<float>(), k, ikb);
                    } else {
                        const float zws = (k == 2 || k == ikb) ? 0.0f : wm;
                        x = (zo[k] - zws * xnext) / cpt_piv(cpt_tab<float>(), k, ikb);
                    }
                    A3(zwz, i, j, k) = (w * x - upz) * wm;
                    xnext = x;
                }
                // the low-order trend and first guess, as k_fct_h
                const float e3t = e3_c, tm = TMASK(float, i, j, k);
                const float ztra = -(((s_c + upz) - upz_kp1) * r1_e1e2t);
                A3(ta, i, j, k) = a_c + ztra / e3t * tm;
                A3(zwi, i, j, k) = (e3t * tb_c + zp2dt * ztra) / e3t * tm;
                upz_kp1 = upz;
                cur = nx;
                nx = nn;
            }
        }
    }
};
template <int TRP, int KMAX, int TA0>
__global__ __launch_bounds__(NH_BLOCK, 2) void k_fct_vz_reg(DevDomain d, double p2dt, const float *__restrict__ pw,
                                                          const float *__restrict__ ptb, const float *__restrict__ ptn,
                                                          float *__restrict__ pta, float *__restrict__ pzwi,
                                                          float *__restrict__ pzwz)
{
    S_fct_vz_reg<TRP, KMAX, TA0>::run(vb_hw(), d, p2dt, pw, ptb, ptn, pta, pzwi, pzwz);
}

// ring and bottom level of a trend that starts from zero (fused step)
template <typename T>
struct S_zero_ring {
    static constexpr int NTHR = NH_BLOCK, WAVES = 1;
    __device__ __forceinline__ static void run(const VB &vb_, const DevDomain &d, T *__restrict__ pta, int ntr)
    {
        THREAD_IJ_V(vb_, 1, 1);
        if (i > d.jpi || j > d.jpj) return;
        const bool ring = !is_interior(d, i, j);
        for (int t = 0; t < ntr; t++) {
            T *ta = pta + t * d.n3;
            if (ring)
                for (int k = 1; k <= d.jpk; k++) A3(ta, i, j, k) = T(0);
            else
                A3(ta, i, j, d.jpk) = T(0);
        }
    }
};
template <typename T>
KERNEL void k_zero_ring(DevDomain d, T *__restrict__ pta, int ntr)
{
    S_zero_ring<T>::run(vb_hw(), d, pta, ntr);
}

// the chain's flux kernels for one tracer with the transports given: pta +=
// the low-order trend; zwi, zwx, zwy, zwz after their lbc_lnk (the first
// stage of tra_adv_fct, as the frozen-limiter entry points need it)
template <typename T>
static int launch_fct_fluxes1(const DevDomain &d, double p2dt, const T *pu, const T *pv, const T *pw, const T *ptb,
                              const T *ptn, T *pta, int kn_h, int kn_v, T *zwi, T *zwx, T *zwy, T *zwz,
                              const char *nm)
{
    Grid2 g3 = grid2(d.jpi - 1, d.jpj - 1), gi = grid2(d.jpi - 2, d.jpj - 2);
    g3.grid.z = (unsigned)(d.jpk - 1);
    const int nt = 1;
    if (kn_h == 4) LAUNCHB(nm, (k_fct_h<0, T, 1, 0, 1>), (S_fct_h<0, T, 1, 0, 1>), g3, d, p2dt, pu, pv, pw, ptb, ptn, pta, zwi, zwx, zwy, nt);
    else LAUNCHB(nm, (k_fct_h<0, T, 0, 0, 1>), (S_fct_h<0, T, 0, 0, 1>), g3, d, p2dt, pu, pv, pw, ptb, ptn, pta, zwi, zwx, zwy, nt);
    if (kn_v == 4) LAUNCHB(nm, (k_fct_v<0, T, 1, 1>), (S_fct_v<0, T, 1, 1>), gi, d, pw, ptb, ptn, zwz, nt);
    else LAUNCHB(nm, (k_fct_v<0, T, 0, 1>), (S_fct_v<0, T, 0, 1>), gi, d, pw, ptb, ptn, zwz, nt);
    int st;
    if ((st = launch_lbc<T>(zwi, d.jpk, 'T', T(0)))) return st;
    if ((st = launch_lbc<T>(zwx, d.jpk, 'U', T(0)))) return st;
    if ((st = launch_lbc<T>(zwy, d.jpk, 'V', T(0)))) return st;
    if ((st = launch_lbc<T>(zwz, d.jpk, 'W', T(0)))) return st;
    return 0;
}

template <int TRP, typename T>
static int launch_fct_pairs(const DevDomain &d, double p2dt, const T *pu, const T *pv, const T *pw, const T *ptb,
                            const T *ptn, T *pta, int kjpt, int kn_fct_h, int kn_fct_v, int ta0)
{
    // scratch slot pairs (two tracers of either storage type each)
    T *zwi = (T *)nh::scratch3(0), *zwx = (T *)nh::scratch3(2), *zwy = (T *)nh::scratch3(4);
    T *zwz = (T *)nh::scratch3(6), *bu = (T *)nh::scratch3(8), *bd = (T *)nh::scratch3(10);
    const char *nm = TRP ? "step_tra" : "tra_adv_fct";
    Grid2 gi = grid2(d.jpi - 2, d.jpj - 2), gk = gi;
    gk.grid.z = (unsigned)((d.jpk - 1 + FctKB<T>::beta - 1) / FctKB<T>::beta);
    const size_t n3 = d.n3;
    int st;
    for (int t0 = 0; t0 < kjpt; t0 += 2) {
        const int nt = kjpt - t0 < 2 ? kjpt - t0 : 2;
        const size_t o = (size_t)t0 * n3;
        if (!d.iper && !d.jper && nh::tune(nh::T_FCT_FUSED)) {
            // closed domain: the whole chain in one kernel per tile and tracer
            // (zwz's slots hold the compact scheme's forward elimination)
            if ((st = launch_fct_fused<TRP, T>(d, p2dt, nt, pu, pv, pw, ptb + o, ptn + o, pta + o, kn_fct_h,
                                               kn_fct_v, ta0, zwz, nm)))
                return st;
            continue;
        }
        Grid2 g3 = grid2(d.jpi - 1, d.jpj - 1);
        g3.grid.z = (unsigned)(d.jpk - 1);
        if (!d.iper && !d.jper && nh::tune(nh::T_FCT_VLIM)) {
            // closed domain: horizontal fluxes, then the compact scheme's
            // forward sweep, then the vertical fluxes, first guess and limiter
            // in one bottom-up march (zwi holds s, zwz the forward sweep)
#define FH_LAUNCH2(H, N) \
    LAUNCHB(nm, (k_fct_hh<TRP, T, H, N>), (S_fct_h<TRP, T, H, 1, N, 0>), g3, d, p2dt, pu, pv, pw, ptb + o, ptn + o, pta + o, zwi, zwx, zwy, nt)
            if (kn_fct_h == 4) { if (nt == 2) FH_LAUNCH2(1, 2); else FH_LAUNCH2(1, 1); }
            else { if (nt == 2) FH_LAUNCH2(0, 2); else FH_LAUNCH2(0, 1); }
#undef FH_LAUNCH2
            if (kn_fct_v == 4 && (st = launch_fct_cpt_fwd<T>(d, nt, ptn + o, zwz, nm))) return st;
            const int nk = nt * d.jpk;
            if ((st = launch_lbc<T>(zwx, nk, 'U', T(0)))) return st;
            if ((st = launch_lbc<T>(zwy, nk, 'V', T(0)))) return st;
            if ((st = launch_fct_vlim<TRP, T>(d, p2dt, nt, kn_fct_v, ta0, pw, ptb + o, zwi,
                                              kn_fct_v == 4 ? zwz : ptn + o, zwx, zwy, pta + o, nm)))
                return st;
            continue;
        }
        bool vdone = false;
        if constexpr (sizeof(T) == 4) {
            if (kn_fct_v == 4 && d.jpk <= FCTV_REG_KMAX) {
                // horizontal fluxes per cell, then the vertical fluxes, the
                // low-order trend and the first guess per column
#define FH_LAUNCH2(H, N) \
    LAUNCHB(nm, (k_fct_hh<TRP, T, H, N>), (S_fct_h<TRP, T, H, 1, N, 0>), g3, d, p2dt, pu, pv, pw, ptb + o, ptn + o, pta + o, zwi, zwx, zwy, nt)
                if (kn_fct_h == 4) { if (nt == 2) FH_LAUNCH2(1, 2); else FH_LAUNCH2(1, 1); }
                else { if (nt == 2) FH_LAUNCH2(0, 2); else FH_LAUNCH2(0, 1); }
#undef FH_LAUNCH2
                Grid2 gv = gi;
                gv.grid.z = (unsigned)nt;
                if (ta0) LAUNCHB(nm, (k_fct_vz_reg<TRP, FCTV_REG_KMAX, 1>), (S_fct_vz_reg<TRP, FCTV_REG_KMAX, 1>), gv, d, p2dt, pw, ptb + o, ptn + o, pta + o, zwi, zwz);
                else LAUNCHB(nm, (k_fct_vz_reg<TRP, FCTV_REG_KMAX, 0>), (S_fct_vz_reg<TRP, FCTV_REG_KMAX, 0>), gv, d, p2dt, pw, ptb + o, ptn + o, pta + o, zwi, zwz);
                vdone = true;
            }
        }
        if (!vdone) {
#define FH_LAUNCH2(H, Z, N) \
    LAUNCHB(nm, (k_fct_h<TRP, T, H, Z, N>), (S_fct_h<TRP, T, H, Z, N>), g3, d, p2dt, pu, pv, pw, ptb + o, ptn + o, pta + o, zwi, zwx, zwy, nt)
#define FH_LAUNCH(H, Z) \
    do { if (nt == 2) FH_LAUNCH2(H, Z, 2); else FH_LAUNCH2(H, Z, 1); } while (0)
            if (kn_fct_h == 4) {
                if (ta0) FH_LAUNCH(1, 1); else FH_LAUNCH(1, 0);
            } else {
                if (ta0) FH_LAUNCH(0, 1); else FH_LAUNCH(0, 0);
            }
#undef FH_LAUNCH
#undef FH_LAUNCH2
        }
        if (vdone) {
        } else if (kn_fct_v == 4) {
            if (nt == 2) LAUNCHB(nm, (k_fct_v<TRP, T, 1, 2>), (S_fct_v<TRP, T, 1, 2>), gi, d, pw, ptb + o, ptn + o, zwz, nt);
            else LAUNCHB(nm, (k_fct_v<TRP, T, 1, 1>), (S_fct_v<TRP, T, 1, 1>), gi, d, pw, ptb + o, ptn + o, zwz, nt);
        } else {
            if (nt == 2) LAUNCHB(nm, (k_fct_v<TRP, T, 0, 2>), (S_fct_v<TRP, T, 0, 2>), gi, d, pw, ptb + o, ptn + o, zwz, nt);
            else LAUNCHB(nm, (k_fct_v<TRP, T, 0, 1>), (S_fct_v<TRP, T, 0, 1>), gi, d, pw, ptb + o, ptn + o, zwz, nt);
        }
        const int nk = nt * d.jpk;
        if ((st = launch_lbc<T>(zwi, nk, 'T', T(0)))) return st;
        if ((st = launch_lbc<T>(zwx, nk, 'U', T(0)))) return st;
        if ((st = launch_lbc<T>(zwy, nk, 'V', T(0)))) return st;
        if ((st = launch_lbc<T>(zwz, nk, 'W', T(0)))) return st;
        if (!d.iper && !d.jper) {
            // closed domain: betas and limited divergence fused, both tracers in one launch
            constexpr int TY = FctLimTY<T>::value, NTH = FctLimTY<T>::nth;
            Grid2 gl = grid2(d.jpi - 2, d.jpj - 2);
            gl.block = dim3(LIM_TX, NTH / LIM_TX, 1);
            gl.grid.y = (unsigned)((d.jpj - 2 + TY - 1) / TY);
            gl.grid.z = (unsigned)nt;
            LAUNCHB(nm, (k_fct_lim<T, TY, NTH>), (S_fct_lim<T, TY, NTH>), gl, d, p2dt, ptb + o, zwi, zwx, zwy, zwz, pta + o);
            continue;
        }
        for (int t = 0; t < nt; t++) {
            const size_t q = (size_t)t * n3;
            LAUNCHB(nm, (k_fct_beta<T, FctKB<T>::beta>), (S_fct_beta<T, FctKB<T>::beta>), gk, d, p2dt, ptb + o + q, zwi + q, zwx + q, zwy + q, zwz + q, bu + q, bd + q);
        }
        if ((st = launch_lbc<T>(bu, nk, 'T', T(0)))) return st;
        if ((st = launch_lbc<T>(bd, nk, 'T', T(0)))) return st;
        for (int t = 0; t < nt; t++) {
            const size_t q = (size_t)t * n3;
            LAUNCHB(nm, (k_fct_final<T>), (S_fct_final<T>), gi, d, zwx + q, zwy + q, zwz + q, bu + q, bd + q, pta + o + q);
        }
    }
    if (ta0) {
        Grid2 ga = grid2(d.jpi, d.jpj);
        LAUNCHB(nm, k_zero_ring<T>, (S_zero_ring<T>), ga, d, pta, kjpt);
    }
    return 0;
}



// tra_adv (TRA/traadv.F90:77-175, FCT): the transports of un, vn, wn formed
// where used and tra_adv_fct added to pta
template <typename T>
int launch_tra_adv(double p2dt, const T *un, const T *vn, const T *wn, const T *ptb, const T *ptn, T *pta, int kjpt,
                   int kn_fct_h, int kn_fct_v)
{
    const DevDomain &d = nh::dom();
    if (d.jpk <= STEP_MAXK)
        return launch_fct_pairs<1, T>(d, p2dt, un, vn, wn, ptb, ptn, pta, kjpt, kn_fct_h, kn_fct_v, 0);
    T *zun = (T *)nh::scratch3(nh::S_ZUN), *zvn = (T *)nh::scratch3(nh::S_ZVN), *zwn = (T *)nh::scratch3(nh::S_ZWN);
    int st = launch_tra_adv_trp<T>(un, vn, wn, zun, zvn, zwn);
    if (st) return st;
    return launch_tra_adv_fct<T>(p2dt, zun, zvn, zwn, ptb, ptn, pta, kjpt, kn_fct_h, kn_fct_v);
}
template int launch_tra_adv<double>(double, const double *, const double *, const double *, const double *,
                                    const double *, double *, int, int, int);
template int launch_tra_adv<float>(double, const float *, const float *, const float *, const float *,
                                   const float *, float *, int, int, int);

template <typename T>
int launch_step_tra(double p2dt, double z2dt, double p2dt_zdf, const T *un, const T *vn, const T *wn, const T *tsb,
                    const T *tsn, T *tsa, int kn_fct_h, int kn_fct_v, const T *ahtu, const T *ahtv, const T *uslp,
                    const T *vslp, const T *wslpi, const T *wslpj, const T *avt, T *ah_wslp2, T *akz)
{
    const DevDomain &d = nh::dom();
    if (d.jpk > STEP_MAXK) return nh::set_error(NEMO_HIP_ERR_UNSUPPORTED, "step_tra: jpk = %d > %d", d.jpk, STEP_MAXK);
    if (!d.ln_traldf_lap) return nh::set_error(NEMO_HIP_ERR_UNSUPPORTED, "step_tra: bilaplacian diffusion is not fused");
    if (d.ln_zdfddm) return nh::set_error(NEMO_HIP_ERR_UNSUPPORTED, "step_tra: ln_zdfddm = T is not fused");
    int st;
    // FCT on transports formed where used, trend from zero
    if ((st = launch_fct_pairs<1, T>(d, p2dt, un, vn, wn, tsb, tsn, tsa, 2, kn_fct_h, kn_fct_v, 1))) return st;
    if ((st = launch_tra_ldf_iso<T>(0, z2dt, ahtu, ahtv, tsb, tsb, tsa, 2, 1, uslp, vslp, wslpi, wslpj, ah_wslp2, akz)))
        return st;
    // tra_zdf: the shared matrix on avt plus the iso-neutral coefficient, salinity reset
    double *zwt = (double *)nh::scratch3(8);
    const T *add = d.ln_traldf_msc ? (const T *)akz : (const T *)ah_wslp2;
    Grid2 gi = grid2(d.jpi - 2, d.jpj - 2);
    return zdf_group<T, 2>("step_tra", d, gi, p2dt_zdf, tsb, tsa, 0, avt, 1, add, zwt, 1);
}
template int launch_step_tra<double>(double, double, double, const double *, const double *, const double *,
                                     const double *, const double *, double *, int, int, const double *,
                                     const double *, const double *, const double *, const double *, const double *,
                                     const double *, double *, double *);
template int launch_step_tra<float>(double, double, double, const float *, const float *, const float *,
                                    const float *, const float *, float *, int, int, const float *, const float *,
                                    const float *, const float *, const float *, const float *, const float *, float *,
                                    float *);
