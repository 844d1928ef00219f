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

// ---------------------------------------------------------------------------
// a3: effective transport, TRA/traadv.F90:100-124
// ---------------------------------------------------------------------------
template <typename T>
__global__ void k_tra_adv_trp(DevDomain d, const T *un, const T *vn, const T *wn, T *zun, T *zvn, T *zwn)
{
    THREAD_IJ(1, 1);
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

template <typename T>
int launch_tra_adv_trp(const T *un, const T *vn, const T *wn, T *zun, T *zvn, T *zwn)
{
    const DevDomain &d = nh::dom();
    Grid2 g = grid2(d.jpi, d.jpj);
    LAUNCH("tra_adv_trp", k_tra_adv_trp<T>, g, d, un, vn, wn, zun, zvn, zwn);
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
__global__ void k_fct_low(DevDomain d, double p2dt, const T *pun, const T *pvn, const T *pwn, const T *tb, T *ta,
                          T *zwi)
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
    if (!lbc_src(d, 'T', i, j, si, sj)) return T(0);
    const T r1_6 = T(1) / T(6);
    if (xdir) {
        const T a = (A3(tn, si + 1, sj, k) - A3(tn, si, sj, k)) * UMASK(T, si, sj, k);
        const T b = (A3(tn, si, sj, k) - A3(tn, si - 1, sj, k)) * UMASK(T, si - 1, sj, k);
        return (a + b) * r1_6;
    } else {
        const T a = (A3(tn, si, sj + 1, k) - A3(tn, si, sj, k)) * VMASK(T, si, sj, k);
        const T b = (A3(tn, si, sj, k) - A3(tn, si, sj - 1, k)) * VMASK(T, si, sj - 1, k);
        return (a + b) * r1_6;
    }
}

template <typename T>
__global__ void k_fct_ad(DevDomain d, int kn_fct_h, int kn_fct_v, const T *pun, const T *pvn, const T *pwn,
                         const T *tb, const T *tn, T *zwx, T *zwy, T *zwz, double *zwt)
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
template <typename T>
__global__ void k_fct_beta(DevDomain d, double p2dt, const T *pbef, const T *paft, const T *paa, const T *pbb,
                           const T *pcc, T *zbetup, T *zbetdo)
{
    THREAD_IJ(2, 2);
    if (i > d.jpi - 1 || j > d.jpj - 1) return;
    const double zrtrn = 1.e-15;
    const size_t ij = IJ(i, j);
    const double e1e2t = d.md.e1e2t[ij];
    const int jpk = d.jpk;
    for (int k = 1; k <= jpk - 1; k++) {
        const int ikm1 = k - 1 > 1 ? k - 1 : 1;
        double zup = zbup(d, pbef, paft, i, j, k);
        zup = fmax2(zup, zbup(d, pbef, paft, i - 1, j, k));
        zup = fmax2(zup, zbup(d, pbef, paft, i + 1, j, k));
        zup = fmax2(zup, zbup(d, pbef, paft, i, j - 1, k));
        zup = fmax2(zup, zbup(d, pbef, paft, i, j + 1, k));
        zup = fmax2(zup, zbup(d, pbef, paft, i, j, ikm1));
        zup = fmax2(zup, zbup(d, pbef, paft, i, j, k + 1));
        double zdo = zbdo(d, pbef, paft, i, j, k);
        zdo = fmin2(zdo, zbdo(d, pbef, paft, i - 1, j, k));
        zdo = fmin2(zdo, zbdo(d, pbef, paft, i + 1, j, k));
        zdo = fmin2(zdo, zbdo(d, pbef, paft, i, j - 1, k));
        zdo = fmin2(zdo, zbdo(d, pbef, paft, i, j + 1, k));
        zdo = fmin2(zdo, zbdo(d, pbef, paft, i, j, ikm1));
        zdo = fmin2(zdo, zbdo(d, pbef, paft, i, j, k + 1));
        const double a_im1 = A3(paa, i - 1, j, k), a_i = A3(paa, i, j, k);
        const double b_jm1 = A3(pbb, i, j - 1, k), b_j = A3(pbb, i, j, k);
        const double c_kp1 = A3(pcc, i, j, k + 1), c_k = A3(pcc, i, j, k);
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
    }
    A3(zbetup, i, j, jpk) = T(0);
    A3(zbetdo, i, j, jpk) = T(0);
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
// paa at (p,q) after nonosc's limiting and its lbc_lnk('U')
template <typename T>
__device__ __forceinline__ double lim_x(const DevDomain &d, const T *paa, const T *bup, const T *bdo, int p, int q, int k)
{
    int sp, sq;
    if (!lbc_src(d, 'U', p, q, sp, sq)) return 0.0;
    return lim_face<T>(A3(paa, sp, sq, k), A3(bdo, sp, sq, k), A3(bup, sp + 1, sq, k), A3(bup, sp, sq, k),
                       A3(bdo, sp + 1, sq, k));
}
template <typename T>
__device__ __forceinline__ double lim_y(const DevDomain &d, const T *pbb, const T *bup, const T *bdo, int p, int q, int k)
{
    int sp, sq;
    if (!lbc_src(d, 'V', p, q, sp, sq)) return 0.0;
    return lim_face<T>(A3(pbb, sp, sq, k), A3(bdo, sp, sq, k), A3(bup, sp, sq + 1, k), A3(bup, sp, sq, k),
                       A3(bdo, sp, sq + 1, k));
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
__global__ void k_fct_final(DevDomain d, const T *paa, const T *pbb, const T *pcc, const T *bup, const T *bdo, T *ta)
{
    THREAD_IJ(2, 2);
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

template <typename T>
int launch_tra_adv_fct(double p2dt, const T *pun, const T *pvn, const T *pwn, const T *ptb, const T *ptn, T *pta,
                       int kjpt, int kn_fct_h, int kn_fct_v)
{
    const DevDomain &d = nh::dom();
    T *zwi = (T *)nh::scratch3(nh::S_ZWI), *zwx = (T *)nh::scratch3(nh::S_ZWX), *zwy = (T *)nh::scratch3(nh::S_ZWY);
    T *zwz = (T *)nh::scratch3(nh::S_ZWZ), *zbu = (T *)nh::scratch3(nh::S_ZBETUP), *zbd = (T *)nh::scratch3(nh::S_ZBETDO);
    double *zwt = (double *)nh::scratch3(nh::S_ZTW);
    Grid2 gi = grid2(d.jpi - 2, d.jpj - 2), gh = grid2(d.jpi - 1, d.jpj - 1);
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
        LAUNCH("tra_adv_fct", (k_fct_beta<T>), gi, d, p2dt, tb, zwi, zwx, zwy, zwz, zbu, zbd);
        if ((st = launch_lbc<T>(zbu, d.jpk, 'T', T(0)))) return st;
        if ((st = launch_lbc<T>(zbd, d.jpk, 'T', T(0)))) return st;
        LAUNCH("tra_adv_fct", (k_fct_final<T>), gi, d, zwx, zwy, zwz, zbu, zbd, ta);
    }
    return 0;
}
template int launch_tra_adv_fct<double>(double, const double *, const double *, const double *, const double *,
                                        const double *, double *, int, int, int);
template int launch_tra_adv_fct<float>(double, const float *, const float *, const float *, const float *,
                                       const float *, float *, int, int, int);

// ---------------------------------------------------------------------------
// a6: interp_4th_cpt alone (traadv_fct.F90:517-616), interior columns,
// levels 2..jpkm1 of pt_out
// ---------------------------------------------------------------------------
__global__ void k_interp_4th_cpt(DevDomain d, const double *tin, double *tout, double *zwt)
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
// a8: tra_ldf_lap, TRA/traldf_lap_blp.F90:99-145 (ln_zps = F); all tracers
// of the call are updated by the same column thread
// ---------------------------------------------------------------------------
template <typename T>
__global__ void k_tra_ldf_lap(DevDomain d, T zsign, const T *pahu, const T *pahv, const T *ptb, T *pta, int kjpt)
{
    THREAD_IJ(2, 2);
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

template <typename T>
int launch_tra_ldf_lap(const T *pahu, const T *pahv, const T *ptb, T *pta, int kjpt, int kpass)
{
    const DevDomain &d = nh::dom();
    Grid2 g = grid2(d.jpi - 2, d.jpj - 2);
    const T zsign = (kpass == 1) ? T(1) : T(-1);
    LAUNCH("tra_ldf_lap", k_tra_ldf_lap<T>, g, d, zsign, pahu, pahv, ptb, pta, kjpt);
    return 0;
}
template int launch_tra_ldf_lap<double>(const double *, const double *, const double *, double *, int, int);
template int launch_tra_ldf_lap<float>(const float *, const float *, const float *, float *, int, int);

// ---------------------------------------------------------------------------
// a9: tra_ldf_iso, TRA/traldf_iso.F90:144-366
// ---------------------------------------------------------------------------
// pass-1 coefficients ah_wslp2 and akz, lines 146-202 (threads 1..jpim1 x 1..jpjm1)
template <typename T>
__global__ void k_ldf_iso_coef(DevDomain d, double z2dt, const T *pahu, const T *pahv, const T *wslpi,
                               const T *wslpj, T *ah_wslp2, T *akz)
{
    THREAD_IJ(1, 1);
    if (i > d.jpi - 1 || j > d.jpj - 1) return;
    const Metrics<T> &m = met<T>(d);
    const bool inner = is_interior(d, i, j);
    const size_t ij = IJ(i, j);
    const T z1_2dt = T(1) / (T)z2dt;
    for (int k = 2; k <= d.jpk - 1; k++) {
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

template <typename T>
__device__ __forceinline__ T iso_zdk1t(const DevDomain &d, const T *tb, int i, int j, int k)
{
    return (A3(tb, i, j, k) - A3(tb, i, j, k + 1)) * WMASK(T, i, j, k + 1);
}
template <typename T>
__device__ __forceinline__ T iso_zdkt(const DevDomain &d, const T *tb, int i, int j, int k)
{
    if (k == 1) return iso_zdk1t(d, tb, i, j, 1);
    return (A3(tb, i, j, k - 1) - A3(tb, i, j, k)) * WMASK(T, i, j, k);
}
// masked horizontal gradients zdit / zdjt, lines 218-225
template <typename T>
__device__ __forceinline__ T iso_zdit(const DevDomain &d, const T *tb, int i, int j, int k)
{
    return (A3(tb, i + 1, j, k) - A3(tb, i, j, k)) * UMASK(T, i, j, k);
}
template <typename T>
__device__ __forceinline__ T iso_zdjt(const DevDomain &d, const T *tb, int i, int j, int k)
{
    return (A3(tb, i, j + 1, k) - A3(tb, i, j, k)) * VMASK(T, i, j, k);
}
// zftu at (p,q,k), lines 257-271
template <typename T>
__device__ __forceinline__ T iso_zftu(const DevDomain &d, const T *pahu, const T *uslp, const T *tb, int p, int q, int k)
{
    const Metrics<T> &m = met<T>(d);
    const size_t pq = IJ(p, q);
    const T zabe1 = A3(pahu, p, q, k) * m.e2_e1u[pq] * E3(m.e3u, pq, k);
    const T zmsku = T(1) / fmax2(WMASK(T, p + 1, q, k) + WMASK(T, p, q, k + 1) + WMASK(T, p + 1, q, k + 1) +
                                     WMASK(T, p, q, k), T(1));
    const T zcof1 = -(A3(pahu, p, q, k) * m.e2u[pq] * A3(uslp, p, q, k) * zmsku);
    return (zabe1 * iso_zdit(d, tb, p, q, k) + zcof1 * (iso_zdkt(d, tb, p + 1, q, k) + iso_zdk1t(d, tb, p, q, k) +
                                                        iso_zdk1t(d, tb, p + 1, q, k) + iso_zdkt(d, tb, p, q, k))) *
           UMASK(T, p, q, k);
}
template <typename T>
__device__ __forceinline__ T iso_zftv(const DevDomain &d, const T *pahv, const T *vslp, const T *tb, int p, int q, int k)
{
    const Metrics<T> &m = met<T>(d);
    const size_t pq = IJ(p, q);
    const T zabe2 = A3(pahv, p, q, k) * m.e1_e2v[pq] * E3(m.e3v, pq, k);
    const T zmskv = T(1) / fmax2(WMASK(T, p, q + 1, k) + WMASK(T, p, q, k + 1) + WMASK(T, p, q + 1, k + 1) +
                                     WMASK(T, p, q, k), T(1));
    const T zcof2 = -(A3(pahv, p, q, k) * m.e1v[pq] * A3(vslp, p, q, k) * zmskv);
    return (zabe2 * iso_zdjt(d, tb, p, q, k) + zcof2 * (iso_zdkt(d, tb, p, q + 1, k) + iso_zdk1t(d, tb, p, q, k) +
                                                        iso_zdk1t(d, tb, p, q + 1, k) + iso_zdkt(d, tb, p, q, k))) *
           VMASK(T, p, q, k);
}
// ztfw at (i,j,k), 2 <= k <= jpkm1, lines 298-357
template <typename T>
__device__ __forceinline__ T iso_ztfw(const DevDomain &d, int kpass, const T *pahu, const T *pahv, const T *wslpi,
                                      const T *wslpj, const T *ah_wslp2, const T *akz, const T *tb, const T *tbb,
                                      int i, int j, int k)
{
    if (k <= 1 || k >= d.jpk) return T(0);
    const Metrics<T> &m = met<T>(d);
    const size_t ij = IJ(i, j);
    const T wm = WMASK(T, i, j, k);
    const T zmsku = wm / fmax2(UMASK(T, i, j, k - 1) + UMASK(T, i - 1, j, k) + UMASK(T, i - 1, j, k - 1) +
                                   UMASK(T, i, j, k), T(1));
    const T zmskv = wm / fmax2(VMASK(T, i, j, k - 1) + VMASK(T, i, j - 1, k) + VMASK(T, i, j - 1, k - 1) +
                                   VMASK(T, i, j, k), T(1));
    const T zahu_w = (A3(pahu, i, j, k - 1) + A3(pahu, i - 1, j, k) + A3(pahu, i - 1, j, k - 1) + A3(pahu, i, j, k)) *
                     zmsku;
    const T zahv_w = (A3(pahv, i, j, k - 1) + A3(pahv, i, j - 1, k) + A3(pahv, i, j - 1, k - 1) + A3(pahv, i, j, k)) *
                     zmskv;
    const T zcoef3 = -(zahu_w * m.e2t[ij] * zmsku * A3(wslpi, i, j, k));
    const T zcoef4 = -(zahv_w * m.e1t[ij] * zmskv * A3(wslpj, i, j, k));
    T ztfw = zcoef3 * (iso_zdit(d, tb, i, j, k - 1) + iso_zdit(d, tb, i - 1, j, k) + iso_zdit(d, tb, i - 1, j, k - 1) +
                       iso_zdit(d, tb, i, j, k)) +
             zcoef4 * (iso_zdjt(d, tb, i, j, k - 1) + iso_zdjt(d, tb, i, j - 1, k) + iso_zdjt(d, tb, i, j - 1, k - 1) +
                       iso_zdjt(d, tb, i, j, k));
    const T e1e2t = m.e1e2t[ij], e3w = E3(m.e3w, ij, k);
    const size_t p = IJK(i, j, k);
    const T dt = A3(tb, i, j, k - 1) - tb[p];
    if (d.ln_traldf_lap) {
        ztfw = ztfw + e1e2t / e3w * wm * (ah_wslp2[p] - akz[p]) * dt;
    } else if (kpass == 1) {
        ztfw = ztfw + ah_wslp2[p] * e1e2t * dt / e3w * wm;
    } else {
        ztfw = ztfw + e1e2t / e3w * wm * (ah_wslp2[p] * dt + akz[p] * (A3(tbb, i, j, k - 1) - tbb[p]));
    }
    return ztfw;
}

template <typename T>
__global__ void k_tra_ldf_iso(DevDomain d, int kpass, T zsign, const T *pahu, const T *pahv, const T *uslp,
                              const T *vslp, const T *wslpi, const T *wslpj, const T *ah_wslp2, const T *akz,
                              const T *tb, const T *tbb, T *ta)
{
    THREAD_IJ(2, 2);
    if (i > d.jpi - 1 || j > d.jpj - 1) return;
    const Metrics<T> &m = met<T>(d);
    const size_t ij = IJ(i, j);
    const T r1_e1e2t = m.r1_e1e2t[ij];
    T ztfw_k = T(0);
    for (int k = 1; k <= d.jpk - 1; k++) {
        const T e3t = E3(m.e3t, ij, k);
        const T hdiv = iso_zftu(d, pahu, uslp, tb, i, j, k) - iso_zftu(d, pahu, uslp, tb, i - 1, j, k) +
                       iso_zftv(d, pahv, vslp, tb, i, j, k) - iso_zftv(d, pahv, vslp, tb, i, j - 1, k);
        const T ztfw_kp1 = iso_ztfw(d, kpass, pahu, pahv, wslpi, wslpj, ah_wslp2, akz, tb, tbb, i, j, k + 1);
        const size_t p = IJK(i, j, k);
        T t = ta[p];
        t = t + zsign * hdiv * r1_e1e2t / e3t;
        t = t + zsign * (ztfw_k - ztfw_kp1) * r1_e1e2t / e3t;
        ta[p] = t;
        ztfw_k = ztfw_kp1;
    }
}

template <typename T>
int launch_tra_ldf_iso(int first, double z2dt, const T *pahu, const T *pahv, const T *ptb, const T *ptbb, T *pta,
                       int kjpt, int kpass, const T *uslp, const T *vslp, const T *wslpi, const T *wslpj, T *ah_wslp2,
                       T *akz)
{
    const DevDomain &d = nh::dom();
    if (kpass == 1) {
        if (first) {
            HIPCHK(hipMemsetAsync(akz, 0, d.n3 * sizeof(T), nh::stream()));
            HIPCHK(hipMemsetAsync(ah_wslp2, 0, d.n3 * sizeof(T), nh::stream()));
        }
        Grid2 gh = grid2(d.jpi - 1, d.jpj - 1);
        LAUNCH("tra_ldf_iso", k_ldf_iso_coef<T>, gh, d, z2dt, pahu, pahv, wslpi, wslpj, ah_wslp2, akz);
        if (!d.ln_traldf_msc)
            HIPCHK(hipMemcpyAsync(akz, ah_wslp2, d.n3 * sizeof(T), hipMemcpyDeviceToDevice, nh::stream()));
    }
    const T zsign = (kpass == 1) ? T(1) : T(-1);
    Grid2 g = grid2(d.jpi - 2, d.jpj - 2);
    for (int jn = 0; jn < kjpt; jn++) {
        const size_t o = (size_t)jn * d.n3;
        LAUNCH("tra_ldf_iso", k_tra_ldf_iso<T>, g, d, kpass, zsign, pahu, pahv, uslp, vslp, wslpi, wslpj, ah_wslp2,
               akz, ptb + o, ptbb + o, pta + o);
    }
    return 0;
}
template int launch_tra_ldf_iso<double>(int, double, const double *, const double *, const double *, const double *,
                                        double *, int, int, const double *, const double *, const double *,
                                        const double *, double *, double *);
template int launch_tra_ldf_iso<float>(int, double, const float *, const float *, const float *, const float *,
                                       float *, int, int, const float *, const float *, const float *, const float *,
                                       float *, float *);

// ---------------------------------------------------------------------------
// a10/a11: tra_zdf_imp (+ the salinity reset of tra_zdf), TRA/trazdf.F90.
// One thread per interior column: the forward sweep builds the matrix on the
// fly (FP64), keeps T_k in a scratch column and substitutes every tracer of
// the group; the backward sweep recomputes the upper diagonal.
// ---------------------------------------------------------------------------
#define ZDF_MAXT 4
template <typename T>
__device__ __forceinline__ double zdf_coef(const DevDomain &d, const T *src, int l_ldfslp, const T *add, int i, int j,
                                           int k)
{
    // zwt(k) of trazdf.F90:153-176: avt/avs, 0 at the surface, plus the
    // iso-neutral contribution on 2..jpkm1
    if (k <= 1) return 0.0;
    double c = (double)A3(src, i, j, k);
    if (l_ldfslp && k <= d.jpk - 1) c = c + (double)A3(add, i, j, k);
    return c;
}

template <typename T>
__global__ void k_tra_zdf(DevDomain d, double p2dt, const T *ptb, T *pta, int jn0, int ntr, const T *src,
                          int l_ldfslp, const T *add, double *zwt, int clamp_jn)
{
    THREAD_IJ(2, 2);
    if (i > d.jpi - 1 || j > d.jpj - 1) return;
    const size_t ij = IJ(i, j);
    const Metrics<T> &m = met<T>(d);
    const int jpk = d.jpk;
    double prev[ZDF_MAXT];
    double zws_prev = 0.0, zwt_prev = 1.0;
    double zc_k = zdf_coef(d, src, l_ldfslp, add, i, j, 1);
    for (int k = 1; k <= jpk - 1; k++) {
        const double zc_kp1 = zdf_coef(d, src, l_ldfslp, add, i, j, k + 1);
        const double e3t = (double)E3(m.e3t, ij, k);
        const double zwi = -(p2dt * zc_k / (double)E3(m.e3w, ij, k));
        const double zws = -(p2dt * zc_kp1 / (double)E3(m.e3w, ij, k + 1));
        const double zwd = e3t - zwi - zws;
        const double zwt_k = (k == 1) ? zwd : zwd - zwi * zws_prev / zwt_prev;
        A3(zwt, i, j, k) = zwt_k;
        const size_t p = IJK(i, j, k);
        for (int n = 0; n < ntr; n++) {
            const size_t o = (size_t)(jn0 + n) * d.n3 + p;
            const double zrhs = e3t * (double)ptb[o] + p2dt * e3t * (double)pta[o];
            const double v = (k == 1) ? zrhs : zrhs - zwi / zwt_prev * prev[n];
            prev[n] = v;
            pta[o] = (T)v;
        }
        zws_prev = zws;
        zwt_prev = zwt_k;
        zc_k = zc_kp1;
    }
    // back substitution, the result is the after tracer
    {
        const double tm = TMASK(double, i, j, jpk - 1);
        for (int n = 0; n < ntr; n++) {
            const double v = prev[n] / zwt_prev * tm;
            prev[n] = v;
            T out = (T)v;
            if (jn0 + n == clamp_jn && out < T(0)) out = T(0.1);
            pta[(size_t)(jn0 + n) * d.n3 + IJK(i, j, jpk - 1)] = out;
        }
    }
    for (int k = jpk - 2; k >= 1; k--) {
        const double zws = -(p2dt * zdf_coef(d, src, l_ldfslp, add, i, j, k + 1) / (double)E3(m.e3w, ij, k + 1));
        const double zwt_k = A3(zwt, i, j, k);
        const double tm = TMASK(double, i, j, k);
        const size_t p = IJK(i, j, k);
        for (int n = 0; n < ntr; n++) {
            const size_t o = (size_t)(jn0 + n) * d.n3 + p;
            const double v = ((double)pta[o] - zws * prev[n]) / zwt_k * tm;
            prev[n] = v;
            T out = (T)v;
            if (jn0 + n == clamp_jn && out < T(0)) out = T(0.1);
            pta[o] = out;
        }
    }
    if (clamp_jn >= jn0 && clamp_jn < jn0 + ntr) {
        T *sa = pta + (size_t)clamp_jn * d.n3;
        if (A3(sa, i, j, jpk) < T(0)) A3(sa, i, j, jpk) = T(0.1);
    }
}

// WHERE( tsa(:,:,:,jp_sal) < 0 ) tsa = 0.1 on the columns the solve does not touch
template <typename T>
__global__ void k_clamp_ring(DevDomain d, T *sa)
{
    THREAD_IJ(1, 1);
    if (i > d.jpi || j > d.jpj || is_interior(d, i, j)) return;
    for (int k = 1; k <= d.jpk; k++)
        if (A3(sa, i, j, k) < T(0)) A3(sa, i, j, k) = T(0.1);
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
    if (is_tra && d.ln_zdfddm) {
        for (int jn = 0; jn < kjpt; jn++)
            LAUNCH("tra_zdf", k_tra_zdf<T>, g, d, p2dt, ptb, pta, jn, 1, jn == 0 ? avt : avs, l_ldfslp, add, zwt,
                   clamp_jn);
    } else {
        const T *src = is_tra ? avt : avs;
        for (int jn = 0; jn < kjpt; jn += ZDF_MAXT) {
            int nt = kjpt - jn < ZDF_MAXT ? kjpt - jn : ZDF_MAXT;
            LAUNCH("tra_zdf", k_tra_zdf<T>, g, d, p2dt, ptb, pta, jn, nt, src, l_ldfslp, add, zwt, clamp_jn);
        }
    }
    if (clamp_sal) {
        Grid2 ga = grid2(d.jpi, d.jpj);
        LAUNCH("tra_zdf", k_clamp_ring<T>, ga, d, pta + d.n3);
    }
    return 0;
}
template int launch_tra_zdf_imp<double>(double, int, const double *, double *, int, const double *, const double *,
                                        int, const double *, int);
template int launch_tra_zdf_imp<float>(double, int, const float *, float *, int, const float *, const float *, int,
                                       const float *, int);
