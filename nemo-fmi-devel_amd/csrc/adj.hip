// adj.hip -- the adjoint of the tracer step for one tracer, FP64: tra_adv_fct
// (orders 2 or 4, the nonosc limiter factors frozen from the forward state),
// tra_ldf_lap (kpass 1), tra_zdf_imp (avt, no slopes) and the leapfrog
// tra_nxt, TRA/traadv_fct.F90, traldf_lap_blp.F90, trazdf.F90, tranxt.F90.
//
// Each kernel is the transpose of the forward routine as the library runs
// it, written as a gather: a thread owns one cell of the adjoint output and
// sums the contributions of the faces around it, so no two threads write
// the same point.  A forward lbc_lnk copies a ring point from its source;
// its transpose adds the ring value to the source and clears the ring,
// which the thread of the source does (it alone reads that ring point).
#include "common.hpp"
#include "kernels.hpp"

namespace {

__device__ __forceinline__ bool interior(const DevDomain &d, int i, int j)
{
    return i >= 2 && i <= d.jpi - 1 && j >= 2 && j <= d.jpj - 1;
}

// the ring points whose lbc source ('T') is the interior point (i,j)
template <typename F>
__device__ __forceinline__ void ring_of(const DevDomain &d, int i, int j, F f)
{
    if (i != 2 && i != d.jpi - 1 && j != 2 && j != d.jpj - 1) return;
    const int ci[3] = {i, 1, d.jpi}, cj[3] = {j, 1, d.jpj};
    for (int a = 0; a < 3; a++)
        for (int b = 0; b < 3; b++) {
            if (a == 0 && b == 0) continue;
            int si, sj;
            if (lbc_src(d, 'T', ci[a], cj[b], si, sj) && si == i && sj == j) f(ci[a], cj[b]);
        }
}

// ---------------------------------------------------------------------------
// tra_adv_fct with frozen factors.  The forward trend at an interior cell
// is  -(dX + dY + dZ) r1/e3t tmask  (upstream fluxes of tb)
// -(dLx + dLy + dLz) r1/e3t  (limited antidiffusive fluxes: factor times
// the high-order flux of tn minus the upstream flux, the horizontal ones
// through lbc_lnk 'U'/'V').  The 4th-order horizontal flux adds
// 0.5 u (zltu(i) - zltu(i+1)), zltu = (ztu(i) + ztu(i-1))/6 at interior
// cells then lbc_lnk 'T', ztu = (tn(i+1) - tn(i)) umask
// (traadv_fct.F90:192-221); the 4th-order vertical one replaces
// (tn(k) + tn(k-1))/2 by the compact interpolation interp_4th_cpt.
// ---------------------------------------------------------------------------
struct FctAd {
    const double *un, *vn, *wn, *cx, *cy, *cz, *mu;
    const double *zl_u, *zl_v;   // kn_h = 4: the adjoint of zltu / zltv after the lbc fold
    const double *hi, *lo;       // the cell factors below, formed once per cell (k_fct_ad_cell)
    int kn_h, kn_v;
};

// mu r1/e3t (hi) and mu r1/e3t tmask (lo) of a cell; zero off the interior.
// k_fct_ad_cell forms them once per cell; the gathers, which need each
// cell's pair for the dozen faces around it, read them back.
__device__ __forceinline__ double cell_hi(const DevDomain &d, const FctAd &a, int i, int j, int k)
{
    if (!interior(d, i, j) || k < 1 || k > d.jpk - 1) return 0.;
    return A3(a.hi, i, j, k);
}
__device__ __forceinline__ double cell_lo(const DevDomain &d, const FctAd &a, int i, int j, int k)
{
    if (!interior(d, i, j) || k < 1 || k > d.jpk - 1) return 0.;
    return A3(a.lo, i, j, k);
}

KERNEL void k_fct_ad_cell(DevDomain d, const double *__restrict__ mu, double *__restrict__ hi,
                          double *__restrict__ lo)
{
    THREAD_IJ(2, 2);
    const int k = (int)blockIdx.z + 1;
    if (i > d.jpi - 1 || j > d.jpj - 1) return;
    const uint32_t ij = IJ32(i, j);
    const double h = A3(mu, i, j, k) * d.md.r1_e1e2t[ij] / E3(d.md.e3t, ij, k);
    A3(hi, i, j, k) = h;
    A3(lo, i, j, k) = h * TMASK(double, i, j, k);
}

// adjoint of the limited x-flux at face p (between cells p and p+1) of row j
__device__ __forceinline__ double ax_face(const DevDomain &d, const FctAd &a, int p, int j, int k)
{
    if (p < 2 || p > d.jpi - 1 || j < 2 || j > d.jpj - 1) return 0.;
    double v = cell_hi(d, a, p + 1, j, k) - cell_hi(d, a, p, j, k);
    if (d.iper && p == d.jpi - 1) v += cell_hi(d, a, 2, j, k);   // the ring face 1 copies face jpi-1
    return A3(a.cx, p, j, k) * v;
}
__device__ __forceinline__ double ay_face(const DevDomain &d, const FctAd &a, int i, int q, int k)
{
    if (i < 2 || i > d.jpi - 1 || q < 2 || q > d.jpj - 1) return 0.;
    double v = cell_hi(d, a, i, q + 1, k) - cell_hi(d, a, i, q, k);
    if (d.jper && q == d.jpj - 1) v += cell_hi(d, a, i, 2, k);
    return A3(a.cy, i, q, k) * v;
}
// adjoint of the limited vertical flux at w-level kk (2..jpk) of an interior column
__device__ __forceinline__ double az_face(const DevDomain &d, const FctAd &a, int i, int j, int kk)
{
    if (!interior(d, i, j) || kk < 2 || kk > d.jpk) return 0.;
    return A3(a.cz, i, j, kk) * (cell_hi(d, a, i, j, kk - 1) - cell_hi(d, a, i, j, kk));
}
// adjoint of the upstream fluxes: the low-order trend minus the antidiffusive part
__device__ __forceinline__ double xb_face(const DevDomain &d, const FctAd &a, int p, int j, int k)
{
    if (p < 1 || p > d.jpi - 1 || j < 1 || j > d.jpj - 1) return 0.;
    return cell_lo(d, a, p + 1, j, k) - cell_lo(d, a, p, j, k) - ax_face(d, a, p, j, k);
}
__device__ __forceinline__ double yb_face(const DevDomain &d, const FctAd &a, int i, int q, int k)
{
    if (i < 1 || i > d.jpi - 1 || q < 1 || q > d.jpj - 1) return 0.;
    return cell_lo(d, a, i, q + 1, k) - cell_lo(d, a, i, q, k) - ay_face(d, a, i, q, k);
}
__device__ __forceinline__ double zb_face(const DevDomain &d, const FctAd &a, int i, int j, int kk)
{
    if (kk < 1 || kk > d.jpk - 1) return 0.;
    double v = -cell_lo(d, a, i, j, kk);
    if (kk >= 2) v += cell_lo(d, a, i, j, kk - 1) - WMASK(double, i, j, kk) * az_face(d, a, i, j, kk);
    return v;
}

KERNEL void k_fct_ad(DevDomain d, FctAd a, double *__restrict__ tb_ad, double *__restrict__ tn_ad)
{
    THREAD_IJ(1, 1);
    const int k = (int)blockIdx.z + 1;
    if (i > d.jpi || j > d.jpj) return;
    double gb = 0., gn = 0.;
    if (i <= d.jpi - 1 && j <= d.jpj - 1) {       // faces i (x) and j (y) of this cell
        const double u = A3(a.un, i, j, k), v = A3(a.vn, i, j, k);
        gn += 0.5 * u * ax_face(d, a, i, j, k) + 0.5 * v * ay_face(d, a, i, j, k);
        gb += 0.5 * (u + fabs(u)) * xb_face(d, a, i, j, k) + 0.5 * (v + fabs(v)) * yb_face(d, a, i, j, k);
    }
    if (i >= 2 && j <= d.jpj - 1) {                // face i-1
        const double u = A3(a.un, i - 1, j, k);
        gn += 0.5 * u * ax_face(d, a, i - 1, j, k);
        gb += 0.5 * (u - fabs(u)) * xb_face(d, a, i - 1, j, k);
    }
    if (j >= 2 && i <= d.jpi - 1) {                // face j-1
        const double v = A3(a.vn, i, j - 1, k);
        gn += 0.5 * v * ay_face(d, a, i, j - 1, k);
        gb += 0.5 * (v - fabs(v)) * yb_face(d, a, i, j - 1, k);
    }
    if (a.kn_h == 4) {   // ztu = dtn umask feeds zltu at cells i and i+1 (r1_6 each)
        const double r1_6 = 1.0 / 6.0;
        if (j <= d.jpj - 1) {
            if (i >= 2)
                gn += (A3(a.zl_u, i - 1, j, k) * r1_6 + A3(a.zl_u, i, j, k) * r1_6) * UMASK(double, i - 1, j, k);
            if (i <= d.jpi - 1)
                gn -= (A3(a.zl_u, i, j, k) * r1_6 + A3(a.zl_u, i + 1, j, k) * r1_6) * UMASK(double, i, j, k);
        }
        if (i <= d.jpi - 1) {
            if (j >= 2)
                gn += (A3(a.zl_v, i, j - 1, k) * r1_6 + A3(a.zl_v, i, j, k) * r1_6) * VMASK(double, i, j - 1, k);
            if (j <= d.jpj - 1)
                gn -= (A3(a.zl_v, i, j, k) * r1_6 + A3(a.zl_v, i, j + 1, k) * r1_6) * VMASK(double, i, j, k);
        }
    }
    // vertical: the antidiffusive flux of w-levels k and k+1 (interior
    // columns); the 4th-order one is k_fct_ad_v4's
    if (a.kn_v == 2 && interior(d, i, j)) {
        if (k >= 2) gn += A3(a.wn, i, j, k) * 0.5 * (az_face(d, a, i, j, k) * WMASK(double, i, j, k));
        if (k + 1 <= d.jpk - 1)
            gn += A3(a.wn, i, j, k + 1) * 0.5 * (az_face(d, a, i, j, k + 1) * WMASK(double, i, j, k + 1));
    }
    // the upstream flux of w-levels k and k+1 (w(1) tb(1) at the surface)
    if (k == 1) {
        gb += A3(a.wn, i, j, 1) * zb_face(d, a, i, j, 1);
    } else {
        const double w = A3(a.wn, i, j, k);
        gb += (w + fabs(w)) * (0.5 * zb_face(d, a, i, j, k) * WMASK(double, i, j, k));
    }
    if (k + 1 <= d.jpk - 1) {
        const double w = A3(a.wn, i, j, k + 1);
        gb += (w - fabs(w)) * (0.5 * zb_face(d, a, i, j, k + 1) * WMASK(double, i, j, k + 1));
    }
    const uint32_t p = IJK32(i, j, k);
    at32(tb_ad, p) += gb;
    at32(tn_ad, p) += gn;
}

// the adjoint of zltu before the lbc: 0.5 u ax of face i minus that of face i-1
__device__ __forceinline__ double zlu_raw(const DevDomain &d, const FctAd &a, int i, int j, int k)
{
    if (j > d.jpj - 1) return 0.;
    double v = 0.;
    if (i >= 2) v = -(0.5 * A3(a.un, i - 1, j, k) * ax_face(d, a, i - 1, j, k));
    if (i <= d.jpi - 1) v += 0.5 * A3(a.un, i, j, k) * ax_face(d, a, i, j, k);
    return v;
}
__device__ __forceinline__ double zlv_raw(const DevDomain &d, const FctAd &a, int i, int j, int k)
{
    if (i > d.jpi - 1) return 0.;
    double v = 0.;
    if (j >= 2) v = -(0.5 * A3(a.vn, i, j - 1, k) * ay_face(d, a, i, j - 1, k));
    if (j <= d.jpj - 1) v += 0.5 * A3(a.vn, i, j, k) * ay_face(d, a, i, j, k);
    return v;
}

// kn_h = 4: zl_u, zl_v = the adjoint of zltu, zltv with the lbc_lnk 'T'
// transposed (ring values added to their interior sources); zero off the
// interior, where the forward computes no zltu of its own
KERNEL void k_fct_ad_l4(DevDomain d, FctAd a, double *__restrict__ zl_u, double *__restrict__ zl_v)
{
    THREAD_IJ(1, 1);
    const int k = (int)blockIdx.z + 1;
    if (i > d.jpi || j > d.jpj) return;
    double lu = 0., lv = 0.;
    if (interior(d, i, j)) {
        lu = zlu_raw(d, a, i, j, k);
        lv = zlv_raw(d, a, i, j, k);
        ring_of(d, i, j, [&](int ri, int rj) {
            lu += zlu_raw(d, a, ri, rj, k);
            lv += zlv_raw(d, a, ri, rj, k);
        });
    }
    A3(zl_u, i, j, k) = lu;
    A3(zl_v, i, j, k) = lv;
}

// kn_v = 4: ztw = interp_4th_cpt(tn) enters the vertical flux as w ztw;
// the transpose of the compact solve (traadv_fct.F90:535-614) on one
// interior column per thread, the substitutions reversed, as two passes
// over the column: top down, the pivots (recomputed as the forward solve
// forms them) and the reversed back substitution, whose results t(k) and
// pivots go to two scratch columns; bottom up, the reversed forward
// substitution and the adjoint of the right-hand side.  Nothing is held
// per level in the thread, so the kernel needs no private memory.
KERNEL void k_fct_ad_v4(DevDomain d, FctAd a, double *__restrict__ tn_ad, double *__restrict__ zwt_s,
                        double *__restrict__ t_s)
{
    THREAD_IJ(2, 2);
    if (i > d.jpi - 1 || j > d.jpj - 1) return;
    const int K = d.jpk - 1, ikt = A2(d.mikt, i, j) + 1, ikb = A2(d.mbkt, i, j);
    // the diagonal, the equal sub- and super-diagonals and the right-hand
    // side coefficient of row k
    auto row = [&](int k, double &zd, double &zs, double &cc) {
        const double wm = WMASK(double, i, j, k);
        zd = 3.0 * wm + 1.0;
        zs = wm;
        cc = 3.0 * wm;
        if (k == ikt || k == ikb) {
            zd = 1.0;
            zs = 0.0;
            cc = 0.5;
        }
    };
    double zwt_p = 0.0, zs_p = 0.0, carry = 0.0;
    for (int k = 2; k <= K; k++) {
        double zd, zs, cc;
        row(k, zd, zs, cc);
        const double zwt = (k == 2) ? zd : zd - zs * zs_p / zwt_p;
        double x = A3(a.wn, i, j, k) * (az_face(d, a, i, j, k) * WMASK(double, i, j, k));
        if (k > 2) x -= carry;
        const double t = x / zwt;
        A3(t_s, i, j, k) = t;
        A3(zwt_s, i, j, k) = zwt;
        carry = zs * t;
        zwt_p = zwt;
        zs_p = zs;
    }
    // r(K) = t(K); r(k-1) = t(k-1) - zs(k) / zwt(k-1) r(k); tn_ad(m) gets
    // cc(m) r(m) and then cc(m+1) r(m+1), in that order
    double zd, zs_k, cc_k;
    row(K, zd, zs_k, cc_k);
    double r_k = A3(t_s, i, j, K);
    double acc = A3(tn_ad, i, j, K) + cc_k * r_k;   // level K: cc(K) r(K) only
    for (int k = K; k >= 2; k--) {
        A3(tn_ad, i, j, k) = acc;
        const double up = cc_k * r_k;                 // cc(k) r(k), also owed to level k-1
        if (k == 2) {
            A3(tn_ad, i, j, 1) = A3(tn_ad, i, j, 1) + up;
            break;
        }
        const double r_km1 = A3(t_s, i, j, k - 1) - zs_k / A3(zwt_s, i, j, k - 1) * r_k;
        double cc_km1;
        row(k - 1, zd, zs_k, cc_km1);
        acc = (A3(tn_ad, i, j, k - 1) + cc_km1 * r_km1) + up;
        r_k = r_km1;
        cc_k = cc_km1;
    }
}

// ---------------------------------------------------------------------------
// tra_ldf_lap, kpass 1: ta += (dtu + dtv) / (e1e2t e3t), tu = aheeu dtb
// ---------------------------------------------------------------------------
__device__ __forceinline__ double lap_cell(const DevDomain &d, const double *mu, int i, int j, int k)
{
    if (!interior(d, i, j)) return 0.;
    const uint32_t ij = IJ32(i, j);
    return A3(mu, i, j, k) / (d.md.e1e2t[ij] * E3(d.md.e3t, ij, k));
}
__device__ __forceinline__ double lap_u(const DevDomain &d, const double *ahu, const double *mu, int p, int j, int k)
{
    if (p < 1 || p > d.jpi - 1 || j > d.jpj - 1) return 0.;
    const uint32_t ij = IJ32(p, j);
    return A3(ahu, p, j, k) * d.md.e2_e1u[ij] * E3(d.md.e3u, ij, k) * (lap_cell(d, mu, p, j, k) - lap_cell(d, mu, p + 1, j, k));
}
__device__ __forceinline__ double lap_v(const DevDomain &d, const double *ahv, const double *mu, int i, int q, int k)
{
    if (q < 1 || q > d.jpj - 1 || i > d.jpi - 1) return 0.;
    const uint32_t ij = IJ32(i, q);
    return A3(ahv, i, q, k) * d.md.e1_e2v[ij] * E3(d.md.e3v, ij, k) * (lap_cell(d, mu, i, q, k) - lap_cell(d, mu, i, q + 1, k));
}

KERNEL void k_ldf_lap_ad(DevDomain d, const double *__restrict__ ahu, const double *__restrict__ ahv,
                         const double *__restrict__ mu, double *__restrict__ tb_ad)
{
    THREAD_IJ(1, 1);
    const int k = (int)blockIdx.z + 1;
    if (i > d.jpi || j > d.jpj) return;
    const double g = lap_u(d, ahu, mu, i - 1, j, k) - lap_u(d, ahu, mu, i, j, k) + lap_v(d, ahv, mu, i, j - 1, k) -
                     lap_v(d, ahv, mu, i, j, k);
    at32(tb_ad, IJK32(i, j, k)) += g;
}

// ---------------------------------------------------------------------------
// tra_zdf_imp, one column per thread: the elimination reversed, as two
// passes over the column.  Top down: the pivots zwt(k) (trazdf.F90:193-234,
// recomputed as the forward solve forms them, kept in a scratch column)
// and the reversed back substitution t(k) = x(k) tmask / zwt(k), x(k+1) -=
// zws(k) t(k), its t(k) written over ta_ad(k).  Bottom up: the reversed
// forward substitution r(k-1) = t(k-1) - zwi(k) / zwt(k-1) r(k) and the
// adjoint of the right-hand side.  No per-level state in the thread.
// ---------------------------------------------------------------------------
KERNEL void k_zdf_imp_ad(DevDomain d, double p2dt, const double *__restrict__ avt, double *__restrict__ ta_ad,
                         double *__restrict__ tb_ad, double *__restrict__ zwt_s)
{
    THREAD_IJ(2, 2);
    if (i > d.jpi - 1 || j > d.jpj - 1) return;
    const Metrics<double> &m = d.md;
    const uint32_t ij = IJ32(i, j);
    const int K = d.jpk - 1;
    auto zwi_of = [&](int k) {
        const double wt_k = (k == 1) ? 0.0 : A3(avt, i, j, k);
        return -p2dt * wt_k / E3(m.e3w, ij, k);
    };
    double zwt_p = 0.0, zws_p = 0.0, carry = 0.0;
    for (int k = 1; k <= K; k++) {
        const double zwi = zwi_of(k);
        const double zws = -p2dt * A3(avt, i, j, k + 1) / E3(m.e3w, ij, k + 1);
        const double zwd = E3(m.e3t, ij, k) - zwi - zws;
        const double zwt = (k == 1) ? zwd : zwd - zwi * zws_p / zwt_p;
        double x = A3(ta_ad, i, j, k);
        if (k > 1) x -= carry;
        const double t = x * TMASK(double, i, j, k) / zwt;
        A3(ta_ad, i, j, k) = t;
        A3(zwt_s, i, j, k) = zwt;
        carry = zws * t;
        zwt_p = zwt;
        zws_p = zws;
    }
    double r = A3(ta_ad, i, j, K);
    for (int k = K; k >= 1; k--) {
        const double e3 = E3(m.e3t, ij, k);
        A3(tb_ad, i, j, k) += e3 * r;
        A3(ta_ad, i, j, k) = p2dt * e3 * r;
        if (k > 1) r = A3(ta_ad, i, j, k - 1) - zwi_of(k) / A3(zwt_s, i, j, k - 1) * r;
    }
}

// ---------------------------------------------------------------------------
// tra_nxt (leapfrog): the lbc_lnk of the new tsb/tsn transposed (folded
// into the sources), then tb' = tn + atfp (ta - 2 tn + tb), tn' = ta
// ---------------------------------------------------------------------------
KERNEL void k_tra_nxt_ad(DevDomain d, double atfp, double *__restrict__ tb_ad, double *__restrict__ tn_ad,
                         double *__restrict__ ta_ad)
{
    THREAD_IJ(1, 1);
    const int k = (int)blockIdx.z + 1;
    if (i > d.jpi || j > d.jpj) return;
    const uint32_t p = IJK32(i, j, k);
    if (!interior(d, i, j)) {
        int si, sj;
        if (!lbc_src(d, 'T', i, j, si, sj)) {   // no source reads it: cleared here
            at32(tb_ad, p) = 0.;
            at32(tn_ad, p) = 0.;
        }
        at32(ta_ad, p) = 0.;
        return;
    }
    double b = at32(tb_ad, p), n = at32(tn_ad, p);
    ring_of(d, i, j, [&](int ri, int rj) {
        const uint32_t q = IJK32(ri, rj, k);
        b += at32(tb_ad, q);
        n += at32(tn_ad, q);
        at32(tb_ad, q) = 0.;
        at32(tn_ad, q) = 0.;
    });
    if (k == d.jpk) {   // level jpk is not filtered: identity
        at32(tb_ad, p) = b;
        at32(tn_ad, p) = n;
        at32(ta_ad, p) = 0.;
        return;
    }
    at32(ta_ad, p) = atfp * b + n;
    at32(tb_ad, p) = atfp * b;
    at32(tn_ad, p) = b - 2.0 * atfp * b;
}

// the transpose of lbc_lnk 'T' on a field (the gradient of a state's interior)
KERNEL void k_lbc_adj(DevDomain d, double *__restrict__ f)
{
    THREAD_IJ(1, 1);
    const int k = (int)blockIdx.z + 1;
    if (i > d.jpi || j > d.jpj) return;
    const uint32_t p = IJK32(i, j, k);
    if (!interior(d, i, j)) {
        int si, sj;
        if (!lbc_src(d, 'T', i, j, si, sj)) at32(f, p) = 0.;
        return;
    }
    double v = at32(f, p);
    ring_of(d, i, j, [&](int ri, int rj) {
        const uint32_t q = IJK32(ri, rj, k);
        v += at32(f, q);
        at32(f, q) = 0.;
    });
    at32(f, p) = v;
}

}  // namespace

int launch_fct_ad(const double *pun, const double *pvn, const double *pwn, int kn_h, int kn_v, const double *cx,
                  const double *cy, const double *cz, const double *ta_ad, double *tb_ad, double *tn_ad)
{
    const DevDomain &d = nh::dom();
    Grid2 g = grid2(d.jpi, d.jpj);
    g.grid.z = (unsigned)(d.jpk - 1);
    double *zl_u = (double *)nh::scratch3(nh::S_ZWI), *zl_v = (double *)nh::scratch3(nh::S_ZWX);
    double *hi = (double *)nh::scratch3(nh::S_ZBETUP), *lo = (double *)nh::scratch3(nh::S_ZBETDO);
    const FctAd a{pun, pvn, pwn, cx, cy, cz, ta_ad, zl_u, zl_v, hi, lo, kn_h, kn_v};
    Grid2 gcell = grid2(d.jpi - 2, d.jpj - 2);
    gcell.grid.z = (unsigned)(d.jpk - 1);
    LAUNCH("tra_adv_fct_ad", k_fct_ad_cell, gcell, d, ta_ad, hi, lo);
    if (kn_h == 4) LAUNCH("tra_adv_fct_ad", k_fct_ad_l4, g, d, a, zl_u, zl_v);
    LAUNCH("tra_adv_fct_ad", k_fct_ad, g, d, a, tb_ad, tn_ad);
    if (kn_v == 4) {
        Grid2 gc = grid2(d.jpi, d.jpj);
        LAUNCH("tra_adv_fct_ad", k_fct_ad_v4, gc, d, a, tn_ad, (double *)nh::scratch3(nh::S_ZWY),
               (double *)nh::scratch3(nh::S_ZWZ));
    }
    return 0;
}

int launch_ldf_lap_ad(const double *pahu, const double *pahv, const double *ta_ad, double *tb_ad)
{
    const DevDomain &d = nh::dom();
    Grid2 g = grid2(d.jpi, d.jpj);
    g.grid.z = (unsigned)(d.jpk - 1);
    LAUNCH("tra_ldf_lap_ad", k_ldf_lap_ad, g, d, pahu, pahv, ta_ad, tb_ad);
    return 0;
}

int launch_zdf_imp_ad(double p2dt, const double *avt, double *ta_ad, double *tb_ad)
{
    const DevDomain &d = nh::dom();
    Grid2 g = grid2(d.jpi, d.jpj);
    LAUNCH("tra_zdf_imp_ad", k_zdf_imp_ad, g, d, p2dt, avt, ta_ad, tb_ad, (double *)nh::scratch3(nh::S_ZWT));
    return 0;
}

int launch_tra_nxt_ad(double *tb_ad, double *tn_ad, double *ta_ad)
{
    const DevDomain &d = nh::dom();
    Grid2 g = grid2(d.jpi, d.jpj);
    g.grid.z = (unsigned)d.jpk;
    LAUNCH("tra_nxt_ad", k_tra_nxt_ad, g, d, d.rn_atfp, tb_ad, tn_ad, ta_ad);
    return 0;
}

int launch_lbc_adj(double *p, int nk)
{
    const DevDomain &d = nh::dom();
    Grid2 g = grid2(d.jpi, d.jpj);
    g.grid.z = (unsigned)nk;
    LAUNCH("lbc_adj", k_lbc_adj, g, d, p);
    return 0;
}
