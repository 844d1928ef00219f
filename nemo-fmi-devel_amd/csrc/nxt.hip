// nxt.hip -- the routines that close a time step around the hot path:
// horizontal divergence (DYN/divhor.F90), sea-level prediction, vertical
// velocity and sea-level swap (DYN/sshwzv.F90), the tracer and momentum
// Asselin filters and swaps (TRA/tranxt.F90, DYN/dynnxt.F90), for the linear
// free surface (e3 fixed in time).
//
// All of them are HBM-bound streams of a few fields.  Arithmetic is done in
// FP64 and rounded once on store, so the FP64 form repeats the reference's
// operations and the FP32 storage form loses nothing beyond its storage
// rounding.
//
// lbc_lnk is folded into the kernels.  A read-only input (un, vn for the
// divergence) lets a ring point compute its value at its lbc source point.
// Fields updated in place (tsb/tsn/tsa, ua) are written to the ring by the
// thread of the interior source point (lbc_targets), so no thread reads a
// value another thread of the same launch overwrites.
#include "common.hpp"
#include "kernels.hpp"

// the ring points whose serial lbc_lnk source ('T', 'U' or 'V' grid) is the
// interior point (i,j): the inverse of lbc_src, at most three points
template <typename F>
__device__ __forceinline__ void lbc_targets(const DevDomain &d, int i, int j, F f)
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

// horizontal divergence at the interior point (i,j), DYN/divhor.F90:79-83.
// The column's metrics are loaded once; at(k) evaluates one level with the
// reference's operation order.
template <typename T>
struct HdivCol {
    double a1, a0, b1, b0, r;
    uint32_t ij;
    const T *un, *vn;
    __device__ __forceinline__ HdivCol(const DevDomain &d, const T *un_, const T *vn_, int i, int j) : un(un_), vn(vn_)
    {
        const Metrics<double> &m = d.md;
        ij = IJ32(i, j);
        a1 = A2(m.e2u, i, j);
        a0 = A2(m.e2u, i - 1, j);
        b1 = A2(m.e1v, i, j);
        b0 = A2(m.e1v, i, j - 1);
        r = A2(m.r1_e1e2t, i, j);
    }
    __device__ __forceinline__ double at(const DevDomain &d, int k) const
    {
        const Metrics<double> &m = d.md;
        const T *u = plane_ptr(un, d.n2, k), *v = plane_ptr(vn, d.n2, k);
        return (a1 * E3(m.e3u, ij, k) * (double)ldp(u, ij) - a0 * E3(m.e3u, ij - 1, k) * (double)ldp(u, ij - 1) +
                b1 * E3(m.e3v, ij, k) * (double)ldp(v, ij) -
                b0 * E3(m.e3v, ij - (uint32_t)d.jpi, k) * (double)ldp(v, ij - (uint32_t)d.jpi)) *
               r / E3(m.e3t, ij, k);
    }
};

// levels per thread of the per-cell closure kernels
#define NXT_KB 8

// ---------------------------------------------------------------------------
// div_hor, DYN/divhor.F90:50-110: hdivn on levels 1..jpkm1 of the interior,
// then lbc_lnk('T') of all jpk levels.  One thread per cell.
// ---------------------------------------------------------------------------
template <typename T>
struct S_div_hor {
    static constexpr int NTHR = NH_BLOCK, WAVES = 1;
    __device__ __forceinline__ static void run(const VB &vb_, const DevDomain &d, const T *__restrict__ un, const T *__restrict__ vn, T *__restrict__ hdivn)
    {
        THREAD_IJ_V(vb_, 1, 1);
        const int k0 = (int)vb_.z * NXT_KB + 1;
        if (i > d.jpi || j > d.jpj) return;
        int si, sj;
        const bool ok = lbc_src(d, 'T', i, j, si, sj);
        const bool ring = si != i || sj != j || !ok;
        const HdivCol<T> c(d, un, vn, ok ? si : 2, ok ? sj : 2);
        const uint32_t ij = IJ32(i, j);
    #pragma unroll
        for (int q = 0; q < NXT_KB; q++) {
            const int k = k0 + q;
            if (k > d.jpk) break;
            T *h = plane_ptr(hdivn, d.n2, k);
            if (k == d.jpk) {   // level jpk is not computed: lbc_lnk only
                if (ring) stp(h, ij, ok ? ldp((const T *)h, c.ij) : T(0));
            } else {
                stp(h, ij, ok ? (T)c.at(d, k) : T(0));
            }
        }
    }
};
template <typename T>
KERNEL void k_div_hor(DevDomain d, const T *__restrict__ un, const T *__restrict__ vn, T *__restrict__ hdivn)
{
    S_div_hor<T>::run(vb_hw(), d, un, vn, hdivn);
}

// ---------------------------------------------------------------------------
// ssh_nxt, DYN/sshwzv.F90:56-126: div_hor, then the after sea level from the
// depth-integrated divergence.  One thread per column: the column computes
// its hdivn (at its lbc source point on the ring), the integral and ssha.
// ---------------------------------------------------------------------------
template <typename T>
struct S_ssh_nxt {
    static constexpr int NTHR = NH_BLOCK, WAVES = 1;
    __device__ __forceinline__ static void run(const VB &vb_, const DevDomain &d, double z2dt, double zcoef, const T *__restrict__ un, const T *__restrict__ vn, T *__restrict__ hdivn, const double *__restrict__ sshb, const double *__restrict__ emp_b, const double *__restrict__ emp, double *__restrict__ ssha)
    {
        THREAD_IJ_V(vb_, 1, 1);
        if (i > d.jpi || j > d.jpj) return;
        int si, sj;
        const bool ok = lbc_src(d, 'T', i, j, si, sj);
        const uint32_t ij = IJ32(i, j);
        const HdivCol<T> c(d, un, vn, ok ? si : 2, ok ? sj : 2);
        double zhdiv = 0.0;
    #pragma unroll 2
        for (int k = 1; k <= d.jpk - 1; k++) {
            const T h = ok ? (T)c.at(d, k) : T(0);
            stp(plane_ptr(hdivn, d.n2, k), ij, h);
            zhdiv = zhdiv + E3(d.md.e3t, ij, k) * (double)h;
        }
        if (si != i || sj != j || !ok) A3(hdivn, i, j, d.jpk) = ok ? A3(hdivn, si, sj, d.jpk) : T(0);
        ssha[ij] = (sshb[ij] - z2dt * (zcoef * (emp_b[ij] + emp[ij]) + zhdiv)) * SSMASK(ij);
    }
};
template <typename T>
KERNEL void k_ssh_nxt(DevDomain d, double z2dt, double zcoef, const T *__restrict__ un, const T *__restrict__ vn,
                      T *__restrict__ hdivn, const double *__restrict__ sshb, const double *__restrict__ emp_b,
                      const double *__restrict__ emp, double *__restrict__ ssha)
{
    S_ssh_nxt<T>::run(vb_hw(), d, z2dt, zcoef, un, vn, hdivn, sshb, emp_b, emp, ssha);
}

// ---------------------------------------------------------------------------
// wzv, DYN/sshwzv.F90:129-213 (z-star / linear free surface branch): w from
// the bottom up.  e3t_a - e3t_b is formed as the reference forms it (zero
// when the scale factors are fixed).  One thread per column.
// ---------------------------------------------------------------------------
template <typename T>
struct S_wzv {
    static constexpr int NTHR = NH_BLOCK, WAVES = 1;
    __device__ __forceinline__ static void run(const VB &vb_, const DevDomain &d, int first, double z1_2dt, const T *__restrict__ hdivn, T *__restrict__ wn)
    {
        THREAD_IJ_V(vb_, 1, 1);
        if (i > d.jpi || j > d.jpj) return;
        const uint32_t ij = IJ32(i, j);
        if (first) A3(wn, i, j, d.jpk) = T(0);
        double w = (double)A3(wn, i, j, d.jpk);
    #pragma unroll 4
        for (int k = d.jpk - 1; k >= 1; k--) {
            const double e3 = E3(d.md.e3t, ij, k);
            w = w - (e3 * (double)A3(hdivn, i, j, k) + z1_2dt * (e3 - e3)) * TMASK(double, i, j, k);
            A3(wn, i, j, k) = (T)w;
        }
    }
};
template <typename T>
KERNEL void k_wzv(DevDomain d, int first, double z1_2dt, const T *__restrict__ hdivn, T *__restrict__ wn)
{
    S_wzv<T>::run(vb_hw(), d, first, z1_2dt, hdivn, wn);
}

// ---------------------------------------------------------------------------
// ssh_swp, DYN/sshwzv.F90:216-267 (linear free surface)
// ---------------------------------------------------------------------------
struct S_ssh_swp {
    static constexpr int NTHR = NH_BLOCK, WAVES = 1;
    __device__ __forceinline__ static void run(const VB &vb_, const DevDomain &d, int euler, double atfp, double *__restrict__ sshb, double *__restrict__ sshn, const double *__restrict__ ssha)
    {
        THREAD_IJ_V(vb_, 1, 1);
        if (i > d.jpi || j > d.jpj) return;
        const uint32_t ij = IJ32(i, j);
        if (!euler) sshb[ij] = sshn[ij] + atfp * (sshb[ij] - 2 * sshn[ij] + ssha[ij]);
        sshn[ij] = ssha[ij];
    }
};
KERNEL void k_ssh_swp(DevDomain d, int euler, double atfp, double *__restrict__ sshb, double *__restrict__ sshn,
                      const double *__restrict__ ssha)
{
    S_ssh_swp::run(vb_hw(), d, euler, atfp, sshb, sshn, ssha);
}

// ---------------------------------------------------------------------------
// tra_nxt, TRA/tranxt.F90:66-192 with tra_nxt_fix, lines 195-237: lbc_lnk of
// tsa; Euler: tsn = tsa on levels 1..jpkm1; leapfrog: Asselin filter on the
// interior, then lbc_lnk of tsb, tsn, tsa.  One thread per cell and tracer
// (grid.z = jpk * 2).
// ---------------------------------------------------------------------------
template <typename T>
struct S_tra_nxt {
    static constexpr int NTHR = NH_BLOCK, WAVES = 1;
    __device__ __forceinline__ static void run(const VB &vb_, const DevDomain &d, int euler, double atfp, T *__restrict__ tsb, T *__restrict__ tsn, T *__restrict__ tsa)
    {
        THREAD_IJ_V(vb_, 1, 1);
        const int k = (int)(vb_.z % (unsigned)d.jpk) + 1;
        const size_t off = (size_t)(vb_.z / (unsigned)d.jpk) * d.n3;
        if (i > d.jpi || j > d.jpj) return;
        T *pb = tsb + off, *pn = tsn + off, *pa = tsa + off;
        if (!is_interior(d, i, j)) {
            int si, sj;
            if (lbc_src(d, 'T', i, j, si, sj)) return;   // written by the thread of (si,sj)
            A3(pa, i, j, k) = T(0);
            if (euler) {
                if (k < d.jpk) A3(pn, i, j, k) = T(0);
            } else {
                A3(pb, i, j, k) = T(0);
                A3(pn, i, j, k) = T(0);
            }
            return;
        }
        const T ta = A3(pa, i, j, k);
        T tb, tn;
        if (euler) {
            tb = A3(pb, i, j, k);   // only copied to ring points by the leapfrog branch
            tn = k < d.jpk ? ta : A3(pn, i, j, k);
            if (k < d.jpk) A3(pn, i, j, k) = tn;
        } else if (k < d.jpk) {
            const double ztn = (double)A3(pn, i, j, k);
            const double ztd = (double)ta - 2.0 * ztn + (double)A3(pb, i, j, k);
            tb = (T)(ztn + atfp * ztd);
            tn = ta;
            A3(pb, i, j, k) = tb;
            A3(pn, i, j, k) = tn;
        } else {
            tb = A3(pb, i, j, k);
            tn = A3(pn, i, j, k);
        }
        lbc_targets(d, i, j, [&](int ti, int tj) {
            A3(pa, ti, tj, k) = ta;
            if (euler) {
                if (k < d.jpk) A3(pn, ti, tj, k) = tn;
            } else {
                A3(pb, ti, tj, k) = tb;
                A3(pn, ti, tj, k) = tn;
            }
        });
    }
};
template <typename T>
KERNEL void k_tra_nxt(DevDomain d, int euler, double atfp, T *__restrict__ tsb, T *__restrict__ tsn,
                      T *__restrict__ tsa)
{
    S_tra_nxt<T>::run(vb_hw(), d, euler, atfp, tsb, tsn, tsa);
}

// ---------------------------------------------------------------------------
// dyn_nxt, DYN/dynnxt.F90:66-377 (linear free surface), in two column
// kernels.
// Part 1 (lines 111-135): with dyn_spg_ts the baroclinic ua/va get the
// barotropic after velocity (ua - depth mean + ua_b) on the interior, written
// also to the ring points the lbc_lnk('U'/'V') of line 139 fills from them;
// ring points without a source are zeroed.  Without ln_bt_fw, un/vn get the
// advective barotropic velocity (lines 127-133), every column on itself.
// ---------------------------------------------------------------------------
template <typename T>
struct S_dyn_nxt1 {
    static constexpr int NTHR = NH_BLOCK, WAVES = 1;
    __device__ __forceinline__ static void run(const VB &vb_, const DevDomain &d, int spg_ts, int bt_fw, T *__restrict__ ua, T *__restrict__ va, T *__restrict__ un, T *__restrict__ vn, const double *__restrict__ ua_b, const double *__restrict__ va_b, const double *__restrict__ un_adv, const double *__restrict__ vn_adv, const double *__restrict__ un_b, const double *__restrict__ vn_b, const double *__restrict__ r1_hu, const double *__restrict__ r1_hv)
    {
        THREAD_IJ_V(vb_, 1, 1);
        if (i > d.jpi || j > d.jpj) return;
        const Metrics<double> &m = d.md;
        const uint32_t ij = IJ32(i, j);
        const int jpk = d.jpk;
        if (spg_ts && !bt_fw) {
            const double cu = un_adv[ij] * r1_hu[ij], cv = vn_adv[ij] * r1_hv[ij];
            for (int k = 1; k <= jpk - 1; k++) {
                A3(un, i, j, k) = (T)(((double)A3(un, i, j, k) - cu + un_b[ij]) * UMASK(double, i, j, k));
                A3(vn, i, j, k) = (T)(((double)A3(vn, i, j, k) - cv + vn_b[ij]) * VMASK(double, i, j, k));
            }
        }
        if (!is_interior(d, i, j)) {
            int si, sj;
            if (lbc_src(d, 'U', i, j, si, sj)) return;   // written by the thread of (si,sj)
            for (int k = 1; k <= jpk; k++) {
                A3(ua, i, j, k) = T(0);
                A3(va, i, j, k) = T(0);
            }
            return;
        }
        double zue = 0.0, zve = 0.0;
        if (spg_ts) {
            zue = E3(m.e3u, ij, 1) * (double)A3(ua, i, j, 1) * UMASK(double, i, j, 1);
            zve = E3(m.e3v, ij, 1) * (double)A3(va, i, j, 1) * VMASK(double, i, j, 1);
    #pragma unroll 4
            for (int k = 2; k <= jpk - 1; k++) {
                zue = zue + E3(m.e3u, ij, k) * (double)A3(ua, i, j, k) * UMASK(double, i, j, k);
                zve = zve + E3(m.e3v, ij, k) * (double)A3(va, i, j, k) * VMASK(double, i, j, k);
            }
            zue = zue * r1_hu[ij];
            zve = zve * r1_hv[ij];
        }
    #pragma unroll 4
        for (int k = 1; k <= jpk; k++) {
            T a = A3(ua, i, j, k), b = A3(va, i, j, k);
            if (spg_ts && k < jpk) {
                a = (T)(((double)a - zue + ua_b[ij]) * UMASK(double, i, j, k));
                b = (T)(((double)b - zve + va_b[ij]) * VMASK(double, i, j, k));
                A3(ua, i, j, k) = a;
                A3(va, i, j, k) = b;
            }
            lbc_targets(d, i, j, [&](int ti, int tj) {
                A3(ua, ti, tj, k) = a;
                A3(va, ti, tj, k) = b;
            });
        }
    }
};
template <typename T>
KERNEL void k_dyn_nxt1(DevDomain d, int spg_ts, int bt_fw, T *__restrict__ ua, T *__restrict__ va,
                       T *__restrict__ un, T *__restrict__ vn, const double *__restrict__ ua_b,
                       const double *__restrict__ va_b, const double *__restrict__ un_adv,
                       const double *__restrict__ vn_adv, const double *__restrict__ un_b,
                       const double *__restrict__ vn_b, const double *__restrict__ r1_hu, const double *__restrict__ r1_hv)
{
    S_dyn_nxt1<T>::run(vb_hw(), d, spg_ts, bt_fw, ua, va, un, vn, ua_b, va_b, un_adv, vn_adv, un_b, vn_b, r1_hu, r1_hv);
}

// Part 2 (lines 147-176, 286-297, 299-347): Euler swap or Asselin filter and
// swap on every column, the forward time-splitting correction of ub/vb, and
// the barotropic now and before velocities.  Each column reads and writes
// only itself.  The depth mean of part 1 multiplies by r1_hu once (zue *
// r1_hu_a) exactly as lines 119-120 do, and the same holds for line 292.
template <typename T>
struct S_dyn_nxt2 {
    static constexpr int NTHR = NH_BLOCK, WAVES = 1;
    __device__ __forceinline__ static void run(const VB &vb_, const DevDomain &d, int euler, int bt_fix, double atfp, T *__restrict__ ub, T *__restrict__ vb, T *__restrict__ un, T *__restrict__ vn, const T *__restrict__ ua, const T *__restrict__ va, double *__restrict__ ub_b, double *__restrict__ vb_b, double *__restrict__ un_b, double *__restrict__ vn_b, const double *__restrict__ r1_hu, const double *__restrict__ r1_hv)
    {
        THREAD_IJ_V(vb_, 1, 1);
        if (i > d.jpi || j > d.jpj) return;
        const Metrics<double> &m = d.md;
        const uint32_t ij = IJ32(i, j);
        const int jpkm1 = d.jpk - 1;
        double sun = 0.0, svn = 0.0, sub = 0.0, svb = 0.0;   // depth sums of un, vn, ub, vb
    #pragma unroll 2
        for (int k = 1; k <= jpkm1; k++) {
            const double a = (double)A3(ua, i, j, k), b = (double)A3(va, i, j, k);
            const double e3u = E3(m.e3u, ij, k), e3v = E3(m.e3v, ij, k);
            const double mu = UMASK(double, i, j, k), mv = VMASK(double, i, j, k);
            double fu, fv;
            if (euler) {
                fu = (double)A3(ub, i, j, k);
                fv = (double)A3(vb, i, j, k);
            } else {
                const double n1 = (double)A3(un, i, j, k), n2 = (double)A3(vn, i, j, k);
                const T zuf = (T)(n1 + atfp * ((double)A3(ub, i, j, k) - 2.0 * n1 + a));
                const T zvf = (T)(n2 + atfp * ((double)A3(vb, i, j, k) - 2.0 * n2 + b));
                A3(ub, i, j, k) = zuf;
                A3(vb, i, j, k) = zvf;
                fu = (double)zuf;
                fv = (double)zvf;
            }
            A3(un, i, j, k) = A3(ua, i, j, k);
            A3(vn, i, j, k) = A3(va, i, j, k);
            const double tu = e3u * (double)A3(ua, i, j, k) * mu, tv = e3v * (double)A3(va, i, j, k) * mv;
            const double bu = e3u * fu * mu, bv = e3v * fv * mv;
            sun = k == 1 ? tu : sun + tu;
            svn = k == 1 ? tv : svn + tv;
            sub = k == 1 ? bu : sub + bu;
            svb = k == 1 ? bv : svb + bv;
        }
        if (!euler && bt_fix) {
            // lines 286-297: the depth mean of ub/vb set to the input un_b/vn_b
            const double cu = sub * r1_hu[ij] - un_b[ij], cv = svb * r1_hv[ij] - vn_b[ij];
    #pragma unroll 4
            for (int k = 1; k <= jpkm1; k++) {
                const double e3u = E3(m.e3u, ij, k), e3v = E3(m.e3v, ij, k);
                const double mu = UMASK(double, i, j, k), mv = VMASK(double, i, j, k);
                const T nb = (T)((double)A3(ub, i, j, k) - cu * mu);
                const T nv = (T)((double)A3(vb, i, j, k) - cv * mv);
                A3(ub, i, j, k) = nb;
                A3(vb, i, j, k) = nv;
                const double bu = e3u * (double)nb * mu, bv = e3v * (double)nv * mv;
                sub = k == 1 ? bu : sub + bu;
                svb = k == 1 ? bv : svb + bv;
            }
        }
        un_b[ij] = sun * r1_hu[ij];
        vn_b[ij] = svn * r1_hv[ij];
        ub_b[ij] = sub * r1_hu[ij];
        vb_b[ij] = svb * r1_hv[ij];
    }
};
template <typename T>
KERNEL void k_dyn_nxt2(DevDomain d, int euler, int bt_fix, double atfp, T *__restrict__ ub, T *__restrict__ vb,
                       T *__restrict__ un, T *__restrict__ vn, const T *__restrict__ ua, const T *__restrict__ va,
                       double *__restrict__ ub_b, double *__restrict__ vb_b, double *__restrict__ un_b,
                       double *__restrict__ vn_b, const double *__restrict__ r1_hu, const double *__restrict__ r1_hv)
{
    S_dyn_nxt2<T>::run(vb_hw(), d, euler, bt_fix, atfp, ub, vb, un, vn, ua, va, ub_b, vb_b, un_b, vn_b, r1_hu, r1_hv);
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
template <typename T>
int launch_div_hor(const T *un, const T *vn, T *hdivn)
{
    const DevDomain &d = nh::dom();
    Grid2 g = grid2(d.jpi, d.jpj);
    g.grid.z = (unsigned)((d.jpk + NXT_KB - 1) / NXT_KB);
    LAUNCHB("div_hor", k_div_hor<T>, (S_div_hor<T>), g, d, un, vn, hdivn);
    return 0;
}

template <typename T>
int launch_ssh_nxt(double z2dt, double zcoef, const T *un, const T *vn, T *hdivn, const double *sshb,
                   const double *emp_b, const double *emp, double *ssha)
{
    const DevDomain &d = nh::dom();
    Grid2 g = grid2(d.jpi, d.jpj);
    LAUNCHB("ssh_nxt", k_ssh_nxt<T>, (S_ssh_nxt<T>), g, d, z2dt, zcoef, un, vn, hdivn, sshb, emp_b, emp, ssha);
    return 0;
}

template <typename T>
int launch_wzv(int first, double z1_2dt, const T *hdivn, T *wn)
{
    const DevDomain &d = nh::dom();
    Grid2 g = grid2(d.jpi, d.jpj);
    LAUNCHB("wzv", k_wzv<T>, (S_wzv<T>), g, d, first, z1_2dt, hdivn, wn);
    return 0;
}

int launch_ssh_swp(int euler, double *sshb, double *sshn, const double *ssha)
{
    const DevDomain &d = nh::dom();
    Grid2 g = grid2(d.jpi, d.jpj);
    LAUNCHB("ssh_swp", k_ssh_swp, (S_ssh_swp), g, d, euler, d.rn_atfp, sshb, sshn, ssha);
    return 0;
}

template <typename T>
int launch_tra_nxt(int euler, T *tsb, T *tsn, T *tsa)
{
    const DevDomain &d = nh::dom();
    Grid2 g = grid2(d.jpi, d.jpj);
    g.grid.z = (unsigned)(2 * d.jpk);
    LAUNCHB("tra_nxt", k_tra_nxt<T>, (S_tra_nxt<T>), g, d, euler, d.rn_atfp, tsb, tsn, tsa);
    return 0;
}

template <typename T>
int launch_dyn_nxt(int euler, T *ub, T *vb, T *un, T *vn, T *ua, T *va, const double *ua_b, const double *va_b,
                   const double *un_adv, const double *vn_adv, double *ub_b, double *vb_b, double *un_b,
                   double *vn_b, const double *r1_hu, const double *r1_hv)
{
    const DevDomain &d = nh::dom();
    Grid2 g = grid2(d.jpi, d.jpj);
    LAUNCHB("dyn_nxt", k_dyn_nxt1<T>, (S_dyn_nxt1<T>), g, d, d.ln_dynspg_ts, d.ln_bt_fw, ua, va, un, vn, ua_b, va_b, un_adv, vn_adv,
           (const double *)un_b, (const double *)vn_b, r1_hu, r1_hv);
    LAUNCHB("dyn_nxt", k_dyn_nxt2<T>, (S_dyn_nxt2<T>), g, d, euler, d.ln_dynspg_ts && d.ln_bt_fw, d.rn_atfp, ub, vb, un, vn,
           (const T *)ua, (const T *)va, ub_b, vb_b, un_b, vn_b, r1_hu, r1_hv);
    return 0;
}

// ---------------------------------------------------------------------------
// stp_ctl, stpctl.F90:113-121: the six maxima as a two-pass reduction
// (per-block partials, then one block), NaN-propagating; and, when a test
// fails, the first Fortran-order position of each reported extremum
// (atomic minimum of the linear index over the matching cells)
// ---------------------------------------------------------------------------
__device__ __forceinline__ double nanmax(double a, double b) { return (a > b || a != a) ? a : b; }

template <typename T>
KERNEL void k_stp_max(DevDomain d, const double *__restrict__ sshn, const T *__restrict__ un,
                      const T *__restrict__ tsn, double *__restrict__ part)
{
    __shared__ double red[6][NH_BLOCK];
    double m[6];
#pragma unroll
    for (int c = 0; c < 6; c++) m[c] = -INFINITY;
    const size_t n2 = d.n2, n3 = d.n3;
    const size_t stride = (size_t)gridDim.x * NH_BLOCK;
    for (size_t q = (size_t)blockIdx.x * NH_BLOCK + threadIdx.x; q < n3; q += stride) {
        if (q < n2) m[0] = nanmax(m[0], fabs(sshn[q]));
        m[1] = nanmax(m[1], fabs((double)un[q]));
        const int k = (int)(q / n2) + 1;
        if (pmask<double, 0>(d.ktu, (uint32_t)(q % n2), k) == 1.0) {
            const double sv = (double)tsn[n3 + q], tv = (double)tsn[q];
            m[2] = nanmax(m[2], -sv);
            m[3] = nanmax(m[3], sv);
            m[4] = nanmax(m[4], -tv);
            m[5] = nanmax(m[5], tv);
        }
    }
#pragma unroll
    for (int c = 0; c < 6; c++) red[c][threadIdx.x] = m[c];
    __syncthreads();
    for (int h = NH_BLOCK / 2; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h)
#pragma unroll
            for (int c = 0; c < 6; c++) red[c][threadIdx.x] = nanmax(red[c][threadIdx.x], red[c][threadIdx.x + h]);
        __syncthreads();
    }
    if (threadIdx.x < 6) part[(size_t)blockIdx.x * 6 + threadIdx.x] = red[threadIdx.x][0];
}

KERNEL void k_stp_final(const double *__restrict__ part, int nb, double *__restrict__ zmax,
                        uint32_t *__restrict__ loc)
{
    __shared__ double red[NH_BLOCK];
    for (int c = 0; c < 6; c++) {
        double m = -INFINITY;
        for (int b = (int)threadIdx.x; b < nb; b += NH_BLOCK) m = nanmax(m, part[(size_t)b * 6 + c]);
        red[threadIdx.x] = m;
        __syncthreads();
        for (int h = NH_BLOCK / 2; h > 0; h >>= 1) {
            if ((int)threadIdx.x < h) red[threadIdx.x] = nanmax(red[threadIdx.x], red[threadIdx.x + h]);
            __syncthreads();
        }
        if (threadIdx.x == 0) zmax[c] = red[0];
        __syncthreads();
    }
    if (threadIdx.x < 4) loc[threadIdx.x] = 0xffffffffu;
}

// loc[0]: |ssh| max, loc[1]: |u| max, loc[2]: S min, loc[3]: S max (linear indices)
__device__ __forceinline__ bool hit(double v, double m) { return v == m || (m != m && v != v); }
template <typename T>
KERNEL void k_stp_loc(DevDomain d, const double *__restrict__ sshn, const T *__restrict__ un,
                      const T *__restrict__ tsn, const double *__restrict__ zmax, uint32_t *__restrict__ loc)
{
    const size_t n2 = d.n2, n3 = d.n3;
    const size_t stride = (size_t)gridDim.x * NH_BLOCK;
    const double m0 = zmax[0], m1 = zmax[1], m2 = zmax[2], m3 = zmax[3];
    for (size_t q = (size_t)blockIdx.x * NH_BLOCK + threadIdx.x; q < n3; q += stride) {
        if (q < n2) {
            const double s = sshn[q];
            if (hit(fabs(s), m0) || (m0 != m0 && s != s)) atomicMin(&loc[0], (uint32_t)q);
        }
        const double u = (double)un[q];
        if (hit(fabs(u), m1) || (m1 != m1 && u != u)) atomicMin(&loc[1], (uint32_t)q);
        const int k = (int)(q / n2) + 1;
        if (pmask<double, 0>(d.ktu, (uint32_t)(q % n2), k) == 1.0) {
            const double sv = (double)tsn[n3 + q];
            if (hit(-sv, m2) || (m2 != m2 && sv != sv)) atomicMin(&loc[2], (uint32_t)q);
            if (hit(sv, m3) || (m3 != m3 && sv != sv)) atomicMin(&loc[3], (uint32_t)q);
        }
    }
}

#define STP_NB 1024
template <typename T>
int launch_stp_ctl(const double *sshn, const T *un, const T *tsn, double *buf, int want_loc)
{
    const DevDomain &d = nh::dom();
    double *part = buf, *zmax = buf + (size_t)STP_NB * 6;
    uint32_t *loc = (uint32_t *)(zmax + 8);
    Grid2 g;
    g.block = dim3(NH_BLOCK, 1, 1);
    g.grid = dim3(STP_NB, 1, 1);
    if (!want_loc) {
        LAUNCH("stp_ctl", k_stp_max<T>, g, d, sshn, un, tsn, part);
        Grid2 g1;
        g1.block = dim3(NH_BLOCK, 1, 1);
        g1.grid = dim3(1, 1, 1);
        LAUNCH("stp_ctl", k_stp_final, g1, (const double *)part, STP_NB, zmax, loc);
    } else {
        LAUNCH("stp_ctl", k_stp_loc<T>, g, d, sshn, un, tsn, (const double *)zmax, loc);
    }
    return 0;
}
size_t stp_ctl_buffer_bytes() { return ((size_t)STP_NB * 6 + 8) * sizeof(double) + 4 * sizeof(uint32_t); }

// ---------------------------------------------------------------------------
// tra_sbc, TRA/trasbc.F90:97-160 (linear free surface): one
// thread per (i,j); every point swaps or resets its trends, the points of
// 2..jpim1 x 2..jpj get the new trend and update level 1 of tsa
// ---------------------------------------------------------------------------
template <typename T>
KERNEL void k_tra_sbc(DevDomain d, int first, double *__restrict__ qns, double *__restrict__ qsr,
                      const double *__restrict__ sfx, const double *__restrict__ emp, const T *__restrict__ tsn,
                      T *__restrict__ tsa, double *__restrict__ sbc, double *__restrict__ sbc_b)
{
    THREAD_IJ(1, 1);
    if (i > d.jpi || j > d.jpj) return;
    const uint32_t ij = IJ32(i, j);
    const size_t n2 = d.n2, n3 = d.n3;
    const double rcp = 3991.86795711963, r1_rau0 = 1.0 / d.rau0, r1_rau0_rcp = 1.0 / (d.rau0 * rcp);
    // lines 97-100: without penetrative solar flux the whole of qsr enters
    // the surface level through qns; with it, tra_qsr takes qsr
    double q = qns[ij];
    if (!d.ln_traqsr) {
        q = q + qsr[ij];
        qns[ij] = q;
        qsr[ij] = 0.0;
    }
    double bt, bs, ct, cs;
    if (first) {
        bt = bs = ct = cs = 0.0;
    } else {
        bt = sbc[ij];
        bs = sbc[n2 + ij];
        ct = bt;
        cs = bs;
    }
    const double zfact = first ? 1.0 : 0.5;
    if (j >= 2 && i >= 2 && i <= d.jpi - 1) {
        ct = r1_rau0_rcp * q;
        cs = r1_rau0 * sfx[ij];
        ct = ct + r1_rau0 * emp[ij] * (double)tsn[ij];
        cs = cs + r1_rau0 * emp[ij] * (double)tsn[n3 + ij];
        const double e3 = E3(d.md.e3t, ij, 1);
        tsa[ij] = (T)((double)tsa[ij] + zfact * (bt + ct) / e3);
        tsa[n3 + ij] = (T)((double)tsa[n3 + ij] + zfact * (bs + cs) / e3);
    }
    sbc_b[ij] = bt;
    sbc_b[n2 + ij] = bs;
    sbc[ij] = ct;
    sbc[n2 + ij] = cs;
}

// ---------------------------------------------------------------------------
// tra_qsr, two bands (ln_qsr_2bd), TRA/traqsr.F90:134-276: one thread per
// (i,j) column.  Every column swaps qsr_hc into qsr_hc_b (or clears it at an
// Euler start); the interior ones form qsr_hc on levels 1..nksr, add the
// trend to tsa and set fraqsr_1lev.  The attenuation of w-level k+1 for
// level k is the attenuation of w-level k for level k+1 (the same
// expression of the same depth), so one pair of exponentials per level.
// mode 0: swap, 1: Euler start (z1_2 = 1), 2: qsr_hc_b given (a restart).
// ---------------------------------------------------------------------------
template <typename T>
KERNEL void k_tra_qsr(DevDomain d, int mode, int nksr, double zz0, double zz1, double xsi0r, double xsi1r,
                      const double *__restrict__ qsr, const T *__restrict__ gdepw, double *__restrict__ hc,
                      double *__restrict__ hc_b, T *__restrict__ tsa, double *__restrict__ fraqsr)
{
    THREAD_IJ(1, 1);
    if (i > d.jpi || j > d.jpj) return;
    if (mode != 2)
        for (int k = 1; k <= d.jpk; k++) A3(hc_b, i, j, k) = mode == 0 ? A3(hc, i, j, k) : 0.0;
    if (i < 2 || i > d.jpi - 1 || j < 2 || j > d.jpj - 1) return;
    const uint32_t ij = IJ32(i, j);
    const double rcp = 3991.86795711963, r1_rau0_rcp = 1.0 / (d.rau0 * rcp);
    const double z1_2 = mode == 1 ? 1.0 : 0.5;
    const double q = qsr[ij];
    double zg = (double)A3(gdepw, i, j, 1);
    double zc0 = zz0 * exp(-zg * xsi0r) + zz1 * exp(-zg * xsi1r);
    for (int k = 1; k <= nksr; k++) {
        zg = (double)A3(gdepw, i, j, k + 1);
        const double zc1 = zz0 * exp(-zg * xsi0r) + zz1 * exp(-zg * xsi1r);
        const double h = q * (zc0 * WMASK(double, i, j, k) - zc1 * WMASK(double, i, j, k + 1));
        const uint32_t p = IJK32(i, j, k);
        at32(hc, p) = h;
        at32(tsa, p) = (T)((double)at32(tsa, p) + z1_2 * (at32(hc_b, p) + h) / E3(d.md.e3t, ij, k));
        if (k == 1) fraqsr[ij] = (q != 0.0) ? h / (r1_rau0_rcp * q) : 1.0;
        zc0 = zc1;
    }
}

template <typename T>
int launch_tra_qsr(int mode, int nksr, double rn_abs, double rn_si0, double rn_si1, const double *qsr,
                   const T *gdepw, double *qsr_hc, double *qsr_hc_b, T *tsa, double *fraqsr_1lev)
{
    const DevDomain &d = nh::dom();
    const double rcp = 3991.86795711963, r1_rau0_rcp = 1.0 / (d.rau0 * rcp);
    const double zz0 = rn_abs * r1_rau0_rcp, zz1 = (1. - rn_abs) * r1_rau0_rcp;
    Grid2 g = grid2(d.jpi, d.jpj);
    LAUNCH("tra_qsr", k_tra_qsr<T>, g, d, mode, nksr, zz0, zz1, 1.0 / rn_si0, 1.0 / rn_si1, qsr, gdepw, qsr_hc,
           qsr_hc_b, tsa, fraqsr_1lev);
    return launch_lbc<double>(fraqsr_1lev, 1, 'T', 0.0);
}

template <typename T>
int launch_tra_sbc(int first, double *qns, double *qsr, const double *sfx, const double *emp, const T *tsn, T *tsa,
                   double *sbc, double *sbc_b)
{
    const DevDomain &d = nh::dom();
    Grid2 g = grid2(d.jpi, d.jpj);
    LAUNCH("tra_sbc", k_tra_sbc<T>, g, d, first, qns, qsr, sfx, emp, tsn, tsa, sbc, sbc_b);
    return 0;
}

// ---------------------------------------------------------------------------
// the depth-mean velocities after a restart, DOM/istate.F90:159-178: every
// column, levels 1..jpkm1 summed in order from zero (linear free surface:
// the before scale factors and inverse depths are the now ones)
// ---------------------------------------------------------------------------
template <typename T>
KERNEL void k_istate_bt(DevDomain d, const T *__restrict__ un, const T *__restrict__ vn, const T *__restrict__ ub,
                        const T *__restrict__ vb, const double *__restrict__ r1_hu, const double *__restrict__ r1_hv,
                        double *__restrict__ un_b, double *__restrict__ vn_b, double *__restrict__ ub_b,
                        double *__restrict__ vb_b)
{
    THREAD_IJ(1, 1);
    if (i > d.jpi || j > d.jpj) return;
    const Metrics<double> &m = d.md;
    const uint32_t ij = IJ32(i, j);
    double a = 0., b = 0., c = 0., e = 0.;
    for (int k = 1; k <= d.jpk - 1; k++) {
        const double eu = E3(m.e3u, ij, k), ev = E3(m.e3v, ij, k);
        const double um = UMASK(double, i, j, k), vm = VMASK(double, i, j, k);
        const uint32_t p = IJK32(i, j, k);
        a = a + eu * (double)at32(un, p) * um;
        b = b + ev * (double)at32(vn, p) * vm;
        c = c + eu * (double)at32(ub, p) * um;
        e = e + ev * (double)at32(vb, p) * vm;
    }
    un_b[ij] = a * r1_hu[ij];
    vn_b[ij] = b * r1_hv[ij];
    ub_b[ij] = c * r1_hu[ij];
    vb_b[ij] = e * r1_hv[ij];
}

template <typename T>
int launch_istate_bt(const T *un, const T *vn, const T *ub, const T *vb, const double *r1_hu, const double *r1_hv,
                     double *un_b, double *vn_b, double *ub_b, double *vb_b)
{
    const DevDomain &d = nh::dom();
    Grid2 g = grid2(d.jpi, d.jpj);
    LAUNCH("rst_read", k_istate_bt<T>, g, d, un, vn, ub, vb, r1_hu, r1_hv, un_b, vn_b, ub_b, vb_b);
    return 0;
}

#define INST(T)                                                                                                  \
    template int launch_istate_bt<T>(const T *, const T *, const T *, const T *, const double *, const double *,  \
                                     double *, double *, double *, double *);                                    \
    template int launch_div_hor<T>(const T *, const T *, T *);                                                   \
    template int launch_ssh_nxt<T>(double, double, const T *, const T *, T *, const double *, const double *,    \
                                   const double *, double *);                                                    \
    template int launch_wzv<T>(int, double, const T *, T *);                                                     \
    template int launch_tra_nxt<T>(int, T *, T *, T *);                                                          \
    template int launch_dyn_nxt<T>(int, T *, T *, T *, T *, T *, T *, const double *, const double *,            \
                                   const double *, const double *, double *, double *, double *, double *,      \
                                   const double *, const double *);                                              \
    template int launch_stp_ctl<T>(const double *, const T *, const T *, double *, int);                        \
    template int launch_tra_qsr<T>(int, int, double, double, double, const double *, const T *, double *,      \
                                   double *, T *, double *);                                                \
    template int launch_tra_sbc<T>(int, double *, double *, const double *, const double *, const T *, T *,       \
                                   double *, double *);
INST(double)
INST(float)
#undef INST
