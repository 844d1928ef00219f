// slp.hip -- the mixed-layer level (ZDF/zdfmxl.F90) and the iso-neutral
// slopes of the standard (Madec) operator (LDF/ldfslp.F90: ldf_slp and
// ldf_slp_mxl), for z-coordinates without partial steps or ice-shelf
// cavities (risfdep = 0, mikt = 1).
//
// Arithmetic is FP64 in both storage forms, rounded once on store, in the
// reference's operation order, so the FP64 form reproduces it bit for bit.
//
// The reference builds the unfiltered slopes in work arrays, applies
// lbc_lnk to them, filters and applies lbc_lnk again.  Here the unfiltered
// slopes go to two FP64 scratch arrays with their ring already equal to what
// lbc_lnk would give (a ring point evaluates the interior formula at its
// lbc source, or is zero on a closed boundary), and the filter kernel does
// the same for the output's ring.  The horizontal and vertical density
// gradients (zgru, zgrv, zdzr) are recomputed from prd and pn2 where used.
#include "common.hpp"
#include "kernels.hpp"

namespace {

constexpr double ZEPS = 1.e-20;

template <typename T>
struct SlpIn {
    const T *prd, *pn2;
};

// zgru / zgrv, LDF/ldfslp.F90:138-145 (defined for i <= jpim1, j <= jpjm1)
template <typename T>
__device__ __forceinline__ double gru(const DevDomain &d, const T *prd, int i, int j, int k)
{
    return UMASK(double, i, j, k) * ((double)A3(prd, i + 1, j, k) - (double)A3(prd, i, j, k));
}
template <typename T>
__device__ __forceinline__ double grv(const DevDomain &d, const T *prd, int i, int j, int k)
{
    return VMASK(double, i, j, k) * ((double)A3(prd, i, j + 1, k) - (double)A3(prd, i, j, k));
}
// zdzr at t-points, lines 162-170 (zero at level 1)
template <typename T>
__device__ __forceinline__ double dzr(const DevDomain &d, const T *prd, const T *pn2, int i, int j, int k)
{
    if (k == 1) return 0.;
    const double zm1_g = -1.0 / d.grav;
    return zm1_g * ((double)A3(prd, i, j, k) + 1.) * ((double)A3(pn2, i, j, k) + (double)A3(pn2, i, j, k + 1)) *
           (1. - 0.5 * TMASK(double, i, j, k + 1));
}
__device__ __forceinline__ double min3(double a, double b, double c) { return fmin(fmin(a, b), c); }
__device__ __forceinline__ double max2(double a, double b) { return a > b ? a : b; }

// omlmask, lines 660-668: 1 inside the mixed layer (k <= nmln - 1)
__device__ __forceinline__ double oml(const int32_t *nmln, uint32_t ij, int k) { return k <= nmln[ij] - 1 ? 1. : 0.; }

// ---------------------------------------------------------------------------
// zdf_mxl, ZDF/zdfmxl.F90:92-125: one thread per column
// ---------------------------------------------------------------------------
template <typename T>
KERNEL void k_zdf_mxl(DevDomain d, const T *__restrict__ rn2b, const T *__restrict__ gdept,
                      const T *__restrict__ gdepw, int nlb10, double zN2_c, int32_t *__restrict__ nmln,
                      double *__restrict__ hmlp, double *__restrict__ hmlpt)
{
    THREAD_IJ(1, 1);
    if (i > d.jpi || j > d.jpj) return;
    const uint32_t ij = IJ32(i, j);
    const int ikt = d.mbkt[ij];
    double h = 0.;
    int nm = nlb10;
    for (int k = nlb10; k <= d.jpk - 1; k++) {
        h = h + max2((double)A3(rn2b, i, j, k), 0.) * E3(d.md.e3w, ij, k);
        if (h < zN2_c) nm = (k < ikt ? k : ikt) + 1;
    }
    const double ss = d.ssmask[ij];
    nmln[ij] = nm;
    hmlp[ij] = (double)A3(gdepw, i, j, nm) * ss;
    hmlpt[ij] = (double)A3(gdept, i, j, nm - 1) * ss;
}

// ---------------------------------------------------------------------------
// ldf_slp_mxl, LDF/ldfslp.F90:682-725, per interior column: the u/v slopes
// just below the mixed layer already divided by the depth of its last
// t-point (zslpml_hmlpu/v, lines 188-195) and the w slopes wslpiml/wslpjml
// ---------------------------------------------------------------------------
template <typename T>
KERNEL void k_slp_mxl(DevDomain d, SlpIn<T> in, const int32_t *__restrict__ nmln, const double *__restrict__ hmlpt,
                      double z1_slpmax, double *__restrict__ zsu, double *__restrict__ zsv, double *__restrict__ wiml,
                      double *__restrict__ wjml)
{
    THREAD_IJ(2, 2);
    if (i > d.jpi - 1 || j > d.jpj - 1) return;
    const Metrics<double> &m = d.md;
    const T *prd = in.prd, *pn2 = in.pn2;
    const uint32_t ij = IJ32(i, j);
    const int jpkm1 = d.jpk - 1;
    const int nm = nmln[ij], nme = nmln[ij + 1], nmn = nmln[ij + (uint32_t)d.jpi];
    const int iku = min(max(1, max(nm, nme)), jpkm1);
    const int ikv = min(max(1, max(nm, nmn)), jpkm1);
    double zbu = 0.5 * (dzr(d, prd, pn2, i, j, iku) + dzr(d, prd, pn2, i + 1, j, iku));
    double zbv = 0.5 * (dzr(d, prd, pn2, i, j, ikv) + dzr(d, prd, pn2, i, j + 1, ikv));
    const double zau = gru(d, prd, i, j, iku) * m.r1_e1u[ij];
    const double zav = grv(d, prd, i, j, ikv) * m.r1_e2v[ij];
    zbu = min3(zbu, -z1_slpmax * fabs(zau), -7.e+3 / E3(m.e3u, ij, iku) * fabs(zau));
    zbv = min3(zbv, -z1_slpmax * fabs(zav), -7.e+3 / E3(m.e3v, ij, ikv) * fabs(zav));
    const double uml = zau / (zbu - ZEPS) * UMASK(double, i, j, iku);
    const double vml = zav / (zbv - ZEPS) * VMASK(double, i, j, ikv);
    zsu[ij] = uml / max2(max2(hmlpt[ij], hmlpt[ij + 1]), 5.);
    zsv[ij] = vml / max2(max2(hmlpt[ij], hmlpt[ij + (uint32_t)d.jpi]), 5.);
    const int ik = min(nm + 1, d.jpk), ikm1 = max(1, ik - 1);
    const double zm1_2g = -0.5 / d.grav;
    const double zbw = zm1_2g * (double)A3(pn2, i, j, ik) * ((double)A3(prd, i, j, ik) + (double)A3(prd, i, j, ikm1) + 2.);
    const double zci = max2(UMASK(double, i - 1, j, ik) + UMASK(double, i, j, ik) + UMASK(double, i - 1, j, ikm1) +
                                UMASK(double, i, j, ikm1),
                            ZEPS) * m.e1t[ij];
    const double zcj = max2(VMASK(double, i, j - 1, ik) + VMASK(double, i, j, ik) + VMASK(double, i, j - 1, ikm1) +
                                VMASK(double, i, j, ikm1),
                            ZEPS) * m.e2t[ij];
    const double tm = TMASK(double, i, j, ik);
    const double zai = (gru(d, prd, i - 1, j, ik) + gru(d, prd, i, j, ik) + gru(d, prd, i - 1, j, ikm1) +
                        gru(d, prd, i, j, ikm1)) / zci * tm;
    const double zaj = (grv(d, prd, i, j - 1, ik) + grv(d, prd, i, j, ik) + grv(d, prd, i, j - 1, ikm1) +
                        grv(d, prd, i, j, ikm1)) / zcj * tm;
    const double e3w = E3(m.e3w, ij, ik);
    const double zbi = min3(zbw, -100. * fabs(zai), -7.e+3 / e3w * fabs(zai));
    const double zbj = min3(zbw, -100. * fabs(zaj), -7.e+3 / e3w * fabs(zaj));
    wiml[ij] = zai / (zbi - ZEPS) * tm;
    wjml[ij] = zaj / (zbj - ZEPS) * tm;
}

// ---------------------------------------------------------------------------
// I. unfiltered u/v slopes, lines 197-219 (zwz, zww), levels 2..jpkm1, every
// point: ring points take their lbc source's value (lbc_lnk, line 231)
// ---------------------------------------------------------------------------
template <typename T>
KERNEL void k_slp_uv_raw(DevDomain d, SlpIn<T> in, const T *__restrict__ gdept, const int32_t *__restrict__ nmln,
                         const double *__restrict__ zsu, const double *__restrict__ zsv, double z1_slpmax,
                         double *__restrict__ zwz, double *__restrict__ zww)
{
    THREAD_IJ(1, 1);
    const int k = (int)blockIdx.z + 2;
    if (i > d.jpi || j > d.jpj) return;
    const uint32_t p = IJK32(i, j, k);
    int si, sj;
    if (!lbc_src(d, 'U', i, j, si, sj)) {
        at32(zwz, p) = 0.;
        at32(zww, p) = 0.;
        return;
    }
    const Metrics<double> &m = d.md;
    const T *prd = in.prd, *pn2 = in.pn2;
    const uint32_t ij = IJ32(si, sj);
    const double zau = gru(d, prd, si, sj, k) * m.r1_e1u[ij];
    const double zav = grv(d, prd, si, sj, k) * m.r1_e2v[ij];
    const double dc = dzr(d, prd, pn2, si, sj, k);
    double zbu = 0.5 * (dc + dzr(d, prd, pn2, si + 1, sj, k));
    double zbv = 0.5 * (dc + dzr(d, prd, pn2, si, sj + 1, k));
    zbu = min3(zbu, -z1_slpmax * fabs(zau), -7.e+3 / E3(m.e3u, ij, k) * fabs(zau));
    zbv = min3(zbv, -z1_slpmax * fabs(zav), -7.e+3 / E3(m.e3v, ij, k) * fabs(zav));
    const double oc = oml(nmln, ij, k);
    const double zfi = max2(oc, oml(nmln, ij + 1, k));
    const double zfj = max2(oc, oml(nmln, ij + (uint32_t)d.jpi, k));
    const double gc = (double)A3(gdept, si, sj, k);
    const double zdepu = 0.5 * ((gc + (double)A3(gdept, si + 1, sj, k)) - E3(m.e3u, ij, d.miku[ij]));
    const double zdepv = 0.5 * ((gc + (double)A3(gdept, si, sj + 1, k)) - E3(m.e3v, ij, d.mikv[ij]));
    at32(zwz, p) = ((1. - zfi) * zau / (zbu - ZEPS) + zfi * zdepu * zsu[ij]) * UMASK(double, si, sj, k);
    at32(zww, p) = ((1. - zfj) * zav / (zbv - ZEPS) + zfj * zdepv * zsv[ij]) * VMASK(double, si, sj, k);
}

// II. unfiltered w slopes, lines 277-304, levels 2..jpkm1, every point (lbc_lnk 'T', line 315)
template <typename T>
KERNEL void k_slp_w_raw(DevDomain d, SlpIn<T> in, const T *__restrict__ gdepw, const int32_t *__restrict__ nmln,
                        const double *__restrict__ hmlp, const double *__restrict__ wiml,
                        const double *__restrict__ wjml, double *__restrict__ zwz, double *__restrict__ zww)
{
    THREAD_IJ(1, 1);
    const int k = (int)blockIdx.z + 2;
    if (i > d.jpi || j > d.jpj) return;
    const uint32_t p = IJK32(i, j, k);
    int si, sj;
    if (!lbc_src(d, 'T', i, j, si, sj)) {
        at32(zwz, p) = 0.;
        at32(zww, p) = 0.;
        return;
    }
    const Metrics<double> &m = d.md;
    const T *prd = in.prd, *pn2 = in.pn2;
    const uint32_t ij = IJ32(si, sj);
    const double zm1_2g = -0.5 / d.grav;
    const double zbw = zm1_2g * (double)A3(pn2, si, sj, k) *
                       ((double)A3(prd, si, sj, k) + (double)A3(prd, si, sj, k - 1) + 2.);
    const double zci = max2(UMASK(double, si - 1, sj, k) + UMASK(double, si, sj, k) + UMASK(double, si - 1, sj, k - 1) +
                                UMASK(double, si, sj, k - 1),
                            ZEPS) * m.e1t[ij];
    const double zcj = max2(VMASK(double, si, sj - 1, k) + VMASK(double, si, sj, k - 1) +
                                VMASK(double, si, sj - 1, k - 1) + VMASK(double, si, sj, k),
                            ZEPS) * m.e2t[ij];
    const double wm = WMASK(double, si, sj, k);
    const double zai = (gru(d, prd, si - 1, sj, k) + gru(d, prd, si, sj, k - 1) + gru(d, prd, si - 1, sj, k - 1) +
                        gru(d, prd, si, sj, k)) / zci * wm;
    const double zaj = (grv(d, prd, si, sj - 1, k) + grv(d, prd, si, sj, k - 1) + grv(d, prd, si, sj - 1, k - 1) +
                        grv(d, prd, si, sj, k)) / zcj * wm;
    const double e3w = E3(m.e3w, ij, k);
    const double zbi = min3(zbw, -100. * fabs(zai), -7.e+3 / e3w * fabs(zai));
    const double zbj = min3(zbw, -100. * fabs(zaj), -7.e+3 / e3w * fabs(zaj));
    const double zfk = max2(oml(nmln, ij, k), oml(nmln, ij, k - 1));
    const int ikt = d.mikt[ij];
    const double gwt = (double)A3(gdepw, si, sj, ikt);
    const double zck = ((double)A3(gdepw, si, sj, k) - gwt) / max2(hmlp[ij] - gwt, 10.);
    at32(zwz, p) = (zai / (zbi - ZEPS) * (1. - zfk) + zck * wiml[ij] * zfk) * wm;
    at32(zww, p) = (zaj / (zbj - ZEPS) * (1. - zfk) + zck * wjml[ij] * zfk) * wm;
}

// the 1-2-1 x 1-2-1 Shapiro sum of lines 234-251 / 318-337 (before its 1/16)
__device__ __forceinline__ double shapiro(const DevDomain &d, const double *z, int i, int j, int k)
{
    return A3(z, i - 1, j - 1, k) + A3(z, i + 1, j - 1, k) + A3(z, i - 1, j + 1, k) + A3(z, i + 1, j + 1, k) +
           2. * (A3(z, i, j - 1, k) + A3(z, i - 1, j, k) + A3(z, i + 1, j, k) + A3(z, i, j + 1, k)) +
           4. * A3(z, i, j, k);
}

// ---------------------------------------------------------------------------
// filter and coastal taper of the u/v slopes (lines 232-271) and the final
// lbc_lnk (line 361), every point and level: levels 1 and jpk only take the
// lbc of their interior values
// ---------------------------------------------------------------------------
template <typename T>
KERNEL void k_slp_uv_filt(DevDomain d, const double *__restrict__ zwz, const double *__restrict__ zww,
                          T *__restrict__ uslp, T *__restrict__ vslp)
{
    THREAD_IJ(1, 1);
    const int k = (int)blockIdx.z + 1;
    if (i > d.jpi || j > d.jpj) return;
    const uint32_t p = IJK32(i, j, k);
    int si, sj;
    const bool src = lbc_src(d, 'U', i, j, si, sj);
    const bool ring = !is_interior(d, i, j);
    if (k == 1 || k == d.jpk) {
        if (!ring) return;
        at32(uslp, p) = src ? A3(uslp, si, sj, k) : (T)0;
        at32(vslp, p) = src ? A3(vslp, si, sj, k) : (T)0;
        return;
    }
    if (!src) {
        at32(uslp, p) = (T)0;
        at32(vslp, p) = (T)0;
        return;
    }
    const double z1_16 = 1.0 / 16.;
    double u = z1_16 * shapiro(d, zwz, si, sj, k);
    double v = z1_16 * shapiro(d, zww, si, sj, k);
    u = u * (UMASK(double, si, sj + 1, k) + UMASK(double, si, sj - 1, k)) * 0.5 *
        (UMASK(double, si, sj, k) + UMASK(double, si, sj, k + 1)) * 0.5;
    v = v * (VMASK(double, si + 1, sj, k) + VMASK(double, si - 1, sj, k)) * 0.5 *
        (VMASK(double, si, sj, k) + VMASK(double, si, sj, k + 1)) * 0.5;
    at32(uslp, p) = (T)u;
    at32(vslp, p) = (T)v;
}

// filter and topographic taper of the w slopes (lines 316-356), final lbc_lnk
template <typename T>
KERNEL void k_slp_w_filt(DevDomain d, const double *__restrict__ zwz, const double *__restrict__ zww,
                         T *__restrict__ wslpi, T *__restrict__ wslpj)
{
    THREAD_IJ(1, 1);
    const int k = (int)blockIdx.z + 1;
    if (i > d.jpi || j > d.jpj) return;
    const uint32_t p = IJK32(i, j, k);
    int si, sj;
    const bool src = lbc_src(d, 'W', i, j, si, sj);
    const bool ring = !is_interior(d, i, j);
    if (k == 1 || k == d.jpk) {
        if (!ring) return;
        at32(wslpi, p) = src ? A3(wslpi, si, sj, k) : (T)0;
        at32(wslpj, p) = src ? A3(wslpj, si, sj, k) : (T)0;
        return;
    }
    if (!src) {
        at32(wslpi, p) = (T)0;
        at32(wslpj, p) = (T)0;
        return;
    }
    const double zcofw = WMASK(double, si, sj, k) * (1.0 / 16.);
    double wi = shapiro(d, zwz, si, sj, k) * zcofw;
    double wj = shapiro(d, zww, si, sj, k) * zcofw;
    const double zck = (UMASK(double, si, sj, k) + UMASK(double, si - 1, sj, k)) *
                       (VMASK(double, si, sj, k) + VMASK(double, si, sj - 1, k)) * 0.25;
    at32(wslpi, p) = (T)(wi * zck);
    at32(wslpj, p) = (T)(wj * zck);
}

}  // namespace

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
template <typename T>
int launch_zdf_mxl(const T *rn2b, const T *gdept, const T *gdepw, int nlb10, double rho_c, int32_t *nmln,
                   double *hmlp, double *hmlpt)
{
    const DevDomain &d = nh::dom();
    const double zN2_c = d.grav * rho_c * (1.0 / d.rau0);
    Grid2 g = grid2(d.jpi, d.jpj);
    LAUNCH("zdf_mxl", k_zdf_mxl<T>, g, d, rn2b, gdept, gdepw, nlb10, zN2_c, nmln, hmlp, hmlpt);
    return 0;
}

template <typename T>
int launch_ldf_slp(const T *prd, const T *pn2, const T *gdept, const T *gdepw, const int32_t *nmln,
                   const double *hmlp, const double *hmlpt, double rn_slpmax, T *uslp, T *vslp, T *wslpi, T *wslpj)
{
    const DevDomain &d = nh::dom();
    if (d.jpk < 3) return 0;
    const double z1_slpmax = 1. / rn_slpmax;
    const SlpIn<T> in{prd, pn2};
    double *zsu = nh::scratch2(NH_S2_SLP), *zsv = nh::scratch2(NH_S2_SLP + 1);
    double *wiml = nh::scratch2(NH_S2_SLP + 2), *wjml = nh::scratch2(NH_S2_SLP + 3);
    double *zwz = (double *)nh::scratch3(nh::S_ZWX), *zww = (double *)nh::scratch3(nh::S_ZWY);
    Grid2 g2 = grid2(d.jpi, d.jpj);
    LAUNCH("ldf_slp", k_slp_mxl<T>, g2, d, in, nmln, hmlpt, z1_slpmax, zsu, zsv, wiml, wjml);
    Grid2 gi = grid2(d.jpi, d.jpj);
    gi.grid.z = (unsigned)(d.jpk - 2);
    Grid2 ga = grid2(d.jpi, d.jpj);
    ga.grid.z = (unsigned)d.jpk;
    LAUNCH("ldf_slp", k_slp_uv_raw<T>, gi, d, in, gdept, nmln, zsu, zsv, z1_slpmax, zwz, zww);
    LAUNCH("ldf_slp", k_slp_uv_filt<T>, ga, d, (const double *)zwz, (const double *)zww, uslp, vslp);
    LAUNCH("ldf_slp", k_slp_w_raw<T>, gi, d, in, gdepw, nmln, hmlp, (const double *)wiml, (const double *)wjml, zwz,
           zww);
    LAUNCH("ldf_slp", k_slp_w_filt<T>, ga, d, (const double *)zwz, (const double *)zww, wslpi, wslpj);
    return 0;
}

#define INST(T)                                                                                                 \
    template int launch_zdf_mxl<T>(const T *, const T *, const T *, int, double, int32_t *, double *, double *); \
    template int launch_ldf_slp<T>(const T *, const T *, const T *, const T *, const int32_t *, const double *,   \
                                   const double *, double, T *, T *, T *, T *);
INST(double)
INST(float)
