// slp.hip -- the mixed-layer level (ZDF/zdfmxl.F90) and the iso-neutral
// slopes of the standard (Madec) operator (LDF/ldfslp.F90: ldf_slp and
// ldf_slp_mxl), for z-coordinates without partial steps or ice-shelf
// cavities (risfdep = 0, mikt = 1).
//
// Arithmetic is FP64 in both storage forms, rounded once on store, in the
// reference's operation order, so the FP64 form reproduces it bit for bit.
//
// The reference builds the unfiltered slopes in work arrays, applies
// lbc_lnk to them, filters and applies lbc_lnk again.  Here one kernel per
// level builds the unfiltered slopes of a tile and its halo in LDS (a halo
// point on the ring evaluates the interior formula at its lbc source, or is
// zero on a closed boundary, which is what lbc_lnk gives), filters them and
// writes the interior; the final lbc_lnk fills the output's ring.  The
// horizontal and vertical density gradients (zgru, zgrv, zdzr) are
// recomputed from prd and pn2 where used.
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
// I. unfiltered u/v slopes at the interior point (si, sj), level k, lines
// 197-219 (zwz, zww).  P(i, j, kk) and N(i, j, kk) give prd and pn2 at the
// levels used (prd: k, k-1; pn2: k, k+1), from global memory or from LDS.
// ---------------------------------------------------------------------------
template <class P, class N>
__device__ __forceinline__ double a_dzr(const DevDomain &d, const P &prd, const N &pn2, int i, int j, int k)
{
    const double zm1_g = -1.0 / d.grav;
    return zm1_g * (prd(i, j, k) + 1.) * (pn2(i, j, k) + pn2(i, j, k + 1)) * (1. - 0.5 * TMASK(double, i, j, k + 1));
}
template <class P>
__device__ __forceinline__ double a_gru(const DevDomain &d, const P &prd, int i, int j, int k)
{
    return UMASK(double, i, j, k) * (prd(i + 1, j, k) - prd(i, j, k));
}
template <class P>
__device__ __forceinline__ double a_grv(const DevDomain &d, const P &prd, int i, int j, int k)
{
    return VMASK(double, i, j, k) * (prd(i, j + 1, k) - prd(i, j, k));
}

template <typename T, class P, class N>
__device__ __forceinline__ void raw_uv(const DevDomain &d, const P &prd, const N &pn2, const T *gdept,
                                       const int32_t *nmln, const double *zsu, const double *zsv, double z1_slpmax,
                                       int si, int sj, int k, double &zwz, double &zww)
{
    const Metrics<double> &m = d.md;
    const uint32_t ij = IJ32(si, sj);
    const double zau = a_gru(d, prd, si, sj, k) * m.r1_e1u[ij];
    const double zav = a_grv(d, prd, si, sj, k) * m.r1_e2v[ij];
    const double dc = a_dzr(d, prd, pn2, si, sj, k);
    double zbu = 0.5 * (dc + a_dzr(d, prd, pn2, si + 1, sj, k));
    double zbv = 0.5 * (dc + a_dzr(d, prd, pn2, si, sj + 1, k));
    zbu = min3(zbu, -z1_slpmax * fabs(zau), -7.e+3 / E3(m.e3u, ij, k) * fabs(zau));
    zbv = min3(zbv, -z1_slpmax * fabs(zav), -7.e+3 / E3(m.e3v, ij, k) * fabs(zav));
    const double oc = oml(nmln, ij, k);
    const double zfi = max2(oc, oml(nmln, ij + 1, k));
    const double zfj = max2(oc, oml(nmln, ij + (uint32_t)d.jpi, k));
    const double gc = (double)A3(gdept, si, sj, k);
    const double zdepu = 0.5 * ((gc + (double)A3(gdept, si + 1, sj, k)) - E3(m.e3u, ij, d.miku[ij]));
    const double zdepv = 0.5 * ((gc + (double)A3(gdept, si, sj + 1, k)) - E3(m.e3v, ij, d.mikv[ij]));
    zwz = ((1. - zfi) * zau / (zbu - ZEPS) + zfi * zdepu * zsu[ij]) * UMASK(double, si, sj, k);
    zww = ((1. - zfj) * zav / (zbv - ZEPS) + zfj * zdepv * zsv[ij]) * VMASK(double, si, sj, k);
}

// II. unfiltered w slopes at the interior point (si, sj), level k, lines 277-304
template <typename T, class P, class N>
__device__ __forceinline__ void raw_w(const DevDomain &d, const P &prd, const N &pn2, const T *gdepw,
                                      const int32_t *nmln, const double *hmlp, const double *wiml,
                                      const double *wjml, int si, int sj, int k, double &zwz, double &zww)
{
    const Metrics<double> &m = d.md;
    const uint32_t ij = IJ32(si, sj);
    const double zm1_2g = -0.5 / d.grav;
    const double zbw = zm1_2g * pn2(si, sj, k) * (prd(si, sj, k) + prd(si, sj, k - 1) + 2.);
    const double zci = max2(UMASK(double, si - 1, sj, k) + UMASK(double, si, sj, k) + UMASK(double, si - 1, sj, k - 1) +
                                UMASK(double, si, sj, k - 1),
                            ZEPS) * m.e1t[ij];
    const double zcj = max2(VMASK(double, si, sj - 1, k) + VMASK(double, si, sj, k - 1) +
                                VMASK(double, si, sj - 1, k - 1) + VMASK(double, si, sj, k),
                            ZEPS) * m.e2t[ij];
    const double wm = WMASK(double, si, sj, k);
    const double zai = (a_gru(d, prd, si - 1, sj, k) + a_gru(d, prd, si, sj, k - 1) +
                        a_gru(d, prd, si - 1, sj, k - 1) + a_gru(d, prd, si, sj, k)) / zci * wm;
    const double zaj = (a_grv(d, prd, si, sj - 1, k) + a_grv(d, prd, si, sj, k - 1) +
                        a_grv(d, prd, si, sj - 1, k - 1) + a_grv(d, prd, si, sj, k)) / zcj * wm;
    const double e3w = E3(m.e3w, ij, k);
    const double zbi = min3(zbw, -100. * fabs(zai), -7.e+3 / e3w * fabs(zai));
    const double zbj = min3(zbw, -100. * fabs(zaj), -7.e+3 / e3w * fabs(zaj));
    const double zfk = max2(oml(nmln, ij, k), oml(nmln, ij, k - 1));
    const int ikt = d.mikt[ij];
    const double gwt = (double)A3(gdepw, si, sj, ikt);
    const double zck = ((double)A3(gdepw, si, sj, k) - gwt) / max2(hmlp[ij] - gwt, 10.);
    zwz = (zai / (zbi - ZEPS) * (1. - zfk) + zck * wiml[ij] * zfk) * wm;
    zww = (zaj / (zbj - ZEPS) * (1. - zfk) + zck * wjml[ij] * zfk) * wm;
}

// prd or pn2 read from global memory
template <typename T>
struct GlobalField {
    const T *p;
    const DevDomain &d;
    __device__ __forceinline__ double operator()(int i, int j, int k) const { return (double)A3(p, i, j, k); }
};

// ---------------------------------------------------------------------------
// one level of a SLP_TX x SLP_TY tile: the four unfiltered slopes of the
// tile and a one-point halo in LDS (a halo point on the domain's ring takes
// its lbc source's value, or zero on a closed side: the reference's
// lbc_lnk of zwz/zww, lines 231 and 315), then the Shapiro filters and
// tapers of lines 232-271 and 316-356 for the tile's interior points.  The
// output's ring is the final lbc_lnk (line 361), a separate pass.
// ---------------------------------------------------------------------------
#define SLP_TX 64
#define SLP_TY 4
template <typename T>
__global__ __launch_bounds__(SLP_TX * SLP_TY) void k_slp(DevDomain d, SlpIn<T> in, const T *__restrict__ gdept,
                                                          const T *__restrict__ gdepw,
                                                          const int32_t *__restrict__ nmln,
                                                          const double *__restrict__ zsu,
                                                          const double *__restrict__ zsv,
                                                          const double *__restrict__ hmlp,
                                                          const double *__restrict__ wiml,
                                                          const double *__restrict__ wjml, double z1_slpmax,
                                                          T *__restrict__ uslp, T *__restrict__ vslp,
                                                          T *__restrict__ wslpi, T *__restrict__ wslpj)
{
    constexpr int LX = SLP_TX + 2, LY = SLP_TY + 2;      // the slopes: tile + 1
    constexpr int SX = SLP_TX + 4, SY = SLP_TY + 4;      // prd, pn2: tile + 2
    __shared__ double su[LY][LX], sv[LY][LX], swi[LY][LX], swj[LY][LX];
    __shared__ double p0[SY][SX], pm[SY][SX], n0[SY][SX], np[SY][SX];
    const int k = (int)blockIdx.z + 2;
    const int ib = 2 + (int)blockIdx.x * SLP_TX, jb = 2 + (int)blockIdx.y * SLP_TY;
    const int tid = (int)(threadIdx.y * SLP_TX + threadIdx.x);
    // prd (k, k-1) and pn2 (k, k+1) of the tile and a two-point halo, as stored
    for (int q = tid; q < SX * SY; q += SLP_TX * SLP_TY) {
        const int li = q % SX, lj = q / SX;
        const int i = ib - 2 + li, j = jb - 2 + lj;
        double a = 0., b = 0., c = 0., e = 0.;
        if (i >= 1 && i <= d.jpi && j >= 1 && j <= d.jpj) {
            const uint32_t p = IJK32(i, j, k);
            a = (double)at32(in.prd, p);
            b = (double)at32(in.prd, p - (uint32_t)d.n2);
            c = (double)at32(in.pn2, p);
            e = (double)at32(in.pn2, p + (uint32_t)d.n2);
        }
        p0[lj][li] = a;
        pm[lj][li] = b;
        n0[lj][li] = c;
        np[lj][li] = e;
    }
    __syncthreads();
    const auto lprd = [&](int i, int j, int kk) {
        return kk == k ? p0[j - jb + 2][i - ib + 2] : pm[j - jb + 2][i - ib + 2];
    };
    const auto ln2 = [&](int i, int j, int kk) {
        return kk == k ? n0[j - jb + 2][i - ib + 2] : np[j - jb + 2][i - ib + 2];
    };
    const GlobalField<T> gprd{in.prd, d}, gn2{in.pn2, d};
    for (int q = tid; q < LX * LY; q += SLP_TX * SLP_TY) {
        const int li = q % LX, lj = q / LX;
        const int i = ib - 1 + li, j = jb - 1 + lj;
        double a = 0., b = 0., c = 0., e = 0.;
        int si, sj;
        if (i <= d.jpi && j <= d.jpj && lbc_src(d, 'U', i, j, si, sj)) {
            if (si == i && sj == j) {   // an interior point: its neighbours are staged
                raw_uv<T>(d, lprd, ln2, gdept, nmln, zsu, zsv, z1_slpmax, si, sj, k, a, b);
                raw_w<T>(d, lprd, ln2, gdepw, nmln, hmlp, wiml, wjml, si, sj, k, c, e);
            } else {                    // a ring point: its lbc source, from global memory
                raw_uv<T>(d, gprd, gn2, gdept, nmln, zsu, zsv, z1_slpmax, si, sj, k, a, b);
                raw_w<T>(d, gprd, gn2, gdepw, nmln, hmlp, wiml, wjml, si, sj, k, c, e);
            }
        }
        su[lj][li] = a;
        sv[lj][li] = b;
        swi[lj][li] = c;
        swj[lj][li] = e;
    }
    __syncthreads();
    const int li = (int)threadIdx.x + 1, lj = (int)threadIdx.y + 1;
    const int i = ib - 1 + li, j = jb - 1 + lj;
    if (i > d.jpi - 1 || j > d.jpj - 1) return;
    auto shapiro = [&](const double (*z)[LX]) {
        return z[lj - 1][li - 1] + z[lj - 1][li + 1] + z[lj + 1][li - 1] + z[lj + 1][li + 1] +
               2. * (z[lj - 1][li] + z[lj][li - 1] + z[lj][li + 1] + z[lj + 1][li]) + 4. * z[lj][li];
    };
    const double z1_16 = 1.0 / 16.;
    double u = z1_16 * shapiro(su);
    double v = z1_16 * shapiro(sv);
    u = u * (UMASK(double, i, j + 1, k) + UMASK(double, i, j - 1, k)) * 0.5 *
        (UMASK(double, i, j, k) + UMASK(double, i, j, k + 1)) * 0.5;
    v = v * (VMASK(double, i + 1, j, k) + VMASK(double, i - 1, j, k)) * 0.5 *
        (VMASK(double, i, j, k) + VMASK(double, i, j, k + 1)) * 0.5;
    const double zcofw = WMASK(double, i, j, k) * z1_16;
    const double zck = (UMASK(double, i, j, k) + UMASK(double, i - 1, j, k)) *
                       (VMASK(double, i, j, k) + VMASK(double, i, j - 1, k)) * 0.25;
    const uint32_t p = IJK32(i, j, k);
    at32(uslp, p) = (T)u;
    at32(vslp, p) = (T)v;
    at32(wslpi, p) = (T)(shapiro(swi) * zcofw * zck);
    at32(wslpj, p) = (T)(shapiro(swj) * zcofw * zck);
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
    Grid2 g2 = grid2(d.jpi, d.jpj);
    LAUNCH("ldf_slp", k_slp_mxl<T>, g2, d, in, nmln, hmlpt, z1_slpmax, zsu, zsv, wiml, wjml);
    Grid2 g;
    g.block = dim3(SLP_TX, SLP_TY, 1);
    g.grid = dim3((unsigned)((d.jpi - 2 + SLP_TX - 1) / SLP_TX), (unsigned)((d.jpj - 2 + SLP_TY - 1) / SLP_TY),
                  (unsigned)(d.jpk - 2));
    LAUNCH("ldf_slp", k_slp<T>, g, d, in, gdept, gdepw, nmln, (const double *)zsu, (const double *)zsv, hmlp,
           (const double *)wiml, (const double *)wjml, z1_slpmax, uslp, vslp, wslpi, wslpj);
    int st;
    if ((st = launch_lbc<T>(uslp, d.jpk, 'U', T(0))) || (st = launch_lbc<T>(vslp, d.jpk, 'V', T(0))) ||
        (st = launch_lbc<T>(wslpi, d.jpk, 'W', T(0))) || (st = launch_lbc<T>(wslpj, d.jpk, 'W', T(0))))
        return st;
    return 0;
}

#define INST(T)                                                                                                 \
    template int launch_zdf_mxl<T>(const T *, const T *, const T *, int, double, int32_t *, double *, double *); \
    template int launch_ldf_slp<T>(const T *, const T *, const T *, const T *, const int32_t *, const double *,   \
                                   const double *, double, T *, T *, T *, T *);
INST(double)
INST(float)
