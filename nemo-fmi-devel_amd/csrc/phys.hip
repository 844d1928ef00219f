// phys.hip -- density (the polynomial TEOS-10 / EOS-80 and the simplified
// equations of state of TRA/eosbn2.F90),
// the hydrostatic pressure gradient in z-coordinates (DYN/dynhpg.F90) and
// the iso-level laplacian viscosity (DYN/dynldf_lap_blp.F90).
//
// Arithmetic is FP64 in both storage forms, rounded once on store, in the
// reference's operation order, so the FP64 form reproduces it bit for bit.
#include "common.hpp"
#include "eos_poly.h"
#include "kernels.hpp"

// {TEOS-10, EOS-80} coefficient tables; a kernel reads one scheme with
// compile-time offsets, so the loads are wave-uniform scalar loads
__constant__ const nemo_eos_poly_t EOS_POLY[2] = NEMO_EOS_POLY_TABLES;

// one depth coefficient (zn0..zn3) of TRA/eosbn2.F90:246-266 / 600-637:
// total degree N in (zs, zt), nested over zt; each zt-power's salinity
// polynomial c0 + zs (c1 + zs (c2 + ...)) adds its constant term last
template <int N>
__device__ __forceinline__ double poly_st(const double (&c)[7][7], double zt, double zs)
{
    double acc = c[N][0];
#pragma unroll
    for (int j = N - 1; j >= 0; j--) {
        double ts = c[j][N - j];
#pragma unroll
        for (int i = N - j - 1; i >= 1; i--) ts = ts * zs + c[j][i];
        acc = acc * zt + ts * zs + c[j][0];
    }
    return acc;
}

// zn = ((zn3 zh + zn2) zh + zn1) zh + zn0 (TRA/eosbn2.F90:268), degrees
// (D0, D1, D2, D3) of the four depth coefficients
template <int D0, int D1, int D2, int D3>
__device__ __forceinline__ double poly_z(const double (&c)[4][7][7], double zt, double zs, double zh, double &zn0)
{
    const double z3 = poly_st<D3>(c[3], zt, zs), z2 = poly_st<D2>(c[2], zt, zs);
    const double z1 = poly_st<D1>(c[1], zt, zs);
    zn0 = poly_st<D0>(c[0], zt, zs);
    return ((z3 * zh + z2) * zh + z1) * zh + zn0;
}

// ---------------------------------------------------------------------------
// eos_insitu (POT = 0), TRA/eosbn2.F90:236-274, and eos_insitu_pot (POT = 1),
// lines 399-437: polynomial TEOS-10 (SCH = 0) / EOS-80 (SCH = 1).  One
// thread per cell, levels 1..jpkm1.
// ---------------------------------------------------------------------------
template <typename T, int POT, int SCH>
KERNEL void k_eos_poly(DevDomain d, const T *__restrict__ pts, T *__restrict__ prd, T *__restrict__ prhop,
                       const T *__restrict__ pdep)
{
    THREAD_IJ(1, 1);
    const int k = (int)blockIdx.z + 1;
    if (i > d.jpi || j > d.jpj) return;
    const nemo_eos_poly_t &P = EOS_POLY[SCH];
    const uint32_t p = IJK32(i, j, k);
    const double r1_rau0 = 1.0 / d.rau0;
    const double zh = (double)at32(pdep, p) * P.r1_Z0;
    const double zt = (double)at32(pts, p) * P.r1_T0;
    const double zs = sqrt(fabs((double)at32(pts + d.n3, p) + P.rdeltaS) * P.r1_S0);
    const double ztm = TMASK(double, i, j, k);
    double zn0;
    const double zn = poly_z<6, 4, 2, 1>(P.rho, zt, zs, zh, zn0);
    if (POT) at32(prhop, p) = (T)(zn0 * ztm);
    at32(prd, p) = (T)((zn * r1_rau0 - 1.) * ztm);
}

// rab_3d, TRA/eosbn2.F90:588-643: alpha and beta of the polynomial schemes
template <typename T, int SCH>
KERNEL void k_rab_poly(DevDomain d, const T *__restrict__ pts, T *__restrict__ pab, const T *__restrict__ gdept)
{
    THREAD_IJ(1, 1);
    const int k = (int)blockIdx.z + 1;
    if (i > d.jpi || j > d.jpj) return;
    const nemo_eos_poly_t &P = EOS_POLY[SCH];
    const uint32_t p = IJK32(i, j, k);
    const double r1_rau0 = 1.0 / d.rau0;
    const double zh = (double)at32(gdept, p) * P.r1_Z0;
    const double zt = (double)at32(pts, p) * P.r1_T0;
    const double zs = sqrt(fabs((double)at32(pts + d.n3, p) + P.rdeltaS) * P.r1_S0);
    const double ztm = TMASK(double, i, j, k);
    double z0;
    const double za = poly_z<5, 3, 1, 0>(P.alp, zt, zs, zh, z0);
    at32(pab, p) = (T)(za * r1_rau0 * ztm);
    const double zb = poly_z<5, 3, 1, 0>(P.bet, zt, zs, zh, z0);
    at32(pab + d.n3, p) = (T)(zb / zs * r1_rau0 * ztm);
}

// ---------------------------------------------------------------------------
// eos_insitu (POT = 0), TRA/eosbn2.F90:276-299, and eos_insitu_pot (POT = 1),
// lines 440-462, S-EOS.  One thread per cell, levels 1..jpkm1.
// ---------------------------------------------------------------------------
template <typename T, int POT>
KERNEL void k_eos(DevDomain d, SeosPar e, const T *__restrict__ pts, T *__restrict__ prd, T *__restrict__ prhop,
                  const T *__restrict__ pdep)
{
    THREAD_IJ(1, 1);
    const int k = (int)blockIdx.z + 1;
    if (i > d.jpi || j > d.jpj) return;
    const uint32_t p = IJK32(i, j, k);
    const double r1_rau0 = 1.0 / d.rau0;
    const double zt = (double)at32(pts, p) - 10.;
    const double zs = (double)at32(pts + d.n3, p) - 35.;
    const double zh = (double)at32(pdep, p);
    const double ztm = TMASK(double, i, j, k);
    if (POT) {
        double zn = -e.a0 * (1. + 0.5 * e.lambda1 * zt) * zt + e.b0 * (1. - 0.5 * e.lambda2 * zs) * zs -
                    e.nu * zt * zs;
        at32(prhop, p) = (T)((d.rau0 + zn) * ztm);
        zn = zn - (e.a0 * e.mu1 * zt + e.b0 * e.mu2 * zs) * zh;
        at32(prd, p) = (T)(zn * r1_rau0 * ztm);
    } else {
        const double zn = -e.a0 * (1. + 0.5 * e.lambda1 * zt + e.mu1 * zh) * zt +
                          e.b0 * (1. - 0.5 * e.lambda2 * zs - e.mu2 * zh) * zs - e.nu * zt * zs;
        at32(prd, p) = (T)(zn * r1_rau0 * ztm);
    }
}

// rab_3d, TRA/eosbn2.F90:645-662 (S-EOS): alpha and beta at t-points
template <typename T>
KERNEL void k_rab(DevDomain d, SeosPar e, const T *__restrict__ pts, T *__restrict__ pab,
                  const T *__restrict__ gdept)
{
    THREAD_IJ(1, 1);
    const int k = (int)blockIdx.z + 1;
    if (i > d.jpi || j > d.jpj) return;
    const uint32_t p = IJK32(i, j, k);
    const double r1_rau0 = 1.0 / d.rau0;
    const double zt = (double)at32(pts, p) - 10.;
    const double zs = (double)at32(pts + d.n3, p) - 35.;
    const double zh = (double)at32(gdept, p);
    const double ztm = TMASK(double, i, j, k);
    double zn = e.a0 * (1. + e.lambda1 * zt + e.mu1 * zh) + e.nu * zs;
    at32(pab, p) = (T)(zn * r1_rau0 * ztm);
    zn = e.b0 * (1. - e.lambda2 * zs - e.mu2 * zh) - e.nu * zt;
    at32(pab + d.n3, p) = (T)(zn * r1_rau0 * ztm);
}

// bn2, TRA/eosbn2.F90:914-925: N^2 at w-points 2..jpkm1 (grid.z = jpk - 2)
template <typename T>
KERNEL void k_bn2(DevDomain d, const T *__restrict__ pts, const T *__restrict__ pab, T *__restrict__ pn2,
                  const T *__restrict__ gdept, const T *__restrict__ gdepw)
{
    THREAD_IJ(1, 1);
    const int k = (int)blockIdx.z + 2;
    if (i > d.jpi || j > d.jpj) return;
    const uint32_t p = IJK32(i, j, k), q = IJK32(i, j, k - 1);
    const size_t n3 = d.n3;
    const double gt = (double)at32(gdept, p);
    const double zrw = ((double)at32(gdepw, p) - gt) / ((double)at32(gdept, q) - gt);
    const double zaw = (double)at32(pab, p) * (1. - zrw) + (double)at32(pab, q) * zrw;
    const double zbw = (double)at32(pab + n3, p) * (1. - zrw) + (double)at32(pab + n3, q) * zrw;
    at32(pn2, p) = (T)(d.grav *
                       (zaw * ((double)at32(pts, q) - (double)at32(pts, p)) -
                        zbw * ((double)at32(pts + n3, q) - (double)at32(pts + n3, p))) /
                       E3(d.md.e3w, IJ32(i, j), k) * WMASK(double, i, j, k));
}

// ---------------------------------------------------------------------------
// hpg_zco, DYN/dynhpg.F90:286-320: per interior column, the pressure
// gradient integrated from the surface in an FP64 register
// ---------------------------------------------------------------------------
template <typename T>
KERNEL void k_hpg_zco(DevDomain d, const T *__restrict__ rhd, T *__restrict__ ua, T *__restrict__ va)
{
    THREAD_IJ(2, 2);
    if (i > d.jpi - 1 || j > d.jpj - 1) return;
    const Metrics<double> &m = d.md;
    const uint32_t ij = IJ32(i, j), ie = ij + 1u, jn = ij + (uint32_t)d.jpi;
    const size_t plane = d.n2;
    const double zcoef0 = -d.grav * 0.5;
    const double r1u = m.r1_e1u[ij], r1v = m.r1_e2v[ij];
    double zhpi, zhpj;
    {
        const double zcoef1 = zcoef0 * E3(m.e3w, ij, 1);
        const double rc = (double)ldp(rhd, ij);
        zhpi = zcoef1 * ((double)ldp(rhd, ie) - rc) * r1u;
        zhpj = zcoef1 * ((double)ldp(rhd, jn) - rc) * r1v;
        stp(ua, ij, (T)((double)ldp((const T *)ua, ij) + zhpi));
        stp(va, ij, (T)((double)ldp((const T *)va, ij) + zhpj));
    }
    double rc_m = (double)ldp(rhd, ij), re_m = (double)ldp(rhd, ie), rn_m = (double)ldp(rhd, jn);
#pragma unroll 4
    for (int k = 2; k <= d.jpk - 1; k++) {
        const T *r = plane_ptr(rhd, plane, k);
        const double rc = (double)ldp(r, ij), re = (double)ldp(r, ie), rn = (double)ldp(r, jn);
        const double zcoef1 = zcoef0 * E3(m.e3w, ij, k);
        zhpi = zhpi + zcoef1 * ((re + re_m) - (rc + rc_m)) * r1u;
        zhpj = zhpj + zcoef1 * ((rn + rn_m) - (rc + rc_m)) * r1v;
        T *uk = plane_ptr(ua, plane, k), *vk = plane_ptr(va, plane, k);
        stp(uk, ij, (T)((double)ldp((const T *)uk, ij) + zhpi));
        stp(vk, ij, (T)((double)ldp((const T *)vk, ij) + zhpj));
        rc_m = rc;
        re_m = re;
        rn_m = rn;
    }
}

// ---------------------------------------------------------------------------
// dyn_ldf_lap, DYN/dynldf_lap_blp.F90:72-100: the f-point curls and t-point
// divergences the trends use, each evaluated as the reference evaluates
// zcur / zdiv
// ---------------------------------------------------------------------------
template <typename T>
__device__ __forceinline__ double ldf_cur(const DevDomain &d, const T *pub, const T *pvb, const T *ahmf, int p, int q,
                                          int k)
{
    const Metrics<double> &m = d.md;
    return (double)A3(ahmf, p, q, k) * E3(m.e3f, IJ32(p, q), k) * A2(m.r1_e1e2f, p, q) *
           (A2(m.e2v, p + 1, q) * (double)A3(pvb, p + 1, q, k) - A2(m.e2v, p, q) * (double)A3(pvb, p, q, k) -
            A2(m.e1u, p, q + 1) * (double)A3(pub, p, q + 1, k) + A2(m.e1u, p, q) * (double)A3(pub, p, q, k));
}
template <typename T>
__device__ __forceinline__ double ldf_div(const DevDomain &d, const T *pub, const T *pvb, const T *ahmt, int p, int q,
                                          int k)
{
    const Metrics<double> &m = d.md;
    return (double)A3(ahmt, p, q, k) * A2(m.r1_e1e2t, p, q) / E3(m.e3t, IJ32(p, q), k) *
           (A2(m.e2u, p, q) * E3(m.e3u, IJ32(p, q), k) * (double)A3(pub, p, q, k) -
            A2(m.e2u, p - 1, q) * E3(m.e3u, IJ32(p - 1, q), k) * (double)A3(pub, p - 1, q, k) +
            A2(m.e1v, p, q) * E3(m.e3v, IJ32(p, q), k) * (double)A3(pvb, p, q, k) -
            A2(m.e1v, p, q - 1) * E3(m.e3v, IJ32(p, q - 1), k) * (double)A3(pvb, p, q - 1, k));
}

// one level of a LDF_TX x LDF_TY tile: the curls (f-points) and divergences
// (t-points) the tile's trends use are evaluated once each into LDS, then
// every interior cell forms its two trends from them
#define LDF_TX 64
#define LDF_TY 4
template <typename T>
__global__ __launch_bounds__(LDF_TX * LDF_TY) void k_dyn_ldf_lap(DevDomain d, double zsign, const T *__restrict__ pub,
                                                                  const T *__restrict__ pvb, T *__restrict__ pua,
                                                                  T *__restrict__ pva, const T *__restrict__ ahmt,
                                                                  const T *__restrict__ ahmf)
{
    constexpr int LX = LDF_TX + 1, LY = LDF_TY + 1;
    __shared__ double scur[LY][LX], sdiv[LY][LX];
    const int k = (int)blockIdx.z + 1;
    const int ib = 2 + (int)blockIdx.x * LDF_TX, jb = 2 + (int)blockIdx.y * LDF_TY;
    const int tid = (int)(threadIdx.y * LDF_TX + threadIdx.x);
    for (int q = tid; q < LX * LY; q += LDF_TX * LDF_TY) {
        const int li = q % LX, lj = q / LX;
        const int pc = ib - 1 + li, qc = jb - 1 + lj;   // f-point (i-1 .. i+TX-1, j-1 .. j+TY-1)
        const int pd = ib + li, qd = jb + lj;           // t-point (i .. i+TX, j .. j+TY)
        scur[lj][li] = (pc <= d.jpi - 1 && qc <= d.jpj - 1) ? ldf_cur(d, pub, pvb, ahmf, pc, qc, k) : 0.;
        sdiv[lj][li] = (pd <= d.jpi && qd <= d.jpj) ? ldf_div(d, pub, pvb, ahmt, pd, qd, k) : 0.;
    }
    __syncthreads();
    const int li = (int)threadIdx.x, lj = (int)threadIdx.y;
    const int i = ib + li, j = jb + lj;
    if (i > d.jpi - 1 || j > d.jpj - 1) return;
    const Metrics<double> &m = d.md;
    const double cur = scur[lj + 1][li + 1], cur_s = scur[lj][li + 1], cur_w = scur[lj + 1][li];
    const double div = sdiv[lj][li], div_e = sdiv[lj][li + 1], div_n = sdiv[lj + 1][li];
    const uint32_t ij = IJ32(i, j);
    A3(pua, i, j, k) = (T)((double)A3(pua, i, j, k) +
                           zsign * (-((cur - cur_s) * m.r1_e2u[ij] / E3(m.e3u, ij, k)) + (div_e - div) * m.r1_e1u[ij]));
    A3(pva, i, j, k) = (T)((double)A3(pva, i, j, k) +
                           zsign * ((cur - cur_w) * m.r1_e1v[ij] / E3(m.e3v, ij, k) + (div_n - div) * m.r1_e2v[ij]));
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
template <typename T>
int launch_eos(int neos, const SeosPar &e, int pot, const T *pts, T *prd, T *prhop, const T *pdep)
{
    const DevDomain &d = nh::dom();
    Grid2 g = grid2(d.jpi, d.jpj);
    g.grid.z = (unsigned)(d.jpk - 1);
    if (neos == NEMO_NP_TEOS10) {
        if (pot) LAUNCH("eos", (k_eos_poly<T, 1, NEMO_EOS_POLY_TEOS10>), g, d, pts, prd, prhop, pdep);
        else LAUNCH("eos", (k_eos_poly<T, 0, NEMO_EOS_POLY_TEOS10>), g, d, pts, prd, prhop, pdep);
    } else if (neos == NEMO_NP_EOS80) {
        if (pot) LAUNCH("eos", (k_eos_poly<T, 1, NEMO_EOS_POLY_EOS80>), g, d, pts, prd, prhop, pdep);
        else LAUNCH("eos", (k_eos_poly<T, 0, NEMO_EOS_POLY_EOS80>), g, d, pts, prd, prhop, pdep);
    } else {
        if (pot) LAUNCH("eos", (k_eos<T, 1>), g, d, e, pts, prd, prhop, pdep);
        else LAUNCH("eos", (k_eos<T, 0>), g, d, e, pts, prd, prhop, pdep);
    }
    return 0;
}

template <typename T>
int launch_rab(int neos, const SeosPar &e, const T *pts, T *pab, const T *gdept)
{
    const DevDomain &d = nh::dom();
    Grid2 g = grid2(d.jpi, d.jpj);
    g.grid.z = (unsigned)(d.jpk - 1);
    if (neos == NEMO_NP_TEOS10) LAUNCH("rab", (k_rab_poly<T, NEMO_EOS_POLY_TEOS10>), g, d, pts, pab, gdept);
    else if (neos == NEMO_NP_EOS80) LAUNCH("rab", (k_rab_poly<T, NEMO_EOS_POLY_EOS80>), g, d, pts, pab, gdept);
    else LAUNCH("rab", k_rab<T>, g, d, e, pts, pab, gdept);
    return 0;
}

template <typename T>
int launch_bn2(const T *pts, const T *pab, T *pn2, const T *gdept, const T *gdepw)
{
    const DevDomain &d = nh::dom();
    if (d.jpk < 3) return 0;
    Grid2 g = grid2(d.jpi, d.jpj);
    g.grid.z = (unsigned)(d.jpk - 2);
    LAUNCH("bn2", k_bn2<T>, g, d, pts, pab, pn2, gdept, gdepw);
    return 0;
}

template <typename T>
int launch_hpg_zco(const T *rhd, T *ua, T *va)
{
    const DevDomain &d = nh::dom();
    Grid2 g = grid2(d.jpi - 2, d.jpj - 2);
    LAUNCH("dyn_hpg", k_hpg_zco<T>, g, d, rhd, ua, va);
    return 0;
}

template <typename T>
int launch_dyn_ldf_lap(double zsign, const T *pub, const T *pvb, T *pua, T *pva, const T *ahmt, const T *ahmf)
{
    const DevDomain &d = nh::dom();
    Grid2 g;
    g.block = dim3(LDF_TX, LDF_TY, 1);
    g.grid = dim3((unsigned)((d.jpi - 2 + LDF_TX - 1) / LDF_TX), (unsigned)((d.jpj - 2 + LDF_TY - 1) / LDF_TY),
                  (unsigned)(d.jpk - 1));
    LAUNCH("dyn_ldf", k_dyn_ldf_lap<T>, g, d, zsign, pub, pvb, pua, pva, ahmt, ahmf);
    return 0;
}

#define INST(T)                                                                                      \
    template int launch_eos<T>(int, const SeosPar &, int, const T *, T *, T *, const T *);            \
    template int launch_rab<T>(int, const SeosPar &, const T *, T *, const T *);                      \
    template int launch_bn2<T>(const T *, const T *, T *, const T *, const T *);                      \
    template int launch_hpg_zco<T>(const T *, T *, T *);                                              \
    template int launch_dyn_ldf_lap<T>(double, const T *, const T *, T *, T *, const T *, const T *);
INST(double)
INST(float)
#undef INST
