// spg.hip -- split-explicit free surface, DYN/dynspg_ts.F90:121-1310, on gfx950.
//
// Configuration implemented: ln_linssh = T, ln_bt_fw = T, no wetting/drying,
// tides, open boundaries, AGRIF, atmospheric pressure or Stokes drift
// (every branch those switches select is absent from the path).
//
// Phase 1 (3D -> 2D) is one column kernel (vertical sums of the trends,
// removal of their mean, bottom baroclinic velocity) plus one 2D kernel (the
// barotropic Coriolis / drag / wind forcing).  Phase 2 runs the `icycle`
// dependent sub-steps as two 2D kernels each:
//   S2 (k_spg_s2)  the after sea level (flux divergence, lbc fill), its ssha
//                  sum and the AB3-AM4 interpolated level zsshp2_e;
//   S3 (k_spg_s3)  the after velocities (Coriolis, drag, pressure gradient,
//                  lbc fill) and the ua_b sum, then the next sub-step's
//                  extrapolated velocities, fluxes and un_adv sum, pointwise.
// Only the first extrapolation (k_spg_s1) has a kernel of its own.  (One
// kernel per sub-step that recomputes zsshp2_e at the east and north
// neighbours moves as many bytes and runs slower; one that shares it through
// LDS, S2 and S3 separated by a workgroup barrier, runs no faster either:
// 2.77 ms per step at 1024x1024 against 2.78 for the two kernels.)  The sub-step levels rotate
// through buffers instead of being copied: three for the velocities, four for
// the sea level.  All barotropic state and arithmetic is FP64 in both storage
// forms.  Phase 3 is one 2D kernel and one column kernel.
#include "common.hpp"
#include "kernels.hpp"

namespace {
struct SpgWork {
    double *zu_frc, *zv_frc, *zwxb, *zwyb, *zCdU_u, *zCdU_v, *zssh_frc, *zwz, *zhf, *ftne, *ftnw, *ftse, *ftsw;
    double *ua_e, *va_e, *zwx, *zwy, *zsshp2_e;
};
// the next sub-step's extrapolation, folded into the velocity kernel: its
// coefficients and weight, the velocity level two steps back, and the
// second set of ua_e/va_e/zwx/zwy buffers it writes (the current set is still
// read by neighbouring points of this sub-step)
struct SpgNext {
    int on;
    double za1, za2, za3, wgt2;
    const double *ub_before, *vb_before;
    double *ua_e, *va_e, *zwx, *zwy, *un_adv, *vn_adv;
};
}

// barotropic Coriolis trend at u/v points, phase 1 (lines 390-443) and
// phase 2 (lines 1000-1056); zwx/zwy are the fluxes, ue/ve the velocities
// used by the t-point (ENT) scheme, h the t-point depth
__device__ __forceinline__ void bt_coriolis(const DevDomain &d, int i, int j, const double *zwx, const double *zwy,
                                            const double *zwz, const double *ftne, const double *ftnw,
                                            const double *ftse, const double *ftsw, const double *ue,
                                            const double *ve, const double *ht, double z1_hu, double z1_hv,
                                            double &zu_trd, double &zv_trd)
{
    const Metrics<double> &m = d.md;
    const double r1_4 = 0.25, r1_8 = 0.125, r1_12 = 1.0 / 12.0;
    switch (d.nvor_scheme) {
    case NEMO_NP_ENT: {
        zu_trd = r1_4 * m.r1_e1e2u[IJ(i, j)] * z1_hu *
                 (m.e1e2t[IJ(i + 1, j)] * A2(ht, i + 1, j) * m.ff_t[IJ(i + 1, j)] * (A2(ve, i + 1, j) + A2(ve, i + 1, j - 1)) +
                  m.e1e2t[IJ(i, j)] * A2(ht, i, j) * m.ff_t[IJ(i, j)] * (A2(ve, i, j) + A2(ve, i, j - 1)));
        zv_trd = -(r1_4 * m.r1_e1e2v[IJ(i, j)] * z1_hv *
                   (m.e1e2t[IJ(i, j + 1)] * A2(ht, i, j + 1) * m.ff_t[IJ(i, j + 1)] * (A2(ue, i, j + 1) + A2(ue, i - 1, j + 1)) +
                    m.e1e2t[IJ(i, j)] * A2(ht, i, j) * m.ff_t[IJ(i, j)] * (A2(ue, i, j) + A2(ue, i - 1, j))));
        break;
    }
    case NEMO_NP_ENE:
    case NEMO_NP_MIX: {
        const double zy1 = (A2(zwy, i, j - 1) + A2(zwy, i + 1, j - 1)) * m.r1_e1u[IJ(i, j)];
        const double zy2 = (A2(zwy, i, j) + A2(zwy, i + 1, j)) * m.r1_e1u[IJ(i, j)];
        const double zx1 = (A2(zwx, i - 1, j) + A2(zwx, i - 1, j + 1)) * m.r1_e2v[IJ(i, j)];
        const double zx2 = (A2(zwx, i, j) + A2(zwx, i, j + 1)) * m.r1_e2v[IJ(i, j)];
        zu_trd = r1_4 * (A2(zwz, i, j - 1) * zy1 + A2(zwz, i, j) * zy2);
        zv_trd = -(r1_4 * (A2(zwz, i - 1, j) * zx1 + A2(zwz, i, j) * zx2));
        break;
    }
    case NEMO_NP_ENS: {
        const double zy1 = r1_8 * (A2(zwy, i, j - 1) + A2(zwy, i + 1, j - 1) + A2(zwy, i, j) + A2(zwy, i + 1, j)) *
                           m.r1_e1u[IJ(i, j)];
        const double zx1 = -(r1_8 * (A2(zwx, i - 1, j) + A2(zwx, i - 1, j + 1) + A2(zwx, i, j) + A2(zwx, i, j + 1)) *
                             m.r1_e2v[IJ(i, j)]);
        zu_trd = zy1 * (A2(zwz, i, j - 1) + A2(zwz, i, j));
        zv_trd = zx1 * (A2(zwz, i - 1, j) + A2(zwz, i, j));
        break;
    }
    default: {   // EET, EEN
        zu_trd = r1_12 * m.r1_e1u[IJ(i, j)] *
                 (A2(ftne, i, j) * A2(zwy, i, j) + A2(ftnw, i + 1, j) * A2(zwy, i + 1, j) +
                  A2(ftse, i, j) * A2(zwy, i, j - 1) + A2(ftsw, i + 1, j) * A2(zwy, i + 1, j - 1));
        zv_trd = -(r1_12 * m.r1_e2v[IJ(i, j)] *
                   (A2(ftsw, i, j + 1) * A2(zwx, i - 1, j + 1) + A2(ftse, i, j + 1) * A2(zwx, i, j + 1) +
                    A2(ftnw, i, j) * A2(zwx, i - 1, j) + A2(ftne, i, j) * A2(zwx, i, j)));
        break;
    }
    }
}

// ---------------------------------------------------------------------------
// Coriolis pre-computations, lines 241-342 (recomputed every call: with a
// linear free surface they are constant, so this equals the kt == nit000 set)
// ---------------------------------------------------------------------------
struct S_spg_cor1 {
    static constexpr int NTHR = NH_BLOCK, WAVES = 1;
    __device__ __forceinline__ static void run(const VB &vb_, const DevDomain &d, const double * __restrict__ ht_n, SpgWork w)
    {
        THREAD_IJ_V(vb_, 1, 1);
        if (i > d.jpi || j > d.jpj) return;
        const Metrics<double> &m = d.md;
        const size_t ij = IJ(i, j);
        switch (d.nvor_scheme) {
        case NEMO_NP_EEN:
            if (i <= d.jpi - 1 && j <= d.jpj - 1) {
                double z;
                const double s = A2(ht_n, i, j + 1) + A2(ht_n, i + 1, j + 1) + A2(ht_n, i, j) + A2(ht_n, i + 1, j);
                if (d.nn_een_e3f == 0) {
                    z = s * 0.25;
                } else {
                    z = s / fmax2(1.0, A2(d.ssmask, i, j + 1) + A2(d.ssmask, i + 1, j + 1) + A2(d.ssmask, i, j) +
                                           A2(d.ssmask, i + 1, j));
                }
                if (z != 0.0) z = m.ff_f[ij] / z;
                w.zhf[ij] = z;      // zwz before lbc_lnk
            }
            break;
        case NEMO_NP_ENE:
        case NEMO_NP_ENS:
        case NEMO_NP_MIX:
            if (j <= d.jpj - 1) {
                // zhf = 0 * (1 - umask*umask) + sum_k e3f * umask(j) * umask(j+1)
                double z = 0.0;
                z = z * (1.0 - UMASK(double, i, j, 1) * UMASK(double, i, j + 1, 1));
                for (int k = 1; k <= d.jpk - 1; k++)
                    z = z + E3(m.e3f, ij, k) * UMASK(double, i, j, k) * UMASK(double, i, j + 1, k);
                w.zhf[ij] = z;
            }
            break;
        default: break;
        }
    }
};
KERNEL void k_spg_cor1(DevDomain d, const double * __restrict__ ht_n, SpgWork w)
{
    S_spg_cor1::run(vb_hw(), d, ht_n, w);
}

struct S_spg_cor2 {
    static constexpr int NTHR = NH_BLOCK, WAVES = 1;
    __device__ __forceinline__ static void run(const VB &vb_, const DevDomain &d, const double * __restrict__ ht_n, SpgWork w)
    {
        THREAD_IJ_V(vb_, 1, 1);
        if (i > d.jpi || j > d.jpj) return;
        const Metrics<double> &m = d.md;
        const size_t ij = IJ(i, j);
        int si, sj;
        switch (d.nvor_scheme) {
        case NEMO_NP_ENE:
        case NEMO_NP_ENS:
        case NEMO_NP_MIX: {
            double zhf = lbc_src(d, 'F', i, j, si, sj) ? A2(w.zhf, si, sj) : 0.0;
            double z = 0.0;
            if (zhf != 0.0) z = 1.0 / zhf;
            w.zwz[ij] = m.ff_f[ij] * z;
            break;
        }
        case NEMO_NP_EEN:
            w.zwz[ij] = lbc_src(d, 'F', i, j, si, sj) ? A2(w.zhf, si, sj) : 0.0;
            break;
        case NEMO_NP_EET:
            if (i >= 2 && j >= 2) {
                const double ss = d.ssmask[ij];
                const double z1_ht = ss / (A2(ht_n, i, j) + 1.0 - ss);
                const double *f = m.ff_f;
                A2(w.ftne, i, j) = (A2(f, i - 1, j) + A2(f, i, j) + A2(f, i, j - 1)) * z1_ht;
                A2(w.ftnw, i, j) = (A2(f, i - 1, j - 1) + A2(f, i - 1, j) + A2(f, i, j)) * z1_ht;
                A2(w.ftse, i, j) = (A2(f, i, j) + A2(f, i, j - 1) + A2(f, i - 1, j - 1)) * z1_ht;
                A2(w.ftsw, i, j) = (A2(f, i, j - 1) + A2(f, i - 1, j - 1) + A2(f, i - 1, j)) * z1_ht;
            } else if (i == 1) {
                A2(w.ftne, i, j) = A2(w.ftnw, i, j) = A2(w.ftse, i, j) = A2(w.ftsw, i, j) = 0.0;
            }
            break;
        default: break;
        }
    }
};
KERNEL void k_spg_cor2(DevDomain d, const double * __restrict__ ht_n, SpgWork w)
{
    S_spg_cor2::run(vb_hw(), d, ht_n, w);
}

struct S_spg_cor3 {
    static constexpr int NTHR = NH_BLOCK, WAVES = 1;
    __device__ __forceinline__ static void run(const VB &vb_, const DevDomain &d, SpgWork w)
    {
        // EEN triads from the lbc-filled zwz, lines 267-275
        THREAD_IJ_V(vb_, 1, 1);
        if (i > d.jpi || j > d.jpj) return;
        if (i == 1) {
            A2(w.ftne, i, j) = A2(w.ftnw, i, j) = A2(w.ftse, i, j) = A2(w.ftsw, i, j) = 0.0;
            return;
        }
        if (j < 2) return;
        const double *z = w.zwz;
        A2(w.ftne, i, j) = A2(z, i - 1, j) + A2(z, i, j) + A2(z, i, j - 1);
        A2(w.ftnw, i, j) = A2(z, i - 1, j - 1) + A2(z, i - 1, j) + A2(z, i, j);
        A2(w.ftse, i, j) = A2(z, i, j) + A2(z, i, j - 1) + A2(z, i - 1, j - 1);
        A2(w.ftsw, i, j) = A2(z, i, j - 1) + A2(z, i - 1, j - 1) + A2(z, i - 1, j);
    }
};
KERNEL void k_spg_cor3(DevDomain d, SpgWork w)
{
    S_spg_cor3::run(vb_hw(), d, w);
}

// ---------------------------------------------------------------------------
// Phase 1a: vertical sums of the trends and their removal (lines 358-378);
// bottom baroclinic velocities (lines 519-527)
// ---------------------------------------------------------------------------
template <typename T>
struct S_spg_p1a {
    static constexpr int NTHR = NH_BLOCK, WAVES = 1;
    __device__ __forceinline__ static void run(const VB &vb_, const DevDomain &d, T * __restrict__ ua, T * __restrict__ va, const T * __restrict__ un, const T * __restrict__ vn, const double * __restrict__ r1_hu_n, const double * __restrict__ r1_hv_n, const double * __restrict__ un_b, const double * __restrict__ vn_b, SpgWork w)
    {
        THREAD_IJ_V(vb_, 1, 1);
        if (i > d.jpi || j > d.jpj) return;
        const Metrics<T> &m = met<T>(d);
        const size_t ij = IJ(i, j);
        const int jpk = d.jpk;
        double su = 0.0, sv = 0.0;
        for (int k = 1; k <= jpk - 1; k++) {
            const size_t p = IJK(i, j, k);
            su = su + (double)E3(m.e3u, ij, k) * (double)ua[p] * UMASK(double, i, j, k);
            sv = sv + (double)E3(m.e3v, ij, k) * (double)va[p] * VMASK(double, i, j, k);
        }
        su = su * r1_hu_n[ij];
        sv = sv * r1_hv_n[ij];
        w.zu_frc[ij] = su;
        w.zv_frc[ij] = sv;
        if (!is_interior(d, i, j)) return;
        for (int k = 1; k <= jpk - 1; k++) {
            const size_t p = IJK(i, j, k);
            ua[p] = (T)((double)ua[p] - su * UMASK(double, i, j, k));
            va[p] = (T)((double)va[p] - sv * VMASK(double, i, j, k));
        }
        w.zwxb[ij] = (double)A3(un, i, j, d.mbku[ij]) - un_b[ij];
        w.zwyb[ij] = (double)A3(vn, i, j, d.mbkv[ij]) - vn_b[ij];
    }
};
template <typename T>
KERNEL void k_spg_p1a(DevDomain d, T * __restrict__ ua, T * __restrict__ va, const T * __restrict__ un, const T * __restrict__ vn, const double * __restrict__ r1_hu_n,
                          const double * __restrict__ r1_hv_n, const double * __restrict__ un_b, const double * __restrict__ vn_b, SpgWork w)
{
    S_spg_p1a<T>::run(vb_hw(), d, ua, va, un, vn, r1_hu_n, r1_hv_n, un_b, vn_b, w);
}

// Phase 1b: barotropic forcing (lines 386-639)
struct S_spg_p1b {
    static constexpr int NTHR = NH_BLOCK, WAVES = 1;
    __device__ __forceinline__ static void run(const VB &vb_, const DevDomain &d, const double * __restrict__ un_b, const double * __restrict__ vn_b, const double * __restrict__ hu_n, const double * __restrict__ hv_n, const double * __restrict__ r1_hu_n, const double * __restrict__ r1_hv_n, const double * __restrict__ ht_n, const double * __restrict__ rCdU_bot, const double * __restrict__ utau, const double * __restrict__ vtau, const double * __restrict__ emp, const double * __restrict__ rnf, SpgWork w)
    {
        THREAD_IJ_V(vb_, 1, 1);
        if (i > d.jpi || j > d.jpj) return;
        const Metrics<double> &m = d.md;
        const size_t ij = IJ(i, j);
        const double r1_rau0 = 1.0 / d.rau0, r1_2 = 0.5;
        // zssh_frc = r1_rau0 * ( emp - rnf + fwfisf ), fwfisf = 0 without cavities
        w.zssh_frc[ij] = r1_rau0 * (emp[ij] - (rnf ? rnf[ij] : 0.0) + 0.0);
        if (!is_interior(d, i, j)) return;
        w.zCdU_u[ij] = r1_2 * (A2(rCdU_bot, i + 1, j) + A2(rCdU_bot, i, j));
        w.zCdU_v[ij] = r1_2 * (A2(rCdU_bot, i, j + 1) + A2(rCdU_bot, i, j));
        // now fluxes zwx = un_b * hu_n * e2u at the 6 points the scheme reads
        double zwx_[2][3], zwy_[3][2];   // zwx(i-1..i, j..j+1) / zwy(i..i+1, j-1..j)
        for (int a = 0; a < 2; a++)
            for (int b = 0; b < 2; b++) {
                const int p = i - 1 + a, q = j + b;
                zwx_[a][b] = A2(un_b, p, q) * A2(hu_n, p, q) * m.e2u[IJ(p, q)];
                const int p2 = i + a, q2 = j - 1 + b;
                zwy_[a][b] = A2(vn_b, p2, q2) * A2(hv_n, p2, q2) * m.e1v[IJ(p2, q2)];
            }
        double zu_trd, zv_trd;
        {
            const double *zwz = w.zwz;
            switch (d.nvor_scheme) {
            case NEMO_NP_ENT: {
                const double r1_4 = 0.25;
                zu_trd = r1_4 * m.r1_e1e2u[ij] * r1_hu_n[ij] *
                         (m.e1e2t[IJ(i + 1, j)] * A2(ht_n, i + 1, j) * m.ff_t[IJ(i + 1, j)] * (A2(vn_b, i + 1, j) + A2(vn_b, i + 1, j - 1)) +
                          m.e1e2t[ij] * A2(ht_n, i, j) * m.ff_t[ij] * (A2(vn_b, i, j) + A2(vn_b, i, j - 1)));
                zv_trd = -(r1_4 * m.r1_e1e2v[ij] * r1_hv_n[ij] *
                           (m.e1e2t[IJ(i, j + 1)] * A2(ht_n, i, j + 1) * m.ff_t[IJ(i, j + 1)] * (A2(un_b, i, j + 1) + A2(un_b, i - 1, j + 1)) +
                            m.e1e2t[ij] * A2(ht_n, i, j) * m.ff_t[ij] * (A2(un_b, i, j) + A2(un_b, i - 1, j))));
                break;
            }
            case NEMO_NP_ENE:
            case NEMO_NP_MIX: {
                const double zy1 = (zwy_[0][0] + zwy_[1][0]) * m.r1_e1u[ij];
                const double zy2 = (zwy_[0][1] + zwy_[1][1]) * m.r1_e1u[ij];
                const double zx1 = (zwx_[0][0] + zwx_[0][1]) * m.r1_e2v[ij];
                const double zx2 = (zwx_[1][0] + zwx_[1][1]) * m.r1_e2v[ij];
                zu_trd = 0.25 * (A2(zwz, i, j - 1) * zy1 + A2(zwz, i, j) * zy2);
                zv_trd = -(0.25 * (A2(zwz, i - 1, j) * zx1 + A2(zwz, i, j) * zx2));
                break;
            }
            case NEMO_NP_ENS: {
                const double zy1 = 0.125 * (zwy_[0][0] + zwy_[1][0] + zwy_[0][1] + zwy_[1][1]) * m.r1_e1u[ij];
                const double zx1 = -(0.125 * (zwx_[0][0] + zwx_[0][1] + zwx_[1][0] + zwx_[1][1]) * m.r1_e2v[ij]);
                zu_trd = zy1 * (A2(zwz, i, j - 1) + A2(zwz, i, j));
                zv_trd = zx1 * (A2(zwz, i - 1, j) + A2(zwz, i, j));
                break;
            }
            default: {
                const double r1_12 = 1.0 / 12.0;
                zu_trd = r1_12 * m.r1_e1u[ij] *
                         (A2(w.ftne, i, j) * zwy_[0][1] + A2(w.ftnw, i + 1, j) * zwy_[1][1] + A2(w.ftse, i, j) * zwy_[0][0] +
                          A2(w.ftsw, i + 1, j) * zwy_[1][0]);
                zv_trd = -(r1_12 * m.r1_e2v[ij] *
                           (A2(w.ftsw, i, j + 1) * zwx_[0][1] + A2(w.ftse, i, j + 1) * zwx_[1][1] +
                            A2(w.ftnw, i, j) * zwx_[0][0] + A2(w.ftne, i, j) * zwx_[1][0]));
                break;
            }
            }
        }
        double zu = w.zu_frc[ij], zv = w.zv_frc[ij];
        zu = zu - zu_trd * d.ssumask[ij];
        zv = zv - zv_trd * d.ssvmask[ij];
        zu = zu + r1_hu_n[ij] * r1_2 * (A2(rCdU_bot, i + 1, j) + A2(rCdU_bot, i, j)) * w.zwxb[ij];
        zv = zv + r1_hv_n[ij] * r1_2 * (A2(rCdU_bot, i, j + 1) + A2(rCdU_bot, i, j)) * w.zwyb[ij];
        zu = zu + r1_rau0 * utau[ij] * r1_hu_n[ij];
        zv = zv + r1_rau0 * vtau[ij] * r1_hv_n[ij];
        w.zu_frc[ij] = zu;
        w.zv_frc[ij] = zv;
    }
};
KERNEL void k_spg_p1b(DevDomain d, const double * __restrict__ un_b, const double * __restrict__ vn_b, const double * __restrict__ hu_n,
                          const double * __restrict__ hv_n, const double * __restrict__ r1_hu_n, const double * __restrict__ r1_hv_n, const double * __restrict__ ht_n,
                          const double * __restrict__ rCdU_bot, const double * __restrict__ utau, const double * __restrict__ vtau, const double * __restrict__ emp,
                          const double * __restrict__ rnf, SpgWork w)
{
    S_spg_p1b::run(vb_hw(), d, un_b, vn_b, hu_n, hv_n, r1_hu_n, r1_hv_n, ht_n, rCdU_bot, utau, vtau, emp, rnf, w);
}

// ---------------------------------------------------------------------------
// Phase 2 sub-steps
// ---------------------------------------------------------------------------
struct S_spg_s1 {
    static constexpr int NTHR = NH_BLOCK, WAVES = 1;
    __device__ __forceinline__ static void run(const VB &vb_, const DevDomain &d, double za1, double za2, double za3, double wgt2, const double * __restrict__ un_e, const double * __restrict__ ub_e, const double * __restrict__ ubb_e, const double * __restrict__ vn_e, const double * __restrict__ vb_e, const double * __restrict__ vbb_e, const double * __restrict__ hu_n, const double * __restrict__ hv_n, double * __restrict__ un_adv, double * __restrict__ vn_adv, SpgWork w)
    {
        THREAD_IJ_V(vb_, 1, 1);
        if (i > d.jpi || j > d.jpj) return;
        const Metrics<double> &m = d.md;
        const size_t ij = IJ(i, j);
        const double ue = za1 * un_e[ij] + za2 * ub_e[ij] + za3 * ubb_e[ij];
        const double ve = za1 * vn_e[ij] + za2 * vb_e[ij] + za3 * vbb_e[ij];
        if (d.nvor_scheme == NEMO_NP_ENT) {   // only the t-point Coriolis reads them
            w.ua_e[ij] = ue;
            w.va_e[ij] = ve;
        }
        const double zwx = m.e2u[ij] * ue * hu_n[ij];
        const double zwy = m.e1v[ij] * ve * hv_n[ij];
        w.zwx[ij] = zwx;
        w.zwy[ij] = zwy;
        un_adv[ij] = un_adv[ij] + wgt2 * zwx * m.r1_e2u[ij];
        vn_adv[ij] = vn_adv[ij] + wgt2 * zwy * m.r1_e1v[ij];
    }
};
KERNEL void k_spg_s1(DevDomain d, double za1, double za2, double za3, double wgt2, const double * __restrict__ un_e,
                         const double * __restrict__ ub_e, const double * __restrict__ ubb_e, const double * __restrict__ vn_e, const double * __restrict__ vb_e,
                         const double * __restrict__ vbb_e, const double * __restrict__ hu_n, const double * __restrict__ hv_n, double * __restrict__ un_adv, double * __restrict__ vn_adv,
                         SpgWork w)
{
    S_spg_s1::run(vb_hw(), d, za1, za2, za3, wgt2, un_e, ub_e, ubb_e, vn_e, vb_e, vbb_e, hu_n, hv_n, un_adv, vn_adv, w);
}

// the sub-step's sea-level levels and the AB3-AM4 interpolation weights
// (lines 930-940)
struct SpgSsh {
    double za0, za1, za2, za3;
    const double *sshn_e, *sshb_e, *sshbb_e;
};

// after sea level of the sub-step at (p,q) after lbc_lnk('T'), lines 905-925:
// the flux divergence evaluated at the fill's source point
__device__ __forceinline__ double spg_ssha(const DevDomain &d, double rdtbt, const SpgSsh &c, const SpgWork &w, int p,
                                           int q)
{
    int si, sj;
    const bool ok = lbc_src_safe(d, 'T', p, q, si, sj);
    const size_t s = IJ(si, sj);
    const double zhdiv = (A2(w.zwx, si, sj) - A2(w.zwx, si - 1, sj) + A2(w.zwy, si, sj) - A2(w.zwy, si, sj - 1)) *
                         d.md.r1_e1e2t[s];
    const double a = (c.sshn_e[s] - rdtbt * (w.zssh_frc[s] + zhdiv)) * SSMASK(s);
    return ok ? a : 0.0;
}
// zsshp2_e at (p,q) from the after level there
__device__ __forceinline__ double spg_p2(const SpgSsh &c, double ssha_e, size_t pq)
{
    return c.za0 * ssha_e + c.za1 * c.sshn_e[pq] + c.za2 * c.sshb_e[pq] + c.za3 * c.sshbb_e[pq];
}

// after velocities at (i,j), lines 1000-1110; p2c, p2e, p2n are zsshp2_e at
// (i,j), (i+1,j) and (i,j+1)
__device__ __forceinline__ void spg_new_vel(const DevDomain &d, int i, int j, double rdtbt, const double *un_e,
                                            const double *vn_e, const double *r1_hu_n, const double *r1_hv_n,
                                            const double *hu_n, const double *hv_n, const double *ht_n,
                                            const SpgWork &w, double p2c, double p2e, double p2n, double &ua,
                                            double &va)
{
    const Metrics<double> &m = d.md;
    const size_t ij = IJ(i, j);
    double z1_hu = 0.0, z1_hv = 0.0;
    if (d.nvor_scheme == NEMO_NP_ENT) {
        const double su = SSUMASK(ij), sv = SSVMASK(ij);
        z1_hu = su / (hu_n[ij] + 1.0 - su);
        z1_hv = sv / (hv_n[ij] + 1.0 - sv);
    }
    double zu_trd, zv_trd;
    bt_coriolis(d, i, j, w.zwx, w.zwy, w.zwz, w.ftne, w.ftnw, w.ftse, w.ftsw, w.ua_e, w.va_e, ht_n, z1_hu, z1_hv,
                zu_trd, zv_trd);
    zu_trd = zu_trd + w.zCdU_u[ij] * un_e[ij] * r1_hu_n[ij];
    zv_trd = zv_trd + w.zCdU_v[ij] * vn_e[ij] * r1_hv_n[ij];
    const double zu_spg = -(d.grav * (p2e - p2c) * m.r1_e1u[ij]);
    const double zv_spg = -(d.grav * (p2n - p2c) * m.r1_e2v[ij]);
    const double zwx = (1.0 - d.rn_scal_load) * zu_spg;
    const double zwy = (1.0 - d.rn_scal_load) * zv_spg;
    ua = (un_e[ij] + rdtbt * (zwx + zu_trd + w.zu_frc[ij])) * SSUMASK(ij);
    va = (vn_e[ij] + rdtbt * (zwy + zv_trd + w.zv_frc[ij])) * SSVMASK(ij);
}
// S2: after sea level at (i,j), its time sum and zsshp2_e
__device__ __forceinline__ void spg_s2_point(const DevDomain &d, double rdtbt, double wgt1, const SpgSsh &c,
                                             double *__restrict__ ssha_to, double *__restrict__ ssha,
                                             const SpgWork &w, int i, int j)
{
    const size_t ij = IJ(i, j);
    const double ssha_e = spg_ssha(d, rdtbt, c, w, i, j);
    ssha_to[ij] = ssha_e;
    // a zero weight leaves the sum (+0 from its reset) as it is
    if (wgt1 != 0.0) ssha[ij] = ssha[ij] + wgt1 * ssha_e;
    w.zsshp2_e[ij] = spg_p2(c, ssha_e, ij);
}
struct S_spg_s2 {
    static constexpr int NTHR = NH_BLOCK, WAVES = 1;
    __device__ __forceinline__ static void run(const VB &vb_, const DevDomain &d, double rdtbt, double wgt1, SpgSsh c, double *__restrict__ ssha_to, double *__restrict__ ssha, SpgWork w)
    {
        THREAD_IJ_V(vb_, 1, 1);
        if (i > d.jpi || j > d.jpj) return;
        spg_s2_point(d, rdtbt, wgt1, c, ssha_to, ssha, w, i, j);
    }
};
KERNEL void k_spg_s2(DevDomain d, double rdtbt, double wgt1, SpgSsh c, double *__restrict__ ssha_to,
                     double *__restrict__ ssha, SpgWork w)
{
    S_spg_s2::run(vb_hw(), d, rdtbt, wgt1, c, ssha_to, ssha, w);
}

__device__ __forceinline__ void spg_new_vel_p2(const DevDomain &d, int p, int q, double rdtbt, const double *un_e,
                                               const double *vn_e, const double *r1_hu_n, const double *r1_hv_n,
                                               const double *hu_n, const double *hv_n, const double *ht_n,
                                               const SpgWork &w, double &ua, double &va)
{
    spg_new_vel(d, p, q, rdtbt, un_e, vn_e, r1_hu_n, r1_hv_n, hu_n, hv_n, ht_n, w, A2(w.zsshp2_e, p, q),
                A2(w.zsshp2_e, p + 1, q), A2(w.zsshp2_e, p, q + 1), ua, va);
}

// S3: after velocities and the next sub-step's extrapolation
__device__ __forceinline__ void spg_s3_point(const DevDomain &d, double rdtbt, double wgt1,
                                             const double *__restrict__ un_e, const double *__restrict__ vn_e,
                                             double *__restrict__ ubb_to_a, double *__restrict__ vbb_to_a,
                                             const double *__restrict__ r1_hu_n, const double *__restrict__ r1_hv_n,
                                             const double *__restrict__ hu_n, const double *__restrict__ hv_n,
                                             const double *__restrict__ ht_n, double *__restrict__ ua_b,
                                             double *__restrict__ va_b, const SpgWork &w, const SpgNext &nx, int i,
                                             int j)
{
    const size_t ij = IJ(i, j);
    // the after velocities at (i,j) after lbc_lnk('U') / ('V'): evaluated at
    // their source points; on the interior both sources are (i,j) and one
    // evaluation serves both components
    int ui, uj, vi, vj;
    const bool oku = lbc_src_safe(d, 'U', i, j, ui, uj);
    const bool okv = lbc_src_safe(d, 'V', i, j, vi, vj);
    double ua1, va1, ua2, va2;
    spg_new_vel_p2(d, ui, uj, rdtbt, un_e, vn_e, r1_hu_n, r1_hv_n, hu_n, hv_n, ht_n, w, ua1, va1);
    if (ui == vi && uj == vj) {
        ua2 = ua1;
        va2 = va1;
    } else {
        spg_new_vel_p2(d, vi, vj, rdtbt, un_e, vn_e, r1_hu_n, r1_hv_n, hu_n, hv_n, ht_n, w, ua2, va2);
    }
    const double ua = oku ? ua1 : 0.0, va = okv ? va2 : 0.0;
    const double ub_now = un_e[ij], vb_now = vn_e[ij];
    ubb_to_a[ij] = ua;
    vbb_to_a[ij] = va;
    if (wgt1 != 0.0) {   // a zero weight leaves the sums (+0 from their reset) as they are
        ua_b[ij] = ua_b[ij] + wgt1 * ua;
        va_b[ij] = va_b[ij] + wgt1 * va;
    }
    if (!nx.on) return;
    // the next sub-step's extrapolation, fluxes and un_adv sum (pointwise in
    // the levels now, before and after of this point), as k_spg_s1 computes them
    const Metrics<double> &m = d.md;
    const double ue = nx.za1 * ua + nx.za2 * ub_now + nx.za3 * nx.ub_before[ij];
    const double ve = nx.za1 * va + nx.za2 * vb_now + nx.za3 * nx.vb_before[ij];
    if (d.nvor_scheme == NEMO_NP_ENT) {   // only the t-point Coriolis reads them
        nx.ua_e[ij] = ue;
        nx.va_e[ij] = ve;
    }
    const double zwx = m.e2u[ij] * ue * hu_n[ij];
    const double zwy = m.e1v[ij] * ve * hv_n[ij];
    nx.zwx[ij] = zwx;
    nx.zwy[ij] = zwy;
    nx.un_adv[ij] = nx.un_adv[ij] + nx.wgt2 * zwx * m.r1_e2u[ij];
    nx.vn_adv[ij] = nx.vn_adv[ij] + nx.wgt2 * zwy * m.r1_e1v[ij];
}
struct S_spg_s3 {
    static constexpr int NTHR = NH_BLOCK, WAVES = 1;
    __device__ __forceinline__ static void run(const VB &vb_, const DevDomain &d, double rdtbt, double wgt1, const double *__restrict__ un_e, const double *__restrict__ vn_e, double *__restrict__ ubb_to_a, double *__restrict__ vbb_to_a, const double *__restrict__ r1_hu_n, const double *__restrict__ r1_hv_n, const double *__restrict__ hu_n, const double *__restrict__ hv_n, const double *__restrict__ ht_n, double *__restrict__ ua_b, double *__restrict__ va_b, SpgWork w, SpgNext nx)
    {
        THREAD_IJ_V(vb_, 1, 1);
        if (i > d.jpi || j > d.jpj) return;
        spg_s3_point(d, rdtbt, wgt1, un_e, vn_e, ubb_to_a, vbb_to_a, r1_hu_n, r1_hv_n, hu_n, hv_n, ht_n, ua_b, va_b, w,
                     nx, i, j);
    }
};
KERNEL void k_spg_s3(DevDomain d, double rdtbt, double wgt1, const double *__restrict__ un_e,
                     const double *__restrict__ vn_e, double *__restrict__ ubb_to_a, double *__restrict__ vbb_to_a,
                     const double *__restrict__ r1_hu_n, const double *__restrict__ r1_hv_n,
                     const double *__restrict__ hu_n, const double *__restrict__ hv_n, const double *__restrict__ ht_n,
                     double *__restrict__ ua_b, double *__restrict__ va_b, SpgWork w, SpgNext nx)
{
    S_spg_s3::run(vb_hw(), d, rdtbt, wgt1, un_e, vn_e, ubb_to_a, vbb_to_a, r1_hu_n, r1_hv_n, hu_n, hv_n, ht_n, ua_b, va_b, w, nx);
}

// ---------------------------------------------------------------------------
// Phase 2 as one cooperative launch: every sub-step's S2 and S3 over the
// workgroups' tiles, separated by grid barriers, with the level buffers
// rotating in the kernel as the host loop rotates them.  Every workgroup is
// resident (the grid is sized from the occupancy), so a barrier cannot wait
// for a workgroup that has not started; each wait is still bounded, and a
// barrier that gives up sets *status, which the library reports.
// ---------------------------------------------------------------------------
#define SPG_MAXIT 96
#define SPG_SPIN_LIMIT (1u << 20)
struct SpgLoop {
    int icycle, ll_init, ntx, ntiles;
    double rdtbt;
    double wgt1[SPG_MAXIT], wgt2[SPG_MAXIT];
    double *ube[3], *vbe[3], *sse[4];
    double *ua_e2[2], *va_e2[2], *zwx2[2], *zwy2[2];
    int ru0[3], rs0[4];
    const double *r1_hu_n, *r1_hv_n, *hu_n, *hv_n, *ht_n;
    double *ua_b, *va_b, *ssha, *un_adv, *vn_adv;
};

// bar[0]: arrivals, bar[1]: generation.  The last workgroup to arrive resets
// the count and moves the generation on; the others wait for that.  Device-
// scope fences around the barrier make the workgroups' stores of one phase
// visible to the loads of the next in every XCD.
__device__ __forceinline__ void spg_grid_barrier(unsigned *bar, int *status, unsigned nblocks)
{
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned *count = bar, *gen = bar + 1;
        const unsigned g = __hip_atomic_load(gen, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __threadfence();
        const unsigned a = __hip_atomic_fetch_add(count, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT) + 1u;
        if (a == nblocks) {
            __hip_atomic_store(count, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_fetch_add(gen, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
        } else {
            unsigned it = 0;
            while (__hip_atomic_load(gen, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) == g) {
                if (++it > SPG_SPIN_LIMIT) {
                    __hip_atomic_store(status, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    break;
                }
                __builtin_amdgcn_s_sleep(1);
            }
        }
        __threadfence();
    }
    __syncthreads();
}

// NTH threads per workgroup, tiles of 64 x NTH/64 points
template <int NTH>
__global__ __launch_bounds__(NTH) void k_spg_loop(DevDomain d, SpgWork w, SpgLoop L, unsigned *bar, int *status)
{
    constexpr int TYL = NTH / 64;
    int ru[3] = {L.ru0[0], L.ru0[1], L.ru0[2]}, rs[4] = {L.rs0[0], L.rs0[1], L.rs0[2], L.rs0[3]};
    const unsigned nb = gridDim.x;
    const int tx = (int)(threadIdx.x & 63u), ty = (int)(threadIdx.x >> 6);
    for (int jn = 1; jn <= L.icycle; jn++) {
        const int cur = (jn - 1) & 1, nxt = jn & 1;
        SpgWork wc = w;
        wc.ua_e = L.ua_e2[cur]; wc.va_e = L.va_e2[cur]; wc.zwx = L.zwx2[cur]; wc.zwy = L.zwy2[cur];
        double za0, za1, za2, za3;
        if (jn == 1 && L.ll_init) { za0 = 1.0; za1 = 0.0; za2 = 0.0; za3 = 0.0; }
        else if (jn == 2 && L.ll_init) { za0 = 1.0833333333333; za1 = -0.1666666666666; za2 = 0.0833333333333; za3 = 0.0; }
        else { za0 = 0.614; za1 = 0.285; za2 = 0.088; za3 = 0.013; }
        const SpgSsh c{za0, za1, za2, za3, L.sse[rs[0]], L.sse[rs[1]], L.sse[rs[2]]};
        const double wgt1 = L.wgt1[jn - 1];
        for (int t = (int)blockIdx.x; t < L.ntiles; t += (int)nb) {
            const int i = (t % L.ntx) * 64 + tx + 1, j = (t / L.ntx) * TYL + ty + 1;
            if (i <= d.jpi && j <= d.jpj) spg_s2_point(d, L.rdtbt, wgt1, c, L.sse[rs[3]], L.ssha, wc, i, j);
        }
        spg_grid_barrier(bar, status, nb);
        SpgNext nx{};
        nx.on = jn < L.icycle;
        if (nx.on) {
            if (jn + 1 < 3 && L.ll_init) { nx.za1 = 1.0; nx.za2 = 0.0; nx.za3 = 0.0; }
            else { nx.za1 = 1.781105; nx.za2 = -1.06221; nx.za3 = 0.281105; }
            nx.wgt2 = L.wgt2[jn];
            nx.ub_before = L.ube[ru[1]];
            nx.vb_before = L.vbe[ru[1]];
            nx.ua_e = L.ua_e2[nxt]; nx.va_e = L.va_e2[nxt]; nx.zwx = L.zwx2[nxt]; nx.zwy = L.zwy2[nxt];
            nx.un_adv = L.un_adv; nx.vn_adv = L.vn_adv;
        }
        for (int t = (int)blockIdx.x; t < L.ntiles; t += (int)nb) {
            const int i = (t % L.ntx) * 64 + tx + 1, j = (t / L.ntx) * TYL + ty + 1;
            if (i <= d.jpi && j <= d.jpj)
                spg_s3_point(d, L.rdtbt, wgt1, L.ube[ru[0]], L.vbe[ru[0]], L.ube[ru[2]], L.vbe[ru[2]], L.r1_hu_n,
                             L.r1_hv_n, L.hu_n, L.hv_n, L.ht_n, L.ua_b, L.va_b, wc, nx, i, j);
        }
        if (jn < L.icycle) spg_grid_barrier(bar, status, nb);
        int q;
        q = ru[2]; ru[2] = ru[1]; ru[1] = ru[0]; ru[0] = q;
        q = rs[3]; rs[3] = rs[2]; rs[2] = rs[1]; rs[1] = rs[0]; rs[0] = q;
    }
}

// ---------------------------------------------------------------------------
// Phase 3, lines 1214-1269
// ---------------------------------------------------------------------------
struct S_spg_p3a {
    static constexpr int NTHR = NH_BLOCK, WAVES = 1;
    __device__ __forceinline__ static void run(const VB &vb_, const DevDomain &d, int first_euler, double * __restrict__ un_adv, double * __restrict__ vn_adv, double * __restrict__ ub2_b, double * __restrict__ vb2_b, double * __restrict__ un_bf, double * __restrict__ vn_bf)
    {
        THREAD_IJ_V(vb_, 1, 1);
        if (i > d.jpi || j > d.jpj) return;
        const size_t ij = IJ(i, j);
        const double r1_2 = 0.5, atfp = d.rn_atfp;
        const double zwx = un_adv[ij], zwy = vn_adv[ij];
        if (!first_euler) {
            un_adv[ij] = r1_2 * (ub2_b[ij] + zwx - atfp * un_bf[ij]);
            vn_adv[ij] = r1_2 * (vb2_b[ij] + zwy - atfp * vn_bf[ij]);
            un_bf[ij] = atfp * un_bf[ij] + (zwx - ub2_b[ij]);
            vn_bf[ij] = atfp * vn_bf[ij] + (zwy - vb2_b[ij]);
        } else {
            un_bf[ij] = 0.0;
            vn_bf[ij] = 0.0;
        }
        ub2_b[ij] = zwx;
        vb2_b[ij] = zwy;
    }
};
KERNEL void k_spg_p3a(DevDomain d, int first_euler, double * __restrict__ un_adv, double * __restrict__ vn_adv, double * __restrict__ ub2_b,
                          double * __restrict__ vb2_b, double * __restrict__ un_bf, double * __restrict__ vn_bf)
{
    S_spg_p3a::run(vb_hw(), d, first_euler, un_adv, vn_adv, ub2_b, vb2_b, un_bf, vn_bf);
}

template <typename T>
struct S_spg_p3b {
    static constexpr int NTHR = NH_BLOCK, WAVES = 1;
    __device__ __forceinline__ static void run(const VB &vb_, const DevDomain &d, double r1_2dt_b, T * __restrict__ ua, T * __restrict__ va, T * __restrict__ un, T * __restrict__ vn, const double * __restrict__ ua_b, const double * __restrict__ va_b, const double * __restrict__ ub_b, const double * __restrict__ vb_b, const double * __restrict__ un_adv, const double * __restrict__ vn_adv, const double * __restrict__ r1_hu_n, const double * __restrict__ r1_hv_n, const double * __restrict__ un_b, const double * __restrict__ vn_b)
    {
        THREAD_IJ_V(vb_, 1, 1);
        if (i > d.jpi || j > d.jpj) return;
        const size_t ij = IJ(i, j);
        const double du = (ua_b[ij] - ub_b[ij]) * r1_2dt_b, dv = (va_b[ij] - vb_b[ij]) * r1_2dt_b;
        const double cu = un_adv[ij] * r1_hu_n[ij], cv = vn_adv[ij] * r1_hv_n[ij];
        const double ubn = un_b[ij], vbn = vn_b[ij];
        for (int k = 1; k <= d.jpk - 1; k++) {
            const size_t p = IJK(i, j, k);
            ua[p] = (T)((double)ua[p] + du);
            va[p] = (T)((double)va[p] + dv);
            un[p] = (T)(((double)un[p] + cu - ubn) * UMASK(double, i, j, k));
            vn[p] = (T)(((double)vn[p] + cv - vbn) * VMASK(double, i, j, k));
        }
    }
};
template <typename T>
KERNEL void k_spg_p3b(DevDomain d, double r1_2dt_b, T * __restrict__ ua, T * __restrict__ va, T * __restrict__ un, T * __restrict__ vn, const double * __restrict__ ua_b,
                          const double * __restrict__ va_b, const double * __restrict__ ub_b, const double * __restrict__ vb_b, const double * __restrict__ un_adv,
                          const double * __restrict__ vn_adv, const double * __restrict__ r1_hu_n, const double * __restrict__ r1_hv_n, const double * __restrict__ un_b,
                          const double * __restrict__ vn_b)
{
    S_spg_p3b<T>::run(vb_hw(), d, r1_2dt_b, ua, va, un, vn, ua_b, va_b, ub_b, vb_b, un_adv, vn_adv, r1_hu_n, r1_hv_n, un_b, vn_b);
}

// phase 2 initialisation, lines 665-702, and the reset of the time sums: the
// now levels from un_b, vn_b, sshn; with ll_init the before levels at zero
// (one kernel instead of memset/memcpy calls, so the call can be recorded
// into a graph)
struct S_spg_init {
    static constexpr int NTHR = NH_BLOCK, WAVES = 1;
    __device__ __forceinline__ static void run(const VB &vb_, const DevDomain &d, int ll_init, const double *__restrict__ un_b, const double *__restrict__ vn_b, const double *__restrict__ sshn, double *__restrict__ ub0, double *__restrict__ vb0, double *__restrict__ ss0, double *__restrict__ ub1, double *__restrict__ vb1, double *__restrict__ ss1, double *__restrict__ ub2, double *__restrict__ vb2, double *__restrict__ ss2, double *__restrict__ ua_b, double *__restrict__ va_b, double *__restrict__ ssha, double *__restrict__ un_adv, double *__restrict__ vn_adv)
    {
        THREAD_IJ_V(vb_, 1, 1);
        if (i > d.jpi || j > d.jpj) return;
        const size_t ij = IJ(i, j);
        if (ll_init) {
            ub1[ij] = 0.0; vb1[ij] = 0.0; ss1[ij] = 0.0;
            ub2[ij] = 0.0; vb2[ij] = 0.0; ss2[ij] = 0.0;
        }
        ss0[ij] = sshn[ij];
        ub0[ij] = un_b[ij];
        vb0[ij] = vn_b[ij];
        ua_b[ij] = 0.0; va_b[ij] = 0.0; ssha[ij] = 0.0; un_adv[ij] = 0.0; vn_adv[ij] = 0.0;
    }
};
KERNEL void k_spg_init(DevDomain d, int ll_init, const double *__restrict__ un_b, const double *__restrict__ vn_b,
                       const double *__restrict__ sshn, double *__restrict__ ub0, double *__restrict__ vb0,
                       double *__restrict__ ss0, double *__restrict__ ub1, double *__restrict__ vb1,
                       double *__restrict__ ss1, double *__restrict__ ub2, double *__restrict__ vb2,
                       double *__restrict__ ss2, double *__restrict__ ua_b, double *__restrict__ va_b,
                       double *__restrict__ ssha, double *__restrict__ un_adv, double *__restrict__ vn_adv)
{
    S_spg_init::run(vb_hw(), d, ll_init, un_b, vn_b, sshn, ub0, vb0, ss0, ub1, vb1, ss1, ub2, vb2, ss2, ua_b, va_b, ssha, un_adv, vn_adv);
}

// ---------------------------------------------------------------------------
// host driver
// ---------------------------------------------------------------------------

int spg_ts_weights(const DevDomain &d, int ll_av, int ll_fw, int *jpit, double *zwgt1, double *zwgt2)
{
    const int nn_baro = d.nn_baro;
    for (int jn = 0; jn < 3 * nn_baro; jn++) zwgt1[jn] = zwgt2[jn] = 0.0;
    const int jic = ll_fw ? nn_baro : 2 * nn_baro;
    if (ll_av) {
        if (d.nn_bt_flt == 0) {
            zwgt1[jic - 1] = 1.0;
            *jpit = jic;
        } else if (d.nn_bt_flt == 1 || d.nn_bt_flt == 2) {
            // float(): default REAL arithmetic, DYN/dynspg_ts.F90:1350 and 1359
            const float lim = d.nn_bt_flt == 1 ? 0.5f : 1.0f;
            for (int jn = 1; jn <= 3 * nn_baro; jn++) {
                const float za1 = fabsf((float)(jn - jic)) / (float)nn_baro;
                if ((double)za1 < (double)lim) {
                    zwgt1[jn - 1] = 1.0;
                    *jpit = jn;
                }
            }
        } else {
            return nh::set_error(NEMO_HIP_ERR_UNSUPPORTED, "ts_wgt: nn_bt_flt = %d", d.nn_bt_flt);
        }
    } else {
        zwgt1[jic - 1] = 1.0;
        *jpit = jic;
    }
    for (int jn = 1; jn <= *jpit; jn++)
        for (int ji = jn; ji <= *jpit; ji++) zwgt2[jn - 1] = zwgt2[jn - 1] + zwgt1[ji - 1];
    double s1 = 0.0, s2 = 0.0;
    for (int jn = 1; jn <= *jpit; jn++) {
        s1 += zwgt1[jn - 1];
        s2 += zwgt2[jn - 1];
    }
    const double za1 = 1.0 / s1, za2 = 1.0 / s2;
    for (int jn = 1; jn <= *jpit; jn++) {
        zwgt1[jn - 1] = zwgt1[jn - 1] * za1;
        zwgt2[jn - 1] = zwgt2[jn - 1] * za2;
    }
    return 0;
}

// phase 2 by one cooperative launch of k_spg_loop: returns false (and
// launches nothing) when the tuning switch is off, a graph is being
// recorded, the weights do not fit the kernel's table or the device cannot
// hold the grid resident; the caller then runs the two-kernel loop.
// Switch value 1: one 1024-thread workgroup per CU (the fewest workgroups
// at each barrier); 2: 256-thread workgroups, as many as stay resident.
static bool spg_loop_coop(const DevDomain &d, const SpgWork &w, const nemo_hip_spg_ts_t *st, int icycle, int ll_init,
                          double rdtbt, const double *wgt1, const double *wgt2, double *const *ube,
                          double *const *vbe, double *const *sse, double *const *ua_e2, double *const *va_e2,
                          double *const *zwx2, double *const *zwy2, const int *ru, const int *rs)
{
    const int mode = nh::tune(nh::T_SPG_PERSISTENT);
    if (!mode || nh::capturing() || icycle > SPG_MAXIT || icycle < 1) return false;
    const int nth = mode == 1 ? 1024 : NH_BLOCK;
    const void *kern = mode == 1 ? (const void *)k_spg_loop<1024> : (const void *)k_spg_loop<NH_BLOCK>;
    unsigned *bar = nh::spg_barrier();
    int *status = (int *)(bar + 2);
    int dev = 0, coop = 0, ncu = 0, per_cu = 0;
    if (hipGetDevice(&dev) != hipSuccess) return false;
    if (hipDeviceGetAttribute(&coop, hipDeviceAttributeCooperativeLaunch, dev) != hipSuccess || !coop) return false;
    if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return false;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, nth, 0) != hipSuccess || per_cu < 1)
        return false;
    if (mode == 1) per_cu = 1;
    SpgLoop L{};
    L.icycle = icycle;
    L.ll_init = ll_init;
    L.ntx = (d.jpi + 63) / 64;
    L.ntiles = L.ntx * ((d.jpj + nth / 64 - 1) / (nth / 64));
    L.rdtbt = rdtbt;
    for (int q = 0; q < icycle; q++) {
        L.wgt1[q] = wgt1[q];
        L.wgt2[q] = wgt2[q];
    }
    for (int q = 0; q < 3; q++) { L.ube[q] = ube[q]; L.vbe[q] = vbe[q]; L.ru0[q] = ru[q]; }
    for (int q = 0; q < 4; q++) { L.sse[q] = sse[q]; L.rs0[q] = rs[q]; }
    for (int q = 0; q < 2; q++) { L.ua_e2[q] = ua_e2[q]; L.va_e2[q] = va_e2[q]; L.zwx2[q] = zwx2[q]; L.zwy2[q] = zwy2[q]; }
    L.r1_hu_n = st->r1_hu_n; L.r1_hv_n = st->r1_hv_n; L.hu_n = st->hu_n; L.hv_n = st->hv_n; L.ht_n = st->ht_n;
    L.ua_b = st->ua_b; L.va_b = st->va_b; L.ssha = st->ssha; L.un_adv = st->un_adv; L.vn_adv = st->vn_adv;
    const unsigned grid = (unsigned)(per_cu * ncu < L.ntiles ? per_cu * ncu : L.ntiles);
    DevDomain dd = d;
    SpgWork ww = w;
    void *args[] = {(void *)&dd, (void *)&ww, (void *)&L, (void *)&bar, (void *)&status};
    hipEvent_t ev0, ev1;
    const bool timed = nh::timing_pair("dyn_spg_ts", "k_spg_loop", &ev0, &ev1);
    if (timed) (void)hipEventRecord(ev0, nh::stream());
    const hipError_t e = hipLaunchCooperativeKernel(kern, dim3(grid), dim3(nth), args, 0, nh::stream());
    if (timed) (void)hipEventRecord(ev1, nh::stream());
    if (e != hipSuccess) {
        (void)hipGetLastError();
        return false;
    }
    nh::spg_loop_launched();
    return true;
}

template <typename T>
int launch_dyn_spg_ts(int kt, const nemo_hip_spg_ts_t *st, int fused)
{
    const DevDomain &d = nh::dom();
    if (!d.ln_bt_fw) return nh::set_error(NEMO_HIP_ERR_UNSUPPORTED, "dyn_spg_ts: ln_bt_fw = F is not implemented");
    T *ua = (T *)st->ua, *va = (T *)st->va, *un = (T *)st->un, *vn = (T *)st->vn;
    SpgWork w;
    double *s2[32];
    for (int q = 0; q < 32; q++) s2[q] = nh::scratch2(q);
    w.zu_frc = s2[0]; w.zv_frc = s2[1]; w.zwxb = s2[2]; w.zwyb = s2[3]; w.zCdU_u = s2[4]; w.zCdU_v = s2[5];
    w.zssh_frc = s2[6]; w.zwz = s2[7]; w.zhf = s2[8]; w.ftne = s2[9]; w.ftnw = s2[10]; w.ftse = s2[11];
    w.ftsw = s2[12]; w.ua_e = s2[13]; w.va_e = s2[14]; w.zwx = s2[15]; w.zwy = s2[16]; w.zsshp2_e = s2[17];
    double *ube[3] = {s2[18], s2[19], s2[20]}, *vbe[3] = {s2[21], s2[22], s2[23]}, *sse[4] = {s2[24], s2[25], s2[26], nh::scratch2(33)};

    const double z2dt_bf = (kt == d.nit000 && d.neuler == 0) ? d.rdt : 2.0 * d.rdt;
    const double r1_2dt_b = 1.0 / z2dt_bf;
    const double rdtbt = d.rdt / (double)d.nn_baro;
    int ll_init = d.ln_bt_av;
    if (kt == d.nit000 && d.neuler == 0) ll_init = 1;
    double wgt1[1024], wgt2[1024];
    int icycle = 0;
    if (3 * d.nn_baro > 1024) return nh::set_error(NEMO_HIP_ERR_ARG, "nn_baro too large");
    int rc = spg_ts_weights(d, d.ln_bt_av, 1, &icycle, wgt1, wgt2);
    if (rc) return rc;

    Grid2 g = grid2(d.jpi, d.jpj);
    LAUNCHB("dyn_spg_ts", k_spg_cor1, (S_spg_cor1), g, d, st->ht_n, w);
    LAUNCHB("dyn_spg_ts", k_spg_cor2, (S_spg_cor2), g, d, st->ht_n, w);
    if (d.nvor_scheme == NEMO_NP_EEN) LAUNCHB("dyn_spg_ts", k_spg_cor3, (S_spg_cor3), g, d, w);
    if (!fused)
        LAUNCHB("dyn_spg_ts", k_spg_p1a<T>, (S_spg_p1a<T>), g, d, ua, va, (const T *)un, (const T *)vn, st->r1_hu_n, st->r1_hv_n,
               st->un_b, st->vn_b, w);
    LAUNCHB("dyn_spg_ts", k_spg_p1b, (S_spg_p1b), g, d, st->un_b, st->vn_b, st->hu_n, st->hv_n, st->r1_hu_n, st->r1_hv_n, st->ht_n,
           st->rCdU_bot, st->utau, st->vtau, st->emp, st->rnf, w);

    // phase 2 initialisation, lines 665-702.  The before levels are set
    // afresh when ll_init, so the buffers restart from the first rotation
    // (the same values in the same buffers at every such call, which a
    // recorded graph relies on); without it the levels of the last call
    // carry over and the rotation continues
    int *ru = nh::spg_rot_u(), *rs = nh::spg_rot_s();
    if (ll_init) {
        for (int q = 0; q < 3; q++) ru[q] = q;
        for (int q = 0; q < 4; q++) rs[q] = q;
    } else if (nh::capturing()) {
        return nh::set_error(NEMO_HIP_ERR_STATE,
                             "dyn_spg_ts: ln_bt_av = F carries the barotropic levels between calls; not recordable");
    }
    LAUNCHB("dyn_spg_ts", k_spg_init, (S_spg_init), g, d, ll_init, st->un_b, st->vn_b, st->sshn, ube[ru[0]], vbe[ru[0]], sse[rs[0]],
           ube[ru[1]], vbe[ru[1]], sse[rs[1]], ube[ru[2]], vbe[ru[2]], sse[rs[2]], st->ua_b, st->va_b, st->ssha,
           st->un_adv, st->vn_adv);

    // two sets of the sub-step fluxes: sub-step jn reads set (jn-1)&1 and its
    // velocity kernel writes the extrapolation of sub-step jn+1 into the other
    double *ua_e2[2] = {w.ua_e, nh::scratch2(29)}, *va_e2[2] = {w.va_e, nh::scratch2(30)};
    double *zwx2[2] = {w.zwx, nh::scratch2(31)}, *zwy2[2] = {w.zwy, nh::scratch2(32)};
    auto ab3 = [&](int jn, double &za1, double &za2, double &za3) {
        if (jn < 3 && ll_init) { za1 = 1.0; za2 = 0.0; za3 = 0.0; }
        else { za1 = 1.781105; za2 = -1.06221; za3 = 0.281105; }
    };
    {
        double za1, za2, za3;
        ab3(1, za1, za2, za3);
        LAUNCHB("dyn_spg_ts", k_spg_s1, (S_spg_s1), g, d, za1, za2, za3, wgt2[0], ube[ru[0]], ube[ru[1]], ube[ru[2]],
               vbe[ru[0]], vbe[ru[1]], vbe[ru[2]], st->hu_n, st->hv_n, st->un_adv, st->vn_adv, w);
    }
    if (spg_loop_coop(d, w, st, icycle, ll_init, rdtbt, wgt1, wgt2, ube, vbe, sse, ua_e2, va_e2, zwx2, zwy2, ru,
                      rs)) {
        // the kernel rotated the levels icycle times; so do the host's indices
        for (int jn = 1; jn <= icycle; jn++) {
            int t;
            t = ru[2]; ru[2] = ru[1]; ru[1] = ru[0]; ru[0] = t;
            t = rs[3]; rs[3] = rs[2]; rs[2] = rs[1]; rs[1] = rs[0]; rs[0] = t;
        }
        icycle = 0;   // the loop below has nothing left to do
    }
    for (int jn = 1; jn <= icycle; jn++) {
        const int cur = (jn - 1) & 1, nxt = jn & 1;
        SpgWork wc = w;
        wc.ua_e = ua_e2[cur]; wc.va_e = va_e2[cur]; wc.zwx = zwx2[cur]; wc.zwy = zwy2[cur];
        double za0, za1, za2, za3;
        if (jn == 1 && ll_init) { za0 = 1.0; za1 = 0.0; za2 = 0.0; za3 = 0.0; }
        else if (jn == 2 && ll_init) { za0 = 1.0833333333333; za1 = -0.1666666666666; za2 = 0.0833333333333; za3 = 0.0; }
        else { za0 = 0.614; za1 = 0.285; za2 = 0.088; za3 = 0.013; }
        const SpgSsh c{za0, za1, za2, za3, sse[rs[0]], sse[rs[1]], sse[rs[2]]};
        LAUNCHB("dyn_spg_ts", k_spg_s2, (S_spg_s2), g, d, rdtbt, wgt1[jn - 1], c, sse[rs[3]], st->ssha, wc);
        SpgNext nx{};
        nx.on = jn < icycle;
        if (nx.on) {
            ab3(jn + 1, nx.za1, nx.za2, nx.za3);
            nx.wgt2 = wgt2[jn];
            // after the rotation: now <- the level written here, before <- now, before-before <- before
            nx.ub_before = ube[ru[1]];
            nx.vb_before = vbe[ru[1]];
            nx.ua_e = ua_e2[nxt]; nx.va_e = va_e2[nxt]; nx.zwx = zwx2[nxt]; nx.zwy = zwy2[nxt];
            nx.un_adv = st->un_adv; nx.vn_adv = st->vn_adv;
        }
        LAUNCHB("dyn_spg_ts", k_spg_s3, (S_spg_s3), g, d, rdtbt, wgt1[jn - 1], ube[ru[0]], vbe[ru[0]], ube[ru[2]], vbe[ru[2]],
               st->r1_hu_n, st->r1_hv_n, st->hu_n, st->hv_n, st->ht_n, st->ua_b, st->va_b, wc, nx);
        // swap: bb <- b, b <- n, n <- a (the after velocities were written
        // over bb; the after sea level into the fourth buffer)
        int t;
        t = ru[2]; ru[2] = ru[1]; ru[1] = ru[0]; ru[0] = t;
        t = rs[3]; rs[3] = rs[2]; rs[2] = rs[1]; rs[1] = rs[0]; rs[0] = t;
    }
    LAUNCHB("dyn_spg_ts", k_spg_p3a, (S_spg_p3a), g, d, (kt == d.nit000 && d.neuler == 0) ? 1 : 0, st->un_adv, st->vn_adv, st->ub2_b,
           st->vb2_b, st->un_bf, st->vn_bf);
    if (!fused)
        LAUNCHB("dyn_spg_ts", k_spg_p3b<T>, (S_spg_p3b<T>), g, d, r1_2dt_b, ua, va, un, vn, st->ua_b, st->va_b, st->ub_b, st->vb_b,
               st->un_adv, st->vn_adv, st->r1_hu_n, st->r1_hv_n, st->un_b, st->vn_b);
    return 0;
}
template int launch_dyn_spg_ts<double>(int, const nemo_hip_spg_ts_t *, int);
template int launch_dyn_spg_ts<float>(int, const nemo_hip_spg_ts_t *, int);


