// dyn.hip -- momentum operators of the ocean time step on gfx950.
//
//   dyn_keg       DYN/dynkeg.F90:46-192    kinetic-energy gradient (C2, Hollingsworth)
//   dyn_zad       DYN/dynzad.F90:38-121    vertical advection (vector form)
//   dyn_adv_cen2  DYN/dynadv_cen2.F90:37-144 flux-form advection
//   dyn_vor       DYN/dynvor.F90:98-828    vorticity: ENS, ENE, MIX, ENT, EET, EEN
//   dyn_zdf       DYN/dynzdf.F90:47-499    implicit vertical diffusion + drag
//
// Same conventions as tra.hip: one thread per (i,j) column, k-marched,
// coalesced along i; the lbc_lnk fills that separate reference loop nests
// are folded in through lbc_src; expression order follows the reference.
#include "common.hpp"
#include "kernels.hpp"

enum { np_COR = 1, np_RVO = 2, np_MET = 3, np_CRV = 4, np_CME = 5 };

// ---------------------------------------------------------------------------
// vector-form advection, dyn_keg then dyn_zad, fused per element
// (ua <- (ua - keg) - zad, the order of the two reference calls)
// ---------------------------------------------------------------------------
template <typename T>
__device__ __forceinline__ T keg_c2(const DevDomain &d, const T *un, const T *vn, int p, int q, int k)
{
    const T u1 = A3(un, p - 1, q, k), u2 = A3(un, p, q, k), v1 = A3(vn, p, q - 1, k), v2 = A3(vn, p, q, k);
    const T zu = u1 * u1 + u2 * u2;
    const T zv = v1 * v1 + v2 * v2;
    return T(0.25) * (zv + zu);
}
template <typename T>
__device__ __forceinline__ T keg_hw(const DevDomain &d, const T *un, const T *vn, int p, int q, int k)
{
    int sp, sq;
    if (!lbc_src(d, 'T', p, q, sp, sq)) return T(0);     // lbc_lnk('dynkeg', zhke, 'T')
    p = sp;
    q = sq;
    const T r1_48 = T(1) / T(48);
    const T a = A3(un, p - 1, q - 1, k) + A3(un, p - 1, q + 1, k);
    const T b = A3(un, p, q - 1, k) + A3(un, p, q + 1, k);
    const T u1 = A3(un, p - 1, q, k), u2 = A3(un, p, q, k);
    const T zu = T(8) * (u1 * u1 + u2 * u2) + a * a + b * b;
    const T c = A3(vn, p - 1, q - 1, k) + A3(vn, p + 1, q - 1, k);
    const T e = A3(vn, p - 1, q, k) + A3(vn, p + 1, q, k);
    const T v1 = A3(vn, p, q - 1, k), v2 = A3(vn, p, q, k);
    const T zv = T(8) * (v1 * v1 + v2 * v2) + c * c + e * e;
    return r1_48 * (zv + zu);
}

template <typename T>
struct S_dyn_keg_zad {
    static constexpr int NTHR = NH_BLOCK, WAVES = 1;
    __device__ __forceinline__ static void run(const VB &vb_, const DevDomain &d, int do_keg, int kscheme, int do_zad, const T * __restrict__ un, const T * __restrict__ vn, const T * __restrict__ wn, T * __restrict__ ua, T * __restrict__ va)
    {
        THREAD_IJ_V(vb_, 2, 2);
        if (i > d.jpi - 1 || j > d.jpj - 1) return;
        const Metrics<T> &m = met<T>(d);
        const size_t ij = IJ(i, j);
        const T e1u = m.e1u[ij], e2v = m.e2v[ij], r1_e1e2u = m.r1_e1e2u[ij], r1_e1e2v = m.r1_e1e2v[ij];
        const T e1e2t = m.e1e2t[ij], e1e2t_e = m.e1e2t[IJ(i + 1, j)], e1e2t_n = m.e1e2t[IJ(i, j + 1)];
        T zwuw_k = T(0), zwvw_k = T(0);    // surface flux zero, dynzad.F90:93-94
        for (int k = 1; k <= d.jpk - 1; k++) {
            const size_t p = IJK(i, j, k);
            T a = ua[p], b = va[p];
            if (do_keg) {
                T h, he, hn;
                if (kscheme == 0) {
                    h = keg_c2(d, un, vn, i, j, k);
                    he = keg_c2(d, un, vn, i + 1, j, k);
                    hn = keg_c2(d, un, vn, i, j + 1, k);
                } else {
                    h = keg_hw(d, un, vn, i, j, k);
                    he = keg_hw(d, un, vn, i + 1, j, k);
                    hn = keg_hw(d, un, vn, i, j + 1, k);
                }
                a = a - (he - h) / e1u;
                b = b - (hn - h) / e2v;
            }
            if (do_zad) {
                T zwuw_kp1 = T(0), zwvw_kp1 = T(0);
                if (k + 1 <= d.jpk - 1) {
                    const int kk = k + 1;
                    const T zww = T(0.25) * e1e2t * A3(wn, i, j, kk);
                    const T zww_e = T(0.25) * e1e2t_e * A3(wn, i + 1, j, kk);
                    const T zww_n = T(0.25) * e1e2t_n * A3(wn, i, j + 1, kk);
                    zwuw_kp1 = (zww_e + zww) * (A3(un, i, j, kk - 1) - A3(un, i, j, kk));
                    zwvw_kp1 = (zww_n + zww) * (A3(vn, i, j, kk - 1) - A3(vn, i, j, kk));
                }
                a = a - (zwuw_k + zwuw_kp1) * r1_e1e2u / E3(m.e3u, ij, k);
                b = b - (zwvw_k + zwvw_kp1) * r1_e1e2v / E3(m.e3v, ij, k);
                zwuw_k = zwuw_kp1;
                zwvw_k = zwvw_kp1;
            }
            ua[p] = a;
            va[p] = b;
        }
    }
};
template <typename T>
KERNEL void k_dyn_keg_zad(DevDomain d, int do_keg, int kscheme, int do_zad, const T * __restrict__ un, const T * __restrict__ vn, const T * __restrict__ wn,
                              T * __restrict__ ua, T * __restrict__ va)
{
    S_dyn_keg_zad<T>::run(vb_hw(), d, do_keg, kscheme, do_zad, un, vn, wn, ua, va);
}

// ---------------------------------------------------------------------------
// dyn_adv_cen2, horizontal then vertical flux divergence per element
// ---------------------------------------------------------------------------
template <typename T>
__device__ __forceinline__ T cen2_zfu(const DevDomain &d, const T *un, int p, int q, int k)
{
    const Metrics<T> &m = met<T>(d);
    const size_t pq = IJ(p, q);
    return T(0.25) * m.e2u[pq] * E3(m.e3u, pq, k) * A3(un, p, q, k);
}
template <typename T>
__device__ __forceinline__ T cen2_zfv(const DevDomain &d, const T *vn, int p, int q, int k)
{
    const Metrics<T> &m = met<T>(d);
    const size_t pq = IJ(p, q);
    return T(0.25) * m.e1v[pq] * E3(m.e3v, pq, k) * A3(vn, p, q, k);
}
template <typename T>
__device__ __forceinline__ T cen2_zfw(const DevDomain &d, const T *wn, int p, int q, int k)
{
    return T(0.25) * met<T>(d).e1e2t[IJ(p, q)] * A3(wn, p, q, k);
}

template <typename T>
struct S_dyn_adv_cen2 {
    static constexpr int NTHR = NH_BLOCK, WAVES = 1;
    __device__ __forceinline__ static void run(const VB &vb_, const DevDomain &d, const T * __restrict__ un, const T * __restrict__ vn, const T * __restrict__ wn, T * __restrict__ ua, T * __restrict__ va)
    {
        THREAD_IJ_V(vb_, 2, 2);
        if (i > d.jpi - 1 || j > d.jpj - 1) return;
        const Metrics<T> &m = met<T>(d);
        const size_t ij = IJ(i, j);
        const T r1_e1e2u = m.r1_e1e2u[ij], r1_e1e2v = m.r1_e1e2v[ij];
        const int jpk = d.jpk;
        // surface fluxes (linear free surface), dynadv_cen2.F90:105-112
        T zfu_uw_k = T(0.5) * (m.e1e2t[ij] * A3(wn, i, j, 1) + m.e1e2t[IJ(i + 1, j)] * A3(wn, i + 1, j, 1)) * A3(un, i, j, 1);
        T zfv_vw_k = T(0.5) * (m.e1e2t[ij] * A3(wn, i, j, 1) + m.e1e2t[IJ(i, j + 1)] * A3(wn, i, j + 1, 1)) * A3(vn, i, j, 1);
        for (int k = 1; k <= jpk - 1; k++) {
            // horizontal: zfu_t(i+1), zfu_t(i), zfv_f(i,j), zfv_f(i,j-1)
            const T fu_im1 = cen2_zfu(d, un, i - 1, j, k), fu_i = cen2_zfu(d, un, i, j, k), fu_ip1 = cen2_zfu(d, un, i + 1, j, k);
            const T fu_jp1 = cen2_zfu(d, un, i, j + 1, k), fu_im1jp1 = cen2_zfu(d, un, i - 1, j + 1, k);
            const T fv_i = cen2_zfv(d, vn, i, j, k), fv_ip1 = cen2_zfv(d, vn, i + 1, j, k);
            const T fv_jm1 = cen2_zfv(d, vn, i, j - 1, k), fv_ip1jm1 = cen2_zfv(d, vn, i + 1, j - 1, k);
            const T fv_jp1 = cen2_zfv(d, vn, i, j + 1, k), fv_im1 = cen2_zfv(d, vn, i - 1, j, k);
            const T u_i = A3(un, i, j, k), u_ip1 = A3(un, i + 1, j, k), u_im1 = A3(un, i - 1, j, k);
            const T u_jp1 = A3(un, i, j + 1, k), u_jm1 = A3(un, i, j - 1, k);
            const T v_j = A3(vn, i, j, k), v_ip1 = A3(vn, i + 1, j, k), v_jp1 = A3(vn, i, j + 1, k);
            const T v_im1 = A3(vn, i - 1, j, k), v_jm1 = A3(vn, i, j - 1, k);
            const T zfu_t_ip1 = (fu_i + fu_ip1) * (u_i + u_ip1);
            const T zfu_t_i = (fu_im1 + fu_i) * (u_im1 + u_i);
            const T zfv_f_j = (fv_i + fv_ip1) * (u_i + u_jp1);
            const T zfv_f_jm1 = (fv_jm1 + fv_ip1jm1) * (u_jm1 + u_i);
            const T zfu_f_i = (fu_i + fu_jp1) * (v_j + v_ip1);
            const T zfu_f_im1 = (fu_im1 + fu_im1jp1) * (v_im1 + v_j);
            const T zfv_t_jp1 = (fv_i + fv_jp1) * (v_j + v_jp1);
            const T zfv_t_j = (fv_jm1 + fv_i) * (v_jm1 + v_j);
            const size_t p = IJK(i, j, k);
            T a = ua[p] - (zfu_t_ip1 - zfu_t_i + zfv_f_j - zfv_f_jm1) * r1_e1e2u / E3(m.e3u, ij, k);
            T b = va[p] - (zfu_f_i - zfu_f_im1 + zfv_t_jp1 - zfv_t_j) * r1_e1e2v / E3(m.e3v, ij, k);
            // vertical, dynadv_cen2.F90:113-133
            T zfu_uw_kp1 = T(0), zfv_vw_kp1 = T(0);
            if (k + 1 <= jpk - 1) {
                const int kk = k + 1;
                const T fw = cen2_zfw(d, wn, i, j, kk);
                zfu_uw_kp1 = (fw + cen2_zfw(d, wn, i + 1, j, kk)) * (A3(un, i, j, kk) + A3(un, i, j, kk - 1));
                zfv_vw_kp1 = (fw + cen2_zfw(d, wn, i, j + 1, kk)) * (A3(vn, i, j, kk) + A3(vn, i, j, kk - 1));
            }
            a = a - (zfu_uw_k - zfu_uw_kp1) * r1_e1e2u / E3(m.e3u, ij, k);
            b = b - (zfv_vw_k - zfv_vw_kp1) * r1_e1e2v / E3(m.e3v, ij, k);
            ua[p] = a;
            va[p] = b;
            zfu_uw_k = zfu_uw_kp1;
            zfv_vw_k = zfv_vw_kp1;
        }
    }
};
template <typename T>
KERNEL void k_dyn_adv_cen2(DevDomain d, const T * __restrict__ un, const T * __restrict__ vn, const T * __restrict__ wn, T * __restrict__ ua, T * __restrict__ va)
{
    S_dyn_adv_cen2<T>::run(vb_hw(), d, un, vn, wn, ua, va);
}

template <typename T>
int launch_dyn_adv(int which, int kscheme, const T *un, const T *vn, const T *wn, T *ua, T *va)
{
    // which: 1 keg only, 2 zad only, 3 keg+zad (vector form), 4 cen2
    const DevDomain &d = nh::dom();
    Grid2 g = grid2(d.jpi - 2, d.jpj - 2);
    if (which == 4) {
        LAUNCHB("dyn_adv_cen2", k_dyn_adv_cen2<T>, (S_dyn_adv_cen2<T>), g, d, un, vn, wn, ua, va);
    } else {
        const char *nm = which == 1 ? "dyn_keg" : which == 2 ? "dyn_zad" : "dyn_adv";
        LAUNCHB(nm, k_dyn_keg_zad<T>, (S_dyn_keg_zad<T>), g, d, (which & 1), kscheme, (which & 2) ? 1 : 0, un, vn, wn, ua, va);
    }
    return 0;
}
template int launch_dyn_adv<double>(int, int, const double *, const double *, const double *, double *, double *);
template int launch_dyn_adv<float>(int, int, const float *, const float *, const float *, float *, float *);

// ---------------------------------------------------------------------------
// vorticity term
// ---------------------------------------------------------------------------
template <typename T>
__device__ __forceinline__ T curl_f(const DevDomain &d, const T *pu, const T *pv, int p, int q, int k)
{
    const Metrics<T> &m = met<T>(d);
    return m.e2v[IJ(p + 1, q)] * A3(pv, p + 1, q, k) - m.e2v[IJ(p, q)] * A3(pv, p, q, k) -
           m.e1u[IJ(p, q + 1)] * A3(pu, p, q + 1, k) + m.e1u[IJ(p, q)] * A3(pu, p, q, k);
}
// f-point vorticity of vor_ene / vor_ens / vor_eeT (before any lbc), p<=jpim1, q<=jpjm1
template <typename T>
__device__ __forceinline__ T fvor(const DevDomain &d, int kvor, const T *pu, const T *pv, int p, int q, int k)
{
    const Metrics<T> &m = met<T>(d);
    const size_t pq = IJ(p, q);
    switch (kvor) {
    case np_COR: return m.ff_f[pq];
    case np_RVO: return curl_f(d, pu, pv, p, q, k) * m.r1_e1e2f[pq];
    case np_MET:
        return (A3(pv, p + 1, q, k) + A3(pv, p, q, k)) * m.di_e2v_2e1e2f[pq] -
               (A3(pu, p, q + 1, k) + A3(pu, p, q, k)) * m.dj_e1u_2e1e2f[pq];
    case np_CRV: return m.ff_f[pq] + curl_f(d, pu, pv, p, q, k) * m.r1_e1e2f[pq];
    default:
        return m.ff_f[pq] + (A3(pv, p + 1, q, k) + A3(pv, p, q, k)) * m.di_e2v_2e1e2f[pq] -
               (A3(pu, p, q + 1, k) + A3(pu, p, q, k)) * m.dj_e1u_2e1e2f[pq];
    }
}
// vor_een's zwz: vorticity times 1/e3f, DYN/dynvor.F90:580-643
template <typename T>
__device__ __forceinline__ T fvor_een(const DevDomain &d, int kvor, const T *pu, const T *pv, int p, int q, int k)
{
    const Metrics<T> &m = met<T>(d);
    const size_t pq = IJ(p, q);
    const T t1 = TMASK(T, p, q + 1, k), t2 = TMASK(T, p + 1, q + 1, k), t3 = TMASK(T, p, q, k), t4 = TMASK(T, p + 1, q, k);
    const T ze3f = E3(m.e3t, IJ(p, q + 1), k) * t1 + E3(m.e3t, IJ(p + 1, q + 1), k) * t2 + E3(m.e3t, pq, k) * t3 +
                   E3(m.e3t, IJ(p + 1, q), k) * t4;
    T z1;
    if (ze3f != T(0)) z1 = (d.nn_een_e3f == 0 ? T(4) : (t1 + t2 + t3 + t4)) / ze3f;
    else z1 = T(0);
    switch (kvor) {
    case np_COR: return m.ff_f[pq] * z1;
    case np_RVO: return curl_f(d, pu, pv, p, q, k) * m.r1_e1e2f[pq] * z1;
    case np_MET:
        return ((A3(pv, p + 1, q, k) + A3(pv, p, q, k)) * m.di_e2v_2e1e2f[pq] -
                (A3(pu, p, q + 1, k) + A3(pu, p, q, k)) * m.dj_e1u_2e1e2f[pq]) * z1;
    case np_CRV: return (m.ff_f[pq] + curl_f(d, pu, pv, p, q, k) * m.r1_e1e2f[pq]) * z1;
    default:
        return (m.ff_f[pq] + (A3(pv, p + 1, q, k) + A3(pv, p, q, k)) * m.di_e2v_2e1e2f[pq] -
                (A3(pu, p, q + 1, k) + A3(pu, p, q, k)) * m.dj_e1u_2e1e2f[pq]) * z1;
    }
}
// zwz after lbc_lnk('F') for the EEN/EET/ENT schemes
template <typename T>
__device__ __forceinline__ T zwz_f(const DevDomain &d, int scheme, int kvor, const T *pu, const T *pv, int p, int q, int k)
{
    int sp, sq;
    if (!lbc_src(d, 'F', p, q, sp, sq)) return T(0);
    if (scheme == NEMO_NP_EEN) return fvor_een(d, kvor, pu, pv, sp, sq, k);
    if (scheme == NEMO_NP_ENT) return curl_f(d, pu, pv, sp, sq, k) * met<T>(d).r1_e1e2f[IJ(sp, sq)];
    return fvor(d, kvor, pu, pv, sp, sq, k);
}

template <typename T>
__device__ __forceinline__ void vor_ene_pt(const DevDomain &d, int kvor, const T *pu, const T *pv, int i, int j, int k,
                                           T &a, T &b)
{
    const Metrics<T> &m = met<T>(d);
    const T zy1 = m.e1v[IJ(i, j - 1)] * A3(pv, i, j - 1, k) + m.e1v[IJ(i + 1, j - 1)] * A3(pv, i + 1, j - 1, k);
    const T zy2 = m.e1v[IJ(i, j)] * A3(pv, i, j, k) + m.e1v[IJ(i + 1, j)] * A3(pv, i + 1, j, k);
    const T zx1 = m.e2u[IJ(i - 1, j)] * A3(pu, i - 1, j, k) + m.e2u[IJ(i - 1, j + 1)] * A3(pu, i - 1, j + 1, k);
    const T zx2 = m.e2u[IJ(i, j)] * A3(pu, i, j, k) + m.e2u[IJ(i, j + 1)] * A3(pu, i, j + 1, k);
    const T z_s = fvor(d, kvor, pu, pv, i, j - 1, k), z_c = fvor(d, kvor, pu, pv, i, j, k),
            z_w = fvor(d, kvor, pu, pv, i - 1, j, k);
    const T r1_4 = T(0.25);
    a = a + r1_4 * m.r1_e1u[IJ(i, j)] * (z_s * zy1 + z_c * zy2);
    b = b - r1_4 * m.r1_e2v[IJ(i, j)] * (z_w * zx1 + z_c * zx2);
}
template <typename T>
__device__ __forceinline__ void vor_ens_pt(const DevDomain &d, int kvor, const T *pu, const T *pv, int i, int j, int k,
                                           T &a, T &b)
{
    const Metrics<T> &m = met<T>(d);
    const T r1_8 = T(0.125);
    const T zuav = r1_8 * m.r1_e1u[IJ(i, j)] *
                   (m.e1v[IJ(i, j - 1)] * A3(pv, i, j - 1, k) + m.e1v[IJ(i + 1, j - 1)] * A3(pv, i + 1, j - 1, k) +
                    m.e1v[IJ(i, j)] * A3(pv, i, j, k) + m.e1v[IJ(i + 1, j)] * A3(pv, i + 1, j, k));
    const T zvau = -(r1_8 * m.r1_e2v[IJ(i, j)] *
                     (m.e2u[IJ(i - 1, j)] * A3(pu, i - 1, j, k) + m.e2u[IJ(i - 1, j + 1)] * A3(pu, i - 1, j + 1, k) +
                      m.e2u[IJ(i, j)] * A3(pu, i, j, k) + m.e2u[IJ(i, j + 1)] * A3(pu, i, j + 1, k)));
    const T z_s = fvor(d, kvor, pu, pv, i, j - 1, k), z_c = fvor(d, kvor, pu, pv, i, j, k),
            z_w = fvor(d, kvor, pu, pv, i - 1, j, k);
    a = a + zuav * (z_s + z_c);
    b = b + zvau * (z_w + z_c);
}
// EEN (use_e3f) and EET: triads of f-point vorticity around t-points
template <typename T>
__device__ __forceinline__ void vor_een_pt(const DevDomain &d, int scheme, int kvor, const T *pu, const T *pv, int i,
                                           int j, int k, T &a, T &b)
{
    const Metrics<T> &m = met<T>(d);
    // zwz(i-1..i+1, j-1..j+1) except (i+1,j+1)
    T z[3][3];
    for (int dj = -1; dj <= 1; dj++)
        for (int di = -1; di <= 1; di++)
            z[dj + 1][di + 1] = (di == 1 && dj == 1) ? T(0) : zwz_f(d, scheme, kvor, pu, pv, i + di, j + dj, k);
#define ZW(di, dj) z[(dj) + 1][(di) + 1]
    // triads at t-point (i+oi, j+oj)
    auto tne = [&](int oi, int oj) { return ZW(oi - 1, oj) + ZW(oi, oj) + ZW(oi, oj - 1); };
    auto tnw = [&](int oi, int oj) { return ZW(oi - 1, oj - 1) + ZW(oi - 1, oj) + ZW(oi, oj); };
    auto tse = [&](int oi, int oj) { return ZW(oi, oj) + ZW(oi, oj - 1) + ZW(oi - 1, oj - 1); };
    auto tsw = [&](int oi, int oj) { return ZW(oi, oj - 1) + ZW(oi - 1, oj - 1) + ZW(oi - 1, oj); };
    T ne00 = tne(0, 0), nw10 = tnw(1, 0), se00 = tse(0, 0), sw10 = tsw(1, 0);
    T sw01 = tsw(0, 1), se01 = tse(0, 1), nw00 = tnw(0, 0);
    if (scheme == NEMO_NP_EET) {
        const T z00 = T(1) / E3(m.e3t, IJ(i, j), k), z10 = T(1) / E3(m.e3t, IJ(i + 1, j), k),
                z01 = T(1) / E3(m.e3t, IJ(i, j + 1), k);
        ne00 = ne00 * z00; nw10 = nw10 * z10; se00 = se00 * z00; sw10 = sw10 * z10;
        sw01 = sw01 * z01; se01 = se01 * z01; nw00 = nw00 * z00;
    }
#undef ZW
    auto zwx = [&](int p, int q) { return m.e2u[IJ(p, q)] * E3(m.e3u, IJ(p, q), k) * A3(pu, p, q, k); };
    auto zwy = [&](int p, int q) { return m.e1v[IJ(p, q)] * E3(m.e3v, IJ(p, q), k) * A3(pv, p, q, k); };
    const T r1_12 = T(1) / T(12);
    const T zua = r1_12 * m.r1_e1u[IJ(i, j)] *
                  (ne00 * zwy(i, j) + nw10 * zwy(i + 1, j) + se00 * zwy(i, j - 1) + sw10 * zwy(i + 1, j - 1));
    const T zva = -(r1_12 * m.r1_e2v[IJ(i, j)] *
                    (sw01 * zwx(i - 1, j + 1) + se01 * zwx(i, j + 1) + nw00 * zwx(i - 1, j) + ne00 * zwx(i, j)));
    a = a + zua;
    b = b + zva;
}
// ENT: t-point vorticity times volume, DYN/dynvor.F90:225-317
template <typename T>
__device__ __forceinline__ T ent_zwt(const DevDomain &d, int kvor, const T *pu, const T *pv, int p, int q, int k)
{
    const Metrics<T> &m = met<T>(d);
    const size_t pq = IJ(p, q);
    const T e3t = E3(m.e3t, pq, k);
    const T r1_4 = T(0.25);
    switch (kvor) {
    case np_COR: return m.ff_t[pq] * m.e1e2t[pq] * e3t;
    case np_RVO: {
        const T s = zwz_f(d, NEMO_NP_ENT, kvor, pu, pv, p - 1, q, k) + zwz_f(d, NEMO_NP_ENT, kvor, pu, pv, p, q, k) +
                    zwz_f(d, NEMO_NP_ENT, kvor, pu, pv, p - 1, q - 1, k) + zwz_f(d, NEMO_NP_ENT, kvor, pu, pv, p, q - 1, k);
        return r1_4 * s * m.e1e2t[pq] * e3t;
    }
    case np_MET:
        return ((A3(pv, p, q, k) + A3(pv, p, q - 1, k)) * m.di_e2u_2[pq] -
                (A3(pu, p, q, k) + A3(pu, p - 1, q, k)) * m.dj_e1v_2[pq]) * e3t;
    case np_CRV: {
        const T s = zwz_f(d, NEMO_NP_ENT, kvor, pu, pv, p - 1, q, k) + zwz_f(d, NEMO_NP_ENT, kvor, pu, pv, p, q, k) +
                    zwz_f(d, NEMO_NP_ENT, kvor, pu, pv, p - 1, q - 1, k) + zwz_f(d, NEMO_NP_ENT, kvor, pu, pv, p, q - 1, k);
        return (m.ff_t[pq] + r1_4 * s) * m.e1e2t[pq] * e3t;
    }
    default:
        return (m.ff_t[pq] * m.e1e2t[pq] + (A3(pv, p, q, k) + A3(pv, p, q - 1, k)) * m.di_e2u_2[pq] -
                (A3(pu, p, q, k) + A3(pu, p - 1, q, k)) * m.dj_e1v_2[pq]) * e3t;
    }
}
template <typename T>
__device__ __forceinline__ void vor_ent_pt(const DevDomain &d, int kvor, const T *pu, const T *pv, int i, int j, int k,
                                           T &a, T &b)
{
    const Metrics<T> &m = met<T>(d);
    const T r1_4 = T(0.25);
    const T zc = ent_zwt(d, kvor, pu, pv, i, j, k), ze = ent_zwt(d, kvor, pu, pv, i + 1, j, k),
            zn = ent_zwt(d, kvor, pu, pv, i, j + 1, k);
    a = a + r1_4 * m.r1_e1e2u[IJ(i, j)] / E3(m.e3u, IJ(i, j), k) *
                (ze * (A3(pv, i + 1, j, k) + A3(pv, i + 1, j - 1, k)) + zc * (A3(pv, i, j, k) + A3(pv, i, j - 1, k)));
    b = b - r1_4 * m.r1_e1e2v[IJ(i, j)] / E3(m.e3v, IJ(i, j), k) *
                (zn * (A3(pu, i, j + 1, k) + A3(pu, i - 1, j + 1, k)) + zc * (A3(pu, i, j, k) + A3(pu, i - 1, j, k)));
}

// The scheme and the vorticity kinds are template parameters: each
// instantiation carries only its own stencil, which keeps register use (and so
// occupancy) at what that scheme needs.
template <typename T, int SCHEME, int NCOR, int NRVM, int NTOT>
struct S_dyn_vor {
    static constexpr int NTHR = NH_BLOCK, WAVES = 1;
    __device__ __forceinline__ static void run(const VB &vb_, const DevDomain &d, const T * __restrict__ un, const T * __restrict__ vn, T * __restrict__ ua, T * __restrict__ va)
    {
        THREAD_IJ_V(vb_, 2, 2);
        if (i > d.jpi - 1 || j > d.jpj - 1) return;
        for (int k = 1; k <= d.jpk - 1; k++) {
            const size_t p = IJK(i, j, k);
            T a = ua[p], b = va[p];
            switch (SCHEME) {
            case NEMO_NP_ENE: vor_ene_pt(d, NTOT, un, vn, i, j, k, a, b); break;
            case NEMO_NP_ENS: vor_ens_pt(d, NTOT, un, vn, i, j, k, a, b); break;
            case NEMO_NP_MIX:
                vor_ens_pt(d, NRVM, un, vn, i, j, k, a, b);
                vor_ene_pt(d, NCOR, un, vn, i, j, k, a, b);
                break;
            case NEMO_NP_ENT: vor_ent_pt(d, NTOT, un, vn, i, j, k, a, b); break;
            default: vor_een_pt(d, SCHEME, NTOT, un, vn, i, j, k, a, b); break;
            }
            ua[p] = a;
            va[p] = b;
        }
    }
};
template <typename T, int SCHEME, int NCOR, int NRVM, int NTOT>
KERNEL void k_dyn_vor(DevDomain d, const T * __restrict__ un, const T * __restrict__ vn, T * __restrict__ ua,
                      T * __restrict__ va)
{
    S_dyn_vor<T, SCHEME, NCOR, NRVM, NTOT>::run(vb_hw(), d, un, vn, ua, va);
}

template <typename T, int SCHEME>
static int launch_dyn_vor_s(const DevDomain &d, const T *un, const T *vn, T *ua, T *va)
{
    Grid2 g = grid2(d.jpi - 2, d.jpj - 2);
    // (nrvm, ntot) from the momentum advection form, DYN/dynvor.F90:903-939
    switch (d.n_dynadv) {
    case NEMO_NP_LIN_DYN:
        LAUNCHB("dyn_vor", (k_dyn_vor<T, SCHEME, np_COR, np_COR, np_COR>), (S_dyn_vor<T, SCHEME, np_COR, np_COR, np_COR>), g, d, un, vn, ua, va);
        break;
    case NEMO_NP_VEC_C2:
        LAUNCHB("dyn_vor", (k_dyn_vor<T, SCHEME, np_COR, np_RVO, np_CRV>), (S_dyn_vor<T, SCHEME, np_COR, np_RVO, np_CRV>), g, d, un, vn, ua, va);
        break;
    default:
        LAUNCHB("dyn_vor", (k_dyn_vor<T, SCHEME, np_COR, np_MET, np_CME>), (S_dyn_vor<T, SCHEME, np_COR, np_MET, np_CME>), g, d, un, vn, ua, va);
        break;
    }
    return 0;
}

template <typename T>
int launch_dyn_vor(const T *un, const T *vn, T *ua, T *va)
{
    const DevDomain &d = nh::dom();
    switch (d.nvor_scheme) {
    case NEMO_NP_ENE: return launch_dyn_vor_s<T, NEMO_NP_ENE>(d, un, vn, ua, va);
    case NEMO_NP_ENS: return launch_dyn_vor_s<T, NEMO_NP_ENS>(d, un, vn, ua, va);
    case NEMO_NP_MIX: return launch_dyn_vor_s<T, NEMO_NP_MIX>(d, un, vn, ua, va);
    case NEMO_NP_ENT: return launch_dyn_vor_s<T, NEMO_NP_ENT>(d, un, vn, ua, va);
    case NEMO_NP_EET: return launch_dyn_vor_s<T, NEMO_NP_EET>(d, un, vn, ua, va);
    case NEMO_NP_EEN: return launch_dyn_vor_s<T, NEMO_NP_EEN>(d, un, vn, ua, va);
    default: return nh::set_error(NEMO_HIP_ERR_UNSUPPORTED, "dyn_vor: nvor_scheme = %d", d.nvor_scheme);
    }
}
template int launch_dyn_vor<double>(const double *, const double *, double *, double *);
template int launch_dyn_vor<float>(const float *, const float *, float *, float *);

// ---------------------------------------------------------------------------
// dyn_zdf, DYN/dynzdf.F90:107-485 (ln_zad_Aimp = F, iso-level momentum
// mixing, r_vvl = 0).  All columns apply the leapfrog right-hand side; the
// interior ones then solve the u and v columns (FP64 elimination).
// ---------------------------------------------------------------------------
// `pre(k, x)` maps the stored trend x at level k to the value the solve
// starts from (the identity for dyn_zdf alone; the barotropic corrections
// of dyn_spg_ts in the fused step).
struct PreIdentity {
    template <typename T>
    __device__ __forceinline__ T operator()(int, T x) const { return x; }
};
template <typename T, typename Pre = PreIdentity>
__device__ __forceinline__ void zdf_column(const DevDomain &d, double r2dt, int comp, int i, int j,
                                           const T *__restrict__ pb, T *__restrict__ pa, const T *__restrict__ avm,
                                           const double *__restrict__ rCdU_bot, const double *__restrict__ pa_b,
                                           const double *__restrict__ tau, const double *__restrict__ tau_b,
                                           double *__restrict__ zwd_s, Pre pre = Pre())
{
    const Metrics<T> &m = met<T>(d);
    const int jpk = d.jpk;
    const size_t plane = d.n2;
    const uint32_t ij = IJ32(i, j);
    const uint32_t ijn = comp ? IJ32(i, j + 1) : IJ32(i + 1, j);     // the other t-point of the u/v face
    const T *e3x = comp ? m.e3v : m.e3u;
    const T *e3xw = comp ? m.e3vw : m.e3uw;
    const KRange &kr = comp ? d.kV : d.kU;
    const int x0 = kr.k0[ij], x1 = kr.k1[ij];
    auto xmask = [&](int k) { return (k >= x0 && k <= x1) ? 1.0 : 0.0; };
    auto wxmask = [&](int k) { return (k >= x0 && k <= x1 && (k == 1 || k - 1 >= x0)) ? 1.0 : 0.0; };
    const double r_vvl = 0.0;
    const bool inner = is_interior(d, i, j);
    const bool rm_b = d.ln_drgimp && d.ln_dynspg_ts;
    const double ab = rm_b ? pa_b[ij] : 0.0;
    if (!inner) {
        for (int k = 1; k <= jpk - 1; k++) {
            const double xm = xmask(k);
            double a = ((double)ldp(plane_ptr(pb, plane, k), ij) +
                        r2dt * (double)pre(k, ldp((const T *)plane_ptr(pa, plane, k), ij))) * xm;
            if (rm_b) a = (a - ab) * xm;
            stp(plane_ptr(pa, plane, k), ij, (T)a);
        }
        return;
    }
    const int ikb = comp ? d.mbkv[ij] : d.mbku[ij];
    const double csum = rCdU_bot[ijn] + rCdU_bot[ij];
    const double zdt = r2dt * 0.5;
    // off-diagonal at w-level k (k >= 2), lines 208-243
    auto zzw = [&](int k, double ze3) {
        const T *ak = plane_ptr(avm, plane, k);
        return -(zdt * ((double)ldp(ak, ijn) + (double)ldp(ak, ij)) / (ze3 * (double)E3(e3xw, ij, k)) * wxmask(k));
    };
    // zzw from loaded operands: avm at the face's two t-points, e3xw
    auto zzw_v = [&](int k, double ze3, T a_n, T a_c, T e3w) {
        return -(zdt * ((double)a_n + (double)a_c) / (ze3 * (double)e3w) * wxmask(k));
    };
    auto ld_avm = [&](int k, T &a_n, T &a_c, T &e3w) {   // level k <= jpk
        const T *ak = plane_ptr(avm, plane, k);
        a_n = ldp(ak, ijn);
        a_c = ldp(ak, ij);
        e3w = E3(e3xw, ij, k);
    };
    double zws_prev, zwd_prev, a_prev;
    // the operands of level k (pb, pa, e3x) and of w-level k + 1 (avm, e3xw)
    // are loaded one level ahead
    T w1n, w1c, w1e;   // w-level k + 1 of the current level
    ld_avm(2, w1n, w1c, w1e);
    {   // k = 1: surface row with the wind stress
        const double xm = xmask(1);
        double a = ((double)ldp(pb, ij) + r2dt * (double)pre(1, ldp((const T *)pa, ij))) * xm;
        if (rm_b) a = (a - ab) * xm;
        const double ze3 = (1.0 - r_vvl) * (double)E3(e3x, ij, 1) + r_vvl * (double)E3(e3x, ij, 1);
        if (rm_b && 1 == ikb) a = a + r2dt * 0.5 * csum * ab / ze3;
        const double zws = zzw_v(2, ze3, w1n, w1c, w1e);
        double zwd = 1.0 - zws;
        if (d.ln_drgimp && 1 == ikb) zwd = zwd - r2dt * 0.5 * csum / ze3;
        a = a + r2dt * 0.5 * (tau_b[ij] + tau[ij]) / (ze3 * d.rau0) * xm;
        stp(pa, ij, (T)a);
        zws_prev = zws;
        zwd_prev = zwd;
        a_prev = a;
    }
    {
        T b_c = ldp(plane_ptr(pb, plane, 2), ij), t_c = ldp((const T *)plane_ptr(pa, plane, 2), ij);
        T e_c = E3(e3x, ij, 2);
        T w0n = w1n, w0c = w1c, w0e = w1e;   // w-level k
        ld_avm(3 <= jpk ? 3 : jpk, w1n, w1c, w1e);
#pragma unroll 4
        for (int k = 2; k <= jpk - 1; k++) {
            // level k + 1's operands (within the arrays: levels <= jpk)
            const int kn = k + 1, kw = k + 2 <= jpk ? k + 2 : jpk;
            const T b_n = ldp(plane_ptr(pb, plane, kn), ij), t_n = ldp((const T *)plane_ptr(pa, plane, kn), ij);
            const T e_n = E3(e3x, ij, kn);
            T w2n, w2c, w2e;
            ld_avm(kw, w2n, w2c, w2e);
            const double xm = xmask(k);
            double a = ((double)b_c + r2dt * (double)pre(k, t_c)) * xm;
            if (rm_b) a = (a - ab) * xm;
            const double ze3 = (1.0 - r_vvl) * (double)e_c + r_vvl * (double)e_c;
            if (rm_b && k == ikb) a = a + r2dt * 0.5 * csum * ab / ze3;
            const double zwi = zzw_v(k, ze3, w0n, w0c, w0e);
            const double zws = zzw_v(k + 1, ze3, w1n, w1c, w1e);
            double zwd = 1.0 - zwi - zws;
            if (d.ln_drgimp && k == ikb) zwd = zwd - r2dt * 0.5 * csum / ze3;
            zwd = zwd - zwi * zws_prev / zwd_prev;
            a = a - zwi / zwd_prev * a_prev;
            if (k % ZDF_SEG == 0) stp(plane_ptr(zwd_s, plane, k), ij, zwd);   // checkpoint pivots only
            stp(plane_ptr(pa, plane, k), ij, (T)a);
            zws_prev = zws;
            zwd_prev = zwd;
            a_prev = a;
            b_c = b_n; t_c = t_n; e_c = e_n;
            w0n = w1n; w0c = w1c; w0e = w1e;
            w1n = w2n; w1c = w2c; w1e = w2e;
        }
    }
    double x = a_prev / zwd_prev;
    stp(plane_ptr(pa, plane, jpk - 1), ij, (T)x);
    // levels jpk-2 .. 1 by segments of ZDF_SEG levels, bottom segment first:
    // the forward sweep kept the pivot of every ZDF_SEG-th level only, so a
    // segment first recomputes its pivots from the one above it (the
    // operations of the forward sweep), then substitutes back.  Each
    // segment's operands are loaded together, from levels within the arrays.
    const int kl = jpk - 2;
    for (int s = (kl + ZDF_SEG - 1) / ZDF_SEG - 1; s >= 0; s--) {
        const int k0 = s * ZDF_SEG + 1;
        T e[ZDF_SEG], t[ZDF_SEG], wn[ZDF_SEG], wc[ZDF_SEG], we[ZDF_SEG];   // level k, w-level k + 1
#pragma unroll
        for (int q = 0; q < ZDF_SEG; q++) {
            const int k = k0 + q <= kl ? k0 + q : kl;
            e[q] = E3(e3x, ij, k);
            t[q] = ldp((const T *)plane_ptr(pa, plane, k), ij);
            ld_avm(k + 1, wn[q], wc[q], we[q]);
        }
        double zwd_p = 0.0, zws_p = 0.0;
        T an = T(0), ac = T(0), aw = T(0);   // w-level k0
        if (s > 0) {
            zwd_p = ldp((const double *)plane_ptr(zwd_s, plane, k0 - 1), ij);
            const T ep = E3(e3x, ij, k0 - 1);
            ld_avm(k0, an, ac, aw);
            const double ze3p = (1.0 - r_vvl) * (double)ep + r_vvl * (double)ep;
            zws_p = zzw_v(k0, ze3p, an, ac, aw);
        }
        double piv[ZDF_SEG], zsv[ZDF_SEG];
#pragma unroll
        for (int q = 0; q < ZDF_SEG; q++) {
            const int k = k0 + q;
            const double ze3 = (1.0 - r_vvl) * (double)e[q] + r_vvl * (double)e[q];
            const double zws = zzw_v(k + 1, ze3, wn[q], wc[q], we[q]);
            double zwd;
            if (k == 1) {
                zwd = 1.0 - zws;
                if (d.ln_drgimp && 1 == ikb) zwd = zwd - r2dt * 0.5 * csum / ze3;
            } else {
                const double zwi = q == 0 ? zzw_v(k, ze3, an, ac, aw) : zzw_v(k, ze3, wn[q - 1], wc[q - 1], we[q - 1]);
                zwd = 1.0 - zwi - zws;
                if (d.ln_drgimp && k == ikb) zwd = zwd - r2dt * 0.5 * csum / ze3;
                zwd = zwd - zwi * zws_p / zwd_p;
            }
            piv[q] = zwd;
            zsv[q] = zws;
            zws_p = zws;
            zwd_p = zwd;
        }
#pragma unroll
        for (int q = ZDF_SEG - 1; q >= 0; q--) {
            const int k = k0 + q;
            if (k <= kl) {
                x = ((double)t[q] - zsv[q] * x) / piv[q];
                stp(plane_ptr(pa, plane, k), ij, (T)x);
            }
        }
    }
}

// The same column solve in the FP32 storage form with jpk <= KMAX: the
// forward-substituted values (rounded to storage, as zdf_column stores
// them) stay in registers instead of passing through pa, and the FP64
// pivots are kept at every ZREG_SEG-th level, the others recomputed segment
// by segment in the back substitution from a second read of avm.  Every
// loop over levels is unrolled to KMAX so the register array is indexed by
// constants; the level count only guards.  Results equal zdf_column's bit
// for bit.
#define ZREG_SEG 8
template <int KMAX, typename Pre = PreIdentity>
__device__ __forceinline__ void zdf_column_reg(const DevDomain &d, double r2dt, int comp, int i, int j,
                                               const float *__restrict__ pb, float *__restrict__ pa,
                                               const float *__restrict__ avm, const double *__restrict__ rCdU_bot,
                                               const double *__restrict__ pa_b, const double *__restrict__ tau,
                                               const double *__restrict__ tau_b, Pre pre = Pre())
{
    const Metrics<float> &m = met<float>(d);
    const int jpk = d.jpk;
    const size_t plane = d.n2;
    const uint32_t ij = IJ32(i, j);
    const uint32_t ijn = comp ? IJ32(i, j + 1) : IJ32(i + 1, j);
    const float *e3x = comp ? m.e3v : m.e3u;
    const float *e3xw = comp ? m.e3vw : m.e3uw;
    const KRange &kr = comp ? d.kV : d.kU;
    const int x0 = kr.k0[ij], x1 = kr.k1[ij];
    auto xmask = [&](int k) { return (k >= x0 && k <= x1) ? 1.0 : 0.0; };
    auto wxmask = [&](int k) { return (k >= x0 && k <= x1 && (k == 1 || k - 1 >= x0)) ? 1.0 : 0.0; };
    const double r_vvl = 0.0;
    const bool rm_b = d.ln_drgimp && d.ln_dynspg_ts;
    const double ab = rm_b ? pa_b[ij] : 0.0;
    if (!is_interior(d, i, j)) {
        for (int k = 1; k <= jpk - 1; k++) {
            const double xm = xmask(k);
            double a = ((double)ldp(plane_ptr(pb, plane, k), ij) +
                        r2dt * (double)pre(k, ldp((const float *)plane_ptr(pa, plane, k), ij))) * xm;
            if (rm_b) a = (a - ab) * xm;
            stp(plane_ptr(pa, plane, k), ij, (float)a);
        }
        return;
    }
    const int ikb = comp ? d.mbkv[ij] : d.mbku[ij];
    const double csum = rCdU_bot[ijn] + rCdU_bot[ij];
    const double zdt = r2dt * 0.5;
    auto zzw = [&](int k, double ze3) {
        const float *ak = plane_ptr(avm, plane, k);
        return -(zdt * ((double)ldp(ak, ijn) + (double)ldp(ak, ij)) / (ze3 * (double)E3(e3xw, ij, k)) * wxmask(k));
    };
    auto ze3_of = [&](int k) { return (1.0 - r_vvl) * (double)E3(e3x, ij, k) + r_vvl * (double)E3(e3x, ij, k); };
    // the eliminated diagonal of level k >= 2 from the previous level's
    // upper diagonal and pivot (lines 208-243 and the forward elimination)
    auto pivot = [&](int k, double ze3, double zws, double zws_p, double zwd_p) {
        const double zwi = zzw(k, ze3);
        double zwd = 1.0 - zwi - zws;
        if (d.ln_drgimp && k == ikb) zwd = zwd - r2dt * 0.5 * csum / ze3;
        return zwd - zwi * zws_p / zwd_p;
    };
    constexpr int NSEG = (KMAX + ZREG_SEG - 1) / ZREG_SEG;
    float fw[KMAX];        // level k at index k - 1
    double ck[NSEG];       // the pivot of level s * ZREG_SEG at index s
    double zws_prev, zwd_prev, a_prev;
    {   // k = 1: surface row with the wind stress
        const double xm = xmask(1);
        double a = ((double)ldp(pb, ij) + r2dt * (double)pre(1, ldp((const float *)pa, ij))) * xm;
        if (rm_b) a = (a - ab) * xm;
        const double ze3 = ze3_of(1);
        if (rm_b && 1 == ikb) a = a + r2dt * 0.5 * csum * ab / ze3;
        const double zws = zzw(2, ze3);
        double zwd = 1.0 - zws;
        if (d.ln_drgimp && 1 == ikb) zwd = zwd - r2dt * 0.5 * csum / ze3;
        a = a + r2dt * 0.5 * (tau_b[ij] + tau[ij]) / (ze3 * d.rau0) * xm;
        fw[0] = (float)a;
        zws_prev = zws;
        zwd_prev = zwd;
        a_prev = a;
    }
#pragma unroll
    for (int k = 2; k < KMAX; k++) {
        if (k <= jpk - 1) {
            const double xm = xmask(k);
            double a = ((double)ldp(plane_ptr(pb, plane, k), ij) +
                        r2dt * (double)pre(k, ldp((const float *)plane_ptr(pa, plane, k), ij))) * xm;
            if (rm_b) a = (a - ab) * xm;
            const double ze3 = ze3_of(k);
            if (rm_b && k == ikb) a = a + r2dt * 0.5 * csum * ab / ze3;
            const double zwi = zzw(k, ze3);
            const double zws = zzw(k + 1, ze3);
            const double zwd = pivot(k, ze3, zws, zws_prev, zwd_prev);
            a = a - zwi / zwd_prev * a_prev;
            if (k % ZREG_SEG == 0) ck[k / ZREG_SEG] = zwd;
            fw[k - 1] = (float)a;
            zws_prev = zws;
            zwd_prev = zwd;
            a_prev = a;
        }
    }
    double x = a_prev / zwd_prev;
    stp(plane_ptr(pa, plane, jpk - 1), ij, (float)x);
    const int kl = jpk - 2;
#pragma unroll
    for (int s = NSEG - 1; s >= 0; s--) {
        const int k0 = s * ZREG_SEG + 1;
        if (k0 <= kl) {
            double piv[ZREG_SEG], zsv[ZREG_SEG];
            double zwd_p = 0.0, zws_p = 0.0;
            if (s > 0) {
                zwd_p = ck[s];
                zws_p = zzw(k0, ze3_of(k0 - 1));
            }
#pragma unroll
            for (int q = 0; q < ZREG_SEG; q++) {
                const int k = k0 + q;
                if (k < KMAX && k <= kl) {
                    const double ze3 = ze3_of(k);
                    const double zws = zzw(k + 1, ze3);
                    double zwd;
                    if (k == 1) {
                        zwd = 1.0 - zws;
                        if (d.ln_drgimp && 1 == ikb) zwd = zwd - r2dt * 0.5 * csum / ze3;
                    } else {
                        zwd = pivot(k, ze3, zws, zws_p, zwd_p);
                    }
                    piv[q] = zwd;
                    zsv[q] = zws;
                    zws_p = zws;
                    zwd_p = zwd;
                }
            }
#pragma unroll
            for (int q = ZREG_SEG - 1; q >= 0; q--) {
                const int k = k0 + q;
                if (k < KMAX && k <= kl) {
                    x = ((double)fw[k - 1] - zsv[q] * x) / piv[q];
                    stp(plane_ptr(pa, plane, k), ij, (float)x);
                }
            }
        }
    }
}
// the levels zdf_column_reg holds in registers (jpk <= ZDF_REG_KMAX): the
// 75 of the full-size configurations
#define ZDF_REG_KMAX 76

template <typename T, bool REG>
struct S_dyn_zdf {
    static constexpr int NTHR = NH_BLOCK, WAVES = 2;
    __device__ __forceinline__ static void run(const VB &vb_, const DevDomain &d, double r2dt, const T * __restrict__ ub, const T * __restrict__ vb, T * __restrict__ ua, T * __restrict__ va, const T * __restrict__ avm, const double * __restrict__ rCdU_bot, const double * __restrict__ ua_b, const double * __restrict__ va_b, const double * __restrict__ utau, const double * __restrict__ utau_b, const double * __restrict__ vtau, const double * __restrict__ vtau_b, double * __restrict__ zwd_s)
    {
        THREAD_IJ_V(vb_, 1, 1);
        if (i > d.jpi || j > d.jpj) return;
        if constexpr (REG) {
            zdf_column_reg<ZDF_REG_KMAX>(d, r2dt, 0, i, j, ub, ua, avm, rCdU_bot, ua_b, utau, utau_b);
            zdf_column_reg<ZDF_REG_KMAX>(d, r2dt, 1, i, j, vb, va, avm, rCdU_bot, va_b, vtau, vtau_b);
        } else {
            zdf_column(d, r2dt, 0, i, j, ub, ua, avm, rCdU_bot, ua_b, utau, utau_b, zwd_s);
            zdf_column(d, r2dt, 1, i, j, vb, va, avm, rCdU_bot, va_b, vtau, vtau_b, zwd_s);
        }
    }
};
template <typename T, bool REG>
KERNEL_W(2) void k_dyn_zdf(DevDomain d, double r2dt, const T * __restrict__ ub, const T * __restrict__ vb, T * __restrict__ ua, T * __restrict__ va, const T * __restrict__ avm,
                          const double * __restrict__ rCdU_bot, const double * __restrict__ ua_b, const double * __restrict__ va_b, const double * __restrict__ utau,
                          const double * __restrict__ utau_b, const double * __restrict__ vtau, const double * __restrict__ vtau_b, double * __restrict__ zwd_s)
{
    S_dyn_zdf<T, REG>::run(vb_hw(), d, r2dt, ub, vb, ua, va, avm, rCdU_bot, ua_b, va_b, utau, utau_b, vtau, vtau_b, zwd_s);
}
// the register form applies: FP32 storage, jpk <= ZDF_REG_KMAX, switch zdf_reg
template <typename T>
static bool zdf_reg(const DevDomain &d)
{
    return sizeof(T) == 4 && d.jpk <= ZDF_REG_KMAX && nh::tune(nh::T_ZDF_REG);
}

template <typename T>
int launch_dyn_zdf(double r2dt, const T *ub, const T *vb, T *ua, T *va, const T *avm, const double *rCdU_bot,
                   const double *ua_b, const double *va_b, const double *utau, const double *utau_b,
                   const double *vtau, const double *vtau_b)
{
    const DevDomain &d = nh::dom();
    Grid2 g = grid2(d.jpi, d.jpj);
    double *zwd = (double *)nh::scratch3(nh::S_ZWT);
    // forward-substituted values stay in ua/va at storage precision
    if constexpr (sizeof(T) == 4) {
        if (zdf_reg<T>(d)) {
            LAUNCHB("dyn_zdf", (k_dyn_zdf<T, true>), (S_dyn_zdf<T, true>), g, d, r2dt, ub, vb, ua, va, avm, rCdU_bot, ua_b, va_b, utau,
                   utau_b, vtau, vtau_b, zwd);
            return 0;
        }
    }
    LAUNCHB("dyn_zdf", (k_dyn_zdf<T, false>), (S_dyn_zdf<T, false>), g, d, r2dt, ub, vb, ua, va, avm, rCdU_bot, ua_b, va_b, utau, utau_b,
           vtau, vtau_b, zwd);
    return 0;
}
template int launch_dyn_zdf<double>(double, const double *, const double *, double *, double *, const double *,
                                    const double *, const double *, const double *, const double *, const double *,
                                    const double *, const double *);
template int launch_dyn_zdf<float>(double, const float *, const float *, float *, float *, const float *,
                                   const double *, const double *, const double *, const double *, const double *,
                                   const double *, const double *);

// ---------------------------------------------------------------------------
// Fused momentum step: ua = va = 0; dyn_adv; dyn_vor; dyn_spg_ts; dyn_zdf
// (the order of step.F90:189-257) in two 3D passes around the barotropic
// sub-steps:
//   k_step_dyn1  advection + vorticity trends written once, with the vertical
//                sums of dyn_spg_ts phase 1 (DYN/dynspg_ts.F90:358-378) and the
//                bottom velocities (519-527) accumulated in the same k-march;
//   k_step_dyn2  phase 3 (lines 1214-1269) applied to un/vn and to the trend
//                on the fly, then the dyn_zdf solve of the column.
// Every stored intermediate of the separate calls is rounded to the storage
// type at the same point here, so the fused step equals the sequence of
// separate calls bit for bit in either storage form.
// ---------------------------------------------------------------------------
template <typename T, int ADV, int KEG, int SCHEME, int NCOR, int NRVM, int NTOT>
struct S_step_dyn1 {
    static constexpr int NTHR = NH_BLOCK, WAVES = 1;
    __device__ __forceinline__ static void run(const VB &vb_, const DevDomain &d, const T *__restrict__ un, const T *__restrict__ vn, const T *__restrict__ wn, T *__restrict__ ua, T *__restrict__ va, const double *__restrict__ r1_hu_n, const double *__restrict__ r1_hv_n, const double *__restrict__ un_b, const double *__restrict__ vn_b, double *__restrict__ zu_frc, double *__restrict__ zv_frc, double *__restrict__ su_s, double *__restrict__ sv_s, double *__restrict__ zwxb, double *__restrict__ zwyb)
    {
        THREAD_IJ_V(vb_, 1, 1);
        if (i > d.jpi || j > d.jpj) return;
        const size_t ij = IJ(i, j);
        const int jpk = d.jpk;
        if (!is_interior(d, i, j)) {
            for (int k = 1; k <= jpk; k++) {
                A3(ua, i, j, k) = T(0);
                A3(va, i, j, k) = T(0);
            }
            double su = 0.0, sv = 0.0;
            su = su * r1_hu_n[ij];
            sv = sv * r1_hv_n[ij];
            zu_frc[ij] = su; su_s[ij] = su;
            zv_frc[ij] = sv; sv_s[ij] = sv;
            return;
        }
        const Metrics<T> &m = met<T>(d);
        const T e1u = m.e1u[ij], e2v = m.e2v[ij], r1_e1e2u = m.r1_e1e2u[ij], r1_e1e2v = m.r1_e1e2v[ij];
        const T e1e2t = m.e1e2t[ij], e1e2t_e = m.e1e2t[IJ(i + 1, j)], e1e2t_n = m.e1e2t[IJ(i, j + 1)];
        // vertical flux carried between levels: zad (dynzad.F90:93-94) or cen2 surface flux (dynadv_cen2.F90:105-112)
        T cu_k = T(0), cv_k = T(0);
        if (ADV == NEMO_NP_FLX_C2) {
            cu_k = T(0.5) * (e1e2t * A3(wn, i, j, 1) + e1e2t_e * A3(wn, i + 1, j, 1)) * A3(un, i, j, 1);
            cv_k = T(0.5) * (e1e2t * A3(wn, i, j, 1) + e1e2t_n * A3(wn, i, j + 1, 1)) * A3(vn, i, j, 1);
        }
        double su = 0.0, sv = 0.0;
        for (int k = 1; k <= jpk - 1; k++) {
            T a = T(0), b = T(0);
            if (ADV == NEMO_NP_VEC_C2) {
                T h, he, hn;
                if (KEG == 0) {
                    h = keg_c2(d, un, vn, i, j, k); he = keg_c2(d, un, vn, i + 1, j, k); hn = keg_c2(d, un, vn, i, j + 1, k);
                } else {
                    h = keg_hw(d, un, vn, i, j, k); he = keg_hw(d, un, vn, i + 1, j, k); hn = keg_hw(d, un, vn, i, j + 1, k);
                }
                a = a - (he - h) / e1u;
                b = b - (hn - h) / e2v;
                T cu_kp1 = T(0), cv_kp1 = T(0);
                if (k + 1 <= jpk - 1) {
                    const int kk = k + 1;
                    const T zww = T(0.25) * e1e2t * A3(wn, i, j, kk);
                    const T zww_e = T(0.25) * e1e2t_e * A3(wn, i + 1, j, kk);
                    const T zww_n = T(0.25) * e1e2t_n * A3(wn, i, j + 1, kk);
                    cu_kp1 = (zww_e + zww) * (A3(un, i, j, kk - 1) - A3(un, i, j, kk));
                    cv_kp1 = (zww_n + zww) * (A3(vn, i, j, kk - 1) - A3(vn, i, j, kk));
                }
                a = a - (cu_k + cu_kp1) * r1_e1e2u / E3(m.e3u, ij, k);
                b = b - (cv_k + cv_kp1) * r1_e1e2v / E3(m.e3v, ij, k);
                cu_k = cu_kp1;
                cv_k = cv_kp1;
            } else if (ADV == NEMO_NP_FLX_C2) {
                const T fu_im1 = cen2_zfu(d, un, i - 1, j, k), fu_i = cen2_zfu(d, un, i, j, k), fu_ip1 = cen2_zfu(d, un, i + 1, j, k);
                const T fu_jp1 = cen2_zfu(d, un, i, j + 1, k), fu_im1jp1 = cen2_zfu(d, un, i - 1, j + 1, k);
                const T fv_i = cen2_zfv(d, vn, i, j, k), fv_ip1 = cen2_zfv(d, vn, i + 1, j, k);
                const T fv_jm1 = cen2_zfv(d, vn, i, j - 1, k), fv_ip1jm1 = cen2_zfv(d, vn, i + 1, j - 1, k);
                const T fv_jp1 = cen2_zfv(d, vn, i, j + 1, k), fv_im1 = cen2_zfv(d, vn, i - 1, j, k);
                const T u_i = A3(un, i, j, k), u_ip1 = A3(un, i + 1, j, k), u_im1 = A3(un, i - 1, j, k);
                const T u_jp1 = A3(un, i, j + 1, k), u_jm1 = A3(un, i, j - 1, k);
                const T v_j = A3(vn, i, j, k), v_ip1 = A3(vn, i + 1, j, k), v_jp1 = A3(vn, i, j + 1, k);
                const T v_im1 = A3(vn, i - 1, j, k), v_jm1 = A3(vn, i, j - 1, k);
                const T zfu_t_ip1 = (fu_i + fu_ip1) * (u_i + u_ip1);
                const T zfu_t_i = (fu_im1 + fu_i) * (u_im1 + u_i);
                const T zfv_f_j = (fv_i + fv_ip1) * (u_i + u_jp1);
                const T zfv_f_jm1 = (fv_jm1 + fv_ip1jm1) * (u_jm1 + u_i);
                const T zfu_f_i = (fu_i + fu_jp1) * (v_j + v_ip1);
                const T zfu_f_im1 = (fu_im1 + fu_im1jp1) * (v_im1 + v_j);
                const T zfv_t_jp1 = (fv_i + fv_jp1) * (v_j + v_jp1);
                const T zfv_t_j = (fv_jm1 + fv_i) * (v_jm1 + v_j);
                a = a - (zfu_t_ip1 - zfu_t_i + zfv_f_j - zfv_f_jm1) * r1_e1e2u / E3(m.e3u, ij, k);
                b = b - (zfu_f_i - zfu_f_im1 + zfv_t_jp1 - zfv_t_j) * r1_e1e2v / E3(m.e3v, ij, k);
                T cu_kp1 = T(0), cv_kp1 = T(0);
                if (k + 1 <= jpk - 1) {
                    const int kk = k + 1;
                    const T fw = cen2_zfw(d, wn, i, j, kk);
                    cu_kp1 = (fw + cen2_zfw(d, wn, i + 1, j, kk)) * (A3(un, i, j, kk) + A3(un, i, j, kk - 1));
                    cv_kp1 = (fw + cen2_zfw(d, wn, i, j + 1, kk)) * (A3(vn, i, j, kk) + A3(vn, i, j, kk - 1));
                }
                a = a - (cu_k - cu_kp1) * r1_e1e2u / E3(m.e3u, ij, k);
                b = b - (cv_k - cv_kp1) * r1_e1e2v / E3(m.e3v, ij, k);
                cu_k = cu_kp1;
                cv_k = cv_kp1;
            }
            switch (SCHEME) {
            case NEMO_NP_ENE: vor_ene_pt(d, NTOT, un, vn, i, j, k, a, b); break;
            case NEMO_NP_ENS: vor_ens_pt(d, NTOT, un, vn, i, j, k, a, b); break;
            case NEMO_NP_MIX:
                vor_ens_pt(d, NRVM, un, vn, i, j, k, a, b);
                vor_ene_pt(d, NCOR, un, vn, i, j, k, a, b);
                break;
            case NEMO_NP_ENT: vor_ent_pt(d, NTOT, un, vn, i, j, k, a, b); break;
            default: vor_een_pt(d, SCHEME, NTOT, un, vn, i, j, k, a, b); break;
            }
            const size_t p = IJK(i, j, k);
            ua[p] = a;
            va[p] = b;
            su = su + (double)E3(m.e3u, ij, k) * (double)a * UMASK(double, i, j, k);
            sv = sv + (double)E3(m.e3v, ij, k) * (double)b * VMASK(double, i, j, k);
        }
        A3(ua, i, j, jpk) = T(0);
        A3(va, i, j, jpk) = T(0);
        su = su * r1_hu_n[ij];
        sv = sv * r1_hv_n[ij];
        zu_frc[ij] = su; su_s[ij] = su;
        zv_frc[ij] = sv; sv_s[ij] = sv;
        zwxb[ij] = (double)A3(un, i, j, d.mbku[ij]) - un_b[ij];
        zwyb[ij] = (double)A3(vn, i, j, d.mbkv[ij]) - vn_b[ij];
    }
};
template <typename T, int ADV, int KEG, int SCHEME, int NCOR, int NRVM, int NTOT>
KERNEL void k_step_dyn1(DevDomain d, const T *__restrict__ un, const T *__restrict__ vn, const T *__restrict__ wn,
                        T *__restrict__ ua, T *__restrict__ va, const double *__restrict__ r1_hu_n,
                        const double *__restrict__ r1_hv_n, const double *__restrict__ un_b,
                        const double *__restrict__ vn_b, double *__restrict__ zu_frc, double *__restrict__ zv_frc,
                        double *__restrict__ su_s, double *__restrict__ sv_s, double *__restrict__ zwxb,
                        double *__restrict__ zwyb)
{
    S_step_dyn1<T, ADV, KEG, SCHEME, NCOR, NRVM, NTOT>::run(vb_hw(), d, un, vn, wn, ua, va, r1_hu_n, r1_hv_n, un_b, vn_b, zu_frc, zv_frc, su_s, sv_s, zwxb, zwyb);
}

// ua after dyn_spg_ts: interior columns remove the vertical mean of the
// trend (phase 1, rounded to storage), every column adds the barotropic
// correction (phase 3, rounded to storage)
template <typename T>
struct PreSpg {
    double s, dd, xm0;
    int x0, x1;
    bool inner;
    __device__ __forceinline__ T operator()(int k, T x) const
    {
        const double xm = (k >= x0 && k <= x1) ? 1.0 : 0.0;
        if (inner) x = (T)((double)x - s * xm);
        return (T)((double)x + dd);
    }
};

template <typename T, bool REG>
struct S_step_dyn2 {
    static constexpr int NTHR = NH_BLOCK, WAVES = 2;
    __device__ __forceinline__ static void run(const VB &vb_, const DevDomain &d, double r2dt, double r1_2dt_b, const T *__restrict__ ub, const T *__restrict__ vb, T *__restrict__ ua, T *__restrict__ va, T *__restrict__ un, T *__restrict__ vn, const T *__restrict__ avm, const double *__restrict__ rCdU_bot, const double *__restrict__ ua_b, const double *__restrict__ va_b, const double *__restrict__ ub_b, const double *__restrict__ vb_b, const double *__restrict__ un_adv, const double *__restrict__ vn_adv, const double *__restrict__ r1_hu_n, const double *__restrict__ r1_hv_n, const double *__restrict__ un_b, const double *__restrict__ vn_b, const double *__restrict__ su_s, const double *__restrict__ sv_s, const double *__restrict__ utau, const double *__restrict__ utau_b, const double *__restrict__ vtau, const double *__restrict__ vtau_b, double *__restrict__ zwd_s)
    {
        THREAD_IJ_V(vb_, 1, 1);
        if (i > d.jpi || j > d.jpj) return;
        const size_t ij = IJ(i, j);
        const size_t plane = d.n2;
        const uint32_t ij32 = IJ32(i, j);
        // phase 3 on the now velocities, lines 1254-1259
        {
            const double cu = un_adv[ij] * r1_hu_n[ij], cv = vn_adv[ij] * r1_hv_n[ij];
            const double ubn = un_b[ij], vbn = vn_b[ij];
            const int u0 = d.kU.k0[ij], u1 = d.kU.k1[ij], v0 = d.kV.k0[ij], v1 = d.kV.k1[ij];
    #pragma unroll 4
            for (int k = 1; k <= d.jpk - 1; k++) {
                T *uk = plane_ptr(un, plane, k), *vk = plane_ptr(vn, plane, k);
                const double um = (k >= u0 && k <= u1) ? 1.0 : 0.0, vm = (k >= v0 && k <= v1) ? 1.0 : 0.0;
                stp(uk, ij32, (T)(((double)ldp((const T *)uk, ij32) + cu - ubn) * um));
                stp(vk, ij32, (T)(((double)ldp((const T *)vk, ij32) + cv - vbn) * vm));
            }
        }
        const bool inner = is_interior(d, i, j);
        PreSpg<T> pu{su_s[ij], (ua_b[ij] - ub_b[ij]) * r1_2dt_b, 0.0, d.kU.k0[ij], d.kU.k1[ij], inner};
        PreSpg<T> pv{sv_s[ij], (va_b[ij] - vb_b[ij]) * r1_2dt_b, 0.0, d.kV.k0[ij], d.kV.k1[ij], inner};
        if constexpr (REG) {
            zdf_column_reg<ZDF_REG_KMAX>(d, r2dt, 0, i, j, ub, ua, avm, rCdU_bot, ua_b, utau, utau_b, pu);
            zdf_column_reg<ZDF_REG_KMAX>(d, r2dt, 1, i, j, vb, va, avm, rCdU_bot, va_b, vtau, vtau_b, pv);
        } else {
            zdf_column(d, r2dt, 0, i, j, ub, ua, avm, rCdU_bot, ua_b, utau, utau_b, zwd_s, pu);
            zdf_column(d, r2dt, 1, i, j, vb, va, avm, rCdU_bot, va_b, vtau, vtau_b, zwd_s, pv);
        }
    }
};
template <typename T, bool REG>
KERNEL_W(2) void k_step_dyn2(DevDomain d, double r2dt, double r1_2dt_b, const T *__restrict__ ub, const T *__restrict__ vb,
                        T *__restrict__ ua, T *__restrict__ va, T *__restrict__ un, T *__restrict__ vn,
                        const T *__restrict__ avm, const double *__restrict__ rCdU_bot,
                        const double *__restrict__ ua_b, const double *__restrict__ va_b,
                        const double *__restrict__ ub_b, const double *__restrict__ vb_b,
                        const double *__restrict__ un_adv, const double *__restrict__ vn_adv,
                        const double *__restrict__ r1_hu_n, const double *__restrict__ r1_hv_n,
                        const double *__restrict__ un_b, const double *__restrict__ vn_b,
                        const double *__restrict__ su_s, const double *__restrict__ sv_s,
                        const double *__restrict__ utau, const double *__restrict__ utau_b,
                        const double *__restrict__ vtau, const double *__restrict__ vtau_b,
                        double *__restrict__ zwd_s)
{
    S_step_dyn2<T, REG>::run(vb_hw(), d, r2dt, r1_2dt_b, ub, vb, ua, va, un, vn, avm, rCdU_bot, ua_b, va_b, ub_b, vb_b, un_adv, vn_adv, r1_hu_n, r1_hv_n, un_b, vn_b, su_s, sv_s, utau, utau_b, vtau, vtau_b, zwd_s);
}

template <typename T, int ADV, int KEG, int SCHEME>
static int launch_step_dyn1_s(const DevDomain &d, Grid2 g, const T *un, const T *vn, const T *wn, T *ua, T *va,
                              const nemo_hip_spg_ts_t *st, double **w)
{
    const double *r1u = st->r1_hu_n, *r1v = st->r1_hv_n, *ub = st->un_b, *vb = st->vn_b;
    switch (ADV) {
    case NEMO_NP_LIN_DYN:
        LAUNCHB("step_dyn", (k_step_dyn1<T, ADV, KEG, SCHEME, np_COR, np_COR, np_COR>), (S_step_dyn1<T, ADV, KEG, SCHEME, np_COR, np_COR, np_COR>), g, d, un, vn, wn, ua, va, r1u,
               r1v, ub, vb, w[0], w[1], w[2], w[3], w[4], w[5]);
        break;
    case NEMO_NP_VEC_C2:
        LAUNCHB("step_dyn", (k_step_dyn1<T, ADV, KEG, SCHEME, np_COR, np_RVO, np_CRV>), (S_step_dyn1<T, ADV, KEG, SCHEME, np_COR, np_RVO, np_CRV>), g, d, un, vn, wn, ua, va, r1u,
               r1v, ub, vb, w[0], w[1], w[2], w[3], w[4], w[5]);
        break;
    default:
        LAUNCHB("step_dyn", (k_step_dyn1<T, ADV, KEG, SCHEME, np_COR, np_MET, np_CME>), (S_step_dyn1<T, ADV, KEG, SCHEME, np_COR, np_MET, np_CME>), g, d, un, vn, wn, ua, va, r1u,
               r1v, ub, vb, w[0], w[1], w[2], w[3], w[4], w[5]);
        break;
    }
    return 0;
}
template <typename T, int SCHEME>
static int launch_step_dyn1_v(const DevDomain &d, Grid2 g, const T *un, const T *vn, const T *wn, T *ua, T *va,
                              const nemo_hip_spg_ts_t *st, double **w)
{
    switch (d.n_dynadv) {
    case NEMO_NP_LIN_DYN: return launch_step_dyn1_s<T, NEMO_NP_LIN_DYN, 0, SCHEME>(d, g, un, vn, wn, ua, va, st, w);
    case NEMO_NP_VEC_C2:
        if (d.nn_dynkeg == 0) return launch_step_dyn1_s<T, NEMO_NP_VEC_C2, 0, SCHEME>(d, g, un, vn, wn, ua, va, st, w);
        return launch_step_dyn1_s<T, NEMO_NP_VEC_C2, 1, SCHEME>(d, g, un, vn, wn, ua, va, st, w);
    case NEMO_NP_FLX_C2: return launch_step_dyn1_s<T, NEMO_NP_FLX_C2, 0, SCHEME>(d, g, un, vn, wn, ua, va, st, w);
    default: return nh::set_error(NEMO_HIP_ERR_UNSUPPORTED, "step_dyn: n_dynadv = %d", d.n_dynadv);
    }
}

template <typename T>
int launch_step_dyn(int kt, double r2dt, const T *ub, const T *vb, const T *wn, const T *avm, const double *utau_b,
                    const double *vtau_b, const nemo_hip_spg_ts_t *st)
{
    const DevDomain &d = nh::dom();
    T *ua = (T *)st->ua, *va = (T *)st->va, *un = (T *)st->un, *vn = (T *)st->vn;
    // the barotropic work arrays the phase-1 kernel of dyn_spg_ts would fill
    // (zu_frc, zv_frc, zwxb, zwyb) plus the two vertical sums phase 3 needs
    double *w[6] = {nh::scratch2(0), nh::scratch2(1), nh::scratch2(27), nh::scratch2(28), nh::scratch2(2),
                    nh::scratch2(3)};
    Grid2 g = grid2(d.jpi, d.jpj);
    int rc;
    switch (d.nvor_scheme) {
    case NEMO_NP_ENE: rc = launch_step_dyn1_v<T, NEMO_NP_ENE>(d, g, un, vn, wn, ua, va, st, w); break;
    case NEMO_NP_ENS: rc = launch_step_dyn1_v<T, NEMO_NP_ENS>(d, g, un, vn, wn, ua, va, st, w); break;
    case NEMO_NP_MIX: rc = launch_step_dyn1_v<T, NEMO_NP_MIX>(d, g, un, vn, wn, ua, va, st, w); break;
    case NEMO_NP_ENT: rc = launch_step_dyn1_v<T, NEMO_NP_ENT>(d, g, un, vn, wn, ua, va, st, w); break;
    case NEMO_NP_EET: rc = launch_step_dyn1_v<T, NEMO_NP_EET>(d, g, un, vn, wn, ua, va, st, w); break;
    default: rc = launch_step_dyn1_v<T, NEMO_NP_EEN>(d, g, un, vn, wn, ua, va, st, w); break;
    }
    if (rc) return rc;
    if ((rc = launch_dyn_spg_ts<T>(kt, st, 1))) return rc;
    const double z2dt_bf = (kt == d.nit000 && d.neuler == 0) ? d.rdt : 2.0 * d.rdt;
    double *zwd = (double *)nh::scratch3(nh::S_ZWT);
    if constexpr (sizeof(T) == 4) {
        if (zdf_reg<T>(d)) {
            LAUNCHB("step_dyn", (k_step_dyn2<T, true>), (S_step_dyn2<T, true>), g, d, r2dt, 1.0 / z2dt_bf, ub, vb, ua, va, un, vn, avm,
                   st->rCdU_bot, st->ua_b, st->va_b, st->ub_b, st->vb_b, st->un_adv, st->vn_adv, st->r1_hu_n,
                   st->r1_hv_n, st->un_b, st->vn_b, w[2], w[3], st->utau, utau_b, st->vtau, vtau_b, zwd);
            return 0;
        }
    }
    LAUNCHB("step_dyn", (k_step_dyn2<T, false>), (S_step_dyn2<T, false>), g, d, r2dt, 1.0 / z2dt_bf, ub, vb, ua, va, un, vn, avm,
           st->rCdU_bot, st->ua_b, st->va_b, st->ub_b, st->vb_b, st->un_adv, st->vn_adv, st->r1_hu_n, st->r1_hv_n,
           st->un_b, st->vn_b, w[2], w[3], st->utau, utau_b, st->vtau, vtau_b, zwd);
    return 0;
}
template int launch_step_dyn<double>(int, double, const double *, const double *, const double *, const double *,
                                     const double *, const double *, const nemo_hip_spg_ts_t *);
template int launch_step_dyn<float>(int, double, const float *, const float *, const float *, const float *,
                                    const double *, const double *, const nemo_hip_spg_ts_t *);
