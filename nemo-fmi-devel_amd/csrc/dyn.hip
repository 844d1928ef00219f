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
// a13 + a14: vector-form advection, dyn_keg then dyn_zad, fused per element
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
__global__ void k_dyn_keg_zad(DevDomain d, int do_keg, int kscheme, int do_zad, const T *un, const T *vn, const T *wn,
                              T *ua, T *va)
{
    THREAD_IJ(2, 2);
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

// ---------------------------------------------------------------------------
// a15: dyn_adv_cen2, horizontal then vertical flux divergence per element
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
__global__ void k_dyn_adv_cen2(DevDomain d, const T *un, const T *vn, const T *wn, T *ua, T *va)
{
    THREAD_IJ(2, 2);
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

template <typename T>
int launch_dyn_adv(int which, int kscheme, const T *un, const T *vn, const T *wn, T *ua, T *va)
{
    // which: 1 keg only, 2 zad only, 3 keg+zad (vector form), 4 cen2
    const DevDomain &d = nh::dom();
    Grid2 g = grid2(d.jpi - 2, d.jpj - 2);
    if (which == 4) {
        LAUNCH("dyn_adv_cen2", k_dyn_adv_cen2<T>, g, d, un, vn, wn, ua, va);
    } else {
        const char *nm = which == 1 ? "dyn_keg" : which == 2 ? "dyn_zad" : "dyn_adv";
        LAUNCH(nm, k_dyn_keg_zad<T>, g, d, (which & 1), kscheme, (which & 2) ? 1 : 0, un, vn, wn, ua, va);
    }
    return 0;
}
template int launch_dyn_adv<double>(int, int, const double *, const double *, const double *, double *, double *);
template int launch_dyn_adv<float>(int, int, const float *, const float *, const float *, float *, float *);

// ---------------------------------------------------------------------------
// a16: vorticity term
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

template <typename T>
__global__ void k_dyn_vor(DevDomain d, int scheme, int ncor, int nrvm, int ntot, const T *un, const T *vn, T *ua, T *va)
{
    THREAD_IJ(2, 2);
    if (i > d.jpi - 1 || j > d.jpj - 1) return;
    for (int k = 1; k <= d.jpk - 1; k++) {
        const size_t p = IJK(i, j, k);
        T a = ua[p], b = va[p];
        switch (scheme) {
        case NEMO_NP_ENE: vor_ene_pt(d, ntot, un, vn, i, j, k, a, b); break;
        case NEMO_NP_ENS: vor_ens_pt(d, ntot, un, vn, i, j, k, a, b); break;
        case NEMO_NP_MIX:
            vor_ens_pt(d, nrvm, un, vn, i, j, k, a, b);
            vor_ene_pt(d, ncor, un, vn, i, j, k, a, b);
            break;
        case NEMO_NP_ENT: vor_ent_pt(d, ntot, un, vn, i, j, k, a, b); break;
        default: vor_een_pt(d, scheme, ntot, un, vn, i, j, k, a, b); break;
        }
        ua[p] = a;
        va[p] = b;
    }
}

template <typename T>
int launch_dyn_vor(const T *un, const T *vn, T *ua, T *va)
{
    const DevDomain &d = nh::dom();
    int ncor = np_COR, nrvm, ntot;
    switch (d.n_dynadv) {
    case NEMO_NP_LIN_DYN: nrvm = np_COR; ntot = np_COR; break;
    case NEMO_NP_VEC_C2: nrvm = np_RVO; ntot = np_CRV; break;
    default: nrvm = np_MET; ntot = np_CME; break;
    }
    Grid2 g = grid2(d.jpi - 2, d.jpj - 2);
    LAUNCH("dyn_vor", k_dyn_vor<T>, g, d, d.nvor_scheme, ncor, nrvm, ntot, un, vn, ua, va);
    return 0;
}
template int launch_dyn_vor<double>(const double *, const double *, double *, double *);
template int launch_dyn_vor<float>(const float *, const float *, float *, float *);

// ---------------------------------------------------------------------------
// a17: dyn_zdf, DYN/dynzdf.F90:107-485 (ln_zad_Aimp = F, iso-level momentum
// mixing, r_vvl = 0).  All columns apply the leapfrog right-hand side; the
// interior ones then solve the u and v columns (FP64 elimination).
// ---------------------------------------------------------------------------
template <typename T>
__device__ void zdf_column(const DevDomain &d, double r2dt, int comp, int i, int j, const T *pb, T *pa,
                           const T *avm, const double *rCdU_bot, const double *pa_b, const double *tau,
                           const double *tau_b, double *zwd_s, double *fwd_s)
{
    const Metrics<T> &m = met<T>(d);
    const int jpk = d.jpk;
    const int di = comp ? 0 : 1, dj = comp ? 1 : 0;
    const size_t ij = IJ(i, j);
    const T *e3x = comp ? m.e3v : m.e3u;
    const T *e3xw = comp ? m.e3vw : m.e3uw;
    const KRange &kr = comp ? d.kV : d.kU;
    const double r_vvl = 0.0;
    const bool inner = is_interior(d, i, j);
    const bool rm_b = d.ln_drgimp && d.ln_dynspg_ts;
    const double ab = rm_b ? pa_b[ij] : 0.0;
    const int ikb = comp ? d.mbkv[ij] : d.mbku[ij];
    double csum = 0.0;
    if (inner) csum = rCdU_bot[IJ(i + di, j + dj)] + rCdU_bot[ij];
    if (!inner) {
        for (int k = 1; k <= jpk - 1; k++) {
            const size_t p = IJK(i, j, k);
            const double xm = kmask<double>(kr, ij, k);
            double a = ((double)pb[p] + r2dt * (double)pa[p]) * xm;
            if (rm_b) a = (a - ab) * xm;
            pa[p] = (T)a;
        }
        return;
    }
    const double zdt = r2dt * 0.5;
    double zws_prev = 0.0, zwd_prev = 1.0, a_prev = 0.0;
    for (int k = 1; k <= jpk - 1; k++) {
        const size_t p = IJK(i, j, k);
        const double xm = kmask<double>(kr, ij, k);
        double a = ((double)pb[p] + r2dt * (double)pa[p]) * xm;
        if (rm_b) a = (a - ab) * xm;
        const double ze3 = (1.0 - r_vvl) * (double)E3(e3x, ij, k) + r_vvl * (double)E3(e3x, ij, k);
        if (rm_b && k == ikb) a = a + r2dt * 0.5 * csum * ab / ze3;
        const double zzwi = -(zdt * ((double)A3(avm, i + di, j + dj, k) + (double)A3(avm, i, j, k)) /
                              (ze3 * (double)E3(e3xw, ij, k)) * wkmask<double>(kr, ij, k));
        const double zzws = -(zdt * ((double)A3(avm, i + di, j + dj, k + 1) + (double)A3(avm, i, j, k + 1)) /
                              (ze3 * (double)E3(e3xw, ij, k + 1)) * wkmask<double>(kr, ij, k + 1));
        double zwi = zzwi, zws = zzws, zwd = 1.0 - zzwi - zzws;
        if (k == 1) {
            zwi = 0.0;
            zwd = 1.0 - zws;
        }
        if (d.ln_drgimp && k == ikb) zwd = zwd - r2dt * 0.5 * csum / ze3;
        if (k >= 2) zwd = zwd - zwi * zws_prev / zwd_prev;
        if (k == 1) a = a + r2dt * 0.5 * (tau_b[ij] + tau[ij]) / (ze3 * d.rau0) * xm;
        else a = a - zwi / zwd_prev * a_prev;
        A3(zwd_s, i, j, k) = zwd;
        if (fwd_s) A3(fwd_s, i, j, k) = a;
        else pa[p] = (T)a;
        zws_prev = zws;
        zwd_prev = zwd;
        a_prev = a;
    }
    double x = a_prev / zwd_prev;
    pa[IJK(i, j, jpk - 1)] = (T)x;
    for (int k = jpk - 2; k >= 1; k--) {
        const double ze3 = (1.0 - r_vvl) * (double)E3(e3x, ij, k) + r_vvl * (double)E3(e3x, ij, k);
        const double zws = -(zdt * ((double)A3(avm, i + di, j + dj, k + 1) + (double)A3(avm, i, j, k + 1)) /
                             (ze3 * (double)E3(e3xw, ij, k + 1)) * wkmask<double>(kr, ij, k + 1));
        const size_t p = IJK(i, j, k);
        const double a = fwd_s ? A3(fwd_s, i, j, k) : (double)pa[p];
        x = (a - zws * x) / A3(zwd_s, i, j, k);
        pa[p] = (T)x;
    }
}

template <typename T>
__global__ void k_dyn_zdf(DevDomain d, double r2dt, const T *ub, const T *vb, T *ua, T *va, const T *avm,
                          const double *rCdU_bot, const double *ua_b, const double *va_b, const double *utau,
                          const double *utau_b, const double *vtau, const double *vtau_b, double *zwd_s,
                          double *fwd_s)
{
    THREAD_IJ(1, 1);
    if (i > d.jpi || j > d.jpj) return;
    zdf_column(d, r2dt, 0, i, j, ub, ua, avm, rCdU_bot, ua_b, utau, utau_b, zwd_s, fwd_s);
    zdf_column(d, r2dt, 1, i, j, vb, va, avm, rCdU_bot, va_b, vtau, vtau_b, zwd_s, fwd_s);
}

template <typename T>
int launch_dyn_zdf(double r2dt, const T *ub, const T *vb, T *ua, T *va, const T *avm, const double *rCdU_bot,
                   const double *ua_b, const double *va_b, const double *utau, const double *utau_b,
                   const double *vtau, const double *vtau_b)
{
    const DevDomain &d = nh::dom();
    Grid2 g = grid2(d.jpi, d.jpj);
    double *zwd = (double *)nh::scratch3(nh::S_ZWT);
    double *fwd = nullptr;   // forward-substituted values stay in ua/va at storage precision
    LAUNCH("dyn_zdf", k_dyn_zdf<T>, g, d, r2dt, ub, vb, ua, va, avm, rCdU_bot, ua_b, va_b, utau, utau_b, vtau, vtau_b,
           zwd, fwd);
    return 0;
}
template int launch_dyn_zdf<double>(double, const double *, const double *, double *, double *, const double *,
                                    const double *, const double *, const double *, const double *, const double *,
                                    const double *, const double *);
template int launch_dyn_zdf<float>(double, const float *, const float *, float *, float *, const float *,
                                   const double *, const double *, const double *, const double *, const double *,
                                   const double *, const double *);
