// fct_vlim.hip -- the vertical half of tra_adv_fct and nonosc in one
// bottom-up k-march per 64 x TY tile and tracer, closed domains
// (TRA/traadv_fct.F90:119-297, 330-428, 577-614).
//
// The flux kernel before it (k_fct_h with VERT = 0) leaves, per cell, the
// horizontal antidiffusive fluxes zwx, zwy and the partial sum s of the
// horizontal upstream-flux divergence; the compact scheme's forward sweep
// (k_fct_cpt_fwd) leaves its eliminated right-hand sides zo.  This march
// then forms, level by level from the bottom:
//   * the vertical upstream flux and, from s, the low-order trend and the
//     first guess zwi on the tile and a 2-point halo (lines 148-170);
//   * the bounds of nonosc from tb and zwi (361-364);
//   * the vertical antidiffusive flux zwz on the tile and a 1-point halo,
//     the compact scheme's back substitution running up each column in a
//     register (255-278, 603-614);
//   * the betas (366-399) and the limited divergence (404-422, 288-297).
// So zwi and zwz never reach HBM, and the vertical upstream flux is formed
// once per column.  Every value is the one the kernel chain computes (same
// expressions in the same order, rounded to the storage type at the same
// points), and the ring values the closed domain's lbc_lnk gives (zero) are
// selected where they apply.
//
// The march runs three stages a level apart: at the step of level k the
// bounds stage works on level k-2, the betas on level k-1 and the limited
// divergence on level k, separated by two barriers.  Every global load is
// issued one step ahead, unconditionally, from addresses clamped into the
// arrays.
#include "fct.hpp"

namespace {


// NTH threads per 64 x TY tile (TY a multiple of NTH / 64)
template <int TRP, typename T, int NV4, int TA0, int TY, int NTH>
struct S_fct_vlim {
    static constexpr int NTHR = NTH, WAVES = NTH == 512 ? 2 : (sizeof(T) == 4 ? 3 : 2);
    __device__ __forceinline__ static void run(const VB &vb_, const DevDomain &d, double p2dt,
                                               const T *__restrict__ pw, const T *__restrict__ ptb_,
                                               const T *__restrict__ ps_, const T *__restrict__ pzo_,
                                               const T *__restrict__ paa_, const T *__restrict__ pbb_,
                                               T *__restrict__ ta_)
    {
        // bounds as in k_fct_lim: FP64 with +-1e40 land sentinels, or FP32
        // with +-inf for FP32 storage
        using ZB = typename std::conditional<sizeof(T) == 4, float, double>::type;
        const ZB zsent = sizeof(T) == 4 ? (ZB)INFINITY : (ZB)1.e+40;
        constexpr int TX = 64;
        constexpr int ZX = TX + 4, ZY = TY + 4, BX = TX + 2, BY = TY + 2;
        constexpr int NZ = (ZX * ZY + NTH - 1) / NTH, NB = (BX * BY + NTH - 1) / NTH;
        constexpr int WY = NTH / TX, NR = TY / WY;
        static_assert(TY % WY == 0, "tile height must be a multiple of the workgroup height");
        __shared__ ZB zup_s[3][ZY][ZX], zdo_s[3][ZY][ZX];   // bounds, level mod 3
        __shared__ T bup_s[2][BY][BX], bdo_s[2][BY][BX];    // betas, level mod 2
        __shared__ T aft_s[3][BY][BX];                      // first guess, level mod 3
        __shared__ T pcc_s[4][BY][BX];                      // zwz, level mod 4
        __shared__ T tra_s[4][TY][TX];                      // low-order trend, level mod 4
        __shared__ T faa_s[BY][BX], fbb_s[BY][BX];          // zwx, zwy of the divergence's level

        const size_t o = (size_t)vb_.z * d.n3;   // vb_.z: the tracer of the pair
        const T *tb = ptb_ + o, *ps = ps_ + o, *pzo = pzo_ + o, *paa = paa_ + o, *pbb = pbb_ + o;
        T *ta = ta_ + o;
        unsigned bx, by;
        xcd_tile_v(vb_, bx, by);
        const int tx = (int)threadIdx.x, ty = (int)threadIdx.y, tid = ty * TX + tx;
        const int i0 = 2 + (int)bx * TX, j0 = 2 + (int)by * TY;
        const int i = i0 + tx;
        const int jpi = d.jpi, jpj = d.jpj, jpk = d.jpk;
        const double zrtrn = 1.e-15;
        const Metrics<T> &m = met<T>(d);
        const T zp2dt = (T)p2dt;

        // ---- bounds stage: this thread's points of the tile + 2 region
        int zpx[NZ], zpy[NZ], zmb[NZ];
        bool zon[NZ], zin[NZ], zinr[NZ], zbe[NZ], zti[NZ];
        uint32_t zkt[NZ], zof[NZ];
        T zr1[NZ], ze12[NZ];
        T ztb[NZ], ztbu[NZ], zs[NZ], zw[NZ], zzo[NZ], ztn[NZ], zupz[NZ], zx[NZ];
    #pragma unroll
        for (int r = 0; r < NZ; r++) {
            const int q = tid + r * NTH;
            zon[r] = q < ZX * ZY;
            zpx[r] = q % ZX;
            zpy[r] = q / ZX;
            const int pi = i0 - 2 + zpx[r], pj = j0 - 2 + zpy[r];
            zin[r] = zon[r] && pi >= 1 && pi <= jpi && pj >= 1 && pj <= jpj;
            zinr[r] = zon[r] && pi >= 2 && pi <= jpi - 1 && pj >= 2 && pj <= jpj - 1;
            zbe[r] = zon[r] && zpx[r] >= 1 && zpx[r] <= BX && zpy[r] >= 1 && zpy[r] <= BY;
            zti[r] = zon[r] && zpx[r] >= 2 && zpx[r] < TX + 2 && zpy[r] >= 2 && zpy[r] < TY + 2;
            zof[r] = IJ32(min(max(pi, 1), jpi), min(max(pj, 1), jpj));
            zkt[r] = zin[r] ? d.ktu[zof[r]] : 0u;
            zmb[r] = zinr[r] ? (int)d.mbkt[zof[r]] : 0;
            zr1[r] = zinr[r] ? m.r1_e1e2t[zof[r]] : T(0);
            ze12[r] = m.e1e2t[zof[r]];
            zupz[r] = T(0);
            zx[r] = T(0);
        }
        auto zwet = [&](int r, int kk) { return kk >= (int)(zkt[r] & 255u) && kk <= (int)((zkt[r] >> 8) & 255u); };
        auto zwwet = [&](int r, int kk) {
            const int a = (int)(zkt[r] & 255u), b = (int)((zkt[r] >> 8) & 255u);
            return kk >= a && kk <= b && (kk == 1 || kk - 1 >= a);
        };
        // the loads of level mm: tb(mm-1), s(mm), w(mm), and zo(mm) (order 4)
        // or tn(mm-1) (order 2)
        auto zb_fetch = [&](int mm) {
            const T *pb = plane_ptr(tb, d.n2, min(max(mm - 1, 1), jpk));
            const T *pq = plane_ptr(ps, d.n2, min(max(mm, 1), jpk - 1));
            const T *pv = plane_ptr(pw, d.n2, min(max(mm, 1), jpk));
            const T *pz = plane_ptr(pzo, d.n2, NV4 ? min(max(mm, 1), jpk) : min(max(mm - 1, 1), jpk));
    #pragma unroll
            for (int r = 0; r < NZ; r++) {
                ztbu[r] = ldp(pb, zof[r]);
                zs[r] = ldp(pq, zof[r]);
                zw[r] = ldp(pv, zof[r]);
                zzo[r] = ldp(pz, zof[r]);
            }
        };
        auto zb_do = [&](int mm) {
            const int sl3 = mm % 3, sl4 = mm & 3;
    #pragma unroll
            for (int r = 0; r < NZ; r++) {
                if (!zon[r]) continue;
                // the vertical upstream flux of w-level mm (tp_upz)
                const T wv = mm >= jpk ? T(0) : (TRP ? ze12[r] * zw[r] : zw[r]);
                const T wm = zwwet(r, mm) ? T(1) : T(0);
                const T zfp = wv + fabs(wv), zfm = wv - fabs(wv);
                const T gen = T(0.5) * (zfp * ztb[r] + zfm * ztbu[r]) * wm;
                const T upz = mm == 1 ? wv * ztb[r] : (mm >= jpk ? T(0) : gen);
                // low-order trend and first guess (interior cells, levels < jpk)
                T zwi = T(0), ztra = T(0);
                if (zinr[r] && mm <= jpk - 1) {
                    const T e3t = E3(m.e3t, zof[r], mm), tm = zwet(r, mm) ? T(1) : T(0);
                    ztra = -(((zs[r] + upz) - zupz[r]) * zr1[r]);
                    zwi = (e3t * ztb[r] + zp2dt * ztra) / e3t * tm;
                }
                // bounds (lines 361-364)
                ZB u = 0, w = 0;
                if (zin[r] && mm <= jpk) {
                    if constexpr (sizeof(T) == 4) {
                        const bool wet = zwet(r, mm);
                        const ZB a = (ZB)ztb[r], b = (ZB)zwi;
                        u = wet ? (a > b ? a : b) : -zsent;
                        w = wet ? (a < b ? a : b) : zsent;
                    } else {
                        const double tm = zwet(r, mm) ? 1.0 : 0.0;
                        const double au = (double)ztb[r] * tm - zsent * (1.0 - tm);
                        const double bu = (double)zwi * tm - zsent * (1.0 - tm);
                        u = au > bu ? au : bu;
                        const double ad = (double)ztb[r] * tm + zsent * (1.0 - tm);
                        const double bd = (double)zwi * tm + zsent * (1.0 - tm);
                        w = ad < bd ? ad : bd;
                    }
                }
                zup_s[sl3][zpy[r]][zpx[r]] = u;
                zdo_s[sl3][zpy[r]][zpx[r]] = w;
                if (zbe[r]) {
                    // the vertical antidiffusive flux of w-level mm
                    T zwz = T(0);
                    if (zinr[r] && mm >= 2 && mm <= jpk - 1) {
                        if constexpr (NV4) {
                            const T piv = cpt_piv(cpt_tab<T>(), mm, zmb[r]);
                            T x;
                            if (mm == jpk - 1) {
                                x = zzo[r] / piv;
                            } else {
                                const T zws = (mm == 2 || mm == zmb[r]) ? T(0) : wm;
                                x = (zzo[r] - zws * zx[r]) / piv;
                            }
                            zx[r] = x;
                            zwz = (wv * x - upz) * wm;
                        } else {
                            zwz = (wv * T(0.5) * (ztn[r] + zzo[r]) - upz) * wm;
                        }
                    }
                    aft_s[sl3][zpy[r] - 1][zpx[r] - 1] = zwi;
                    pcc_s[sl4][zpy[r] - 1][zpx[r] - 1] = zwz;
                }
                if (zti[r]) tra_s[sl4][zpy[r] - 2][zpx[r] - 2] = ztra;
                zupz[r] = upz;
                ztb[r] = ztbu[r];
                if constexpr (!NV4) ztn[r] = zzo[r];
            }
        };

        // ---- betas: this thread's points of the tile + 1 region
        int bpx[NB], bpy[NB];
        bool bon[NB], bin[NB];
        uint32_t bof[NB];
        T fa0[NB], fa1[NB], fb0[NB], fb1[NB], fa1_prev[NB], fb1_prev[NB];
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
        auto beta_fetch = [&](int mm) {
            const int kc = min(max(mm, 1), jpk - 1);
            const T *pa = plane_ptr(paa, d.n2, kc), *pb = plane_ptr(pbb, d.n2, kc);
            const uint32_t w = (uint32_t)jpi;
    #pragma unroll
            for (int r = 0; r < NB; r++) {
                const uint32_t q = bof[r];
                fa0[r] = ldp(pa, q - 1u);
                fa1[r] = ldp(pa, q);
                fb0[r] = ldp(pb, q - w);
                fb1[r] = ldp(pb, q);
                fe3t[r] = E3(d.md.e3t, q, kc);
            }
        };
        // nonosc betas of level mm (lines 366-399, the closed ring zero); also
        // hands zwx, zwy of level mm+1 to the divergence of that level
        auto beta_do = [&](int mm) {
            const int sb = mm & 1;
            const int sa = (mm > 1 ? mm - 1 : 1) % 3, sc = mm % 3, sn = (mm + 1) % 3;
    #pragma unroll
            for (int r = 0; r < NB; r++) {
                if (!bon[r]) continue;
                faa_s[bpy[r]][bpx[r]] = fa1_prev[r];
                fbb_s[bpy[r]][bpx[r]] = fb1_prev[r];
                fa1_prev[r] = fa1[r];
                fb1_prev[r] = fb1[r];
                T bu_t = T(0), bd_t = T(0);
                if (bin[r] && mm <= jpk - 1) {
                    const int zx_ = bpx[r] + 1, zy_ = bpy[r] + 1;
                    ZB zup = zup_s[sc][zy_][zx_];
                    zup = fmax2(zup, zup_s[sc][zy_][zx_ - 1]);
                    zup = fmax2(zup, zup_s[sc][zy_][zx_ + 1]);
                    zup = fmax2(zup, zup_s[sc][zy_ - 1][zx_]);
                    zup = fmax2(zup, zup_s[sc][zy_ + 1][zx_]);
                    zup = fmax2(zup, zup_s[sa][zy_][zx_]);
                    zup = fmax2(zup, zup_s[sn][zy_][zx_]);
                    ZB zdo = zdo_s[sc][zy_][zx_];
                    zdo = fmin2(zdo, zdo_s[sc][zy_][zx_ - 1]);
                    zdo = fmin2(zdo, zdo_s[sc][zy_][zx_ + 1]);
                    zdo = fmin2(zdo, zdo_s[sc][zy_ - 1][zx_]);
                    zdo = fmin2(zdo, zdo_s[sc][zy_ + 1][zx_]);
                    zdo = fmin2(zdo, zdo_s[sa][zy_][zx_]);
                    zdo = fmin2(zdo, zdo_s[sn][zy_][zx_]);
                    const double a_im1 = fa0[r], a_i = fa1[r], b_jm1 = fb0[r], b_j = fb1[r];
                    const double c_k = pcc_s[mm & 3][bpy[r]][bpx[r]], c_kp1 = pcc_s[(mm + 1) & 3][bpy[r]][bpx[r]];
                    const double zpos = fmax2(0.0, a_im1) - fmin2(0.0, a_i) + fmax2(0.0, b_jm1) - fmin2(0.0, b_j) +
                                        fmax2(0.0, c_kp1) - fmin2(0.0, c_k);
                    const double zneg = fmax2(0.0, a_i) - fmin2(0.0, a_im1) + fmax2(0.0, b_j) - fmin2(0.0, b_jm1) +
                                        fmax2(0.0, c_k) - fmin2(0.0, c_kp1);
                    const double zbt = fe1e2t[r] * fe3t[r] / p2dt;
                    const double paft_k = aft_s[sc][bpy[r]][bpx[r]];
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

        // ---- the limited divergence: this thread's cells of its i column
        int jr[NR];
        bool own[NR];
        uint32_t fof[NR], fkt[NR];
        T gta[NR], ge3t[NR], r1_e1e2t[NR];
        double lz_kp1[NR];
    #pragma unroll
        for (int r = 0; r < NR; r++) {
            jr[r] = j0 + ty + WY * r;
            own[r] = (i <= jpi - 1 && jr[r] <= jpj - 1);
            fof[r] = IJ32(min(i, jpi - 1), min(jr[r], jpj - 1));
            fkt[r] = own[r] ? d.ktu[fof[r]] : 0u;
            r1_e1e2t[r] = own[r] ? m.r1_e1e2t[fof[r]] : T(0);
            gta[r] = ge3t[r] = T(1);
            lz_kp1[r] = 0.0;
        }
        auto fin_fetch = [&](int k) {
            const int kc = min(max(k, 1), jpk - 1);
            const T *pt_ = plane_ptr((const T *)ta, d.n2, kc);
    #pragma unroll
            for (int r = 0; r < NR; r++) {
                if constexpr (!TA0) gta[r] = ldp(pt_, fof[r]);
                ge3t[r] = E3(m.e3t, fof[r], kc);
            }
        };
        const int cx = tx + 1;
        const bool okxm = i - 1 >= 2;
        auto fin_do = [&](int k) {
            const int s0 = k & 1, s1 = (k - 1) & 1;
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
                // the limited flux of w-level k+1: zero flux and zero betas
                // below level jpk-1, else the one found at level k+1
                if (k == jpk - 1) lz_kp1[r] = lim_face<T>(0.0, 0.0, bu_c, 0.0, bd_c);
                // of w-level k: the surface flux zwz(1) = 0 is not limited
                const double lz_k =
                    k == 1 ? 0.0
                           : lim_face<T>(pcc_s[k & 3][cy][cx], bd_c, bup_s[s1][cy][cx], bu_c, bdo_s[s1][cy][cx]);
                const int kt0 = (int)(fkt[r] & 255u), kt1 = (int)((fkt[r] >> 8) & 255u);
                const T tm = (k >= kt0 && k <= kt1) ? T(1) : T(0);
                const T a = TA0 ? T(0) : gta[r];
                const T t1 = a + tra_s[k & 3][cy - 1][tx] / ge3t[r] * tm;   // the low-order trend added
                A3(ta, i, j, k) = t1 - ((T)lx_i - (T)lx_im1 + (T)ly_j - (T)ly_jm1 + (T)lz_k - (T)lz_kp1[r]) *
                                           r1_e1e2t[r] / ge3t[r];
                lz_kp1[r] = lz_k;
            }
        };

        // ---- prologue: bounds of levels jpk, jpk-1, jpk-2, betas of jpk-1,
        // the loads of the first step in flight
        const int K = jpk - 1;
        {
            const T *pb = plane_ptr(tb, d.n2, jpk), *pn = plane_ptr(pzo, d.n2, jpk);
    #pragma unroll
            for (int r = 0; r < NZ; r++) {
                ztb[r] = ldp(pb, zof[r]);
                ztn[r] = NV4 ? T(0) : ldp(pn, zof[r]);
            }
        }
        zb_fetch(jpk);
        zb_do(jpk);
        zb_fetch(K);
        zb_do(K);
        zb_fetch(K - 1);
        beta_fetch(K);
        zb_do(K - 1);
        zb_fetch(K - 2);
        __syncthreads();
        beta_do(K);
        beta_fetch(K - 1);
        fin_fetch(K);
        __syncthreads();
        // ---- march: at the top of step k the LDS holds the bounds of levels
        // k-1..k+1, the betas of level k, zwz of levels k-1..k+1 and the
        // trends of levels k-1..k+1; the registers hold the loads of level
        // k-2's bounds, k-1's betas and k's divergence
        for (int k = K; k >= 1; k--) {
            if (k - 2 >= 1) {
                zb_do(k - 2);
                zb_fetch(k - 3);
            }
            __syncthreads();
            if (k - 1 >= 1) {
                beta_do(k - 1);
                beta_fetch(k - 2);
            } else {
                // hand zwx, zwy of level 1 to its divergence
    #pragma unroll
                for (int r = 0; r < NB; r++) {
                    if (!bon[r]) continue;
                    faa_s[bpy[r]][bpx[r]] = fa1_prev[r];
                    fbb_s[bpy[r]][bpx[r]] = fb1_prev[r];
                }
            }
            __syncthreads();
            fin_do(k);
            fin_fetch(k - 1);
        }
    }
};

template <int TRP, typename T, int NV4, int TA0, int TY, int NTH>
__global__ __launch_bounds__(NTH, (S_fct_vlim<TRP, T, NV4, TA0, TY, NTH>::WAVES)) void k_fct_vlim(
    DevDomain d, double p2dt, const T *__restrict__ pw, const T *__restrict__ ptb_,
    const T *__restrict__ ps_, const T *__restrict__ pzo_, const T *__restrict__ paa_, const T *__restrict__ pbb_,
    T *__restrict__ ta_)
{
    S_fct_vlim<TRP, T, NV4, TA0, TY, NTH>::run(vb_hw(), d, p2dt, pw, ptb_, ps_, pzo_, paa_, pbb_, ta_);
}

}  // namespace

// switch fct_vlim 1: 256 threads per 64 x 4 tile; 2: 512 threads per 64 x 8
template <int TRP, typename T, int V, int Z, int TY, int NTH>
static int vl_launch(const DevDomain &d, double p2dt, int nt, const T *pw, const T *ptb, const T *ps, const T *pzo,
                     const T *paa, const T *pbb, T *pta, const char *nm)
{
    Grid2 g = grid2(d.jpi - 2, d.jpj - 2);
    g.block = dim3(64, NTH / 64, 1);
    g.grid.y = (unsigned)((d.jpj - 2 + TY - 1) / TY);
    g.grid.z = (unsigned)nt;
    LAUNCHB(nm, (k_fct_vlim<TRP, T, V, Z, TY, NTH>), (S_fct_vlim<TRP, T, V, Z, TY, NTH>), g, d, p2dt, pw, ptb, ps,
            pzo, paa, pbb, pta);
    return 0;
}

template <int TRP, typename T>
int launch_fct_vlim(const DevDomain &d, double p2dt, int nt, int kn_fct_v, int ta0,
                    const T *pw, const T *ptb, const T *ps, const T *pzo, const T *paa, const T *pbb, T *pta,
                    const char *nm)
{
    if (d.jpk < 3) return nh::set_error(NEMO_HIP_ERR_ARG, "fct: jpk = %d (at least 3)", d.jpk);
    const bool wide = nh::tune(nh::T_FCT_VLIM) == 2;
#define VL_CASE(V, Z)                                                                                           \
    if ((kn_fct_v == 4) == !!V && !!ta0 == !!Z)                                                                 \
        return wide ? vl_launch<TRP, T, V, Z, 8, 512>(d, p2dt, nt, pw, ptb, ps, pzo, paa, pbb, pta, nm)         \
                    : vl_launch<TRP, T, V, Z, 4, 256>(d, p2dt, nt, pw, ptb, ps, pzo, paa, pbb, pta, nm);
    VL_CASE(1, 1) VL_CASE(1, 0) VL_CASE(0, 1) VL_CASE(0, 0)
#undef VL_CASE
    return nh::set_error(NEMO_HIP_ERR_ARG, "fct: vertical order %d", kn_fct_v);
}
#define VL_INST(TRP, T)                                                                                         \
    template int launch_fct_vlim<TRP, T>(const DevDomain &, double, int, int, int,                            \
                                         const T *, const T *, const T *, const T *, const T *, const T *, T *, \
                                         const char *);
VL_INST(0, double)
VL_INST(1, double)
VL_INST(0, float)
VL_INST(1, float)
