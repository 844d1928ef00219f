// fct_fused.hip -- tra_adv_fct on closed domains as one k-marched kernel per
// tile and tracer (TRA/traadv_fct.F90:119-297 with nonosc, 330-428).
//
// The separate chain writes the upstream-corrected first guess zwi, the
// antidiffusive fluxes zwx, zwy, zwz and the trend to HBM and reads them
// back (k_fct_h, k_fct_v, k_lbc, k_fct_lim).  Here a workgroup owns a 64 x TY
// tile of one tracer and marches the levels from the bottom up; every one of
// those intermediates lives in LDS or registers, and HBM sees tb, tn, the
// transports, the trend (read unless it starts from zero) and the trend
// written once.  The one exception is the order-4 compact vertical
// interpolation, whose tridiagonal solve couples a whole column: a column
// pre-pass (k_fct_cpt_fwd) writes its forward elimination, and the march,
// which runs bottom up, does the back substitution in a register as it goes.
//
// Values are computed with the expressions, operands and rounding points of
// the chain, so the results are the same bits.  A level's values at the
// tile's halo are recomputed by each tile that needs them (the chain's
// lbc_lnk on a closed domain is a zero ring, selected where it applies):
//   ring 0 (the tile)        the limited divergence, the low-order trend
//   ring <= 1                the Zalesak betas, zwz, zwi carried down
//   ring <= 2                zwi and the bounds zbup / zbdo
//   x in [-3,TX+1] and y in [-3,TY+1]   upstream face fluxes (east, north)
//   ring 3                   tb and tn, loaded for the stencils
// At level it of the march: the stencil stage A computes the upstream face
// fluxes, zwi, the bounds and the antidiffusive fluxes of level it; stage B
// the betas of level it+1 (which need the bounds one level up and down);
// stage C the limited divergence of level it+2 (which needs the betas of
// levels it+1 .. it+3).  LDS holds a plane of tb and tn, one of the upstream
// fluxes (then of the limited fluxes), three planes of bounds and two of
// antidiffusive fluxes and of betas, by level.
#include "common.hpp"
#include "kernels.hpp"
#include "fct.hpp"

#define FU_TX 64
template <typename T> struct FuTY {
    static constexpr int value = sizeof(T) == 4 ? 8 : 4;   // 64x8 tiles for FP32, 64x4 for FP64
};
// waves per SIMD the LDS leaves room for (two workgroups per CU for FP32,
// one for FP64), for a workgroup of NTH threads
template <typename T, int NTH> struct FuWaves {
    static constexpr int value = (sizeof(T) == 4 ? 2 : 1) * NTH / 256;
};

// point q of a tile's load rectangle [-3, TX+2] x [-3, TY+2]: the tile
// first, then the rings around it (bottom row, top row, left column, right
// column of each), so the points of ring <= r are q < (TX+2r)(TY+2r)
template <int TX, int TY>
__device__ __forceinline__ void fu_xy(int q, int &x, int &y)
{
    if (q < TX * TY) {
        x = q % TX;
        y = q / TX;
        return;
    }
    int base = TX * TY, L = 1;
#pragma unroll
    for (int c = 1; c < 3; c++) {
        const int W = TX + 2 * L, H = TY + 2 * L;
        if (q >= base + W * H - (W - 2) * (H - 2)) {
            base += W * H - (W - 2) * (H - 2);
            L++;
        }
    }
    const int W = TX + 2 * L, H = TY + 2 * L;
    int r = q - base;
    if (r < W) { x = -L + r; y = -L; return; }
    r -= W;
    if (r < W) { x = -L + r; y = TY - 1 + L; return; }
    r -= W;
    if (r < H - 2) { x = -L; y = -L + 1 + r; return; }
    r -= H - 2;
    x = TX - 1 + L;
    y = -L + 1 + r;
}

// the forward elimination of interp_4th_cpt (traadv_fct.F90:577-597) per
// interior column: zout(k), k = 2..jpk-1, for the back substitution of the
// fused march (the same operations as k_fct_v)
template <typename T, int NT>
struct S_fct_cpt_fwd {
    static constexpr int NTHR = NH_BLOCK, WAVES = 1;
    __device__ __forceinline__ static void run(const VB &vb_, const DevDomain &d, const T *__restrict__ ptn, T *__restrict__ pzo)
    {
        THREAD_IJ_V(vb_, 2, 2);
        if (i > d.jpi - 1 || j > d.jpj - 1) return;
        const size_t n3 = d.n3;
        const int jpk = d.jpk;
        const int ikb = d.mbkt[IJ(i, j)];
    #pragma unroll
        for (int t = 0; t < NT; t++) {
            const T *tn = ptn + t * n3;
            T *zo = pzo + t * n3;
            T zout_prev = T(0), tn_km1 = A3(tn, i, j, 1);
    #pragma unroll 4
            for (int k = 2; k <= jpk - 1; k++) {
                const T tnk = A3(tn, i, j, k);
                T zwrm, zout, zwi_c = T(0);
                if (k == 2 || k == ikb) {
                    zwrm = T(0.5) * (tn_km1 + tnk);
                } else {
                    const T wm = WMASK(T, i, j, k);
                    zwi_c = wm;
                    zwrm = T(3) * wm * (tnk + tn_km1);
                }
                if (k == 2) zout = zwrm;
                else zout = zwrm - zwi_c / cpt_piv(cpt_tab<T>(), k - 1, ikb) * zout_prev;
                A3(zo, i, j, k) = zout;
                zout_prev = zout;
                tn_km1 = tnk;
            }
        }
    }
};
template <typename T, int NT>
KERNEL void k_fct_cpt_fwd(DevDomain d, const T *__restrict__ ptn, T *__restrict__ pzo)
{
    S_fct_cpt_fwd<T, NT>::run(vb_hw(), d, ptn, pzo);
}

// TRP: transports formed from un, vn, wn (1) or read (0); NH4 / NV4: order 4
// horizontally / vertically; TA0: the trend starts from zero (not read);
// NTH threads share the points of the tile's load rectangle
template <int TRP, typename T, int NH4, int NV4, int TA0, int TY, int NTH>
struct S_fct_fused {
    static constexpr int NTHR = NTH, WAVES = FuWaves<T, NTH>::value;
    __device__ __forceinline__ static void run(const VB &vb_, const DevDomain &d, double p2dt, int ntr, const T *__restrict__ pu, const T *__restrict__ pv, const T *__restrict__ pw, const T *__restrict__ ptb_, const T *__restrict__ ptn_, const T *__restrict__ pzo_, T *__restrict__ pta_)
    {
        using ZB = typename std::conditional<sizeof(T) == 4, float, double>::type;
        constexpr int TX = FU_TX;
        constexpr int PX = TX + 6, PY = TY + 6, LP = PX * PY;
        constexpr int NS = (LP + NTH - 1) / NTH;
        constexpr int NTL = TX * TY, NBR = (TX + 2) * (TY + 2), NZR = (TX + 4) * (TY + 4);
        const ZB zsent = sizeof(T) == 4 ? (ZB)INFINITY : (ZB)1.e+40;
        __shared__ T s_tb[LP], s_tn[LP];
        __shared__ T s_hu[LP], s_hv[LP];               // upstream fluxes, then the limited fluxes
        __shared__ ZB s_zup[3][LP], s_zdo[3][LP];      // bounds, level mod 3
        __shared__ T s_zwx[2][LP], s_zwy[2][LP];       // antidiffusive fluxes, level mod 2
        __shared__ T s_bup[2][LP], s_bdo[2][LP];       // betas, level mod 2

        const int jpi = d.jpi, jpj = d.jpj, jpk = d.jpk;
        const size_t n2 = d.n2, n3 = d.n3;
        const Metrics<T> &m = met<T>(d);
        const double zrtrn = 1.e-15;
        const T zp2dt = (T)p2dt, r1_6 = T(1) / T(6);

        // tile and tracer: the tracers of a tile are consecutive in the XCD-aware
        // order (workgroups b and b + 8), so they share the transports in one L2
        const unsigned nb = vb_.gx, b = vb_.x;
        const unsigned l = (nb & 7u) ? b : (b & 7u) * (nb >> 3) + (b >> 3);
        const int tr = (int)(l % (unsigned)ntr), tile = (int)(l / (unsigned)ntr);
        const int ntx = (jpi - 2 + TX - 1) / TX;
        const int i0 = 2 + (tile % ntx) * TX, j0 = 2 + (tile / ntx) * TY;
        const T *tb = ptb_ + (size_t)tr * n3, *tn = ptn_ + (size_t)tr * n3;
        const T *zo = NV4 ? pzo_ + (size_t)tr * n3 : nullptr;
        T *ta = pta_ + (size_t)tr * n3;
        const int tid = (int)threadIdx.x;

        // ---- this thread's points: slot r holds point q = tid + 256 r
        int sx[NS], sy[NS], sl[NS];        // tile coordinates, LDS index
        int gi[NS], gj[NS];
        uint32_t of[NS];                   // clamped in-plane offset
        bool vT[NS], vB[NS], vZ[NS], vP[NS], vL[NS];   // tile cell owned; ring<=1; ring<=2; face owner; loaded
        bool inD[NS], inI[NS];             // inside the domain; interior 2..jpim1 x 2..jpjm1
        bool fX[NS], fY[NS];               // owner of an antidiffusive x / y face the betas use
        uint32_t wt[NS], wv[NS];           // packed masks (T0|T1|U0|U1, V0|V1|F0|F1) of the column
        uint32_t wtw[NS], wte[NS], wvs[NS], wvn[NS];   // ... of the west / east / south / north neighbours
    #pragma unroll
        for (int r = 0; r < NS; r++) {
            const int q = tid + r * NTH;
            vL[r] = q < LP;
            int x = 0, y = 0;
            if (vL[r]) fu_xy<TX, TY>(q, x, y);
            sx[r] = x;
            sy[r] = y;
            sl[r] = (x + 3) + PX * (y + 3);
            gi[r] = i0 + x;
            gj[r] = j0 + y;
            inD[r] = vL[r] && gi[r] >= 1 && gi[r] <= jpi && gj[r] >= 1 && gj[r] <= jpj;
            inI[r] = vL[r] && gi[r] >= 2 && gi[r] <= jpi - 1 && gj[r] >= 2 && gj[r] <= jpj - 1;
            vT[r] = q < NTL && gi[r] <= jpi - 1 && gj[r] <= jpj - 1;
            vB[r] = q < NBR;
            vZ[r] = q < NZR;
            vP[r] = vL[r] && x <= TX + 1 && y <= TY + 1;
            fX[r] = vL[r] && x >= -2 && x <= TX && y >= -1 && y <= TY;
            fY[r] = vL[r] && x >= -1 && x <= TX && y >= -2 && y <= TY;
            const int ci = min(max(gi[r], 1), jpi), cj = min(max(gj[r], 1), jpj);
            of[r] = (uint32_t)(cj - 1) * (uint32_t)jpi + (uint32_t)(ci - 1);
            wt[r] = inD[r] ? d.ktu[of[r]] : 0u;
            wv[r] = inD[r] ? d.kvf[of[r]] : 0u;
            const uint32_t ow = (uint32_t)(cj - 1) * (uint32_t)jpi + (uint32_t)(max(ci - 1, 1) - 1);
            const uint32_t oe = (uint32_t)(cj - 1) * (uint32_t)jpi + (uint32_t)(min(ci + 1, jpi) - 1);
            const uint32_t os = (uint32_t)(max(cj - 1, 1) - 1) * (uint32_t)jpi + (uint32_t)(ci - 1);
            const uint32_t on = (uint32_t)(min(cj + 1, jpj) - 1) * (uint32_t)jpi + (uint32_t)(ci - 1);
            wtw[r] = d.ktu[ow];
            wte[r] = d.ktu[oe];
            wvs[r] = d.kvf[os];
            wvn[r] = d.kvf[on];
        }
        // mask of packed interval word w (shift SH) at level k
        auto pm = [](uint32_t w, int sh, int k) {
            const uint32_t v = w >> sh;
            return k >= (int)(v & 255u) && k <= (int)((v >> 8) & 255u);
        };
        auto wmk = [&](uint32_t w, int k) {   // w-mask: tmask(k) tmask(k-1)
            const int a = (int)(w & 255u), bb = (int)((w >> 8) & 255u);
            return k >= a && k <= bb && (k == 1 || k - 1 >= a);
        };

        // per-slot metrics
        T e2u[NS], e1v[NS], e1e2t[NS], r1e[NS];
        double e1e2t_d[NS];
        int ikb[NS];
    #pragma unroll
        for (int r = 0; r < NS; r++) {
            e2u[r] = m.e2u[of[r]];
            e1v[r] = m.e1v[of[r]];
            e1e2t[r] = m.e1e2t[of[r]];
            r1e[r] = m.r1_e1e2t[of[r]];
            e1e2t_d[r] = d.md.e1e2t[of[r]];
            ikb[r] = d.mbkt[of[r]];
        }

        // ---- loads of a level.  Every global load of the march is issued one
        // level (tb two) before the level that uses it and only moved between
        // registers in between: the wait counter of the loads is in order, so a
        // load issued in a stage and used at once would also wait for every load
        // of the next level in flight.  The raw values are loaded; the transports
        // are formed at their level (the same products, in the same order).
        // Levels are clamped into 1..jpk-1 for the transports, 1..jpk for e3t.
        auto kcl = [&](int k) { return min(max(k, 1), jpk - 1); };
        auto ld_raw = [&](const T *p, int r, int k) { return ldp(plane_ptr(p, n2, kcl(k)), of[r]); };

        // ---- registers carried down the march
        T tbc[NS], tbx[NS], tbxx[NS], tnc[NS], tnx[NS];
        T uc[NS], ux[NS], vc[NS], vx[NS], wc[NS], wx[NS], zoc[NS], zox[NS], tac[NS], tax[NS];
        T eux[NS], evx[NS];                 // e3u, e3v of the prefetched transport level (TRP)
        // e3t at the level in flight (it until the loads of level it are issued,
        // then it-1), it, it+1, it+2; and in FP64 for the FP32 form
        T e3x[NS], e3c[NS], e3p1[NS], e3p2[NS];
        double e3dx[NS], e3dc[NS], e3dp1[NS];
        T upzb[NS], hu[NS], hv[NS];
        T zwi0[NS], zwi1[NS];
        T zx0[NS], zx1[NS], zx2[NS], zy0[NS], zy1[NS], zy2[NS];
        T zz0[NS], zz1[NS], zz2[NS], xcar[NS];
        T tl0[NS], tl1[NS], tl2[NS];
        double lzc[NS];
    #pragma unroll
        for (int r = 0; r < NS; r++) {
            const int kt = jpk, k1 = max(jpk - 1, 1);
            tbc[r] = ldp(plane_ptr(tb, n2, kt), of[r]);
            tbx[r] = ldp(plane_ptr(tb, n2, k1), of[r]);
            tnc[r] = ldp(plane_ptr(tn, n2, kt), of[r]);
            tnx[r] = ldp(plane_ptr(tn, n2, k1), of[r]);
            ux[r] = ld_raw(pu, r, jpk - 1);
            vx[r] = ld_raw(pv, r, jpk - 1);
            wx[r] = r * NTH < NZR ? ld_raw(pw, r, jpk - 1) : T(0);
            eux[r] = TRP ? E3(m.e3u, of[r], kcl(jpk - 1)) : T(0);
            evx[r] = TRP ? E3(m.e3v, of[r], kcl(jpk - 1)) : T(0);
            e3x[r] = E3(m.e3t, of[r], kt);      // the first level of the march
            e3dx[r] = sizeof(T) == 8 ? 0.0 : E3(d.md.e3t, of[r], kt);
            e3c[r] = e3p1[r] = e3p2[r] = T(1);
            e3dc[r] = e3dp1[r] = 1.0;
            zox[r] = (NV4 && r * NTH < NBR) ? ldp(plane_ptr(zo, n2, k1), of[r]) : T(0);
            tax[r] = (!TA0 && r * NTH < NTL) ? ldp(plane_ptr((const T *)ta, n2, k1), of[r]) : T(0);
            uc[r] = vc[r] = wc[r] = zoc[r] = tac[r] = tbxx[r] = T(0);
            upzb[r] = hu[r] = hv[r] = T(0);
            zwi0[r] = zwi1[r] = zx0[r] = zx1[r] = zx2[r] = zy0[r] = zy1[r] = zy2[r] = T(0);
            zz0[r] = zz1[r] = zz2[r] = xcar[r] = tl0[r] = tl1[r] = tl2[r] = T(0);
            lzc[r] = 0.0;
        }

        for (int it = jpk; it >= -1; it--) {
            const int mB = it + 1, nC = it + 2;
            const bool doA = it >= 1 && it <= jpk - 1, doB = mB >= 1 && mB <= jpk - 1;
            const bool doC = nC >= 1 && nC <= jpk - 1;
            // ---- the planes of level it; loads of the levels the next iterations need
            if (it >= 1) {
    #pragma unroll
                for (int r = 0; r < NS; r++) {
                    if (!vL[r]) continue;
                    s_tb[sl[r]] = tbc[r];
                    s_tn[sl[r]] = tnc[r];
                }
            }
            {
                const int k2 = max(it - 2, 1), k1 = max(it - 1, 1);
    #pragma unroll
                for (int r = 0; r < NS; r++) {
                    // level it from the registers loaded one level earlier
                    if (TRP) {
                        uc[r] = e2u[r] * eux[r] * ux[r];
                        vc[r] = e1v[r] * evx[r] * vx[r];
                        wc[r] = e1e2t[r] * wx[r];
                    } else {
                        uc[r] = ux[r];
                        vc[r] = vx[r];
                        wc[r] = wx[r];
                    }
                    zoc[r] = zox[r];
                    tac[r] = tax[r];
                    // loads of the next levels
                    tbxx[r] = ldp(plane_ptr(tb, n2, k2), of[r]);
                    ux[r] = ld_raw(pu, r, it - 1);
                    vx[r] = ld_raw(pv, r, it - 1);
                    if (TRP) {
                        eux[r] = E3(m.e3u, of[r], kcl(it - 1));
                        evx[r] = E3(m.e3v, of[r], kcl(it - 1));
                    }
                    if (r * NTH < NZR) wx[r] = ld_raw(pw, r, it - 1);
                    if (NV4 && r * NTH < NBR) zox[r] = ldp(plane_ptr(zo, n2, k1), of[r]);
                    // the trend only at the tile's own cells (other tiles write theirs)
                    if (!TA0 && r * NTH < NTL) tax[r] = ldp(plane_ptr((const T *)ta, n2, k1), of[r]);
                    e3c[r] = e3x[r];
                    e3x[r] = E3(m.e3t, of[r], k1);
                    if (sizeof(T) == 4) {
                        e3dc[r] = e3dx[r];
                        e3dx[r] = E3(d.md.e3t, of[r], k1);
                    }
                }
            }
            __syncthreads();
            // ---- A1: upstream fluxes of the east and north faces (lines 123-145)
            if (doA) {
    #pragma unroll
                for (int r = 0; r < NS; r++) {
                    if (!vP[r]) continue;
                    const T u = uc[r], v = vc[r], tbo = tbc[r];
                    hu[r] = T(0.5) * ((u + fabs(u)) * tbo + (u - fabs(u)) * s_tb[sl[r] + 1]);
                    hv[r] = T(0.5) * ((v + fabs(v)) * tbo + (v - fabs(v)) * s_tb[sl[r] + PX]);
                    s_hu[sl[r]] = hu[r];
                    s_hv[sl[r]] = hv[r];
                }
            }
            __syncthreads();
            // ---- A2: first guess zwi and low-order trend (158-170), the bounds
            // (361-364), the antidiffusive fluxes (180-278)
            if (it >= 1) {
                const int sb3 = it % 3, sb2 = it & 1;
    #pragma unroll
                for (int r = 0; r < NS; r++) {
                    if (!vZ[r]) continue;
                    const bool wet = pm(wt[r], 0, it);
                    T zwi = T(0), upzt = T(0);
                    if (doA) {
                        // vertical upstream flux of the top face: the linear free
                        // surface flux at level 1, w-masked upstream below
                        const T w = wc[r];
                        if (it == 1) upzt = w * tbc[r];
                        else upzt = T(0.5) * ((w + fabs(w)) * tbc[r] + (w - fabs(w)) * tbx[r]) *
                                    (wmk(wt[r], it) ? T(1) : T(0));
                        if (inI[r]) {
                            const T ztra = -((hu[r] - s_hu[sl[r] - 1] + hv[r] - s_hv[sl[r] - PX] + upzt - upzb[r]) *
                                             r1e[r]);
                            const T e3 = e3c[r];
                            const T tm = wet ? T(1) : T(0);
                            zwi = (e3 * tbc[r] + zp2dt * ztra) / e3 * tm;
                            if (vT[r]) tl0[r] = (TA0 ? T(0) : tac[r]) + ztra / e3 * tm;
                        }
                    }
                    // bounds of the point (zero off the domain, as never used)
                    ZB bu = 0, bd = 0;
                    if (inD[r]) {
                        if constexpr (sizeof(T) == 4) {
                            const ZB a = (ZB)tbc[r], bq = (ZB)zwi;
                            bu = wet ? (a > bq ? a : bq) : -zsent;
                            bd = wet ? (a < bq ? a : bq) : zsent;
                        } else {
                            const double tm = wet ? 1.0 : 0.0;
                            const double au = (double)tbc[r] * tm - zsent * (1.0 - tm);
                            const double bu2 = (double)zwi * tm - zsent * (1.0 - tm);
                            bu = au > bu2 ? au : bu2;
                            const double ad = (double)tbc[r] * tm + zsent * (1.0 - tm);
                            const double bd2 = (double)zwi * tm + zsent * (1.0 - tm);
                            bd = ad < bd2 ? ad : bd2;
                        }
                    }
                    s_zup[sb3][sl[r]] = bu;
                    s_zdo[sb3][sl[r]] = bd;
                    zwi0[r] = zwi;
                    upzb[r] = upzt;
                    if (!doA) {
                        zz0[r] = T(0);     // level jpk: no flux through the bottom
                        continue;
                    }
                    // antidiffusive x and y fluxes of the point's faces (zero on
                    // the closed ring, the lbc_lnk of line 280)
                    if (fX[r]) {
                        T f = T(0);
                        if (gi[r] >= 2 && gi[r] <= jpi - 1 && gj[r] >= 2 && gj[r] <= jpj - 1) {
                            const int c = sl[r];
                            const T c2u = s_tn[c] + s_tn[c + 1];
                            if (NH4) {
                                // zltu (lines 195-209) of the cell and of its east neighbour
                                const T um_w = pm(wtw[r], 16, it) ? T(1) : T(0), um_c = pm(wt[r], 16, it) ? T(1) : T(0);
                                const T um_e = pm(wte[r], 16, it) ? T(1) : T(0);
                                const T za = (s_tn[c + 1] - s_tn[c]) * um_c, zb = (s_tn[c] - s_tn[c - 1]) * um_w;
                                const T zl0 = (za + zb) * r1_6;
                                const T zc = (s_tn[c + 2] - s_tn[c + 1]) * um_e, zd = (s_tn[c + 1] - s_tn[c]) * um_c;
                                const T zl1 = gi[r] + 1 <= jpi - 1 ? (zc + zd) * r1_6 : T(0);
                                f = T(0.5) * uc[r] * (c2u + zl0 - zl1) - hu[r];
                            } else {
                                f = T(0.5) * uc[r] * c2u - hu[r];
                            }
                        }
                        zx0[r] = f;
                        s_zwx[sb2][sl[r]] = f;
                    }
                    if (fY[r]) {
                        T f = T(0);
                        if (gi[r] >= 2 && gi[r] <= jpi - 1 && gj[r] >= 2 && gj[r] <= jpj - 1) {
                            const int c = sl[r];
                            const T c2v = s_tn[c] + s_tn[c + PX];
                            if (NH4) {
                                const T vm_s = pm(wvs[r], 0, it) ? T(1) : T(0), vm_c = pm(wv[r], 0, it) ? T(1) : T(0);
                                const T vm_n = pm(wvn[r], 0, it) ? T(1) : T(0);
                                const T za = (s_tn[c + PX] - s_tn[c]) * vm_c, zb = (s_tn[c] - s_tn[c - PX]) * vm_s;
                                const T zl0 = (za + zb) * r1_6;
                                const T zc = (s_tn[c + 2 * PX] - s_tn[c + PX]) * vm_n, zd = (s_tn[c + PX] - s_tn[c]) * vm_c;
                                const T zl1 = gj[r] + 1 <= jpj - 1 ? (zc + zd) * r1_6 : T(0);
                                f = T(0.5) * vc[r] * (c2v + zl0 - zl1) - hv[r];
                            } else {
                                f = T(0.5) * vc[r] * c2v - hv[r];
                            }
                        }
                        zy0[r] = f;
                        s_zwy[sb2][sl[r]] = f;
                    }
                    // antidiffusive flux of the top w-face (lines 255-278), zero
                    // at the surface and off the interior columns
                    if (vB[r]) {
                        T f = T(0);
                        if (NV4 && it <= jpk - 1 && it >= 2) {
                            const int kb = ikb[r];
                            T x;
                            if (it == jpk - 1) {
                                x = zoc[r] / cpt_piv(cpt_tab<T>(), it, kb);
                            } else {
                                const T zws = (it == 2 || it == kb) ? T(0) : (wmk(wt[r], it) ? T(1) : T(0));
                                x = (zoc[r] - zws * xcar[r]) / cpt_piv(cpt_tab<T>(), it, kb);
                            }
                            xcar[r] = x;
                            if (inI[r]) f = (wc[r] * x - upzt) * (wmk(wt[r], it) ? T(1) : T(0));
                        } else if (!NV4 && it >= 2 && inI[r]) {
                            f = (wc[r] * T(0.5) * (tnc[r] + tnx[r]) - upzt) * (wmk(wt[r], it) ? T(1) : T(0));
                        }
                        zz0[r] = f;
                    }
                }
            }
            __syncthreads();
            // ---- B: betas of level mB (366-399), zero off the interior and at jpk
            if (doB || mB == jpk) {
                const int sc = mB % 3, sa = (mB > 1 ? mB - 1 : 1) % 3, sn = (mB + 1) % 3, s2 = mB & 1;
    #pragma unroll
                for (int r = 0; r < NS; r++) {
                    if (!vB[r]) continue;
                    T bu_t = T(0), bd_t = T(0);
                    if (doB && inI[r]) {
                        const int c = sl[r];
                        ZB zup = s_zup[sc][c];
                        zup = fmax2(zup, s_zup[sc][c - 1]);
                        zup = fmax2(zup, s_zup[sc][c + 1]);
                        zup = fmax2(zup, s_zup[sc][c - PX]);
                        zup = fmax2(zup, s_zup[sc][c + PX]);
                        zup = fmax2(zup, s_zup[sa][c]);
                        zup = fmax2(zup, s_zup[sn][c]);
                        ZB zdo = s_zdo[sc][c];
                        zdo = fmin2(zdo, s_zdo[sc][c - 1]);
                        zdo = fmin2(zdo, s_zdo[sc][c + 1]);
                        zdo = fmin2(zdo, s_zdo[sc][c - PX]);
                        zdo = fmin2(zdo, s_zdo[sc][c + PX]);
                        zdo = fmin2(zdo, s_zdo[sa][c]);
                        zdo = fmin2(zdo, s_zdo[sn][c]);
                        const double a_im1 = s_zwx[s2][c - 1], a_i = zx1[r], b_jm1 = s_zwy[s2][c - PX], b_j = zy1[r];
                        const double c_k = zz1[r], c_kp1 = zz2[r];
                        const double zpos = fmax2(0.0, a_im1) - fmin2(0.0, a_i) + fmax2(0.0, b_jm1) - fmin2(0.0, b_j) +
                                            fmax2(0.0, c_kp1) - fmin2(0.0, c_k);
                        const double zneg = fmax2(0.0, a_i) - fmin2(0.0, a_im1) + fmax2(0.0, b_j) - fmin2(0.0, b_jm1) +
                                            fmax2(0.0, c_k) - fmin2(0.0, c_kp1);
                        const double e3d = sizeof(T) == 8 ? (double)e3p1[r] : e3dp1[r];   // level mB
                        const double zbt = e1e2t_d[r] * e3d / p2dt;
                        const double paft_k = zwi1[r];
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
                    s_bup[s2][sl[r]] = bu_t;
                    s_bdo[s2][sl[r]] = bd_t;
                }
            }
            __syncthreads();
            // ---- C: limited face fluxes of level nC (404-422), then the corrected
            // divergence of the tile's cells (288-297)
            const int s2c = nC & 1, s2p = (nC - 1) & 1;
            if (doC) {
    #pragma unroll
                for (int r = 0; r < NS; r++) {
                    if (!vB[r]) continue;
                    const int c = sl[r];
                    if (fX[r] && sx[r] <= TX - 1)
                        s_hu[c] = (T)lim_face<T>(zx2[r], s_bdo[s2c][c], s_bup[s2c][c + 1], s_bup[s2c][c],
                                                 s_bdo[s2c][c + 1]);
                    if (fY[r] && sy[r] <= TY - 1)
                        s_hv[c] = (T)lim_face<T>(zy2[r], s_bdo[s2c][c], s_bup[s2c][c + PX], s_bup[s2c][c],
                                                 s_bdo[s2c][c + PX]);
                }
            }
            __syncthreads();
            if (doC) {
    #pragma unroll
                for (int r = 0; r < NS; r++) {
                    if (!vT[r]) continue;
                    const int c = sl[r];
                    const double bu_c = s_bup[s2c][c], bd_c = s_bdo[s2c][c];
                    const T lx_i = s_hu[c], lx_im1 = gi[r] - 1 >= 2 ? s_hu[c - 1] : T(0);
                    const T ly_j = s_hv[c], ly_jm1 = gj[r] - 1 >= 2 ? s_hv[c - PX] : T(0);
                    double lz_n, lz_np1;
                    if (nC == 1) lz_n = (double)zz2[r];   // the surface flux, not limited
                    else lz_n = lim_face<T>(zz2[r], bd_c, s_bup[s2p][c], bu_c, s_bdo[s2p][c]);
                    if (nC == jpk - 1) lz_np1 = lim_face<T>(T(0), 0.0, bu_c, 0.0, bd_c);
                    else lz_np1 = lzc[r];
                    const T e3 = e3p2[r];     // level nC
                    A3(ta, gi[r], gj[r], nC) =
                        tl2[r] - (lx_i - lx_im1 + ly_j - ly_jm1 + (T)lz_n - (T)lz_np1) * r1e[r] / e3;
                    lzc[r] = lz_n;
                }
            }
            // ---- down one level
    #pragma unroll
            for (int r = 0; r < NS; r++) {
                tbc[r] = tbx[r];
                tbx[r] = tbxx[r];
                tnc[r] = tnx[r];
                zwi1[r] = zwi0[r];
                zx2[r] = zx1[r];
                zx1[r] = zx0[r];
                zy2[r] = zy1[r];
                zy1[r] = zy0[r];
                zz2[r] = zz1[r];
                zz1[r] = zz0[r];
                tl2[r] = tl1[r];
                tl1[r] = tl0[r];
                e3p2[r] = e3p1[r];
                e3p1[r] = e3c[r];
                e3dp1[r] = e3dc[r];
            }
            {
                const int k2 = max(it - 2, 1);
    #pragma unroll
                for (int r = 0; r < NS; r++) tnx[r] = ldp(plane_ptr(tn, n2, k2), of[r]);
            }
        }
    }
};
template <int TRP, typename T, int NH4, int NV4, int TA0, int TY, int NTH>
__global__ __launch_bounds__(NTH, (FuWaves<T, NTH>::value)) void k_fct_fused(
    DevDomain d, double p2dt, int ntr, const T *__restrict__ pu, const T *__restrict__ pv,
    const T *__restrict__ pw, const T *__restrict__ ptb_, const T *__restrict__ ptn_, const T *__restrict__ pzo_,
    T *__restrict__ pta_)
{
    S_fct_fused<TRP, T, NH4, NV4, TA0, TY, NTH>::run(vb_hw(), d, p2dt, ntr, pu, pv, pw, ptb_, ptn_, pzo_, pta_);
}

// fct_fused switch value 1: 256 threads per workgroup; 2: 512
template <int TRP, typename T, int NH4, int NV4, int TA0>
static int fu_launch(const DevDomain &d, double p2dt, int nt, const T *pu, const T *pv,
                     const T *pw, const T *ptb, const T *ptn, const T *pzo, T *pta, const char *nm)
{
    constexpr int TY = FuTY<T>::value;
    const int ntx = (d.jpi - 2 + FU_TX - 1) / FU_TX, nty = (d.jpj - 2 + TY - 1) / TY;
    Grid2 g;
    g.grid = dim3((unsigned)(ntx * nty * nt), 1, 1);
    if (nh::tune(nh::T_FCT_FUSED) == 2) {
        g.block = dim3(512, 1, 1);
        LAUNCHB(nm, (k_fct_fused<TRP, T, NH4, NV4, TA0, TY, 512>), (S_fct_fused<TRP, T, NH4, NV4, TA0, TY, 512>), g, d, p2dt, nt, pu, pv, pw, ptb, ptn, pzo,
               pta);
    } else {
        g.block = dim3(256, 1, 1);
        LAUNCHB(nm, (k_fct_fused<TRP, T, NH4, NV4, TA0, TY, 256>), (S_fct_fused<TRP, T, NH4, NV4, TA0, TY, 256>), g, d, p2dt, nt, pu, pv, pw, ptb, ptn, pzo,
               pta);
    }
    return 0;
}

template <typename T>
int launch_fct_cpt_fwd(const DevDomain &d, int nt, const T *ptn, T *pzo, const char *nm)
{
    Grid2 gi = grid2(d.jpi - 2, d.jpj - 2);
    if (nt == 2) LAUNCHB(nm, (k_fct_cpt_fwd<T, 2>), (S_fct_cpt_fwd<T, 2>), gi, d, ptn, pzo);
    else LAUNCHB(nm, (k_fct_cpt_fwd<T, 1>), (S_fct_cpt_fwd<T, 1>), gi, d, ptn, pzo);
    return 0;
}
template int launch_fct_cpt_fwd<double>(const DevDomain &, int, const double *, double *,
                                        const char *);
template int launch_fct_cpt_fwd<float>(const DevDomain &, int, const float *, float *,
                                       const char *);

// tra_adv_fct of nt (1 or 2) tracers on a closed domain; pzo: scratch for
// the pre-pass of the order-4 vertical scheme (nt arrays of n3)
template <int TRP, typename T>
int launch_fct_fused(const DevDomain &d, double p2dt, int nt, const T *pu, const T *pv,
                     const T *pw, const T *ptb, const T *ptn, T *pta, int kn_fct_h, int kn_fct_v, int ta0, T *pzo,
                     const char *nm)
{
    if (kn_fct_v == 4) {
        Grid2 gi = grid2(d.jpi - 2, d.jpj - 2);
        if (nt == 2) LAUNCHB(nm, (k_fct_cpt_fwd<T, 2>), (S_fct_cpt_fwd<T, 2>), gi, d, ptn, pzo);
        else LAUNCHB(nm, (k_fct_cpt_fwd<T, 1>), (S_fct_cpt_fwd<T, 1>), gi, d, ptn, pzo);
    }
#define FU_CASE(H, V, Z) \
    if (kn_fct_h == (H ? 4 : 2) && kn_fct_v == (V ? 4 : 2) && !!ta0 == !!Z) \
        return fu_launch<TRP, T, H, V, Z>(d, p2dt, nt, pu, pv, pw, ptb, ptn, pzo, pta, nm);
    FU_CASE(1, 1, 1) FU_CASE(1, 1, 0) FU_CASE(1, 0, 1) FU_CASE(1, 0, 0)
    FU_CASE(0, 1, 1) FU_CASE(0, 1, 0) FU_CASE(0, 0, 1) FU_CASE(0, 0, 0)
#undef FU_CASE
    return nh::set_error(NEMO_HIP_ERR_ARG, "fct: orders %d/%d", kn_fct_h, kn_fct_v);
}
#define FU_INST(TRP, T)                                                                                           \
    template int launch_fct_fused<TRP, T>(const DevDomain &, double, int, const T *,                             \
                                          const T *, const T *, const T *, const T *, T *, int, int, int, T *,    \
                                          const char *);
FU_INST(0, double)
FU_INST(1, double)
FU_INST(0, float)
FU_INST(1, float)
