// dom_lbc.hip -- masks / k-indices expansion (DOM/dommsk.F90, DOM/domzgr.F90)
// and the serial lateral boundary fill (LBC/lbc_lnk_generic.h90).
#include "common.hpp"
#include "kernels.hpp"

// ---------------------------------------------------------------------------
// dom_msk: expand the per-column level intervals to NEMO's REAL 0/1 masks.
// One thread per (i,j) column, k-marched, coalesced along i.
// ---------------------------------------------------------------------------
KERNEL void k_dom_msk(DevDomain d, double * __restrict__ tmask, double * __restrict__ umask, double * __restrict__ vmask, double * __restrict__ fmask, double * __restrict__ wmask,
                          double * __restrict__ wumask, double * __restrict__ wvmask, double * __restrict__ ssmask, double * __restrict__ ssumask, double * __restrict__ ssvmask)
{
    THREAD_IJ(1, 1);
    if (i > d.jpi || j > d.jpj) return;
    const size_t ij = IJ(i, j);
    for (int k = 1; k <= d.jpk; k++) {
        if (tmask) A3(tmask, i, j, k) = kmask<double>(d.kT, ij, k);
        if (umask) A3(umask, i, j, k) = kmask<double>(d.kU, ij, k);
        if (vmask) A3(vmask, i, j, k) = kmask<double>(d.kV, ij, k);
        if (fmask) A3(fmask, i, j, k) = kmask<double>(d.kF, ij, k);
        if (wmask) A3(wmask, i, j, k) = wkmask<double>(d.kT, ij, k);
        if (wumask) A3(wumask, i, j, k) = wkmask<double>(d.kU, ij, k);
        if (wvmask) A3(wvmask, i, j, k) = wkmask<double>(d.kV, ij, k);
    }
    if (ssmask) ssmask[ij] = d.ssmask[ij];
    if (ssumask) ssumask[ij] = d.ssumask[ij];
    if (ssvmask) ssvmask[ij] = d.ssvmask[ij];
}

int launch_dom_msk(double *tmask, double *umask, double *vmask, double *fmask, double *wmask, double *wumask,
                   double *wvmask, double *ssmask, double *ssumask, double *ssvmask)
{
    const DevDomain &d = nh::dom();
    Grid2 g = grid2(d.jpi, d.jpj);
    LAUNCH("dom_msk", k_dom_msk, g, d, tmask, umask, vmask, fmask, wmask, wumask, wvmask, ssmask, ssumask, ssvmask);
    return 0;
}

// ---------------------------------------------------------------------------
// lbc_lnk on (jpi,jpj,nk): one block per level; the east-west fill completes
// (barrier) before the north-south fill copies whole rows.
// ---------------------------------------------------------------------------
template <typename T>
struct S_lbc {
    static constexpr int NTHR = NH_BLOCK, WAVES = 1;
    __device__ __forceinline__ static void run(const VB &vb_, const DevDomain &d, T * __restrict__ p, int nk, char nat, T pval)
    {
        const int k = (int)vb_.x + 1;
        if (k > nk) return;
        for (int j = (int)threadIdx.x + 1; j <= d.jpj; j += blockDim.x) {
            if (d.iper) {
                T a = A3(p, d.jpi - 1, j, k), b = A3(p, 2, j, k);
                A3(p, 1, j, k) = a;
                A3(p, d.jpi, j, k) = b;
            } else {
                if (nat != 'F') A3(p, 1, j, k) = pval;
                A3(p, d.jpi, j, k) = pval;
            }
        }
        __syncthreads();
        for (int i = (int)threadIdx.x + 1; i <= d.jpi; i += blockDim.x) {
            if (d.jper) {
                T a = A3(p, i, d.jpj - 1, k), b = A3(p, i, 2, k);
                A3(p, i, 1, k) = a;
                A3(p, i, d.jpj, k) = b;
            } else {
                if (nat != 'F') A3(p, i, 1, k) = pval;
                A3(p, i, d.jpj, k) = pval;
            }
        }
    }
};
template <typename T>
KERNEL void k_lbc(DevDomain d, T * __restrict__ p, int nk, char nat, T pval)
{
    S_lbc<T>::run(vb_hw(), d, p, nk, nat, pval);
}

template <typename T>
int launch_lbc(T *p, int nk, char nat, T pval)
{
    const DevDomain &d = nh::dom();
    Grid2 g;
    g.grid = dim3((unsigned)nk, 1, 1);
    g.block = dim3(NH_BLOCK, 1, 1);
    LAUNCHB("lbc_lnk", k_lbc<T>, (S_lbc<T>), g, d, p, nk, nat, pval);
    return 0;
}
template int launch_lbc<double>(double *, int, char, double);
template int launch_lbc<float>(float *, int, char, float);
