// gsum.hip -- compensated (double-double) global sums on gfx950.
//
//   glob_sum, glob_sum_full   lib_fortran.F90:300-332 (DDPDD),
//                             lib_fortran_generic.h90 (FUNCTION_GLOBSUM)
//   dia_hsb                   DIA/diahsb.F90:60-239, dia_hsb_rst :278-307
//
// The reference accumulates ptab * mask over (k, j, i) in loop order into a
// double-double (sum, error) pair with DDPDD, Knuth's two-sum with the error
// terms carried (He & Ding 2001), and returns the sum part.  Here every
// thread accumulates its elements with the same DDPDD, the pairs of a wave
// are combined by lane shuffles, the waves of a block in LDS, and the blocks'
// partial pairs by one final block: the same compensated sum in a different
// order.  Its sum part is the double nearest the exact sum unless the sum
// cancels to within about 1e-16 of the sum of magnitudes, as with the
// reference's sequential order, so both agree with an exact (math.fsum) sum
// of the same terms.  The library is built with -ffp-contract=off, which the
// two-sum needs (a fused multiply-add would not change it, but a contracted
// error term would).
//
// Masks (DOM/dommsk.F90:213-238), one domain, halo width 1: tmask_h is 0 on
// the outer ring and 1 inside; tmask_i = ssmask * tmask_h.
#include "common.hpp"
#include "kernels.hpp"

// DDPDD: (br, bi) <- (ar, ai) + (br, bi), lib_fortran.F90:322-328
__device__ __forceinline__ void ddpdd(double ar, double ai, double &br, double &bi)
{
    const double zt1 = ar + br;
    const double zerr = zt1 - ar;
    const double zt2 = ((br - zerr) + (ar - (zt1 - zerr))) + ai + bi;
    br = zt1 + zt2;
    bi = zt2 - ((zt1 + zt2) - zt1);
}

// combine the pairs of the 64 lanes of a wave (every lane ends with the total)
__device__ __forceinline__ void dd_wave(double &h, double &l)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        const double oh = __shfl_xor(h, off), ol = __shfl_xor(l, off);
        ddpdd(oh, ol, h, l);
    }
}

// tmask_h (kind 1) or tmask_i (kind 0) of column ij
__device__ __forceinline__ double gs_mask(const DevDomain &d, uint32_t ij, int kind)
{
    const int i = (int)(ij % (uint32_t)d.jpi) + 1, j = (int)(ij / (uint32_t)d.jpi) + 1;
    const double h = is_interior(d, i, j) ? 1.0 : 0.0;
    return kind ? h : SSMASK(ij) * h;
}

// block reduction of NS pairs into part[blockIdx.x * NS + s] (hi, lo)
template <int NS>
__device__ __forceinline__ void dd_block(double (&h)[NS], double (&l)[NS], double2 *__restrict__ part)
{
    __shared__ double sh[NH_BLOCK / 64][NS], sl[NH_BLOCK / 64][NS];
    const int lane = (int)threadIdx.x & 63, w = (int)threadIdx.x >> 6;
#pragma unroll
    for (int s = 0; s < NS; s++) {
        dd_wave(h[s], l[s]);
        if (lane == 0) {
            sh[w][s] = h[s];
            sl[w][s] = l[s];
        }
    }
    __syncthreads();
    if (threadIdx.x < NS) {
        const int s = (int)threadIdx.x;
        double bh = sh[0][s], bl = sl[0][s];
        for (int q = 1; q < NH_BLOCK / 64; q++) ddpdd(sh[q][s], sl[q][s], bh, bl);
        part[(size_t)blockIdx.x * NS + s] = make_double2(bh, bl);
    }
}

// the NS sums over nb blocks' partial pairs, one block; out[s] = sum part
template <int NS>
KERNEL void k_dd_final(const double2 *__restrict__ part, int nb, double *__restrict__ out)
{
    double h[NS], l[NS];
#pragma unroll
    for (int s = 0; s < NS; s++) h[s] = l[s] = 0.0;
    for (int b = (int)threadIdx.x; b < nb; b += NH_BLOCK)
#pragma unroll
        for (int s = 0; s < NS; s++) {
            const double2 v = part[(size_t)b * NS + s];
            ddpdd(v.x, v.y, h[s], l[s]);
        }
    // one block: its pairs go after the partials, and thread s, which
    // wrote pair s, passes on its sum part
    double2 *tot = (double2 *)(part + (size_t)nb * NS);
    dd_block<NS>(h, l, tot);
    if (threadIdx.x < NS) out[threadIdx.x] = tot[threadIdx.x].x;
}

// glob_sum of ptab(:,:,1:nk) * mask, ptab in the storage type T
template <typename T>
KERNEL void k_gsum_part(DevDomain d, const T *__restrict__ p, int nk, int kind, double2 *__restrict__ part)
{
    double h[1] = {0.0}, l[1] = {0.0};
    const uint32_t n2 = (uint32_t)d.n2, n = n2 * (uint32_t)nk;
    const uint32_t stride = gridDim.x * NH_BLOCK;
    for (uint32_t q = blockIdx.x * NH_BLOCK + threadIdx.x; q < n; q += stride) {
        const double ztmp = (double)p[q] * gs_mask(d, q % n2, kind);
        ddpdd(ztmp, 0.0, h[0], l[0]);
    }
    dd_block<1>(h, l, part);
}

#define GSUM_NB 1024
// the partial pairs of GSUM_NB blocks and the total, in doubles
#define GSUM_SCRATCH_DOUBLES (2 * (GSUM_NB + 1))

template <typename T>
int launch_glob_sum(const T *p, int nk, int kind, double *scratch, double *out)
{
    const DevDomain &d = nh::dom();
    Grid2 g;
    g.block = dim3(NH_BLOCK, 1, 1);
    const size_t n = d.n2 * (size_t)nk;
    const size_t nb = (n + NH_BLOCK * 8 - 1) / (NH_BLOCK * 8);
    g.grid = dim3((unsigned)(nb < GSUM_NB ? (nb > 0 ? nb : 1) : GSUM_NB), 1, 1);
    double2 *part = (double2 *)scratch;
    LAUNCH("glob_sum", k_gsum_part<T>, g, d, p, nk, kind, part);
    Grid2 g1;
    g1.block = dim3(NH_BLOCK, 1, 1);
    g1.grid = dim3(1, 1, 1);
    LAUNCH("glob_sum", k_dd_final<1>, g1, (const double2 *)part, (int)g.grid.x, out);
    return 0;
}
template int launch_glob_sum<double>(const double *, int, int, double *, double *);
template int launch_glob_sum<float>(const float *, int, int, double *, double *);
size_t glob_sum_scratch_bytes() { return GSUM_SCRATCH_DOUBLES * sizeof(double); }

// ---------------------------------------------------------------------------
// the glob_sum that allocates nothing (nemo_hip_glob_sum_noalloc): its
// partial pairs live in this file-scope device array, 16 KiB per GPU, whose
// device address nemo_hip_init takes once per device, with that device
// current.  Host code uses only that address: the array's name is a host
// stand-in.  Calls on one GPU share the array, so calls that overlap on one
// GPU race; nothing in it needs clearing (every launch writes the partials
// that the final block then reads).
// ---------------------------------------------------------------------------
static __device__ __align__(256) double scratch[GSUM_SCRATCH_DOUBLES];
#define GSUM_MAXDEV 64
static double *g_scratch_dev[GSUM_MAXDEV];

int glob_sum_scratch_init(int device)
{
    if (device < 0 || device >= GSUM_MAXDEV)
        return nh::set_error(NEMO_HIP_ERR_ARG, "glob_sum: device %d beyond %d", device, GSUM_MAXDEV);
    if (g_scratch_dev[device]) return 0;
    void *ptr = nullptr;
    HIPCHK(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    g_scratch_dev[device] = (double *)ptr;
    return 0;
}
double *glob_sum_static_scratch(int device)
{
    return (device >= 0 && device < GSUM_MAXDEV) ? g_scratch_dev[device] : nullptr;
}

// ---------------------------------------------------------------------------
// dia_hsb, DIA/diahsb.F90:88-189 (linear free surface, no ice-shelf cavity,
// no runoff, no geothermal flux): the thirteen global sums of one call.
//   0  -(emp - rnf) surf                   glob_sum       (:93)
//   1  sbc_tsc(T) surf, 2 sbc_tsc(S) surf  glob_sum       (:94-95)
//   3  qsr surf (only with ln_traqsr)      glob_sum       (:102)
//   4  surf wn(1) tsb(1,T), 5 (S)          glob_sum       (:115-119)
//   6  surf sshn - surf_ini ssh_ini        glob_sum_full  (:137)
//   7  surf (tsn(1,T) sshn - ssh_hc_loc_ini), 8 (S)  glob_sum_full (:149-153)
//   9  sum_k (surf e3t_n - surf_ini e3t_ini) tmask            (:158-159)
//  10  sum_k (surf e3t_n tsn(T) - surf_ini hc_loc_ini) tmask  (:161-163)
//  11  sum_k (surf e3t_n tsn(S) - surf_ini sc_loc_ini) tmask  (:165-167)
//  12  sum_k surf e3t_n tmask                                 (:187-189)
// surf = e1e2t tmask_i (dia_hsb_init, :411).  Every 3D sum runs over levels
// 1..jpkm1.  The reference first multiplies tsn and tsb by tmask in place
// (:88-89); the sums take tmask-weighted terms wherever those values enter,
// so the sums are the same (for finite values) and the inputs stay as given.
// All arithmetic is FP64, also from FP32 storage (the precision form keeps
// these sums in FP64).
// ---------------------------------------------------------------------------
#define HSB_NS 13
template <typename T>
KERNEL void k_hsb_part(DevDomain d, int ln_traqsr, const double *__restrict__ emp, const double *__restrict__ rnf,
                       const double *__restrict__ sbc, const double *__restrict__ qsr,
                       const double *__restrict__ sshn, const T *__restrict__ tsn, const T *__restrict__ tsb,
                       const T *__restrict__ wn, const double *__restrict__ surf_ini,
                       const double *__restrict__ ssh_ini, const double *__restrict__ ssh_hc,
                       const double *__restrict__ ssh_sc, const double *__restrict__ e3t_ini,
                       const double *__restrict__ hc_ini, const double *__restrict__ sc_ini,
                       double2 *__restrict__ part)
{
    double h[HSB_NS], l[HSB_NS];
#pragma unroll
    for (int s = 0; s < HSB_NS; s++) h[s] = l[s] = 0.0;
    const uint32_t n2 = (uint32_t)d.n2;
    const size_t n3 = d.n3;
    const uint32_t stride = gridDim.x * NH_BLOCK;
    for (uint32_t ij = blockIdx.x * NH_BLOCK + threadIdx.x; ij < n2; ij += stride) {
        const double mi = gs_mask(d, ij, 0), mh = gs_mask(d, ij, 1);
        const double surf = d.md.e1e2t[ij] * mi;
        const double sn = sshn[ij];
        ddpdd(-(emp[ij] - rnf[ij]) * surf * mi, 0.0, h[0], l[0]);
        ddpdd(sbc[ij] * surf * mi, 0.0, h[1], l[1]);
        ddpdd(sbc[n2 + ij] * surf * mi, 0.0, h[2], l[2]);
        if (ln_traqsr) ddpdd(qsr[ij] * surf * mi, 0.0, h[3], l[3]);
        const double t1 = pmask<double, 0>(d.ktu, ij, 1);
        const double w1 = (double)wn[ij];
        ddpdd(surf * w1 * ((double)tsb[ij] * t1) * mi, 0.0, h[4], l[4]);
        ddpdd(surf * w1 * ((double)tsb[n3 + ij] * t1) * mi, 0.0, h[5], l[5]);
        const double sfi = surf_ini[ij];
        ddpdd((surf * sn - sfi * ssh_ini[ij]) * mh, 0.0, h[6], l[6]);
        ddpdd(surf * ((double)tsn[ij] * t1 * sn - ssh_hc[ij]) * mh, 0.0, h[7], l[7]);
        ddpdd(surf * ((double)tsn[n3 + ij] * t1 * sn - ssh_sc[ij]) * mh, 0.0, h[8], l[8]);
        for (int k = 1; k <= d.jpk - 1; k++) {
            const double tm = pmask<double, 0>(d.ktu, ij, k);
            const size_t p = (size_t)(k - 1) * n2 + ij;
            const double e3 = E3(d.md.e3t, ij, k);
            const double se3 = surf * e3;
            const double tt = (double)tsn[p] * tm, ts = (double)tsn[n3 + p] * tm;
            ddpdd((se3 - sfi * e3t_ini[p]) * tm * mh, 0.0, h[9], l[9]);
            ddpdd((se3 * tt - sfi * hc_ini[p]) * tm * mh, 0.0, h[10], l[10]);
            ddpdd((se3 * ts - sfi * sc_ini[p]) * tm * mh, 0.0, h[11], l[11]);
            ddpdd(se3 * tm * mh, 0.0, h[12], l[12]);
        }
    }
    dd_block<HSB_NS>(h, l, part);
}

#define HSB_NB 1024
size_t dia_hsb_scratch_bytes() { return (size_t)(HSB_NB + 1) * HSB_NS * sizeof(double2); }

template <typename T>
int launch_dia_hsb(int ln_traqsr, const double *emp, const double *rnf, const double *sbc, const double *qsr,
                   const double *sshn, const T *tsn, const T *tsb, const T *wn, const double *surf_ini,
                   const double *ssh_ini, const double *ssh_hc, const double *ssh_sc, const double *e3t_ini,
                   const double *hc_ini, const double *sc_ini, double *scratch, double *sums)
{
    const DevDomain &d = nh::dom();
    Grid2 g;
    g.block = dim3(NH_BLOCK, 1, 1);
    const size_t nb = (d.n2 + NH_BLOCK - 1) / NH_BLOCK;
    g.grid = dim3((unsigned)(nb < HSB_NB ? nb : HSB_NB), 1, 1);
    double2 *part = (double2 *)scratch;
    LAUNCH("dia_hsb", k_hsb_part<T>, g, d, ln_traqsr, emp, rnf, sbc, qsr, sshn, tsn, tsb, wn, surf_ini, ssh_ini,
           ssh_hc, ssh_sc, e3t_ini, hc_ini, sc_ini, part);
    Grid2 g1;
    g1.block = dim3(NH_BLOCK, 1, 1);
    g1.grid = dim3(1, 1, 1);
    LAUNCH("dia_hsb", k_dd_final<HSB_NS>, g1, (const double2 *)part, (int)g.grid.x, sums);
    return 0;
}

// the initial state of dia_hsb_rst without a restart (diahsb.F90:278-307):
// surf_ini, ssh_ini, e3t_ini, hc_loc_ini, sc_loc_ini over all jpk levels,
// ssh_hc_loc_ini, ssh_sc_loc_ini
template <typename T>
KERNEL void k_hsb_init(DevDomain d, const double *__restrict__ sshn, const T *__restrict__ tsn,
                       double *__restrict__ surf_ini, double *__restrict__ ssh_ini, double *__restrict__ ssh_hc,
                       double *__restrict__ ssh_sc, double *__restrict__ e3t_ini, double *__restrict__ hc_ini,
                       double *__restrict__ sc_ini)
{
    const uint32_t n2 = (uint32_t)d.n2;
    const size_t n3 = d.n3;
    const uint32_t stride = gridDim.x * NH_BLOCK;
    for (uint32_t ij = blockIdx.x * NH_BLOCK + threadIdx.x; ij < n2; ij += stride) {
        surf_ini[ij] = d.md.e1e2t[ij] * gs_mask(d, ij, 0);
        const double sn = sshn[ij];
        ssh_ini[ij] = sn;
        ssh_hc[ij] = (double)tsn[ij] * sn;
        ssh_sc[ij] = (double)tsn[n3 + ij] * sn;
        for (int k = 1; k <= d.jpk; k++) {
            const double tm = pmask<double, 0>(d.ktu, ij, k);
            const size_t p = (size_t)(k - 1) * n2 + ij;
            const double e3 = E3(d.md.e3t, ij, k);
            e3t_ini[p] = e3 * tm;
            hc_ini[p] = (double)tsn[p] * e3 * tm;
            sc_ini[p] = (double)tsn[n3 + p] * e3 * tm;
        }
    }
}

template <typename T>
int launch_dia_hsb_init(const double *sshn, const T *tsn, double *surf_ini, double *ssh_ini, double *ssh_hc,
                        double *ssh_sc, double *e3t_ini, double *hc_ini, double *sc_ini)
{
    const DevDomain &d = nh::dom();
    Grid2 g;
    g.block = dim3(NH_BLOCK, 1, 1);
    const size_t nb = (d.n2 + NH_BLOCK - 1) / NH_BLOCK;
    g.grid = dim3((unsigned)(nb < HSB_NB ? nb : HSB_NB), 1, 1);
    LAUNCH("dia_hsb", k_hsb_init<T>, g, d, sshn, tsn, surf_ini, ssh_ini, ssh_hc, ssh_sc, e3t_ini, hc_ini, sc_ini);
    return 0;
}

#define HSB_INST(T)                                                                                              \
    template int launch_dia_hsb<T>(int, const double *, const double *, const double *, const double *,          \
                                   const double *, const T *, const T *, const T *, const double *,              \
                                   const double *, const double *, const double *, const double *,               \
                                   const double *, const double *, double *, double *);                          \
    template int launch_dia_hsb_init<T>(const double *, const T *, double *, double *, double *, double *,       \
                                        double *, double *, double *);
HSB_INST(double)
HSB_INST(float)

// ---------------------------------------------------------------------------
// a streaming copy of n 16-byte words, the measured bandwidth ceiling a
// streaming kernel on this device meets (benchmark utility): each thread
// moves four words per pass, loads issued before stores, grid-stride
// ---------------------------------------------------------------------------
KERNEL void k_stream_copy(const uint4 *__restrict__ src, uint4 *__restrict__ dst, size_t n)
{
    const size_t stride = (size_t)gridDim.x * NH_BLOCK * 4;
    for (size_t q = (size_t)blockIdx.x * NH_BLOCK * 4 + threadIdx.x; q < n; q += stride) {
        uint4 a[4];
#pragma unroll
        for (int u = 0; u < 4; u++) {
            const size_t e = q + (size_t)u * NH_BLOCK;
            a[u] = e < n ? src[e] : make_uint4(0, 0, 0, 0);
        }
#pragma unroll
        for (int u = 0; u < 4; u++) {
            const size_t e = q + (size_t)u * NH_BLOCK;
            if (e < n) dst[e] = a[u];
        }
    }
}

int launch_stream_copy(const void *src, void *dst, size_t bytes, int blocks)
{
    Grid2 g;
    g.block = dim3(NH_BLOCK, 1, 1);
    g.grid = dim3((unsigned)blocks, 1, 1);
    LAUNCH("stream_copy", k_stream_copy, g, (const uint4 *)src, (uint4 *)dst, bytes / 16);
    return 0;
}
