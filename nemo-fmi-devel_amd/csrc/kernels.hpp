// kernels.hpp -- host launchers of the device kernels (all enqueue on the
// library stream and return a nemo_hip status code).
#pragma once
#include "common.hpp"

int launch_dom_msk(double *tmask, double *umask, double *vmask, double *fmask, double *wmask, double *wumask,
                   double *wvmask, double *ssmask, double *ssumask, double *ssvmask);
template <typename T> int launch_lbc(T *p, int nk, char nat, T pval);

template <typename T> int launch_tra_adv_trp(const T *un, const T *vn, const T *wn, T *zun, T *zvn, T *zwn);
template <typename T>
int launch_tra_adv_fct(double p2dt, const T *pun, const T *pvn, const T *pwn, const T *ptb, const T *ptn, T *pta,
                       int kjpt, int kn_fct_h, int kn_fct_v);
int launch_interp_4th_cpt(const double *tin, double *tout);
template <typename T>
int launch_tra_adv(double p2dt, const T *un, const T *vn, const T *wn, const T *ptb, const T *ptn, T *pta, int kjpt,
                   int kn_fct_h, int kn_fct_v);
template <typename T> int launch_tra_ldf_lap(const T *pahu, const T *pahv, const T *ptb, T *pta, int kjpt, int kpass);
template <typename T>
int launch_tra_ldf_iso(int first, double z2dt, const T *pahu, const T *pahv, const T *ptb, const T *ptbb, T *pta,
                       int kjpt, int kpass, const T *uslp, const T *vslp, const T *wslpi, const T *wslpj, T *ah_wslp2,
                       T *akz);
template <typename T>
int launch_tra_zdf_imp(double p2dt, int is_tra, const T *ptb, T *pta, int kjpt, const T *avt, const T *avs,
                       int l_ldfslp, const T *add, int clamp_sal);

template <typename T> int launch_dyn_adv(int which, int kscheme, const T *un, const T *vn, const T *wn, T *ua, T *va);
template <typename T> int launch_dyn_vor(const T *un, const T *vn, T *ua, T *va);
template <typename T>
int launch_dyn_zdf(double r2dt, const T *ub, const T *vb, T *ua, T *va, const T *avm, const double *rCdU_bot,
                   const double *ua_b, const double *va_b, const double *utau, const double *utau_b,
                   const double *vtau, const double *vtau_b);
int spg_ts_weights(const DevDomain &d, int ll_av, int ll_fw, int *jpit, double *zwgt1, double *zwgt2);
// fused = 1: the 3D parts of phases 1 and 3 are left to the fused step kernels
template <typename T> int launch_dyn_spg_ts(int kt, const nemo_hip_spg_ts_t *st, int fused = 0);
template <typename T>
int launch_step_dyn(int kt, double r2dt, const T *ub, const T *vb, const T *wn, const T *avm, const double *utau_b,
                    const double *vtau_b, const nemo_hip_spg_ts_t *st);
template <typename T>
int launch_step_tra(double p2dt, double z2dt, double p2dt_zdf, const T *un, const T *vn, const T *wn, const T *tsb,
                    const T *tsn, T *tsa, int kn_fct_h, int kn_fct_v, const T *ahtu, const T *ahtv, const T *uslp,
                    const T *vslp, const T *wslpi, const T *wslpj, const T *avt, T *ah_wslp2, T *akz);

// time-step closure (nxt.hip)
template <typename T> int launch_div_hor(const T *un, const T *vn, T *hdivn);
template <typename T>
int launch_ssh_nxt(double z2dt, double zcoef, const T *un, const T *vn, T *hdivn, const double *sshb,
                   const double *emp_b, const double *emp, double *ssha);
template <typename T> int launch_wzv(int first, double z1_2dt, const T *hdivn, T *wn);
int launch_ssh_swp(int euler, double *sshb, double *sshn, const double *ssha);
template <typename T> int launch_tra_nxt(int euler, T *tsb, T *tsn, T *tsa);
template <typename T>
int launch_dyn_nxt(int euler, T *ub, T *vb, T *un, T *vn, T *ua, T *va, const double *ua_b, const double *va_b,
                   const double *un_adv, const double *vn_adv, double *ub_b, double *vb_b, double *un_b,
                   double *vn_b, const double *r1_hu, const double *r1_hv);

// density, pressure gradient, lateral viscosity (phys.hip)
struct SeosPar {   // simplified EOS coefficients, TRA/eosbn2.F90:89-95
    double a0, b0, lambda1, lambda2, mu1, mu2, nu;
};
template <typename T> int launch_eos(int neos, const SeosPar &e, int pot, const T *pts, T *prd, T *prhop, const T *pdep);
template <typename T> int launch_rab(int neos, const SeosPar &e, const T *pts, T *pab, const T *gdept);
int launch_fct_frozen(int mode, double p2dt, const double *pun, const double *pvn, const double *pwn,
                      const double *tb, const double *tn, int kn_h, int kn_v, double *cx, double *cy, double *cz,
                      double *pta);
int launch_fct_ad(const double *pun, const double *pvn, const double *pwn, int kn_h, int kn_v, const double *cx,
                  const double *cy, const double *cz, const double *ta_ad, double *tb_ad, double *tn_ad);
int launch_ldf_lap_ad(const double *pahu, const double *pahv, const double *ta_ad, double *tb_ad);
int launch_zdf_imp_ad(double p2dt, const double *avt, double *ta_ad, double *tb_ad);
int launch_tra_nxt_ad(double *tb_ad, double *tn_ad, double *ta_ad);
int launch_lbc_adj(double *p, int nk);
template <typename T>
int launch_istate_bt(const T *un, const T *vn, const T *ub, const T *vb, const double *r1_hu, const double *r1_hv,
                     double *un_b, double *vn_b, double *ub_b, double *vb_b);
template <typename T>
int launch_zdf_mxl(const T *rn2b, const T *gdept, const T *gdepw, int nlb10, double rho_c, int32_t *nmln,
                   double *hmlp, double *hmlpt);
template <typename T>
int launch_ldf_slp(const T *prd, const T *pn2, const T *gdept, const T *gdepw, const int32_t *nmln,
                   const double *hmlp, const double *hmlpt, double rn_slpmax, T *uslp, T *vslp, T *wslpi, T *wslpj);
template <typename T> int launch_bn2(const T *pts, const T *pab, T *pn2, const T *gdept, const T *gdepw);
template <typename T> int launch_hpg_zco(const T *rhd, T *ua, T *va);
template <typename T>
int launch_dyn_ldf_lap(double zsign, const T *pub, const T *pvb, T *pua, T *pva, const T *ahmt, const T *ahmf);
template <typename T> int launch_stp_ctl(const double *sshn, const T *un, const T *tsn, double *buf, int want_loc);
size_t stp_ctl_buffer_bytes();
template <typename T>
int launch_tra_qsr(int mode, int nksr, double rn_abs, double rn_si0, double rn_si1, const double *qsr,
                   const T *gdepw, double *qsr_hc, double *qsr_hc_b, T *tsa, double *fraqsr_1lev);
template <typename T>
int launch_tra_sbc(int first, double *qns, double *qsr, const double *sfx, const double *emp, const T *tsn, T *tsa,
                   double *sbc, double *sbc_b);

// compensated global sums (gsum.hip)
template <typename T> int launch_glob_sum(const T *p, int nk, int kind, double *scratch, double *out);
size_t glob_sum_scratch_bytes();
int glob_sum_scratch_init(int device);     // the static scratch's address on this device
double *glob_sum_static_scratch(int device);
template <typename T>
int launch_dia_hsb(int ln_traqsr, const double *emp, const double *rnf, const double *sbc, const double *qsr,
                   const double *sshn, const T *tsn, const T *tsb, const T *wn, const double *surf_ini,
                   const double *ssh_ini, const double *ssh_hc, const double *ssh_sc, const double *e3t_ini,
                   const double *hc_ini, const double *sc_ini, double *scratch, double *sums);
template <typename T>
int launch_dia_hsb_init(const double *sshn, const T *tsn, double *surf_ini, double *ssh_ini, double *ssh_hc,
                        double *ssh_sc, double *e3t_ini, double *hc_ini, double *sc_ini);
size_t dia_hsb_scratch_bytes();
int launch_stream_copy(const void *src, void *dst, size_t bytes, int blocks);
