/*
 * nemo_hip.h -- C-ABI of the MI355X (gfx950) implementation of the NEMO 4.0
 * ocean time-step hot path.
 *
 * Every entry point is extern "C", takes plain pointers and sizes, returns an
 * int status (NEMO_HIP_OK == 0) and never throws or aborts across the ABI.  On
 * failure the message is available through nemo_hip_last_error().
 *
 * Array conventions (identical to NEMO's Fortran arrays, so a Fortran caller
 * passes C_LOC() of its own arrays with no transposition):
 *   - 3D fields are (jpi,jpj,jpk) column-major, i fastest: element (i,j,k)
 *     (1-based) lives at offset (i-1) + jpi*((j-1) + jpj*(k-1));
 *   - 4D tracer fields are (jpi,jpj,jpk,kjpt), tracer index slowest;
 *   - 2D fields are (jpi,jpj).
 *
 * Pointer spaces.  With nemo_hip_set_pointer_mode(NEMO_HIP_PTR_HOST) (the
 * default, and the drop-in mode a Fortran model uses) array arguments are host
 * memory: the library stages inputs into HBM, runs the kernels and copies the
 * outputs back before returning.  With NEMO_HIP_PTR_DEVICE they are device
 * pointers already resident in HBM; calls are then asynchronous on the
 * library's stream and nemo_hip_sync() waits for them.
 *
 * Domain state that NEMO keeps in module variables (sizes, scale factors,
 * masks, k-indices, namelist switches) is registered once with
 * nemo_hip_domain_register(); the 3D masks are not passed, they are derived on
 * the device from k_top/k_bot exactly as DOM/dommsk.F90 and DOM/domzgr.F90 do.
 *
 * Module arrays a reference routine reads or writes without having them in its
 * dummy-argument list (avt, the iso-neutral slopes, ah_wslp2/akz, ...) appear
 * as extra trailing pointer arguments.
 *
 * All references below are relative to the NEMO 4.0 source tree (src/OCE).
 */
#ifndef NEMO_HIP_H
#define NEMO_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------------ */
/* status codes                                                             */
/* ------------------------------------------------------------------------ */
#define NEMO_HIP_OK              0
#define NEMO_HIP_ERR_ARG         1   /* bad argument / inconsistent sizes       */
#define NEMO_HIP_ERR_UNSUPPORTED 2   /* namelist option not implemented         */
#define NEMO_HIP_ERR_STATE       3   /* not initialised / no domain registered  */
#define NEMO_HIP_ERR_HIP         4   /* HIP runtime error                       */
#define NEMO_HIP_ERR_NUMERIC     5   /* stp_ctl-style blow-up detected          */
#define NEMO_HIP_ERR_IO          6   /* restart file missing, unreadable or not
                                        matching the registered domain          */

#define NEMO_HIP_PTR_HOST   0
#define NEMO_HIP_PTR_DEVICE 1

/* vorticity schemes: DYN/dynvor.F90:65-70 */
#define NEMO_NP_ENS 0
#define NEMO_NP_ENE 1
#define NEMO_NP_ENT 2
#define NEMO_NP_EET 3
#define NEMO_NP_EEN 4
#define NEMO_NP_MIX 5
/* momentum advection forms: DYN/dynadv.F90:41-44 */
#define NEMO_NP_LIN_DYN 0
#define NEMO_NP_VEC_C2  1
#define NEMO_NP_FLX_C2  2
#define NEMO_NP_FLX_UBS 3

/* ------------------------------------------------------------------------ */
/* domain: the module state of par_oce.F90, dom_oce.F90, phycst.F90 and the   */
/* namelist switches the hot path reads                                     */
/* ------------------------------------------------------------------------ */
typedef struct nemo_hip_domain {
    /* sizes, par_oce.F90:50-76 */
    int32_t jpi, jpj, jpk;
    /* lateral boundary type, DOM/dom_oce.F90:68 (0 closed, 1 cyclic E-W,
       2 cyclic N-S, 7 cyclic both; north folds 3-6 are not supported) */
    int32_t jperio;
    /* vertical coordinate / free surface, DOM/dom_oce.F90 */
    int32_t ln_linssh;          /* must be 1: e3 fixed in time              */
    int32_t ln_isfcav;          /* must be 0: no ice-shelf cavities         */
    int32_t ln_zps;             /* must be 0: full steps                    */
    int32_t ln_sco;             /* must be 0                                */
    /* tracer physics, LDF/ldftra.F90, ZDF/zdf_oce.F90 */
    int32_t ln_traldf_lap;      /* laplacian (1) or bilaplacian (0)         */
    int32_t ln_traldf_msc;      /* method of stabilizing correction         */
    int32_t ln_zdfddm;          /* double diffusion: separate S matrix      */
    int32_t ln_zad_Aimp;        /* must be 0                                */
    /* momentum, DYN/dynadv.F90, DYN/dynkeg.F90, DYN/dynvor.F90 */
    int32_t n_dynadv;           /* NEMO_NP_*_DYN / NEMO_NP_VEC_C2 / ...     */
    int32_t nn_dynkeg;          /* 0 C2, 1 Hollingsworth                    */
    int32_t nvor_scheme;        /* NEMO_NP_ENS ... NEMO_NP_MIX              */
    int32_t nn_een_e3f;         /* 0 or 1                                   */
    /* vertical momentum diffusion, ZDF/zdfdrg.F90 */
    int32_t ln_drgimp;          /* implicit bottom friction                 */
    /* split-explicit free surface, DYN/dynspg_ts.F90 (namdyn_spg)          */
    int32_t ln_dynspg_ts;
    int32_t ln_bt_fw, ln_bt_av, nn_bt_flt, nn_baro;
    /* time stepping, DOM/dom_oce.F90:62-63 */
    int32_t neuler, nit000;
    /* penetrative solar flux, namsbc (TRA/trasbc.F90:97-100): 1 keeps qsr
       for tra_qsr, 0 merges it into qns in tra_sbc                         */
    int32_t ln_traqsr;
    double  rdt;                /* baroclinic time step [s]                 */
    double  rn_atfp;            /* Asselin filter parameter                 */
    double  rau0;               /* reference density, TRA/eosbn2.F90:1244   */
    double  grav;               /* DOM/phycst.F90:38                        */
    double  rn_bt_alpha;        /* must be 0                                */
    double  rn_scal_load;       /* 0 without tides (SBC/tideini.F90:84)     */
    /* first / last wet level per column (0 = land), DOM/dommsk.F90 input */
    const int32_t *k_top, *k_bot;
    /* horizontal scale factors and their derived forms, DOM/domhgr.F90:133-150 */
    const double *e1t, *e2t, *e1u, *e2u, *e1v, *e2v, *e1f, *e2f;
    const double *e1e2t, *r1_e1e2t, *e1e2u, *r1_e1e2u, *e1e2v, *r1_e1e2v;
    const double *e1e2f, *r1_e1e2f;
    const double *e2_e1u, *e1_e2v, *r1_e1u, *r1_e2u, *r1_e1v, *r1_e2v;
    /* Coriolis parameter at f- and t-points */
    const double *ff_f, *ff_t;
    /* vertical scale factors (now == before == after when ln_linssh) */
    const double *e3t_n, *e3u_n, *e3v_n, *e3w_n, *e3f_n, *e3uw_n, *e3vw_n;
} nemo_hip_domain_t;

/* ------------------------------------------------------------------------ */
/* split-explicit surface pressure gradient: the module state dyn_spg_ts    */
/* reads and writes (oce.F90, sbc_oce.F90, zdfdrg.F90, dynspg_ts.F90:73)    */
/* ------------------------------------------------------------------------ */
typedef struct nemo_hip_spg_ts {
    /* 3D (jpi,jpj,jpk): in/out momentum trends and now velocities */
    void *ua, *va, *un, *vn;
    /* 2D (jpi,jpj), always FP64 */
    const double *un_b, *vn_b, *ub_b, *vb_b;       /* now / before barotropic   */
    const double *sshn;                            /* now sea surface height    */
    const double *hu_n, *hv_n, *r1_hu_n, *r1_hv_n, *ht_n;
    const double *rCdU_bot;                        /* bottom drag (t-points)    */
    const double *utau, *vtau;                     /* wind stress               */
    const double *emp, *rnf;                       /* fresh-water fluxes        */
    double *ua_b, *va_b, *ssha;                    /* out: after barotropic     */
    double *un_adv, *vn_adv;                       /* out: advective transport  */
    double *ub2_b, *vb2_b, *un_bf, *vn_bf;         /* in/out: time-filter state */
} nemo_hip_spg_ts_t;

/* ------------------------------------------------------------------------ */
/* equation of state: the namelist state of eos_init (TRA/eosbn2.F90:1218),  */
/* nameos parameters with the defaults of TRA/eosbn2.F90:89-95               */
/* ------------------------------------------------------------------------ */
#define NEMO_NP_TEOS10 (-1)
#define NEMO_NP_EOS80  0
#define NEMO_NP_SEOS   1
typedef struct nemo_hip_eos {
    int32_t neos;               /* NEMO_NP_TEOS10 / NEMO_NP_EOS80 (the
                                   polynomials of TRA/eosbn2.F90:1266-1647;
                                   the rn_* fields are unused) or
                                   NEMO_NP_SEOS                             */
    int32_t reserved0;
    double  rn_a0, rn_b0;       /* thermal / saline expansion               */
    double  rn_lambda1, rn_lambda2;   /* cabbeling in T^2, S^2              */
    double  rn_mu1, rn_mu2;     /* thermobaric in T, S                      */
    double  rn_nu;              /* cabbeling in T*S                         */
} nemo_hip_eos_t;

/* ------------------------------------------------------------------------ */
/* restart state: the variables of rst_write / rst_read                     */
/* (IOM/restart.F90:133-191, 240-309) and ts_rst (DYN/dynspg_ts.F90:        */
/* 1391-1463), as a directory holding manifest.json and one raw array per   */
/* variable (<name>.bin, Fortran order, little-endian "f8" or "f4")         */
/* ------------------------------------------------------------------------ */
typedef struct nemo_hip_rst {
    int32_t real_bytes;         /* 8: the 3D fields are double; 4: float
                                   (the _sp storage form)                   */
    int32_t kqsr_b;             /* out of rst_read: 1 when qsr_hc_b was read
                                   (pass it to nemo_hip_tra_qsr at nit000,
                                   TRA/traqsr.F90:136-139), else 0          */
    void   *ub, *vb, *tsb;      /* before fields; tsb (jpi,jpj,jpk,2) holds
                                   the variables tb and sb                  */
    void   *un, *vn, *tsn;      /* now fields (tn, sn)                      */
    void   *rhop;               /* potential density                        */
    double *sshb, *sshn;        /* (jpi,jpj)                                */
    double *ub2_b, *vb2_b, *un_bf, *vn_bf;   /* ts_rst, (jpi,jpj)           */
    /* on read only, optional: the depth-mean velocities DOM/istate.F90:
       159-178 derives after a restart, from r1_hu_n / r1_hv_n            */
    double *un_b, *vn_b, *ub_b, *vb_b;
    const double *r1_hu_n, *r1_hv_n;
    /* optional, the two-band tra_qsr's variables (TRA/traqsr.F90:287-292,
       424-425): rst_write stores qsr_hc (jpi,jpj,jpk) under the name
       qsr_hc_b and fraqsr_1lev (jpi,jpj); rst_read reads them into qsr_hc_b
       and fraqsr_1lev when the file holds them.  NULL skips them.          */
    const double *qsr_hc;
    double *qsr_hc_b, *fraqsr_1lev;
} nemo_hip_rst_t;

/* ------------------------------------------------------------------------ */
/* library control                                                          */
/* ------------------------------------------------------------------------ */
int nemo_hip_init(int device);
int nemo_hip_finalize(void);
/* copies the last error message (NUL terminated, truncated to len) */
int nemo_hip_last_error(char *buf, int len);
int nemo_hip_set_pointer_mode(int mode);
int nemo_hip_sync(void);
/* the library's HIP stream, as an opaque handle (hipStream_t) */
void *nemo_hip_stream(void);

/* Resident fields for device mode (no counterpart routine in the reference:
   this is the field bind / push / pull of a resident-mode caller, which keeps
   NEMO's prognostic arrays in HBM between calls and touches host memory only
   for I/O, restarts and diagnostics).  alloc returns a device address that
   the *_DEVICE-mode entry points accept; push copies a host array to it and
   returns when the copy is done; pull waits for the work queued on the
   library stream, then copies the device array to host memory; free
   releases an address alloc returned (others are rejected). */
int nemo_hip_field_alloc(size_t bytes, void **dev);
int nemo_hip_field_free(void *dev);
int nemo_hip_field_push(void *dev, const void *host, size_t bytes);
int nemo_hip_field_pull(void *host, const void *dev, size_t bytes);

/* replaces the module initialisation done by dom_init (DOM/domain.F90) for
   the arrays the hot path reads; call again if the domain changes */
int nemo_hip_domain_register(const nemo_hip_domain_t *dom);

/* Contexts (no counterpart routine in the reference, which runs one domain
   per process): an ensemble of independent domains in one process, each
   with its own registered domain, scratch and the state the reference keeps
   in module variables between calls (r2dt, the barotropic levels).  Context
   0 exists from nemo_hip_init and lives until nemo_hip_finalize; create
   returns a new, empty one; select makes a context current for every call
   that follows (registration and entry points); destroy frees a context
   that is not current. */
int nemo_hip_context_create(int *ctx);
int nemo_hip_context_select(int ctx);
int nemo_hip_context_destroy(int ctx);

/* Graphs (no counterpart routine in the reference: they serve callers that
   step the same fields many times, such as an ensemble).  begin records the
   launches of the device-mode calls that follow on the library stream, in
   any contexts; end turns the record into an executable graph and returns
   its id; launch replays it on the library stream; free releases it.  A
   recorded call replays with the arguments and time step it was recorded
   with, so recording refuses what would change between replays: the first
   two steps (kt <= nit000 + 1), the first tra_ldf_iso call (kt = kit000),
   dyn_spg_ts without time averaging (ln_bt_av = F), host-pointer mode, and
   calls that synchronise or allocate. */
int nemo_hip_graph_begin(void);
int nemo_hip_graph_end(int *graph);
int nemo_hip_graph_launch(int graph);
int nemo_hip_graph_free(int graph);

/* Plans (no counterpart routine in the reference: the batching form of an
   ensemble).  begin starts tracing: the device-mode calls that follow, in
   any contexts, are checked and recorded as they would be for a graph but
   nothing is launched.  end builds the plan: the contexts that traced calls
   are its members; members whose launch sequences match (the same kernels
   and workgroup shapes, as members with the same options have) form a
   group, and every position of a group's sequence becomes ONE launch over
   the work list of all its members' workgroups, each running with its own
   member's domain, fields and grid (for the kernels written to batch, which
   are those of nemo_hip_step_dyn, nemo_hip_step_tra and the time-step
   closure; any other traced launch replays once per member).  launch issues
   the plan on the library stream (also while a graph is recorded, so a
   graph of a plan holds one node per batched launch); info reports the
   members, groups, launches per replay and how many of them are batched;
   free releases it.  The rules of graph recording apply to the traced
   calls, and the fields they name must stay allocated while the plan
   exists. */
int nemo_hip_plan_begin(void);
int nemo_hip_plan_end(int *plan);
int nemo_hip_plan_launch(int plan);
int nemo_hip_plan_info(int plan, int *nmembers, int *ngroups, int *nops, int *nbatched, int *nlaunch);
int nemo_hip_plan_free(int plan);

/* per-routine GPU timers keyed by reference routine name (timing.F90:89-212):
   when enabled every kernel launch is bracketed by HIP events on the
   library's stream and accumulated by name */
int nemo_hip_timing_enable(int on);
/* restrict the timers to the launches of one kernel, keyed "routine/kernel"
   (e.g. "step_tra/k_fct_h"); NULL or "" times every launch.  Lets a caller
   time one kernel inside a production loop at the cost of three events per
   launch of it */
int nemo_hip_timing_only(const char *key);
int nemo_hip_timing_reset(void);
/* total milliseconds and launch count accumulated for `name` */
int nemo_hip_timing_get(const char *name, double *total_ms, int64_t *count);

/* ------------------------------------------------------------------------ */
/* masks and k-indices, DOM/dommsk.F90:134-210 and DOM/domzgr.F90:267-316 */
/* every output pointer may be NULL                                         */
/* ------------------------------------------------------------------------ */
int nemo_hip_dom_msk(double *tmask, double *umask, double *vmask, double *fmask,
                     double *wmask, double *wumask, double *wvmask,
                     double *ssmask, double *ssumask, double *ssvmask,
                     int32_t *mikt, int32_t *mbkt, int32_t *miku, int32_t *mikv,
                     int32_t *mbku, int32_t *mbkv);

/* lbc_lnk, serial path, LBC/lbclnk.F90:81-89 + LBC/lbc_lnk_generic.h90:48-108.
   ptab is (jpi,jpj,nk); cd_nat in 'T','U','V','W','F'; psgn only matters at a
   north fold (unsupported); pval is the closed-boundary value. */
int nemo_hip_lbc_lnk(double *ptab, int nk, char cd_nat, double psgn, double pval);

/* ------------------------------------------------------------------------ */
/* tracers                                                                  */
/* ------------------------------------------------------------------------ */
/* effective transport of tra_adv, TRA/traadv.F90:100-124 */
int nemo_hip_tra_adv_trp(const double *un, const double *vn, const double *wn,
                         double *zun, double *zvn, double *zwn);

/* tra_adv, TRA/traadv.F90:77-175 with the FCT scheme (nn_fct_h, nn_fct_v in
   {2,4}): r2dt of kt (lines 95-97, kept by the library as the module keeps
   it), the effective transports of un, vn, wn (lines 100-124, formed where
   the fluxes use them, not stored) and tra_adv_fct (line 150) adding the
   trend of the kjpt tracers to pta */
int nemo_hip_tra_adv(int kt, const double *un, const double *vn, const double *wn,
                     const double *ptb, const double *ptn, double *pta, int kjpt,
                     int kn_fct_h, int kn_fct_v);
int nemo_hip_tra_adv_sp(int kt, const float *un, const float *vn, const float *wn,
                        const float *ptb, const float *ptn, float *pta, int kjpt,
                        int kn_fct_h, int kn_fct_v);

/* FCT advection, TRA/traadv_fct.F90:54-80 (kn_fct_h, kn_fct_v in {2,4}) */
int nemo_hip_tra_adv_fct(int kt, int kit000, const char *cdtype, double p2dt,
                         const double *pun, const double *pvn, const double *pwn,
                         const double *ptb, const double *ptn, double *pta,
                         int kjpt, int kn_fct_h, int kn_fct_v);

/* 4th order compact T->w interpolation, TRA/traadv_fct.F90:517 */
int nemo_hip_interp_4th_cpt(const double *pt_in, double *pt_out);

/* iso-level laplacian, TRA/traldf_lap_blp.F90:47-76 */
int nemo_hip_tra_ldf_lap(int kt, int kit000, const char *cdtype,
                         const double *pahu, const double *pahv,
                         const double *pgu, const double *pgv,
                         const double *pgui, const double *pgvi,
                         const double *ptb, double *pta, int kjpt, int kpass);

/* rotated (iso-neutral) laplacian, TRA/traldf_iso.F90:50-103.
   Trailing arguments: slopes of LDF/ldfslp.F90:57-58 (in) and the module
   arrays ah_wslp2, akz of LDF/ldfslp.F90:64-65 (written when kpass == 1). */
int nemo_hip_tra_ldf_iso(int kt, int kit000, const char *cdtype,
                         const double *pahu, const double *pahv,
                         const double *pgu, const double *pgv,
                         const double *pgui, const double *pgvi,
                         const double *ptb, const double *ptbb, double *pta,
                         int kjpt, int kpass,
                         const double *uslp, const double *vslp,
                         const double *wslpi, const double *wslpj,
                         double *ah_wslp2, double *akz);

/* implicit vertical diffusion, TRA/trazdf.F90:109-135.  pta is the trend
   on entry and the after tracer on exit.  avs may equal avt.  l_ldfslp != 0
   adds ah_wslp2 (or akz when ln_traldf_msc) to the diffusivity. */
int nemo_hip_tra_zdf_imp(int kt, int kit000, const char *cdtype, double p2dt,
                         const double *ptb, double *pta, int kjpt,
                         const double *avt, const double *avs, int l_ldfslp,
                         const double *ah_wslp2, const double *akz);

/* tra_zdf, TRA/trazdf.F90:46-106: r2dt from kt, tra_zdf_imp on (T,S),
   then negative salinity reset to 0.1 */
int nemo_hip_tra_zdf(int kt, const double *tsb, double *tsa,
                     const double *avt, const double *avs, int l_ldfslp,
                     const double *ah_wslp2, const double *akz);

/* ------------------------------------------------------------------------ */
/* momentum                                                                 */
/* ------------------------------------------------------------------------ */
/* DYN/dynkeg.F90:46-192 (kscheme 0 C2, 1 Hollingsworth) */
int nemo_hip_dyn_keg(int kt, int kscheme, const double *un, const double *vn,
                     double *ua, double *va);
/* DYN/dynzad.F90:38-121 */
int nemo_hip_dyn_zad(int kt, const double *un, const double *vn, const double *wn,
                     double *ua, double *va);
/* DYN/dynadv_cen2.F90:37-144 */
int nemo_hip_dyn_adv_cen2(int kt, const double *un, const double *vn,
                          const double *wn, double *ua, double *va);
/* DYN/dynadv.F90:55-85, dispatch on the registered n_dynadv */
int nemo_hip_dyn_adv(int kt, const double *un, const double *vn, const double *wn,
                     double *ua, double *va);
/* DYN/dynvor.F90:98-186, scheme from the registered nvor_scheme */
int nemo_hip_dyn_vor(int kt, const double *un, const double *vn,
                     double *ua, double *va);

/* DYN/dynzdf.F90:47-499.  ua/va are trends on entry and after
   velocities on exit. */
int nemo_hip_dyn_zdf(int kt, const double *ub, const double *vb,
                     double *ua, double *va, const double *avm,
                     const double *rCdU_bot, const double *ua_b, const double *va_b,
                     const double *utau, const double *utau_b,
                     const double *vtau, const double *vtau_b);

/* DYN/dynspg_ts.F90:121-1310 (3D arrays in st are double*) */
int nemo_hip_dyn_spg_ts(int kt, nemo_hip_spg_ts_t *st);
/* ts_wgt, DYN/dynspg_ts.F90:1313-1388: weights and cycle length */
int nemo_hip_ts_wgt(int ll_av, int ll_fw, int *jpit, double *zwgt1, double *zwgt2);

/* ------------------------------------------------------------------------ */
/* density, pressure gradient, lateral viscosity                             */
/* ------------------------------------------------------------------------ */
/* eos_init, TRA/eosbn2.F90:1218-1700: the equation of state of the current
   context (call after nemo_hip_domain_register) */
int nemo_hip_eos_register(const nemo_hip_eos_t *eos);
/* eos_insitu, TRA/eosbn2.F90:188-301: prd = (rho - rau0) / rau0 on
   levels 1..jpkm1 from pts (jpi,jpj,jpk,2) and the depth pdep (jpi,jpj,jpk) */
int nemo_hip_eos_insitu(const double *pts, double *prd, const double *pdep);
/* eos_insitu_pot, TRA/eosbn2.F90:304-468: prd and the surface-referenced
   potential density prhop */
int nemo_hip_eos_insitu_pot(const double *pts, double *prd, double *prhop, const double *pdep);
/* rab_3d, TRA/eosbn2.F90:566-676: thermal / haline expansion ratios pab
   (jpi,jpj,jpk,2) at the t-point depths gdept_n */
int nemo_hip_rab_3d(const double *pts, double *pab, const double *gdept);
/* bn2, TRA/eosbn2.F90:889-933: N^2 at w-points, levels 2..jpkm1 */
int nemo_hip_bn2(const double *pts, const double *pab, double *pn2,
                 const double *gdept, const double *gdepw);
/* rst_write, IOM/restart.F90:133-191, with ts_rst WRITE
   (DYN/dynspg_ts.F90:1440-1451): writes rdt, kt and the fields of rst into
   the directory dir (created if missing).  ln_bt_av = F (the barotropic
   *_e levels) is not supported. */
int nemo_hip_rst_write(const char *dir, int kt, const nemo_hip_rst_t *rst);
/* rst_read, IOM/restart.F90:240-309, with ts_rst READ (dynspg_ts.F90:
   1402-1437): fills the fields of rst from dir.  *neuler is 0 (Euler
   restart: before fields set to the now fields, the ts_rst fields to zero)
   when the file's rdt differs from the domain's or it holds no before
   fields, else 1; *kt is the step the file was written at.  When un_b is
   not NULL the four depth means are computed as DOM/istate.F90:159-178. */
int nemo_hip_rst_read(const char *dir, nemo_hip_rst_t *rst, int *kt, int *neuler);
/* Tangent linear and adjoint of the tracer step, one tracer, FP64: the
   step tra_adv_fct (orders 2 or 4) + tra_ldf_lap (kpass 1) + tra_zdf_imp (avt,
   no slopes) + leapfrog tra_nxt, with the nonosc limiter factors frozen
   from the forward state (the step is then linear in the tracer).  The
   adjoint entry points have no counterpart in the reference; each is the
   transpose of the routine named, as this library runs it.  Adjoint arrays
   (*_ad) are accumulated unless stated.
   tra_adv_fct_coef: the factors nonosc (TRA/traadv_fct.F90:404-426) gives
   the state (tb, tn): cx, cy at interior faces, cz(k) for the face above
   level k (k = 2..jpk).  pun/pvn/pwn are the transports tra_adv_fct takes. */
int nemo_hip_tra_adv_fct_coef(double p2dt, const double *pun, const double *pvn, const double *pwn, const double *ptb,
                              const double *ptn, int kn_fct_h, int kn_fct_v, double *pcx, double *pcy, double *pcz);
/* pta += the FCT trend (orders kn_fct_h, kn_fct_v: 2 or 4) with the factors
   given (tra_adv_fct itself when they are the state's own) */
int nemo_hip_tra_adv_fct_tl(double p2dt, const double *pun, const double *pvn, const double *pwn, const double *ptb,
                            const double *ptn, int kn_fct_h, int kn_fct_v, const double *pcx, const double *pcy,
                            const double *pcz, double *pta);
/* ptb_ad, ptn_ad += the transpose of tra_adv_fct_tl applied to pta_ad */
int nemo_hip_tra_adv_fct_ad(const double *pun, const double *pvn, const double *pwn, int kn_fct_h, int kn_fct_v,
                            const double *pcx, const double *pcy, const double *pcz, const double *pta_ad,
                            double *ptb_ad, double *ptn_ad);
/* ptb_ad += the transpose of tra_ldf_lap (kpass 1) applied to pta_ad */
int nemo_hip_tra_ldf_lap_ad(const double *pahu, const double *pahv, const double *pta_ad, double *ptb_ad);
/* tra_zdf_imp: on entry pta_ad is the adjoint of the after field, on exit
   that of the trend; ptb_ad += the adjoint of tb */
int nemo_hip_tra_zdf_imp_ad(double p2dt, const double *avt, double *pta_ad, double *ptb_ad);
/* leapfrog tra_nxt (kt > nit000): on entry ptb_ad, ptn_ad are the adjoints
   of the new tsb, tsn; on exit those of the old ones, and pta_ad (set)
   that of tsa; ring points are folded into their lbc sources */
int nemo_hip_tra_nxt_ad(double *ptb_ad, double *ptn_ad, double *pta_ad);
/* the transpose of lbc_lnk 'T' on nk levels: ring values added to their
   sources, the ring cleared */
int nemo_hip_lbc_adj(double *p, int nk);
/* zdf_mxl, ZDF/zdfmxl.F90:64-140: the mixed-layer base level nmln
   (jpi,jpj), its w-depth hmlp and the depth hmlpt of its last t-point, from
   rn2b (jpi,jpj,jpk) and the rho_c density criterion (0.01 in the
   reference, ZDF/zdfmxl.F90:34); nlb10 is the shallowest w-level below 10 m
   (DOM/domzgr.F90:144-145).  The turbocline depth hmld (a diagnostic) is
   not computed.  2-D outputs are FP64 in both storage forms. */
int nemo_hip_zdf_mxl(int kt, const double *rn2b, const double *gdept, const double *gdepw, int nlb10, double rho_c,
                     int32_t *nmln, double *hmlp, double *hmlpt);
/* ldf_slp, LDF/ldfslp.F90:83-374 with ldf_slp_mxl (620-729): iso-neutral
   slopes uslp, vslp, wslpi, wslpj (jpi,jpj,jpk) from the in situ density
   prd and N^2 pn2 (the reference passes rhd and rn2b, OCE/step.F90:150),
   with the mixed layer of nemo_hip_zdf_mxl and the nam_traldf limit
   rn_slpmax.  Levels 2..jpkm1 are computed; levels 1 and jpk keep their
   interior values and every level's ring is set by lbc_lnk. */
int nemo_hip_ldf_slp(int kt, const double *prd, const double *pn2, const double *gdept, const double *gdepw,
                     const int32_t *nmln, const double *hmlp, const double *hmlpt, double rn_slpmax, double *uslp,
                     double *vslp, double *wslpi, double *wslpj);
/* hpg_zco, DYN/dynhpg.F90:258-323 (dyn_hpg with ln_hpg_zco): the hydrostatic
   pressure gradient of the in situ density anomaly rhd added to ua, va */
int nemo_hip_hpg_zco(int kt, const double *rhd, double *ua, double *va);
/* dyn_ldf_lap, DYN/dynldf_lap_blp.F90:37-104: the iso-level laplacian
   (rot/div form) of pub, pvb added to pua, pva; ahmt, ahmf are the
   viscosities of LDF/ldfdyn.F90 (masked by tmask, fmask) */
int nemo_hip_dyn_ldf_lap(int kt, const double *pub, const double *pvb, double *pua, double *pva,
                         int kpass, const double *ahmt, const double *ahmf);

/* tra_qsr, TRA/traqsr.F90:80-304, two bands (ln_qsr_2bd = T): the
   penetrative solar heat content qsr_hc (3D, FP64) of levels 1..nksr,
   added to the temperature trend tsa with its value of the step before,
   qsr_hc_b, and fraqsr_1lev (2D) = the fraction absorbed in level 1.
   nksr is trc_oce_ext_lev(rn_si1, 100) of trc_oce.F90, computed once by
   the caller's tra_qsr_init (traqsr.F90:411).  At kt = nit000, kqsr_b = 1
   says qsr_hc_b holds the restart's field (leapfrog); kqsr_b = 0 clears it
   (Euler).  A run with tra_qsr registers ln_traqsr = 1, so tra_sbc leaves
   qsr to it.  The
   exponentials are the device's exp: FP64 results agree with a host libm
   to a few units in the last place, not bit for bit. */
int nemo_hip_tra_qsr(int kt, int kqsr_b, int nksr, double rn_abs, double rn_si0, double rn_si1, const double *qsr,
                     const double *gdepw, double *qsr_hc, double *qsr_hc_b, double *tsa, double *fraqsr_1lev);
int nemo_hip_tra_qsr_sp(int kt, int kqsr_b, int nksr, double rn_abs, double rn_si0, double rn_si1, const double *qsr,
                        const float *gdepw, double *qsr_hc, double *qsr_hc_b, float *tsa, double *fraqsr_1lev);

/* tra_sbc, TRA/trasbc.F90:62-238, linear free surface, no ice shelf, runoff
   or assimilation terms: with the domain's ln_traqsr = 0, qns += qsr and
   qsr = 0 (lines 97-100); with ln_traqsr = 1 both are left for tra_qsr; the swap of the surface content trends
   sbc_tsc_b <- sbc_tsc and the new sbc_tsc (jpi,jpj,2) from qns, sfx and the
   concentration/dilution of emp; the trend added to level 1 of tsa.  At
   kt = nit000 the run starts from rest of the forcing (no restart): both
   trends reset and zfact = 1.  rcp = 3991.86795711963 (TRA/eosbn2.F90:1245). */
int nemo_hip_tra_sbc(int kt, double *qns, double *qsr, const double *sfx, const double *emp,
                     const double *tsn, double *tsa, double *sbc_tsc, double *sbc_tsc_b);
int nemo_hip_tra_sbc_sp(int kt, double *qns, double *qsr, const double *sfx, const double *emp,
                        const float *tsn, float *tsa, double *sbc_tsc, double *sbc_tsc_b);

/* stp_ctl, stpctl.F90:55-190: the run control of one time step.  zmax[0..5]
   = max|sshn|, max|un|, max(-S), max(S), max(-T), max(T) (the tracers over
   tmask = 1), as lines 113-121 define them; a NaN anywhere in sshn, un or S
   makes the corresponding maximum NaN.  When one of the tests of lines
   149-155 fails, *kindic = -3, loc holds the first (Fortran order) positions
   of max|ssh| (i,j), max|u| (i,j,k), min S and max S (i,j,k), 1-based, and
   the call returns NEMO_HIP_ERR_NUMERIC with the reference's message;
   otherwise *kindic is unchanged.  The call waits for the queued work. */
int nemo_hip_stp_ctl(int kt, const double *sshn, const double *un, const double *tsn,
                     double zmax[6], int32_t loc[11], int *kindic);
int nemo_hip_stp_ctl_sp(int kt, const double *sshn, const float *un, const float *tsn,
                        double zmax[6], int32_t loc[11], int *kindic);

/* A streaming device-to-device copy of bytes (a multiple of 16, both
   addresses 16-byte aligned) with `blocks` workgroups, on the library stream
   (device pointers, asynchronous).  No counterpart in the reference: the
   benchmark's measured bandwidth ceiling for a streaming kernel. */
int nemo_hip_stream_copy(const void *src, void *dst, size_t bytes, int blocks);

/* Tuning switches (no counterpart in the reference): the kernel variants
   and tile shapes a caller or the project's autotuner may choose, process
   wide.  Every choice gives the same results bit for bit.  Keys:
     "fct_fused"  1 or 2: tra_adv_fct on closed domains as one kernel per
                  tile and tracer (intermediates on chip), with 256 (1) or
                  512 (2) threads per tile; 0 (default): the flux / limiter
                  chain through scratch arrays.
     "spg_persistent"  1 or 2: dyn_spg_ts's sub-steps in one cooperative
                  launch with grid barriers (1: one 1024-thread workgroup
                  per CU; 2: 256-thread workgroups, as many as stay
                  resident; bounded waits: a barrier that gives up makes
                  the next nemo_hip_sync fail); 0 (default), or while a
                  graph is recorded: two launches per sub-step.
     "zdf_reg"    1: in the FP32 storage form with jpk <= 76, the column
                  solves of tra_zdf and dyn_zdf keep the forward-
                  substituted values on chip; 0 (default): written to the
                  trend array and read back.
     "fct_vlim"   1 or 2: in the chain (fct_fused 0) on closed domains,
                  the vertical upstream and antidiffusive fluxes, the first
                  guess and the limiter in one bottom-up march per tile and
                  tracer (256 threads per 64 x 4 tile (1) or 512 per 64 x 8
                  (2)); 0 (default): as the kernels of the chain.
   Unknown keys and values are errors. */
int nemo_hip_tune_set(const char *key, int value);
int nemo_hip_tune_get(const char *key, int *value);

/* glob_sum and glob_sum_full (lib_fortran_generic.h90 FUNCTION_GLOBSUM,
   lib_fortran.F90:300-332 DDPDD): the compensated (double-double) sum of
   ptab(i,j,k) * mask(i,j) over levels 1..nk of a (jpi,jpj,nk) array, mask =
   tmask_i (kind 0, glob_sum) or tmask_h (kind 1, glob_sum_full; both from
   DOM/dommsk.F90:213-238).  The terms and the sum are FP64 in both storage
   forms.  psum: one double in the pointer space of the call (in device mode
   a device address, and the call is asynchronous on the library stream). */
int nemo_hip_glob_sum(const double *ptab, int nk, int kind, double *psum);
int nemo_hip_glob_sum_sp(const float *ptab, int nk, int kind, double *psum);
/* nemo_hip_glob_sum that allocates nothing.  No counterpart in the reference:
   it is for callers that cannot allocate device memory (code under HIP graph
   capture, serving loops, harnesses that forbid hipMalloc).  Device-pointer
   mode only (ptab and psum device addresses; asynchronous on the library
   stream).  Its partial sums live in a static device array of 16 KiB that
   every GPU holds until the process ends and that all calls on one GPU
   share: calls that overlap on one GPU race.  At most
   NEMO_HIP_GLOB_SUM_NOALLOC_MAX values (jpi*jpj*nk). */
#define NEMO_HIP_GLOB_SUM_NOALLOC_MAX 2147483647
int nemo_hip_glob_sum_noalloc(const double *ptab, int nk, int kind, double *psum);

/* dia_hsb, DIA/diahsb.F90:60-239 (linear free surface, no ice-shelf cavity,
   runoff or geothermal flux): the heat, salt and volume budget of the run
   from thirteen compensated global sums, all FP64 (also in the precision
   form).  nemo_hip_dia_hsb_init is dia_hsb_init with dia_hsb_rst('READ')
   without a restart (:278-307, :411-412): it writes the initial-state
   arrays from tsn and sshn, sets surf_tot and zeroes the accumulated forcing.
   nemo_hip_dia_hsb adds the step's forcing to frc_* and fills sums[] (in the
   order of gsum.hip's table) and the diagnostics iom_put writes (:221-233).
   The reference also masks tsn and tsb in place (:88-89); this call leaves
   them as given and masks the terms instead, which gives the same sums.
   Pointers follow the pointer mode; the scalars are host memory and the
   calls synchronise. */
typedef struct nemo_hip_hsb {
    int32_t real_bytes;                    /* 8: tsn, tsb, wn double; 4: float */
    int32_t reserved0;
    const void *tsn, *tsb;                 /* (jpi,jpj,jpk,2)                  */
    const void *wn;                        /* (jpi,jpj,jpk)                    */
    const double *sshn, *emp, *rnf, *qsr;  /* (jpi,jpj); qsr only with ln_traqsr */
    const double *sbc_tsc;                 /* (jpi,jpj,2)                      */
    double *surf_ini, *ssh_ini, *ssh_hc_loc_ini, *ssh_sc_loc_ini;   /* (jpi,jpj)      */
    double *e3t_ini, *hc_loc_ini, *sc_loc_ini;                       /* (jpi,jpj,jpk)  */
    double surf_tot;                       /* set by init                      */
    double frc_v, frc_t, frc_s, frc_wn_t, frc_wn_s;   /* accumulated forcing   */
    double sums[13];                       /* out: the call's global sums      */
    double bgfrcvol, bgfrctem, bgfrchfx, bgfrcsal, bgtemper, bgsaline;
    double bgheatco, bgheatfx, bgsaltco, bgvolssh, bgmistem, bgmissal;   /* out */
} nemo_hip_hsb_t;
int nemo_hip_dia_hsb_init(nemo_hip_hsb_t *h);
int nemo_hip_dia_hsb(int kt, nemo_hip_hsb_t *h);

/* ------------------------------------------------------------------------ */
/* time-step closure around the hot path (linear free surface)              */
/* ------------------------------------------------------------------------ */
/* div_hor, DYN/divhor.F90:50-110: hdivn (jpi,jpj,jpk) on levels 1..jpkm1,
   then lbc_lnk('T') */
int nemo_hip_div_hor(int kt, const double *un, const double *vn, double *hdivn);
/* ssh_nxt, DYN/sshwzv.F90:56-126: div_hor into hdivn, then the after sea
   level ssha from sshb, emp_b, emp (2D, FP64) */
int nemo_hip_ssh_nxt(int kt, const double *un, const double *vn, double *hdivn,
                     const double *sshb, const double *emp_b, const double *emp,
                     double *ssha);
/* wzv, DYN/sshwzv.F90:129-213: wn from hdivn (wn(jpk) = 0 set at nit000) */
int nemo_hip_wzv(int kt, const double *hdivn, double *wn);
/* ssh_swp, DYN/sshwzv.F90:216-267: Asselin filter of sshb, sshn = ssha */
int nemo_hip_ssh_swp(int kt, double *sshb, double *sshn, const double *ssha);
/* tra_nxt, TRA/tranxt.F90:66-192 (tra_nxt_fix, lines 195-237) on the two
   active tracers (jpi,jpj,jpk,2) */
int nemo_hip_tra_nxt(int kt, double *tsb, double *tsn, double *tsa);
/* dyn_nxt, DYN/dynnxt.F90:66-377.  ua_b/va_b (and with ln_bt_fw = F
   un_adv/vn_adv) are the dyn_spg_ts outputs; ub_b, vb_b, un_b, vn_b are
   updated; r1_hu_n/r1_hv_n stand for r1_hu_a and r1_hu_b, which equal them
   when the scale factors are fixed */
int nemo_hip_dyn_nxt(int kt, double *ub, double *vb, double *un, double *vn,
                     double *ua, double *va, const double *ua_b, const double *va_b,
                     const double *un_adv, const double *vn_adv,
                     double *ub_b, double *vb_b, double *un_b, double *vn_b,
                     const double *r1_hu_n, const double *r1_hv_n);

/* ------------------------------------------------------------------------ */
/* Fused steps on device-resident fields (pointer mode NEMO_HIP_PTR_DEVICE). */
/* Each equals, bit for bit, the sequence of the separate entry points in   */
/* the order of the time step, step.F90:179-257.                            */
/* ------------------------------------------------------------------------ */
/* ua = va = 0; dyn_adv (DYN/dynadv.F90:55); dyn_vor (DYN/dynvor.F90:98);
   dyn_spg_ts (DYN/dynspg_ts.F90:121) on st; dyn_zdf (DYN/dynzdf.F90:47) with
   ub, vb, avm, st->rCdU_bot, st->ua_b/va_b, st->utau/vtau and utau_b/vtau_b.
   The advection and vorticity trends read st->un/st->vn before dyn_spg_ts
   updates them.  st->ua/st->va need no initial value. */
int nemo_hip_step_dyn(int kt, const double *ub, const double *vb, const double *wn,
                      const double *avm, const double *utau_b, const double *vtau_b,
                      nemo_hip_spg_ts_t *st);
/* tsa = 0; tra_adv (TRA/traadv.F90:77, transports of un/vn/wn + tra_adv_fct
   with kn_fct_h/kn_fct_v); tra_ldf_iso (TRA/traldf_iso.F90:50, laplacian,
   kpass 1, ptbb = tsb, writes ah_wslp2 and with ln_traldf_msc akz on the
   interior); tra_zdf (TRA/trazdf.F90:46, avt for both tracers, iso-neutral
   coefficient, salinity reset).  Two tracers (T, S) in tsb/tsn/tsa. */
int nemo_hip_step_tra(int kt, const double *un, const double *vn, const double *wn,
                      const double *tsb, const double *tsn, double *tsa,
                      int kn_fct_h, int kn_fct_v,
                      const double *ahtu, const double *ahtv,
                      const double *uslp, const double *vslp,
                      const double *wslpi, const double *wslpj,
                      double *ah_wslp2, double *akz, const double *avt);

/* ------------------------------------------------------------------------ */
/* FP32-storage precision form: same arguments with float 3D fields.         */
/* Stencils run in FP32; the Thomas eliminations of tra_zdf/dyn_zdf, the     */
/* barotropic subcycles and all 2D barotropic state stay FP64.               */
/* ------------------------------------------------------------------------ */
int nemo_hip_tra_adv_trp_sp(const float *un, const float *vn, const float *wn,
                            float *zun, float *zvn, float *zwn);
int nemo_hip_tra_adv_fct_sp(int kt, int kit000, const char *cdtype, double p2dt,
                            const float *pun, const float *pvn, const float *pwn,
                            const float *ptb, const float *ptn, float *pta,
                            int kjpt, int kn_fct_h, int kn_fct_v);
int nemo_hip_tra_ldf_lap_sp(int kt, int kit000, const char *cdtype,
                            const float *pahu, const float *pahv,
                            const float *ptb, float *pta, int kjpt, int kpass);
int nemo_hip_tra_ldf_iso_sp(int kt, int kit000, const char *cdtype,
                            const float *pahu, const float *pahv,
                            const float *ptb, const float *ptbb, float *pta,
                            int kjpt, int kpass,
                            const float *uslp, const float *vslp,
                            const float *wslpi, const float *wslpj,
                            float *ah_wslp2, float *akz);
int nemo_hip_tra_zdf_sp(int kt, const float *tsb, float *tsa,
                        const float *avt, const float *avs, int l_ldfslp,
                        const float *ah_wslp2, const float *akz);
int nemo_hip_dyn_adv_sp(int kt, const float *un, const float *vn, const float *wn,
                        float *ua, float *va);
int nemo_hip_dyn_vor_sp(int kt, const float *un, const float *vn,
                        float *ua, float *va);
int nemo_hip_dyn_zdf_sp(int kt, const float *ub, const float *vb,
                        float *ua, float *va, const float *avm,
                        const double *rCdU_bot, const double *ua_b, const double *va_b,
                        const double *utau, const double *utau_b,
                        const double *vtau, const double *vtau_b);
/* 3D arrays in st are float* */
int nemo_hip_dyn_spg_ts_sp(int kt, nemo_hip_spg_ts_t *st);
int nemo_hip_step_dyn_sp(int kt, const float *ub, const float *vb, const float *wn,
                         const float *avm, const double *utau_b, const double *vtau_b,
                         nemo_hip_spg_ts_t *st);
int nemo_hip_step_tra_sp(int kt, const float *un, const float *vn, const float *wn,
                         const float *tsb, const float *tsn, float *tsa,
                         int kn_fct_h, int kn_fct_v,
                         const float *ahtu, const float *ahtv,
                         const float *uslp, const float *vslp,
                         const float *wslpi, const float *wslpj,
                         float *ah_wslp2, float *akz, const float *avt);
/* density, pressure gradient, lateral viscosity: float 3D fields; the
   arithmetic is FP64, rounded once on store */
int nemo_hip_eos_insitu_sp(const float *pts, float *prd, const float *pdep);
int nemo_hip_eos_insitu_pot_sp(const float *pts, float *prd, float *prhop, const float *pdep);
int nemo_hip_zdf_mxl_sp(int kt, const float *rn2b, const float *gdept, const float *gdepw, int nlb10, double rho_c,
                        int32_t *nmln, double *hmlp, double *hmlpt);
int nemo_hip_ldf_slp_sp(int kt, const float *prd, const float *pn2, const float *gdept, const float *gdepw,
                        const int32_t *nmln, const double *hmlp, const double *hmlpt, double rn_slpmax, float *uslp,
                        float *vslp, float *wslpi, float *wslpj);
int nemo_hip_rab_3d_sp(const float *pts, float *pab, const float *gdept);
int nemo_hip_bn2_sp(const float *pts, const float *pab, float *pn2,
                    const float *gdept, const float *gdepw);
int nemo_hip_hpg_zco_sp(int kt, const float *rhd, float *ua, float *va);
int nemo_hip_dyn_ldf_lap_sp(int kt, const float *pub, const float *pvb, float *pua, float *pva,
                            int kpass, const float *ahmt, const float *ahmf);
/* time-step closure: float 3D fields, FP64 2D fields and arithmetic */
int nemo_hip_div_hor_sp(int kt, const float *un, const float *vn, float *hdivn);
int nemo_hip_ssh_nxt_sp(int kt, const float *un, const float *vn, float *hdivn,
                        const double *sshb, const double *emp_b, const double *emp,
                        double *ssha);
int nemo_hip_wzv_sp(int kt, const float *hdivn, float *wn);
int nemo_hip_tra_nxt_sp(int kt, float *tsb, float *tsn, float *tsa);
int nemo_hip_dyn_nxt_sp(int kt, float *ub, float *vb, float *un, float *vn,
                        float *ua, float *va, const double *ua_b, const double *va_b,
                        const double *un_adv, const double *vn_adv,
                        double *ub_b, double *vb_b, double *un_b, double *vn_b,
                        const double *r1_hu_n, const double *r1_hv_n);

#ifdef __cplusplus
}
#endif
#endif /* NEMO_HIP_H */
