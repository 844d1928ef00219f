MODULE nemo_hip_mod
   !!======================================================================
   !!                     ***  MODULE  nemo_hip_mod  ***
   !! Drop-in replacements of NEMO 4.0 hot-path routines that run on an
   !! MI355X through the C-ABI of include/nemo_hip.h.
   !!
   !! Each public routine keeps the dummy-argument list of the routine it
   !! replaces (paths relative to src/OCE):
   !!   tra_adv_fct    TRA/traadv_fct.F90:54        interp_4th_cpt TRA/traadv_fct.F90:517
   !!   tra_ldf_lap    TRA/traldf_lap_blp.F90:47    tra_ldf_iso    TRA/traldf_iso.F90:50
   !!   tra_zdf_imp    TRA/trazdf.F90:109
   !!   dyn_keg        DYN/dynkeg.F90:46            dyn_zad        DYN/dynzad.F90:38
   !!   dyn_adv_cen2   DYN/dynadv_cen2.F90:37       dyn_vor        DYN/dynvor.F90:98
   !!   dyn_zdf        DYN/dynzdf.F90:47            dyn_spg_ts     DYN/dynspg_ts.F90:121
   !!   div_hor        DYN/divhor.F90:50            ssh_nxt        DYN/sshwzv.F90:56
   !!   wzv            DYN/sshwzv.F90:129           ssh_swp        DYN/sshwzv.F90:216
   !!   tra_nxt        TRA/tranxt.F90:66            dyn_nxt        DYN/dynnxt.F90:66
   !!   eos_insitu     TRA/eosbn2.F90:188           eos_insitu_pot TRA/eosbn2.F90:304
   !!   rab_3d         TRA/eosbn2.F90:566           bn2            TRA/eosbn2.F90:889
   !!   hpg_zco        DYN/dynhpg.F90:258           dyn_ldf_lap    DYN/dynldf_lap_blp.F90:37
   !!   tra_sbc        TRA/trasbc.F90:62            stp_ctl        stpctl.F90:55
   !!   zdf_mxl        ZDF/zdfmxl.F90:64            ldf_slp        LDF/ldfslp.F90:83
   !!   rst_write      IOM/restart.F90:133          rst_read       IOM/restart.F90:240
   !! and the module-global step routines that dispatch to them:
   !!   tra_adv        TRA/traadv.F90:77            tra_ldf        TRA/traldf.F90:49
   !!   tra_zdf        TRA/trazdf.F90:46            dyn_adv        DYN/dynadv.F90:55
   !!   dyn_spg        DYN/dynspg.F90:55
   !! The arrays those routines take from NEMO modules (oce, zdf_oce, ldfslp,
   !! sbc_oce, zdfdrg, dynspg_ts, dom_oce) are bound once, after allocation, with the
   !! nemo_hip_bind_* routines: module arrays never move, so their addresses
   !! stay valid for the run.  A nonzero status is turned into ctl_stop by
   !! the caller-supplied error hook (default: print and ERROR STOP).
   !!======================================================================
   USE, INTRINSIC :: ISO_C_BINDING
   IMPLICIT NONE
   PRIVATE

   INTEGER, PARAMETER :: wp = C_DOUBLE

   PUBLIC :: nemo_hip_start, nemo_hip_stop, nemo_hip_domain
   PUBLIC :: nemo_hip_bind_zdf, nemo_hip_bind_ldfslp, nemo_hip_bind_dyn, nemo_hip_bind_spg
   PUBLIC :: tra_adv_fct, interp_4th_cpt, tra_ldf_lap, tra_ldf_iso, tra_zdf_imp
   PUBLIC :: dyn_keg, dyn_zad, dyn_adv_cen2, dyn_vor, dyn_zdf, dyn_spg_ts
   PUBLIC :: nemo_hip_bind_nxt, div_hor, ssh_nxt, wzv, ssh_swp, tra_nxt, dyn_nxt
   PUBLIC :: nemo_hip_eos_t, nemo_hip_eos_init, nemo_hip_bind_phys
   PUBLIC :: eos_insitu, eos_insitu_pot, rab_3d, bn2, hpg_zco, dyn_ldf_lap, tra_sbc, tra_qsr, stp_ctl
   PUBLIC :: nemo_hip_bind_mxl, zdf_mxl, ldf_slp
   PUBLIC :: nemo_hip_rst_t, rst_write, rst_read
   PUBLIC :: nemo_hip_bind_tra, tra_adv, tra_ldf, tra_zdf, dyn_adv, dyn_spg
   !! ldftra.F90:79-80: the operators tra_ldf dispatches to
   INTEGER, PARAMETER, PUBLIC :: np_lap = 10, np_lap_i = 11
   PUBLIC :: nemo_hip_domain_t, nemo_hip_spg_ts_t
   PUBLIC :: nemo_hip_field_alloc, nemo_hip_field_free, nemo_hip_field_push, nemo_hip_field_pull

   !! mirror of nemo_hip_domain_t (same field order and kinds)
   TYPE, BIND(C) :: nemo_hip_domain_t
      INTEGER(C_INT32_T) :: jpi, jpj, jpk, jperio
      INTEGER(C_INT32_T) :: ln_linssh, ln_isfcav, ln_zps, ln_sco
      INTEGER(C_INT32_T) :: ln_traldf_lap, ln_traldf_msc, ln_zdfddm, ln_zad_Aimp
      INTEGER(C_INT32_T) :: n_dynadv, nn_dynkeg, nvor_scheme, nn_een_e3f
      INTEGER(C_INT32_T) :: ln_drgimp
      INTEGER(C_INT32_T) :: ln_dynspg_ts, ln_bt_fw, ln_bt_av, nn_bt_flt, nn_baro
      INTEGER(C_INT32_T) :: neuler, nit000, ln_traqsr
      REAL(C_DOUBLE)     :: rdt, rn_atfp, rau0, grav, rn_bt_alpha, rn_scal_load
      TYPE(C_PTR) :: k_top, k_bot
      TYPE(C_PTR) :: e1t, e2t, e1u, e2u, e1v, e2v, e1f, e2f
      TYPE(C_PTR) :: e1e2t, r1_e1e2t, e1e2u, r1_e1e2u, e1e2v, r1_e1e2v, e1e2f, r1_e1e2f
      TYPE(C_PTR) :: e2_e1u, e1_e2v, r1_e1u, r1_e2u, r1_e1v, r1_e2v
      TYPE(C_PTR) :: ff_f, ff_t
      TYPE(C_PTR) :: e3t_n, e3u_n, e3v_n, e3w_n, e3f_n, e3uw_n, e3vw_n
   END TYPE nemo_hip_domain_t

   !! mirror of nemo_hip_eos_t (namelist nameos of the simplified EOS)
   TYPE, BIND(C) :: nemo_hip_eos_t
      INTEGER(C_INT32_T) :: neos, reserved0
      REAL(C_DOUBLE)     :: rn_a0, rn_b0, rn_lambda1, rn_lambda2, rn_mu1, rn_mu2, rn_nu
   END TYPE nemo_hip_eos_t

   !! mirror of nemo_hip_spg_ts_t
   TYPE, BIND(C) :: nemo_hip_spg_ts_t
      TYPE(C_PTR) :: ua, va, un, vn
      TYPE(C_PTR) :: un_b, vn_b, ub_b, vb_b, sshn, hu_n, hv_n, r1_hu_n, r1_hv_n, ht_n
      TYPE(C_PTR) :: rCdU_bot, utau, vtau, emp, rnf
      TYPE(C_PTR) :: ua_b, va_b, ssha, un_adv, vn_adv
      TYPE(C_PTR) :: ub2_b, vb2_b, un_bf, vn_bf
   END TYPE nemo_hip_spg_ts_t

   TYPE, BIND(C) :: nemo_hip_rst_t
      INTEGER(C_INT32_T) :: real_bytes, kqsr_b
      TYPE(C_PTR) :: ub, vb, tsb, un, vn, tsn, rhop, sshb, sshn, ub2_b, vb2_b, un_bf, vn_bf
      TYPE(C_PTR) :: un_b, vn_b, ub_b, vb_b, r1_hu_n, r1_hv_n
      TYPE(C_PTR) :: qsr_hc, qsr_hc_b, fraqsr_1lev
   END TYPE nemo_hip_rst_t

   !! module arrays bound at initialisation
   TYPE(C_PTR), SAVE :: p_avt = C_NULL_PTR, p_avs = C_NULL_PTR, p_avm = C_NULL_PTR
   TYPE(C_PTR), SAVE :: p_uslp = C_NULL_PTR, p_vslp = C_NULL_PTR, p_wslpi = C_NULL_PTR, p_wslpj = C_NULL_PTR
   TYPE(C_PTR), SAVE :: p_ah_wslp2 = C_NULL_PTR, p_akz = C_NULL_PTR
   TYPE(C_PTR), SAVE :: p_un = C_NULL_PTR, p_vn = C_NULL_PTR, p_wn = C_NULL_PTR, p_ub = C_NULL_PTR, p_vb = C_NULL_PTR
   TYPE(C_PTR), SAVE :: p_ua = C_NULL_PTR, p_va = C_NULL_PTR
   TYPE(C_PTR), SAVE :: p_rCdU_bot = C_NULL_PTR, p_ua_b = C_NULL_PTR, p_va_b = C_NULL_PTR
   TYPE(C_PTR), SAVE :: p_utau = C_NULL_PTR, p_utau_b = C_NULL_PTR, p_vtau = C_NULL_PTR, p_vtau_b = C_NULL_PTR
   TYPE(C_PTR), SAVE :: p_hdivn = C_NULL_PTR, p_sshb = C_NULL_PTR, p_sshn = C_NULL_PTR, p_ssha = C_NULL_PTR
   TYPE(C_PTR), SAVE :: p_emp_b = C_NULL_PTR, p_emp = C_NULL_PTR
   TYPE(C_PTR), SAVE :: p_tsb = C_NULL_PTR, p_tsn = C_NULL_PTR, p_tsa = C_NULL_PTR
   TYPE(C_PTR), SAVE :: p_ub_b = C_NULL_PTR, p_vb_b = C_NULL_PTR, p_un_b = C_NULL_PTR, p_vn_b = C_NULL_PTR
   TYPE(C_PTR), SAVE :: p_un_adv = C_NULL_PTR, p_vn_adv = C_NULL_PTR, p_r1_hu_n = C_NULL_PTR, p_r1_hv_n = C_NULL_PTR
   TYPE(C_PTR), SAVE :: p_gdept = C_NULL_PTR, p_gdepw = C_NULL_PTR, p_rhd = C_NULL_PTR
   TYPE(C_PTR), SAVE :: p_rn2b = C_NULL_PTR, p_nmln = C_NULL_PTR, p_hmlp = C_NULL_PTR, p_hmlpt = C_NULL_PTR
   INTEGER(C_INT), SAVE :: m_nlb10 = 0
   REAL(wp), SAVE :: m_rho_c = 0.01_wp, m_slpmax = 0.01_wp
   TYPE(C_PTR), SAVE :: p_ahmt = C_NULL_PTR, p_ahmf = C_NULL_PTR
   TYPE(C_PTR), SAVE :: p_qns = C_NULL_PTR, p_qsr = C_NULL_PTR, p_sfx = C_NULL_PTR
   TYPE(C_PTR), SAVE :: p_sbc_tsc = C_NULL_PTR, p_sbc_tsc_b = C_NULL_PTR
   INTEGER, SAVE :: l_ldfslp = 0
   !! the traadv / ldftra namelist choices and arrays tra_adv and tra_ldf use
   INTEGER, SAVE :: m_nn_fct_h = 4, m_nn_fct_v = 4, m_nldf_tra = np_lap, m_nit000 = 1
   TYPE(C_PTR), SAVE :: p_ahtu = C_NULL_PTR, p_ahtv = C_NULL_PTR
   TYPE(nemo_hip_spg_ts_t), SAVE :: spg_state

   INTERFACE
      INTEGER(C_INT) FUNCTION c_init(device) BIND(C, name="nemo_hip_init")
         IMPORT :: C_INT
         INTEGER(C_INT), VALUE :: device
      END FUNCTION
      INTEGER(C_INT) FUNCTION c_finalize() BIND(C, name="nemo_hip_finalize")
         IMPORT :: C_INT
      END FUNCTION
      INTEGER(C_INT) FUNCTION c_last_error(buf, len) BIND(C, name="nemo_hip_last_error")
         IMPORT :: C_INT, C_CHAR
         CHARACTER(KIND=C_CHAR), DIMENSION(*) :: buf
         INTEGER(C_INT), VALUE :: len
      END FUNCTION
      INTEGER(C_INT) FUNCTION c_field_alloc(nbytes, dev) BIND(C, name="nemo_hip_field_alloc")
         IMPORT :: C_INT, C_SIZE_T, C_PTR
         INTEGER(C_SIZE_T), VALUE :: nbytes
         TYPE(C_PTR), INTENT(out) :: dev
      END FUNCTION
      INTEGER(C_INT) FUNCTION c_field_free(dev) BIND(C, name="nemo_hip_field_free")
         IMPORT :: C_INT, C_PTR
         TYPE(C_PTR), VALUE :: dev
      END FUNCTION
      INTEGER(C_INT) FUNCTION c_field_push(dev, host, nbytes) BIND(C, name="nemo_hip_field_push")
         IMPORT :: C_INT, C_SIZE_T, C_PTR
         TYPE(C_PTR), VALUE :: dev, host
         INTEGER(C_SIZE_T), VALUE :: nbytes
      END FUNCTION
      INTEGER(C_INT) FUNCTION c_field_pull(host, dev, nbytes) BIND(C, name="nemo_hip_field_pull")
         IMPORT :: C_INT, C_SIZE_T, C_PTR
         TYPE(C_PTR), VALUE :: host, dev
         INTEGER(C_SIZE_T), VALUE :: nbytes
      END FUNCTION
      INTEGER(C_INT) FUNCTION c_domain_register(dom) BIND(C, name="nemo_hip_domain_register")
         IMPORT :: C_INT, nemo_hip_domain_t
         TYPE(nemo_hip_domain_t), INTENT(in) :: dom
      END FUNCTION
      INTEGER(C_INT) FUNCTION c_tra_adv_fct(kt, kit000, cdtype, p2dt, pun, pvn, pwn, ptb, ptn, pta, kjpt,   &
         &                                  kn_fct_h, kn_fct_v) BIND(C, name="nemo_hip_tra_adv_fct")
         IMPORT :: C_INT, C_CHAR, C_DOUBLE, C_PTR
         INTEGER(C_INT), VALUE :: kt, kit000, kjpt, kn_fct_h, kn_fct_v
         CHARACTER(KIND=C_CHAR), DIMENSION(*) :: cdtype
         REAL(C_DOUBLE), VALUE :: p2dt
         TYPE(C_PTR), VALUE :: pun, pvn, pwn, ptb, ptn, pta
      END FUNCTION
      INTEGER(C_INT) FUNCTION c_tra_adv(kt, un, vn, wn, ptb, ptn, pta, kjpt, kn_fct_h, kn_fct_v)   &
         &                              BIND(C, name="nemo_hip_tra_adv")
         IMPORT :: C_INT, C_PTR
         INTEGER(C_INT), VALUE :: kt, kjpt, kn_fct_h, kn_fct_v
         TYPE(C_PTR), VALUE :: un, vn, wn, ptb, ptn, pta
      END FUNCTION
      INTEGER(C_INT) FUNCTION c_tra_zdf(kt, tsb, tsa, avt, avs, l_ldfslp, ah_wslp2, akz) BIND(C, name="nemo_hip_tra_zdf")
         IMPORT :: C_INT, C_PTR
         INTEGER(C_INT), VALUE :: kt, l_ldfslp
         TYPE(C_PTR), VALUE :: tsb, tsa, avt, avs, ah_wslp2, akz
      END FUNCTION
      INTEGER(C_INT) FUNCTION c_dyn_adv(kt, un, vn, wn, ua, va) BIND(C, name="nemo_hip_dyn_adv")
         IMPORT :: C_INT, C_PTR
         INTEGER(C_INT), VALUE :: kt
         TYPE(C_PTR), VALUE :: un, vn, wn, ua, va
      END FUNCTION
      INTEGER(C_INT) FUNCTION c_interp_4th_cpt(pin, pout) BIND(C, name="nemo_hip_interp_4th_cpt")
         IMPORT :: C_INT, C_PTR
         TYPE(C_PTR), VALUE :: pin, pout
      END FUNCTION
      INTEGER(C_INT) FUNCTION c_tra_ldf_lap(kt, kit000, cdtype, pahu, pahv, pgu, pgv, pgui, pgvi, ptb, pta,   &
         &                                  kjpt, kpass) BIND(C, name="nemo_hip_tra_ldf_lap")
         IMPORT :: C_INT, C_CHAR, C_PTR
         INTEGER(C_INT), VALUE :: kt, kit000, kjpt, kpass
         CHARACTER(KIND=C_CHAR), DIMENSION(*) :: cdtype
         TYPE(C_PTR), VALUE :: pahu, pahv, pgu, pgv, pgui, pgvi, ptb, pta
      END FUNCTION
      INTEGER(C_INT) FUNCTION c_tra_ldf_iso(kt, kit000, cdtype, pahu, pahv, pgu, pgv, pgui, pgvi, ptb, ptbb, pta, &
         &                                  kjpt, kpass, uslp, vslp, wslpi, wslpj, ah_wslp2, akz)             &
         &                                  BIND(C, name="nemo_hip_tra_ldf_iso")
         IMPORT :: C_INT, C_CHAR, C_PTR
         INTEGER(C_INT), VALUE :: kt, kit000, kjpt, kpass
         CHARACTER(KIND=C_CHAR), DIMENSION(*) :: cdtype
         TYPE(C_PTR), VALUE :: pahu, pahv, pgu, pgv, pgui, pgvi, ptb, ptbb, pta
         TYPE(C_PTR), VALUE :: uslp, vslp, wslpi, wslpj, ah_wslp2, akz
      END FUNCTION
      INTEGER(C_INT) FUNCTION c_tra_zdf_imp(kt, kit000, cdtype, p2dt, ptb, pta, kjpt, avt, avs, l_ldfslp,   &
         &                                  ah_wslp2, akz) BIND(C, name="nemo_hip_tra_zdf_imp")
         IMPORT :: C_INT, C_CHAR, C_DOUBLE, C_PTR
         INTEGER(C_INT), VALUE :: kt, kit000, kjpt, l_ldfslp
         CHARACTER(KIND=C_CHAR), DIMENSION(*) :: cdtype
         REAL(C_DOUBLE), VALUE :: p2dt
         TYPE(C_PTR), VALUE :: ptb, pta, avt, avs, ah_wslp2, akz
      END FUNCTION
      INTEGER(C_INT) FUNCTION c_dyn_keg(kt, kscheme, un, vn, ua, va) BIND(C, name="nemo_hip_dyn_keg")
         IMPORT :: C_INT, C_PTR
         INTEGER(C_INT), VALUE :: kt, kscheme
         TYPE(C_PTR), VALUE :: un, vn, ua, va
      END FUNCTION
      INTEGER(C_INT) FUNCTION c_dyn_zad(kt, un, vn, wn, ua, va) BIND(C, name="nemo_hip_dyn_zad")
         IMPORT :: C_INT, C_PTR
         INTEGER(C_INT), VALUE :: kt
         TYPE(C_PTR), VALUE :: un, vn, wn, ua, va
      END FUNCTION
      INTEGER(C_INT) FUNCTION c_dyn_adv_cen2(kt, un, vn, wn, ua, va) BIND(C, name="nemo_hip_dyn_adv_cen2")
         IMPORT :: C_INT, C_PTR
         INTEGER(C_INT), VALUE :: kt
         TYPE(C_PTR), VALUE :: un, vn, wn, ua, va
      END FUNCTION
      INTEGER(C_INT) FUNCTION c_dyn_vor(kt, un, vn, ua, va) BIND(C, name="nemo_hip_dyn_vor")
         IMPORT :: C_INT, C_PTR
         INTEGER(C_INT), VALUE :: kt
         TYPE(C_PTR), VALUE :: un, vn, ua, va
      END FUNCTION
      INTEGER(C_INT) FUNCTION c_dyn_zdf(kt, ub, vb, ua, va, avm, rCdU_bot, ua_b, va_b, utau, utau_b, vtau, vtau_b) &
         &                              BIND(C, name="nemo_hip_dyn_zdf")
         IMPORT :: C_INT, C_PTR
         INTEGER(C_INT), VALUE :: kt
         TYPE(C_PTR), VALUE :: ub, vb, ua, va, avm, rCdU_bot, ua_b, va_b, utau, utau_b, vtau, vtau_b
      END FUNCTION
      INTEGER(C_INT) FUNCTION c_dyn_spg_ts(kt, st) BIND(C, name="nemo_hip_dyn_spg_ts")
         IMPORT :: C_INT, nemo_hip_spg_ts_t
         INTEGER(C_INT), VALUE :: kt
         TYPE(nemo_hip_spg_ts_t), INTENT(inout) :: st
      END FUNCTION
      INTEGER(C_INT) FUNCTION c_div_hor(kt, un, vn, hdivn) BIND(C, name="nemo_hip_div_hor")
         IMPORT :: C_INT, C_PTR
         INTEGER(C_INT), VALUE :: kt
         TYPE(C_PTR), VALUE :: un, vn, hdivn
      END FUNCTION
      INTEGER(C_INT) FUNCTION c_ssh_nxt(kt, un, vn, hdivn, sshb, emp_b, emp, ssha) BIND(C, name="nemo_hip_ssh_nxt")
         IMPORT :: C_INT, C_PTR
         INTEGER(C_INT), VALUE :: kt
         TYPE(C_PTR), VALUE :: un, vn, hdivn, sshb, emp_b, emp, ssha
      END FUNCTION
      INTEGER(C_INT) FUNCTION c_wzv(kt, hdivn, wn) BIND(C, name="nemo_hip_wzv")
         IMPORT :: C_INT, C_PTR
         INTEGER(C_INT), VALUE :: kt
         TYPE(C_PTR), VALUE :: hdivn, wn
      END FUNCTION
      INTEGER(C_INT) FUNCTION c_ssh_swp(kt, sshb, sshn, ssha) BIND(C, name="nemo_hip_ssh_swp")
         IMPORT :: C_INT, C_PTR
         INTEGER(C_INT), VALUE :: kt
         TYPE(C_PTR), VALUE :: sshb, sshn, ssha
      END FUNCTION
      INTEGER(C_INT) FUNCTION c_tra_nxt(kt, tsb, tsn, tsa) BIND(C, name="nemo_hip_tra_nxt")
         IMPORT :: C_INT, C_PTR
         INTEGER(C_INT), VALUE :: kt
         TYPE(C_PTR), VALUE :: tsb, tsn, tsa
      END FUNCTION
      INTEGER(C_INT) FUNCTION c_dyn_nxt(kt, ub, vb, un, vn, ua, va, ua_b, va_b, un_adv, vn_adv, ub_b, vb_b,   &
         &                              un_b, vn_b, r1_hu_n, r1_hv_n) BIND(C, name="nemo_hip_dyn_nxt")
         IMPORT :: C_INT, C_PTR
         INTEGER(C_INT), VALUE :: kt
         TYPE(C_PTR), VALUE :: ub, vb, un, vn, ua, va, ua_b, va_b, un_adv, vn_adv, ub_b, vb_b, un_b, vn_b
         TYPE(C_PTR), VALUE :: r1_hu_n, r1_hv_n
      END FUNCTION
      INTEGER(C_INT) FUNCTION c_eos_register(eos) BIND(C, name="nemo_hip_eos_register")
         IMPORT :: C_INT, nemo_hip_eos_t
         TYPE(nemo_hip_eos_t), INTENT(in) :: eos
      END FUNCTION
      INTEGER(C_INT) FUNCTION c_eos_insitu(pts, prd, pdep) BIND(C, name="nemo_hip_eos_insitu")
         IMPORT :: C_INT, C_PTR
         TYPE(C_PTR), VALUE :: pts, prd, pdep
      END FUNCTION
      INTEGER(C_INT) FUNCTION c_eos_insitu_pot(pts, prd, prhop, pdep) BIND(C, name="nemo_hip_eos_insitu_pot")
         IMPORT :: C_INT, C_PTR
         TYPE(C_PTR), VALUE :: pts, prd, prhop, pdep
      END FUNCTION
      INTEGER(C_INT) FUNCTION c_rab_3d(pts, pab, gdept) BIND(C, name="nemo_hip_rab_3d")
         IMPORT :: C_INT, C_PTR
         TYPE(C_PTR), VALUE :: pts, pab, gdept
      END FUNCTION
      INTEGER(C_INT) FUNCTION c_bn2(pts, pab, pn2, gdept, gdepw) BIND(C, name="nemo_hip_bn2")
         IMPORT :: C_INT, C_PTR
         TYPE(C_PTR), VALUE :: pts, pab, pn2, gdept, gdepw
      END FUNCTION
      INTEGER(C_INT) FUNCTION c_rst_write(dir, kt, rst) BIND(C, name="nemo_hip_rst_write")
         IMPORT :: C_INT, C_CHAR, nemo_hip_rst_t
         CHARACTER(KIND=C_CHAR), DIMENSION(*), INTENT(in) :: dir
         INTEGER(C_INT), VALUE :: kt
         TYPE(nemo_hip_rst_t), INTENT(in) :: rst
      END FUNCTION
      INTEGER(C_INT) FUNCTION c_rst_read(dir, rst, kt, neuler) BIND(C, name="nemo_hip_rst_read")
         IMPORT :: C_INT, C_CHAR, nemo_hip_rst_t
         CHARACTER(KIND=C_CHAR), DIMENSION(*), INTENT(in) :: dir
         TYPE(nemo_hip_rst_t), INTENT(inout) :: rst
         INTEGER(C_INT), INTENT(out) :: kt, neuler
      END FUNCTION
      INTEGER(C_INT) FUNCTION c_zdf_mxl(kt, rn2b, gdept, gdepw, nlb10, rho_c, nmln, hmlp, hmlpt) &
         &                          BIND(C, name="nemo_hip_zdf_mxl")
         IMPORT :: C_INT, C_PTR, C_DOUBLE
         INTEGER(C_INT), VALUE :: kt, nlb10
         REAL(C_DOUBLE), VALUE :: rho_c
         TYPE(C_PTR), VALUE :: rn2b, gdept, gdepw, nmln, hmlp, hmlpt
      END FUNCTION
      INTEGER(C_INT) FUNCTION c_ldf_slp(kt, prd, pn2, gdept, gdepw, nmln, hmlp, hmlpt, rn_slpmax, uslp, vslp, &
         &                              wslpi, wslpj) BIND(C, name="nemo_hip_ldf_slp")
         IMPORT :: C_INT, C_PTR, C_DOUBLE
         INTEGER(C_INT), VALUE :: kt
         REAL(C_DOUBLE), VALUE :: rn_slpmax
         TYPE(C_PTR), VALUE :: prd, pn2, gdept, gdepw, nmln, hmlp, hmlpt, uslp, vslp, wslpi, wslpj
      END FUNCTION
      INTEGER(C_INT) FUNCTION c_hpg_zco(kt, rhd, ua, va) BIND(C, name="nemo_hip_hpg_zco")
         IMPORT :: C_INT, C_PTR
         INTEGER(C_INT), VALUE :: kt
         TYPE(C_PTR), VALUE :: rhd, ua, va
      END FUNCTION
      INTEGER(C_INT) FUNCTION c_dyn_ldf_lap(kt, pub, pvb, pua, pva, kpass, ahmt, ahmf) BIND(C, name="nemo_hip_dyn_ldf_lap")
         IMPORT :: C_INT, C_PTR
         INTEGER(C_INT), VALUE :: kt, kpass
         TYPE(C_PTR), VALUE :: pub, pvb, pua, pva, ahmt, ahmf
      END FUNCTION
      INTEGER(C_INT) FUNCTION c_tra_qsr(kt, kqsr_b, nksr, rn_abs, rn_si0, rn_si1, qsr, gdepw, qsr_hc, qsr_hc_b, tsa,  &
         &                              fraqsr_1lev) BIND(C, name="nemo_hip_tra_qsr")
         IMPORT :: C_INT, C_PTR, C_DOUBLE
         INTEGER(C_INT), VALUE :: kt, kqsr_b, nksr
         REAL(C_DOUBLE), VALUE :: rn_abs, rn_si0, rn_si1
         TYPE(C_PTR), VALUE :: qsr, gdepw, qsr_hc, qsr_hc_b, tsa, fraqsr_1lev
      END FUNCTION
      INTEGER(C_INT) FUNCTION c_tra_sbc(kt, qns, qsr, sfx, emp, tsn, tsa, sbc_tsc, sbc_tsc_b) BIND(C, name="nemo_hip_tra_sbc")
         IMPORT :: C_INT, C_PTR
         INTEGER(C_INT), VALUE :: kt
         TYPE(C_PTR), VALUE :: qns, qsr, sfx, emp, tsn, tsa, sbc_tsc, sbc_tsc_b
      END FUNCTION
      INTEGER(C_INT) FUNCTION c_stp_ctl(kt, sshn, un, tsn, zmax, loc, kindic) BIND(C, name="nemo_hip_stp_ctl")
         IMPORT :: C_INT, C_PTR, C_INT32_T, C_DOUBLE
         INTEGER(C_INT), VALUE :: kt
         TYPE(C_PTR), VALUE :: sshn, un, tsn
         REAL(C_DOUBLE), INTENT(out) :: zmax(6)
         INTEGER(C_INT32_T), INTENT(out) :: loc(11)
         INTEGER(C_INT), INTENT(inout) :: kindic
      END FUNCTION
   END INTERFACE

CONTAINS

   SUBROUTINE check( kstat, cdname )
      !! turns a library status into a stop, the way ctl_stop ends a NEMO run
      INTEGER(C_INT)  , INTENT(in) :: kstat
      CHARACTER(len=*), INTENT(in) :: cdname
      CHARACTER(KIND=C_CHAR) :: buf(512)
      CHARACTER(len=512) :: msg
      INTEGER :: i, ist
      IF( kstat == 0 ) RETURN
      ist = c_last_error( buf, 512_C_INT )
      msg = ' '
      DO i = 1, 512
         IF( buf(i) == C_NULL_CHAR ) EXIT
         msg(i:i) = buf(i)
      END DO
      WRITE(*,'(a,a,a,i0,a,a)') 'nemo_hip: ', cdname, ' failed (status ', kstat, '): ', TRIM(msg)
      ERROR STOP 'nemo_hip'
   END SUBROUTINE check

   FUNCTION cstr( cd ) RESULT( c )
      CHARACTER(len=*), INTENT(in) :: cd
      CHARACTER(KIND=C_CHAR) :: c(LEN(cd)+1)
      INTEGER :: i
      DO i = 1, LEN(cd)
         c(i) = cd(i:i)
      END DO
      c(LEN(cd)+1) = C_NULL_CHAR
   END FUNCTION cstr

   SUBROUTINE nemo_hip_start( kdevice )
      INTEGER, INTENT(in) :: kdevice
      CALL check( c_init( INT(kdevice, C_INT) ), 'nemo_hip_init' )
   END SUBROUTINE nemo_hip_start

   SUBROUTINE nemo_hip_stop
      CALL check( c_finalize(), 'nemo_hip_finalize' )
   END SUBROUTINE nemo_hip_stop

   SUBROUTINE nemo_hip_domain( sdom )
      !! sdom: sizes, switches and C_LOC of the dom_oce arrays (see nemo_hip.h)
      TYPE(nemo_hip_domain_t), INTENT(in) :: sdom
      CALL check( c_domain_register( sdom ), 'nemo_hip_domain_register' )
      m_nit000 = sdom%nit000
   END SUBROUTINE nemo_hip_domain

   SUBROUTINE nemo_hip_bind_tra( kn_fct_h, kn_fct_v, knldf_tra, pahtu, pahtv )
      !! the traadv namelist orders (nn_fct_h, nn_fct_v), ldftra's operator
      !! (nldf_tra: np_lap or np_lap_i) and diffusivities (ahtu, ahtv) that
      !! tra_adv(kt) and tra_ldf(kt) use; the tracers are those bound with
      !! nemo_hip_bind_nxt (tsb, tsn, tsa), the velocities with nemo_hip_bind_dyn
      INTEGER, INTENT(in) :: kn_fct_h, kn_fct_v, knldf_tra
      REAL(wp), DIMENSION(:,:,:), TARGET, CONTIGUOUS, INTENT(in) :: pahtu, pahtv
      m_nn_fct_h = kn_fct_h ; m_nn_fct_v = kn_fct_v ; m_nldf_tra = knldf_tra
      p_ahtu = C_LOC( pahtu ) ; p_ahtv = C_LOC( pahtv )
   END SUBROUTINE nemo_hip_bind_tra

   !!----------------------------------------------------------------------
   !! resident fields (device mode): a device mirror per prognostic array,
   !! pushed once, stepped in HBM, pulled for output and restarts
   !!----------------------------------------------------------------------
   FUNCTION nemo_hip_field_alloc( kbytes ) RESULT( pdev )
      INTEGER(C_SIZE_T), INTENT(in) :: kbytes
      TYPE(C_PTR) :: pdev
      CALL check( c_field_alloc( kbytes, pdev ), 'nemo_hip_field_alloc' )
   END FUNCTION nemo_hip_field_alloc

   SUBROUTINE nemo_hip_field_free( pdev )
      TYPE(C_PTR), INTENT(in) :: pdev
      CALL check( c_field_free( pdev ), 'nemo_hip_field_free' )
   END SUBROUTINE nemo_hip_field_free

   SUBROUTINE nemo_hip_field_push( pdev, ptab )
      !! whole array, host -> device
      TYPE(C_PTR), INTENT(in) :: pdev
      REAL(wp), DIMENSION(..), TARGET, CONTIGUOUS, INTENT(in) :: ptab
      CALL check( c_field_push( pdev, C_LOC(ptab), INT(SIZE(ptab), C_SIZE_T) * 8_C_SIZE_T ), 'nemo_hip_field_push' )
   END SUBROUTINE nemo_hip_field_push

   SUBROUTINE nemo_hip_field_pull( ptab, pdev )
      !! whole array, device -> host, after the work queued on the library stream
      REAL(wp), DIMENSION(..), TARGET, CONTIGUOUS, INTENT(inout) :: ptab
      TYPE(C_PTR), INTENT(in) :: pdev
      CALL check( c_field_pull( C_LOC(ptab), pdev, INT(SIZE(ptab), C_SIZE_T) * 8_C_SIZE_T ), 'nemo_hip_field_pull' )
   END SUBROUTINE nemo_hip_field_pull

   SUBROUTINE nemo_hip_bind_zdf( pavt, pavs, pavm )
      !! zdf_oce module arrays (avt, avs, avm)
      REAL(wp), DIMENSION(:,:,:), TARGET, CONTIGUOUS, INTENT(in) :: pavt, pavs, pavm
      p_avt = C_LOC( pavt ) ; p_avs = C_LOC( pavs ) ; p_avm = C_LOC( pavm )
   END SUBROUTINE nemo_hip_bind_zdf

   SUBROUTINE nemo_hip_bind_ldfslp( puslp, pvslp, pwslpi, pwslpj, pah_wslp2, pakz )
      !! ldfslp module arrays; binding them switches on l_ldfslp for tra_zdf_imp
      REAL(wp), DIMENSION(:,:,:), TARGET, CONTIGUOUS, INTENT(in) :: puslp, pvslp, pwslpi, pwslpj, pah_wslp2, pakz
      p_uslp = C_LOC( puslp ) ; p_vslp = C_LOC( pvslp ) ; p_wslpi = C_LOC( pwslpi ) ; p_wslpj = C_LOC( pwslpj )
      p_ah_wslp2 = C_LOC( pah_wslp2 ) ; p_akz = C_LOC( pakz )
      l_ldfslp = 1
   END SUBROUTINE nemo_hip_bind_ldfslp

   SUBROUTINE nemo_hip_bind_dyn( pun, pvn, pwn, pub, pvb, pua, pva, prCdU_bot, pua_b, pva_b, putau, putau_b, pvtau, pvtau_b )
      !! oce (un, vn, wn, ub, vb, ua, va), zdfdrg (rCdU_bot), oce (ua_b, va_b), sbc_oce (utau, vtau, ...)
      REAL(wp), DIMENSION(:,:,:), TARGET, CONTIGUOUS, INTENT(in) :: pun, pvn, pwn, pub, pvb, pua, pva
      REAL(wp), DIMENSION(:,:)  , TARGET, CONTIGUOUS, INTENT(in) :: prCdU_bot, pua_b, pva_b, putau, putau_b, pvtau, pvtau_b
      p_un = C_LOC( pun ) ; p_vn = C_LOC( pvn ) ; p_wn = C_LOC( pwn ) ; p_ub = C_LOC( pub ) ; p_vb = C_LOC( pvb )
      p_ua = C_LOC( pua ) ; p_va = C_LOC( pva )
      p_rCdU_bot = C_LOC( prCdU_bot ) ; p_ua_b = C_LOC( pua_b ) ; p_va_b = C_LOC( pva_b )
      p_utau = C_LOC( putau ) ; p_utau_b = C_LOC( putau_b ) ; p_vtau = C_LOC( pvtau ) ; p_vtau_b = C_LOC( pvtau_b )
   END SUBROUTINE nemo_hip_bind_dyn

   SUBROUTINE nemo_hip_bind_nxt( phdivn, psshb, psshn, pssha, pemp_b, pemp, ptsb, ptsn, ptsa,   &
      &                          pub_b, pvb_b, pun_b, pvn_b, pun_adv, pvn_adv, pr1_hu_n, pr1_hv_n )
      !! oce (hdivn, sshb/n/a, tsb/n/a, ub_b ... vn_b), sbc_oce (emp_b, emp), dynspg_ts (un_adv, vn_adv),
      !! dom_oce (r1_hu_n, r1_hv_n) for the time-step closure
      REAL(wp), DIMENSION(:,:,:)  , TARGET, CONTIGUOUS, INTENT(in) :: phdivn
      REAL(wp), DIMENSION(:,:)    , TARGET, CONTIGUOUS, INTENT(in) :: psshb, psshn, pssha, pemp_b, pemp
      REAL(wp), DIMENSION(:,:,:,:), TARGET, CONTIGUOUS, INTENT(in) :: ptsb, ptsn, ptsa
      REAL(wp), DIMENSION(:,:)    , TARGET, CONTIGUOUS, INTENT(in) :: pub_b, pvb_b, pun_b, pvn_b, pun_adv, pvn_adv
      REAL(wp), DIMENSION(:,:)    , TARGET, CONTIGUOUS, INTENT(in) :: pr1_hu_n, pr1_hv_n
      p_hdivn = C_LOC( phdivn ) ; p_sshb = C_LOC( psshb ) ; p_sshn = C_LOC( psshn ) ; p_ssha = C_LOC( pssha )
      p_emp_b = C_LOC( pemp_b ) ; p_emp = C_LOC( pemp )
      p_tsb = C_LOC( ptsb ) ; p_tsn = C_LOC( ptsn ) ; p_tsa = C_LOC( ptsa )
      p_ub_b = C_LOC( pub_b ) ; p_vb_b = C_LOC( pvb_b ) ; p_un_b = C_LOC( pun_b ) ; p_vn_b = C_LOC( pvn_b )
      p_un_adv = C_LOC( pun_adv ) ; p_vn_adv = C_LOC( pvn_adv )
      p_r1_hu_n = C_LOC( pr1_hu_n ) ; p_r1_hv_n = C_LOC( pr1_hv_n )
   END SUBROUTINE nemo_hip_bind_nxt

   SUBROUTINE nemo_hip_eos_init( seos )
      !! eos_init (TRA/eosbn2.F90:1218): seos%neos selects the polynomial
      !! TEOS-10 (-1) or EOS-80 (0), or the simplified EOS (1) with the
      !! nameos parameters in the rn_* fields
      TYPE(nemo_hip_eos_t), INTENT(in) :: seos
      CALL check( c_eos_register( seos ), 'eos_init' )
   END SUBROUTINE nemo_hip_eos_init

   SUBROUTINE nemo_hip_bind_mxl( prn2b, knmln, phmlp, phmlpt, knlb10, prho_c, pslpmax )
      !! zdfmxl (nmln, hmlp, hmlpt) and zdf_oce (rn2b) arrays, dom_oce nlb10,
      !! the density criterion rho_c and the nam_traldf slope limit rn_slpmax
      REAL(wp), DIMENSION(:,:,:), TARGET, CONTIGUOUS, INTENT(in) :: prn2b
      INTEGER(C_INT32_T), DIMENSION(:,:), TARGET, CONTIGUOUS, INTENT(in) :: knmln
      REAL(wp), DIMENSION(:,:), TARGET, CONTIGUOUS, INTENT(in) :: phmlp, phmlpt
      INTEGER, INTENT(in) :: knlb10
      REAL(wp), INTENT(in) :: prho_c, pslpmax
      p_rn2b = C_LOC( prn2b ) ; p_nmln = C_LOC( knmln ) ; p_hmlp = C_LOC( phmlp ) ; p_hmlpt = C_LOC( phmlpt )
      m_nlb10 = knlb10 ; m_rho_c = prho_c ; m_slpmax = pslpmax
   END SUBROUTINE nemo_hip_bind_mxl

   SUBROUTINE nemo_hip_bind_phys( pgdept, pgdepw, prhd, pahmt, pahmf, pqns, pqsr, psfx, psbc_tsc, psbc_tsc_b )
      !! dom_oce (gdept_n, gdepw_n), oce (rhd), ldfdyn (ahmt, ahmf), sbc_oce (qns, qsr, sfx, sbc_tsc, sbc_tsc_b)
      REAL(wp), DIMENSION(:,:,:), TARGET, CONTIGUOUS, INTENT(in) :: pgdept, pgdepw, prhd, pahmt, pahmf
      REAL(wp), DIMENSION(:,:)  , TARGET, CONTIGUOUS, INTENT(in) :: pqns, pqsr, psfx
      REAL(wp), DIMENSION(:,:,:), TARGET, CONTIGUOUS, INTENT(in) :: psbc_tsc, psbc_tsc_b
      p_gdept = C_LOC( pgdept ) ; p_gdepw = C_LOC( pgdepw ) ; p_rhd = C_LOC( prhd )
      p_ahmt = C_LOC( pahmt ) ; p_ahmf = C_LOC( pahmf )
      p_qns = C_LOC( pqns ) ; p_qsr = C_LOC( pqsr ) ; p_sfx = C_LOC( psfx )
      p_sbc_tsc = C_LOC( psbc_tsc ) ; p_sbc_tsc_b = C_LOC( psbc_tsc_b )
   END SUBROUTINE nemo_hip_bind_phys

   SUBROUTINE nemo_hip_bind_spg( sstate )
      !! C_LOC of every array dyn_spg_ts reads or writes (oce, dom_oce, sbc_oce, dynspg_ts)
      TYPE(nemo_hip_spg_ts_t), INTENT(in) :: sstate
      spg_state = sstate
   END SUBROUTINE nemo_hip_bind_spg

   !!----------------------------------------------------------------------
   !! drop-in routines
   !!----------------------------------------------------------------------
   SUBROUTINE tra_adv( kt )
      !! TRA/traadv.F90:77 with the FCT scheme: r2dt, the effective transports
      !! (formed on the device where the fluxes use them) and tra_adv_fct on
      !! (T,S), added to tsa
      INTEGER, INTENT(in) ::   kt
      CALL check( c_tra_adv( INT(kt,C_INT), p_un, p_vn, p_wn, p_tsb, p_tsn, p_tsa, 2_C_INT,                    &
         &                   INT(m_nn_fct_h,C_INT), INT(m_nn_fct_v,C_INT) ), 'tra_adv' )
   END SUBROUTINE tra_adv

   SUBROUTINE tra_ldf( kt )
      !! TRA/traldf.F90:49: the laplacian operator nldf_tra selects, kpass 1
      INTEGER, INTENT(in) ::   kt
      SELECT CASE( m_nldf_tra )
      CASE( np_lap )
         CALL check( c_tra_ldf_lap( INT(kt,C_INT), INT(m_nit000,C_INT), cstr('TRA'), p_ahtu, p_ahtv,             &
            &        p_ahtu, p_ahtu, p_ahtu, p_ahtu, p_tsb, p_tsa, 2_C_INT, 1_C_INT ), 'tra_ldf' )
      CASE( np_lap_i )
         CALL check( c_tra_ldf_iso( INT(kt,C_INT), INT(m_nit000,C_INT), cstr('TRA'), p_ahtu, p_ahtv,             &
            &        p_ahtu, p_ahtu, p_ahtu, p_ahtu, p_tsb, p_tsb, p_tsa, 2_C_INT, 1_C_INT,                         &
            &        p_uslp, p_vslp, p_wslpi, p_wslpj, p_ah_wslp2, p_akz ), 'tra_ldf' )
      CASE DEFAULT
         CALL check( 2_C_INT, 'tra_ldf: only np_lap and np_lap_i are on the device' )
      END SELECT
   END SUBROUTINE tra_ldf

   SUBROUTINE tra_zdf( kt )
      !! TRA/trazdf.F90:46: r2dt, tra_zdf_imp on (T,S) and the reset of negative salinity
      INTEGER, INTENT(in) ::   kt
      CALL check( c_tra_zdf( INT(kt,C_INT), p_tsb, p_tsa, p_avt, p_avs, INT(l_ldfslp,C_INT), p_ah_wslp2, p_akz ),   &
         &        'tra_zdf' )
   END SUBROUTINE tra_zdf

   SUBROUTINE dyn_adv( kt )
      !! DYN/dynadv.F90:55: vector form (dyn_keg + dyn_zad) or flux form
      !! (dyn_adv_cen2), as the registered n_dynadv and nn_dynkeg select
      INTEGER, INTENT(in) ::   kt
      CALL check( c_dyn_adv( INT(kt,C_INT), p_un, p_vn, p_wn, p_ua, p_va ), 'dyn_adv' )
   END SUBROUTINE dyn_adv

   SUBROUTINE dyn_spg( kt )
      !! DYN/dynspg.F90:55 with ln_dynspg_ts and no atmospheric pressure, tide
      !! or ice loading (the only trend it adds is dyn_spg_ts's)
      INTEGER, INTENT(in) ::   kt
      CALL dyn_spg_ts( kt )
   END SUBROUTINE dyn_spg

   SUBROUTINE tra_adv_fct( kt, kit000, cdtype, p2dt, pun, pvn, pwn, ptb, ptn, pta, kjpt, kn_fct_h, kn_fct_v )
      INTEGER                            , INTENT(in   ) ::   kt, kit000, kjpt, kn_fct_h, kn_fct_v
      CHARACTER(len=3)                   , INTENT(in   ) ::   cdtype
      REAL(wp)                           , INTENT(in   ) ::   p2dt
      REAL(wp), DIMENSION(:,:,:)  , TARGET, CONTIGUOUS, INTENT(in   ) ::   pun, pvn, pwn
      REAL(wp), DIMENSION(:,:,:,:), TARGET, CONTIGUOUS, INTENT(in   ) ::   ptb, ptn
      REAL(wp), DIMENSION(:,:,:,:), TARGET, CONTIGUOUS, INTENT(inout) ::   pta
      CALL check( c_tra_adv_fct( INT(kt,C_INT), INT(kit000,C_INT), cstr(cdtype), REAL(p2dt,C_DOUBLE),         &
         &        C_LOC(pun), C_LOC(pvn), C_LOC(pwn), C_LOC(ptb), C_LOC(ptn), C_LOC(pta),                         &
         &        INT(kjpt,C_INT), INT(kn_fct_h,C_INT), INT(kn_fct_v,C_INT) ), 'tra_adv_fct' )
   END SUBROUTINE tra_adv_fct

   SUBROUTINE interp_4th_cpt( pt_in, pt_out )
      REAL(wp), DIMENSION(:,:,:), TARGET, CONTIGUOUS, INTENT(in   ) ::   pt_in
      REAL(wp), DIMENSION(:,:,:), TARGET, CONTIGUOUS, INTENT(inout) ::   pt_out
      CALL check( c_interp_4th_cpt( C_LOC(pt_in), C_LOC(pt_out) ), 'interp_4th_cpt' )
   END SUBROUTINE interp_4th_cpt

   SUBROUTINE tra_ldf_lap( kt, kit000, cdtype, pahu, pahv, pgu, pgv, pgui, pgvi, ptb, pta, kjpt, kpass )
      INTEGER                            , INTENT(in   ) ::   kt, kit000, kjpt, kpass
      CHARACTER(len=3)                   , INTENT(in   ) ::   cdtype
      REAL(wp), DIMENSION(:,:,:)  , TARGET, CONTIGUOUS, INTENT(in   ) ::   pahu, pahv
      REAL(wp), DIMENSION(:,:,:)  , TARGET, CONTIGUOUS, INTENT(in   ) ::   pgu, pgv, pgui, pgvi
      REAL(wp), DIMENSION(:,:,:,:), TARGET, CONTIGUOUS, INTENT(in   ) ::   ptb
      REAL(wp), DIMENSION(:,:,:,:), TARGET, CONTIGUOUS, INTENT(inout) ::   pta
      CALL check( c_tra_ldf_lap( INT(kt,C_INT), INT(kit000,C_INT), cstr(cdtype), C_LOC(pahu), C_LOC(pahv),     &
         &        C_LOC(pgu), C_LOC(pgv), C_LOC(pgui), C_LOC(pgvi), C_LOC(ptb), C_LOC(pta),                       &
         &        INT(kjpt,C_INT), INT(kpass,C_INT) ), 'tra_ldf_lap' )
   END SUBROUTINE tra_ldf_lap

   SUBROUTINE tra_ldf_iso( kt, kit000, cdtype, pahu, pahv, pgu, pgv, pgui, pgvi, ptb, ptbb, pta, kjpt, kpass )
      INTEGER                            , INTENT(in   ) ::   kt, kit000, kjpt, kpass
      CHARACTER(len=3)                   , INTENT(in   ) ::   cdtype
      REAL(wp), DIMENSION(:,:,:)  , TARGET, CONTIGUOUS, INTENT(in   ) ::   pahu, pahv
      REAL(wp), DIMENSION(:,:,:)  , TARGET, CONTIGUOUS, INTENT(in   ) ::   pgu, pgv, pgui, pgvi
      REAL(wp), DIMENSION(:,:,:,:), TARGET, CONTIGUOUS, INTENT(in   ) ::   ptb, ptbb
      REAL(wp), DIMENSION(:,:,:,:), TARGET, CONTIGUOUS, INTENT(inout) ::   pta
      CALL check( c_tra_ldf_iso( INT(kt,C_INT), INT(kit000,C_INT), cstr(cdtype), C_LOC(pahu), C_LOC(pahv),     &
         &        C_LOC(pgu), C_LOC(pgv), C_LOC(pgui), C_LOC(pgvi), C_LOC(ptb), C_LOC(ptbb), C_LOC(pta),          &
         &        INT(kjpt,C_INT), INT(kpass,C_INT), p_uslp, p_vslp, p_wslpi, p_wslpj, p_ah_wslp2, p_akz ),      &
         &        'tra_ldf_iso' )
   END SUBROUTINE tra_ldf_iso

   SUBROUTINE tra_zdf_imp( kt, kit000, cdtype, p2dt, ptb, pta, kjpt )
      INTEGER                            , INTENT(in   ) ::   kt, kit000, kjpt
      CHARACTER(len=3)                   , INTENT(in   ) ::   cdtype
      REAL(wp)                           , INTENT(in   ) ::   p2dt
      REAL(wp), DIMENSION(:,:,:,:), TARGET, CONTIGUOUS, INTENT(in   ) ::   ptb
      REAL(wp), DIMENSION(:,:,:,:), TARGET, CONTIGUOUS, INTENT(inout) ::   pta
      CALL check( c_tra_zdf_imp( INT(kt,C_INT), INT(kit000,C_INT), cstr(cdtype), REAL(p2dt,C_DOUBLE),         &
         &        C_LOC(ptb), C_LOC(pta), INT(kjpt,C_INT), p_avt, p_avs, INT(l_ldfslp,C_INT),                      &
         &        p_ah_wslp2, p_akz ), 'tra_zdf_imp' )
   END SUBROUTINE tra_zdf_imp

   SUBROUTINE dyn_keg( kt, kscheme )
      INTEGER, INTENT(in) ::   kt, kscheme
      CALL check( c_dyn_keg( INT(kt,C_INT), INT(kscheme,C_INT), p_un, p_vn, p_ua, p_va ), 'dyn_keg' )
   END SUBROUTINE dyn_keg

   SUBROUTINE dyn_zad( kt )
      INTEGER, INTENT(in) ::   kt
      CALL check( c_dyn_zad( INT(kt,C_INT), p_un, p_vn, p_wn, p_ua, p_va ), 'dyn_zad' )
   END SUBROUTINE dyn_zad

   SUBROUTINE dyn_adv_cen2( kt )
      INTEGER, INTENT(in) ::   kt
      CALL check( c_dyn_adv_cen2( INT(kt,C_INT), p_un, p_vn, p_wn, p_ua, p_va ), 'dyn_adv_cen2' )
   END SUBROUTINE dyn_adv_cen2

   SUBROUTINE dyn_vor( kt )
      INTEGER, INTENT(in) ::   kt
      CALL check( c_dyn_vor( INT(kt,C_INT), p_un, p_vn, p_ua, p_va ), 'dyn_vor' )
   END SUBROUTINE dyn_vor

   SUBROUTINE dyn_zdf( kt )
      INTEGER, INTENT(in) ::   kt
      CALL check( c_dyn_zdf( INT(kt,C_INT), p_ub, p_vb, p_ua, p_va, p_avm, p_rCdU_bot, p_ua_b, p_va_b,        &
         &        p_utau, p_utau_b, p_vtau, p_vtau_b ), 'dyn_zdf' )
   END SUBROUTINE dyn_zdf

   SUBROUTINE dyn_spg_ts( kt )
      INTEGER, INTENT(in) ::   kt
      CALL check( c_dyn_spg_ts( INT(kt,C_INT), spg_state ), 'dyn_spg_ts' )
   END SUBROUTINE dyn_spg_ts

   SUBROUTINE div_hor( kt )
      INTEGER, INTENT(in) ::   kt
      CALL check( c_div_hor( INT(kt,C_INT), p_un, p_vn, p_hdivn ), 'div_hor' )
   END SUBROUTINE div_hor

   SUBROUTINE ssh_nxt( kt )
      INTEGER, INTENT(in) ::   kt
      CALL check( c_ssh_nxt( INT(kt,C_INT), p_un, p_vn, p_hdivn, p_sshb, p_emp_b, p_emp, p_ssha ), 'ssh_nxt' )
   END SUBROUTINE ssh_nxt

   SUBROUTINE wzv( kt )
      INTEGER, INTENT(in) ::   kt
      CALL check( c_wzv( INT(kt,C_INT), p_hdivn, p_wn ), 'wzv' )
   END SUBROUTINE wzv

   SUBROUTINE ssh_swp( kt )
      INTEGER, INTENT(in) ::   kt
      CALL check( c_ssh_swp( INT(kt,C_INT), p_sshb, p_sshn, p_ssha ), 'ssh_swp' )
   END SUBROUTINE ssh_swp

   SUBROUTINE tra_nxt( kt )
      INTEGER, INTENT(in) ::   kt
      CALL check( c_tra_nxt( INT(kt,C_INT), p_tsb, p_tsn, p_tsa ), 'tra_nxt' )
   END SUBROUTINE tra_nxt

   SUBROUTINE dyn_nxt( kt )
      INTEGER, INTENT(in) ::   kt
      CALL check( c_dyn_nxt( INT(kt,C_INT), p_ub, p_vb, p_un, p_vn, p_ua, p_va, p_ua_b, p_va_b,               &
         &        p_un_adv, p_vn_adv, p_ub_b, p_vb_b, p_un_b, p_vn_b, p_r1_hu_n, p_r1_hv_n ), 'dyn_nxt' )
   END SUBROUTINE dyn_nxt

   SUBROUTINE eos_insitu( pts, prd, pdep )
      REAL(wp), DIMENSION(:,:,:,:), TARGET, CONTIGUOUS, INTENT(in   ) ::   pts
      REAL(wp), DIMENSION(:,:,:)  , TARGET, CONTIGUOUS, INTENT(  out) ::   prd
      REAL(wp), DIMENSION(:,:,:)  , TARGET, CONTIGUOUS, INTENT(in   ) ::   pdep
      CALL check( c_eos_insitu( C_LOC(pts), C_LOC(prd), C_LOC(pdep) ), 'eos_insitu' )
   END SUBROUTINE eos_insitu

   SUBROUTINE eos_insitu_pot( pts, prd, prhop, pdep )
      REAL(wp), DIMENSION(:,:,:,:), TARGET, CONTIGUOUS, INTENT(in   ) ::   pts
      REAL(wp), DIMENSION(:,:,:)  , TARGET, CONTIGUOUS, INTENT(  out) ::   prd, prhop
      REAL(wp), DIMENSION(:,:,:)  , TARGET, CONTIGUOUS, INTENT(in   ) ::   pdep
      CALL check( c_eos_insitu_pot( C_LOC(pts), C_LOC(prd), C_LOC(prhop), C_LOC(pdep) ), 'eos_insitu_pot' )
   END SUBROUTINE eos_insitu_pot

   SUBROUTINE rab_3d( pts, pab )
      REAL(wp), DIMENSION(:,:,:,:), TARGET, CONTIGUOUS, INTENT(in   ) ::   pts
      REAL(wp), DIMENSION(:,:,:,:), TARGET, CONTIGUOUS, INTENT(  out) ::   pab
      CALL check( c_rab_3d( C_LOC(pts), C_LOC(pab), p_gdept ), 'rab_3d' )
   END SUBROUTINE rab_3d

   SUBROUTINE bn2( pts, pab, pn2 )
      REAL(wp), DIMENSION(:,:,:,:), TARGET, CONTIGUOUS, INTENT(in   ) ::   pts, pab
      REAL(wp), DIMENSION(:,:,:)  , TARGET, CONTIGUOUS, INTENT(  out) ::   pn2
      CALL check( c_bn2( C_LOC(pts), C_LOC(pab), C_LOC(pn2), p_gdept, p_gdepw ), 'bn2' )
   END SUBROUTINE bn2

   FUNCTION rst_state( prhop ) RESULT( r )
      !! the restart variables from the bound oce and dynspg_ts arrays
      TYPE(C_PTR), INTENT(in) :: prhop
      TYPE(nemo_hip_rst_t) :: r
      r%real_bytes = 8 ; r%kqsr_b = 0
      r%qsr_hc = C_NULL_PTR ; r%qsr_hc_b = C_NULL_PTR ; r%fraqsr_1lev = C_NULL_PTR
      r%ub = p_ub ; r%vb = p_vb ; r%tsb = p_tsb ; r%un = p_un ; r%vn = p_vn ; r%tsn = p_tsn ; r%rhop = prhop
      r%sshb = p_sshb ; r%sshn = p_sshn
      r%ub2_b = spg_state%ub2_b ; r%vb2_b = spg_state%vb2_b ; r%un_bf = spg_state%un_bf ; r%vn_bf = spg_state%vn_bf
      r%un_b = spg_state%un_b ; r%vn_b = spg_state%vn_b ; r%ub_b = spg_state%ub_b ; r%vb_b = spg_state%vb_b
      r%r1_hu_n = spg_state%r1_hu_n ; r%r1_hv_n = spg_state%r1_hv_n
   END FUNCTION rst_state

   SUBROUTINE rst_write( kt, cdir, prhop, pqsr_hc, pfraqsr_1lev )
      !! IOM/restart.F90:133 with ts_rst WRITE: the bound before and now fields,
      !! sea levels, barotropic time-filter fields and prhop into directory cdir;
      !! with tra_qsr, its qsr_hc (as qsr_hc_b) and fraqsr_1lev (traqsr.F90:291-292)
      INTEGER, INTENT(in) :: kt
      CHARACTER(len=*), INTENT(in) :: cdir
      REAL(wp), DIMENSION(:,:,:), TARGET, CONTIGUOUS, INTENT(in) :: prhop
      REAL(wp), DIMENSION(:,:,:), TARGET, CONTIGUOUS, INTENT(in), OPTIONAL :: pqsr_hc
      REAL(wp), DIMENSION(:,:)  , TARGET, CONTIGUOUS, INTENT(in), OPTIONAL :: pfraqsr_1lev
      TYPE(nemo_hip_rst_t) :: r
      r = rst_state( C_LOC(prhop) )
      IF( PRESENT(pqsr_hc) )        r%qsr_hc = C_LOC(pqsr_hc)
      IF( PRESENT(pfraqsr_1lev) )   r%fraqsr_1lev = C_LOC(pfraqsr_1lev)
      CALL check( c_rst_write( cstr(TRIM(cdir)), INT(kt, C_INT), r ), 'rst_write' )
   END SUBROUTINE rst_write

   SUBROUTINE rst_read( cdir, prhop, kt, kneuler, pqsr_hc_b, pfraqsr_1lev, kqsr_b )
      !! IOM/restart.F90:240 with ts_rst READ and the depth means of
      !! DOM/istate.F90:159-178; kneuler = 0 for an Euler restart; with
      !! tra_qsr, qsr_hc_b and fraqsr_1lev (traqsr.F90:136-139, 424-425) and
      !! kqsr_b = 1 when the file held qsr_hc_b (tra_qsr's argument at nit000)
      CHARACTER(len=*), INTENT(in) :: cdir
      REAL(wp), DIMENSION(:,:,:), TARGET, CONTIGUOUS, INTENT(inout) :: prhop
      INTEGER, INTENT(out) :: kt, kneuler
      REAL(wp), DIMENSION(:,:,:), TARGET, CONTIGUOUS, INTENT(inout), OPTIONAL :: pqsr_hc_b
      REAL(wp), DIMENSION(:,:)  , TARGET, CONTIGUOUS, INTENT(inout), OPTIONAL :: pfraqsr_1lev
      INTEGER, INTENT(out), OPTIONAL :: kqsr_b
      TYPE(nemo_hip_rst_t) :: r
      INTEGER(C_INT) :: ikt, ine
      r = rst_state( C_LOC(prhop) )
      IF( PRESENT(pqsr_hc_b) )      r%qsr_hc_b = C_LOC(pqsr_hc_b)
      IF( PRESENT(pfraqsr_1lev) )   r%fraqsr_1lev = C_LOC(pfraqsr_1lev)
      CALL check( c_rst_read( cstr(TRIM(cdir)), r, ikt, ine ), 'rst_read' )
      kt = ikt ; kneuler = ine
      IF( PRESENT(kqsr_b) )   kqsr_b = r%kqsr_b
   END SUBROUTINE rst_read

   SUBROUTINE zdf_mxl( kt )
      !! ZDF/zdfmxl.F90:64: nmln, hmlp, hmlpt from rn2b (hmld is not computed)
      INTEGER, INTENT(in) :: kt
      CALL check( c_zdf_mxl( kt, p_rn2b, p_gdept, p_gdepw, m_nlb10, m_rho_c, p_nmln, p_hmlp, p_hmlpt ), 'zdf_mxl' )
   END SUBROUTINE zdf_mxl

   SUBROUTINE ldf_slp( kt, prd, pn2 )
      !! LDF/ldfslp.F90:83: uslp, vslp, wslpi, wslpj (the bound ldfslp arrays)
      INTEGER, INTENT(in) :: kt
      REAL(wp), DIMENSION(:,:,:), TARGET, CONTIGUOUS, INTENT(in) :: prd, pn2
      CALL check( c_ldf_slp( kt, C_LOC(prd), C_LOC(pn2), p_gdept, p_gdepw, p_nmln, p_hmlp, p_hmlpt, m_slpmax,  &
         &                   p_uslp, p_vslp, p_wslpi, p_wslpj ), 'ldf_slp' )
   END SUBROUTINE ldf_slp

   SUBROUTINE hpg_zco( kt )
      INTEGER, INTENT(in) ::   kt
      CALL check( c_hpg_zco( INT(kt,C_INT), p_rhd, p_ua, p_va ), 'hpg_zco' )
   END SUBROUTINE hpg_zco

   SUBROUTINE dyn_ldf_lap( kt, pub, pvb, pua, pva, kpass )
      INTEGER                         , INTENT(in   ) ::   kt, kpass
      REAL(wp), DIMENSION(:,:,:), TARGET, CONTIGUOUS, INTENT(in   ) ::   pub, pvb
      REAL(wp), DIMENSION(:,:,:), TARGET, CONTIGUOUS, INTENT(inout) ::   pua, pva
      CALL check( c_dyn_ldf_lap( INT(kt,C_INT), C_LOC(pub), C_LOC(pvb), C_LOC(pua), C_LOC(pva), INT(kpass,C_INT), &
         &        p_ahmt, p_ahmf ), 'dyn_ldf_lap' )
   END SUBROUTINE dyn_ldf_lap

   SUBROUTINE tra_sbc( kt )
      INTEGER, INTENT(in) ::   kt
      CALL check( c_tra_sbc( INT(kt,C_INT), p_qns, p_qsr, p_sfx, p_emp, p_tsn, p_tsa, p_sbc_tsc, p_sbc_tsc_b ),   &
         &        'tra_sbc' )
   END SUBROUTINE tra_sbc

   SUBROUTINE tra_qsr( kt, kqsr_b, knksr, pabs, psi0, psi1, pqsr_hc, pqsr_hc_b, pfraqsr_1lev )
      !! TRA/traqsr.F90:80, two bands: the traqsr module's arrays and namelist values, nksr of tra_qsr_init
      INTEGER , INTENT(in) ::   kt, kqsr_b, knksr
      REAL(wp), INTENT(in) ::   pabs, psi0, psi1
      REAL(wp), DIMENSION(:,:,:), TARGET, CONTIGUOUS, INTENT(inout) ::   pqsr_hc, pqsr_hc_b
      REAL(wp), DIMENSION(:,:)  , TARGET, CONTIGUOUS, INTENT(inout) ::   pfraqsr_1lev
      CALL check( c_tra_qsr( INT(kt,C_INT), INT(kqsr_b,C_INT), INT(knksr,C_INT), REAL(pabs,C_DOUBLE),         &
         &                   REAL(psi0,C_DOUBLE), REAL(psi1,C_DOUBLE), p_qsr, p_gdepw, C_LOC(pqsr_hc),     &
         &                   C_LOC(pqsr_hc_b), p_tsa, C_LOC(pfraqsr_1lev) ), 'tra_qsr' )
   END SUBROUTINE tra_qsr

   SUBROUTINE stp_ctl( kt, kindic )
      INTEGER, INTENT(in   ) ::   kt
      INTEGER, INTENT(inout) ::   kindic
      REAL(C_DOUBLE) :: zmax(6)
      INTEGER(C_INT32_T) :: iloc(11)
      INTEGER(C_INT) :: ikind
      ikind = INT(kindic, C_INT)
      CALL check( c_stp_ctl( INT(kt,C_INT), p_sshn, p_un, p_tsn, zmax, iloc, ikind ), 'stp_ctl' )
      kindic = ikind
   END SUBROUTINE stp_ctl

END MODULE nemo_hip_mod
