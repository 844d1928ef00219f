PROGRAM drive_routines
   !!======================================================================
   !! Calls hot-path routines through nemo_hip_mod the way a NEMO build
   !! does: NEMO's own host arrays, the default host-pointer mode, the
   !! reference argument lists.  Domain and state come from an unformatted
   !! stream file written by tests/test_gpu_fortran.py (sizes, switches,
   !! k_top/k_bot, the 24 horizontal metrics and 7 vertical scale factors
   !! in nemo_hip_domain_t order, then the fields); the results go to a
   !! second stream file:  tra_ldf_lap, tra_zdf_imp, tra_adv_fct outputs, then
   !! the module-global step routines as step.F90 calls them: tsa from zero
   !! through tra_adv, tra_ldf, tra_zdf; ua, va from zero through dyn_adv,
   !! dyn_vor, dyn_zdf; hdivn of div_hor.
   !!
   !! usage: drive_routines <input> <output>
   !!======================================================================
   USE, INTRINSIC :: ISO_C_BINDING
   USE nemo_hip_mod
   IMPLICIT NONE
   INTEGER, PARAMETER :: wp = C_DOUBLE
   TYPE(nemo_hip_domain_t) :: sdom
   INTEGER(C_INT32_T) :: ih(25)
   REAL(wp) :: rh(6), p2dt
   INTEGER :: jpi, jpj, jpk, iu, kt
   INTEGER(C_INT32_T), ALLOCATABLE, TARGET :: ktop(:,:), kbot(:,:)
   REAL(wp), ALLOCATABLE, TARGET :: m2(:,:,:), e3(:,:,:,:)
   REAL(wp), ALLOCATABLE, TARGET :: ahtu(:,:,:), ahtv(:,:,:), avt(:,:,:), avs(:,:,:), avm(:,:,:)
   REAL(wp), ALLOCATABLE, TARGET :: zun(:,:,:), zvn(:,:,:), zwn(:,:,:), zgu(:,:,:)
   REAL(wp), ALLOCATABLE, TARGET :: tsb(:,:,:,:), tsn(:,:,:,:), ta_lap(:,:,:,:), ta_zdf(:,:,:,:), ta_fct(:,:,:,:)
   REAL(wp), ALLOCATABLE, TARGET :: un(:,:,:), vn(:,:,:), wn(:,:,:), ub(:,:,:), vb(:,:,:), ua(:,:,:), va(:,:,:)
   REAL(wp), ALLOCATABLE, TARGET :: hdivn(:,:,:), tsa(:,:,:,:), f2(:,:,:)
   CHARACTER(len=1024) :: cin, cout

   CALL GET_COMMAND_ARGUMENT( 1, cin )
   CALL GET_COMMAND_ARGUMENT( 2, cout )
   OPEN( NEWUNIT=iu, FILE=TRIM(cin), ACCESS='STREAM', FORM='UNFORMATTED', STATUS='OLD', ACTION='READ' )
   READ(iu) ih, rh
   jpi = ih(1) ; jpj = ih(2) ; jpk = ih(3)
   ALLOCATE( ktop(jpi,jpj), kbot(jpi,jpj), m2(jpi,jpj,24), e3(jpi,jpj,jpk,7) )
   ALLOCATE( ahtu(jpi,jpj,jpk), ahtv(jpi,jpj,jpk), avt(jpi,jpj,jpk), avs(jpi,jpj,jpk), avm(jpi,jpj,jpk) )
   ALLOCATE( zun(jpi,jpj,jpk), zvn(jpi,jpj,jpk), zwn(jpi,jpj,jpk), zgu(jpi,jpj,2) )
   ALLOCATE( tsb(jpi,jpj,jpk,2), tsn(jpi,jpj,jpk,2), ta_lap(jpi,jpj,jpk,2), ta_zdf(jpi,jpj,jpk,2) )
   ALLOCATE( ta_fct(jpi,jpj,jpk,2) )
   READ(iu) ktop, kbot, m2, e3, ahtu, ahtv, avt, avs, avm, zun, zvn, zwn, tsb, tsn, ta_lap, kt, p2dt
   ALLOCATE( un(jpi,jpj,jpk), vn(jpi,jpj,jpk), wn(jpi,jpj,jpk), ub(jpi,jpj,jpk), vb(jpi,jpj,jpk) )
   ALLOCATE( ua(jpi,jpj,jpk), va(jpi,jpj,jpk), hdivn(jpi,jpj,jpk), tsa(jpi,jpj,jpk,2), f2(jpi,jpj,18) )
   ! f2: rCdU_bot utau utau_b vtau vtau_b, then zeros for ua_b va_b and the closure's 2D arrays
   f2(:,:,:) = 0._wp
   READ(iu) un, vn, wn, ub, vb, f2(:,:,1:5)
   CLOSE(iu)
   ta_zdf(:,:,:,:) = ta_lap(:,:,:,:)
   ta_fct(:,:,:,:) = ta_lap(:,:,:,:)
   zgu(:,:,:) = 0._wp

   ! the dom_oce sizes, switches and arrays, as dom_init would hand them over
   sdom%jpi = ih(1) ; sdom%jpj = ih(2) ; sdom%jpk = ih(3) ; sdom%jperio = ih(4)
   sdom%ln_linssh = ih(5) ; sdom%ln_isfcav = ih(6) ; sdom%ln_zps = ih(7) ; sdom%ln_sco = ih(8)
   sdom%ln_traldf_lap = ih(9) ; sdom%ln_traldf_msc = ih(10) ; sdom%ln_zdfddm = ih(11) ; sdom%ln_zad_Aimp = ih(12)
   sdom%n_dynadv = ih(13) ; sdom%nn_dynkeg = ih(14) ; sdom%nvor_scheme = ih(15) ; sdom%nn_een_e3f = ih(16)
   sdom%ln_drgimp = ih(17)
   sdom%ln_dynspg_ts = ih(18) ; sdom%ln_bt_fw = ih(19) ; sdom%ln_bt_av = ih(20) ; sdom%nn_bt_flt = ih(21)
   sdom%nn_baro = ih(22) ; sdom%neuler = ih(23) ; sdom%nit000 = ih(24) ; sdom%ln_traqsr = ih(25)
   sdom%rdt = rh(1) ; sdom%rn_atfp = rh(2) ; sdom%rau0 = rh(3) ; sdom%grav = rh(4)
   sdom%rn_bt_alpha = rh(5) ; sdom%rn_scal_load = rh(6)
   sdom%k_top = C_LOC( ktop ) ; sdom%k_bot = C_LOC( kbot )
   sdom%e1t      = C_LOC( m2(1,1, 1) ) ; sdom%e2t      = C_LOC( m2(1,1, 2) )
   sdom%e1u      = C_LOC( m2(1,1, 3) ) ; sdom%e2u      = C_LOC( m2(1,1, 4) )
   sdom%e1v      = C_LOC( m2(1,1, 5) ) ; sdom%e2v      = C_LOC( m2(1,1, 6) )
   sdom%e1f      = C_LOC( m2(1,1, 7) ) ; sdom%e2f      = C_LOC( m2(1,1, 8) )
   sdom%e1e2t    = C_LOC( m2(1,1, 9) ) ; sdom%r1_e1e2t = C_LOC( m2(1,1,10) )
   sdom%e1e2u    = C_LOC( m2(1,1,11) ) ; sdom%r1_e1e2u = C_LOC( m2(1,1,12) )
   sdom%e1e2v    = C_LOC( m2(1,1,13) ) ; sdom%r1_e1e2v = C_LOC( m2(1,1,14) )
   sdom%e1e2f    = C_LOC( m2(1,1,15) ) ; sdom%r1_e1e2f = C_LOC( m2(1,1,16) )
   sdom%e2_e1u   = C_LOC( m2(1,1,17) ) ; sdom%e1_e2v   = C_LOC( m2(1,1,18) )
   sdom%r1_e1u   = C_LOC( m2(1,1,19) ) ; sdom%r1_e2u   = C_LOC( m2(1,1,20) )
   sdom%r1_e1v   = C_LOC( m2(1,1,21) ) ; sdom%r1_e2v   = C_LOC( m2(1,1,22) )
   sdom%ff_f     = C_LOC( m2(1,1,23) ) ; sdom%ff_t     = C_LOC( m2(1,1,24) )
   sdom%e3t_n  = C_LOC( e3(1,1,1,1) ) ; sdom%e3u_n  = C_LOC( e3(1,1,1,2) ) ; sdom%e3v_n  = C_LOC( e3(1,1,1,3) )
   sdom%e3w_n  = C_LOC( e3(1,1,1,4) ) ; sdom%e3f_n  = C_LOC( e3(1,1,1,5) )
   sdom%e3uw_n = C_LOC( e3(1,1,1,6) ) ; sdom%e3vw_n = C_LOC( e3(1,1,1,7) )

   CALL nemo_hip_start( 0 )
   CALL nemo_hip_domain( sdom )
   CALL nemo_hip_bind_zdf( avt, avs, avm )
   ! the three calls of tra_ldf (traldf.F90:70), tra_zdf (trazdf.F90:77), tra_adv (traadv.F90:150)
   CALL tra_ldf_lap( kt, 1, 'TRA', ahtu, ahtv, zgu, zgu, zgu, zgu, tsb, ta_lap, 2, 1 )
   CALL tra_zdf_imp( kt, 1, 'TRA', p2dt, tsb, ta_zdf, 2 )
   CALL tra_adv_fct( kt, 1, 'TRA', p2dt, zun, zvn, zwn, tsb, tsn, ta_fct, 2, 4, 4 )
   ! the module arrays of oce, zdfdrg, sbc_oce, dynspg_ts and ldftra, bound once
   CALL nemo_hip_bind_dyn( un, vn, wn, ub, vb, ua, va, f2(:,:,1), f2(:,:,6), f2(:,:,7), f2(:,:,2), f2(:,:,3),   &
      &                    f2(:,:,4), f2(:,:,5) )
   CALL nemo_hip_bind_nxt( hdivn, f2(:,:,8), f2(:,:,9), f2(:,:,10), f2(:,:,11), f2(:,:,12), tsb, tsn, tsa,       &
      &                    f2(:,:,13), f2(:,:,14), f2(:,:,15), f2(:,:,16), f2(:,:,17), f2(:,:,18), f2(:,:,8),      &
      &                    f2(:,:,8) )
   CALL nemo_hip_bind_tra( 4, 4, np_lap, ahtu, ahtv )
   ! step.F90:234-257: the tracer trend from zero, and :179-204 the momentum trend
   tsa(:,:,:,:) = 0._wp
   CALL tra_adv( kt )
   CALL tra_ldf( kt )
   CALL tra_zdf( kt )
   ua(:,:,:) = 0._wp ; va(:,:,:) = 0._wp
   CALL dyn_adv( kt )
   CALL dyn_vor( kt )
   CALL dyn_zdf( kt )
   CALL div_hor( kt )
   CALL nemo_hip_stop

   OPEN( NEWUNIT=iu, FILE=TRIM(cout), ACCESS='STREAM', FORM='UNFORMATTED', STATUS='REPLACE', ACTION='WRITE' )
   WRITE(iu) ta_lap, ta_zdf, ta_fct, tsa, ua, va, hdivn
   CLOSE(iu)
   WRITE(*,'(a)') 'drive_routines: done'
END PROGRAM drive_routines
