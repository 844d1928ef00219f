"""ctypes binding of liboracle.so -- TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and the cpu_baseline leg of bench.py may
use this module; it is the checker, never the thing measured or shipped.
Functions take numpy float64 arrays in NEMO's Fortran layout and mirror the
C-ABI entry points of include/nemo_hip.h (same names without the prefix).
"""
import ctypes as C
import os
import subprocess
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# NEMO_ORACLE_O3=1 loads the -O3 build (the timed CPU baseline); the tests use the -O2 one
LIB = os.path.join(HERE, "liboracle_o3.so" if os.environ.get("NEMO_ORACLE_O3") == "1" else "liboracle.so")
sys.path.insert(0, os.path.join(os.path.dirname(HERE), "nemo-fmi-devel_amd"))
from nemo_hip import abi  # noqa: E402

_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE, os.path.basename(LIB)], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        _lib = C.CDLL(LIB)
    return _lib


def _p(a):
    if a is None:
        return None
    assert isinstance(a, np.ndarray) and (a.flags.f_contiguous or a.ndim == 1), "Fortran-ordered array expected"
    return C.c_void_p(a.ctypes.data)


def _chk(st, name):
    if st != 0:
        raise RuntimeError(f"oracle {name} returned status {st}")


class Oracle:
    """Oracle bound to one registered domain (the oracle keeps module state in C)."""

    def __init__(self, dom):
        self.dom = dom
        self.s = dom.struct()
        L = lib()
        _chk(L.ora_domain_register(C.byref(self.s)), "domain_register")
        L.ora_spg_ts_reset()
        self.L = L

    # a2 ------------------------------------------------------------------
    def dom_msk(self):
        d = self.dom
        f3 = lambda: np.zeros((d.jpi, d.jpj, d.jpk), order="F")
        f2 = lambda: np.zeros((d.jpi, d.jpj), order="F")
        i2 = lambda: np.zeros((d.jpi, d.jpj), dtype=np.int32, order="F")
        out = dict(tmask=f3(), umask=f3(), vmask=f3(), fmask=f3(), wmask=f3(), wumask=f3(), wvmask=f3(),
                   ssmask=f2(), ssumask=f2(), ssvmask=f2(), mikt=i2(), mbkt=i2(), miku=i2(), mikv=i2(),
                   mbku=i2(), mbkv=i2())
        _chk(self.L.ora_dom_msk(*[_p(out[k]) for k in out]), "dom_msk")
        return out

    # a19 -----------------------------------------------------------------
    def lbc_lnk(self, ptab, nat, psgn=1.0, pval=0.0):
        nk = ptab.shape[2] if ptab.ndim == 3 else 1
        _chk(self.L.ora_lbc_lnk(_p(ptab), C.c_int(nk), C.c_char(nat.encode()), C.c_double(psgn),
                                C.c_double(pval)), "lbc_lnk")

    # tracers ---------------------------------------------------------------
    def tra_adv_trp(self, un, vn, wn, zun, zvn, zwn):
        _chk(self.L.ora_tra_adv_trp(_p(un), _p(vn), _p(wn), _p(zun), _p(zvn), _p(zwn)), "tra_adv_trp")

    def tra_adv_r2dt(self, kt):
        self.L.ora_tra_adv_r2dt.restype = C.c_double
        return self.L.ora_tra_adv_r2dt(C.c_int(kt))

    def tra_adv_fct(self, kt, kit000, cdtype, p2dt, pun, pvn, pwn, ptb, ptn, pta, kjpt, kn_fct_h, kn_fct_v):
        _chk(self.L.ora_tra_adv_fct(C.c_int(kt), C.c_int(kit000), cdtype.encode(), C.c_double(p2dt), _p(pun),
                                    _p(pvn), _p(pwn), _p(ptb), _p(ptn), _p(pta), C.c_int(kjpt),
                                    C.c_int(kn_fct_h), C.c_int(kn_fct_v)), "tra_adv_fct")

    def interp_4th_cpt(self, pt_in, pt_out):
        _chk(self.L.ora_interp_4th_cpt(_p(pt_in), _p(pt_out)), "interp_4th_cpt")

    def tra_ldf_lap(self, kt, kit000, cdtype, pahu, pahv, ptb, pta, kjpt, kpass):
        _chk(self.L.ora_tra_ldf_lap(C.c_int(kt), C.c_int(kit000), cdtype.encode(), _p(pahu), _p(pahv), None,
                                    None, None, None, _p(ptb), _p(pta), C.c_int(kjpt), C.c_int(kpass)),
             "tra_ldf_lap")

    def tra_ldf_iso(self, kt, kit000, cdtype, pahu, pahv, ptb, ptbb, pta, kjpt, kpass, uslp, vslp, wslpi,
                    wslpj, ah_wslp2, akz):
        _chk(self.L.ora_tra_ldf_iso(C.c_int(kt), C.c_int(kit000), cdtype.encode(), _p(pahu), _p(pahv), None,
                                    None, None, None, _p(ptb), _p(ptbb), _p(pta), C.c_int(kjpt), C.c_int(kpass),
                                    _p(uslp), _p(vslp), _p(wslpi), _p(wslpj), _p(ah_wslp2), _p(akz)),
             "tra_ldf_iso")

    def tra_zdf_imp(self, kt, kit000, cdtype, p2dt, ptb, pta, kjpt, avt, avs, l_ldfslp=0, ah_wslp2=None,
                    akz=None):
        _chk(self.L.ora_tra_zdf_imp(C.c_int(kt), C.c_int(kit000), cdtype.encode(), C.c_double(p2dt), _p(ptb),
                                    _p(pta), C.c_int(kjpt), _p(avt), _p(avs), C.c_int(l_ldfslp), _p(ah_wslp2),
                                    _p(akz)), "tra_zdf_imp")

    def tra_zdf(self, kt, tsb, tsa, avt, avs, l_ldfslp=0, ah_wslp2=None, akz=None):
        _chk(self.L.ora_tra_zdf(C.c_int(kt), _p(tsb), _p(tsa), _p(avt), _p(avs), C.c_int(l_ldfslp),
                                _p(ah_wslp2), _p(akz)), "tra_zdf")

    # momentum ----------------------------------------------------------------
    def dyn_keg(self, kt, kscheme, un, vn, ua, va):
        _chk(self.L.ora_dyn_keg(C.c_int(kt), C.c_int(kscheme), _p(un), _p(vn), _p(ua), _p(va)), "dyn_keg")

    def dyn_zad(self, kt, un, vn, wn, ua, va):
        _chk(self.L.ora_dyn_zad(C.c_int(kt), _p(un), _p(vn), _p(wn), _p(ua), _p(va)), "dyn_zad")

    def dyn_adv_cen2(self, kt, un, vn, wn, ua, va):
        _chk(self.L.ora_dyn_adv_cen2(C.c_int(kt), _p(un), _p(vn), _p(wn), _p(ua), _p(va)), "dyn_adv_cen2")

    def dyn_adv(self, kt, un, vn, wn, ua, va):
        _chk(self.L.ora_dyn_adv(C.c_int(kt), _p(un), _p(vn), _p(wn), _p(ua), _p(va)), "dyn_adv")

    def dyn_vor(self, kt, un, vn, ua, va):
        _chk(self.L.ora_dyn_vor(C.c_int(kt), _p(un), _p(vn), _p(ua), _p(va)), "dyn_vor")

    def dyn_zdf(self, kt, ub, vb, ua, va, avm, rCdU_bot, ua_b, va_b, utau, utau_b, vtau, vtau_b):
        _chk(self.L.ora_dyn_zdf(C.c_int(kt), _p(ub), _p(vb), _p(ua), _p(va), _p(avm), _p(rCdU_bot), _p(ua_b),
                                _p(va_b), _p(utau), _p(utau_b), _p(vtau), _p(vtau_b)), "dyn_zdf")

    def ts_wgt(self, ll_av, ll_fw):
        n = 3 * self.dom.opts["nn_baro"]
        w1, w2 = np.zeros(n), np.zeros(n)
        jpit = C.c_int(0)
        _chk(self.L.ora_ts_wgt(C.c_int(ll_av), C.c_int(ll_fw), C.byref(jpit), _p(w1), _p(w2)), "ts_wgt")
        return jpit.value, w1, w2

    def dyn_spg_ts(self, kt, fields):
        """fields: dict of numpy arrays named as in nemo_hip_spg_ts_t (modified in place)."""
        s = abi.SpgTsStruct()
        for n, _ in abi.SpgTsStruct._fields_:
            a = fields.get(n)
            setattr(s, n, None if a is None else a.ctypes.data)
        _chk(self.L.ora_dyn_spg_ts(C.c_int(kt), C.byref(s)), "dyn_spg_ts")

    # equation of state, pressure gradient, lateral viscosity ---------------
    def eos_register(self, eos):
        """eos: abi.EosStruct"""
        _chk(self.L.ora_eos_register(C.byref(eos)), "eos_register")

    def eos_poly_table(self, scheme):
        """The coefficient table of polynomial scheme 0 (TEOS-10) or 1 (EOS-80):
        dict of rdeltaS, r1_S0, r1_T0, r1_Z0 and rho/alp/bet arrays [k][j][i]."""
        buf = np.zeros(4 + 3 * 196)
        _chk(self.L.ora_eos_poly_table(C.c_int(scheme), _p(buf)), "eos_poly_table")
        t = dict(zip(["rdeltaS", "r1_S0", "r1_T0", "r1_Z0"], buf[:4]))
        for n, name in enumerate(["rho", "alp", "bet"]):
            t[name] = buf[4 + 196 * n: 4 + 196 * (n + 1)].reshape(4, 7, 7).copy()
        return t

    def eos_insitu(self, pts, prd, pdep):
        _chk(self.L.ora_eos_insitu(_p(pts), _p(prd), _p(pdep)), "eos_insitu")

    def eos_insitu_pot(self, pts, prd, prhop, pdep):
        _chk(self.L.ora_eos_insitu_pot(_p(pts), _p(prd), _p(prhop), _p(pdep)), "eos_insitu_pot")

    def rab_3d(self, pts, pab, gdept):
        _chk(self.L.ora_rab_3d(_p(pts), _p(pab), _p(gdept)), "rab_3d")

    def bn2(self, pts, pab, pn2, gdept, gdepw):
        _chk(self.L.ora_bn2(_p(pts), _p(pab), _p(pn2), _p(gdept), _p(gdepw)), "bn2")

    def zdf_mxl(self, kt, rn2b, gdept, gdepw, nlb10, rho_c=0.01):
        """(nmln int32, hmlp, hmlpt) of zdf_mxl."""
        jpi, jpj = rn2b.shape[:2]
        nmln = np.zeros((jpi, jpj), dtype=np.int32, order="F")
        hmlp, hmlpt = np.zeros((jpi, jpj), order="F"), np.zeros((jpi, jpj), order="F")
        _chk(self.L.ora_zdf_mxl(C.c_int(kt), _p(rn2b), _p(gdept), _p(gdepw), C.c_int(nlb10), C.c_double(rho_c),
                                nmln.ctypes.data_as(C.POINTER(C.c_int)), _p(hmlp), _p(hmlpt)), "zdf_mxl")
        return nmln, hmlp, hmlpt

    def ldf_slp(self, kt, prd, pn2, gdept, gdepw, nmln, hmlp, hmlpt, rn_slpmax, uslp, vslp, wslpi, wslpj):
        nm = np.asfortranarray(nmln, dtype=np.int32)
        _chk(self.L.ora_ldf_slp(C.c_int(kt), _p(prd), _p(pn2), _p(gdept), _p(gdepw),
                                nm.ctypes.data_as(C.POINTER(C.c_int)), _p(hmlp), _p(hmlpt), C.c_double(rn_slpmax),
                                _p(uslp), _p(vslp), _p(wslpi), _p(wslpj)), "ldf_slp")

    # tangent linear and adjoint of the tracer step (one tracer) -------------
    def fct_coef(self, p2dt, pun, pvn, pwn, tb, tn, cx, cy, cz, kn=(2, 2)):
        _chk(self.L.ora_fct_coef(C.c_double(p2dt), _p(pun), _p(pvn), _p(pwn), _p(tb), _p(tn), C.c_int(kn[0]),
                                 C.c_int(kn[1]), _p(cx), _p(cy), _p(cz)), "fct_coef")

    def fct_tl(self, p2dt, pun, pvn, pwn, tb, tn, cx, cy, cz, ta, kn=(2, 2)):
        _chk(self.L.ora_fct_tl(C.c_double(p2dt), _p(pun), _p(pvn), _p(pwn), _p(tb), _p(tn), C.c_int(kn[0]),
                               C.c_int(kn[1]), _p(cx), _p(cy), _p(cz), _p(ta)), "fct_tl")

    def fct_ad(self, pun, pvn, pwn, cx, cy, cz, ta_ad, tb_ad, tn_ad, kn=(2, 2)):
        _chk(self.L.ora_fct_ad(_p(pun), _p(pvn), _p(pwn), C.c_int(kn[0]), C.c_int(kn[1]), _p(cx), _p(cy), _p(cz),
                               _p(ta_ad), _p(tb_ad), _p(tn_ad)), "fct_ad")

    def ldf_lap_ad(self, pahu, pahv, ta_ad, tb_ad):
        _chk(self.L.ora_ldf_lap_ad(_p(pahu), _p(pahv), _p(ta_ad), _p(tb_ad)), "ldf_lap_ad")

    def zdf_imp_ad(self, p2dt, avt, ta_ad, tb_ad):
        _chk(self.L.ora_zdf_imp_ad(C.c_double(p2dt), _p(avt), _p(ta_ad), _p(tb_ad)), "zdf_imp_ad")

    def tra_nxt_ad(self, tb_ad, tn_ad, ta_ad):
        _chk(self.L.ora_tra_nxt_ad(_p(tb_ad), _p(tn_ad), _p(ta_ad)), "tra_nxt_ad")

    def lbc_adj(self, p, nk, nat):
        _chk(self.L.ora_lbc_adj(_p(p), C.c_int(nk), C.c_char(nat.encode())), "lbc_adj")

    def hpg_zco(self, kt, rhd, ua, va):
        _chk(self.L.ora_hpg_zco(C.c_int(kt), _p(rhd), _p(ua), _p(va)), "hpg_zco")

    def dyn_ldf_lap(self, kt, pub, pvb, pua, pva, kpass, ahmt, ahmf):
        _chk(self.L.ora_dyn_ldf_lap(C.c_int(kt), _p(pub), _p(pvb), _p(pua), _p(pva), C.c_int(kpass), _p(ahmt),
                                    _p(ahmf)), "dyn_ldf_lap")

    def stp_ctl(self, kt, sshn, un, tsn):
        """(status, zmax[6], loc[11], kindic) of the oracle's stp_ctl."""
        zmax = np.zeros(6)
        loc = np.zeros(11, np.int32)
        kind = C.c_int(0)
        st = self.L.ora_stp_ctl(C.c_int(kt), _p(sshn), _p(un), _p(tsn), _p(zmax), _p(loc), C.byref(kind))
        return st, zmax, loc, kind.value

    def glob_sum(self, ptab, nk, kind=0):
        """glob_sum (kind 0, tmask_i) / glob_sum_full (kind 1, tmask_h) of ptab(:,:,1:nk)"""
        f = self.L.ora_glob_sum
        f.restype = C.c_double
        return float(f(_p(ptab), C.c_int(nk), C.c_int(kind)))

    def dia_hsb_init(self, tsn, sshn):
        """dia_hsb_init without a restart: (surf_tot, dict of the initial-state arrays)"""
        d = self.dom
        s2, s3 = (d.jpi, d.jpj), (d.jpi, d.jpj, d.jpk)
        a = {n: np.zeros(s2, order="F") for n in ("surf_ini", "ssh_ini", "ssh_hc_loc_ini", "ssh_sc_loc_ini")}
        a.update({n: np.zeros(s3, order="F") for n in ("e3t_ini", "hc_loc_ini", "sc_loc_ini")})
        f = self.L.ora_dia_hsb_init
        f.restype = C.c_double
        tot = f(_p(tsn), _p(sshn), _p(a["surf_ini"]), _p(a["ssh_ini"]), _p(a["ssh_hc_loc_ini"]),
                _p(a["ssh_sc_loc_ini"]), _p(a["e3t_ini"]), _p(a["hc_loc_ini"]), _p(a["sc_loc_ini"]))
        return float(tot), a

    def dia_hsb_sums(self, tsn, tsb, wn, sshn, emp, rnf, qsr, sbc_tsc, ini):
        """the thirteen global sums of one dia_hsb call"""
        sums = np.zeros(13)
        q = qsr if qsr is not None else emp
        _chk(self.L.ora_dia_hsb_sums(_p(tsn), _p(tsb), _p(wn), _p(sshn), _p(emp), _p(rnf), _p(q), _p(sbc_tsc),
                                     _p(ini["surf_ini"]), _p(ini["ssh_ini"]), _p(ini["ssh_hc_loc_ini"]),
                                     _p(ini["ssh_sc_loc_ini"]), _p(ini["e3t_ini"]), _p(ini["hc_loc_ini"]),
                                     _p(ini["sc_loc_ini"]), _p(sums)), "dia_hsb_sums")
        return sums

    def qsr_nksr(self, gdepw0, prldex=23.0, pqsr_frc=100.0):
        """trc_oce_ext_lev(prldex, pqsr_frc): nksr of the two-band tra_qsr"""
        return int(self.L.ora_qsr_nksr(C.c_double(prldex), C.c_double(pqsr_frc), _p(gdepw0)))

    def tra_qsr(self, kt, nksr, qsr, gdepw, qsr_hc, qsr_hc_b, tsa, fraqsr_1lev, kqsr_b=0, rn_abs=0.58, rn_si0=0.35,
                rn_si1=23.0):
        _chk(self.L.ora_tra_qsr(C.c_int(kt), C.c_int(kqsr_b), C.c_int(nksr), C.c_double(rn_abs), C.c_double(rn_si0),
                                C.c_double(rn_si1), _p(qsr), _p(gdepw), _p(qsr_hc), _p(qsr_hc_b), _p(tsa),
                                _p(fraqsr_1lev)), "tra_qsr")

    def tra_sbc(self, kt, qns, qsr, sfx, emp, tsn, tsa, sbc_tsc, sbc_tsc_b):
        _chk(self.L.ora_tra_sbc(C.c_int(kt), _p(qns), _p(qsr), _p(sfx), _p(emp), _p(tsn), _p(tsa), _p(sbc_tsc),
                                _p(sbc_tsc_b)), "tra_sbc")

    # time-step closure ----------------------------------------------------
    def div_hor(self, kt, un, vn, hdivn):
        _chk(self.L.ora_div_hor(C.c_int(kt), _p(un), _p(vn), _p(hdivn)), "div_hor")

    def ssh_nxt(self, kt, un, vn, hdivn, sshb, emp_b, emp, ssha):
        _chk(self.L.ora_ssh_nxt(C.c_int(kt), _p(un), _p(vn), _p(hdivn), _p(sshb), _p(emp_b), _p(emp), _p(ssha)),
             "ssh_nxt")

    def wzv(self, kt, hdivn, wn):
        _chk(self.L.ora_wzv(C.c_int(kt), _p(hdivn), _p(wn)), "wzv")

    def ssh_swp(self, kt, sshb, sshn, ssha):
        _chk(self.L.ora_ssh_swp(C.c_int(kt), _p(sshb), _p(sshn), _p(ssha)), "ssh_swp")

    def tra_nxt(self, kt, tsb, tsn, tsa):
        _chk(self.L.ora_tra_nxt(C.c_int(kt), _p(tsb), _p(tsn), _p(tsa)), "tra_nxt")

    def dyn_nxt(self, kt, ub, vb, un, vn, ua, va, ua_b, va_b, un_adv, vn_adv, ub_b, vb_b, un_b, vn_b, r1_hu_n,
                r1_hv_n):
        _chk(self.L.ora_dyn_nxt(C.c_int(kt), _p(ub), _p(vb), _p(un), _p(vn), _p(ua), _p(va), _p(ua_b), _p(va_b),
                                _p(un_adv), _p(vn_adv), _p(ub_b), _p(vb_b), _p(un_b), _p(vn_b), _p(r1_hu_n),
                                _p(r1_hv_n)), "dyn_nxt")
