"""Host-side mirror of NEMO's hot-path routines over the gfx950 C-ABI.

Every method keeps the reference routine's name and dummy-argument order
(e.g. ``tra_adv_fct(kt, kit000, cdtype, p2dt, pun, pvn, pwn, ptb, ptn, pta,
kjpt, kn_fct_h, kn_fct_v)``, TRA/traadv_fct.F90:54) followed by the module
arrays the routine reads, and raises ``NemoHipError`` carrying the library's
message where the reference would call ctl_stop.

Arrays are either numpy arrays in Fortran order (host-pointer mode: the
library stages them through HBM) or 1-D torch tensors on the GPU holding the
same Fortran-ordered data (device mode: zero copy, asynchronous on the
library's stream).  There is no CPU fallback: without libnemo_hip.so every
call raises.
"""
import ctypes as C
import os

import numpy as np

from . import abi
from .abi import (NEMO_HIP_PTR_DEVICE, NEMO_HIP_PTR_HOST, NP_EEN, NP_EET, NP_ENE, NP_ENS, NP_ENT,  # noqa: F401
                  NP_MIX, NP_FLX_C2, NP_LIN_DYN, NP_VEC_C2)

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(os.path.dirname(HERE), "libnemo_hip.so")
HEADER = os.path.join(os.path.dirname(os.path.dirname(HERE)), "include", "nemo_hip.h")


class NemoHipError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"nemo_hip error {code}: {msg}")
        self.code = code


def load_library(path=LIB_PATH):
    if not os.path.exists(path):
        raise NemoHipError(abi.NEMO_HIP_ERR_STATE,
                           f"{path} is missing: build it with __graft_entry__.build() (no CPU fallback exists)")
    # One HIP runtime per process: the library binds to whichever
    # libamdhip64 is loaded first.  When torch is installed its bundled
    # runtime is loaded before the library, so that device tensors and the
    # library share it (loading the system runtime first leaves torch
    # without a device).
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    return C.CDLL(path)


def _is_torch(a):
    return type(a).__module__.startswith("torch")


class DeviceField:
    """A resident field of nemo_hip_field_alloc: a device address and its size
    in bytes; device-mode entry points accept it where they accept a tensor."""

    def __init__(self, ptr, nbytes):
        self.ptr, self.nbytes = ptr, nbytes


class NemoHip:
    """One process-wide handle on the library (the library itself is a singleton)."""

    def __init__(self, device=0, mode=NEMO_HIP_PTR_DEVICE, path=LIB_PATH):
        self.L = load_library(path)
        self.L.nemo_hip_stream.restype = C.c_void_p
        self._check(self.L.nemo_hip_init(C.c_int(device)))
        self.set_pointer_mode(mode)
        self.dom = None

    # ------------------------------------------------------------------
    def last_error(self):
        buf = C.create_string_buffer(1024)
        self.L.nemo_hip_last_error(buf, C.c_int(1024))
        return buf.value.decode(errors="replace")

    def _check(self, rc):
        if rc != 0:
            raise NemoHipError(rc, self.last_error())

    def set_pointer_mode(self, mode):
        self._check(self.L.nemo_hip_set_pointer_mode(C.c_int(mode)))
        self.mode = mode

    def sync(self):
        self._check(self.L.nemo_hip_sync())

    def stream(self):
        return self.L.nemo_hip_stream()

    def finalize(self):
        self._check(self.L.nemo_hip_finalize())

    def register(self, dom):
        """dom: configs.Domain (host arrays); the library copies what it needs."""
        self._dstruct = dom.struct()
        self._check(self.L.nemo_hip_domain_register(C.byref(self._dstruct)))
        self.dom = dom

    # pointers ----------------------------------------------------------
    def _p(self, a, dtype=np.float64):
        if a is None:
            return None
        if isinstance(a, DeviceField):
            if self.mode != NEMO_HIP_PTR_DEVICE:
                raise NemoHipError(abi.NEMO_HIP_ERR_ARG, "resident field passed in host-pointer mode")
            return C.c_void_p(a.ptr)
        if _is_torch(a):
            if self.mode != NEMO_HIP_PTR_DEVICE:
                raise NemoHipError(abi.NEMO_HIP_ERR_ARG, "torch tensor passed in host-pointer mode")
            if not a.is_cuda or not a.is_contiguous():
                raise NemoHipError(abi.NEMO_HIP_ERR_ARG, "device tensors must be contiguous GPU tensors")
            return C.c_void_p(a.data_ptr())
        if self.mode != NEMO_HIP_PTR_HOST:
            raise NemoHipError(abi.NEMO_HIP_ERR_ARG, "numpy array passed in device-pointer mode")
        if a.dtype != dtype:
            raise NemoHipError(abi.NEMO_HIP_ERR_ARG, f"expected {dtype}, got {a.dtype}")
        if not (a.flags.f_contiguous or a.ndim == 1):
            raise NemoHipError(abi.NEMO_HIP_ERR_ARG, "host arrays must be Fortran-contiguous")
        return C.c_void_p(a.ctypes.data)

    # contexts and graphs -------------------------------------------------
    def context_create(self):
        c = C.c_int()
        self._check(self.L.nemo_hip_context_create(C.byref(c)))
        return c.value

    def context_select(self, ctx):
        self._check(self.L.nemo_hip_context_select(C.c_int(ctx)))

    def context_destroy(self, ctx):
        self._check(self.L.nemo_hip_context_destroy(C.c_int(ctx)))

    def graph_begin(self):
        self._check(self.L.nemo_hip_graph_begin())

    def graph_end(self):
        g = C.c_int()
        self._check(self.L.nemo_hip_graph_end(C.byref(g)))
        return g.value

    def graph_launch(self, g):
        self._check(self.L.nemo_hip_graph_launch(C.c_int(g)))

    def graph_free(self, g):
        self._check(self.L.nemo_hip_graph_free(C.c_int(g)))

    # plans: one traced step of every member, replayed as batched launches
    def plan_begin(self):
        self._check(self.L.nemo_hip_plan_begin())

    def plan_end(self):
        p = C.c_int()
        self._check(self.L.nemo_hip_plan_end(C.byref(p)))
        return p.value

    def plan_launch(self, p):
        self._check(self.L.nemo_hip_plan_launch(C.c_int(p)))

    def plan_info(self, p):
        v = [C.c_int() for _ in range(5)]
        self._check(self.L.nemo_hip_plan_info(C.c_int(p), *[C.byref(x) for x in v]))
        return dict(zip(("members", "groups", "ops", "batched", "launches"), (x.value for x in v)))

    def plan_free(self, p):
        self._check(self.L.nemo_hip_plan_free(C.c_int(p)))

    # resident fields -----------------------------------------------------
    def field_alloc(self, nbytes):
        p = C.c_void_p()
        self._check(self.L.nemo_hip_field_alloc(C.c_size_t(nbytes), C.byref(p)))
        return DeviceField(p.value, nbytes)

    def field_free(self, f):
        self._check(self.L.nemo_hip_field_free(C.c_void_p(f.ptr if isinstance(f, DeviceField) else f)))

    def field_push(self, f, host, nbytes=None):
        """Copy a Fortran-ordered host array into the field (whole array unless nbytes)."""
        host = np.asfortranarray(host)
        n = host.nbytes if nbytes is None else nbytes
        self._check(self.L.nemo_hip_field_push(C.c_void_p(f.ptr), C.c_void_p(host.ctypes.data), C.c_size_t(n)))

    def field_pull(self, host, f, nbytes=None):
        """Copy the field into a Fortran-contiguous host array, after the queued work."""
        if not (host.flags.f_contiguous or host.ndim == 1):
            raise NemoHipError(abi.NEMO_HIP_ERR_ARG, "host arrays must be Fortran-contiguous")
        n = host.nbytes if nbytes is None else nbytes
        self._check(self.L.nemo_hip_field_pull(C.c_void_p(host.ctypes.data), C.c_void_p(f.ptr), C.c_size_t(n)))

    # timing ------------------------------------------------------------
    def timing_enable(self, on=True):
        self._check(self.L.nemo_hip_timing_enable(C.c_int(1 if on else 0)))

    def timing_only(self, key=None):
        """Time only the launches of one kernel, keyed "routine/kernel"; None times every launch."""
        self._check(self.L.nemo_hip_timing_only(key.encode() if key else None))

    def timing_reset(self):
        self._check(self.L.nemo_hip_timing_reset())

    def timing_get(self, name):
        ms, n = C.c_double(0), C.c_int64(0)
        self._check(self.L.nemo_hip_timing_get(name.encode(), C.byref(ms), C.byref(n)))
        return ms.value, n.value

    # masks and lateral boundaries ----------------------------------------------------------
    def dom_msk(self, tmask=None, umask=None, vmask=None, fmask=None, wmask=None, wumask=None, wvmask=None,
                ssmask=None, ssumask=None, ssvmask=None, mikt=None, mbkt=None, miku=None, mikv=None, mbku=None,
                mbkv=None):
        f = self._p
        i = lambda a: self._p(a, np.int32)
        self._check(self.L.nemo_hip_dom_msk(f(tmask), f(umask), f(vmask), f(fmask), f(wmask), f(wumask), f(wvmask),
                                            f(ssmask), f(ssumask), f(ssvmask), i(mikt), i(mbkt), i(miku), i(mikv),
                                            i(mbku), i(mbkv)))

    def lbc_lnk(self, ptab, nk, cd_nat, psgn=1.0, pval=0.0):
        self._check(self.L.nemo_hip_lbc_lnk(self._p(ptab), C.c_int(nk), C.c_char(cd_nat.encode()),
                                            C.c_double(psgn), C.c_double(pval)))

    # tracers -------------------------------------------------------------
    def tra_adv_trp(self, un, vn, wn, zun, zvn, zwn, sp=False):
        fn = self.L.nemo_hip_tra_adv_trp_sp if sp else self.L.nemo_hip_tra_adv_trp
        dt = np.float32 if sp else np.float64
        p = lambda a: self._p(a, dt)
        self._check(fn(p(un), p(vn), p(wn), p(zun), p(zvn), p(zwn)))

    def tra_adv(self, kt, un, vn, wn, ptb, ptn, pta, kjpt, kn_fct_h, kn_fct_v, sp=False):
        """tra_adv (FCT): r2dt of kt, transports of un, vn, wn formed where used"""
        fn = self.L.nemo_hip_tra_adv_sp if sp else self.L.nemo_hip_tra_adv
        dt = np.float32 if sp else np.float64
        p = lambda a: self._p(a, dt)
        self._check(fn(C.c_int(kt), p(un), p(vn), p(wn), p(ptb), p(ptn), p(pta), C.c_int(kjpt), C.c_int(kn_fct_h),
                       C.c_int(kn_fct_v)))

    def tra_adv_fct(self, kt, kit000, cdtype, p2dt, pun, pvn, pwn, ptb, ptn, pta, kjpt, kn_fct_h, kn_fct_v, sp=False):
        fn = self.L.nemo_hip_tra_adv_fct_sp if sp else self.L.nemo_hip_tra_adv_fct
        dt = np.float32 if sp else np.float64
        p = lambda a: self._p(a, dt)
        self._check(fn(C.c_int(kt), C.c_int(kit000), cdtype.encode(), C.c_double(p2dt), p(pun), p(pvn), p(pwn),
                       p(ptb), p(ptn), p(pta), C.c_int(kjpt), C.c_int(kn_fct_h), C.c_int(kn_fct_v)))

    def interp_4th_cpt(self, pt_in, pt_out):
        self._check(self.L.nemo_hip_interp_4th_cpt(self._p(pt_in), self._p(pt_out)))

    def tra_ldf_lap(self, kt, kit000, cdtype, pahu, pahv, ptb, pta, kjpt, kpass, sp=False):
        dt = np.float32 if sp else np.float64
        p = lambda a: self._p(a, dt)
        if sp:
            rc = self.L.nemo_hip_tra_ldf_lap_sp(C.c_int(kt), C.c_int(kit000), cdtype.encode(), p(pahu), p(pahv),
                                                p(ptb), p(pta), C.c_int(kjpt), C.c_int(kpass))
        else:
            rc = self.L.nemo_hip_tra_ldf_lap(C.c_int(kt), C.c_int(kit000), cdtype.encode(), p(pahu), p(pahv), None,
                                             None, None, None, p(ptb), p(pta), C.c_int(kjpt), C.c_int(kpass))
        self._check(rc)

    def tra_ldf_iso(self, kt, kit000, cdtype, pahu, pahv, ptb, ptbb, pta, kjpt, kpass, uslp, vslp, wslpi, wslpj,
                    ah_wslp2, akz, sp=False):
        dt = np.float32 if sp else np.float64
        p = lambda a: self._p(a, dt)
        if sp:
            rc = self.L.nemo_hip_tra_ldf_iso_sp(C.c_int(kt), C.c_int(kit000), cdtype.encode(), p(pahu), p(pahv),
                                                p(ptb), p(ptbb), p(pta), C.c_int(kjpt), C.c_int(kpass), p(uslp),
                                                p(vslp), p(wslpi), p(wslpj), p(ah_wslp2), p(akz))
        else:
            rc = self.L.nemo_hip_tra_ldf_iso(C.c_int(kt), C.c_int(kit000), cdtype.encode(), p(pahu), p(pahv), None,
                                             None, None, None, p(ptb), p(ptbb), p(pta), C.c_int(kjpt), C.c_int(kpass),
                                             p(uslp), p(vslp), p(wslpi), p(wslpj), p(ah_wslp2), p(akz))
        self._check(rc)

    def tra_zdf_imp(self, kt, kit000, cdtype, p2dt, ptb, pta, kjpt, avt, avs, l_ldfslp=0, ah_wslp2=None, akz=None):
        p = self._p
        self._check(self.L.nemo_hip_tra_zdf_imp(C.c_int(kt), C.c_int(kit000), cdtype.encode(), C.c_double(p2dt),
                                                p(ptb), p(pta), C.c_int(kjpt), p(avt), p(avs), C.c_int(l_ldfslp),
                                                p(ah_wslp2), p(akz)))

    def tra_zdf(self, kt, tsb, tsa, avt, avs, l_ldfslp=0, ah_wslp2=None, akz=None, sp=False):
        fn = self.L.nemo_hip_tra_zdf_sp if sp else self.L.nemo_hip_tra_zdf
        dt = np.float32 if sp else np.float64
        p = lambda a: self._p(a, dt)
        self._check(fn(C.c_int(kt), p(tsb), p(tsa), p(avt), p(avs), C.c_int(l_ldfslp), p(ah_wslp2), p(akz)))

    # momentum ------------------------------------------------------------
    def dyn_keg(self, kt, kscheme, un, vn, ua, va):
        p = self._p
        self._check(self.L.nemo_hip_dyn_keg(C.c_int(kt), C.c_int(kscheme), p(un), p(vn), p(ua), p(va)))

    def dyn_zad(self, kt, un, vn, wn, ua, va):
        p = self._p
        self._check(self.L.nemo_hip_dyn_zad(C.c_int(kt), p(un), p(vn), p(wn), p(ua), p(va)))

    def dyn_adv_cen2(self, kt, un, vn, wn, ua, va):
        p = self._p
        self._check(self.L.nemo_hip_dyn_adv_cen2(C.c_int(kt), p(un), p(vn), p(wn), p(ua), p(va)))

    def dyn_adv(self, kt, un, vn, wn, ua, va, sp=False):
        fn = self.L.nemo_hip_dyn_adv_sp if sp else self.L.nemo_hip_dyn_adv
        dt = np.float32 if sp else np.float64
        p = lambda a: self._p(a, dt)
        self._check(fn(C.c_int(kt), p(un), p(vn), p(wn), p(ua), p(va)))

    def dyn_vor(self, kt, un, vn, ua, va, sp=False):
        fn = self.L.nemo_hip_dyn_vor_sp if sp else self.L.nemo_hip_dyn_vor
        dt = np.float32 if sp else np.float64
        p = lambda a: self._p(a, dt)
        self._check(fn(C.c_int(kt), p(un), p(vn), p(ua), p(va)))

    def dyn_zdf(self, kt, ub, vb, ua, va, avm, rCdU_bot, ua_b, va_b, utau, utau_b, vtau, vtau_b, sp=False):
        fn = self.L.nemo_hip_dyn_zdf_sp if sp else self.L.nemo_hip_dyn_zdf
        dt = np.float32 if sp else np.float64
        p = lambda a: self._p(a, dt)
        q = self._p
        self._check(fn(C.c_int(kt), p(ub), p(vb), p(ua), p(va), p(avm), q(rCdU_bot), q(ua_b), q(va_b), q(utau),
                       q(utau_b), q(vtau), q(vtau_b)))

    def ts_wgt(self, ll_av, ll_fw):
        n = 3 * self.dom.opts["nn_baro"]
        w1, w2 = np.zeros(n), np.zeros(n)
        jpit = C.c_int(0)
        self._check(self.L.nemo_hip_ts_wgt(C.c_int(ll_av), C.c_int(ll_fw), C.byref(jpit),
                                           w1.ctypes.data_as(C.c_void_p), w2.ctypes.data_as(C.c_void_p)))
        return jpit.value, w1, w2

    def dyn_spg_ts(self, kt, fields, sp=False):
        """fields: dict named as nemo_hip_spg_ts_t (3D: ua, va, un, vn; 2D FP64 others)."""
        s = abi.SpgTsStruct()
        for n, _ in abi.SpgTsStruct._fields_:
            a = fields.get(n)
            if a is None:
                setattr(s, n, None)
                continue
            dt = np.float32 if (sp and n in abi.SPG_3D_FIELDS) else np.float64
            setattr(s, n, self._p(a, dt).value)
        fn = self.L.nemo_hip_dyn_spg_ts_sp if sp else self.L.nemo_hip_dyn_spg_ts
        self._check(fn(C.c_int(kt), C.byref(s)))

    # fused steps ----------------------------------------------------------
    def step_dyn(self, kt, ub, vb, wn, avm, utau_b, vtau_b, fields, sp=False):
        """ua = va = 0; dyn_adv; dyn_vor; dyn_spg_ts; dyn_zdf on resident fields.
        fields: dict named as nemo_hip_spg_ts_t (3D: ua, va, un, vn)."""
        s = abi.SpgTsStruct()
        dt = np.float32 if sp else np.float64
        for n, _ in abi.SpgTsStruct._fields_:
            a = fields.get(n)
            if a is None:
                setattr(s, n, None)
                continue
            setattr(s, n, self._p(a, dt if n in abi.SPG_3D_FIELDS else np.float64).value)
        p = lambda a: self._p(a, dt)
        q = self._p
        fn = self.L.nemo_hip_step_dyn_sp if sp else self.L.nemo_hip_step_dyn
        self._check(fn(C.c_int(kt), p(ub), p(vb), p(wn), p(avm), q(utau_b), q(vtau_b), C.byref(s)))

    def step_tra(self, kt, un, vn, wn, tsb, tsn, tsa, kn_fct_h, kn_fct_v, ahtu, ahtv, uslp, vslp, wslpi, wslpj,
                 ah_wslp2, akz, avt, sp=False):
        """tsa = 0; tra_adv (transport + FCT); tra_ldf_iso; tra_zdf on resident fields."""
        dt = np.float32 if sp else np.float64
        p = lambda a: self._p(a, dt)
        fn = self.L.nemo_hip_step_tra_sp if sp else self.L.nemo_hip_step_tra
        self._check(fn(C.c_int(kt), p(un), p(vn), p(wn), p(tsb), p(tsn), p(tsa), C.c_int(kn_fct_h),
                       C.c_int(kn_fct_v), p(ahtu), p(ahtv), p(uslp), p(vslp), p(wslpi), p(wslpj), p(ah_wslp2),
                       p(akz), p(avt)))


    def tra_qsr(self, kt, nksr, qsr, gdepw, qsr_hc, qsr_hc_b, tsa, fraqsr_1lev, kqsr_b=0, rn_abs=0.58, rn_si0=0.35,
                rn_si1=23.0, sp=False):
        """tra_qsr, two bands (namelist_ref's rn_abs, rn_si0, rn_si1 by default)"""
        fn = self.L.nemo_hip_tra_qsr_sp if sp else self.L.nemo_hip_tra_qsr
        dt = np.float32 if sp else np.float64
        p = self._p
        self._check(fn(C.c_int(kt), C.c_int(kqsr_b), C.c_int(nksr), C.c_double(rn_abs), C.c_double(rn_si0),
                       C.c_double(rn_si1), p(qsr), p(gdepw, dt), p(qsr_hc), p(qsr_hc_b), p(tsa, dt),
                       p(fraqsr_1lev)))

    def tra_sbc(self, kt, qns, qsr, sfx, emp, tsn, tsa, sbc_tsc, sbc_tsc_b, sp=False):
        fn = self.L.nemo_hip_tra_sbc_sp if sp else self.L.nemo_hip_tra_sbc
        dt = np.float32 if sp else np.float64
        q = self._p
        self._check(fn(C.c_int(kt), q(qns), q(qsr), q(sfx), q(emp), q(tsn, dt), q(tsa, dt), q(sbc_tsc),
                       q(sbc_tsc_b)))

    def stream_copy(self, src, dst, nbytes, blocks):
        self._check(self.L.nemo_hip_stream_copy(self._p(src), self._p(dst), C.c_size_t(nbytes), C.c_int(blocks)))

    # tuning switches -------------------------------------------------------------
    def tune_set(self, key, value):
        self._check(self.L.nemo_hip_tune_set(str(key).encode(), C.c_int(int(value))))

    def tune_get(self, key):
        v = C.c_int()
        self._check(self.L.nemo_hip_tune_get(str(key).encode(), C.byref(v)))
        return v.value

    # compensated global sums -----------------------------------------------------
    def glob_sum(self, ptab, nk, kind=0, out=None, sp=False):
        """glob_sum (kind 0) / glob_sum_full (kind 1) of ptab(:,:,1:nk).  Device
        mode: out is a one-element FP64 device tensor (returned; the call is
        asynchronous).  Host mode: returns the sum."""
        fn = self.L.nemo_hip_glob_sum_sp if sp else self.L.nemo_hip_glob_sum
        dt = np.float32 if sp else np.float64
        if self.mode == NEMO_HIP_PTR_DEVICE:
            self._check(fn(self._p(ptab, dt), C.c_int(nk), C.c_int(kind), self._p(out)))
            return out
        r = np.zeros(1)
        self._check(fn(self._p(ptab, dt), C.c_int(nk), C.c_int(kind), self._p(r)))
        return float(r[0])

    def glob_sum_noalloc(self, ptab, nk, kind, out):
        """nemo_hip_glob_sum_noalloc: FP64 device arrays only; out is a
        one-element FP64 device tensor, written asynchronously."""
        self._check(self.L.nemo_hip_glob_sum_noalloc(self._p(ptab), C.c_int(nk), C.c_int(kind), self._p(out)))
        return out

    def _hsb(self, fields, sp, state=None):
        h = abi.HsbStruct() if state is None else state
        h.real_bytes = 4 if sp else 8
        dt3 = np.float32 if sp else np.float64
        for n in abi.HSB_PTR_FIELDS:
            a = fields.get(n)
            if a is not None:
                setattr(h, n, self._p(a, dt3 if n in ("tsn", "tsb", "wn") else np.float64).value)
        return h

    def dia_hsb_init(self, sp=False, **fields):
        """dia_hsb_init: fills the *_ini arrays given; returns the nemo_hip_hsb_t
        state (surf_tot set, frc_* zero) to pass to dia_hsb."""
        h = self._hsb(fields, sp)
        self._check(self.L.nemo_hip_dia_hsb_init(C.byref(h)))
        return h

    def dia_hsb(self, kt, state, sp=False, **fields):
        """dia_hsb: updates state (frc_*, sums, diagnostics) in place; returns it"""
        h = self._hsb(fields, sp, state)
        self._check(self.L.nemo_hip_dia_hsb(C.c_int(kt), C.byref(h)))
        return h

    # run control ----------------------------------------------------------------
    def stp_ctl(self, kt, sshn, un, tsn, sp=False):
        """stp_ctl: returns (status, zmax[6], loc[11], kindic, message).  status
        is 0, or NEMO_HIP_ERR_NUMERIC when a test fails (the reference's
        ctl_stop); other errors raise."""
        fn = self.L.nemo_hip_stp_ctl_sp if sp else self.L.nemo_hip_stp_ctl
        dt = np.float32 if sp else np.float64
        zmax = np.zeros(6)
        loc = np.zeros(11, np.int32)
        kind = C.c_int(0)
        rc = fn(C.c_int(kt), self._p(sshn), self._p(un, dt), self._p(tsn, dt), zmax.ctypes.data_as(C.c_void_p),
                loc.ctypes.data_as(C.c_void_p), C.byref(kind))
        if rc not in (0, abi.NEMO_HIP_ERR_NUMERIC):
            self._check(rc)
        return rc, zmax, loc, kind.value, (self.last_error() if rc else "")

    # density, pressure gradient, lateral viscosity -----------------------------
    def eos_register(self, eos=None):
        """eos: abi.EosStruct (default: the simplified EOS with its defaults)."""
        self._eos = eos if eos is not None else abi.seos()
        self._check(self.L.nemo_hip_eos_register(C.byref(self._eos)))

    def eos_insitu(self, pts, prd, pdep, sp=False):
        fn = self.L.nemo_hip_eos_insitu_sp if sp else self.L.nemo_hip_eos_insitu
        dt = np.float32 if sp else np.float64
        self._check(fn(self._p(pts, dt), self._p(prd, dt), self._p(pdep, dt)))

    def eos_insitu_pot(self, pts, prd, prhop, pdep, sp=False):
        fn = self.L.nemo_hip_eos_insitu_pot_sp if sp else self.L.nemo_hip_eos_insitu_pot
        dt = np.float32 if sp else np.float64
        self._check(fn(self._p(pts, dt), self._p(prd, dt), self._p(prhop, dt), self._p(pdep, dt)))

    def rab_3d(self, pts, pab, gdept, sp=False):
        fn = self.L.nemo_hip_rab_3d_sp if sp else self.L.nemo_hip_rab_3d
        dt = np.float32 if sp else np.float64
        self._check(fn(self._p(pts, dt), self._p(pab, dt), self._p(gdept, dt)))

    def bn2(self, pts, pab, pn2, gdept, gdepw, sp=False):
        fn = self.L.nemo_hip_bn2_sp if sp else self.L.nemo_hip_bn2
        dt = np.float32 if sp else np.float64
        p = lambda a: self._p(a, dt)
        self._check(fn(p(pts), p(pab), p(pn2), p(gdept), p(gdepw)))

    # tangent linear and adjoint of the tracer step (one tracer, FP64) --------
    def tra_adv_fct_coef(self, p2dt, pun, pvn, pwn, ptb, ptn, cx, cy, cz, kn=(2, 2)):
        p = self._p
        self._check(self.L.nemo_hip_tra_adv_fct_coef(C.c_double(p2dt), p(pun), p(pvn), p(pwn), p(ptb), p(ptn),
                                                     C.c_int(kn[0]), C.c_int(kn[1]), p(cx), p(cy), p(cz)))

    def tra_adv_fct_tl(self, p2dt, pun, pvn, pwn, ptb, ptn, cx, cy, cz, pta, kn=(2, 2)):
        p = self._p
        self._check(self.L.nemo_hip_tra_adv_fct_tl(C.c_double(p2dt), p(pun), p(pvn), p(pwn), p(ptb), p(ptn),
                                                   C.c_int(kn[0]), C.c_int(kn[1]), p(cx), p(cy), p(cz), p(pta)))

    def tra_adv_fct_ad(self, pun, pvn, pwn, cx, cy, cz, ta_ad, tb_ad, tn_ad, kn=(2, 2)):
        p = self._p
        self._check(self.L.nemo_hip_tra_adv_fct_ad(p(pun), p(pvn), p(pwn), C.c_int(kn[0]), C.c_int(kn[1]), p(cx),
                                                   p(cy), p(cz), p(ta_ad), p(tb_ad), p(tn_ad)))

    def tra_ldf_lap_ad(self, pahu, pahv, ta_ad, tb_ad):
        p = self._p
        self._check(self.L.nemo_hip_tra_ldf_lap_ad(p(pahu), p(pahv), p(ta_ad), p(tb_ad)))

    def tra_zdf_imp_ad(self, p2dt, avt, ta_ad, tb_ad):
        p = self._p
        self._check(self.L.nemo_hip_tra_zdf_imp_ad(C.c_double(p2dt), p(avt), p(ta_ad), p(tb_ad)))

    def tra_nxt_ad(self, tb_ad, tn_ad, ta_ad):
        p = self._p
        self._check(self.L.nemo_hip_tra_nxt_ad(p(tb_ad), p(tn_ad), p(ta_ad)))

    def lbc_adj(self, a, nk):
        self._check(self.L.nemo_hip_lbc_adj(self._p(a), C.c_int(nk)))

    def _rst(self, fields, sp):
        r = abi.RstStruct()
        r.real_bytes = 4 if sp else 8
        dt3 = np.float32 if sp else np.float64
        for n in abi.RST_FIELDS:
            a = fields.get(n)
            if a is not None:
                fp64 = n in ("sshb", "sshn", "ub2_b", "vb2_b", "un_bf", "vn_bf", "un_b", "vn_b", "ub_b", "vb_b",
                             "r1_hu_n", "r1_hv_n", "qsr_hc", "qsr_hc_b", "fraqsr_1lev")
                setattr(r, n, self._p(a, np.float64 if fp64 else dt3).value)
        return r

    def rst_write(self, path, kt, sp=False, **fields):
        """rst_write: the fields named as in nemo_hip_rst_t into the directory path."""
        r = self._rst(fields, sp)
        self._check(self.L.nemo_hip_rst_write(str(path).encode(), C.c_int(kt), C.byref(r)))

    def rst_read(self, path, sp=False, **fields):
        """rst_read: fills the fields; returns (kt, neuler).  self.kqsr_b is
        then 1 when the file held qsr_hc_b (and it was asked for), else 0."""
        r = self._rst(fields, sp)
        kt, ne = C.c_int(), C.c_int()
        self._check(self.L.nemo_hip_rst_read(str(path).encode(), C.byref(r), C.byref(kt), C.byref(ne)))
        self.kqsr_b = r.kqsr_b
        return kt.value, ne.value

    def zdf_mxl(self, kt, rn2b, gdept, gdepw, nlb10, nmln, hmlp, hmlpt, rho_c=0.01, sp=False):
        """zdf_mxl: nmln (int32), hmlp, hmlpt (FP64, jpi x jpj) from rn2b."""
        fn = self.L.nemo_hip_zdf_mxl_sp if sp else self.L.nemo_hip_zdf_mxl
        dt = np.float32 if sp else np.float64
        self._check(fn(C.c_int(kt), self._p(rn2b, dt), self._p(gdept, dt), self._p(gdepw, dt), C.c_int(nlb10),
                       C.c_double(rho_c), self._p(nmln, np.int32), self._p(hmlp), self._p(hmlpt)))

    def ldf_slp(self, kt, prd, pn2, gdept, gdepw, nmln, hmlp, hmlpt, rn_slpmax, uslp, vslp, wslpi, wslpj,
                sp=False):
        """ldf_slp: the iso-neutral slopes (updated in place)."""
        fn = self.L.nemo_hip_ldf_slp_sp if sp else self.L.nemo_hip_ldf_slp
        dt = np.float32 if sp else np.float64
        self._check(fn(C.c_int(kt), self._p(prd, dt), self._p(pn2, dt), self._p(gdept, dt), self._p(gdepw, dt),
                       self._p(nmln, np.int32), self._p(hmlp), self._p(hmlpt), C.c_double(rn_slpmax),
                       self._p(uslp, dt), self._p(vslp, dt), self._p(wslpi, dt), self._p(wslpj, dt)))

    def hpg_zco(self, kt, rhd, ua, va, sp=False):
        fn = self.L.nemo_hip_hpg_zco_sp if sp else self.L.nemo_hip_hpg_zco
        dt = np.float32 if sp else np.float64
        self._check(fn(C.c_int(kt), self._p(rhd, dt), self._p(ua, dt), self._p(va, dt)))

    def dyn_ldf_lap(self, kt, pub, pvb, pua, pva, kpass, ahmt, ahmf, sp=False):
        fn = self.L.nemo_hip_dyn_ldf_lap_sp if sp else self.L.nemo_hip_dyn_ldf_lap
        dt = np.float32 if sp else np.float64
        p = lambda a: self._p(a, dt)
        self._check(fn(C.c_int(kt), p(pub), p(pvb), p(pua), p(pva), C.c_int(kpass), p(ahmt), p(ahmf)))

    # time-step closure ------------------------------------------------------
    def div_hor(self, kt, un, vn, hdivn, sp=False):
        fn = self.L.nemo_hip_div_hor_sp if sp else self.L.nemo_hip_div_hor
        dt = np.float32 if sp else np.float64
        p = lambda a: self._p(a, dt)
        self._check(fn(C.c_int(kt), p(un), p(vn), p(hdivn)))

    def ssh_nxt(self, kt, un, vn, hdivn, sshb, emp_b, emp, ssha, sp=False):
        fn = self.L.nemo_hip_ssh_nxt_sp if sp else self.L.nemo_hip_ssh_nxt
        dt = np.float32 if sp else np.float64
        p = lambda a: self._p(a, dt)
        q = self._p
        self._check(fn(C.c_int(kt), p(un), p(vn), p(hdivn), q(sshb), q(emp_b), q(emp), q(ssha)))

    def wzv(self, kt, hdivn, wn, sp=False):
        fn = self.L.nemo_hip_wzv_sp if sp else self.L.nemo_hip_wzv
        dt = np.float32 if sp else np.float64
        self._check(fn(C.c_int(kt), self._p(hdivn, dt), self._p(wn, dt)))

    def ssh_swp(self, kt, sshb, sshn, ssha):
        self._check(self.L.nemo_hip_ssh_swp(C.c_int(kt), self._p(sshb), self._p(sshn), self._p(ssha)))

    def tra_nxt(self, kt, tsb, tsn, tsa, sp=False):
        fn = self.L.nemo_hip_tra_nxt_sp if sp else self.L.nemo_hip_tra_nxt
        dt = np.float32 if sp else np.float64
        p = lambda a: self._p(a, dt)
        self._check(fn(C.c_int(kt), p(tsb), p(tsn), p(tsa)))

    def dyn_nxt(self, kt, ub, vb, un, vn, ua, va, ua_b, va_b, un_adv, vn_adv, ub_b, vb_b, un_b, vn_b, r1_hu_n,
                r1_hv_n, sp=False):
        fn = self.L.nemo_hip_dyn_nxt_sp if sp else self.L.nemo_hip_dyn_nxt
        dt = np.float32 if sp else np.float64
        p = lambda a: self._p(a, dt)
        q = self._p
        self._check(fn(C.c_int(kt), p(ub), p(vb), p(un), p(vn), p(ua), p(va), q(ua_b), q(va_b), q(un_adv),
                       q(vn_adv), q(ub_b), q(vb_b), q(un_b), q(vn_b), q(r1_hu_n), q(r1_hv_n)))


_handle = None


def get(device=0, mode=NEMO_HIP_PTR_DEVICE):
    """The process-wide handle (created on first use)."""
    global _handle
    if _handle is None:
        _handle = NemoHip(device, mode)
    else:
        _handle.set_pointer_mode(mode)
    return _handle


def exported_symbols(header=HEADER):
    """Names of the entry points declared in include/nemo_hip.h."""
    import re
    txt = open(header).read()
    return sorted(set(re.findall(r"\b(?:int|void\s*\*)\s+(nemo_hip_\w+)\s*\(", txt)))
