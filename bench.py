#!/usr/bin/env python3
"""Benchmark of the NEMO ocean time-step hot path on one MI355X.

Workload: 1024 x 1024 x 75 seeded basin (about 30% land, lognormal bathymetry
on full-step z-levels, 10 km grid), rdt = 600 s, nn_baro = 30, FCT 4/4,
iso-neutral laplacian, ENE vorticity / vector-form momentum advection,
implicit vertical diffusion and bottom drag.  One "step" is the hot path of
step.F90 over the whole domain, with every input already resident in HBM:

  ua = va = 0;  dyn_adv;  dyn_vor;  dyn_spg_ts;  dyn_zdf
  tsa = 0;      tra_adv (transport + tra_adv_fct);  tra_ldf_iso;  tra_zdf

Two storage forms are timed: the plain FP64 form and the FP32-storage form
(FP32 stencils; FP64 tridiagonal eliminations and barotropic subcycles).
value = grid-cell updates/s of the FP32 form = jpi*jpj*jpk / t_step.

Also reported: per-routine / per-kernel times from HIP events on the
library's stream, the roofline of the dominant kernel, the step's fraction of
HBM peak from the algorithmic bytes, and the CPU oracle (plain C restatement
of the reference loops, one core) timed on a bounded sub-domain.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for p in (ROOT, os.path.join(ROOT, "nemo-fmi-devel_amd")):
    if p not in sys.path:
        sys.path.insert(0, p)

METRIC = "grid-cell updates/s per step + % HBM peak, 1024\u00b2\u00d775, FP32 vs FP64 form"
HBM_PEAK = 8.0e12

# Algorithmic HBM bytes per grid cell of each routine (3D arrays the routine's
# arithmetic reads plus those it writes, once each, at storage width; masks,
# 1-D e3 tables and 2D metrics count 0; temporaries count 0).
#   dyn_adv (keg+zad)  r un vn wn ua va            w ua va          7
#   dyn_vor            r un vn ua va               w ua va          6
#   dyn_spg_ts         r ua va un vn (+ua va again for phase 3)     w ua va un vn   8
#   dyn_zdf            r ub vb ua va avm           w ua va          7
#   tra_adv_trp        r un vn wn                  w zun zvn zwn    6
#   tra_adv_fct (x2)   r zun zvn zwn tb tn ta      w ta             3 + 2*4 = 11
#   tra_ldf_iso        r ahtu ahtv uslp vslp wslpi wslpj tb*2 ta*2  w ta*2 ah_wslp2 akz  14
#   tra_zdf            r avt ah_wslp2 tb*2 ta*2    w ta*2           8
#   zeroing ua va tsa                              w 4              4
ROUTINE_ARRAYS = {"dyn_adv": 7, "dyn_vor": 6, "dyn_spg_ts": 8, "dyn_zdf": 7, "tra_adv_trp": 6,
                  "tra_adv_fct": 11, "tra_ldf_iso": 14, "tra_zdf": 8, "zero": 4}
def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--size", default="1024x1024x75")
    ap.add_argument("--forms", default="fp64,fp32,fp64_fused,fp32_fused")
    ap.add_argument("--cpu-sample", default="512x512x75")
    ap.add_argument("--cpu-steps", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-workers", type=int, default=16, metavar="N",
                    help="processes of the all-cores CPU baseline (one sub-domain each, as NEMO's MPI ranks; "
                         "capped at the cores this process may use)")
    ap.add_argument("--cpu-worker-sample", default="256x256x75")
    ap.add_argument("--ensemble", type=int, default=24, metavar="MEMBERS",
                    help="also step an ensemble of MEMBERS domains (i x j drawn from 256x320 .. 768x640, 56 levels, "
                         "FP32 fused step, one library context each), eagerly and as one recorded HIP graph per "
                         "step: reported apart, never the headline value")
    ap.add_argument("--host-mode", type=int, default=0, metavar="STEPS",
                    help="also time STEPS steps of the FP64 routines in host-pointer mode (NEMO's host arrays, "
                         "every call staged over PCIe): reported apart, never the headline value")
    ap.add_argument("--whole-step", default="fp32_fused,fp64_fused", metavar="FORMS",
                    help="also time, for these fused forms, the whole time step: the hot path closed by ssh_nxt, "
                         "wzv, div_hor, tra_nxt, dyn_nxt and ssh_swp (reported apart; '' skips)")
    ap.add_argument("--physics-step", default="fp32,fp64", metavar="FORMS",
                    help="also time, for these forms, the whole step with its physics: the iso-neutral slopes of "
                         "the state (rab_3d, bn2, zdf_mxl, eos_insitu, ldf_slp), the density, hpg_zco, "
                         "dyn_ldf_lap, tra_sbc and the separate hot-path routines (TEOS-10)")
    ap.add_argument("--adjoint", default="512x512x75", metavar="SIZE",
                    help="also time the adjoint form at this size ('' to skip): the gradient of an SST misfit "
                         "after --adjoint-steps tracer steps, checkpoints every 10 steps")
    ap.add_argument("--adjoint-steps", type=int, default=50)
    ap.add_argument("--adjoint-dirs", type=int, default=64, metavar="N",
                    help="finite-difference directions for the adjoint gradient check")
    ap.add_argument("--seed", type=int, default=20261204)
    ap.add_argument("--timing-steps", type=int, default=2)
    ap.add_argument("--tuning", default=os.path.join(ROOT, "nemo-fmi-devel_amd", "nemo_hip", "tuning.json"),
                    metavar="PATH", help="settings of the library's switches per workload (nemo_hip.autotune); the "
                    "setting saved for a form's workload is applied before that form is timed")
    ap.add_argument("--autotune", default="", metavar="PATH",
                    help="first tune the switches on each form's step (nemo_hip.autotune) and save the settings to "
                         "PATH (then applied, as from --tuning)")
    ap.add_argument("--tune", default="", metavar="KEY=V,...",
                    help="library tuning switches (nemo_hip_tune_set) to set before timing, e.g. spg_persistent=1")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a device-to-device copy of MIB MiB (three times): the reference kernel the "
                         "FETCH_SIZE / WRITE_SIZE counters are scaled by")
    return ap.parse_args()


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def build_case(jpi, jpj, jpk, seed):
    from nemo_hip import configs
    opts = dict(nvor_scheme=1, n_dynadv=1, nn_dynkeg=0, ln_traldf_msc=0, nn_baro=30, rdt=600.0, neuler=0, nit000=1)
    d = configs.basin(jpi, jpj, jpk, seed=seed, opts=opts)
    return d


class Stepper:
    """The hot path of one time step, in one storage form, on resident arrays.
    fused=False: the separate drop-in routines in step order; fused=True: the
    two fused step entry points (same results, bit for bit)."""

    def __init__(self, h, d, S, sp, fused=False, physics=False):
        import torch
        self.h, self.d, self.sp, self.fused = h, d, sp, fused
        dt = torch.float32 if sp else torch.float64
        n3 = d.jpi * d.jpj * d.jpk
        c = lambda k: S[k].to(dt, copy=True).contiguous()   # the steps update their own copies
        self.f3 = {k: c(k) for k in ("un", "vn", "wn", "ub", "vb", "avt", "avm", "ahtu", "ahtv", "uslp", "vslp",
                                     "wslpi", "wslpj")}
        self.f3["avs"] = self.f3["avt"]
        self.tsb, self.tsn = c("tsb"), c("tsn")
        z = lambda n: torch.zeros(n, dtype=dt, device="cuda")
        self.tsa, self.ua, self.va = z(2 * n3), z(n3), z(n3)
        self.zun, self.zvn, self.zwn = (None, None, None) if fused else (z(n3), z(n3), z(n3))
        self.ah, self.akz = z(n3), z(n3)
        self.f2 = {k: S[k].clone() for k in ("un_b", "vn_b", "ub_b", "vb_b", "sshn", "hu_n", "hv_n", "r1_hu_n",
                                             "r1_hv_n", "ht_n", "rCdU_bot", "utau", "vtau", "emp", "rnf", "ub2_b",
                                             "vb2_b", "un_bf", "vn_bf", "utau_b", "vtau_b")}
        n2 = d.jpi * d.jpj
        for k in ("ua_b", "va_b", "ssha", "un_adv", "vn_adv"):
            self.f2[k] = torch.zeros(n2, dtype=torch.float64, device="cuda")
        # the momentum fields dyn_spg_ts updates; the tracer transport reads the original un/vn
        self.un_dyn, self.vn_dyn = self.f3["un"].clone(), self.f3["vn"].clone()
        self.spg = dict(self.f2)
        self.spg.update(ua=self.ua, va=self.va, un=self.un_dyn, vn=self.vn_dyn)
        self.r2dt = 2.0 * d.opts["rdt"]
        # the closure's fields: divergence work array and the before sea level
        self.hdivn = z(n3)
        self.sshb = self.f2["sshn"].clone()
        if physics:
            self._physics_fields(S, dt, z, n2, n3)

    def _physics_fields(self, S, dt, z, n2, n3):
        """The fields of the physics step: depths, expansion ratios, N^2, the
        mixed layer, densities, a 2000 m2/s lateral viscosity and analytic
        surface fluxes (TEOS-10 equation of state)."""
        import numpy as np
        import torch
        from nemo_hip import abi, configs
        d, h = self.d, self.h
        dev = lambda a, t: torch.from_numpy(np.ascontiguousarray(np.asarray(a).ravel(order="F"))).to("cuda", t)
        gt, gw = configs.depths(d)
        self.gdept, self.gdepw = dev(gt, dt), dev(gw, dt)
        del gt, gw
        t, _, _, _, f = configs.masks(d)
        self.ahmt, self.ahmf = dev(2000.0 * t, dt), dev(2000.0 * f, dt)
        del t, f
        self.pab, self.rn2b, self.rhd, self.rhop = z(2 * n3), z(n3), z(n3), z(n3)
        self.nmln = torch.zeros(n2, dtype=torch.int32, device="cuda")
        f64 = lambda: torch.zeros(n2, dtype=torch.float64, device="cuda")
        self.hmlp, self.hmlpt, self.sbc, self.sbc_b = f64(), f64(), torch.zeros(2 * n2, dtype=torch.float64,
                                                                              device="cuda"), \
            torch.zeros(2 * n2, dtype=torch.float64, device="cuda")
        y = np.arange(d.jpj) / max(d.jpj - 1, 1)
        self.qns0 = dev(np.broadcast_to((-40.0 + 80.0 * np.cos(np.pi * y))[None, :], (d.jpi, d.jpj)), torch.float64)
        self.qsr0 = torch.full((n2,), 30.0, dtype=torch.float64, device="cuda")
        self.sfx = f64()
        self.qns, self.qsr = self.qns0.clone(), self.qsr0.clone()
        self.nlb10 = configs.nlb10(d)
        h.eos_register(abi.teos10())

    def physics_step(self, kt):
        """The whole time step with its physics, step.F90:127-279: the iso-
        neutral slopes of the before state (rab_3d, bn2, zdf_mxl, eos_insitu,
        ldf_slp), the now density, the pressure gradient and lateral
        viscosity, the surface tracer fluxes, the separate hot-path routines
        and the closure."""
        h, f, f2, sp = self.h, self.f3, self.f2, self.sp
        h.rab_3d(self.tsb, self.pab, self.gdept, sp=sp)
        h.bn2(self.tsb, self.pab, self.rn2b, self.gdept, self.gdepw, sp=sp)
        h.zdf_mxl(kt, self.rn2b, self.gdept, self.gdepw, self.nlb10, self.nmln, self.hmlp, self.hmlpt, sp=sp)
        h.eos_insitu(self.tsb, self.rhd, self.gdept, sp=sp)
        h.ldf_slp(kt, self.rhd, self.rn2b, self.gdept, self.gdepw, self.nmln, self.hmlp, self.hmlpt, 0.01, f["uslp"],
                  f["vslp"], f["wslpi"], f["wslpj"], sp=sp)
        h.ssh_nxt(kt, self.un_dyn, self.vn_dyn, self.hdivn, self.sshb, f2["emp"], f2["emp"], f2["ssha"], sp=sp)
        h.wzv(kt, self.hdivn, f["wn"], sp=sp)
        h.eos_insitu_pot(self.tsn, self.rhd, self.rhop, self.gdept, sp=sp)
        self.ua.zero_()
        self.va.zero_()
        h.dyn_adv(kt, self.un_dyn, self.vn_dyn, f["wn"], self.ua, self.va, sp=sp)
        h.dyn_vor(kt, self.un_dyn, self.vn_dyn, self.ua, self.va, sp=sp)
        h.dyn_ldf_lap(kt, f["ub"], f["vb"], self.ua, self.va, 1, self.ahmt, self.ahmf, sp=sp)
        h.hpg_zco(kt, self.rhd, self.ua, self.va, sp=sp)
        h.dyn_spg_ts(kt, self.spg, sp=sp)
        h.div_hor(kt, self.un_dyn, self.vn_dyn, self.hdivn, sp=sp)
        h.wzv(kt, self.hdivn, f["wn"], sp=sp)
        h.dyn_zdf(kt, f["ub"], f["vb"], self.ua, self.va, f["avm"], f2["rCdU_bot"], f2["ua_b"], f2["va_b"],
                  f2["utau"], f2["utau_b"], f2["vtau"], f2["vtau_b"], sp=sp)
        self.tsa.zero_()
        self.qns.copy_(self.qns0)
        self.qsr.copy_(self.qsr0)
        h.tra_sbc(kt, self.qns, self.qsr, self.sfx, f2["emp"], self.tsn, self.tsa, self.sbc, self.sbc_b, sp=sp)
        h.tra_adv_trp(self.un_dyn, self.vn_dyn, f["wn"], self.zun, self.zvn, self.zwn, sp=sp)
        h.tra_adv_fct(kt, 1, "TRA", self.r2dt, self.zun, self.zvn, self.zwn, self.tsb, self.tsn, self.tsa, 2, 4, 4,
                      sp=sp)
        h.tra_ldf_iso(kt, 1, "TRA", f["ahtu"], f["ahtv"], self.tsb, self.tsb, self.tsa, 2, 1, f["uslp"], f["vslp"],
                      f["wslpi"], f["wslpj"], self.ah, self.akz, sp=sp)
        h.tra_zdf(kt, self.tsb, self.tsa, f["avt"], f["avs"], 1, self.ah, self.akz, sp=sp)
        h.tra_nxt(kt, self.tsb, self.tsn, self.tsa, sp=sp)
        h.dyn_nxt(kt, f["ub"], f["vb"], self.un_dyn, self.vn_dyn, self.ua, self.va, f2["ua_b"], f2["va_b"],
                  f2["un_adv"], f2["vn_adv"], f2["ub_b"], f2["vb_b"], f2["un_b"], f2["vn_b"], f2["r1_hu_n"],
                  f2["r1_hv_n"], sp=sp)
        h.ssh_swp(kt, self.sshb, f2["sshn"], f2["ssha"])

    def whole_step(self, kt):
        """The whole time step, step.F90:161-279: the fused hot path closed by
        the sea-level, vertical-velocity and time-filter routines, so every
        prognostic field (u, v, w, T, S, ssh) advances on the device."""
        h, f, f2, sp = self.h, self.f3, self.f2, self.sp
        h.ssh_nxt(kt, self.un_dyn, self.vn_dyn, self.hdivn, self.sshb, f2["emp"], f2["emp"], f2["ssha"], sp=sp)
        h.wzv(kt, self.hdivn, f["wn"], sp=sp)
        h.step_dyn(kt, f["ub"], f["vb"], f["wn"], f["avm"], f2["utau_b"], f2["vtau_b"], self.spg, sp=sp)
        h.div_hor(kt, self.un_dyn, self.vn_dyn, self.hdivn, sp=sp)
        h.wzv(kt, self.hdivn, f["wn"], sp=sp)
        h.step_tra(kt, self.un_dyn, self.vn_dyn, f["wn"], self.tsb, self.tsn, self.tsa, 4, 4, f["ahtu"], f["ahtv"],
                   f["uslp"], f["vslp"], f["wslpi"], f["wslpj"], self.ah, self.akz, f["avt"], sp=sp)
        h.tra_nxt(kt, self.tsb, self.tsn, self.tsa, sp=sp)
        h.dyn_nxt(kt, f["ub"], f["vb"], self.un_dyn, self.vn_dyn, self.ua, self.va, f2["ua_b"], f2["va_b"],
                  f2["un_adv"], f2["vn_adv"], f2["ub_b"], f2["vb_b"], f2["un_b"], f2["vn_b"], f2["r1_hu_n"],
                  f2["r1_hv_n"], sp=sp)
        h.ssh_swp(kt, self.sshb, f2["sshn"], f2["ssha"])

    def step(self, kt):
        h, f, sp = self.h, self.f3, self.sp
        if self.fused:
            h.step_dyn(kt, f["ub"], f["vb"], f["wn"], f["avm"], self.f2["utau_b"], self.f2["vtau_b"], self.spg, sp=sp)
            h.step_tra(kt, f["un"], f["vn"], f["wn"], self.tsb, self.tsn, self.tsa, 4, 4, f["ahtu"], f["ahtv"],
                       f["uslp"], f["vslp"], f["wslpi"], f["wslpj"], self.ah, self.akz, f["avt"], sp=sp)
            return
        # step.F90:179-180 and 234: trends start at zero
        for a in (self.ua, self.va, self.tsa):
            if hasattr(a, "zero_"):
                a.zero_()
            else:
                a.fill(0.0)
        h.dyn_adv(kt, self.un_dyn, self.vn_dyn, f["wn"], self.ua, self.va, sp=sp)
        h.dyn_vor(kt, self.un_dyn, self.vn_dyn, self.ua, self.va, sp=sp)
        h.dyn_spg_ts(kt, self.spg, sp=sp)
        h.dyn_zdf(kt, f["ub"], f["vb"], self.ua, self.va, f["avm"], self.f2["rCdU_bot"], self.f2["ua_b"],
                  self.f2["va_b"], self.f2["utau"], self.f2["utau_b"], self.f2["vtau"], self.f2["vtau_b"], sp=sp)
        h.tra_adv_trp(f["un"], f["vn"], f["wn"], self.zun, self.zvn, self.zwn, sp=sp)
        h.tra_adv_fct(kt, 1, "TRA", self.r2dt, self.zun, self.zvn, self.zwn, self.tsb, self.tsn, self.tsa, 2, 4, 4,
                      sp=sp)
        h.tra_ldf_iso(kt, 1, "TRA", f["ahtu"], f["ahtv"], self.tsb, self.tsb, self.tsa, 2, 1, f["uslp"], f["vslp"],
                      f["wslpi"], f["wslpj"], self.ah, self.akz, sp=sp)
        h.tra_zdf(kt, self.tsb, self.tsa, f["avt"], f["avs"], 1, self.ah, self.akz, sp=sp)


def autotune_form(h, d, S, form, a, log):
    """tune the library's switches on one form's step at the bench workload
    and save the setting to a.autotune under "<form> <size>" """
    import torch
    from nemo_hip import autotune
    st = Stepper(h, d, S, form.startswith("fp32"), form.endswith("_fused"))
    kt = [2]

    def step():
        st.step(kt[0])
        kt[0] += 1
    with torch.cuda.stream(torch.cuda.ExternalStream(h.stream())):
        step()
        step()
        best, ms, trials = autotune.tune(h, step, steps=3, warmup=1, log=log)
    autotune.save(a.autotune, f"{form} {a.size}", best, ms, trials,
                  note="bench.py Stepper.step on the seeded basin, coordinate descent, 3 timed steps per trial")
    log(f"autotune {form}: {best} {ms:.3f} ms/step")
    del st
    torch.cuda.empty_cache()


def time_form(h, d, S, sp, fused, steps, warmup, timing_steps, live=False, whole=False, physics=False):
    """Warm up, time every kernel over `timing_steps` steps with HIP events,
    then time `steps` steps with the wall clock.  live=True also brackets the
    launches of the dominant kernel (the one with the most time per step that
    has a byte count) with HIP events on the library stream inside the timed
    region; its measured launch durations are returned for the roofline."""
    import torch
    st = Stepper(h, d, S, sp, fused, physics=physics)
    step = st.physics_step if physics else st.whole_step if whole else st.step
    ext = torch.cuda.ExternalStream(h.stream())
    with torch.cuda.stream(ext):
        kt = 2
        for _ in range(warmup):
            step(kt)
            kt += 1
        h.sync()
        # per-routine / per-kernel HIP-event timers over a short run
        h.timing_reset()
        h.timing_only(None)
        h.timing_enable(True)
        for _ in range(timing_steps):
            step(kt)
            kt += 1
        h.sync()
        h.timing_enable(False)
        ktimes = kernel_times(h)
        key = None
        if live:
            cand = {k: v[0] for k, v in ktimes.items() if "/" in k and v[1] > 0 and k.split("/")[1] in KERNEL_BYTES}
            key = max(cand, key=cand.get) if cand else None
            if key:
                h.timing_reset()
                h.timing_only(key)
                h.timing_enable(True)
        h.sync()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            step(kt)
            kt += 1
        h.sync()
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        live_res = None
        if key:
            h.timing_enable(False)
            tot, cnt = h.timing_get(key)
            h.timing_only(None)
            live_res = {"kernel": key, "total_ms": tot, "launches": cnt}
    ms = (t1 - t0) * 1e3 / steps
    finite = bool(torch.isfinite(st.tsa).all().item() and torch.isfinite(st.ua).all().item())
    del st
    torch.cuda.empty_cache()
    return ms, finite, ktimes, live_res


def time_host_mode(h, d, S, steps):
    """ms per step of the separate FP64 routines called on host arrays (the
    drop-in path a NEMO build without resident fields takes)."""
    import nemo_hip
    import torch
    st = Stepper(h, d, S, False, False)
    host = lambda t: t.detach().cpu().numpy().copy()
    for k in st.f3:
        st.f3[k] = host(st.f3[k])
    st.tsb, st.tsn, st.tsa, st.ua, st.va = (host(t) for t in (st.tsb, st.tsn, st.tsa, st.ua, st.va))
    st.zun, st.zvn, st.zwn, st.ah, st.akz = (host(t) for t in (st.zun, st.zvn, st.zwn, st.ah, st.akz))
    st.f2 = {k: host(v) for k, v in st.f2.items()}
    st.un_dyn, st.vn_dyn = host(st.un_dyn), host(st.vn_dyn)
    st.spg = dict(st.f2)
    st.spg.update(ua=st.ua, va=st.va, un=st.un_dyn, vn=st.vn_dyn)
    torch.cuda.synchronize()
    h.set_pointer_mode(nemo_hip.NEMO_HIP_PTR_HOST)
    try:
        st.step(2)
        t0 = time.perf_counter()
        for kt in range(3, 3 + steps):
            st.step(kt)
        ms = (time.perf_counter() - t0) * 1e3 / steps
    finally:
        h.set_pointer_mode(nemo_hip.NEMO_HIP_PTR_DEVICE)
    return ms


def time_adjoint(h, size, steps, ckpt, ndir, seed, kn=(4, 4)):
    """The adjoint form: J = 1/2 sum over wet surface points of (SST - target)^2
    after `steps` tracer steps (FCT of orders kn, laplacian, implicit diffusion,
    leapfrog filter; velocities and diffusivities fixed), target = SST +
    N(0, 0.1 degC); its gradient with respect to the initial temperature by
    the adjoint with the limiter factors frozen, the forward state kept every
    `ckpt` steps and each segment recomputed before its backward sweep.
    Checked by a dot-product test of one step and, along `ndir` seeded
    directions, against the tangent linear model it transposes (the forward
    trajectory's limiter factors at every step) and against central
    differences of that model (the finite-difference check); central
    differences of the nonlinear model (8 directions) measure what freezing
    the limiter costs."""
    import numpy as np
    import torch
    from nemo_hip import configs
    jpi, jpj, jpk = [int(x) for x in size.split("x")]
    ctx = h.context_create()
    h.context_select(ctx)
    ext = torch.cuda.ExternalStream(h.stream())
    try:
        with torch.cuda.stream(ext):
            d = build_case(jpi, jpj, jpk, seed)
            h.register(restarted_opts(d))
            S = configs.make_state_device(d, seed=seed + 1)
            n3 = jpi * jpj * jpk
            f64 = lambda: torch.zeros(n3, dtype=torch.float64, device="cuda")
            zun, zvn, zwn = f64(), f64(), f64()
            h.tra_adv_trp(S["un"], S["vn"], S["wn"], zun, zvn, zwn)
            p2dt, kt = 2.0 * d.opts["rdt"], 5
            ahtu, ahtv, avt = S["ahtu"], S["ahtv"], S["avt"]
            T0 = S["tsn"][:n3].clone()
            h.lbc_lnk(T0, jpk, "T")
            B, N, A = (torch.zeros(2 * n3, dtype=torch.float64, device="cuda") for _ in range(3))
            c = [f64() for _ in range(3)]

            def step(tb, tn, cc, own):
                """one tracer step from (tb, tn); own: the factors of the state
                itself (the nonlinear step: tra_adv_fct, which equals the
                frozen-factor trend with the state's own factors bit for bit),
                else the frozen factors cc"""
                ta = A[:n3]
                ta.zero_()
                if own:
                    h.tra_adv_fct(kt, 1, "TRA", p2dt, zun, zvn, zwn, tb, tn, ta, 1, *kn)
                else:
                    h.tra_adv_fct_tl(p2dt, zun, zvn, zwn, tb, tn, *cc, ta, kn=kn)
                h.tra_ldf_lap(kt, 1, "TRA", ahtu, ahtv, tb, ta, 1, 1)
                h.tra_zdf_imp(kt, 1, "TRA", p2dt, tb, ta, 1, avt, avt)
                B[:n3].copy_(tb)
                N[:n3].copy_(tn)
                h.tra_nxt(kt, B, N, A)
                return B[:n3].clone(), N[:n3].clone()

            def run(T, keep=None, frozen=None):
                """the nonlinear model, or with frozen: the linear model with
                those factors at every step"""
                b, n = T.clone(), T.clone()
                for q in range(steps):
                    if keep is not None and q % ckpt == 0:
                        keep.append((b.clone(), n.clone()))
                    b, n = step(b, n, frozen[q] if frozen else c, frozen is None)
                return n

            t, *_ = configs.masks(d)
            wgt = np.zeros((jpi, jpj, jpk))
            wgt[1:-1, 1:-1, 0] = t[1:-1, 1:-1, 0]
            wgt = torch.from_numpy(wgt.ravel(order="F")).to("cuda")
            del t
            nN = run(T0)
            rng = np.random.default_rng(seed + 2)
            target = nN + 0.1 * torch.from_numpy(rng.standard_normal(n3)).to("cuda")
            J = lambda n: 0.5 * float(torch.sum(wgt * (n - target) ** 2))
            # timed: forward with checkpoints, then the segmented adjoint sweep
            h.sync()
            t0 = time.perf_counter()
            chk = []
            nN = run(T0, keep=chk)
            h.sync()
            t1 = time.perf_counter()
            lb, ln = torch.zeros_like(nN), wgt * (nN - target)
            la = f64()
            for s0 in reversed(range(0, steps, ckpt)):
                b, n = chk[s0 // ckpt]
                seg = []
                for q in range(s0, min(s0 + ckpt, steps)):
                    cc = [f64() for _ in range(3)]
                    h.tra_adv_fct_coef(p2dt, zun, zvn, zwn, b, n, *cc, kn=kn)
                    seg.append(cc)
                    b, n = step(b, n, None, True)
                for cc in reversed(seg):
                    h.tra_nxt_ad(lb, ln, la)
                    h.tra_zdf_imp_ad(p2dt, avt, la, lb)
                    h.tra_ldf_lap_ad(ahtu, ahtv, la, lb)
                    h.tra_adv_fct_ad(zun, zvn, zwn, *cc, la, lb, ln, kn=kn)
                del seg
            grad = lb + ln
            h.lbc_adj(grad, jpk)
            h.sync()
            t2 = time.perf_counter()
            # dot-product test of one frozen step (factors of the initial state)
            h.tra_adv_fct_coef(p2dt, zun, zvn, zwn, T0, T0, *c, kn=kn)
            xb, xn, yb, yn = (torch.from_numpy(rng.standard_normal(n3)).to("cuda") for _ in range(4))
            b1, n1 = step(xb, xn, c, False)
            gb, gn = yb.clone(), yn.clone()
            h.tra_nxt_ad(gb, gn, la)
            h.tra_zdf_imp_ad(p2dt, avt, la, gb)
            h.tra_ldf_lap_ad(ahtu, ahtv, la, gb)
            h.tra_adv_fct_ad(zun, zvn, zwn, *c, la, gb, gn, kn=kn)
            lhs = float(torch.dot(b1, yb) + torch.dot(n1, yn))
            rhs = float(torch.dot(xb, gb) + torch.dot(xn, gn))
            dot_err = abs(lhs - rhs) / max(abs(lhs), abs(rhs))
            # the gradient against the tangent linear model it transposes (the
            # forward trajectory's limiter factors at every step): <g, d> =
            # <r, L d>, r the final misfit, L d the linear model run on d (the
            # exact limit of central differences of that model); and against
            # central differences of the nonlinear model (8 directions), the
            # error of freezing the limiter
            frozen, b, n = [], T0.clone(), T0.clone()
            for q in range(steps):
                cc = [f64() for _ in range(3)]
                h.tra_adv_fct_coef(p2dt, zun, zvn, zwn, b, n, *cc, kn=kn)
                frozen.append(cc)
                b, n = step(b, n, cc, False)
            r = wgt * (n - target)
            interior = torch.zeros((jpk, jpj, jpi), dtype=torch.float64, device="cuda")
            interior[:-1, 1:-1, 1:-1] = 1.0
            interior = interior.reshape(-1)
            errs, errs_nl, errs_fd, growth = [], [], [], []
            t3 = time.perf_counter()
            for q in range(ndir):
                dirn = torch.from_numpy(np.random.default_rng(seed + 100 + q).standard_normal(n3)).to("cuda") * interior
                h.lbc_lnk(dirn, jpk, "T")
                ad = float(torch.dot(grad, dirn))
                Ld = run(dirn, frozen=frozen)
                tl = float(torch.dot(r, Ld))
                errs.append(abs(tl - ad) / max(abs(tl), 1e-300))
                growth.append(float(torch.linalg.vector_norm(Ld) / torch.linalg.vector_norm(dirn)))
                # central differences of the model the adjoint transposes (the
                # forward trajectory's factors frozen): J is quadratic in the
                # initial state there, so the differences carry rounding only;
                # the step moves the final state by 1e-3 of the direction
                eps = 1e-3 / max(1.0, growth[-1])
                fd = (J(run(T0 + eps * dirn, frozen=frozen)) - J(run(T0 - eps * dirn, frozen=frozen))) / (2 * eps)
                errs_fd.append(abs(fd - ad) / max(abs(fd), 1e-300))
                if q < 8:
                    eps = 1e-3
                    fd_nl = (J(run(T0 + eps * dirn)) - J(run(T0 - eps * dirn))) / (2 * eps)
                    errs_nl.append(abs(fd_nl - ad) / max(abs(fd_nl), 1e-300))
            t4 = time.perf_counter()
            del frozen
    finally:
        h.context_select(0)
        torch.cuda.empty_cache()
        h.context_destroy(ctx)
    errs, errs_nl, errs_fd = np.array(errs), np.array(errs_nl), np.array(errs_fd)
    ndir = len(errs)
    return {"workload": f"{jpi}x{jpj}x{jpk} FP64 tracer step (FCT {kn[0]}/{kn[1]} with the limiter frozen, laplacian, implicit "
                        f"diffusion, leapfrog filter), {steps} steps, checkpoints every {ckpt}",
            "forward_ms_per_step": (t1 - t0) * 1e3 / steps,
            "adjoint_ms_per_step": (t2 - t1) * 1e3 / steps,
            "gradient_s": t2 - t0, "dot_product_relerr": dot_err, "directions": ndir,
            "tl_relerr_median": float(np.median(errs)) if ndir else None,
            "tl_relerr_max": float(errs.max()) if ndir else None, "check_s": t4 - t3,
            "tl_growth_median": float(np.median(growth)) if ndir else None,
            "fd_directions": len(errs_fd),
            "fd_relerr_median": float(np.median(errs_fd)) if len(errs_fd) else None,
            "fd_relerr_max": float(errs_fd.max()) if len(errs_fd) else None,
            "fd_nonlinear_directions": len(errs_nl),
            "fd_nonlinear_relerr_median": float(np.median(errs_nl)) if len(errs_nl) else None}


def restarted_opts(d):
    """the domain with a leapfrog start (neuler = 1): every step is a leapfrog step"""
    from dataclasses import replace
    d2 = replace(d, opts=dict(d.opts, neuler=1))
    d2.e3t_1d, d2.e3w_1d = d.e3t_1d, d.e3w_1d
    return d2


def time_ensemble(h, members, steps, seed, ij_lo=(256, 320), ij_hi=(768, 640), jpk=56):
    """One ensemble step of `members` domains of uneven sizes (i x j drawn
    from ij_lo .. ij_hi, jpk levels, FP32 fused step), timed four ways:
    issued call by call; replayed from one graph of those calls; as a plan
    (one batched launch per position of the members' launch sequence); and
    replayed from a graph of the plan.  Returns a dict of ms per step, the
    launches per step and the ensemble's cells."""
    import numpy as np
    import torch
    from nemo_hip import configs
    rng = np.random.default_rng(seed)
    ctxs, st, cells = [], [], 0
    for m in range(members):
        jpi = int(rng.integers(ij_lo[0], ij_hi[0] + 1)) // 8 * 8
        jpj = int(rng.integers(ij_lo[1], ij_hi[1] + 1)) // 8 * 8
        c = h.context_create()
        ctxs.append(c)
        h.context_select(c)
        d = build_case(jpi, jpj, jpk, seed + 7 * m)
        h.register(d)
        st.append(Stepper(h, d, configs.make_state_device(d, seed=seed + 7 * m + 1), True, True))
        cells += jpi * jpj * jpk
    ext = torch.cuda.ExternalStream(h.stream())

    def timed(fn):
        fn()
        h.sync()
        t0 = time.perf_counter()
        for _ in range(steps):
            fn()
        h.sync()
        return (time.perf_counter() - t0) * 1e3 / steps

    out = {"members": members, "cells": cells, "jpk": jpk, "ij_range": [list(ij_lo), list(ij_hi)]}
    try:
        with torch.cuda.stream(ext):
            for kt in (1, 2, 3):
                for c, s_ in zip(ctxs, st):
                    h.context_select(c)
                    s_.step(kt)
            h.sync()
            kt = [4]

            def eager():
                for c, s_ in zip(ctxs, st):
                    h.context_select(c)
                    s_.step(kt[0])
                kt[0] += 1
            out["eager_ms_per_step"] = timed(eager)
            h.graph_begin()
            eager()
            g = h.graph_end()
            out["graph_ms_per_step"] = timed(lambda: h.graph_launch(g))
            h.graph_free(g)
            h.plan_begin()
            eager()
            p = h.plan_end()
            info = h.plan_info(p)
            out["plan"] = info
            out["plan_ms_per_step"] = timed(lambda: h.plan_launch(p))
            h.graph_begin()
            h.plan_launch(p)
            g = h.graph_end()
            out["plan_graph_ms_per_step"] = timed(lambda: h.graph_launch(g))
            h.graph_free(g)
            h.plan_free(p)
    finally:
        h.context_select(0)
        del st
        torch.cuda.empty_cache()
        for c in ctxs:
            h.context_destroy(c)
    best = min(out[k] for k in ("graph_ms_per_step", "plan_ms_per_step", "plan_graph_ms_per_step"))
    out["cell_updates_per_s"] = cells / (best * 1e-3)
    return out


# the physics step's routines beside the hot path (their timer names)
PHYSICS_ROUTINES = ["rab", "bn2", "zdf_mxl", "eos", "ldf_slp", "dyn_hpg", "dyn_ldf", "tra_sbc"]


def kernel_times(h):
    """{name: (total_ms, count)} of every timer the library accumulated."""
    import ctypes as C
    out = {}
    names = ["dyn_adv", "dyn_vor", "dyn_spg_ts", "dyn_zdf", "tra_adv_trp", "tra_adv_fct", "tra_ldf_iso", "tra_zdf",
             "lbc_lnk", "step_dyn", "step_tra", "div_hor", "ssh_nxt", "wzv", "ssh_swp", "tra_nxt", "dyn_nxt"] + \
        PHYSICS_ROUTINES
    kern = {
        "dyn_adv": ["k_dyn_keg_zad"], "dyn_vor": ["k_dyn_vor"], "dyn_zdf": ["k_dyn_zdf"],
        "dyn_spg_ts": ["k_spg_cor1", "k_spg_cor2", "k_spg_init", "k_spg_p1a", "k_spg_p1b", "k_spg_s1", "k_spg_s2", "k_spg_s3",
                       "k_spg_loop", "k_spg_p3a", "k_spg_p3b"],
        "tra_adv_trp": ["k_tra_adv_trp"],
        "tra_adv_fct": ["k_fct_fused", "k_fct_cpt_fwd", "k_fct_h", "k_fct_hh", "k_fct_v", "k_fct_vz_reg", "k_fct_lim", "k_fct_vlim",
                        "k_fct_beta", "k_fct_final", "k_fct_low", "k_fct_ad"],
        "tra_ldf_iso": ["k_ldf_iso_coef", "k_tra_ldf_iso", "k_fill3", "k_copy3"], "tra_zdf": ["k_tra_zdf", "k_tra_zdf_reg", "k_clamp_ring"],
        "lbc_lnk": ["k_lbc"],
        "step_dyn": ["k_step_dyn1", "k_step_dyn2"],
        "step_tra": ["k_fct_fused", "k_fct_cpt_fwd", "k_fct_h", "k_fct_hh", "k_fct_v", "k_fct_vz_reg", "k_fct_lim", "k_fct_vlim",
                     "k_fct_beta", "k_fct_final", "k_zero_ring", "k_tra_zdf", "k_tra_zdf_reg"],
        "div_hor": ["k_div_hor"], "ssh_nxt": ["k_ssh_nxt"], "wzv": ["k_wzv"], "ssh_swp": ["k_ssh_swp"],
        "tra_nxt": ["k_tra_nxt"], "dyn_nxt": ["k_dyn_nxt1", "k_dyn_nxt2"],
        "ldf_slp": ["k_slp_mxl", "k_slp_uv_raw", "k_slp_uv_filt", "k_slp_w_raw", "k_slp_w_filt"],
    }
    for n in names:
        out[n] = h.timing_get(n)
        for k in kern.get(n, []):
            key = f"{n}/{k}"
            out[key] = h.timing_get(key)
    return out


def cpu_info():
    """CPU model and the number of cores this process may use."""
    model = None
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    try:
        cores = len(os.sched_getaffinity(0))
    except AttributeError:
        cores = os.cpu_count()
    return model, cores, os.cpu_count()


def _cpu_worker(args):
    """One all-cores baseline process: the -O3 oracle stepping its own
    sub-domain (NEMO runs one MPI rank per core; the halo exchanges it would
    add are left out, in the CPU's favour)."""
    sample, steps, seed = args
    os.environ["NEMO_ORACLE_O3"] = "1"
    for p in (ROOT, os.path.join(ROOT, "nemo-fmi-devel_amd")):
        if p not in sys.path:
            sys.path.insert(0, p)
    return cpu_baseline(sample, steps, seed, timed_only=True)


def cpu_allcores(sample, steps, seed, workers):
    """Aggregate cells/s of `workers` processes, each the -O3 oracle on its
    own sub-domain of the same physics, started together."""
    import multiprocessing as mp
    ctx = mp.get_context("spawn")
    with ctx.Pool(workers) as pool:
        times = pool.map(_cpu_worker, [(sample, steps, seed + 100 * w) for w in range(workers)])
    jpi, jpj, jpk = [int(x) for x in sample.split("x")]
    cells = jpi * jpj * jpk * steps * workers
    return cells / max(times), max(times)


def cpu_baseline(sample, steps, seed, timed_only=False):
    """The oracle (one core, no FMA) on a bounded sub-domain of the same
    physics: cells/s of the same step sequence (timed_only: the seconds)."""
    import numpy as np
    from nemo_hip import configs
    from oracle import oracle as ora
    jpi, jpj, jpk = [int(x) for x in sample.split("x")]
    d = build_case(jpi, jpj, jpk, seed)
    S = configs.make_state(d, seed=seed + 1)
    o = ora.Oracle(d)
    s3 = (jpi, jpj, jpk)
    F = {k: S[k].copy(order="F") for k in ("un_b", "vn_b", "ub_b", "vb_b", "sshn", "hu_n", "hv_n", "r1_hu_n",
                                           "r1_hv_n", "ht_n", "rCdU_bot", "utau", "vtau", "emp", "rnf", "ub2_b",
                                           "vb2_b", "un_bf", "vn_bf")}
    for k in ("ua_b", "va_b", "ssha", "un_adv", "vn_adv"):
        F[k] = np.zeros(s3[:2], order="F")
    un, vn = S["un"].copy(order="F"), S["vn"].copy(order="F")
    z = [np.zeros(s3, order="F") for _ in range(3)]
    ah, akz = np.zeros(s3, order="F"), np.zeros(s3, order="F")
    t0 = time.perf_counter()
    for kt in range(2, 2 + steps):
        ua, va = np.zeros(s3, order="F"), np.zeros(s3, order="F")
        tsa = np.zeros(s3 + (2,), order="F")
        o.dyn_adv(kt, S["un"], S["vn"], S["wn"], ua, va)
        o.dyn_vor(kt, S["un"], S["vn"], ua, va)
        F.update(ua=ua, va=va, un=un, vn=vn)
        o.dyn_spg_ts(kt, F)
        o.dyn_zdf(kt, S["ub"], S["vb"], ua, va, S["avm"], S["rCdU_bot"], F["ua_b"], F["va_b"], S["utau"],
                  S["utau_b"], S["vtau"], S["vtau_b"])
        o.tra_adv_trp(S["un"], S["vn"], S["wn"], *z)
        o.tra_adv_fct(kt, 1, "TRA", 1200.0, *z, S["tsb"], S["tsn"], tsa, 2, 4, 4)
        o.tra_ldf_iso(kt, 1, "TRA", S["ahtu"], S["ahtv"], S["tsb"], S["tsb"], tsa, 2, 1, S["uslp"], S["vslp"],
                      S["wslpi"], S["wslpj"], ah, akz)
        o.tra_zdf(kt, S["tsb"], tsa, S["avt"], S["avs"], 1, ah, akz)
    t = time.perf_counter() - t0
    if timed_only:
        return t
    cells = jpi * jpj * jpk
    model, cores, ncpu = cpu_info()
    return {"value": cells * steps / t, "unit": "grid-cell updates/s", "cores": 1, "kind": "port",
            "sample": f"{jpi}x{jpj}x{jpk} sub-domain of the same seeded physics, {steps} steps, "
                      f"{t:.1f} s, plain C oracle -O2 -ffp-contract=off, 1 thread",
            "cpu_model": model, "cores_available": cores, "nproc": ncpu}


TRAFFIC_TABLE = os.path.join(ROOT, "profiles", "hbm_traffic.md")


def traffic_per_launch(kname, form, jpi, jpj, jpk):
    """HBM bytes per launch of a kernel, from the table of calibrated
    rocprofv3 FETCH_SIZE / WRITE_SIZE counts (one run per counter, each scaled
    by the same counter's reading for a copy kernel of known size); None when
    the kernel, form and domain have no row."""
    if not os.path.exists(TRAFFIC_TABLE):
        return None
    kernel = kname.split("/")[-1]
    want = f"{jpi}x{jpj}x{jpk}"
    for line in open(TRAFFIC_TABLE):
        cells = [c.strip() for c in line.strip().strip("|").split("|")]
        if len(cells) >= 8 and cells[0] == form and cells[1] == want and cells[2] == kernel:
            try:
                return int(float(cells[7]))
            except ValueError:
                return None
    return None


def stream_copy(mib=2048, reps=10):
    """Device-to-device copy bandwidth (read + write bytes / time) of torch's
    copy kernel over `mib` MiB: the measured ceiling a streaming kernel meets."""
    import torch
    n = mib * (1 << 20) // 4
    a = torch.ones(n, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        b.copy_(a)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / reps
    del a, b
    torch.cuda.empty_cache()
    return 2 * n * 4 / (ms * 1e-3) / 1e9


def stream_copy_hand(h, mib=2048, reps=10):
    """The same copy with the library's hand-written streaming kernel
    (16-byte loads and stores, grid-stride), best over a few grid sizes;
    timed with HIP events on the library stream."""
    import torch
    n = mib * (1 << 20) // 4
    a = torch.ones(n, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    ext = torch.cuda.ExternalStream(h.stream())
    best = 0.0
    for blocks in (1024, 2048, 4096, 8192):
        with torch.cuda.stream(ext):
            for _ in range(2):
                h.stream_copy(a, b, n * 4, blocks)
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record(ext)
            for _ in range(reps):
                h.stream_copy(a, b, n * 4, blocks)
            e1.record(ext)
            e1.synchronize()
        best = max(best, 2 * n * 4 / (e0.elapsed_time(e1) / reps * 1e-3) / 1e9)
    assert torch.equal(a, b)
    del a, b
    torch.cuda.empty_cache()
    return best


def copy_calibration(mib):
    """A plain copy kernel of known size, for scaling the HBM byte counters."""
    import torch
    n = mib * (1 << 20) // 4
    src = torch.rand(n, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_calibration_bytes_read": n * 4, "copy_calibration_bytes_written": n * 4, "launches": 3}))


def main():
    a = parse()
    if a.copy_calibration:
        copy_calibration(a.copy_calibration)
        return
    if a.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {a.gpus} is not supported "
              f"(the path shards as NEMO's (i,j) domain decomposition, see DESIGN.md)", file=sys.stderr)
        sys.exit(2)
    import torch
    import nemo_hip
    from nemo_hip import configs
    jpi, jpj, jpk = [int(x) for x in a.size.split("x")]
    cells = jpi * jpj * jpk
    t0 = time.perf_counter()
    d = build_case(jpi, jpj, jpk, a.seed)
    log(f"domain {jpi}x{jpj}x{jpk} built in {time.perf_counter() - t0:.1f} s, wet fraction "
        f"{float((d.arrays['k_top'] > 0).mean()):.3f}")
    h = nemo_hip.get(0, nemo_hip.NEMO_HIP_PTR_DEVICE)
    from nemo_hip import autotune
    defaults = {k: h.tune_get(k) for k in autotune.SWITCHES}
    overrides = {}
    for kv in [x for x in a.tune.split(",") if x]:
        key, val = kv.split("=")
        overrides[key.strip()] = int(val)
    tuning = {}

    def set_tuning(form):
        """the library defaults, the setting saved for the form's workload,
        then the --tune overrides"""
        saved = autotune.load(a.autotune or a.tuning, f"{form} {a.size}") or {}
        setting = {**defaults, **saved, **overrides}
        autotune.apply(h, setting)
        tuning[form] = setting
        if saved or overrides:
            log(f"tuning of {form}: {setting}")
    h.register(d)
    t0 = time.perf_counter()
    S = configs.make_state_device(d, seed=a.seed + 1)
    torch.cuda.synchronize()
    log(f"state generated on the device in {time.perf_counter() - t0:.1f} s")
    res = {}
    forms = a.forms.split(",")
    head = next((f for f in ("fp32_fused", "fp32") if f in forms), forms[0])
    live = None
    if a.autotune:
        for form in forms:
            autotune_form(h, d, S, form, a, log)
    for form in forms:
        sp = form.startswith("fp32")
        fused = form.endswith("_fused")
        set_tuning(form)
        ms, finite, kt, lv = time_form(h, d, S, sp, fused, a.steps, a.warmup, a.timing_steps, live=(form == head))
        if form == head:
            live = lv
        res[form] = {"ms_per_step": ms, "cell_updates_per_s": cells / (ms * 1e-3), "finite": finite,
                     "routines_ms_per_step": {k: v[0] / a.timing_steps for k, v in kt.items()
                                              if "/" not in k and v[1] > 0},
                     "kernels_ms_per_step": {k: v[0] / a.timing_steps for k, v in kt.items()
                                             if "/" in k and v[1] > 0},
                     "kernel_launches_per_step": {k: v[1] // a.timing_steps for k, v in kt.items() if v[1] > 0}}
        width = 4 if sp else 8
        alg = sum(ROUTINE_ARRAYS.values()) * width * cells
        res[form]["algorithmic_bytes_per_step"] = alg
        res[form]["hbm_frac_step"] = alg / (ms * 1e-3) / HBM_PEAK
        log(f"{form}: {ms:.2f} ms/step, {cells / (ms * 1e-3):.3e} cell-updates/s, "
            f"HBM fraction of step {res[form]['hbm_frac_step']:.3f}, finite={finite}")
        for k, v in sorted(res[form]["routines_ms_per_step"].items(), key=lambda x: -x[1]):
            log(f"   {k:14s} {v:8.3f} ms")
    whole = {}
    for form in [f for f in a.whole_step.split(",") if f]:
        set_tuning(form)
        ms, finite, kt, _ = time_form(h, d, S, form.startswith("fp32"), True, a.steps, a.warmup, a.timing_steps,
                                      whole=True)
        closure = ("div_hor", "ssh_nxt", "wzv", "ssh_swp", "tra_nxt", "dyn_nxt")
        whole[form] = {"ms_per_step": ms, "cell_updates_per_s": cells / (ms * 1e-3), "finite": finite,
                       "closure_ms_per_step": {k: kt[k][0] / a.timing_steps for k in closure if kt[k][1] > 0},
                       "closure_kernels_ms_per_step": {k: v[0] / a.timing_steps for k, v in kt.items()
                                                       if "/" in k and k.split("/")[0] in closure and v[1] > 0}}
        log(f"whole step {form}: {ms:.2f} ms/step, {cells / (ms * 1e-3):.3e} cell-updates/s, finite={finite}, "
            f"closure {sum(whole[form]['closure_ms_per_step'].values()):.3f} ms")
    autotune.apply(h, dict(defaults, **overrides))   # the remaining workloads run with the library defaults
    physics = {}
    for form in [f for f in a.physics_step.split(",") if f]:
        ms, finite, kt, _ = time_form(h, d, S, form.startswith("fp32"), False, a.steps, a.warmup, a.timing_steps,
                                      physics=True)
        phys = {k: kt[k][0] / a.timing_steps for k in PHYSICS_ROUTINES if kt[k][1] > 0}
        physics[form] = {"ms_per_step": ms, "cell_updates_per_s": cells / (ms * 1e-3), "finite": finite,
                         "physics_ms_per_step": phys,
                         "ldf_slp_kernels_ms_per_step": {k: v[0] / a.timing_steps for k, v in kt.items()
                                                         if k.startswith("ldf_slp/") and v[1] > 0}}
        log(f"physics step {form}: {ms:.2f} ms/step, {cells / (ms * 1e-3):.3e} cell-updates/s, finite={finite}, "
            f"physics routines {sum(phys.values()):.3f} ms: " +
            ", ".join(f"{k} {v:.3f}" for k, v in sorted(phys.items(), key=lambda x: -x[1])))
    # roofline of the dominant kernel of the headline form: its launches timed
    # with HIP events inside the timed region
    width = 4 if head.startswith("fp32") else 8
    roof = None
    if live and live["launches"] > 0:
        kname = live["kernel"]
        avg_ms = live["total_ms"] / live["launches"]
        alg_bytes = KERNEL_BYTES[kname.split("/")[1]](width, jpi, jpj, jpk, head)
        roof = {"kernel": kname, "bound": "hbm", "achieved": alg_bytes / (avg_ms * 1e-3) / 1e9,
                "peak": HBM_PEAK / 1e9, "unit": "GB/s", "frac": alg_bytes / (avg_ms * 1e-3) / HBM_PEAK,
                "traffic": traffic_per_launch(kname, head, jpi, jpj, jpk), "avg_launch_ms": avg_ms,
                "launches_timed": live["launches"], "algorithmic_bytes_per_launch": alg_bytes}
        roof["traffic_over_algorithmic"] = roof["traffic"] / alg_bytes if roof["traffic"] else None
    host_ms = None
    if a.host_mode:
        host_ms = time_host_mode(h, d, S, a.host_mode)
        log(f"host-pointer mode (FP64 routines, PCIe-inclusive): {host_ms:.1f} ms/step, "
            f"{cells / (host_ms * 1e-3):.3e} cell-updates/s")
    ens = None
    adj = None
    if a.adjoint:
        adj = time_adjoint(h, a.adjoint, a.adjoint_steps, 10, a.adjoint_dirs, a.seed + 21)
        log(f"adjoint {adj['workload']}: forward {adj['forward_ms_per_step']:.2f} ms/step, adjoint sweep "
            f"{adj['adjoint_ms_per_step']:.2f} ms/step, gradient {adj['gradient_s']:.2f} s; dot product "
            f"{adj['dot_product_relerr']:.1e}; against the tangent linear over {adj['directions']} directions: "
            f"median {adj['tl_relerr_median']:.2e}, max {adj['tl_relerr_max']:.2e} (growth of a perturbation "
            f"{adj['tl_growth_median']:.2e}); against FD of the frozen-limiter model over {adj['fd_directions']}: "
            f"median {adj['fd_relerr_median']:.2e}, max {adj['fd_relerr_max']:.2e}; against FD of the nonlinear "
            f"model over {adj['fd_nonlinear_directions']}: median {adj['fd_nonlinear_relerr_median']:.2e} "
            f"({adj['check_s']:.1f} s)")
    if a.ensemble:
        ens = {}
        for name, kw in (("regional", {}), ("small", dict(ij_lo=(48, 48), ij_hi=(128, 128)))):
            e = time_ensemble(h, a.ensemble, 5, a.seed + 11, **kw)
            ens[name] = e
            pi = e["plan"]
            log(f"ensemble of {a.ensemble} {name} members (FP32 fused, {e['cells']:.3e} cells): "
                f"{e['eager_ms_per_step']:.1f} ms/step call by call, {e['graph_ms_per_step']:.1f} from one graph, "
                f"{e['plan_ms_per_step']:.1f} as a plan ({pi['launches']} launches per step, {pi['batched']} of "
                f"{pi['ops']} positions batched over {pi['groups']} group(s)), {e['plan_graph_ms_per_step']:.1f} "
                f"from a graph of the plan; {e['cell_updates_per_s']:.3e} cell-updates/s")
    copy_gbs = stream_copy()
    copy_hand = stream_copy_hand(h)
    log(f"device copy (2 GiB, read + write): torch {copy_gbs:.0f} GB/s, hand-written kernel {copy_hand:.0f} GB/s")
    kt_, kb_ = d.arrays["k_top"], d.arrays["k_bot"]
    wet_cells = int(((kb_ - kt_ + 1) * (kt_ > 0)).sum())
    cpu = None
    if not a.no_cpu:
        t0 = time.perf_counter()
        cpu = cpu_baseline(a.cpu_sample, a.cpu_steps, a.seed)
        log(f"cpu baseline {cpu['value']:.3e} cells/s ({time.perf_counter() - t0:.1f} s incl. setup)")
        nw = max(1, min(a.cpu_workers, cpu["cores_available"] or 1))
        if nw > 1:
            t0 = time.perf_counter()
            v, tmax = cpu_allcores(a.cpu_worker_sample, a.cpu_steps, a.seed, nw)
            cpu["all_cores"] = {"value": v, "unit": "grid-cell updates/s", "cores": nw, "kind": "port",
                                "sample": f"{nw} processes, each a {a.cpu_worker_sample} sub-domain of the same "
                                          f"physics for {a.cpu_steps} steps, plain C oracle -O3 "
                                          f"-ffp-contract=off, slowest process {tmax:.1f} s (one process per "
                                          f"core as NEMO's MPI ranks, halo exchanges not counted)"}
            log(f"cpu baseline, {nw} processes: {v:.3e} cells/s ({time.perf_counter() - t0:.1f} s incl. setup)")
    h_ms = res[head]["ms_per_step"]
    line = {
        "metric": METRIC,
        "value": res[head]["cell_updates_per_s"],
        "unit": "grid-cell updates/s",
        "n_gpus": 1, "steps": a.steps, "warmup": a.warmup, "ms_per_step": h_ms,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32 storage (f64 tridiagonal + barotropic)" if head.startswith("fp32") else "f64",
        "data": "synthetic: seeded land mask, lognormal bathymetry, streamfunction velocities",
        "config": {"workload": f"NEMO hot-path step (dyn_adv+dyn_vor+dyn_spg_ts+dyn_zdf+tra_adv_fct+tra_ldf_iso"
                               f"+tra_zdf), {jpi}x{jpj}x{jpk}, FCT 4/4, nn_baro 30, rdt 600 s",
                   "jpi": jpi, "jpj": jpj, "jpk": jpk, "form": head},
        "hbm_frac_step": res[head]["hbm_frac_step"],
        "wet_cells": wet_cells,
        "wet_cell_updates_per_s": wet_cells / (h_ms * 1e-3),
        "stream_copy_GBps": copy_gbs,
        "stream_copy_hand_GBps": copy_hand,
        "host_mode_fp64_ms_per_step": host_ms,
        "ensemble": ens,
        "adjoint": adj,
        "whole_step": whole,
        "physics_step": physics,
        "forms": res,
        "roofline": roof,
        "cpu_baseline": cpu,
        "tuning": tuning,
    }
    if "fp64" in res:
        for f in ("fp32", "fp32_fused", "fp64_fused"):
            if f in res:
                line[f"{f}_speedup_vs_fp64"] = res["fp64"]["ms_per_step"] / res[f]["ms_per_step"]
    print(json.dumps(line))


# Algorithmic bytes per launch of each library kernel, for the roofline of the
# dominant one and for the traffic table: every 3D array of the reference
# routine's inputs and outputs that the kernel reads or writes counts once per
# interior cell at storage width w; arrays that only pass between the kernels
# of one routine (FCT's zwi, zwx, zwy, zwz, betas and compact-scheme sweep,
# the Thomas pivots) count 0, and so do masks, 1-D e3 tables and 2D metrics;
# the barotropic sub-steps count their 2D FP64 arrays.  Arguments: w, jpi,
# jpj, jpk and the form ("fp64", "fp32", "fp64_fused", "fp32_fused").
def _b(n_arrays_interior):
    return lambda w, i, j, k, f: n_arrays_interior * w * (i - 2) * (j - 2) * (k - 1)


def _fused(f):
    return f.endswith("_fused")


KERNEL_BYTES = {
    "k_tra_zdf": _b(8),          # r avt ah_wslp2 tb*2 ta*2, w ta*2
    "k_tra_zdf_reg": _b(8),
    # r ahtu ahtv uslp vslp wslpi wslpj ah_wslp2 akz tb*2 ta*2, w ta*2
    "k_tra_ldf_iso": _b(14),
    "k_ldf_iso_coef": _b(5),     # r ahtu ahtv wslpi wslpj, w ah_wslp2
    "k_dyn_zdf": _b(7),          # r ub vb ua va avm, w ua va
    "k_dyn_vor": _b(6),          # r un vn ua va, w ua va
    "k_dyn_keg_zad": _b(7),      # r un vn wn ua va, w ua va
    # FCT, two tracers per launch: r pun pvn pwn (un vn wn in the fused step)
    # tb*2 tn*2 ta*2 (no ta read in the fused step: the trend starts at zero),
    # w ta*2
    "k_fct_fused": lambda w, i, j, k, f: (9 if _fused(f) else 11) * w * (i - 2) * (j - 2) * (k - 1),
    "k_fct_cpt_fwd": _b(2),      # r tn*2 (its sweep is a temporary)
    "k_fct_h": lambda w, i, j, k, f: (9 if _fused(f) else 11) * w * (i - 2) * (j - 2) * (k - 1),
    "k_fct_v": _b(5),            # r pwn tb*2 tn*2
    "k_fct_hh": _b(6),           # r pun pvn tb*2 tn*2
    # r pwn tb*2 tn*2 ta*2 (not in the fused step), w ta*2
    "k_fct_vz_reg": lambda w, i, j, k, f: (7 if _fused(f) else 9) * w * (i - 2) * (j - 2) * (k - 1),
    "k_fct_lim": _b(6),          # two tracers: r tb ta, w ta
    "k_fct_final": _b(2),        # r ta, w ta
    "k_fct_beta": _b(1),         # r tb
    "k_fct_ad": _b(5),           # r pun pvn pwn tb tn
    "k_fct_low": _b(7),          # r pun pvn pwn tb ta, w ta zwi
    "k_spg_p1a": _b(6),          # r ua va un vn, w ua va
    "k_spg_p3b": _b(8),          # r ua va un vn, w ua va un vn
    "k_tra_adv_trp": lambda w, i, j, k, f: 6 * w * i * j * (k - 1),   # r un vn wn, w zun zvn zwn
    # barotropic sub-step kernels (2D, FP64 in both forms; ENE scheme)
    # S3: reads zwx zwy zwz zCdU_u zCdU_v un_e vn_e zu_frc zv_frc zsshp2_e
    # ub_e(-1) vb_e(-1) hu_n hv_n r1_hu_n r1_hv_n r1_e1u r1_e2v e2u e1v r1_e2u
    # r1_e1v and the sums ua_b va_b un_adv vn_adv; writes zwx' zwy' and the
    # four sums (ua_e va_e only for the t-point scheme): 26 + 6 arrays
    "k_spg_s3": lambda w, i, j, k, f: 32 * 8 * i * j,
    # S2: reads zwx zwy zssh_frc r1_e1e2t sshn_e sshb_e sshbb_e ssha; writes
    # ssha_e ssha zsshp2_e: 8 + 3 arrays
    "k_spg_s2": lambda w, i, j, k, f: 11 * 8 * i * j,
    # fused momentum step
    "k_step_dyn1": _b(5),        # r un vn wn, w ua va
    "k_step_dyn2": _b(11),       # r ua va un vn ub vb avm, w ua va un vn
    # time-step closure
    "k_div_hor": _b(3),          # r un vn, w hdivn
    "k_ssh_nxt": lambda w, i, j, k, f: 3 * w * (i - 2) * (j - 2) * (k - 1) + 4 * 8 * i * j,   # + sshb emp_b emp ssha
    "k_wzv": _b(2),              # r hdivn, w wn
    "k_tra_nxt": _b(10),         # two tracers: r tsb tsn tsa, w tsb tsn
    "k_dyn_nxt1": _b(4),         # r ua va, w ua va
    "k_dyn_nxt2": _b(10),        # r ua va un vn ub vb, w un vn ub vb
}

if __name__ == "__main__":
    main()
