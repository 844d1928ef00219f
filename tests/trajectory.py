"""A whole NEMO time step, step.F90:161-279, for a configuration without an
equation of state, pressure gradient, lateral viscosity or surface module
(constant forcing, the before forcing equal to the now forcing):

    ssh_nxt, wzv                      first guess of ssha, now w
    ua = va = 0, dyn_adv, dyn_vor,
    dyn_spg_ts                        barotropic sub-steps, ssha, ua_b, ...
    div_hor, wzv                      w of the corrected now velocities
    dyn_zdf                           after velocities
    tsa = 0, tra_adv (fct),
    tra_ldf_iso, tra_zdf              after tracers
    tra_nxt, dyn_nxt, ssh_swp         Asselin filters and swaps

With physics=True the step also has the iso-neutral slopes, a density and
its pressure gradient, a lateral viscosity and the surface tracer fluxes,
step.F90:127-150, 165, 191-193 and 238: rab_3d, bn2, zdf_mxl, eos_insitu
and ldf_slp of the before tracers first (the slopes tra_ldf_iso uses),
eos_insitu_pot of the now tracers after wzv (the simplified EOS unless
`eos` names another),
dyn_ldf_lap of the before velocities (ahm = AHM m2/s) and hpg_zco between
dyn_vor and dyn_spg_ts, and tra_sbc (an analytic heat flux, the generator's
fresh-water flux) before tra_adv.

oracle_run repeats it with the oracle's separate routines; gpu_run with the
library on device-resident fields (the fused step_dyn / step_tra entry points,
or with physics the separate momentum routines, plus the closure routines),
in FP64 or the FP32 storage form.  Both return the prognostic state after the
steps listed in `marks`.
"""
import numpy as np

SPG_2D = ["un_b", "vn_b", "ub_b", "vb_b", "sshn", "hu_n", "hv_n", "r1_hu_n", "r1_hv_n", "ht_n", "rCdU_bot", "utau",
          "vtau", "emp", "rnf", "ub2_b", "vb2_b", "un_bf", "vn_bf"]
FIX_3D = ["avm", "avt", "ahtu", "ahtv", "uslp", "vslp", "wslpi", "wslpj"]
OUT_3D = ["un", "vn", "ub", "vb", "wn"]
OUT_4D = ["tsn", "tsb"]
OUT_2D = ["sshn", "sshb", "un_b", "vn_b", "ub_b", "vb_b", "ub2_b", "un_bf"]
KN_FCT = (4, 4)
AHM = 2000.0
RN_SLPMAX = 0.01


def r2dt(d, kt):
    o = d.opts
    return o["rdt"] if (o["neuler"] == 0 and kt == o["nit000"]) else 2.0 * o["rdt"]


def initial(S):
    """The prognostic fields of the generator's state, with sshb = sshn and
    the before forcing equal to the now forcing."""
    I = {k: S[k] for k in SPG_2D + FIX_3D + ["un", "vn", "ub", "vb", "wn", "tsb", "tsn"]}
    I["sshb"] = S["sshn"]
    I["emp_b"], I["utau_b"], I["vtau_b"] = S["emp"], S["utau"], S["vtau"]
    return I


def surface_fluxes(d):
    """Analytic non-solar and solar heat fluxes (W/m2), salt flux 0."""
    y = np.arange(d.jpj) / max(d.jpj - 1, 1)
    qns = np.asfortranarray(np.broadcast_to((-40.0 + 80.0 * np.cos(np.pi * y))[None, :], (d.jpi, d.jpj)).copy())
    qsr = np.asfortranarray(np.full((d.jpi, d.jpj), 30.0))
    return qns, qsr, np.zeros((d.jpi, d.jpj), order="F")


def viscosity(d):
    """ahmt, ahmf of a constant lateral viscosity AHM, masked by tmask / fmask."""
    from nemo_hip import configs
    t, u, v, w, f = configs.masks(d)
    return np.asfortranarray(AHM * t), np.asfortranarray(AHM * f)


def oracle_run(d, S, marks, physics=False, eos=None):
    from oracle.oracle import Oracle
    from nemo_hip import abi, configs
    o = Oracle(d)
    if physics:
        o.eos_register(eos or abi.seos())
        gdept, gdepw = configs.depths(d)
        nlb10 = configs.nlb10(d)
        ahmt, ahmf = viscosity(d)
        qns0, qsr0, sfx = surface_fluxes(d)
        sbc, sbc_b = np.zeros((d.jpi, d.jpj, 2), order="F"), np.zeros((d.jpi, d.jpj, 2), order="F")
    s2, s3, s4 = (d.jpi, d.jpj), (d.jpi, d.jpj, d.jpk), (d.jpi, d.jpj, d.jpk, 2)
    X = {k: np.array(v, order="F", copy=True) for k, v in initial(S).items()}
    for k in ("ua_b", "va_b", "ssha", "un_adv", "vn_adv"):
        X[k] = np.zeros(s2, order="F")
    hdivn = np.zeros(s3, order="F")
    ah, akz = np.zeros(s3, order="F"), np.zeros(s3, order="F")
    out = {}
    for kt in range(1, max(marks) + 1):
        if physics:
            pab, rn2b, rhd = np.zeros(s4, order="F"), np.zeros(s3, order="F"), np.zeros(s3, order="F")
            o.rab_3d(X["tsb"], pab, gdept)
            o.bn2(X["tsb"], pab, rn2b, gdept, gdepw)
            nmln, hmlp, hmlpt = o.zdf_mxl(kt, rn2b, gdept, gdepw, nlb10)
            o.eos_insitu(X["tsb"], rhd, gdept)
            o.ldf_slp(kt, rhd, rn2b, gdept, gdepw, nmln, hmlp, hmlpt, RN_SLPMAX, X["uslp"], X["vslp"], X["wslpi"],
                      X["wslpj"])
        o.ssh_nxt(kt, X["un"], X["vn"], hdivn, X["sshb"], X["emp_b"], X["emp"], X["ssha"])
        o.wzv(kt, hdivn, X["wn"])
        if physics:
            rhd, rhop = np.zeros(s3, order="F"), np.zeros(s3, order="F")
            o.eos_insitu_pot(X["tsn"], rhd, rhop, gdept)
        ua, va = np.zeros(s3, order="F"), np.zeros(s3, order="F")
        o.dyn_adv(kt, X["un"], X["vn"], X["wn"], ua, va)
        o.dyn_vor(kt, X["un"], X["vn"], ua, va)
        if physics:
            o.dyn_ldf_lap(kt, X["ub"], X["vb"], ua, va, 1, ahmt, ahmf)
            o.hpg_zco(kt, rhd, ua, va)
        F = {k: X[k] for k in SPG_2D + ["ua_b", "va_b", "ssha", "un_adv", "vn_adv"]}
        F.update(ua=ua, va=va, un=X["un"], vn=X["vn"])
        o.dyn_spg_ts(kt, F)
        o.div_hor(kt, X["un"], X["vn"], hdivn)
        o.wzv(kt, hdivn, X["wn"])
        o.dyn_zdf(kt, X["ub"], X["vb"], ua, va, X["avm"], X["rCdU_bot"], X["ua_b"], X["va_b"], X["utau"],
                  X["utau_b"], X["vtau"], X["vtau_b"])
        z = [np.zeros(s3, order="F") for _ in range(3)]
        tsa = np.zeros(s4, order="F")
        if physics:
            qns, qsr = qns0.copy(order="F"), qsr0.copy(order="F")
            o.tra_sbc(kt, qns, qsr, sfx, X["emp"], X["tsn"], tsa, sbc, sbc_b)
        o.tra_adv_trp(X["un"], X["vn"], X["wn"], *z)
        o.tra_adv_fct(kt, 1, "TRA", r2dt(d, kt), *z, X["tsb"], X["tsn"], tsa, 2, *KN_FCT)
        o.tra_ldf_iso(kt, 1, "TRA", X["ahtu"], X["ahtv"], X["tsb"], X["tsb"], tsa, 2, 1, X["uslp"], X["vslp"],
                      X["wslpi"], X["wslpj"], ah, akz)
        o.tra_zdf(kt, X["tsb"], tsa, X["avt"], X["avt"], 1, ah, akz)
        o.tra_nxt(kt, X["tsb"], X["tsn"], tsa)
        o.dyn_nxt(kt, X["ub"], X["vb"], X["un"], X["vn"], ua, va, X["ua_b"], X["va_b"], X["un_adv"], X["vn_adv"],
                  X["ub_b"], X["vb_b"], X["un_b"], X["vn_b"], X["r1_hu_n"], X["r1_hv_n"])
        o.ssh_swp(kt, X["sshb"], X["sshn"], X["ssha"])
        if kt in marks:
            out[kt] = {k: X[k].copy(order="F") for k in OUT_3D + OUT_4D + OUT_2D}
    return out


class HotStep:
    """One member's device-resident state and the hot-path whole step of
    gpu_run without physics.  step() issues library calls only (nothing on
    torch's streams), so an ensemble of members can record one step of each
    into a HIP graph and replay it.  X0 (device tensors by field name)
    replaces the initial state where given."""

    def __init__(self, h, d, S, sp=False, X0=None):
        import torch
        self.h, self.d, self.sp = h, d, sp
        dt = torch.float32 if sp else torch.float64
        self.dt = dt
        n2, n3 = d.jpi * d.jpj, d.jpi * d.jpj * d.jpk
        ndim = {k: len(np.shape(v)) for k, v in initial(S).items()}
        dev = lambda a, t: torch.from_numpy(np.ascontiguousarray(np.asarray(a).ravel(order="F"))).to("cuda", t)
        X = {k: dev(v, dt if ndim[k] >= 3 else torch.float64) for k, v in initial(S).items()}
        for k in ("ua_b", "va_b", "ssha", "un_adv", "vn_adv"):
            X[k] = torch.zeros(n2, dtype=torch.float64, device="cuda")
        X.update(X0 or {})
        z = lambda n: torch.zeros(n, dtype=dt, device="cuda")
        self.z = z
        self.hdivn, self.ua, self.va, self.tsa, self.ah, self.akz = z(n3), z(n3), z(n3), z(2 * n3), z(n3), z(n3)
        self.F = {k: X[k] for k in SPG_2D + ["ua_b", "va_b", "ssha", "un_adv", "vn_adv"]}
        self.F.update(ua=self.ua, va=self.va, un=X["un"], vn=X["vn"])
        self.X = X
        s2, s3, s4 = (d.jpi, d.jpj), (d.jpi, d.jpj, d.jpk), (d.jpi, d.jpj, d.jpk, 2)
        self.shape = {k: (s4 if k in OUT_4D else s3 if k in OUT_3D else s2) for k in OUT_3D + OUT_4D + OUT_2D}

    def open(self, kt):
        """ssh_nxt and wzv: the first guess of ssha and the now w"""
        h, X, sp = self.h, self.X, self.sp
        h.ssh_nxt(kt, X["un"], X["vn"], self.hdivn, X["sshb"], X["emp_b"], X["emp"], X["ssha"], sp=sp)
        h.wzv(kt, self.hdivn, X["wn"], sp=sp)

    def middle(self, kt):
        """div_hor and wzv of the corrected now velocities"""
        self.h.div_hor(kt, self.X["un"], self.X["vn"], self.hdivn, sp=self.sp)
        self.h.wzv(kt, self.hdivn, self.X["wn"], sp=self.sp)

    def close(self, kt):
        """tra_nxt, dyn_nxt and ssh_swp"""
        h, X, sp = self.h, self.X, self.sp
        h.tra_nxt(kt, X["tsb"], X["tsn"], self.tsa, sp=sp)
        h.dyn_nxt(kt, X["ub"], X["vb"], X["un"], X["vn"], self.ua, self.va, X["ua_b"], X["va_b"], X["un_adv"],
                  X["vn_adv"], X["ub_b"], X["vb_b"], X["un_b"], X["vn_b"], X["r1_hu_n"], X["r1_hv_n"], sp=sp)
        h.ssh_swp(kt, X["sshb"], X["sshn"], X["ssha"])

    def step(self, kt):
        h, X, sp = self.h, self.X, self.sp
        self.open(kt)
        h.step_dyn(kt, X["ub"], X["vb"], X["wn"], X["avm"], X["utau_b"], X["vtau_b"], self.F, sp=sp)
        self.middle(kt)
        h.step_tra(kt, X["un"], X["vn"], X["wn"], X["tsb"], X["tsn"], self.tsa, *KN_FCT, X["ahtu"], X["ahtv"],
                   X["uslp"], X["vslp"], X["wslpi"], X["wslpj"], self.ah, self.akz, X["avt"], sp=sp)
        self.close(kt)

    def fields(self):
        """the prognostic fields, FP64 host arrays (call after a synchronisation)"""
        return {k: self.X[k].cpu().numpy().astype(np.float64).reshape(s, order="F") for k, s in self.shape.items()}


def gpu_run(h, d, S, marks, sp=False, physics=False, eos=None, kt0=1, X0=None, keep=None):
    """h: the library handle in device-pointer mode with d registered.
    Steps kt0 .. max(marks); X0 (device tensors by field name) replaces the
    initial state where given; `keep`, a dict, receives the final fields."""
    import torch
    from nemo_hip import abi, configs
    st = HotStep(h, d, S, sp, X0)
    X, F, z, dt = st.X, st.F, st.z, st.dt
    ua, va, tsa, ah, akz = st.ua, st.va, st.tsa, st.ah, st.akz
    n2, n3 = d.jpi * d.jpj, d.jpi * d.jpj * d.jpk
    dev = lambda a, t: torch.from_numpy(np.ascontiguousarray(np.asarray(a).ravel(order="F"))).to("cuda", t)
    if physics:
        h.eos_register(eos or abi.seos())
        gdept, gdepw = (dev(a, dt) for a in configs.depths(d))
        nlb10 = configs.nlb10(d)
        pab, rn2b = z(2 * n3), z(n3)
        nmln = torch.zeros(n2, dtype=torch.int32, device="cuda")
        hmlp, hmlpt = (torch.zeros(n2, dtype=torch.float64, device="cuda") for _ in range(2))
        ahmt, ahmf = (dev(a, dt) for a in viscosity(d))
        rhd, rhop = z(n3), z(n3)
        qns0, qsr0, sfx = (dev(a, torch.float64) for a in surface_fluxes(d))
        qns, qsr = qns0.clone(), qsr0.clone()
        sbc, sbc_b = torch.zeros(2 * n2, dtype=torch.float64, device="cuda"), \
            torch.zeros(2 * n2, dtype=torch.float64, device="cuda")
        zun, zvn, zwn = z(n3), z(n3), z(n3)
    out = {}
    # torch work (the trend reset) is queued on the library stream, in order with the library's launches
    lib_stream = torch.cuda.ExternalStream(h.stream())
    for kt in range(kt0, max(marks) + 1):
        if not physics:
            st.step(kt)
        else:
            h.rab_3d(X["tsb"], pab, gdept, sp=sp)
            h.bn2(X["tsb"], pab, rn2b, gdept, gdepw, sp=sp)
            h.zdf_mxl(kt, rn2b, gdept, gdepw, nlb10, nmln, hmlp, hmlpt, sp=sp)
            h.eos_insitu(X["tsb"], rhd, gdept, sp=sp)
            h.ldf_slp(kt, rhd, rn2b, gdept, gdepw, nmln, hmlp, hmlpt, RN_SLPMAX, X["uslp"], X["vslp"], X["wslpi"],
                      X["wslpj"], sp=sp)
            st.open(kt)
            h.eos_insitu_pot(X["tsn"], rhd, rhop, gdept, sp=sp)
            with torch.cuda.stream(lib_stream):
                ua.zero_()
                va.zero_()
            h.dyn_adv(kt, X["un"], X["vn"], X["wn"], ua, va, sp=sp)
            h.dyn_vor(kt, X["un"], X["vn"], ua, va, sp=sp)
            h.dyn_ldf_lap(kt, X["ub"], X["vb"], ua, va, 1, ahmt, ahmf, sp=sp)
            h.hpg_zco(kt, rhd, ua, va, sp=sp)
            h.dyn_spg_ts(kt, F, sp=sp)
            st.middle(kt)
            h.dyn_zdf(kt, X["ub"], X["vb"], ua, va, X["avm"], X["rCdU_bot"], X["ua_b"], X["va_b"], X["utau"],
                      X["utau_b"], X["vtau"], X["vtau_b"], sp=sp)
            with torch.cuda.stream(lib_stream):
                tsa.zero_()
                qns.copy_(qns0)
                qsr.copy_(qsr0)
            h.tra_sbc(kt, qns, qsr, sfx, X["emp"], X["tsn"], tsa, sbc, sbc_b, sp=sp)
            h.tra_adv_trp(X["un"], X["vn"], X["wn"], zun, zvn, zwn, sp=sp)
            h.tra_adv_fct(kt, 1, "TRA", r2dt(d, kt), zun, zvn, zwn, X["tsb"], X["tsn"], tsa, 2, *KN_FCT, sp=sp)
            h.tra_ldf_iso(kt, 1, "TRA", X["ahtu"], X["ahtv"], X["tsb"], X["tsb"], tsa, 2, 1, X["uslp"], X["vslp"],
                          X["wslpi"], X["wslpj"], ah, akz, sp=sp)
            h.tra_zdf(kt, X["tsb"], tsa, X["avt"], X["avt"], 1, ah, akz, sp=sp)
            st.close(kt)
        if kt in marks:
            torch.cuda.synchronize()
            h.sync()
            out[kt] = st.fields()
    if keep is not None:
        keep.update(X)
    return out
