"""Error of the FP32-storage form against the FP64 oracle along a trajectory.

Workload: the 32x22x31 GYRE-like closed basin (f = 1e-4 + 2e-11 y, flat
4500 m, tanh levels, T = 4 + 21 exp(-z/800) and S = 35 plus seeded N(0, 0.01)
noise), FCT 2/2, laplacian tracer diffusion 1000 m2/s, Kz = 1e-5 m2/s,
dyn_spg_ts with nn_baro = 30, rdt = 1800 s, Euler first step, the analytic
wind stress of the state generator.

Each step runs the hot path in step order (ua = va = 0, dyn_adv, dyn_vor,
dyn_spg_ts, dyn_zdf; tsa = 0, transports, tra_adv_fct, tra_ldf_lap, tra_zdf)
and then advances the state with the leapfrog of tra_nxt/dyn_nxt: the
Robert-Asselin filtered now level becomes the before level,
xb <- xn + atfp (xb - 2 xn + xa) with rn_atfp = 0.1 (TRA/tranxt.F90:221-235,
DYN/dynnxt.F90:199-211, fixed volume), and the after level becomes the now
level (the sea-level update is outside the path).  Tracers are carried by the
initial, continuity-consistent (u, v, w).

The FP64 trajectory is the oracle; the FP32 trajectory is the library's
*_sp entry points.  Reported per field: max |x32 - x64| / max |x64| over the
domain after the listed steps (tracers: T and S of the now level; u, v: now
velocities; ssh: the barotropic after sea level ssha of the step).

Run directly (needs a GPU):  python tests/precision_study.py [steps...]
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
for p in (ROOT, os.path.join(ROOT, "nemo-fmi-devel_amd")):
    if p not in sys.path:
        sys.path.insert(0, p)

SPG_2D = ["un_b", "vn_b", "ub_b", "vb_b", "sshn", "hu_n", "hv_n", "r1_hu_n", "r1_hv_n", "ht_n", "rCdU_bot", "utau",
          "vtau", "emp", "rnf", "ub2_b", "vb2_b", "un_bf", "vn_bf"]


# seeded streamfunction flow (0.2 m/s scale) with a 5 % divergent part and w
# from discrete continuity: vertical Courant number about 0.2 at 2 rdt (a 20 %
# divergent part reaches 0.8 near the surface, where the unfiltered
# computational mode of the FCT leapfrog is no longer damped)
UMAX, DIVERGENT = 0.2, 0.05


def setup(seed=20261202):
    from nemo_hip import configs
    d = configs.gyre_like(opts={"rdt": 1800.0, "nn_baro": 30})
    S = configs.make_state(d, seed=seed, kz_const=1e-5, aht=1000.0, umax=UMAX, divergent=DIVERGENT)
    return d, S


def r2dt(d, kt):
    o = d.opts
    return o["rdt"] if (o["neuler"] == 0 and kt == o["nit000"]) else 2.0 * o["rdt"]


def oracle_trajectory(d, S, marks):
    from oracle.oracle import Oracle
    o = Oracle(d)
    s3, s2 = (d.jpi, d.jpj, d.jpk), (d.jpi, d.jpj)
    c = lambda k: S[k].copy(order="F")
    tsb, tsn = c("tsb"), c("tsn")
    ub, vb, un, vn = c("ub"), c("vb"), c("un"), c("vn")
    F = {k: c(k) for k in SPG_2D}
    out = {}
    for kt in range(1, max(marks) + 1):
        ua, va = np.zeros(s3, order="F"), np.zeros(s3, order="F")
        for k in ("ua_b", "va_b", "ssha", "un_adv", "vn_adv"):
            F[k] = np.zeros(s2, order="F")
        o.dyn_adv(kt, un, vn, S["wn"], ua, va)
        o.dyn_vor(kt, un, vn, ua, va)
        F.update(ua=ua, va=va, un=un, vn=vn)
        o.dyn_spg_ts(kt, F)
        o.dyn_zdf(kt, ub, vb, ua, va, S["avm"], S["rCdU_bot"], F["ua_b"], F["va_b"], S["utau"], S["utau_b"],
                  S["vtau"], S["vtau_b"])
        z = [np.zeros(s3, order="F") for _ in range(3)]
        tsa = np.zeros(s3 + (2,), order="F")
        o.tra_adv_trp(S["un"], S["vn"], S["wn"], *z)
        o.tra_adv_fct(kt, 1, "TRA", r2dt(d, kt), *z, tsb, tsn, tsa, 2, 2, 2)
        o.tra_ldf_lap(kt, 1, "TRA", S["ahtu"], S["ahtv"], tsb, tsa, 2, 1)
        o.tra_zdf(kt, tsb, tsa, S["avt"], S["avs"], 0)
        atf = d.opts["rn_atfp"]
        tsb = tsn + atf * (tsb - 2.0 * tsn + tsa)
        tsn = tsa
        ub = un + atf * (ub - 2.0 * un + ua)
        vb = vn + atf * (vb - 2.0 * vn + va)
        un, vn = ua, va
        if kt in marks:
            out[kt] = {"T": tsn[..., 0].copy(), "S": tsn[..., 1].copy(), "u": un.copy(), "v": vn.copy(),
                       "ssh": F["ssha"].copy()}
    return out


def gpu_trajectory(d, S, marks):
    import torch
    import nemo_hip
    h = nemo_hip.get(0, nemo_hip.NEMO_HIP_PTR_DEVICE)
    h.register(d)
    s3 = (d.jpi, d.jpj, d.jpk)
    n2, n3 = d.jpi * d.jpj, d.jpi * d.jpj * d.jpk
    f32 = lambda k: torch.from_numpy(np.ascontiguousarray(S[k].ravel(order="F"))).to("cuda", torch.float32)
    f64 = lambda k: torch.from_numpy(np.ascontiguousarray(S[k].ravel(order="F"))).to("cuda", torch.float64)
    tsb, tsn = f32("tsb"), f32("tsn")
    ub, vb, un, vn = f32("ub"), f32("vb"), f32("un"), f32("vn")
    fix = {k: f32(k) for k in ("un", "vn", "wn", "avm", "avt", "avs", "ahtu", "ahtv")}
    F = {k: f64(k) for k in SPG_2D}
    two = {k: f64(k) for k in ("rCdU_bot", "utau", "utau_b", "vtau", "vtau_b")}
    z32 = lambda n: torch.zeros(n, dtype=torch.float32, device="cuda")
    out = {}
    shp = lambda t, s: t.cpu().numpy().astype(np.float64).reshape(s, order="F")
    for kt in range(1, max(marks) + 1):
        ua, va = z32(n3), z32(n3)
        for k in ("ua_b", "va_b", "ssha", "un_adv", "vn_adv"):
            F[k] = torch.zeros(n2, dtype=torch.float64, device="cuda")
        h.dyn_adv(kt, un, vn, fix["wn"], ua, va, sp=True)
        h.dyn_vor(kt, un, vn, ua, va, sp=True)
        F.update(ua=ua, va=va, un=un, vn=vn)
        h.dyn_spg_ts(kt, F, sp=True)
        h.dyn_zdf(kt, ub, vb, ua, va, fix["avm"], two["rCdU_bot"], F["ua_b"], F["va_b"], two["utau"], two["utau_b"],
                  two["vtau"], two["vtau_b"], sp=True)
        zun, zvn, zwn = z32(n3), z32(n3), z32(n3)
        tsa = z32(2 * n3)
        h.tra_adv_trp(fix["un"], fix["vn"], fix["wn"], zun, zvn, zwn, sp=True)
        h.tra_adv_fct(kt, 1, "TRA", r2dt(d, kt), zun, zvn, zwn, tsb, tsn, tsa, 2, 2, 2, sp=True)
        h.tra_ldf_lap(kt, 1, "TRA", fix["ahtu"], fix["ahtv"], tsb, tsa, 2, 1, sp=True)
        h.tra_zdf(kt, tsb, tsa, fix["avt"], fix["avs"], 0, sp=True)
        atf = d.opts["rn_atfp"]
        tsb = tsn + atf * (tsb - 2.0 * tsn + tsa)
        tsn = tsa
        ub = un + atf * (ub - 2.0 * un + ua)
        vb = vn + atf * (vb - 2.0 * vn + va)
        un, vn = ua, va
        if kt in marks:
            torch.cuda.synchronize()
            h.sync()
            t = shp(tsn, s3 + (2,))
            out[kt] = {"T": t[..., 0], "S": t[..., 1], "u": shp(un, s3), "v": shp(vn, s3),
                       "ssh": shp(F["ssha"], s3[:2])}
    return out


def errors(marks=(1, 10, 100)):
    d, S = setup()
    a = oracle_trajectory(d, S, marks)
    b = gpu_trajectory(d, S, marks)
    res = {}
    for kt in marks:
        res[kt] = {}
        for f in ("T", "S", "u", "v", "ssh"):
            ref = a[kt][f]
            res[kt][f] = float(np.abs(b[kt][f] - ref).max() / max(np.abs(ref).max(), 1e-300))
    return res


if __name__ == "__main__":
    marks = tuple(int(x) for x in sys.argv[1:]) or (1, 10, 100)
    print(json.dumps(errors(marks), indent=1))
