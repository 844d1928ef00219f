"""Synthetic NEMO configurations: grids, masks, vertical levels and fields.

Arrays are numpy float64 / int32 in Fortran order with NEMO's shapes
(jpi,jpj[,jpk[,kjpt]]), so their raw memory is exactly what a Fortran caller
would pass through the C-ABI.

Configurations (seeds are numpy PCG64 seeds):
  * ``gyre_like``: closed beta-plane basin, flat bottom, tanh-stretched z-levels
    (the 32x22x31 CPU-scale case);
  * ``basin``: seeded land mask (Gaussian-smoothed noise thresholded to a land
    fraction, largest connected ocean component kept) with lognormal
    bathymetry converted to full-step bottom levels (the 1024x1024x75 case and
    the small parity cases);
and ``make_state`` draws the prognostic fields the hot path consumes.
"""
from dataclasses import dataclass, field
import ctypes as C

import numpy as np

from . import abi

RAU0 = 1026.0          # TRA/eosbn2.F90:1244
GRAV = 9.80665         # DOM/phycst.F90:38


def fzeros(shape):
    return np.zeros(shape, dtype=np.float64, order="F")


@dataclass
class Domain:
    jpi: int
    jpj: int
    jpk: int
    jperio: int = 0
    opts: dict = field(default_factory=dict)
    arrays: dict = field(default_factory=dict)

    # ------------------------------------------------------------------
    def opt(self, name):
        return self.opts[name]

    def struct(self):
        """The nemo_hip_domain_t for this domain (arrays stay owned by self)."""
        s = abi.DomainStruct()
        s.jpi, s.jpj, s.jpk, s.jperio = self.jpi, self.jpj, self.jpk, self.jperio
        for n in abi.DOMAIN_INT_FIELDS[4:]:
            setattr(s, n, int(self.opts.get(n, 0)))
        for n in abi.DOMAIN_REAL_FIELDS:
            setattr(s, n, float(self.opts[n]))
        for n in abi.DOMAIN_KIDX_FIELDS:
            a = self.arrays[n]
            assert a.dtype == np.int32 and a.flags.f_contiguous
            setattr(s, n, a.ctypes.data_as(C.POINTER(C.c_int32)))
        for n in abi.DOMAIN_2D_FIELDS + abi.DOMAIN_3D_FIELDS:
            a = self.arrays[n]
            assert a.dtype == np.float64 and a.flags.f_contiguous, n
            setattr(s, n, a.ctypes.data_as(C.POINTER(C.c_double)))
        self._struct = s          # keep alive
        return s

    # masks derived exactly as DOM/dommsk.F90 does, for the generators below
    def tmask(self):
        kt, kb = self.arrays["k_top"], self.arrays["k_bot"]
        k = np.arange(1, self.jpk + 1)[None, None, :]
        m = (kt[:, :, None] != 0) & (k >= kt[:, :, None]) & (k <= kb[:, :, None])
        return np.asfortranarray(m.astype(np.float64))


DEFAULT_OPTS = dict(
    ln_linssh=1, ln_isfcav=0, ln_zps=0, ln_sco=0,
    ln_traldf_lap=1, ln_traldf_msc=0, ln_zdfddm=0, ln_zad_Aimp=0,
    n_dynadv=abi.NP_VEC_C2, nn_dynkeg=0, nvor_scheme=abi.NP_ENE, nn_een_e3f=1,
    ln_drgimp=1, ln_dynspg_ts=1, ln_bt_fw=1, ln_bt_av=1, nn_bt_flt=1, nn_baro=30,
    neuler=0, nit000=1, ln_traqsr=0,
    rdt=1800.0, rn_atfp=0.1, rau0=RAU0, grav=GRAV, rn_bt_alpha=0.0, rn_scal_load=0.0,
)


def tanh_levels(jpk, depth, stretch=3.0):
    """w-level depths gdepw(1..jpk) of a tanh-stretched z-coordinate (finer near the surface)."""
    s = np.linspace(0.0, 1.0, jpk)
    g = 1.0 - np.tanh(stretch * (1.0 - s)) / np.tanh(stretch)
    gdepw = depth * g
    e3t = np.empty(jpk)
    e3t[:-1] = np.diff(gdepw)
    e3t[-1] = e3t[-2]
    gdept = np.empty(jpk)
    gdept[:-1] = 0.5 * (gdepw[:-1] + gdepw[1:])
    gdept[-1] = gdepw[-1] + 0.5 * e3t[-1]
    e3w = np.empty(jpk)
    e3w[0] = gdept[0]
    e3w[1:] = np.diff(gdept)
    return gdepw, gdept, e3t, e3w


def _finish_domain(jpi, jpj, jpk, jperio, e1, e2, ff_f, ff_t, k_top, k_bot, e3t_1d, e3w_1d, opts):
    """Horizontal metrics as DOM/domhgr.F90:133-150 derives them, and the z-coordinate
    vertical scale factors (every point type shares the level's thickness)."""
    o = dict(DEFAULT_OPTS)
    o.update(opts or {})
    A = {}
    for n in ("e1t", "e1u", "e1v", "e1f"):
        A[n] = np.asfortranarray(np.broadcast_to(e1, (jpi, jpj)).astype(np.float64))
    for n in ("e2t", "e2u", "e2v", "e2f"):
        A[n] = np.asfortranarray(np.broadcast_to(e2, (jpi, jpj)).astype(np.float64))
    A["r1_e1u"] = 1.0 / A["e1u"]
    A["r1_e2u"] = 1.0 / A["e2u"]
    A["r1_e1v"] = 1.0 / A["e1v"]
    A["r1_e2v"] = 1.0 / A["e2v"]
    A["e1e2t"] = A["e1t"] * A["e2t"]
    A["r1_e1e2t"] = 1.0 / A["e1e2t"]
    A["e1e2f"] = A["e1f"] * A["e2f"]
    A["r1_e1e2f"] = 1.0 / A["e1e2f"]
    A["e1e2u"] = A["e1u"] * A["e2u"]
    A["e1e2v"] = A["e1v"] * A["e2v"]
    A["r1_e1e2u"] = 1.0 / A["e1e2u"]
    A["r1_e1e2v"] = 1.0 / A["e1e2v"]
    A["e2_e1u"] = A["e2u"] / A["e1u"]
    A["e1_e2v"] = A["e1v"] / A["e2v"]
    A["ff_f"] = np.asfortranarray(ff_f.astype(np.float64))
    A["ff_t"] = np.asfortranarray(ff_t.astype(np.float64))
    for n in A:
        A[n] = np.asfortranarray(A[n])
    A["k_top"] = np.asfortranarray(k_top.astype(np.int32))
    A["k_bot"] = np.asfortranarray(k_bot.astype(np.int32))
    e3t3 = np.asfortranarray(np.broadcast_to(e3t_1d[None, None, :], (jpi, jpj, jpk)).astype(np.float64))
    e3w3 = np.asfortranarray(np.broadcast_to(e3w_1d[None, None, :], (jpi, jpj, jpk)).astype(np.float64))
    for n in ("e3t_n", "e3u_n", "e3v_n", "e3f_n"):
        A[n] = e3t3.copy(order="F")
    for n in ("e3w_n", "e3uw_n", "e3vw_n"):
        A[n] = e3w3.copy(order="F")
    d = Domain(jpi, jpj, jpk, jperio, o, A)
    d.e3t_1d, d.e3w_1d = e3t_1d, e3w_1d
    return d


def _lbc_int(a, jperio, nat="T"):
    """Serial lbc on a 2D integer field (closed ring -> 0, cyclic -> copies)."""
    a = a.copy()
    iper = jperio in (1, 4, 6, 7)
    jper = jperio in (2, 7)
    if iper:
        a[0, :] = a[-2, :]
        a[-1, :] = a[1, :]
    else:
        a[0, :] = 0
        a[-1, :] = 0
    if jper:
        a[:, 0] = a[:, -2]
        a[:, -1] = a[:, 1]
    else:
        a[:, 0] = 0
        a[:, -1] = 0
    return a


def gyre_like(jpi=32, jpj=22, jpk=31, depth=4500.0, e1=106000.0, opts=None, jperio=0):
    """Closed beta-plane basin with a flat bottom: f = 1e-4 + 2e-11 * y."""
    gdepw, gdept, e3t, e3w = tanh_levels(jpk, depth)
    y = (np.arange(jpj) - 0.5) * e1
    yf = np.arange(jpj) * e1
    ff_t = np.broadcast_to(1.0e-4 + 2.0e-11 * y[None, :], (jpi, jpj))
    ff_f = np.broadcast_to(1.0e-4 + 2.0e-11 * yf[None, :], (jpi, jpj))
    k_top = np.ones((jpi, jpj), np.int32)
    k_bot = np.full((jpi, jpj), jpk - 1, np.int32)
    k_top = _lbc_int(k_top, jperio)
    k_bot = _lbc_int(k_bot, jperio)
    k_bot[k_top == 0] = 0
    return _finish_domain(jpi, jpj, jpk, jperio, e1, e1, ff_f, ff_t, k_top, k_bot, e3t, e3w, opts)


def _smooth(noise, sigma):
    """Periodic Gaussian smoothing through the FFT."""
    kx = np.fft.fftfreq(noise.shape[0])[:, None]
    ky = np.fft.fftfreq(noise.shape[1])[None, :]
    g = np.exp(-2.0 * (np.pi * sigma) ** 2 * (kx ** 2 + ky ** 2))
    return np.real(np.fft.ifft2(np.fft.fft2(noise) * g))


def basin(jpi, jpj, jpk, seed=0, land_frac=0.3, depth_max=5000.0, depth_min=50.0, e1=10000.0,
          jperio=0, opts=None, min_levels=3, sigma=None):
    """Seeded land mask and lognormal bathymetry on full-step z-levels."""
    from scipy import ndimage
    rng = np.random.default_rng(seed)
    sigma = sigma if sigma is not None else max(2.0, min(jpi, jpj) / 24.0)
    field_ = _smooth(rng.standard_normal((jpi, jpj)), sigma)
    thr = np.quantile(field_, land_frac)
    ocean = field_ > thr
    interior = np.zeros_like(ocean)
    interior[1:-1, 1:-1] = True
    ocean &= interior
    lab, nlab = ndimage.label(ocean)
    if nlab > 1:
        sizes = ndimage.sum(ocean, lab, range(1, nlab + 1))
        ocean = lab == (1 + int(np.argmax(sizes)))
    gdepw, gdept, e3t, e3w = tanh_levels(jpk, depth_max)
    bathy = np.exp(rng.normal(np.log(1500.0), 0.8, (jpi, jpj)))
    bathy = np.clip(_smooth(bathy, sigma / 2.0), depth_min, depth_max)
    # full steps: deepest t-level whose centre lies above the bottom
    kb = np.searchsorted(gdept[: jpk - 1], bathy, side="left")
    kb = np.clip(kb, min_levels, jpk - 1).astype(np.int32)
    k_top = np.where(ocean, 1, 0).astype(np.int32)
    k_bot = np.where(ocean, kb, 0).astype(np.int32)
    k_top = _lbc_int(k_top, jperio)
    k_bot = _lbc_int(k_bot, jperio)
    y = (np.arange(jpj) - 0.5) * e1
    yf = np.arange(jpj) * e1
    ff_t = np.broadcast_to(1.0e-4 + 2.0e-11 * y[None, :], (jpi, jpj))
    ff_f = np.broadcast_to(1.0e-4 + 2.0e-11 * yf[None, :], (jpi, jpj))
    o = {"rdt": 600.0}                 # barotropic Courant number < 0.9 at 10 km, 5000 m
    o.update(opts or {})
    d = _finish_domain(jpi, jpj, jpk, jperio, e1, e1, ff_f, ff_t, k_top, k_bot, e3t, e3w, o)
    d.bathy = bathy
    return d


# ----------------------------------------------------------------------
# prognostic fields
# ----------------------------------------------------------------------
def depths(d):
    """t- and w-point depths (jpi,jpj,jpk) consistent with the domain's
    vertical scale factors: gdepw(1) = 0, gdepw(k) = gdepw(k-1) + e3t(k-1);
    gdept(1) = e3w(1), gdept(k) = gdept(k-1) + e3w(k)."""
    e3t, e3w = d.arrays["e3t_n"], d.arrays["e3w_n"]
    gdepw = np.zeros_like(e3t)
    gdepw[:, :, 1:] = np.cumsum(e3t[:, :, :-1], axis=2)
    gdept = np.cumsum(e3w, axis=2)
    return np.asfortranarray(gdept), np.asfortranarray(gdepw)


def nlb10(d):
    """The shallowest w-level below ~10 m (1-based), DOM/domzgr.F90:144-145."""
    gdepw = np.concatenate([[0.0], np.cumsum(d.e3t_1d[:-1])])
    zref = 10.0 - 0.1 * np.min(d.e3w_1d)
    return int(np.argmax(gdepw > zref)) + 1


def masks(d):
    """(tmask, umask, vmask, wmask, fmask) the way DOM/dommsk.F90 builds them (used only
    to shape synthetic fields; the kernels derive their own)."""
    t = d.tmask()
    iper, jper = d.jperio in (1, 4, 6, 7), d.jperio in (2, 7)

    def lbc(a, nat):
        a = a.copy(order="F")
        if iper:
            a[0], a[-1] = a[-2].copy(), a[1].copy()
        else:
            if nat != "F":
                a[0] = 0
            a[-1] = 0
        if jper:
            a[:, 0], a[:, -1] = a[:, -2].copy(), a[:, 1].copy()
        else:
            if nat != "F":
                a[:, 0] = 0
            a[:, -1] = 0
        return a
    t = lbc(t, "T")
    u = fzeros(t.shape)
    v = fzeros(t.shape)
    f = fzeros(t.shape)
    u[:-1, :-1] = t[:-1, :-1] * t[1:, :-1]
    v[:-1, :-1] = t[:-1, :-1] * t[:-1, 1:]
    f[:-1, :-1] = t[:-1, :-1] * t[1:, :-1] * t[:-1, 1:] * t[1:, 1:]
    u, v, f = lbc(u, "U"), lbc(v, "V"), lbc(f, "F")
    w = t.copy(order="F")
    w[:, :, 1:] = t[:, :, 1:] * t[:, :, :-1]
    return t, u, v, w, f


def make_state(d, seed=1, umax=0.2, divergent=0.2, kz_range=(1e-5, 1e-1), kz_const=None,
               aht=1000.0, slope_max=0.01, sst_noise=0.01, fct_orders=(2, 2)):
    """Draw every field the hot path reads.  Velocities come from a seeded
    streamfunction (non-divergent part) plus a divergent part, and w from the
    discrete continuity equation (linear free surface, as DYN/sshwzv.F90 wzv)."""
    rng = np.random.default_rng(seed)
    jpi, jpj, jpk = d.jpi, d.jpj, d.jpk
    A = d.arrays
    t, u, v, w, f = masks(d)
    sig = max(1.5, min(jpi, jpj) / 16.0)
    gdept = np.cumsum(d.e3t_1d) - 0.5 * d.e3t_1d
    S = {}
    # tracers T, S (jpi,jpj,jpk,2)
    tn = fzeros((jpi, jpj, jpk, 2))
    prof = 4.0 + 21.0 * np.exp(-gdept / 800.0)
    tn[:, :, :, 0] = (prof[None, None, :] + sst_noise * rng.standard_normal((jpi, jpj, jpk))) * t
    tn[:, :, :, 1] = (35.0 + sst_noise * rng.standard_normal((jpi, jpj, jpk))) * t
    tb = tn + 0.01 * rng.standard_normal(tn.shape) * t[:, :, :, None]
    S["tsn"] = np.asfortranarray(tn)
    S["tsb"] = np.asfortranarray(tb)
    # velocities
    separable = jpi * jpj * jpk > 4_000_000
    if separable:
        # large grids: two smooth horizontal modes times random vertical
        # structure (the 3D fields of the small cases cost one FFT per level)
        def field3d():
            a, b = _smooth(rng.standard_normal((jpi, jpj)), sig), _smooth(rng.standard_normal((jpi, jpj)), sig)
            za, zb = rng.standard_normal(jpk), rng.standard_normal(jpk)
            return a[:, :, None] * za[None, None, :] + b[:, :, None] * zb[None, None, :]
    else:
        def field3d():
            return np.stack([_smooth(rng.standard_normal((jpi, jpj)), sig) for _ in range(jpk)], axis=2)
    psi = field3d()
    phi = field3d()
    psi *= f
    e1, e2 = A["e1u"][:, :, None], A["e2u"][:, :, None]
    un = fzeros((jpi, jpj, jpk))
    vn = fzeros((jpi, jpj, jpk))
    un[:, 1:] = -(psi[:, 1:] - psi[:, :-1])
    vn[1:, :] = (psi[1:, :] - psi[:-1, :])
    un[:-1, :] += divergent * (phi[1:, :] - phi[:-1, :])
    vn[:, :-1] += divergent * (phi[:, 1:] - phi[:, :-1])
    un *= u
    vn *= v
    sc = umax / max(1e-30, np.sqrt(np.mean(un[u > 0] ** 2)) if np.any(u > 0) else 1.0)
    un *= sc
    vn *= sc
    un = np.asfortranarray(un)
    vn = np.asfortranarray(vn)
    # w from continuity, wzv (linear free surface): wn(k) = wn(k+1) - e3t*hdiv
    e3t = A["e3t_n"]
    zu = A["e2u"][:, :, None] * A["e3u_n"] * un
    zv = A["e1v"][:, :, None] * A["e3v_n"] * vn
    hdiv = fzeros((jpi, jpj, jpk))
    hdiv[1:-1, 1:-1] = ((zu[1:-1, 1:-1] - zu[:-2, 1:-1] + zv[1:-1, 1:-1] - zv[1:-1, :-2])
                        * A["r1_e1e2t"][1:-1, 1:-1, None] / e3t[1:-1, 1:-1])
    wn = fzeros((jpi, jpj, jpk))
    for k in range(jpk - 2, -1, -1):
        wn[:, :, k] = (wn[:, :, k + 1] - e3t[:, :, k] * hdiv[:, :, k]) * t[:, :, k]
    S["un"], S["vn"], S["wn"] = un, vn, np.asfortranarray(wn)
    S["ub"] = np.asfortranarray(un * (1.0 + 0.05 * rng.standard_normal(un.shape)) * u)
    S["vb"] = np.asfortranarray(vn * (1.0 + 0.05 * rng.standard_normal(vn.shape)) * v)
    # vertical mixing coefficients at w-points (masked: no flux through top / bottom)
    if kz_const is not None:
        kz = np.full((jpi, jpj, jpk), kz_const)
        km = np.full((jpi, jpj, jpk), kz_const)
    else:
        lo, hi = np.log(kz_range[0]), np.log(kz_range[1])
        kz = np.exp(rng.uniform(lo, hi, (jpi, jpj, jpk)))
        km = np.exp(rng.uniform(lo, hi, (jpi, jpj, jpk)))
    wk = w.copy(order="F")
    wk[:, :, 0] = 0.0
    S["avt"] = np.asfortranarray(kz * wk)
    S["avs"] = S["avt"]
    S["avm"] = np.asfortranarray(km * wk)
    # lateral diffusivity at u-, v-points (masked)
    S["ahtu"] = np.asfortranarray(aht * u)
    S["ahtv"] = np.asfortranarray(aht * v)
    # iso-neutral slopes (smooth, |slope| <= slope_max), LDF/ldfslp.F90 masking
    def slope(mask):
        s = field3d()
        s = s / max(1e-30, np.abs(s).max()) * slope_max
        return np.asfortranarray(s * mask)
    S["uslp"], S["vslp"], S["wslpi"], S["wslpj"] = slope(u), slope(v), slope(w), slope(w)
    # barotropic fields
    ssm = (t.max(axis=2) > 0).astype(np.float64)
    ssum = (u.max(axis=2) > 0).astype(np.float64)
    ssvm = (v.max(axis=2) > 0).astype(np.float64)
    hu = np.asfortranarray((A["e3u_n"] * u).sum(axis=2))
    hv = np.asfortranarray((A["e3v_n"] * v).sum(axis=2))
    ht = np.asfortranarray((A["e3t_n"] * t).sum(axis=2))
    S["hu_n"], S["hv_n"], S["ht_n"] = hu, hv, ht
    S["r1_hu_n"] = np.asfortranarray(ssum / (hu + 1.0 - ssum))
    S["r1_hv_n"] = np.asfortranarray(ssvm / (hv + 1.0 - ssvm))
    S["un_b"] = np.asfortranarray((A["e3u_n"] * un * u).sum(axis=2) * S["r1_hu_n"])
    S["vn_b"] = np.asfortranarray((A["e3v_n"] * vn * v).sum(axis=2) * S["r1_hv_n"])
    S["ub_b"] = np.asfortranarray((A["e3u_n"] * S["ub"] * u).sum(axis=2) * S["r1_hu_n"])
    S["vb_b"] = np.asfortranarray((A["e3v_n"] * S["vb"] * v).sum(axis=2) * S["r1_hv_n"])
    S["sshn"] = np.asfortranarray(0.1 * _smooth(rng.standard_normal((jpi, jpj)), sig) * ssm)
    S["rCdU_bot"] = np.asfortranarray(-rng.uniform(1e-4, 1e-3, (jpi, jpj)) * ssm)
    ly = jpj * A["e2t"][0, 0]
    yy = (np.arange(jpj) - 0.5) * A["e2t"][0, 0]
    S["utau"] = np.asfortranarray(-0.1 * np.cos(2 * np.pi * yy / ly)[None, :] * u[:, :, 0])
    S["vtau"] = fzeros((jpi, jpj))
    S["utau_b"] = np.asfortranarray(S["utau"] * 0.98)
    S["vtau_b"] = fzeros((jpi, jpj))
    S["emp"] = np.asfortranarray(1e-5 * _smooth(rng.standard_normal((jpi, jpj)), sig) * ssm)
    S["rnf"] = fzeros((jpi, jpj))
    # previous-step time-averaged transports per unit width (m2/s), as
    # dyn_spg_ts stores them in un_adv / ub2_b (DYN/dynspg_ts.F90:1214-1230)
    S["ub2_b"] = np.asfortranarray(S["un_b"] * hu)
    S["vb2_b"] = np.asfortranarray(S["vn_b"] * hv)
    S["un_bf"] = np.asfortranarray(0.01 * S["ub2_b"])
    S["vn_bf"] = np.asfortranarray(0.01 * S["vb2_b"])
    # after-level barotropic velocities used by dyn_zdf (from a dyn_spg_ts call)
    S["ua_b"] = np.asfortranarray(S["un_b"] * 1.01)
    S["va_b"] = np.asfortranarray(S["vn_b"] * 1.01)
    # trends start at a small random value (earlier operators of the step)
    S["tsa"] = np.asfortranarray(1e-6 * rng.standard_normal((jpi, jpj, jpk, 2)) * t[:, :, :, None])
    S["ua"] = np.asfortranarray(1e-6 * rng.standard_normal((jpi, jpj, jpk)) * u)
    S["va"] = np.asfortranarray(1e-6 * rng.standard_normal((jpi, jpj, jpk)) * v)
    S["masks"] = (t, u, v, w, f)
    return S


# ----------------------------------------------------------------------
# the same construction built directly in HBM (benchmark sizes)
# ----------------------------------------------------------------------
def make_state_device(d, seed=1, umax=0.2, divergent=0.2, kz_range=(1e-5, 1e-1), aht=1000.0, slope_max=0.01,
                      sst_noise=0.01, device="cuda"):
    """make_state's fields generated with torch on the GPU: every 3D field is a
    tensor of shape (jpk, jpj, jpi), i.e. NEMO's Fortran layout, flattened by
    .reshape(-1).  Horizontal structure: FFT-smoothed noise; vertical
    structure: two random modes per field (as make_state on large grids).
    Unlike make_state, the streamfunction is tapered to zero near the coast
    and the divergent part bounded by a vertical Courant number of 0.5, so
    the state is within the advection schemes' stability limits at every
    size."""
    import torch
    g = torch.Generator(device=device)
    g.manual_seed(int(seed))
    jpi, jpj, jpk = d.jpi, d.jpj, d.jpk
    f64 = torch.float64
    A = d.arrays
    T2 = lambda a: torch.from_numpy(np.ascontiguousarray(np.asarray(a).T)).to(device)   # (jpj, jpi)
    sig = max(1.5, min(jpi, jpj) / 16.0)

    def randn(*shape):
        return torch.randn(*shape, generator=g, device=device, dtype=f64)

    def smooth2(noise):
        kx = torch.fft.fftfreq(jpi, device=device, dtype=f64)[None, :]
        ky = torch.fft.fftfreq(jpj, device=device, dtype=f64)[:, None]
        gk = torch.exp(-2.0 * (np.pi * sig) ** 2 * (kx ** 2 + ky ** 2))
        return torch.real(torch.fft.ifft2(torch.fft.fft2(noise) * gk))

    def field3d():
        a, b = smooth2(randn(jpj, jpi)), smooth2(randn(jpj, jpi))
        za, zb = randn(jpk), randn(jpk)
        return a[None] * za[:, None, None] + b[None] * zb[:, None, None]

    # masks, DOM/dommsk.F90 (closed or cyclic ring as in the domain's k_top/k_bot)
    kt, kb = T2(A["k_top"]).long(), T2(A["k_bot"]).long()
    k1 = torch.arange(1, jpk + 1, device=device)[:, None, None]
    t = ((kt[None] != 0) & (k1 >= kt[None]) & (k1 <= kb[None])).to(f64)
    iper, jper = d.jperio in (1, 4, 6, 7), d.jperio in (2, 7)

    def lbc(a, nat):
        a = a.clone()
        if iper:
            a[..., 0], a[..., -1] = a[..., -2].clone(), a[..., 1].clone()
        else:
            if nat != "F":
                a[..., 0] = 0
            a[..., -1] = 0
        if jper:
            a[..., 0, :], a[..., -1, :] = a[..., -2, :].clone(), a[..., 1, :].clone()
        else:
            if nat != "F":
                a[..., 0, :] = 0
            a[..., -1, :] = 0
        return a
    t = lbc(t, "T")
    u = torch.zeros_like(t)
    v = torch.zeros_like(t)
    f = torch.zeros_like(t)
    u[:, :-1, :-1] = t[:, :-1, :-1] * t[:, :-1, 1:]
    v[:, :-1, :-1] = t[:, :-1, :-1] * t[:, 1:, :-1]
    f[:, :-1, :-1] = t[:, :-1, :-1] * t[:, :-1, 1:] * t[:, 1:, :-1] * t[:, 1:, 1:]
    u, v, f = lbc(u, "U"), lbc(v, "V"), lbc(f, "F")
    w = t.clone()
    w[1:] = t[1:] * t[:-1]
    S = {}
    gdept = torch.from_numpy(np.cumsum(d.e3t_1d) - 0.5 * d.e3t_1d).to(device)
    prof = 4.0 + 21.0 * torch.exp(-gdept / 800.0)
    tn = torch.empty((2, jpk, jpj, jpi), dtype=f64, device=device)
    tn[0] = (prof[:, None, None] + sst_noise * randn(jpk, jpj, jpi)) * t
    tn[1] = (35.0 + sst_noise * randn(jpk, jpj, jpi)) * t
    S["tsn"] = tn
    S["tsb"] = tn + 0.01 * randn(2, jpk, jpj, jpi) * t[None]
    # The streamfunction is tapered to zero over r = sig/2 cells from the coast
    # (per level, the mean of the f-mask eroded by 1..r cells) rather than cut
    # there: a cut leaves jumps of tens of rms speeds along the coast.
    # Tapered, its velocities stay divergence-free to rounding, coastal cells
    # included, so w comes from the divergent part alone.
    def taper(mask, r=max(4, int(round(sig / 2.0)))):
        e, acc = 1.0 - mask, torch.zeros_like(mask)
        for _ in range(r):
            e = torch.nn.functional.max_pool2d(e[:, None], 3, stride=1, padding=1)[:, 0]
            acc += 1.0 - e
        return acc / r
    psi = field3d() * taper(f)
    phi = field3d()
    e3t = torch.from_numpy(d.e3t_1d).to(device)[:, None, None]
    e3w = torch.from_numpy(d.e3w_1d).to(device)[:, None, None]
    e2u, e1v, r1_e1e2t = T2(A["e2u"]), T2(A["e1v"]), T2(A["r1_e1e2t"])

    def velocities(a_div):
        un = torch.zeros_like(t)
        vn = torch.zeros_like(t)
        un[:, 1:, :] = -(psi[:, 1:, :] - psi[:, :-1, :])
        vn[:, :, 1:] = psi[:, :, 1:] - psi[:, :, :-1]
        un[:, :, :-1] += a_div * (phi[:, :, 1:] - phi[:, :, :-1])
        vn[:, :-1, :] += a_div * (phi[:, 1:, :] - phi[:, :-1, :])
        return un * u, vn * v

    def continuity(un, vn):
        """wn(k) = wn(k+1) - e3t*hdiv (masked), wn(jpk) = 0 (sshwzv.F90 wzv, linear free surface)"""
        zu = e2u[None] * e3t * un
        zv = e1v[None] * e3t * vn
        hdiv = torch.zeros_like(t)
        hdiv[:, 1:-1, 1:-1] = ((zu[:, 1:-1, 1:-1] - zu[:, 1:-1, :-2] + zv[:, 1:-1, 1:-1] - zv[:, :-2, 1:-1])
                               * r1_e1e2t[None, 1:-1, 1:-1] / e3t)
        inc = (e3t * hdiv * t)[:-1]
        wn = torch.zeros_like(t)
        wn[:-1] = torch.flip(torch.cumsum(torch.flip(inc, [0]), 0), [0]) * t[:-1]
        return wn

    un, vn = velocities(divergent)
    sc = umax / torch.sqrt(torch.mean(un[u > 0] ** 2)).item()
    # The divergent part is then scaled down, if need be, so that the vertical
    # Courant number 2 rdt |w| / e3w stays below 0.5 (the thin surface levels
    # of a stretched grid set it): the upstream and centred fluxes of the FCT
    # are stable only below 1.
    p2dt = 2.0 * float(d.opts.get("rdt", 600.0))
    cw = float((continuity(*velocities(divergent)).abs() * (sc * p2dt) / e3w).max())
    a_div = divergent * min(1.0, 0.5 / cw) if cw > 0 else divergent
    un, vn = velocities(a_div)
    un *= sc
    vn *= sc
    wn = continuity(un, vn)
    S["un"], S["vn"], S["wn"] = un, vn, wn
    S["ub"] = un * (1.0 + 0.05 * randn(jpk, jpj, jpi)) * u
    S["vb"] = vn * (1.0 + 0.05 * randn(jpk, jpj, jpi)) * v
    lo, hi = np.log(kz_range[0]), np.log(kz_range[1])
    wk = w.clone()
    wk[0] = 0
    unif = lambda: torch.rand(jpk, jpj, jpi, generator=g, device=device, dtype=f64)
    S["avt"] = torch.exp(lo + (hi - lo) * unif()) * wk
    S["avs"] = S["avt"]
    S["avm"] = torch.exp(lo + (hi - lo) * unif()) * wk
    S["ahtu"] = aht * u
    S["ahtv"] = aht * v

    def slope(mask):
        s = field3d()
        return s / s.abs().max() * slope_max * mask
    S["uslp"], S["vslp"], S["wslpi"], S["wslpj"] = slope(u), slope(v), slope(w), slope(w)
    ssm = (t.amax(0) > 0).to(f64)
    ssum = (u.amax(0) > 0).to(f64)
    ssvm = (v.amax(0) > 0).to(f64)
    hu = (e3t * u).sum(0)
    hv = (e3t * v).sum(0)
    S["hu_n"], S["hv_n"], S["ht_n"] = hu, hv, (e3t * t).sum(0)
    S["r1_hu_n"] = ssum / (hu + 1.0 - ssum)
    S["r1_hv_n"] = ssvm / (hv + 1.0 - ssvm)
    S["un_b"] = (e3t * un * u).sum(0) * S["r1_hu_n"]
    S["vn_b"] = (e3t * vn * v).sum(0) * S["r1_hv_n"]
    S["ub_b"] = (e3t * S["ub"] * u).sum(0) * S["r1_hu_n"]
    S["vb_b"] = (e3t * S["vb"] * v).sum(0) * S["r1_hv_n"]
    S["sshn"] = 0.1 * smooth2(randn(jpj, jpi)) * ssm
    S["rCdU_bot"] = -(1e-4 + 9e-4 * torch.rand(jpj, jpi, generator=g, device=device, dtype=f64)) * ssm
    ly = jpj * float(A["e2t"][0, 0])
    yy = (torch.arange(jpj, device=device, dtype=f64) - 0.5) * float(A["e2t"][0, 0])
    S["utau"] = -0.1 * torch.cos(2 * np.pi * yy / ly)[:, None] * u[0]
    S["vtau"] = torch.zeros_like(hu)
    S["utau_b"] = S["utau"] * 0.98
    S["vtau_b"] = torch.zeros_like(hu)
    S["emp"] = 1e-5 * smooth2(randn(jpj, jpi)) * ssm
    S["rnf"] = torch.zeros_like(hu)
    S["ub2_b"] = S["un_b"] * hu
    S["vb2_b"] = S["vn_b"] * hv
    S["un_bf"] = 0.01 * S["ub2_b"]
    S["vn_bf"] = 0.01 * S["vb2_b"]
    S["ua_b"] = S["un_b"] * 1.01
    S["va_b"] = S["vn_b"] * 1.01
    S["tmask"] = t
    return {k: (v.contiguous().reshape(-1) if k != "tmask" else v) for k, v in S.items()}
