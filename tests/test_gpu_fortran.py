"""The Fortran drop-in boundary end to end: tests/fortran/drive_routines, a
Fortran program built against nemo_hip_mod (nemo-fmi-devel_amd/fortran),
calls tra_ldf_lap, tra_zdf_imp and tra_adv_fct with the reference argument
lists on its own host arrays (host-pointer mode), as a NEMO build would.
Its results must equal the oracle's on the same domain and state, bit for
bit (bar 1e-12).  The program is built by __graft_entry__.build() when a
Fortran compiler is present; it runs here as a child process."""
import os
import subprocess

import numpy as np
import pytest

from nemo_hip import abi, configs
from oracle.oracle import Oracle
from test_gpu_parity import check

HERE = os.path.dirname(os.path.abspath(__file__))
DRIVER = os.path.join(HERE, "fortran", "drive_routines")


def write_input(path, d, S, kt, p2dt, zun, zvn, zwn):
    """The stream file drive_routines reads: header, domain arrays in
    nemo_hip_domain_t order, then the fields (all Fortran order)."""
    ih = [d.jpi, d.jpj, d.jpk, d.jperio] + [int(d.opts.get(n, 0)) for n in abi.DOMAIN_INT_FIELDS[4:]]
    rh = [float(d.opts[n]) for n in abi.DOMAIN_REAL_FIELDS]
    f = lambda a, dt=np.float64: np.asarray(a, dtype=dt).ravel(order="F").tobytes()
    with open(path, "wb") as fh:
        fh.write(np.asarray(ih, dtype=np.int32).tobytes())
        fh.write(np.asarray(rh, dtype=np.float64).tobytes())
        fh.write(f(d.arrays["k_top"], np.int32) + f(d.arrays["k_bot"], np.int32))
        for n in abi.DOMAIN_2D_FIELDS:
            fh.write(f(d.arrays[n]))
        for n in abi.DOMAIN_3D_FIELDS:
            fh.write(f(d.arrays[n]))
        for n in ("ahtu", "ahtv", "avt", "avs", "avm"):
            fh.write(f(S[n]))
        for a in (zun, zvn, zwn, S["tsb"], S["tsn"], S["tsa"]):
            fh.write(f(a))
        fh.write(np.asarray([kt], dtype=np.int32).tobytes())
        fh.write(np.asarray([p2dt], dtype=np.float64).tobytes())
        for n in ("un", "vn", "wn", "ub", "vb", "rCdU_bot", "utau", "utau_b", "vtau", "vtau_b"):
            fh.write(f(S[n]))


@pytest.mark.gpu
@pytest.mark.parametrize("case", ["basin", "cyc7"])
def test_fortran_module_drives_routines(tmp_path, case):
    assert os.path.exists(DRIVER), "drive_routines is not built (needs amdflang at build time)"
    d = (configs.basin(40, 30, 12, seed=3) if case == "basin" else configs.basin(40, 30, 12, seed=4, jperio=7))
    S = configs.make_state(d, seed=17)
    o = Oracle(d)
    s3, s4 = (d.jpi, d.jpj, d.jpk), (d.jpi, d.jpj, d.jpk, 2)
    kt, p2dt = 3, 2 * d.opts["rdt"]
    z = [np.zeros(s3, order="F") for _ in range(3)]
    o.tra_adv_trp(S["un"], S["vn"], S["wn"], *z)
    want_lap, want_zdf, want_fct = (S["tsa"].copy(order="F") for _ in range(3))
    o.tra_ldf_lap(kt, 1, "TRA", S["ahtu"], S["ahtv"], S["tsb"], want_lap, 2, 1)
    o.tra_zdf_imp(kt, 1, "TRA", p2dt, S["tsb"], want_zdf, 2, S["avt"], S["avs"])
    o.tra_adv_fct(kt, 1, "TRA", p2dt, *z, S["tsb"], S["tsn"], want_fct, 2, 4, 4)
    fin, fout = str(tmp_path / "in.bin"), str(tmp_path / "out.bin")
    write_input(fin, d, S, kt, p2dt, *z)
    print(f"running {DRIVER} on case {case}", flush=True)
    r = subprocess.run([DRIVER, fin, fout], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, f"drive_routines failed ({r.returncode}):\n{r.stdout}\n{r.stderr}"
    out = np.fromfile(fout, dtype=np.float64)
    n4, n3 = int(np.prod(s4)), int(np.prod(s3))
    assert out.size == 4 * n4 + 3 * n3
    got = [out[q * n4:(q + 1) * n4].reshape(s4, order="F") for q in range(4)]
    o3 = out[4 * n4:]
    g3 = [o3[q * n3:(q + 1) * n3].reshape(s3, order="F") for q in range(3)]
    check(got[0], want_lap, what="tra_ldf_lap via Fortran")
    check(got[1], want_zdf, what="tra_zdf_imp via Fortran")
    check(got[2], want_fct, what="tra_adv_fct via Fortran")
    # the module-global routines: tra_adv(kt), tra_ldf(kt), tra_zdf(kt) from a zero trend (r2dt = 2 rdt at kt 3)
    tsa = np.zeros(s4, order="F")
    o.tra_adv_fct(kt, 1, "TRA", p2dt, *z, S["tsb"], S["tsn"], tsa, 2, 4, 4)
    o.tra_ldf_lap(kt, 1, "TRA", S["ahtu"], S["ahtv"], S["tsb"], tsa, 2, 1)
    o.tra_zdf(kt, S["tsb"], tsa, S["avt"], S["avs"], 0, np.zeros(s3, order="F"), np.zeros(s3, order="F"))
    check(got[3], tsa, what="tra_adv + tra_ldf + tra_zdf via Fortran")
    # dyn_adv(kt), dyn_vor(kt), dyn_zdf(kt) from a zero trend, div_hor(kt)
    ua, va = np.zeros(s3, order="F"), np.zeros(s3, order="F")
    o.dyn_adv(kt, S["un"], S["vn"], S["wn"], ua, va)
    o.dyn_vor(kt, S["un"], S["vn"], ua, va)
    z2 = np.zeros((d.jpi, d.jpj), order="F")
    o.dyn_zdf(kt, S["ub"], S["vb"], ua, va, S["avm"], S["rCdU_bot"], z2, z2, S["utau"], S["utau_b"], S["vtau"],
              S["vtau_b"])
    check(g3[0], ua, what="dyn_adv + dyn_vor + dyn_zdf ua via Fortran")
    check(g3[1], va, what="dyn_adv + dyn_vor + dyn_zdf va via Fortran")
    hd = np.zeros(s3, order="F")
    o.div_hor(kt, S["un"], S["vn"], hd)
    check(g3[2], hd, what="div_hor via Fortran")
