"""CPU checks of the C-ABI boundary: the library loads, exports every entry
point include/nemo_hip.h declares, and the ctypes mirror of the structures has
the C layout (no GPU call is made)."""
import ctypes as C

import nemo_hip
from nemo_hip import abi
from oracle import oracle


def test_header_declares_entry_points():
    names = nemo_hip.exported_symbols()
    for n in ["nemo_hip_init", "nemo_hip_domain_register", "nemo_hip_tra_adv_fct", "nemo_hip_tra_ldf_iso",
              "nemo_hip_tra_zdf_imp", "nemo_hip_dyn_vor", "nemo_hip_dyn_zdf", "nemo_hip_dyn_spg_ts",
              "nemo_hip_lbc_lnk", "nemo_hip_dom_msk"]:
        assert n in names


def test_library_exports_every_declared_symbol():
    L = nemo_hip.load_library()
    missing = [n for n in nemo_hip.exported_symbols() if not hasattr(L, n)]
    assert not missing, missing


def test_struct_layout_matches_c():
    out = (C.c_size_t * 32)()
    n = oracle.lib().ora_abi_layout(out, 32)
    got = list(out[:n])
    D, S, E, R = abi.DomainStruct, abi.SpgTsStruct, abi.EosStruct, abi.RstStruct
    want = [C.sizeof(D), D.rdt.offset, D.k_top.offset, D.e1t.offset, D.e3t_n.offset, D.e3vw_n.offset,
            C.sizeof(S), S.un_b.offset, S.vn_bf.offset, C.sizeof(E), E.rn_a0.offset, E.rn_nu.offset,
            C.sizeof(R), R.rhop.offset, R.r1_hv_n.offset, R.fraqsr_1lev.offset, D.ln_traqsr.offset,
            R.kqsr_b.offset]
    assert got == want


def test_init_without_gpu_fails_loudly():
    import torch
    if torch.cuda.is_available():
        return
    L = nemo_hip.load_library()
    assert L.nemo_hip_init(0) != 0
    buf = C.create_string_buffer(512)
    L.nemo_hip_last_error(buf, 512)
    assert b"device" in buf.value


def test_fortran_module_binds_declared_entry_points(tmp_path):
    """Every BIND(C) name of the Fortran drop-in module is an entry point the
    header declares, and the module compiles with the image's Fortran
    compiler (skipped when there is none)."""
    import os
    import re
    import shutil
    import subprocess
    import pytest
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    src = os.path.join(root, "nemo-fmi-devel_amd", "fortran", "nemo_hip_mod.F90")
    bound = set(re.findall(r'BIND\(C,\s*name="(\w+)"\)', open(src).read()))
    assert len(bound) >= 15
    missing = sorted(bound - set(nemo_hip.exported_symbols()))
    assert not missing, missing
    fc = shutil.which("amdflang") or shutil.which("flang")
    if not fc:
        pytest.skip("no Fortran compiler")
    r = subprocess.run([fc, "-c", src, "-o", str(tmp_path / "m.o"), "-module-dir", str(tmp_path)],
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
