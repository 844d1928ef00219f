"""Build of the gfx950 HIP library (libnemo_hip.so) and the Fortran drop-in module.

hipcc flags:
  --offload-arch=gfx950             MI355X only
  -ffp-contract=off                 no fused multiply-add: every expression rounds
                                    like the reference Fortran expression
  -Xclang -target-feature -Xclang -packed-fp32-ops
                                    no packed FP32 instructions
  -mllvm -pragma-unroll-threshold=500000
                                    `#pragma unroll` unrolls the column solves
                                    that keep a whole column in registers
                                    fully (their arrays are then indexed by
                                    constants and stay out of scratch memory)
"""
import concurrent.futures as cf
import os
import shutil
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libnemo_hip.so")
SOURCES = ["api.hip", "dom_lbc.hip", "tra.hip", "dyn.hip", "spg.hip", "nxt.hip", "phys.hip", "slp.hip", "adj.hip", "gsum.hip", "fct_fused.hip", "fct_vlim.hip"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = ["--offload-arch=gfx950", "-O3", "-fPIC", "-std=c++17", "-ffp-contract=off",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-mllvm", "-pragma-unroll-threshold=500000",
         "-Wall", "-Wno-unused-variable", "-Wno-unused-function", "-I", CSRC]


def _compile(src):
    obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
    cmd = [HIPCC] + FLAGS + ["-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed on {src}:\n{r.stdout}\n{r.stderr}")
    return obj, r.stderr


def build_hip(verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=min(len(SOURCES), os.cpu_count() or 4)) as ex:
        results = list(ex.map(_compile, SOURCES))
    objs = [o for o, _ in results]
    if verbose:
        for _, err in results:
            if err.strip():
                print(err)
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    return LIB


def build_fortran():
    """Compile the ISO_C_BINDING drop-in module (needs amdflang)."""
    fc = shutil.which("amdflang") or shutil.which("flang")
    src = os.path.join(HERE, "fortran", "nemo_hip_mod.F90")
    if not fc or not os.path.exists(src):
        return None
    os.makedirs(BUILD, exist_ok=True)
    obj = os.path.join(BUILD, "nemo_hip_mod.o")
    r = subprocess.run([fc, "-c", "-O2", "-fPIC", src, "-o", obj, "-module-dir", BUILD],
                       capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"Fortran build failed:\n{r.stdout}\n{r.stderr}")
    return obj


def build_fortran_driver(modobj):
    """The Fortran test program tests/fortran/drive_routines: the drop-in
    module linked against libnemo_hip.so, which it finds next to itself
    through a relative run path."""
    fc = shutil.which("amdflang") or shutil.which("flang")
    root = os.path.dirname(HERE)
    src = os.path.join(root, "tests", "fortran", "drive_routines.F90")
    if not fc or not modobj or not os.path.exists(src):
        return None
    exe = os.path.join(root, "tests", "fortran", "drive_routines")
    cmd = [fc, "-O2", "-I", BUILD, src, modobj, "-o", exe, "-L", HERE, "-lnemo_hip",
           "-Wl,-rpath,$ORIGIN/../../" + os.path.basename(HERE)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"Fortran driver build failed:\n{r.stdout}\n{r.stderr}")
    return exe


def build(verbose=False):
    lib = build_hip(verbose)
    build_fortran_driver(build_fortran())
    return lib


if __name__ == "__main__":
    print(build(verbose=True))
