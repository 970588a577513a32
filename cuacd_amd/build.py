"""Build the native parts in-tree.

* ``cuacd_amd/libcuacd_hip.so`` - every ``csrc/*.hip`` translation unit,
  compiled by hipcc for gfx950 and linked into one C-ABI shared library;
* ``cuacd_amd/_gpu.abi3.so`` - the CPython extension (``csrc/module.c``)
  that exposes the C ABI to Python, linked against the library above;
* ``oracle/liboracle.so`` - the CPU oracle (test infrastructure only).

Run ``python cuacd_amd/build.py`` (or ``__graft_entry__.build()``); it does not import the package.
"""
import concurrent.futures as cf
import os
import subprocess
import sys
import sysconfig

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "cuacd_amd")
CSRC = os.path.join(PKG, "csrc")
OBJ = os.path.join(ROOT, "_build", "obj")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("CUACD_OFFLOAD_ARCH", "gfx950")

HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC",
    # packed-FP32 VALU forms stay off: the compiler must not emit v_pk_*_f32
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-munsafe-fp-atomics",
    "-Wno-unused-result",
    f"-I{os.path.join(ROOT, 'include')}", f"-I{CSRC}",
]


def _run(cmd):
    p = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if p.returncode != 0:
        raise RuntimeError(f"command failed ({p.returncode}): {' '.join(cmd)}\n{p.stdout}")
    return p.stdout


def _compile_hip(src, extra=()):
    os.makedirs(OBJ, exist_ok=True)
    obj = os.path.join(OBJ, os.path.basename(src) + ".o")
    _run([HIPCC, *HIP_FLAGS, *extra, "-c", src, "-o", obj])
    return obj


def build_library(jobs=8, verbose=True):
    srcs = sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith(".hip"))
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(_compile_hip, srcs))
    out = os.path.join(PKG, "libcuacd_hip.so")
    tmp = out + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs])
    os.replace(tmp, out)
    if verbose:
        print(f"[build] {out}")
    return out


def build_extension(verbose=True):
    inc = sysconfig.get_paths()["include"]
    out = os.path.join(PKG, "_gpu.abi3.so")
    tmp = out + ".tmp"
    _run(["gcc", "-O2", "-fPIC", "-shared", "-std=c11", "-Wall", "-Wno-unused-function",
          f"-I{inc}", f"-I{os.path.join(ROOT, 'include')}",
          os.path.join(CSRC, "module.c"), "-o", tmp,
          f"-L{PKG}", "-lcuacd_hip", "-Wl,-rpath,$ORIGIN"])
    os.replace(tmp, out)
    if verbose:
        print(f"[build] {out}")
    return out


def build_oracle(verbose=True):
    odir = os.path.join(ROOT, "oracle")
    srcs = sorted(os.path.join(odir, f) for f in os.listdir(odir) if f.endswith(".c"))
    out = os.path.join(odir, "liboracle.so")
    tmp = out + ".tmp"
    _run(["gcc", "-O2", "-fPIC", "-shared", "-std=c11", "-fopenmp", "-ffp-contract=off",
          "-Wall", "-Wno-unused-function", *srcs, "-o", tmp, "-lm"])
    os.replace(tmp, out)
    if verbose:
        print(f"[build] {out}")
    return out


def build_all(verbose=True):
    build_library(verbose=verbose)
    build_extension(verbose=verbose)
    build_oracle(verbose=verbose)


if __name__ == "__main__":
    what = sys.argv[1:] or ["all"]
    for w in what:
        {"all": build_all, "lib": build_library, "ext": build_extension, "oracle": build_oracle}[w]()
