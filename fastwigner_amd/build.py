"""Build libfastwigner.so in-tree with hipcc.

The device code (csrc/fw_device.h) is embedded into the library as a string and
compiled at run time by hipRTC for the GPU's architecture (gfx950), together
with each propagator's spliced V/K/mask expressions. ``check_device_code()``
compiles the standard configurations for gfx950 without a GPU, which is the
"does the device code build" check, and warms the code-object cache shipped
next to the library.

Usage:  python -m fastwigner_amd.build
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(ROOT, "_build")
LIB = os.path.join(HERE, "libfastwigner.so")
KCACHE = os.path.join(HERE, "_kcache")

SOURCES = ["fw_runtime.cpp", "fw_wigner.cpp", "fw_rho.cpp", "fw_batch.cpp", "fw_fno.cpp"]
EMBED = [("fw_common.h", "fw_common_h_src"), ("fw_device.h", "fw_device_h_src"), ("fw_rho.h", "fw_rho_h_src"),
         ("fw_batch.h", "fw_batch_h_src"), ("fw_fno.h", "fw_fno_h_src"), ("fw_math.h", "fw_math_h_src")]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")


def _embed(header, symbol):
    text = open(os.path.join(CSRC, header)).read()
    delim = "FWSRC"
    assert ")" + delim + '"' not in text
    out = os.path.join(BUILD, symbol + ".cpp")
    body = 'extern const char %s[];\nconst char %s[] = R"%s(%s)%s";\n' % (symbol, symbol, delim, text, delim)
    if not os.path.exists(out) or open(out).read() != body:
        with open(out, "w") as f:
            f.write(body)
    return out


def build_library(verbose=True):
    os.makedirs(BUILD, exist_ok=True)
    srcs = [os.path.join(CSRC, s) for s in SOURCES if os.path.exists(os.path.join(CSRC, s))]
    srcs += [_embed(h, sym) for h, sym in EMBED if os.path.exists(os.path.join(CSRC, h))]
    cmd = [HIPCC, "--offload-arch=gfx950", "-O2", "-std=c++17", "-fPIC", "-shared",
           "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
           "-I", os.path.join(ROOT, "include"), "-I", CSRC,
           *srcs, "-o", LIB + ".tmp", "-lhiprtc", "-lamdhip64"]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(LIB + ".tmp", LIB)
    return LIB


def check_device_code(verbose=True):
    """Compile the device code for gfx950 via hipRTC for the standard grids
    (no GPU needed); populates the in-tree code-object cache."""
    os.environ.setdefault("FASTWIGNER_CACHE_DIR", KCACHE)
    from fastwigner_amd import _lib
    lib = _lib.load(force_reload=True)
    cases = [(256, 256, 0), (1024, 1024, 0), (2048, 2048, 0), (1024, 1024, 1), (512, 512, 0)]
    for nx, npp, prec in cases:
        desc = _lib.wigner_desc(nx, 8.0, npp, 8.0, 0.01, 0.5, prec)
        ex = _lib.wigner_exprs(V="0.5*X*X+0.1*X*X*X*X", K="0.5*P*P", diff_V="X+0.4*X*X*X", diff_K="P")
        rc = lib.fw_wigner_check(desc, ex)
        if rc != 0:
            raise RuntimeError("device code failed to compile for %dx%d: %s" % (nx, npp, lib.fw_last_error().decode()))
        if verbose:
            print("gfx950 device code OK: Wigner %dx%d %s" % (nx, npp, "fp32" if prec else "fp64"), flush=True)


if __name__ == "__main__":
    build_library()
    if "--no-check" not in sys.argv:
        check_device_code()
