"""Build recipe of helix_amd's native libraries (in-tree, no JIT cache).

  libhelix_hip.so  hand-written gfx950 HIP kernels + the C-ABI of include/helix_hip.h  (hipcc)
  libhelix.so      C++17 host: set-up (include/helix_setup.h) and the public API of <helix/helix.h>,
                   calling the kernels only through the C-ABI                         (g++)
  helix_example    the documented drop-in usage with a trace print, examples/legacy_spin_glass_trace.cpp (g++)
  helix_reference_example  the reference's examples/cpp/legacy_spin_glass.cpp compiled unchanged, when present

Run ``python -m helix_amd.build`` (or ``__graft_entry__.build()``).
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
INCLUDE = os.path.join(ROOT, "include")
CSRC = os.path.join(PKG, "csrc")
ARCH = "gfx950"
REFERENCE_EXAMPLE = "/root/reference/examples/cpp/legacy_spin_glass.cpp"

HIP_SOURCES = ["heom_plan.hip", "mlp.hip", "peaks.hip"]
HOST_SOURCES = ["host/heom_setup.cpp", "host/setup_capi.cpp", "host/helix_api.cpp", "host/helix_io.cpp"]

HIPCC_FLAGS = [
    "-O3", "-std=c++17", f"--offload-arch={ARCH}", "-fPIC", "-shared",
    # packed FP32 instructions off for every HIP translation unit
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-command-line-argument",
]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + r.stdout + r.stderr)
        raise RuntimeError(f"build step failed: {cmd[0]} -> {cmd[-1]}")
    noise = [ln for ln in (r.stdout + r.stderr).splitlines() if ln.strip() and "packed-fp32-ops" not in ln]
    if noise:
        sys.stderr.write("\n".join(noise) + "\n")


def hipcc() -> str:
    for c in (shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found")


def build(verbose: bool = False) -> dict:
    out = {}
    hip_lib = os.path.join(PKG, "libhelix_hip.so")
    srcs = [os.path.join(CSRC, s) for s in HIP_SOURCES]
    _run([hipcc(), *HIPCC_FLAGS, f"-I{INCLUDE}", f"-I{CSRC}", *srcs, "-o", hip_lib])
    out["libhelix_hip"] = hip_lib

    host_lib = os.path.join(PKG, "libhelix.so")
    hsrcs = [os.path.join(CSRC, s) for s in HOST_SOURCES]
    _run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-Wall", "-Wextra", f"-I{INCLUDE}",
          f"-I{os.path.join(CSRC, 'host')}", *hsrcs, "-o", host_lib, f"-L{PKG}", "-lhelix_hip",
          "-Wl,-rpath,$ORIGIN"])
    out["libhelix"] = host_lib

    # the `helix` Python module: pybind11 bindings of the public C++ API
    import sysconfig
    import pybind11
    ext = sysconfig.get_config_var("EXT_SUFFIX") or ".so"
    pymod = os.path.join(PKG, "helix" + ext)
    _run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-Wall", f"-I{INCLUDE}", f"-I{pybind11.get_include()}",
          f"-I{sysconfig.get_paths()['include']}", os.path.join(CSRC, "host", "python_module.cpp"), "-o", pymod,
          f"-L{PKG}", "-lhelix", "-lhelix_hip", "-Wl,-rpath,$ORIGIN"])
    out["helix_python_module"] = pymod

    for src, name in (("legacy_spin_glass_trace.cpp", "helix_example"), ("trace_and_checkpoint.cpp", "helix_trace_checkpoint"),
                      ("sparse_system.cpp", "helix_sparse_system")):
        example = os.path.join(ROOT, "examples", src)
        if os.path.exists(example):
            exe = os.path.join(PKG, name)
            _run(["g++", "-O2", "-std=c++17", f"-I{INCLUDE}", example, "-o", exe, f"-L{PKG}", "-lhelix", "-lhelix_hip",
                  "-Wl,-rpath,$ORIGIN"])
            out[name] = exe
    # the reference's own example program, compiled unchanged from where it lies (a drop-in check of the public
    # headers; nothing is copied): only in a tree next to the reference, its binary then travels with the tree
    if os.path.exists(REFERENCE_EXAMPLE):
        exe = os.path.join(PKG, "helix_reference_example")
        _run(["g++", "-O2", "-std=c++17", f"-I{INCLUDE}", REFERENCE_EXAMPLE, "-o", exe, f"-L{PKG}", "-lhelix",
              "-lhelix_hip", "-Wl,-rpath,$ORIGIN"])
        out["helix_reference_example"] = exe
    if verbose:
        for k, v in out.items():
            print(f"built {k}: {v}")
    return out


if __name__ == "__main__":
    build(verbose=True)
