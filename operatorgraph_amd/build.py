"""Builds libpga_amd.so (the C ABI over the gfx950 kernels) in-tree with hipcc.

Flags that matter for parity and for the hardware:
  -ffp-contract=off        no FMA contraction: FP32 results follow the
                           reference's separately rounded products and sums;
  -fhip-fp32-correctly-rounded-divide-sqrt
                           IEEE division (Repeat/Subdivide divide sizes);
  -target-feature -packed-fp32-ops
                           no packed FP32 instructions.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libpga_amd.so")
SOURCES = ["pga_kernels.hip", "pga_adjoint.hip", "pga_lbvh.hip", "pga_loss.hip", "pga_mesh.hip", "pga_api.cpp",
           "pga_spec.cpp"]
HOST_SOURCES = ["pga_compiler.cpp"]   # no device code: built with the host compiler
HEADERS = ["pga_device.h", "pga_ops.h", "pga_kernels.h", "pga_level.h", "pga_spec.h", "pga_spec_host.h"]
# device headers compiled again at run time (hipRTC) for the kernels
# specialised to one program; embedded in the library as text
RTC_HEADERS = [("pga.h", os.path.join(ROOT, "include", "pga.h"))] + \
    [(h, os.path.join(CSRC, h)) for h in ("pga_device.h", "pga_ops.h", "pga_kernels.h", "pga_level.h", "pga_spec.h")]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
CXX = os.environ.get("CXX", "g++")
ARCH = os.environ.get("PGA_OFFLOAD_ARCH", "gfx950")

COMMON_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off", "-fno-fast-math",
    "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-fvisibility=hidden", "-Wall", "-Wno-unused-function", "-Wno-unused-result",
]


def _run(cmd):
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError("build step failed: " + " ".join(cmd))
    return res


def write_rtc_headers(build_dir):
    """The device headers as C++ raw string literals (pga_rtc_headers.inc)."""
    delim = "PGA_RTC_TEXT"
    names, texts = [], []
    for name, path in RTC_HEADERS:
        with open(path, encoding="utf-8") as f:
            text = f.read()
        if ")" + delim + '"' in text:
            raise RuntimeError(f"{path} contains the raw-string delimiter")
        names.append(f'"{name}"')
        texts.append(f'R"{delim}({text}){delim}"')
    out = os.path.join(build_dir, "pga_rtc_headers.inc")
    body = ("// written by operatorgraph_amd/build.py: device headers for hipRTC\n"
            "static const char* const kRtcHeaderNames[] = {" + ", ".join(names) + "};\n"
            "static const char* const kRtcHeaderTexts[] = {\n" + ",\n".join(texts) + "};\n")
    old = open(out, encoding="utf-8").read() if os.path.exists(out) else None
    if old != body:
        with open(out, "w", encoding="utf-8") as f:
            f.write(body)
    return out


def build(verbose=False):
    """Compiles every source for gfx950 into operatorgraph_amd/libpga_amd.so."""
    objs = []
    build_dir = os.path.join(HERE, "_build")
    os.makedirs(build_dir, exist_ok=True)
    write_rtc_headers(build_dir)
    for src in SOURCES:
        obj = os.path.join(build_dir, src + ".o")
        cmd = [HIPCC, f"--offload-arch={ARCH}", *COMMON_FLAGS, "-I", os.path.join(ROOT, "include"),
               "-I", build_dir, "-c", os.path.join(CSRC, src), "-o", obj]
        if src.endswith(".cpp"):
            cmd.insert(1, "-x")
            cmd.insert(2, "hip")
        if verbose:
            print(" ".join(cmd))
        _run(cmd)
        objs.append(obj)
    for src in HOST_SOURCES:
        obj = os.path.join(build_dir, src + ".o")
        cmd = [CXX, "-O2", "-std=c++17", "-ffp-contract=off", "-fPIC", "-fvisibility=hidden", "-Wall", "-Wextra", "-Werror",
               "-I", os.path.join(ROOT, "include"), "-c", os.path.join(CSRC, src), "-o", obj]
        if verbose:
            print(" ".join(cmd))
        _run(cmd)
        objs.append(obj)
    tmp = LIB + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs, "-lhiprtc"])
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
