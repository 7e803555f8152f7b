"""Build libweft_hip.so in-tree (no cmake, no JIT cache).

1. ``hipcc --genco --offload-arch=gfx950`` compiles csrc/weft_kernels.hip into the code
   object ``lib/weft_kernels.hsaco`` (-ffp-contract=off: no fused multiply-add, so the
   float32 arithmetic matches NumPy's separately rounded operations; denormals kept;
   packed FP32 instructions disabled).
2. The code object is embedded into the library through an ``.incbin`` assembly stub.
3. ``g++`` compiles csrc/weft_host.cpp against the HIP runtime and links
   ``lib/libweft_hip.so``.
4. ``g++`` compiles csrc/weft_pack.cpp, the host-side packing of the prepacked scoring
   drop-in, into the CPython extension ``lib/_weft_pack``.

Run ``python -m weft_amd.build`` (or ``__graft_entry__.build()``).
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
import sysconfig

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
LIBDIR = os.path.join(PKG, "lib")
INCLUDE = os.path.join(ROOT, "include")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
ARCH = "gfx950"

HSACO = os.path.join(LIBDIR, "weft_kernels.hsaco")
LIBSO = os.path.join(LIBDIR, "libweft_hip.so")

DEVICE_FLAGS = [
    "-O3", "-std=c++17", "-ffp-contract=off", "-fno-fast-math",
    "-fno-gpu-flush-denormals-to-zero",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def _run(cmd: list[str]) -> None:
    print("+", " ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def build(verbose: bool = True, libdir: str | None = None, defines: tuple[str, ...] = ()) -> str:
    """Build the code object and library into `libdir` (default weft_amd/lib)."""
    libdir = libdir or LIBDIR
    hsaco = os.path.join(libdir, "weft_kernels.hsaco")
    libso = os.path.join(libdir, "libweft_hip.so")
    os.makedirs(libdir, exist_ok=True)
    hipcc = shutil.which("hipcc") or os.path.join(ROCM, "bin", "hipcc")
    _run([hipcc, "--genco", f"--offload-arch={ARCH}", *DEVICE_FLAGS, *[f"-D{d}" for d in defines],
          "-I", INCLUDE, "-I", CSRC, os.path.join(CSRC, "weft_kernels.hip"), "-o", hsaco + ".tmp"])
    os.replace(hsaco + ".tmp", hsaco)
    embed = os.path.join(libdir, "weft_embed.S")
    with open(embed, "w") as f:
        f.write(
            "    .section .rodata\n"
            "    .balign 4096\n"
            "    .global weft_hsaco_start\n"
            "weft_hsaco_start:\n"
            f"    .incbin \"{hsaco}\"\n"
            "    .global weft_hsaco_end\n"
            "weft_hsaco_end:\n"
            "    .byte 0\n"
            "    .section .note.GNU-stack,\"\",@progbits\n"
        )
    _run(["g++", "-O2", "-fPIC", "-shared", "-std=c++17", "-D__HIP_PLATFORM_AMD__",
          "-Wall", "-Wextra", "-Wno-unused-parameter",
          "-I", os.path.join(ROCM, "include"), "-I", INCLUDE, "-I", CSRC,
          os.path.join(CSRC, "weft_host.cpp"), embed,
          "-L", os.path.join(ROCM, "lib"), "-lamdhip64", "-o", libso + ".tmp"])
    os.replace(libso + ".tmp", libso)
    os.remove(embed)
    build_pack(libdir)
    return libso


def pack_path(libdir: str | None = None) -> str:
    return os.path.join(libdir or LIBDIR, "_weft_pack" + sysconfig.get_config_var("EXT_SUFFIX"))


def build_pack(libdir: str | None = None) -> str:
    """The CPython extension that packs the prepacked drop-in's inputs (no HIP code)."""
    import numpy as np

    out = pack_path(libdir)
    _run(["g++", "-O3", "-fPIC", "-shared", "-std=c++17", "-Wall", "-Wextra", "-Wno-unused-parameter",
          "-Wno-missing-field-initializers", "-I", sysconfig.get_paths()["include"], "-I", np.get_include(),
          os.path.join(CSRC, "weft_pack.cpp"), "-pthread", "-o", out + ".tmp"])
    os.replace(out + ".tmp", out)
    return out


if __name__ == "__main__":
    build()
    sys.exit(0)
