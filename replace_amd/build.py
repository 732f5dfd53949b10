"""Builds libreplace_amd.so (HIP, gfx950) and the replace_cli driver in-tree.

  python -m replace_amd.build            # library + CLI
  python -m replace_amd.build --lib      # library only

Flags:
  --offload-arch=gfx950                        MI355X (CDNA4)
  -ffp-contract=off                            no FMA contraction: the reference's
                                               x86-64 build rounds every product
  -Xclang -target-feature -Xclang -packed-fp32-ops
                                               no packed FP32 instructions
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libreplace_amd.so")
BIN = os.path.join(HERE, "bin")
CLI = os.path.join(BIN, "replace_cli")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("RP_OFFLOAD_ARCH", "gfx950")

COMMON = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
    "-I" + os.path.join(ROOT, "include"), "-I" + CSRC,
]

LIB_SOURCES = ["rp_kernels.hip", "rp_ip.hip", "rp_dct_mfma.hip", "rp_noalloc.hip", "rp_host.cpp", "rp_batch.cpp",
               "Replace.cpp"]


def _run(cmd):
    print(" ".join(cmd), flush=True)
    subprocess.check_call(cmd)


def build_lib() -> str:
    """Compiles the sources in parallel (one hipcc per file, at most
    RP_BUILD_JOBS at once, default 8), then links."""
    from concurrent.futures import ThreadPoolExecutor

    objs = [os.path.join(CSRC, src + ".o") for src in LIB_SOURCES]
    cmds = [[HIPCC, *COMMON, "-x", "hip", "-c", os.path.join(CSRC, src), "-o", obj]
            for src, obj in zip(LIB_SOURCES, objs)]
    jobs = max(1, int(os.environ.get("RP_BUILD_JOBS", "8")))
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        for _ in ex.map(_run, cmds):
            pass
    _run([HIPCC, "-shared", f"--offload-arch={ARCH}", "-fPIC", "-o", LIB, *objs, "-lhipblaslt"])
    for o in objs:
        os.remove(o)
    return LIB


def build_cli() -> str:
    os.makedirs(BIN, exist_ok=True)
    src = os.path.join(CSRC, "replace_cli.cpp")
    _run(["g++", "-O2", "-std=c++17", "-I" + os.path.join(ROOT, "include"), src, "-o", CLI,
          "-L" + HERE, "-lreplace_amd", "-Wl,-rpath," + HERE + ",-rpath,$ORIGIN/.."])
    return CLI


def main(argv=None):
    argv = sys.argv[1:] if argv is None else argv
    build_lib()
    if "--lib" not in argv:
        build_cli()


if __name__ == "__main__":
    main()
