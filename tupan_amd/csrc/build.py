"""Build libtupan_hip.so (and libtupan_hip_f32.so) for gfx950 with hipcc.

Usage:  python -m tupan_amd.csrc.build [--jobs N]

Outputs go to tupan_amd/_lib/.  Every translation unit is compiled with
  --offload-arch=gfx950 -O3 -ffp-contract=off
  -Xclang -target-feature -Xclang -packed-fp32-ops
The last pair keeps the compiler from forming packed FP32 instructions (some
of their forms are unreliable on the MI355X fleet this targets); the host
pass warns that it does not know the feature, which is harmless.
-ffp-contract=off keeps the kernels' operation order exactly as written:
every fused multiply-add in the sources is explicit.
"""
import argparse
import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.dirname(HERE)
LIBDIR = os.path.join(PKG, "_lib")
BUILDDIR = os.path.join(PKG, "_build")

SOURCES = ["capi.hip", "pairwise.hip", "sakura.hip", "ensemble.hip", "mlp.hip", "sia.hip",
           "hermite.hip", "sakura_step.hip"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

COMMON = [
    "--offload-arch=" + ARCH,
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-function",
]


def _run(cmd):
    proc = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT,
                          text=True)
    out = "\n".join(line for line in proc.stdout.splitlines()
                    if "packed-fp32-ops" not in line
                    and "not a recognized feature" not in line)
    if proc.returncode != 0:
        raise RuntimeError("command failed (%d): %s\n%s"
                           % (proc.returncode, " ".join(cmd), out))
    return out


def build(jobs=None, verbose=False):
    """Compile every source and link the libraries.  Always recompiles."""
    if not os.path.exists(HIPCC):
        raise RuntimeError("hipcc not found at %s" % HIPCC)
    os.makedirs(LIBDIR, exist_ok=True)
    os.makedirs(BUILDDIR, exist_ok=True)
    objs = []
    cmds = []
    for src in SOURCES:
        obj = os.path.join(BUILDDIR, src.replace(".hip", ".o"))
        objs.append(obj)
        cmds.append([HIPCC] + COMMON + ["-c", os.path.join(HERE, src),
                                        "-o", obj])
    jobs = jobs or min(len(cmds), os.cpu_count() or 1, 8)
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        for out in ex.map(_run, cmds):
            if verbose and out.strip():
                print(out)
    lib = os.path.join(LIBDIR, "libtupan_hip.so")
    tmp = lib + ".tmp"
    _run([HIPCC, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", tmp]
         + objs)
    os.replace(tmp, lib)

    # FP32 reference-ABI library: the same entry point names with float
    # types, forwarding to libtupan_hip.so.
    lib32 = os.path.join(LIBDIR, "libtupan_hip_f32.so")
    tmp32 = lib32 + ".tmp"
    _run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared",
          "-I" + os.path.join(PKG, os.pardir, "include"),
          os.path.join(HERE, "compat_f32.cpp"), "-o", tmp32,
          "-L" + LIBDIR, "-ltupan_hip", "-Wl,-rpath,$ORIGIN"])
    os.replace(tmp32, lib32)
    return lib


def clean():
    shutil.rmtree(BUILDDIR, ignore_errors=True)


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=None)
    ap.add_argument("-v", "--verbose", action="store_true")
    args = ap.parse_args(argv)
    lib = build(args.jobs, args.verbose)
    print(lib)
    return 0


if __name__ == "__main__":
    sys.exit(main())
