"""Build script of the MI355X speckle-masking engine (in-tree, no JIT cache).

Produces, next to the sources:
  speckle-stacker_amd/lib/libsmip_mi355x.so   HIP kernels + C ABI (gfx950)
  speckle-stacker_amd/lib/libsmip_synth.so    host-only synthetic speckle frames
  speckle-stacker_amd/lib/smip_cli            smip-cli compatible driver
  speckle-stacker_amd/lib/smip_dropin_test    C++ drop-in surface test runner
  speckle-stacker_amd/lib/smip_gif_dump       the CLI's GIF decoder, for the tests
  speckle-stacker_amd/lib/smip_frame_dump     the CLI's PNG / GIF frame reading, for the tests
  speckle-stacker_amd/lib/smip_copy_calib     copy kernel calibrating rocprofv3 byte counters

Run ``python speckle-stacker_amd/build.py`` or call ``build()``.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(PKG, "csrc")
LIB = os.path.join(PKG, "lib")
OBJ = os.path.join(PKG, "lib", "obj")
REPO = os.path.dirname(PKG)
INCLUDE = os.path.join(REPO, "include")

HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
ARCH = "gfx950"
# Packed FP32 instructions stay disabled (the compiler would otherwise form
# v_pk_* from plain float code); the triple products use scalar FP32 FMAs.
HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument", f"-I{INCLUDE}",
]
HIP_SOURCES = ["fft.hip", "fft_mixed.hip", "bispec.hip", "reco.hip", "post.hip", "abi.hip", "serve.hip", "tune.hip"]


def _run(cmd: list[str]) -> None:
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout)
        raise RuntimeError(f"build step failed: {' '.join(cmd)}")
    out = "\n".join(l for l in r.stdout.splitlines()
                    if "packed-fp32-ops" not in l and "not a recognized feature" not in l)
    if out.strip():
        sys.stderr.write(out + "\n")


def build_engine(jobs: int = 8) -> str:
    os.makedirs(OBJ, exist_ok=True)
    objs = []
    cmds = []
    for src in HIP_SOURCES:
        o = os.path.join(OBJ, src.replace(".hip", ".o"))
        objs.append(o)
        cmds.append([HIPCC, *HIP_FLAGS, "-c", os.path.join(CSRC, src), "-o", o])
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        list(ex.map(_run, cmds))
    so = os.path.join(LIB, "libsmip_mi355x.so")
    tmp = so + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs])
    os.replace(tmp, so)
    return so


def build_synth() -> str | None:
    src = os.path.join(PKG, "synth", "speckle_synth.cpp")
    if not os.path.exists(src):
        return None
    so = os.path.join(LIB, "libsmip_synth.so")
    _run(["g++", "-std=c++17", "-O3", "-fPIC", "-shared", "-pthread", "-o", so + ".tmp", src])
    os.replace(so + ".tmp", so)
    return so


def build_cli() -> list[str]:
    built = []
    cli = os.path.join(PKG, "cli", "smip_cli.cpp")
    eng = os.path.join(LIB, "libsmip_mi355x.so")
    common = ["-std=c++20", "-O2", f"-I{INCLUDE}", f"-I{PKG}", "-Wl,-rpath,$ORIGIN", f"-L{LIB}", "-lsmip_mi355x"]
    if os.path.exists(cli):
        out = os.path.join(LIB, "smip_cli")
        _run(["g++", cli, "-o", out, *common, "-lz"])
        built.append(out)
    test = os.path.join(REPO, "tests", "cpp", "dropin_test.cpp")
    if os.path.exists(test):
        out = os.path.join(LIB, "smip_dropin_test")
        _run(["g++", test, "-o", out, *common])
        built.append(out)
    dump = os.path.join(REPO, "tests", "cpp", "gif_dump.cpp")
    if os.path.exists(dump):
        out = os.path.join(LIB, "smip_gif_dump")
        _run(["g++", dump, "-o", out, "-std=c++20", "-O2", f"-I{INCLUDE}", f"-I{PKG}", "-lz"])
        built.append(out)
    fdump = os.path.join(REPO, "tests", "cpp", "frame_dump.cpp")
    if os.path.exists(fdump):
        out = os.path.join(LIB, "smip_frame_dump")
        _run(["g++", fdump, "-o", out, "-std=c++20", "-O2", f"-I{INCLUDE}", f"-I{PKG}", "-lz"])
        built.append(out)
    return built


def build_calib() -> str:
    out = os.path.join(LIB, "smip_copy_calib")
    _run([HIPCC, f"--offload-arch={ARCH}", "-O3", "-x", "hip", os.path.join(CSRC, "copy_calib.cpp"), "-o", out])
    return out


def build() -> None:
    os.makedirs(LIB, exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=2) as ex:
        f_eng = ex.submit(build_engine)
        f_syn = ex.submit(build_synth)
        f_eng.result()
        f_syn.result()
    build_cli()
    build_calib()


if __name__ == "__main__":
    build()
    print("built", sorted(os.listdir(LIB)))
