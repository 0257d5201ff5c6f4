"""Build ``libshorefit.so`` (HIP for gfx950) in-tree with hipcc.

    python -m cs_dmri_amd.build

Every translation unit is compiled for gfx950 only, with packed-FP32 VALU
instructions disabled (``-packed-fp32-ops``).
"""

from __future__ import annotations

import os
import shutil
import subprocess
import sys
import time
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
LIB = PKG / "libshorefit.so"
SOURCES = ["setup.hip", "fista.hip", "postfit.hip", "graddev.hip", "adjoint.hip", "volume.hip", "noalloc.hip",
           "api.cpp"]
ARCH = "gfx950"


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found (ROCm is required to build libshorefit.so)")


def build_stamp() -> tuple[str, str]:
    """(git SHA of the sources, with "-dirty" for uncommitted changes, or
    "unknown"; UTC build time in ISO 8601) baked into the library for the
    provenance record of its outputs."""
    sha = "unknown"
    try:
        r = subprocess.run(["git", "-C", str(ROOT), "rev-parse", "HEAD"], capture_output=True, text=True, timeout=10)
        if r.returncode == 0 and r.stdout.strip():
            sha = r.stdout.strip()
            d = subprocess.run(["git", "-C", str(ROOT), "status", "--porcelain", "--untracked-files=no"],
                               capture_output=True, text=True, timeout=10)
            if d.returncode == 0 and d.stdout.strip():
                sha += "-dirty"
    except (OSError, subprocess.SubprocessError):
        pass
    return sha, time.strftime("%Y-%m-%dT%H:%M:%SZ", time.gmtime())


def flags(stamp: tuple[str, str] | None = None) -> list[str]:
    sha, ts = stamp or ("unknown", "unknown")
    return [
        f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-Wall",
        "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
        f"-I{ROOT / 'include'}", f"-I{CSRC}",
        f'-DSHOREFIT_GIT_SHA="{sha}"', f'-DSHOREFIT_BUILD_TIME="{ts}"',
    ]


def build(verbose: bool = False) -> Path:
    cc = hipcc()
    objdir = PKG / "_build"
    objdir.mkdir(exist_ok=True)
    objs = []
    procs = []
    stamp = build_stamp()
    for src in SOURCES:
        path = CSRC / src
        if not path.exists():
            continue
        obj = objdir / (path.stem + ".o")
        cmd = [cc, *flags(stamp), "-x", "hip", "-c", str(path), "-o", str(obj)]
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append((src, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)))
        objs.append(obj)
    failed = []
    for src, p in procs:
        out, _ = p.communicate()
        if p.returncode != 0:
            failed.append(f"--- {src} ---\n{out}")
        elif verbose and out.strip():
            print(out)
    if failed:
        raise RuntimeError("hipcc failed:\n" + "\n".join(failed))
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [cc, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp), *map(str, objs)]
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    return LIB


# Host-code sanitizers (AddressSanitizer + UndefinedBehaviorSanitizer) for
# the C-ABI layer: every flag goes to the host pass only, the device code is
# built as usual.
SAN_HOST_FLAGS = ["-O1", "-g",
                  "-Xarch_host", "-fsanitize=address", "-Xarch_host", "-fsanitize=undefined",
                  "-Xarch_host", "-fno-sanitize-recover=all", "-Xarch_host", "-fno-omit-frame-pointer"]
SAN_LIB = PKG / "_san" / "libshorefit.so"


def asan_runtime() -> Path:
    """The clang AddressSanitizer runtime the sanitized library links against
    (preload it into the host process)."""
    cc = Path(hipcc()).resolve()
    for root in (cc.parent.parent / "llvm" / "lib" / "clang", cc.parent.parent / "lib" / "clang",
                 Path("/opt/rocm/llvm/lib/clang")):
        hits = sorted(root.glob("*/lib/linux/libclang_rt.asan-x86_64.so"))
        if hits:
            return hits[-1]
    raise RuntimeError("libclang_rt.asan-x86_64.so not found next to hipcc")


def build_host_sanitized(verbose: bool = False) -> Path:
    """``_san/libshorefit.so``: api.cpp (argument checks, gradient table,
    shells, caches, progress) built with host sanitizers and linked with the
    kernel objects of the regular build (run ``build()`` first)."""
    cc = hipcc()
    objdir = PKG / "_build"
    kernels = [objdir / (Path(src).stem + ".o") for src in SOURCES if src != "api.cpp"]
    missing = [str(o) for o in kernels if not o.exists()]
    if missing:
        raise RuntimeError(f"kernel objects missing (run the regular build first): {missing}")
    out = SAN_LIB.parent
    out.mkdir(exist_ok=True)
    api = out / "api.o"
    fl = [f for f in flags(build_stamp()) if f != "-O3"]
    cmd = [cc, *fl, *SAN_HOST_FLAGS, "-x", "hip", "-c", str(CSRC / "api.cpp"), "-o", str(api)]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True, capture_output=not verbose)
    tmp = SAN_LIB.with_suffix(".so.tmp")
    link = [cc, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp), str(api), *map(str, kernels),
            "-Xarch_host", "-fsanitize=address", "-Xarch_host", "-fsanitize=undefined", "-shared-libsan"]
    subprocess.run(link, check=True, capture_output=not verbose)
    os.replace(tmp, SAN_LIB)
    return SAN_LIB


if __name__ == "__main__":
    if "--sanitize-host" in sys.argv:
        print(build_host_sanitized(verbose="-v" in sys.argv))
    else:
        print(build(verbose="-v" in sys.argv))
