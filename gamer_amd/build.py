"""Build the gfx950 solver library in-tree: gamer_amd/libgamer_amd_hydro.so (hipcc, no JIT cache)."""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libgamer_amd_hydro.so")
SOURCES = ["hydro_api.hip", "hydro_mhm_rp.hip", "hydro_mhm.hip:32", "hydro_mhm.hip:64", "hydro_grid.hip",
           "hydro_adjoint.hip", "hydro_sync.hip", "hydro_interp.hip"]     # "file:NN": the file for one precision only
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    # no packed-FP32 VALU ops anywhere in device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I" + os.path.join(ROOT, "include"), "-I" + CSRC,
]
# per-source additions.  The MHM_RP kernels march z-planes with four barriers per plane; LLVM's iterative ILP
# scheduler gives the 8-wave kernels a schedule that is 5.1% (FP32) and 9.1% (FP64) faster than the default one
# and 11% / 4% faster than the max-ILP one (interleaved A/B on the 4096-PG workload).  The MHM file is compiled
# once per precision: its FP32 kernels run fastest with the max-ILP scheduler (2.3% ahead of iterative ILP), its
# FP64 kernels with iterative ILP (4.1% ahead of the default).  The adjoint, interpolation and sync sources crash
# the compiler under these strategies and keep the default.
EXTRA = {"hydro_mhm_rp.hip": ["-mllvm", "-amdgpu-sched-strategy=iterative-ilp"],
         "hydro_mhm.hip:32": ["-DMHM_PRECISION=32", "-mllvm", "-amdgpu-sched-strategy=max-ilp"],
         "hydro_mhm.hip:64": ["-DMHM_PRECISION=64", "-mllvm", "-amdgpu-sched-strategy=iterative-ilp"]}


def build(verbose: bool = False) -> str:
    """Compile every HIP source into one shared library (always recompiles)."""
    objs = []
    procs = []
    for src in SOURCES:
        path = src.split(":")[0]
        obj = os.path.join(CSRC, src.replace(".hip", "").replace(":", "_") + ".o")
        cmd = [HIPCC, *FLAGS, *EXTRA.get(src, []), "-c", os.path.join(CSRC, path), "-o", obj]
        cmd.remove("-shared")
        procs.append((cmd, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)))
        objs.append(obj)
    for cmd, p in procs:
        out, _ = p.communicate()
        if p.returncode != 0:
            raise RuntimeError(f"hipcc failed: {' '.join(cmd)}\n{out}")
        if verbose and out.strip():
            print(out)
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", *objs, "-o", tmp]
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{r.stdout}")
    os.replace(tmp, LIB)
    for o in objs:
        os.remove(o)
    return LIB


def build_host_sanitized(out_dir: str) -> str:
    """The same library with its host code (C ABI, argument checks, buffer management) built with AddressSanitizer
    and UndefinedBehaviorSanitizer, every report fatal; device code as in build() but without the scheduler
    options (they do not change host code).  Written to out_dir/libgamer_amd_hydro.so, for the sanitizer test,
    which loads it with the compiler's ASan runtime preloaded.  Never the shipped library."""
    os.makedirs(out_dir, exist_ok=True)
    san = ["-Xarch_host", "-fsanitize=address", "-Xarch_host", "-fsanitize=undefined",
           "-Xarch_host", "-fno-sanitize-recover=all"]
    flags = [f for f in FLAGS if f != "-shared"] + ["-Xarch_host", "-O1", "-Xarch_host", "-g"] + san
    objs, procs = [], []
    for src in SOURCES:
        obj = os.path.join(out_dir, src.replace(".hip", "").replace(":", "_") + ".o")
        extra = [f for f in EXTRA.get(src, []) if f.startswith("-D")]
        cmd = [HIPCC, *flags, *extra, "-c", os.path.join(CSRC, src.split(":")[0]), "-o", obj]
        procs.append((cmd, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)))
        objs.append(obj)
    for cmd, p in procs:
        out, _ = p.communicate()
        if p.returncode != 0:
            raise RuntimeError(f"hipcc failed: {' '.join(cmd)}\n{out}")
    lib = os.path.join(out_dir, "libgamer_amd_hydro.so")
    r = subprocess.run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", *san[:4], "-shared-libsan", *objs,
                        "-o", lib], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout}")
    for o in objs:
        os.remove(o)
    return lib


if __name__ == "__main__":
    if "--sanitize-host" in sys.argv:
        print(build_host_sanitized(sys.argv[sys.argv.index("--sanitize-host") + 1]))
    else:
        print(build(verbose="-v" in sys.argv))
