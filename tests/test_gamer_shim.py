"""The GAMER-side adapter (gamer_amd/gamer_shim/CUAPI_FluidSolver_gfx950.cpp) against GAMER's own declarations.

Compiled with g++ against the reference's headers (include/GAMER.h with -DGPU, so include/Prototype.h declares
the GPU API), for several compile-time variants.  Pinned:
* every CUAPI_* function that GAMER's declarations make its callers reference (a probe translation unit takes
  their addresses) is defined by the adapter under the identical mangled name, so a prototype mismatch, which in
  C++ would silently declare an overload, fails here;
* every library function the adapter calls is exported by gamer_amd/libgamer_amd_hydro.so.
Needs the reference tree and g++; skipped otherwise (the reference does not travel to GPU machines).
"""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF = "/root/reference"
SHIM = os.path.join(ROOT, "gamer_amd", "gamer_shim", "CUAPI_FluidSolver_gfx950.cpp")
LIB = os.path.join(ROOT, "gamer_amd", "libgamer_amd_hydro.so")

pytestmark = pytest.mark.skipif(not (os.path.isfile(os.path.join(REF, "include", "GAMER.h")) and shutil.which("g++")
                                     and shutil.which("nm")),
                                reason="reference headers or toolchain not present")

# GAMER's CUAPI_MemAllocate.cu declares CUAPI_MemAllocate_Fluid and CUAPI_SetMemSize locally
# (src/GPU_API/CUAPI_MemAllocate.cu:9-11), and CPU_EoS_Gamma.cpp declares EoS_SetGPUFunc_Gamma
# (src/EoS/Gamma/CPU_EoS_Gamma.cpp:447-449); the rest is include/Prototype.h:646-690
PROBE = """#include "GAMER.h"
int CUAPI_MemAllocate_Fluid( const int Flu_NPG, const int Pot_NPG, const int Src_NPG, const int GPU_NStream );
void CUAPI_SetMemSize( int &GPU_NStream, int &Flu_GPU_NPGroup, int &Pot_GPU_NPGroup, int &Che_GPU_NPGroup,
                       int &Src_GPU_NPGroup );
void EoS_SetGPUFunc_Gamma( EoS_DE2P_t &, EoS_DP2E_t &, EoS_DP2C_t &, EoS_DE2T_t &, EoS_DT2P_t &, EoS_DE2S_t &,
                           EoS_GENE_t & );
void *probe[] = { (void*)&CUAPI_Asyn_FluidSolver, (void*)&CUAPI_SetDevice, (void*)&CUAPI_SetCache,
                  (void*)&CUAPI_MemFree_Fluid, (void*)&CUAPI_Synchronize, (void*)&CUAPI_MemAllocate_Fluid,
                  (void*)&CUAPI_Asyn_dtSolver, (void*)&CUAPI_DiagnoseDevice, (void*)&CUAPI_MemAllocate,
                  (void*)&CUAPI_SetConstMemory, (void*)&CUAPI_SetConstMemory_EoS, (void*)&CUAPI_SetMemSize,
                  (void*)&EoS_SetGPUFunc_Gamma };
"""
NPROBE = 13

VARIANTS = [("MHM_RP", "HLLC", False), ("MHM_RP", "ROE", True), ("MHM", "HLLE", False), ("MHM", "HLLC", True)]


def _flags(scheme, rsolver, fp64):
    f = ["-std=c++17", "-c", "-w", f"-I{REF}/include", f"-I{ROOT}/include", "-DMODEL=HYDRO", f"-DFLU_SCHEME={scheme}",
         f"-DRSOLVER={rsolver}", "-DLR_SCHEME=PLM", "-DEOS=EOS_GAMMA", "-DNCOMP_PASSIVE_USER=0", "-DPATCH_SIZE=8",
         "-DNLEVEL=10", "-DMAX_PATCH=1000000", "-DOPENMP", "-DSERIAL", "-DRANDOM_NUMBER=RNG_GNU_EXT", "-DGPU",
         "-DGPU_COMPUTE_CAPABILITY=900"]
    return f + (["-DFLOAT8"] if fp64 else [])


def _symbols(obj, kind):
    out = subprocess.run(["nm", obj], check=True, capture_output=True, text=True).stdout
    return {line.split()[-1] for line in out.splitlines() if line.split()[-2:-1] == [kind]}


@pytest.mark.parametrize("scheme,rsolver,fp64", VARIANTS)
def test_shim_defines_gamer_gpu_api(tmp_path, scheme, rsolver, fp64):
    flags = _flags(scheme, rsolver, fp64)
    shim_o, probe_o, probe_c = tmp_path / "shim.o", tmp_path / "probe.o", tmp_path / "probe.cpp"
    probe_c.write_text(PROBE)
    subprocess.run(["g++"] + flags + [SHIM, "-o", str(shim_o)], check=True)
    subprocess.run(["g++"] + flags + [str(probe_c), "-o", str(probe_o)], check=True)
    wanted = {s for s in _symbols(probe_o, "U") if "CUAPI_" in s or "EoS_SetGPUFunc" in s}
    assert len(wanted) == NPROBE, wanted
    missing = wanted - _symbols(shim_o, "T")
    assert not missing, f"GAMER declares but the adapter does not define: {missing}"


def test_shim_calls_only_exported_library_functions(tmp_path):
    if not os.path.isfile(LIB):
        pytest.skip("library not built")
    shim_o = tmp_path / "shim.o"
    subprocess.run(["g++"] + _flags("MHM_RP", "HLLC", False) + [SHIM, "-o", str(shim_o)], check=True)
    called = {s for s in _symbols(shim_o, "U") if s.startswith("hydro_")}
    out = subprocess.run(["nm", "-D", "--defined-only", LIB], check=True, capture_output=True, text=True).stdout
    exported = {line.split()[-1] for line in out.splitlines()}
    assert called and called <= exported, called - exported


@pytest.mark.parametrize("extra,frag", [
    (["-ULR_SCHEME", "-DLR_SCHEME=PPM"], "PLM"),
    (["-UNCOMP_PASSIVE_USER", "-DNCOMP_PASSIVE_USER=1"], "passive"),
    (["-DDUAL_ENERGY=DE_ENPY"], "dual energy"),
    (["-DMHD"], "MHD"),
    (["-UEOS", "-DEOS=EOS_ISOTHERMAL"], "EOS_GAMMA"),
    (["-DGRAVITY"], "self-gravity"),
])
def test_shim_rejects_unsupported_builds(tmp_path, extra, frag):
    """A GAMER build whose arrays the library's layouts do not fit (PPM's FLU_NXT, passive scalars, dual energy, MHD,
    another EoS, self-gravity) must not compile against the adapter: it would overflow the pinned buffers."""
    flags = _flags("MHM_RP", "HLLC", False) + extra
    r = subprocess.run(["g++"] + flags + [SHIM, "-o", str(tmp_path / "shim.o")], capture_output=True, text=True)
    assert r.returncode != 0
    assert frag in r.stderr, r.stderr[-2000:]
