"""The C-ABI library: it loads without a GPU, exports every function include/gamer_amd_hydro.h declares,
and rejects bad arguments with HYDRO_E_* codes before touching the device."""
import ctypes
import os
import re

import pytest

from gamer_amd import hydro

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "gamer_amd_hydro.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:const\s+)?(?:int|void|char|size_t)\s*\*?\s*(hydro_\w+)\s*\(", text, flags=re.M)))


def test_header_declares_expected_api():
    assert set(declared_functions()) == set(hydro.EXPORTS)


def test_library_exports_every_declared_symbol():
    L = hydro.lib()
    for name in declared_functions():
        assert hasattr(L, name), name


def test_adjoint_entry_points_reject_bad_arguments():
    """Only MHM_RP in FP64 is differentiated; a workspace is required; checks run before any device work."""
    L = hydro.lib()
    assert L.hydro_mhm_adjoint_workspace_bytes(0) == 0
    assert L.hydro_mhm_adjoint_workspace_bytes(2) == 2 * L.hydro_mhm_adjoint_workspace_bytes(1) > 0
    ok = dict(npg=4, dt=1e-3, dh=1e-2, in_=16, store_flux=False)
    for kw, frag in [(dict(scheme="MHM"), "MHM_RP"), (dict(precision="fp32"), "FP64")]:
        a = hydro.make_args(**{**ok, **kw})
        assert L.hydro_mhm_adjoint_async(ctypes.byref(a), 16, 16, 16, None) == hydro.HYDRO_E_ARG
        assert frag in L.hydro_last_error().decode()
    a = hydro.make_args(**ok)
    assert L.hydro_mhm_adjoint_async(ctypes.byref(a), 16, 16, None, None) == hydro.HYDRO_E_ARG
    assert "workspace" in L.hydro_last_error().decode()
    assert L.hydro_mhm_adjoint_async(ctypes.byref(a), None, 16, 16, None) == hydro.HYDRO_E_ARG
    assert L.hydro_mhm_tangent_async(ctypes.byref(a), None, None, 16, None) == hydro.HYDRO_E_ARG
    assert L.hydro_prepare_uniform_adjoint_async(16, 16, 20, 16, 16, 3, 0, None) == hydro.HYDRO_E_ARG
    # a batch whose per-phase grids would exceed one launch is an argument error, found before any launch
    big = hydro.make_args(**{**ok, "npg": 1 << 19})
    assert L.hydro_mhm_adjoint_async(ctypes.byref(big), 16, 16, 16, None) == hydro.HYDRO_E_ARG
    assert "does not fit" in L.hydro_last_error().decode()
    big = hydro.make_args(**{**ok, "npg": 1 << 19, "out": 16})
    assert L.hydro_mhm_tangent_async(ctypes.byref(big), 16, 16, 16, None) == hydro.HYDRO_E_ARG
    assert "does not fit" in L.hydro_last_error().decode()


def _call(**kw):
    a = hydro.make_args(**kw)
    return hydro.lib().hydro_mhm_solve_async(ctypes.byref(a), None)


@pytest.mark.parametrize("kw,frag", [
    (dict(rsolver="HLLC"), "in/out"),
    (dict(), "in/out"),
])
def test_null_pointers_rejected(kw, frag):
    rc = _call(npg=4, dt=1e-3, dh=1e-2, **kw)
    assert rc == -1 and frag in hydro.lib().hydro_last_error().decode()


def test_bad_enums_rejected():
    a = hydro.make_args(npg=0, dt=1e-3, dh=1e-2)
    for field, bad in (("rsolver", 7), ("limiter", 5), ("precision", 3), ("scheme", 9), ("minmod_maxiter", -1)):
        b = hydro.HydroSolveArgs.from_buffer_copy(a)
        setattr(b, field, bad)
        assert hydro.lib().hydro_mhm_solve_async(ctypes.byref(b), None) == -1, field
    b = hydro.HydroSolveArgs.from_buffer_copy(a)
    b.gamma = 1.0
    assert hydro.lib().hydro_mhm_solve_async(ctypes.byref(b), None) == -1


def test_empty_batch_is_a_no_op():
    assert _call(npg=0, dt=1e-3, dh=1e-2) == 0


def test_host_path_requires_allocation():
    a = hydro.make_args(npg=1, dt=1e-3, dh=1e-2)
    buf = ctypes.create_string_buffer(8)
    rc = hydro.lib().hydro_asyn_fluid_solver(ctypes.cast(buf, ctypes.c_void_p), ctypes.cast(buf, ctypes.c_void_p),
                                             ctypes.cast(buf, ctypes.c_void_p), ctypes.byref(a), 1)
    assert rc == -2 and "hydro_mem_allocate_fluid" in hydro.lib().hydro_last_error().decode()


def _interp_args(scheme, cstart, crange, csize, reduce_mono=0):
    I3 = ctypes.c_int * 3
    return hydro.HydroInterpArgs(16, 16, 1, I3(*csize), I3(*cstart), I3(*crange), I3(*(2 * r for r in crange)),
                                 I3(0, 0, 0), scheme, 0x1F, 2.0, 0, 0, reduce_mono, 10, 5.0 / 3.0, 0.0,
                                 hydro.PRECISION["fp64"], 16)


@pytest.mark.parametrize("scheme,cstart,crange,csize,reduce_mono,frag", [
    (8, (1, 1, 1), (4, 4, 4), (6, 6, 6), 0, "scheme"),
    (0, (1, 1, 1), (4, 4, 4), (6, 6, 6), 0, "scheme"),
    (3, (1, 1, 1), (4, 4, 4), (6, 6, 6), 1, "reduce_mono"),          # INT_VANLEER
    (1, (1, 1, 1), (4, 4, 4), (6, 6, 6), 1, "reduce_mono"),          # INT_MINMOD3D
    (6, (1, 2, 2), (4, 4, 4), (8, 8, 8), 0, "2 ghost cells"),        # INT_CQUAR needs two coarse ghosts
    (7, (2, 2, 2), (4, 4, 5), (8, 8, 8), 0, "2 ghost cells"),        # INT_QUAR: upper side
    (4, (0, 1, 1), (4, 4, 4), (6, 6, 6), 0, "1 ghost cells"),
])
def test_interpolate_rejects_bad_arguments(scheme, cstart, crange, csize, reduce_mono, frag):
    """Interpolate.cpp:103-114 (ghost widths of Int_Table) and :299-300 (no coefficient to reduce) as errors."""
    L = hydro.lib()
    a = _interp_args(scheme, cstart, crange, csize, reduce_mono)
    assert L.hydro_interpolate_async(ctypes.byref(a), None) == hydro.HYDRO_E_ARG
    assert frag in L.hydro_last_error().decode()


def test_header_is_c99_and_links(tmp_path):
    """The boundary is plain C: a C99 translation unit that includes the header and takes the address of every
    declared entry point compiles and links against the library (no C++ types, no torch)."""
    import shutil
    import subprocess
    cc = shutil.which("gcc")
    if cc is None:
        pytest.skip("gcc not available")
    names = declared_functions()
    src = tmp_path / "use_header.c"
    body = "\n".join(f"   (void *)&{n}," for n in names)
    src.write_text('#include "gamer_amd_hydro.h"\n#include <stdio.h>\n'
                   f"static void *const entry[] = {{\n{body}\n}};\n"
                   'int main( void ) { printf( "%d\\n", (int)( sizeof entry / sizeof entry[0] ) ); return 0; }\n')
    exe = tmp_path / "use_header"
    libdir = os.path.dirname(hydro.LIB_PATH)
    subprocess.run([cc, "-std=c99", "-Wall", "-Werror", f"-I{os.path.join(ROOT, 'include')}",
                    str(src), "-o", str(exe), f"-L{libdir}", "-lgamer_amd_hydro", f"-Wl,-rpath,{libdir}"], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout
    assert int(out) == len(names) > 20


def test_one_hip_runtime_whatever_the_import_order():
    """torch ships its own HIP runtime; ``hydro.lib()`` loads torch first so that the library binds to that copy.
    In a fresh process that asks for the library before importing torch, exactly one libamdhip64 is mapped."""
    import subprocess
    import sys
    pytest.importorskip("torch")
    code = ("import sys; sys.path.insert(0, %r)\n"
            "from gamer_amd import hydro\n"
            "hydro.lib()\n"
            "import torch\n"
            "maps = open('/proc/self/maps').read().splitlines()\n"
            "print(len({l.split()[-1] for l in maps if 'libamdhip64' in l}))\n") % ROOT
    out = subprocess.run([sys.executable, "-c", code], check=True, capture_output=True, text=True, timeout=600)
    assert out.stdout.strip().splitlines()[-1] == "1", out.stdout + out.stderr


class _FakeDev:
    """Duck-typed stand-in for a device tensor: exercises the wrappers' output checks without a GPU."""

    def __init__(self, shape, dtype, device="cuda:0", cuda=True, contiguous=True):
        import torch
        self.shape, self.dtype, self.device = tuple(shape), dtype, torch.device(device)
        self.is_cuda, self._c = cuda, contiguous

    def is_contiguous(self):
        return self._c


def test_check_out_rejects_bad_outputs():
    """Caller-supplied out / flux / dt_patch tensors are checked for device, dtype, shape and contiguity before any
    launch: a wrong one would make the kernels write through a raw pointer past its end."""
    torch = pytest.importorskip("torch")
    f64, dev = torch.float64, torch.device("cuda:0")
    shape = (4, 5, 16, 16, 16)
    hydro._check_out(None, shape, f64, dev, "out")
    hydro._check_out(_FakeDev(shape, f64), shape, f64, dev, "out")
    for bad, frag in [(_FakeDev(shape, torch.float32), "dtype"),
                      (_FakeDev((4, 5, 16, 16, 17), f64), "shape"),
                      (_FakeDev(shape, f64, cuda=False, device="cpu"), "device"),
                      (_FakeDev(shape, f64, device="cuda:1"), "device"),
                      (_FakeDev(shape, f64, contiguous=False), "contiguous")]:
        with pytest.raises(hydro.HydroError, match=frag):
            hydro._check_out(bad, shape, f64, dev, "out")


def test_wrappers_reject_host_tensors():
    torch = pytest.importorskip("torch")
    u = torch.zeros((1, 5, 22, 22, 22), dtype=torch.float64)
    for fn in (lambda: hydro.fluid_solver(u, 1e-3, 1e-2), lambda: hydro.solve_batch(u),
               lambda: hydro.dt_cfl(torch.zeros((1, 5, 8, 8, 8)), 1e-2, 0.5)):
        with pytest.raises(hydro.HydroError):
            fn()


def test_launch_shape_rejected_before_device_work():
    """waves_per_group: 0 (default), 4 or 8 for MHM_RP; 0 or 4 for MHM; anything else is HYDRO_E_ARG."""
    for scheme, w in (("MHM_RP", 2), ("MHM_RP", 16), ("MHM", 8), ("MHM_RP", -4)):
        rc = _call(npg=4, dt=1e-3, dh=1e-2, in_=16, out=16, flux=16, scheme=scheme, waves_per_group=w)
        assert rc == -1 and "waves_per_group" in hydro.lib().hydro_last_error().decode(), (scheme, w)
