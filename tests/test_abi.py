"""The drop-in C ABI: the library loads without a GPU, exports every
function include/*.h declares (and nothing from the C++ internals), and
its host-side pieces that need no device (GCF table, bin-to-plane map,
AIPS blank value) agree with Obit's definitions."""
import ctypes as C
import glob
import os
import re
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        txt = open(h).read()
        txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
        for m in re.finditer(r"^[A-Za-z_][\w\s\*]*?\b([A-Za-z_]\w*)\s*\(", txt, flags=re.M):
            name = m.group(1)
            if name not in ("if", "while", "for", "return", "sizeof"):
                names.add(name)
    return names


def _exports(path):
    out = subprocess.run(["nm", "-D", "--defined-only", path], capture_output=True, text=True, check=True).stdout
    return {ln.split()[-1] for ln in out.splitlines() if " T " in ln}


def test_every_declared_symbol_is_exported():
    from obit_amd import _lib
    lib = _lib.load()
    declared = _declared()
    assert len(declared) >= 50
    exports = _exports(_lib.LIB_PATH)
    missing = sorted(declared - exports)
    assert not missing, missing
    for name in declared:
        getattr(lib, name)
    # nothing but the C ABI leaks out
    assert all(n.startswith("Obit") or n == "CUDAMagicF" for n in exports), sorted(
        n for n in exports if not n.startswith("Obit"))


def test_reference_entry_points_present():
    """The ObitCUDA* entry points Obit's host classes call
    (src/ObitGPUSkyModel.c, src/ObitGPUGrid.c, src/ObitGPUBeamModel.c)."""
    from obit_amd import _lib
    exports = _exports(_lib.LIB_PATH)
    for n in ("ObitCUDASkyModelDFTInit", "ObitCUDASkyModelDFTSetMod", "ObitCUDASkyModelDFTCalc",
              "ObitCUDASkyModelDFTShutdown", "ObitCUDAGridAlloc", "ObitCUDAGridInit",
              "ObitCUDAGridGrid", "ObitCUDAGridFlip", "ObitCUDAGrid2CPU", "ObitCUDAGridShutdown",
              "ObitCUDACheckGPU", "ObitCUDASetGPU", "ObitCUDAResetGPU", "ObitCUDADeviceSynchronize",
              "ObitCUDAStreamCreate", "ObitCUDAStreamDestroy", "ObitCUDAEventCreate",
              "ObitCUDAEventDestroy", "ObitCUDAEventRecord", "ObitCUDAEventSynchronize",
              "ObitCUDAUtilAllocHost", "ObitCUDAUtilFreeHost", "ObitCUDAUtilAllocGPU",
              "ObitCUDAUtilFreeGPU", "ObitCUDAUtilHost2GPU", "ObitCUDAUtilHostAny2GPU",
              "ObitCUDAUtilGPU2Host", "ObitCUDAUtilGPUAny2Host", "ObitCUDAUtilMemory",
              "ObitCUDAUtilPointerType", "CUDAMagicF", "ObitCUDABeamModelDFTInit",
              "ObitCUDABeamModelDFTSetMod", "ObitCUDABeamModelDFTCalc", "ObitCUDABeamModelDFTShutdown"):
        assert n in exports, n


def test_magic_blank():
    from obit_amd import _lib
    v = _lib.load().CUDAMagicF()
    assert np.float32(v).tobytes() == b"INDE"


def test_host_gcf_table_equals_oracle(oracle):
    from obit_amd.gridder import conv_func_sph
    cf, _ = oracle.convfunc()
    h = conv_func_sph()
    assert np.array_equal(h[:701], cf[:701])


def test_planes_from_bins():
    from obit_amd.gridder import plane_of_channel
    b = np.array([0, 10, 20]); e = np.array([9, 19, 29])
    cpu = plane_of_channel(30, b, e, cpu_parity=True)
    gpu = plane_of_channel(30, b, e, cpu_parity=False)
    # CPU: [bCh, eCh) with an inclusive eCh -> last channel of each bin left out
    assert list(cpu[:10]) == [0] * 9 + [103] and list(cpu[10:20]) == [1] * 9 + [103]
    assert list(gpu) == [0] * 10 + [1] * 10 + [2] * 10


def test_status_and_message_without_gpu():
    """An ObitHIP* call made in the wrong order returns a status and a
    message instead of aborting (no device needed)."""
    from obit_amd import _lib
    lib = _lib.load()
    fake = C.c_void_p(12345)
    rc = lib.ObitHIPSkyModelDFTCalcResident(fake)
    assert rc == 4
    assert b"resident" in lib.ObitHIPLastError()


def test_grid_refuses_more_rows_than_the_lists_index():
    """The track lists hold 32-bit row indices: a resident pass of 2^31 rows
    or more is refused before anything touches the device."""
    from obit_amd import _lib
    lib = _lib.load()
    fake = C.c_void_p(12345)
    for fn in (lib.ObitHIPGridGridDevice, lib.ObitHIPGridGridDeviceReuse):
        rc = fn(fake, 0, C.c_void_p(0), C.c_longlong(2 ** 31))
        assert rc == 1
        assert b"2^31" in lib.ObitHIPLastError()


def test_dft_path_query_without_state():
    from obit_amd import _lib
    lib = _lib.load()
    p = C.c_int(-1)
    assert lib.ObitHIPSkyModelDFTGetPath(C.c_void_p(777), C.byref(p)) == 4
    assert p.value == -1


def test_beam_model_struct_layouts(tmp_path):
    """The ctypes mirrors of the beam-model structures have the sizes and
    member offsets include/ObitHIPBeamModel.h gives them under the host C
    compiler (the layout the reference's host class fills)."""
    from obit_amd import _lib
    names = ["GPUBeamVisInfo", "GPUBeamModelInfo", "CUDAFArray", "CUDAImageDesc", "CUDABeamInterp",
             "GPUBeamAntInfo", "GPUBeamInfo"]
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "ObitHIPBeamModel.h"', 'int main(void) {']
    want = []
    for n in names:
        st = getattr(_lib, n)
        lines.append(f'  printf("%zu\\n", sizeof({n}));')
        want.append(C.sizeof(st))
        for f in st._fields_:
            lines.append(f'  printf("%zu\\n", offsetof({n}, {f[0]}));')
            want.append(getattr(st, f[0]).offset)
    lines += ['  return 0;', '}']
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c99", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    got = [int(x) for x in subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split()]
    assert got == want


def test_beam_model_entry_points_need_init():
    """Calc on a GPUBeamInfo that Init never saw exits with EXIT_FAILURE and
    a message (as the reference's checkCudaErrors does); run in a child, with
    no GPU."""
    code = ("import ctypes as C\n"
            "from obit_amd import _lib\n"
            "lib = _lib.load()\n"
            "g = _lib.GPUBeamInfo()\n"
            "vr = (C.c_long * 2)(0, 9)\n"
            "lib.ObitCUDABeamModelDFTCalc(C.byref(g), 0, 1, 0.0, vr)\n")
    r = subprocess.run(["python3", "-c", code], cwd=ROOT, capture_output=True, text=True)
    assert r.returncode == 1
    assert "ObitCUDABeamModelDFTCalc" in r.stderr and "Init" in r.stderr
