"""CPU checks of the C ABI: the library loads, exports every entry point that
include/scft.h declares, and its bcp.input reader agrees with the oracle's
restatement of read_input (code/read_input.cpp:7-242)."""
import ctypes
import os

import pytest

import pyoracle
import scft_amd


def test_library_exports_every_declared_symbol():
    lib = scft_amd.load_library()
    names = scft_amd.exported_symbols()
    assert "scft_create" in names and "scft_calc_H" in names
    for n in names:
        assert hasattr(lib, n), n


def _c_layout(tmp_path, struct, fields):
    """sizeof and offsetof as gcc lays out the header's struct."""
    import subprocess
    inc = os.path.join(os.path.dirname(scft_amd.PKG_DIR), "include")
    src = tmp_path / "probe.c"
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "scft.h"', 'int main(void){',
             f'printf("%zu\\n", sizeof({struct}));']
    lines += [f'printf("%zu\\n", offsetof({struct}, {f}));' for f in fields]
    lines += ['return 0;}']
    src.write_text("\n".join(lines))
    exe = tmp_path / "probe"
    subprocess.run(["gcc", "-I", inc, str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split()
    return int(out[0]), [int(x) for x in out[1:]]


@pytest.mark.parametrize("cls,struct", [(scft_amd.ScftParams, "scft_params"),
                                        (scft_amd.SysInfo, "scft_sysinfo")])
def test_struct_layout_matches_header(tmp_path, cls, struct):
    fields = [f for f, _ in cls._fields_]
    size, offs = _c_layout(tmp_path, struct, fields)
    assert ctypes.sizeof(cls) == size
    assert [getattr(cls, f).offset for f in fields] == offs


@pytest.mark.parametrize("name,dim", [("bcp_dist-0.300.input", 2), ("bcp_angle-45.00.input", 2),
                                      ("bcp_dist-0.300.input", 3)])
def test_reader_matches_oracle(golden_dir, name, dim):
    path = os.path.join(golden_dir, name)
    a = scft_amd.read_input(path, dim)
    b = pyoracle.read_input(path, dim)
    for f, _ in scft_amd.ScftParams._fields_:
        va, vb = getattr(a, f), getattr(b, f)
        if hasattr(va, "__len__"):
            va, vb = list(map(lambda x: list(x) if hasattr(x, "__len__") else x, va)), \
                list(map(lambda x: list(x) if hasattr(x, "__len__") else x, vb))
        assert va == vb, f


def test_reader_rejects_bad_input(tmp_path):
    p = tmp_path / "bcp.input"
    p.write_text("# Main parameters\n30 30 Na Nb\n")
    lib = scft_amd.load_library()
    out = scft_amd.ScftParams()
    assert lib.scft_read_input(str(p).encode(), 2, ctypes.byref(out)) == scft_amd.SCFT_EINVAL
    assert lib.scft_read_input(b"/nonexistent/bcp.input", 2, ctypes.byref(out)) == scft_amd.SCFT_EIO
    assert lib.scft_read_input(str(p).encode(), 4, ctypes.byref(out)) == scft_amd.SCFT_EINVAL


def test_create_without_device_fails_cleanly():
    if os.path.exists("/dev/kfd"):
        pytest.skip("a GPU driver is present")
    lib = scft_amd.load_library()
    p = scft_amd.Workload(grid=(8, 8, 8), box_rg=(1.0, 1.0, 1.0), N=4).params()
    h = ctypes.c_void_p()
    rc = lib.scft_create(ctypes.byref(p), 1, 0, 0, ctypes.byref(h))
    assert rc in (scft_amd.SCFT_EDEVICE, scft_amd.SCFT_EINVAL)
    assert not h.value
