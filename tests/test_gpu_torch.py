"""The libtorch autograd wrapper of the plate energy (torch_ext/scft_torch.cpp).

The GPU check runs in a child process (tests/gpu_torch_child.py) that imports
torch before the library, so that it holds one HIP runtime; this pytest
process has already loaded the library's own runtime.  The CPU checks load
the extension and exercise its error path.
"""
import json
import os
import subprocess
import sys

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(HERE, "..", "aligned-nanoplate-separation-and-rotation_amd")


def _built():
    import glob
    return bool(glob.glob(os.path.join(PKG, "torch_ext", "build", "scft_torch_ext*.so")))


@pytest.mark.skipif(not _built(), reason="the torch extension is not built (__graft_entry__.build())")
def test_extension_loads_and_rejects_null_context():
    import torch
    import scft_torch
    m = scft_torch.load()
    assert hasattr(m, "plate_energy")
    with pytest.raises(RuntimeError):
        m.plate_energy(torch.zeros(2, dtype=torch.float64), 0, 0, 1.0, 0, 0.0)


@pytest.mark.gpu
def test_plate_energy_autograd_on_gpu():
    assert _built(), "the torch extension is not built"
    print("torch child", flush=True)
    r = subprocess.run([sys.executable, "-u", os.path.join(HERE, "gpu_torch_child.py")], capture_output=True,
                       text=True, timeout=300)
    lines = [l for l in r.stdout.splitlines() if l.startswith("TORCH_RESULTS ")]
    print(r.stdout[-3000:], r.stderr[-3000:], flush=True)
    assert lines, "child printed no results"
    res = json.loads(lines[-1][len("TORCH_RESULTS "):])
    assert r.returncode == 0, res
    assert res["gradcheck_fixed_fields"]
    assert len(res["hip_runtimes"]) == 1
