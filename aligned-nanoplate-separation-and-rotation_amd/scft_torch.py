"""libtorch autograd wrapper of the plate energy (torch_ext/scft_torch.cpp).

    import torch                      # first: the process then holds torch's
    import scft_amd, scft_torch       # HIP runtime only, shared by the library
    ctx = scft_amd.Context(params)
    geom = torch.tensor([D, theta], dtype=torch.float64, requires_grad=True)
    H = scft_torch.plate_energy(ctx, geom, max_iter=5000, tol=1e-13)
    H.backward()                      # geom.grad = (dH/dD, dH/dtheta)

build() compiles the extension in-tree (torch.utils.cpp_extension, host
C++ only; it links libscft_amd.so); load() imports the built module and
fails loudly when it is missing.
"""
from __future__ import annotations

import glob
import importlib.machinery
import importlib.util
import os

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "torch_ext", "scft_torch.cpp")
BUILD = os.path.join(HERE, "torch_ext", "build")
NAME = "scft_torch_ext"
_mod = None


def build(verbose: bool = False):
    """Compile the extension into torch_ext/build (no GPU needed)."""
    from torch.utils.cpp_extension import load as _load
    os.makedirs(BUILD, exist_ok=True)
    return _load(name=NAME, sources=[SRC], build_directory=BUILD, verbose=verbose, with_cuda=False,
                 extra_include_paths=[os.path.join(HERE, "..", "include")],
                 extra_cflags=["-O2"],
                 extra_ldflags=[f"-L{HERE}", "-lscft_amd", f"-Wl,-rpath,{HERE}"])


def load():
    """The built extension module (import torch before calling this)."""
    global _mod
    if _mod is None:
        import torch  # noqa: F401  (the extension links libtorch)
        found = glob.glob(os.path.join(BUILD, NAME + "*.so"))
        if not found:
            raise RuntimeError(f"{NAME} is not built: run scft_torch.build() (or __graft_entry__.build())")
        loader = importlib.machinery.ExtensionFileLoader(NAME, found[0])
        spec = importlib.util.spec_from_file_location(NAME, found[0], loader=loader)
        _mod = importlib.util.module_from_spec(spec)
        loader.exec_module(_mod)
    return _mod


def plate_energy(ctx, geom, sys: int = 0, max_iter: int = 0, tol: float = 0.0):
    """Re H of system `sys` of the scft_amd.Context `ctx` with its two plates
    at geom = (D [b units], theta [rad]); with max_iter > 0 the fields are
    relaxed first (chunks of 100 iterations until |dH/H| < tol).  The
    backward pass gives (dH/dD, dH/dtheta)."""
    m = load()
    Lx = float(ctx.params[sys].L[0])
    return m.plate_energy(geom, int(ctx.h.value), sys, Lx, max_iter, tol)
