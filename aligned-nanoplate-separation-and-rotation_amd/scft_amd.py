"""ctypes binding of libscft_amd.so (the C ABI declared in include/scft.h) and
builders for the nanoplate workloads.

The library is the product: HIP kernels for gfx950 behind a plain C ABI.  This
module only moves numpy arrays across that ABI.  It raises if the library is
missing; there is no CPU fallback.

Parameter mapping from polymer units (Rg, chi N, C, kappa N, dt) to the
reference's b units (code/initialize.cpp:55 sets Rg^2 = (N-1)/6 and
code/initialize.cpp:403 uses exp(-k^2/6) bonds):
    Rg = sqrt((N-1)/6) b,  chi = chiN/N,  kappa = kappaN/N,  rho0 = C N / Rg^3,
    lam_imag = lam_real = dt / (rho0 N).
"""
from __future__ import annotations

import ctypes
import math
import os
from dataclasses import dataclass

import numpy as np

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
REPO_DIR = os.path.dirname(PKG_DIR)
LIB_PATH = os.path.join(PKG_DIR, "libscft_amd.so")

SCFT_OK = 0
SCFT_EINVAL = -1
SCFT_ENOMEM = -2
SCFT_EDEVICE = -3
SCFT_ENAN = -4
SCFT_EUNSUPPORTED = -5
SCFT_EIO = -6

PREC_FP64 = 0
PREC_FP32 = 1
OPT_CHECKPOINT = 0x100   # OR into the precision: sqrt(N)-checkpointed propagators

NP_ROD = 2
NP_SQUARE = 3


class ScftParams(ctypes.Structure):
    """Mirror of scft_params (include/scft.h); bcp.input order,
    code/read_input.cpp:31-236."""

    _fields_ = [
        ("dim", ctypes.c_int),
        ("Nda", ctypes.c_int), ("Ndb", ctypes.c_int),
        ("phiH", ctypes.c_double), ("Nah", ctypes.c_int),
        ("a_smear", ctypes.c_double), ("rho0", ctypes.c_double),
        ("chi_ab", ctypes.c_double), ("chi_bc", ctypes.c_double),
        ("chi_ac", ctypes.c_double), ("kappa", ctypes.c_double),
        ("ic_flag", ctypes.c_int * 2), ("ic_pre", ctypes.c_double * 2),
        ("ic_dir", ctypes.c_double * 2), ("ic_period", ctypes.c_double * 2),
        ("Nx", ctypes.c_int * 3), ("L", ctypes.c_double * 3),
        ("lam_imag", ctypes.c_double), ("lam_real", ctypes.c_double),
        ("itermax", ctypes.c_int), ("print_freq", ctypes.c_int),
        ("sample_freq", ctypes.c_int), ("sample_wait", ctypes.c_int),
        ("error_tol", ctypes.c_double),
        ("update_scheme", ctypes.c_int), ("do_CL", ctypes.c_int),
        ("do_film", ctypes.c_int),
        ("wallT", ctypes.c_double), ("wallXi", ctypes.c_double),
        ("top_wall_lamA", ctypes.c_double), ("top_wall_lamB", ctypes.c_double),
        ("bot_wall_lamA", ctypes.c_double), ("bot_wall_lamB", ctypes.c_double),
        ("do_channel", ctypes.c_int), ("hor_dir", ctypes.c_int),
        ("vert_dir", ctypes.c_int),
        ("wall_width", ctypes.c_double), ("channel_width", ctypes.c_double),
        ("channel_xi", ctypes.c_double),
        ("n_exp_nr", ctypes.c_int), ("do_fld_np", ctypes.c_int),
        ("np_type", ctypes.c_int), ("np_chem", ctypes.c_int),
        ("sigma", ctypes.c_double), ("Ng", ctypes.c_int), ("Nu", ctypes.c_int),
        ("np_frac", ctypes.c_double),
        ("L_nr", ctypes.c_double), ("R_nr", ctypes.c_double),
        ("xi_nr", ctypes.c_double),
        ("exp_nr_c", (ctypes.c_double * 3) * 2),
        ("exp_nr_u", (ctypes.c_double * 3) * 2),
        ("seed", ctypes.c_uint64),
    ]

    @property
    def M(self) -> int:
        m = 1
        for d in range(self.dim):
            m *= self.Nx[d]
        return m

    @property
    def N(self) -> int:
        return self.Nda + self.Ndb

    def grid(self) -> tuple:
        return tuple(self.Nx[d] if d < self.dim else 1 for d in range(3))

    def box(self) -> tuple:
        return tuple(self.L[d] if d < self.dim else 1.0 for d in range(3))


class SysInfo(ctypes.Structure):
    _fields_ = [
        ("M", ctypes.c_int), ("N", ctypes.c_int),
        ("V", ctypes.c_double), ("Vf", ctypes.c_double),
        ("dx", ctypes.c_double * 3),
        ("V_1_exp_np", ctypes.c_double), ("np_frac", ctypes.c_double),
        ("phiD", ctypes.c_double), ("phiG", ctypes.c_double),
        ("nP", ctypes.c_double), ("ng_per_np", ctypes.c_double),
        ("nD", ctypes.c_double), ("Rg2", ctypes.c_double),
        ("fD", ctypes.c_double),
    ]


# --------------------------------------------------------------------------
# workloads
# --------------------------------------------------------------------------
@dataclass
class Workload:
    """A nanoplate system in polymer units (lengths in Rg)."""

    grid: tuple = (32, 32, 64)
    box_rg: tuple = (4.0, 4.0, 8.0)
    N: int = 50
    f_A: float = 0.5
    chiN: float = 20.0
    C: float = 50.0
    kappaN: float = 50.0
    plate_side_rg: float = 1.0
    plate_thick_rg: float = 0.2
    plate_xi_rg: float = 0.05
    n_plates: int = 2
    d: float = 0.30          # fractional centre separation along x
    theta_deg: float = 0.0   # rotation of the plate normals about y
    np_type: int = NP_SQUARE
    dt: float = 0.005
    itermax: int = 199
    print_freq: int = 50
    error_tol: float = 1e-10
    update_scheme: int = 1
    do_CL: int = 0
    a_smear: float = 0.2     # b units, as in the published inputs
    seed: int = 12345

    @property
    def rg(self) -> float:
        return math.sqrt((self.N - 1) / 6.0)

    def params(self) -> ScftParams:
        p = ScftParams()
        rg = self.rg
        p.dim = 3 if len(self.grid) == 3 and self.grid[2] > 1 else 2
        p.Nda = int(round(self.f_A * self.N))
        p.Ndb = self.N - p.Nda
        p.phiH = 0.0
        p.Nah = 0
        p.a_smear = self.a_smear
        p.rho0 = self.C * self.N / rg ** 3
        p.chi_ab = self.chiN / self.N
        p.chi_bc = 0.0
        p.chi_ac = 0.0
        p.kappa = self.kappaN / self.N
        p.ic_flag[0], p.ic_flag[1] = 0, 0
        for d in range(3):
            p.Nx[d] = self.grid[d] if d < len(self.grid) else 1
            p.L[d] = self.box_rg[d] * rg if d < len(self.box_rg) else 1.0
        # polymer-unit step dt: with w~ = N w, r~ = r / Rg and H = C H~, the
        # force obeys dH/dw = rho0 dH~/dw~, so w~ <- w~ - dt dH~/dw~ is the
        # reference's w <- w - lam dH/dw with lam = dt / (rho0 N)
        p.lam_imag = self.dt / (p.rho0 * self.N)
        p.lam_real = self.dt / (p.rho0 * self.N)
        p.itermax = self.itermax
        p.print_freq = self.print_freq
        p.sample_freq, p.sample_wait = 50, 0
        p.error_tol = self.error_tol
        p.update_scheme = self.update_scheme
        p.do_CL = self.do_CL
        p.n_exp_nr = self.n_plates
        p.np_type = self.np_type
        p.np_chem = 0
        p.sigma = 0.0
        p.Ng = 0
        p.L_nr = self.plate_thick_rg * rg
        p.R_nr = 0.5 * self.plate_side_rg * rg
        p.xi_nr = self.plate_xi_rg * rg
        th = math.radians(self.theta_deg)
        u = (math.sin(th), 0.0, math.cos(th))
        centres = [0.5 - self.d / 2.0, 0.5 + self.d / 2.0]
        for j in range(2):
            p.exp_nr_c[j][0] = centres[j]
            p.exp_nr_c[j][1] = 0.5
            p.exp_nr_c[j][2] = 0.5
            for d in range(3):
                p.exp_nr_u[j][d] = u[d]
        p.seed = self.seed
        return p


def config(name: str, **over) -> Workload:
    """Workloads of the benchmark configurations (polymer units)."""
    if name == "plain-32x32x64-N50":
        w = Workload(grid=(32, 32, 64), box_rg=(4.0, 4.0, 8.0), N=50, itermax=199)
    elif name == "production-64x64x128-N100":
        w = Workload(grid=(64, 64, 128), box_rg=(4.0, 4.0, 8.0), N=100, itermax=1999)
    elif name == "cl-96x96x192-N100":
        w = Workload(grid=(96, 96, 192), box_rg=(6.0, 6.0, 12.0), N=100, C=5.0,
                     dt=0.001, update_scheme=0, do_CL=1, itermax=9999)
    else:
        raise KeyError(name)
    for k, v in over.items():
        setattr(w, k, v)
    return w


def sweep(name: str, **over) -> list:
    """Workloads of the batched sweeps (one context holds all of them).

    "separation-64": 16 separations d = 0.14, 0.16, ..., 0.44 on the
      production 64 x 64 x 128 grid (N = 100), the headline batch;
    "config3": the 48-system sweep, 33 separations d = 0.14..0.46 (step 0.01)
      at theta = 0 and 15 angles {0..10, 20, 30, 45, 90} deg at d = 0.30,
      x grid cycled 56 / 64 / 72 with Lx = 3.5 / 4.0 / 4.5 Rg, 64 x 128 in y, z.
    """
    out = []
    if name == "separation-64":
        for i in range(16):
            w = config("production-64x64x128-N100", d=round(0.14 + 0.02 * i, 4), seed=1000 + i)
            out.append(w)
    elif name == "config3":
        geoms = [(round(0.14 + 0.01 * i, 4), 0.0) for i in range(33)]
        geoms += [(0.30, float(a)) for a in list(range(11)) + [20, 30, 45, 90]]
        xs = [(56, 3.5), (64, 4.0), (72, 4.5)]
        for i, (d, th) in enumerate(geoms):
            nx, lx = xs[i % 3]
            out.append(Workload(grid=(nx, 64, 128), box_rg=(lx, 4.0, 8.0), N=100, d=d, theta_deg=th,
                                itermax=1999, seed=1000 + i))
    else:
        raise KeyError(name)
    for w in out:
        for k, v in over.items():
            setattr(w, k, v)
    return out


def seeded_fields(p: ScftParams, seed: int, period_rg: float = 4.0,
                  amp: float = 1.0, sigma: float = 0.01):
    """The lamellar start: wAB- = amp cos(2 pi z / period) + N(0, sigma^2);
    w+ and wAB+ = i N(0, sigma^2).  Returns interleaved complex128 arrays
    (w+, wAB+, wAB-) in the reference layout x + Nx (y + Ny z)."""
    nx, ny, nz = p.grid()
    M = nx * ny * nz
    rng = np.random.default_rng(seed)
    rg = math.sqrt((p.N - 1) / 6.0)
    axis = p.dim - 1
    n_along = (nx, ny, nz)[axis]
    coord = np.arange(n_along) * (p.L[axis] / n_along)
    prof = amp * np.cos(2 * np.pi * coord / (period_rg * rg))
    shape = (nz, ny, nx)
    if axis == 2:
        base = np.broadcast_to(prof[:, None, None], shape)
    else:
        base = np.broadcast_to(prof[None, :, None], shape)
    wabm = base.reshape(M) + sigma * rng.standard_normal(M)
    wpl = 1j * sigma * rng.standard_normal(M)
    wabp = 1j * sigma * rng.standard_normal(M)
    return (np.ascontiguousarray(wpl, dtype=np.complex128),
            np.ascontiguousarray(wabp, dtype=np.complex128),
            np.ascontiguousarray(wabm.astype(np.complex128)))


# --------------------------------------------------------------------------
# reference file formats
# --------------------------------------------------------------------------
def _num(x) -> str:
    return repr(float(x)) if isinstance(x, float) else str(x)


def write_input(p: ScftParams, path: str) -> None:
    """Writes `p` as a bcp.input file in the positional layout read_input
    parses (code/read_input.cpp:31-236; the published inputs, e.g.
    separation/dist-0.300/bcp.input, show the section order).  Per-dimension
    lines carry p.dim values; floats are written with full precision."""
    d = p.dim

    def vec(a):
        return " ".join(_num(float(a[k])) for k in range(d))

    def ivec(a):
        return " ".join(str(int(a[k])) for k in range(d))

    L = ["# Main parameters",
         f"{p.Nda} {p.Ndb}  Na, Nb length of each diblock's block",
         f"{_num(p.phiH)} {p.Nah}  phiAH, Nah",
         f"{_num(p.a_smear)}  a, smearing length, 0 for no smearing",
         f"{_num(p.rho0)}  rho0, bulk density",
         "",
         "# Interaction parameters",
         f"{_num(p.chi_ab)}  chiAB #between diblock",
         f"{_num(p.chi_bc)}  chiBC #between part B and the graft",
         f"{_num(p.chi_ac)}  chiAC #between part A and the graft",
         f"{_num(p.kappa)}   kappa, strength of compressibility potential",
         "",
         "#Init field cond flags (0=zero, -1=rand, -2=cosine, -3=sine, -4=factor of 1)"]
    for k in range(2):
        L.append(f"{p.ic_flag[k]} {_num(p.ic_pre[k])} {int(p.ic_dir[k])} {_num(p.ic_period[k])}"
                 f"  flag {k + 1}, prefactor, direction (x=0, y=1, z=2), # periods")
    L += ["",
          "# Simulation parameters",
          f"{ivec(p.Nx)}  Nx, # grid points in each direction",
          f"{vec(p.L)}    L, box dimension in each direction",
          f"{_num(p.lam_imag)} lam_pl, step size on the w+ fields",
          f"{_num(p.lam_real)}  lam_mi, step size on the w- fields",
          f"{p.itermax}    itermax, max. # of iterations",
          f"{p.print_freq}       print_freq",
          f"{p.sample_freq} {p.sample_wait}  sample_freq, sample_wait",
          f"{_num(p.error_tol)}  error_tol",
          f"{p.update_scheme}       Update scheme (0 = Euler, 1=1S)",
          f"{p.do_CL}        do_CL",
          "",
          "# Film parameters",
          f"{p.do_film}        do_film flag, creates walls in dim-1 direction",
          f"{_num(p.wallT)} {_num(p.wallXi)}  Wall thickness, interfacial width",
          f"{_num(p.top_wall_lamA)} {_num(p.top_wall_lamB)}  Top wall interaction parameters, lambda_A, lambda_B",
          f"{_num(p.bot_wall_lamA)} {_num(p.bot_wall_lamB)}  Bottom wall interaction parameters, lambda_A, lambda_B",
          "",
          "# Channel wall parameters",
          f"{p.do_channel} {p.hor_dir} {p.vert_dir}          do_channel flag",
          f"{_num(p.wall_width)} {_num(p.channel_width)} {_num(p.channel_xi)}    wall_width channel_width xi",
          "",
          "# Nanoparticle parameters",
          f"{p.n_exp_nr}       n_exp_nr, number of explicit nanoparticles",
          f"{p.do_fld_np}                do_fld_np, 1 for field-based particles, 0 for no",
          f"{p.np_type}                np_type, 1=sphere, 2=nanorod",
          f"{p.np_chem}                np_chem, 0=neutral, 1=A",
          f"{_num(p.sigma)}          sigma",
          f"{p.Ng}             Ng",
          f"{p.Nu}               Nu, number of points per pi for spherical integration",
          f"{_num(p.np_frac)}              np_frac, volume fraction of field-based nps",
          f"{_num(p.L_nr)} {_num(p.R_nr)} {_num(p.xi_nr)}      L_nr, R_nr, xi_nr for all nps"]
    # read_input always reads two plate blocks (code/read_input.cpp:205)
    for j in range(2):
        L.append(f"{vec(p.exp_nr_c[j])}   Nanorod center x, y, z for np {j + 1}")
        L.append(f"{vec(p.exp_nr_u[j])}  Nanorod orientation vector x, y, z for np {j + 1}")
    with open(path, "w") as f:
        f.write("\n".join(L) + "\n")


def _write_res_grid(path: str, a, dim: int, Nx, L) -> None:
    a = np.asarray(a, dtype=np.complex128)
    nx = Nx[0]
    ny = Nx[1] if dim > 1 else 1
    M = int(np.prod(Nx[:dim]))
    if a.size != M:
        raise ValueError(f"{path}: {a.size} values for a grid of {M} points")
    dx = [L[k] / Nx[k] for k in range(dim)]
    out = []
    for i in range(M):
        n = (i % nx, (i // nx) % ny, i // (nx * ny))
        c = " ".join(repr(float(n[k] * dx[k])) for k in range(dim))
        out.append(f"{c} {float(a[i].real)!r} {float(a[i].imag)!r}")
        if dim == 2 and n[0] == nx - 1:
            out.append("")
    with open(path, "w") as f:
        f.write("\n".join(out) + "\n")


def write_res(path: str, a, p: ScftParams) -> None:
    """Writes a start field as a .res file for read_one_resume_file
    (code/io_utils.cpp:137-177): one point per line, the `dim` coordinates
    n dx then Re Im, x fastest, a blank line after each x row in 2D."""
    _write_res_grid(path, a, p.dim, [p.Nx[k] for k in range(3)], [p.L[k] for k in range(3)])


def bin_to_res(bin_path: str, res_path: str) -> None:
    """Turns a field dump (<name>.p0.bin of write_data_bin) into the .res
    start file read_resume_files reads (code/io_utils.cpp:183-240), so a run
    resumes from the dumps of an earlier one."""
    dim, nx, L, _, data = read_bin(bin_path)
    _write_res_grid(res_path, data, dim, nx, L)


def read_bin(path: str):
    """Reads a <name>.p0.bin dump (code/io_utils.cpp:375-397): returns
    (dim, Nx, L, nprocs, data)."""
    with open(path, "rb") as f:
        raw = f.read()
    dim = int(np.frombuffer(raw, np.int32, 1, 0)[0])
    off = 4
    nx = np.frombuffer(raw, np.int32, dim, off).tolist()
    off += 4 * dim
    L = np.frombuffer(raw, np.float64, dim, off).tolist()
    off += 8 * dim
    nprocs = int(np.frombuffer(raw, np.int32, 1, off)[0])
    ml = int(np.frombuffer(raw, np.int32, 1, off + 4)[0])
    data = np.frombuffer(raw, np.complex128, ml, off + 8).copy()
    return dim, nx, L, nprocs, data


def read_data_dat(path: str) -> np.ndarray:
    """data.dat rows (code/simulate.cpp:132-135): iter, Re H, Im H,
    Re -log Qd, Im -log Qd, -log Qha (2), -log Qp + smwp_min (2), error."""
    rows = [list(map(float, l.split())) for l in open(path) if l.strip()]
    return np.array(rows, dtype=np.float64).reshape(-1, 10)


# --------------------------------------------------------------------------
# library binding
# --------------------------------------------------------------------------
_lib = None


def load_library(path: str | None = None):
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise RuntimeError(f"libscft_amd.so not built at {p}: run __graft_entry__.build()")
    lib = ctypes.CDLL(p)
    P = ctypes.POINTER
    dp = P(ctypes.c_double)
    vp = ctypes.c_void_p
    sig = {
        "scft_read_input": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_int, P(ScftParams)]),
        "scft_create": (ctypes.c_int, [P(ScftParams), ctypes.c_int, ctypes.c_int,
                                       ctypes.c_int, P(vp)]),
        "scft_destroy": (ctypes.c_int, [vp]),
        "scft_get_info": (ctypes.c_int, [vp, ctypes.c_int, P(SysInfo)]),
        "scft_set_fields": (ctypes.c_int, [vp, ctypes.c_int, dp, dp, dp]),
        "scft_get_fields": (ctypes.c_int, [vp, ctypes.c_int, dp, dp, dp]),
        "scft_get_density": (ctypes.c_int, [vp, ctypes.c_int, dp, dp]),
        "scft_get_smeared": (ctypes.c_int, [vp, ctypes.c_int, dp, dp]),
        "scft_get_cavity": (ctypes.c_int, [vp, ctypes.c_int, dp]),
        "scft_get_Q": (ctypes.c_int, [vp, ctypes.c_int, dp]),
        "scft_fft": (ctypes.c_int, [vp, ctypes.c_int, dp, ctypes.c_int]),
        "scft_get_noise": (ctypes.c_int, [vp, ctypes.c_int, dp, dp, dp]),
        "scft_get_averages": (ctypes.c_int, [vp, ctypes.c_int, dp, dp, dp]),
        "scft_get_array": (ctypes.c_int, [vp, ctypes.c_int, ctypes.c_char_p, dp]),
        "scft_set_array": (ctypes.c_int, [vp, ctypes.c_int, ctypes.c_char_p, dp]),
        "scft_get_Qga": (ctypes.c_int, [vp, ctypes.c_int, dp]),
        "scft_set_plates": (ctypes.c_int, [vp, ctypes.c_int, dp, dp]),
        "scft_grad_plates": (ctypes.c_int, [vp, ctypes.c_int, dp, dp]),
        "scft_calc_poly_density": (ctypes.c_int, [vp]),
        "scft_update_1s": (ctypes.c_int, [vp]),
        "scft_update_euler": (ctypes.c_int, [vp]),
        "scft_calc_H": (ctypes.c_int, [vp, dp]),
        "scft_iterate": (ctypes.c_int, [vp, ctypes.c_int]),
        "scft_synchronize": (ctypes.c_int, [vp]),
        "scft_describe": (ctypes.c_int, [vp, ctypes.c_char_p, ctypes.c_int]),
        "scft_time_kernel": (ctypes.c_int, [vp, ctypes.c_int, ctypes.c_int, dp, dp]),
        "scft_last_error": (ctypes.c_char_p, []),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    _lib = lib
    return lib


def exported_symbols() -> list:
    """Names declared with extern "C" linkage in include/scft.h."""
    import re
    text = open(os.path.join(REPO_DIR, "include", "scft.h")).read()
    return sorted(set(re.findall(r"^\s*(?:int|const char\*)\s+(scft_\w+)\(", text, re.M)))


class ScftError(RuntimeError):
    def __init__(self, code: int, what: str):
        msg = _lib.scft_last_error().decode() if _lib is not None else ""
        super().__init__(f"{what} failed with {code}: {msg}")
        self.code = code


def _dptr(a: np.ndarray):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


class Context:
    """One scft_ctx: n systems on one GPU."""

    def __init__(self, params, device: int = 0, precision: int = PREC_FP64):
        self.lib = load_library()
        if isinstance(params, ScftParams):
            params = [params]
        self.params = list(params)
        arr = (ScftParams * len(self.params))(*self.params)
        h = ctypes.c_void_p()
        rc = self.lib.scft_create(arr, len(self.params), device, precision, ctypes.byref(h))
        if rc != SCFT_OK:
            raise ScftError(rc, "scft_create")
        self.h = h
        self.precision = precision

    def _check(self, rc, what):
        if rc != SCFT_OK:
            raise ScftError(rc, what)

    def close(self):
        if getattr(self, "h", None):
            self.lib.scft_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def M(self, sys=0):
        return self.params[sys].M

    def info(self, sys=0) -> SysInfo:
        s = SysInfo()
        self._check(self.lib.scft_get_info(self.h, sys, ctypes.byref(s)), "scft_get_info")
        return s

    def _sized(self, a, n, what):
        if a.size != n:
            raise ValueError(f"{what}: expected {n} values, got {a.size}")
        return a

    def set_fields(self, wpl, wabp, wabm, sys=0):
        a = [self._sized(np.ascontiguousarray(x, dtype=np.complex128), self.M(sys), "set_fields")
             for x in (wpl, wabp, wabm)]
        self._check(self.lib.scft_set_fields(self.h, sys, *[_dptr(x) for x in a]),
                    "scft_set_fields")

    def get_fields(self, sys=0):
        out = [np.empty(self.M(sys), np.complex128) for _ in range(3)]
        self._check(self.lib.scft_get_fields(self.h, sys, *[_dptr(x) for x in out]),
                    "scft_get_fields")
        return out

    def get_density(self, sys=0):
        out = [np.empty(self.M(sys), np.complex128) for _ in range(2)]
        self._check(self.lib.scft_get_density(self.h, sys, *[_dptr(x) for x in out]),
                    "scft_get_density")
        return out

    def get_smeared(self, sys=0):
        out = [np.empty(self.M(sys), np.complex128) for _ in range(2)]
        self._check(self.lib.scft_get_smeared(self.h, sys, *[_dptr(x) for x in out]),
                    "scft_get_smeared")
        return out

    def get_cavity(self, sys=0):
        out = np.empty(self.M(sys), np.complex128)
        self._check(self.lib.scft_get_cavity(self.h, sys, _dptr(out)), "scft_get_cavity")
        return out

    def get_Q(self, sys=0) -> complex:
        out = np.empty(2)
        self._check(self.lib.scft_get_Q(self.h, sys, _dptr(out)), "scft_get_Q")
        return complex(out[0], out[1])

    def set_plates(self, centres, us, sys=0):
        """centres, us: (n_plates, 3) fractional centres and orientations."""
        n = 3 * self.params[sys].n_exp_nr
        c = self._sized(np.ascontiguousarray(np.asarray(centres, dtype=np.float64).reshape(-1)), n, "centres")
        u = self._sized(np.ascontiguousarray(np.asarray(us, dtype=np.float64).reshape(-1)), n, "orientations")
        self._check(self.lib.scft_set_plates(self.h, sys, _dptr(c), _dptr(u)), "scft_set_plates")

    def set_geometry(self, d: float, theta_deg: float, sys=0):
        """Two plates at fractional separation d along x, normals rotated by
        theta about y (the Workload convention)."""
        th = math.radians(theta_deg)
        u = [math.sin(th), 0.0, math.cos(th)]
        self.set_plates([[0.5 - d / 2, 0.5, 0.5], [0.5 + d / 2, 0.5, 0.5]], [u, u], sys=sys)

    def grad_plates(self, axis=(0.0, 1.0, 0.0), sys=0):
        """(n, 4) complex: dH/dc_x, dH/dc_y, dH/dc_z (b units), dH/dtheta per plate."""
        a = self._sized(np.ascontiguousarray(np.asarray(axis, dtype=np.float64).reshape(-1)), 3, "axis")
        out = np.empty(16)
        self._check(self.lib.scft_grad_plates(self.h, sys, _dptr(a), _dptr(out)), "scft_grad_plates")
        g = out[0::2] + 1j * out[1::2]
        return g.reshape(2, 4)

    def plate_forces(self, sys=0):
        """dH/dD and dH/dtheta for the symmetric pair of the Workload
        convention: D = physical centre separation along x (b units), both
        plates rotating about y."""
        g = self.grad_plates((0.0, 1.0, 0.0), sys=sys)
        return 0.5 * (g[1, 0] - g[0, 0]), g[0, 3] + g[1, 3]

    def get_noise(self, sys=0):
        out = [np.empty(self.M(sys), np.complex128) for _ in range(3)]
        self._check(self.lib.scft_get_noise(self.h, sys, *[_dptr(x) for x in out]), "scft_get_noise")
        return out

    def get_averages(self, sys=0):
        """(avg rhoA, avg rhoB, n_samples) of the complex-Langevin run."""
        a = np.empty(self.M(sys), np.complex128)
        b = np.empty(self.M(sys), np.complex128)
        n = np.zeros(1)
        self._check(self.lib.scft_get_averages(self.h, sys, _dptr(a), _dptr(b), _dptr(n)), "scft_get_averages")
        return a, b, float(n[0])

    def array(self, name: str, sys=0):
        """A named complex array of M values (fields, densities, graft sites)."""
        out = np.empty(self.M(sys), np.complex128)
        self._check(self.lib.scft_get_array(self.h, sys, name.encode(), _dptr(out)), "scft_get_array")
        return out

    def set_array(self, name: str, a, sys=0):
        x = np.ascontiguousarray(np.asarray(a, dtype=np.complex128))
        if x.size != self.M(sys):
            raise ValueError(f"{name}: expected {self.M(sys)} values, got {x.size}")
        self._check(self.lib.scft_set_array(self.h, sys, name.encode(), _dptr(x)), "scft_set_array")

    def get_Qga(self, sys=0) -> complex:
        out = np.empty(2)
        self._check(self.lib.scft_get_Qga(self.h, sys, _dptr(out)), "scft_get_Qga")
        return complex(out[0], out[1])

    def fft(self, a, direction: int, sys=0):
        x = self._sized(np.array(a, dtype=np.complex128, copy=True, order="C"), self.M(sys), "fft")
        self._check(self.lib.scft_fft(self.h, sys, _dptr(x), direction), "scft_fft")
        return x

    def calc_poly_density(self):
        self._check(self.lib.scft_calc_poly_density(self.h), "scft_calc_poly_density")

    def update_1s(self):
        self._check(self.lib.scft_update_1s(self.h), "scft_update_1s")

    def update_euler(self):
        self._check(self.lib.scft_update_euler(self.h), "scft_update_euler")

    def calc_H(self):
        out = np.empty(2 * len(self.params))
        self._check(self.lib.scft_calc_H(self.h, _dptr(out)), "scft_calc_H")
        return out[0::2] + 1j * out[1::2]

    def iterate(self, n: int):
        self._check(self.lib.scft_iterate(self.h, n), "scft_iterate")

    def describe(self) -> str:
        buf = ctypes.create_string_buffer(8192)
        self._check(self.lib.scft_describe(self.h, buf, len(buf)), "scft_describe")
        return buf.value.decode()

    def synchronize(self):
        self._check(self.lib.scft_synchronize(self.h), "scft_synchronize")

    def time_kernel(self, which: int, reps: int = 20):
        """(average ms, algorithmic bytes) of one launch, see scft_time_kernel."""
        ms = ctypes.c_double()
        b = ctypes.c_double()
        self._check(self.lib.scft_time_kernel(self.h, which, reps, ctypes.byref(ms), ctypes.byref(b)),
                    "scft_time_kernel")
        return ms.value, b.value


def read_input(path: str, dim: int) -> ScftParams:
    lib = load_library()
    p = ScftParams()
    rc = lib.scft_read_input(path.encode(), dim, ctypes.byref(p))
    if rc != SCFT_OK:
        raise ScftError(rc, "scft_read_input")
    return p
