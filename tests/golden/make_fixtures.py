"""Writes the golden fixtures under tests/golden/.

Run in a container that has the reference checkout at /root/reference (the
input files and LOGs are read as text; nothing of the reference is imported
or executed):

    python tests/golden/make_fixtures.py

Outputs
  log_known_answers.json   init-time scalars printed in every published LOG
                           (V_1_exp_np, nD, np_frac, phiG, Vseg, nD*N) and
                           the converged H of extracted_hamiltonians.dat,
                           with the parameters of each run's bcp.input
  bcp_dist-0.300.input     two published input files, kept verbatim as
  bcp_angle-45.00.input    parser fixtures (data files)
  fft_vectors.npz          numpy.fft (pocketfft) golden vectors, forward
                           e^{-i} and backward e^{+i}, unnormalised
  log_templates.json       the line sequence of two published LOGs with every
                           number replaced by '#' (logfmt.log_template)
  log_last_rows.json       the last three main-loop lines of every published
                           LOG: iter, H, -log(Qd), -log(Qg), err

    python tests/golden/make_fixtures.py --templates   # the two LOG-line fixtures only
"""
import glob
import json
import os
import re
import shutil

import numpy as np

from logfmt import log_template

REF = "/root/reference"
HERE = os.path.dirname(os.path.abspath(__file__))


def parse_log(path):
    text = open(path).read()

    def last(pat):
        m = re.findall(pat, text)
        return float(m[-1]) if m else None

    return {
        "V_1_exp_np": last(r"V_1_exp_np = (\S+)"),
        "nD": last(r"\bnD = (\S+)"),
        "np_frac": last(r"Volume fraction of particles excluding grafts\s*=\s*(\S+)"),
        "phiG": last(r"Volume fraction of grafts\s*=\s*(\S+)"),
        "Vseg": last(r"Total V_segment actual: (\S+)"),
        "nDN": last(r"nD \* N = (\S+)"),
        "int_rho": last(r"integ\(rhoda \+ rhodb\) = (\S+)"),
        "ng_per_np": last(r"# of grafts per np = (\S+)"),
        "V": last(r"\bV = (\S+)"),
    }


def write_log_templates():
    runs = {}
    for sweep, run in (("angle-rotation", "angle-05.00"), ("particle-separation", "dist-0.300")):
        runs[run] = log_template(open(os.path.join(REF, sweep, run, "LOG")).read())
    with open(os.path.join(HERE, "log_templates.json"), "w") as f:
        json.dump(runs, f, indent=1)


def last_rows(path, k=3):
    """The last k main-loop lines of a LOG: iter, H, -log(Qd), -log(Qg), err."""
    pat = re.compile(r"Iter: (\d+), H=(\S+), -log\(Qd\)=(\S+), -log\(Qg\)=(\S+), err=(\S+)")
    rows = [[float(x) for x in m.groups()] for m in map(pat.match, open(path)) if m]
    return rows[-k:]


def write_last_rows():
    out = {}
    for sweep in ("particle-separation", "angle-rotation"):
        for d in sorted(glob.glob(os.path.join(REF, sweep, "*-*"))):
            if os.path.isdir(d):
                out[os.path.basename(d)] = last_rows(os.path.join(d, "LOG"))
    with open(os.path.join(HERE, "log_last_rows.json"), "w") as f:
        json.dump(out, f, indent=1)


def parse_hamiltonians(path):
    rows = []
    for line in open(path):
        parts = line.split()
        if len(parts) >= 2:
            try:
                rows.append([float(parts[0]), float(parts[1])])
            except ValueError:
                pass
    return rows


def main():
    runs = []
    for sweep in ("particle-separation", "angle-rotation"):
        for d in sorted(glob.glob(os.path.join(REF, sweep, "*-*"))):
            if not os.path.isdir(d):
                continue
            runs.append({
                "sweep": sweep,
                "run": os.path.basename(d),
                "input": open(os.path.join(d, "bcp.input")).read(),
                "log": parse_log(os.path.join(d, "LOG")),
            })
    out = {
        "source": "LOG files and extracted_hamiltonians.dat of the published 2D runs",
        "runs": runs,
        "H_separation": parse_hamiltonians(
            os.path.join(REF, "particle-separation", "extracted_hamiltonians.dat")),
        "H_angle": parse_hamiltonians(
            os.path.join(REF, "angle-rotation", "extracted_hamiltonians.dat")),
    }
    with open(os.path.join(HERE, "log_known_answers.json"), "w") as f:
        json.dump(out, f, indent=1)
    shutil.copy(os.path.join(REF, "particle-separation", "dist-0.300", "bcp.input"),
                os.path.join(HERE, "bcp_dist-0.300.input"))
    shutil.copy(os.path.join(REF, "angle-rotation", "angle-45.00", "bcp.input"),
                os.path.join(HERE, "bcp_angle-45.00.input"))

    rng = np.random.default_rng(20240601)
    vec = {}
    for n in (2, 3, 4, 5, 6, 7, 8, 12, 14, 16, 32, 56, 64, 72, 96, 128, 192):
        x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
        vec[f"x{n}"] = x
        vec[f"F{n}"] = np.fft.fft(x)
        vec[f"B{n}"] = np.fft.ifft(x) * n
    for shape in ((4, 6, 8), (16, 8, 8), (8, 12, 14), (1, 14, 10)):
        key = "x".join(map(str, shape))
        a = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
        vec[f"a{key}"] = a
        vec[f"A{key}"] = np.fft.fftn(a)
    np.savez_compressed(os.path.join(HERE, "fft_vectors.npz"), **vec)
    write_log_templates()
    write_last_rows()
    print("wrote", len(runs), "runs and", len(vec), "fft arrays")


if __name__ == "__main__":
    import sys
    if "--templates" in sys.argv:
        write_log_templates()
        write_last_rows()
    else:
        main()
