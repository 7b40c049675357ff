"""Masking of LOG text for structural comparison (fixture writer and tests)."""
import re

_NUM = re.compile(r"[-+]?(?:\d+\.\d*|\.\d+|\d+)(?:[eE][-+]?\d+)?")


def log_template(text):
    """The LOG's line sequence with numbers replaced by '#': per-rank lines
    ("Processor k ...") and the resume-file notices ("Read ...", one per rank
    and file) dropped, each run of main-loop lines ("Iter: ...") collapsed
    to one, the finishing date replaced by <date>."""
    out = []
    for line in text.splitlines():
        if line.startswith("Processor ") or line.startswith("Read "):
            continue
        if line.startswith("The program finished at "):
            out.append("The program finished at <date>")
            continue
        t = _NUM.sub("#", line.rstrip())
        if t.startswith("Iter: ") and out and out[-1] == t:
            continue
        out.append(t)
    return out
