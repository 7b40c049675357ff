"""Writes the bcp.input of one published run (from log_known_answers.json)
into a directory, with the run's published converged H beside it:

    python tests/golden/published_input.py angle-05.00 /tmp/run
    cd /tmp/run && scft_run --dim 2 --no-bin

The published runs restarted from converged .res fields that the snapshot
does not hold, so a run from this input starts from the input's own
init_fields flags and reaches the saddle point by itself.
"""
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))


def published(run):
    d = json.load(open(os.path.join(HERE, "log_known_answers.json")))
    r = {x["run"]: x for x in d["runs"]}[run]
    H = dict((f"angle-{a:05.2f}", h) for h, a in d["H_angle"])
    H.update((f"dist-{x:.3f}", h) for h, x in d["H_separation"])
    return r["input"], H.get(run)


if __name__ == "__main__":
    text, h = published(sys.argv[1])
    os.makedirs(sys.argv[2], exist_ok=True)
    open(os.path.join(sys.argv[2], "bcp.input"), "w").write(text)
    print("published H", h)
