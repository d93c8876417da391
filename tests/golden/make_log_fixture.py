"""Extract the numeric golden values of the reference's shipped run transcript
(src_v2026_2_1/vanadis_log.txt) into vanadis_log_default.json.

Run in a container that has the reference tree:
    python tests/golden/make_log_fixture.py /root/reference/src_v2026_2_1/vanadis_log.txt
The JSON holds data only: per-step time, BiCG iterations, recursive residual,
mass, the diagonal-dominance flag/factor and the mesh sizes.
"""
import json
import os
import re
import sys


def parse(path):
    txt = open(path, encoding="latin-1").read().replace("\r", "")
    steps = []
    cur = None
    for line in txt.splitlines():
        m = re.search(r"nr of elements:\s+(\d+)\s+nodes:\s+(\d+)\s+time:\s+(\d+)", line)
        if m:
            cur = {"elements": int(m.group(1)), "nodes": int(m.group(2)), "t": int(m.group(3))}
            continue
        if cur is None:
            continue
        m = re.search(r"Diagonally dominant matrix \[T/F\]\s+([TF])", line)
        if m:
            cur["diag_dominant"] = m.group(1) == "T"
        m = re.search(r"wspolczynnik dominacji\s+([0-9.Ee+-]+)", line)
        if m:
            cur["ddf"] = float(m.group(1))
            cur["ddf_text"] = m.group(1)
        m = re.search(r"number of iterations:\s+(\d+).*residual norm:\s+([0-9.Ee+-]+)", line)
        if m:
            cur["iters"] = int(m.group(1))
            cur["residual"] = float(m.group(2))
        m = re.search(r"czas=\s+(\d+)\s+masa=\s+([0-9.Ee+-]+)", line)
        if m:
            cur["mass"] = float(m.group(2))
            cur["mass_text"] = m.group(2)
            steps.append(cur)
            cur = None
    return steps


if __name__ == "__main__":
    src = sys.argv[1] if len(sys.argv) > 1 else "/root/reference/src_v2026_2_1/vanadis_log.txt"
    steps = parse(src)
    out = {"source": "src_v2026_2_1/vanadis_log.txt", "config": {
        "Q": 1.0, "K": 40.0, "P": 7e-4, "V": [0.0, -5.0, 0.0], "alpha": 0.01, "dt": 100,
        "nax": 7, "nay": 20, "naz": 20, "tol": 1e-8, "itmax": 1000}, "steps": steps}
    dst = os.path.join(os.path.dirname(os.path.abspath(__file__)), "vanadis_log_default.json")
    with open(dst, "w") as f:
        json.dump(out, f, indent=1)
    print(f"{len(steps)} steps -> {dst}")
