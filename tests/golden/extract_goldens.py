"""Writes finch_goldens.json: known-answer vectors taken from Finch.jl's own
golden files and tests (inputs and expected outputs only).

Run in a container where the reference tree is mounted:
    python tests/golden/extract_goldens.py [/root/reference]
The JSON it writes is committed; tests read only the JSON.

Sources (paths relative to the Finch.jl tree):
  * test/reference64/typical/typical_spmv_csc.txt:1-17,92 — CSC SpMV;
  * test/reference64/typical/typical_merge_{twofinger,leadfollow,gallop}.txt
    (last line) with inputs test/suites/typical_tests.jl:102-111 — merge;
  * test/reference64/parallel/parallel_spmv_atomics.txt:117 with inputs
    test/suites/parallel_tests.jl:414-416 — A^T x on CSC buffers;
  * test/suites/issue_tests.jl:388-404 — 0-based (PlusOneVector) buffers;
  * test/reference64/representation/Dense{SparseList}_representation.txt —
    dense -> Dense(SparseList(Element)) ptr/idx/val.
"""
from __future__ import annotations

import ast
import json
import os
import re
import sys

HERE = os.path.dirname(os.path.abspath(__file__))


def julia_matrix(text: str):
    """Parses a Julia matrix literal like `[0 0 3.3; 1.1 0 0]` (or `Bool[...]`)."""
    body = text[text.index("[") + 1: text.rindex("]")]
    rows = [r.split() for r in body.split(";")]
    conv = lambda t: float(t)
    return [[conv(t) for t in r] for r in rows]


def julia_vector(text: str):
    text = text.strip()
    text = re.sub(r"^[A-Za-z0-9]*\[", "[", text)
    return ast.literal_eval(text) if text != "[]" else []


def main(ref: str) -> dict:
    out: dict = {}
    p = os.path.join(ref, "test/reference64/typical/typical_spmv_csc.txt")
    lines = open(p).read().splitlines()
    A = julia_matrix(lines[0])
    last = [l for l in lines if l.strip()][-1]
    res = re.search(r"Element\{0\.0, Float64, Int64\}\((\[[^\]]*\])\)", last).group(1)
    out["spmv_csc"] = {"source": "test/reference64/typical/typical_spmv_csc.txt:1-17,92",
                       "A_dense": A, "x": [1.0, 2.0, 3.0], "y": julia_vector(res)}

    merges = {}
    for kind in ("twofinger", "leadfollow", "gallop"):
        p = os.path.join(ref, f"test/reference64/typical/typical_merge_{kind}.txt")
        last = [l for l in open(p).read().splitlines() if l.strip()][-1]
        m = re.search(r"\((\[[^\]]*\])\), (\d+), (\[[^\]]*\]), (\[[^\]]*\])\)", last)
        merges[kind] = {"val": julia_vector(m.group(1)), "shape": int(m.group(2)),
                        "ptr": julia_vector(m.group(3)), "idx": julia_vector(m.group(4))}
    out["merge_add"] = {
        "source": "test/reference64/typical/typical_merge_*.txt (last line); inputs "
                  "test/suites/typical_tests.jl:102-111",
        "x": {"ptr": [1, 6], "idx": [1, 3, 5, 7, 9], "val": [2.0, 3.0, 4.0, 5.0, 6.0]},
        "y": {"ptr": [1, 4], "idx": [2, 5, 8], "val": [1.0, 1.0, 1.0]},
        "shape": 10, "z": merges}

    p = os.path.join(ref, "test/reference64/parallel/parallel_spmv_atomics.txt")
    last = [l for l in open(p).read().splitlines() if l.strip()][-1]
    y = julia_vector(re.search(r"Element\{0, Int64, Int64\}\((\[[^\]]*\])\)", last).group(1))
    out["spmv_transpose"] = {
        "source": "test/reference64/parallel/parallel_spmv_atomics.txt:117; inputs "
                  "test/suites/parallel_tests.jl:414-416",
        "A_dense": [[1.0, 2.0], [3.0, 4.0]], "x": [1.0, 1.0], "y": y}

    src = open(os.path.join(ref, "test/suites/issue_tests.jl")).read()
    blk = src[src.index("ptr_c = [0, 3, 3, 5]"):]
    blk = blk[: blk.index("end")]
    ptr_c = julia_vector(re.search(r"ptr_c = (\[[^\]]*\])", blk).group(1))
    idx_c = julia_vector(re.search(r"idx_c = (\[[^\]]*\])", blk).group(1))
    val_c = julia_vector(re.search(r"val_c = (\[[^\]]*\])", blk).group(1))
    dense = julia_matrix(re.search(r"A == (\[[^\]]*\])", blk).group(1))
    out["zero_based"] = {"source": "test/suites/issue_tests.jl:388-404", "m": 4, "n": 3,
                         "ptr": ptr_c, "idx": idx_c, "val": val_c, "A_dense": dense}

    p = os.path.join(ref, "test/reference64/representation/Dense{SparseList}_representation.txt")
    lines = open(p).read().splitlines()
    reps = []
    for i, l in enumerate(lines):
        m = re.match(r"^(\w+): (.*\[.*\])$", l)
        if not m or m.group(1) in ("tensor", "countstored") or i + 1 >= len(lines):
            continue
        t = lines[i + 1]
        tm = re.search(r"Element\{([^,]+), (\w+), Int64\}\(([A-Za-z0-9]*\[[^\]]*\])\), (\d+), "
                       r"(\[[^\]]*\]), ([A-Za-z0-9]*\[[^\]]*\])\), (\d+)\)", t)
        if not tm or tm.group(2) not in ("Float64", "Bool"):
            continue
        rec = {"name": m.group(1), "dense": julia_matrix(m.group(2)),
               "val": [float(v) for v in julia_vector(tm.group(3).replace("Bool", ""))],
               "ptr": julia_vector(tm.group(5)), "idx": julia_vector(tm.group(6))}
        if rec not in reps:
            reps.append(rec)
    out["representation_dense_sparselist"] = {
        "source": "test/reference64/representation/Dense{SparseList}_representation.txt",
        "cases": reps}
    return out


if __name__ == "__main__":
    ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    data = main(ref)
    with open(os.path.join(HERE, "finch_goldens.json"), "w") as f:
        json.dump(data, f, indent=1)
    print("wrote", os.path.join(HERE, "finch_goldens.json"),
          {k: (len(v["cases"]) if "cases" in v else 1) for k, v in data.items()})
