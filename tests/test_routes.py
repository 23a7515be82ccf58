"""The routing table (finch.jl_amd/routes.toml) against the C ABI, the
integration notes and both routers, on the CPU.

Both the Julia glue (FinchGPU.jl: the @finch hook and the compute scheduler)
and the Python mirror (finchgpu/interface.py) pick kernels from this table by
a program's canonical form, so a program written the way Finch's tests write
it must reduce to the route named here.
"""
import os
import re

import pytest

from finchgpu import interface
from test_abi import declared_symbols

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
JL = open(os.path.join(REPO, "finch.jl_amd", "julia", "FinchGPU.jl")).read()

# (program as a user writes it, each operand: csc | csr (a swizzle(A, 2, 1)) |
# dense) -> (handler, op or semiring).
CASES = [
    ("y[i] += A[i, j] * x[j]", ["csc", "dense"], "spmv_csc", None),
    ("y[i] += S[i, j] * x[j]", ["csr", "dense"], "spmv_csr", None),
    ("y[i] += A[j, i] * x[j]", ["csc", "dense"], "spmv_csr", None),   # benchmarks.jl:365-376
    ("C[i, k] += A[i, j] * B[j, k]", ["csr", "dense"], "spmm_csr", None),
    ("C[i, k] += A[i, j] * B[j, k]", ["csc", "dense"], "spmm_csc", None),
    ("C[m, n] += A[m, p] * B[p, n]", ["csr", "dense"], "spmm_csr", None),
    ("C[i, j] += M[i, j] * X[i, k] * Y[k, j]", ["csr", "dense", "dense"], "sddmm_csr", None),
    ("C[i, j] += M[i, j] * X[i, k] * Y[k, j]", ["csc", "dense", "dense"], "sddmm_csc", None),
    ("C[i, j] = A[i, j] * B[i, j]", ["csc", "csc"], "ewise", "mul"),
    ("C[i, j] = A[i, j] + B[i, j]", ["csc", "csc"], "ewise", "add"),
    ("C[i, j] = max(A[i, j], B[i, j])", ["csc", "csc"], "ewise", "max"),
    ("C[i, j] = A[i, j] * B[i, j]", ["csr", "csr"], "ewise_rows", "mul"),
    ("C[i, j] = A[i, j] + B[i, j]", ["csr", "csr"], "ewise_rows", "add"),
    ("C[i, j] = max(A[i, j], B[i, j])", ["csr", "csr"], "ewise_rows", "max"),
    ("C[i, j] = A[i, j] * B[i, j]", ["csc", "csr"], "ewise_transpose_b", "mul"),
    ("C[i, j] = A[i, j] + B[i, j]", ["csc", "csr"], "ewise_transpose_b", "add"),
    ("C[i, j] = max(A[i, j], B[i, j])", ["csc", "csr"], "ewise_transpose_b", "max"),
    ("C[i, j] = A[i, j] * B[i, j]", ["csr", "csc"], "ewise_rows_transpose_b", "mul"),
    ("C[i, j] = A[i, j] + B[i, j]", ["csr", "csc"], "ewise_rows_transpose_b", "add"),
    ("C[i, j] = max(A[i, j], B[i, j])", ["csr", "csc"], "ewise_rows_transpose_b", "max"),
    ("C[i, k] += A[i, j] * B[j, k]", ["csc", "csc"], "spgemm_csc", "plus_times"),
    ("C[i, k] += A[i, j] * B[j, k]", ["csr", "csr"], "spgemm_csr", "plus_times"),
    ("C[i, k] += A[i, j] * B[j, k]", ["csr", "csc"], "spgemm_csr_transpose_b", "plus_times"),
    ("C[i, k] += A[i, j] * B[j, k]", ["csc", "csr"], "spgemm_csc_transpose_b", "plus_times"),
    ("C[i, k] <<min>>= A[i, j] + B[j, k]", ["csc", "csc"], "spgemm_csc", "min_plus"),
    ("C[i, k] <<min>>= A[i, j] + B[j, k]", ["csr", "csr"], "spgemm_csr", "min_plus"),
    ("C[i, k] <<min>>= A[i, j] + B[j, k]", ["csr", "csc"], "spgemm_csr_transpose_b", "min_plus"),
    ("C[i, k] <<min>>= A[i, j] + B[j, k]", ["csc", "csr"], "spgemm_csc_transpose_b", "min_plus"),
]


def _route(program, kinds):
    out, oidx, red, terms, comb = interface.parse(program)
    operands = [("dense", False) if k == "dense" else ("sparse", k == "csr") for k in kinds]
    key, ks = interface.canonical(oidx, red, terms, comb, operands)
    return interface.lookup(key, ks)


@pytest.mark.parametrize("program,kinds,handler,param", CASES)
def test_programs_reach_their_route(program, kinds, handler, param):
    r = _route(program, kinds)
    assert r is not None, program
    assert r["handler"] == handler
    assert r.get("op", r.get("semiring")) == param


def test_every_route_is_reached_and_keys_are_unique():
    keys = [(r["program"], tuple(r["operands"])) for r in interface.routes()]
    assert len(keys) == len(set(keys))
    reached = {id(_route(p, k)) for p, k, _, _ in CASES}
    assert reached == {id(r) for r in interface.routes()}


@pytest.mark.parametrize("program,kinds", [
    ("y[i] += A[i, j] * x[j]", ["dense", "dense"]),          # dense GEMV
    ("C[i, j] = A[i, j] * B[i, j]", ["csc", "dense"]),       # sparse .* dense
    ("C[i, k] <<max>>= A[i, j] + B[j, k]", ["csc", "csc"]),  # max-plus
    ("C[i, k] += A[i, j] * B[j, k]", ["dense", "csc"]),      # dense x sparse
])
def test_unrouted_programs_have_no_route(program, kinds):
    assert _route(program, kinds) is None


def test_route_entries_are_declared_and_documented():
    syms = set(declared_symbols())
    integ = open(os.path.join(REPO, "INTEGRATION.md")).read()
    for r in interface.routes():
        for e in r["entry"]:
            assert e in syms, (r["program"], e)
            assert f"`{e}`" in integ, e
        assert r["reference"]


def test_both_routers_have_every_handler():
    handlers = {r["handler"] for r in interface.routes()}
    assert handlers == set(interface.HANDLERS)
    jl = set(re.findall(r'"(\w+)"\s*=>\s*route_\w+', JL))
    assert handlers == jl
    # The Julia router reads the same file and hooks both of Finch's paths.
    assert "routes.toml" in JL
    assert "function Finch.execute(" in JL
    assert "Finch.set_scheduler!" in JL or "set_scheduler!" in JL
