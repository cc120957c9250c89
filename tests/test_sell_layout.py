"""The SELL-64 layout of the level products, built on the host (exfemus_amd/csrc/sell_host.cpp, the code the
library runs before uploading), checked slot by slot by tests/cpp/sell_layout_check.cpp on structured patterns
(Q2-shaped couplings of radius 1 and 2, 1 to 4 variables per node), natural and random internal orders, row
groups, sort windows of 64 and 4096 rows, the narrowest column class allowed (2, 8, 16 or 32 bits): every CSR entry is held exactly once at a
position of its own row, decodes to its internal column, rows are sorted, padding is marked and decodes inside the
vector, and each group starts a slice."""
import json
import os
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
CHECK = os.path.join(HERE, "cpp", "_build", "sell_layout_check")

CASES = [
    # nx ny nz nv r seed groups W minbits permute
    (66, 9, 5, 1, 1, 1, 1, 4096, 2, 0),      # lines of 64 interior nodes: the lane-predicted 2-bit class
    (10, 8, 6, 1, 1, 1, 1, 4096, 2, 0),
    (66, 9, 5, 1, 2, 8, 3, 4096, 2, 0),
    (7, 9, 5, 4, 2, 7, 11, 64, 2, 1),
    (10, 8, 6, 1, 1, 1, 1, 4096, 8, 0),
    (10, 8, 6, 1, 2, 3, 5, 4096, 8, 1),
    (7, 9, 5, 4, 2, 7, 11, 64, 8, 1),
    (12, 5, 5, 3, 1, 2, 3, 4096, 16, 0),
    (33, 17, 9, 1, 2, 4, 1, 4096, 8, 0),
    (10, 8, 6, 2, 2, 6, 3, 4096, 32, 0),     # absolute columns in every slice
    (80, 80, 40, 1, 1, 5, 4, 4096, 8, 1),   # random order of 256,000 rows: slices with 32-bit columns
]


def _run(args):
    if not os.path.exists(CHECK):
        pytest.fail("layout checker not built: run __graft_entry__.build()")
    out = subprocess.run([CHECK, *map(str, args)], capture_output=True, text=True, timeout=120)
    return json.loads(out.stdout.strip().splitlines()[-1]), out.returncode


@pytest.mark.parametrize("args", CASES)
def test_sell_layout(args):
    res, rc = _run(args)
    assert rc == 0 and res["ok"], res
    assert res["c2"] + res["c8"] + res["c16"] + res["c32"] >= res["slots"] >= res["nnz"]
    nx, ny, nz, nv, r, seed, groups, W, minbits, permute = args
    if minbits >= 8:
        assert res["c2"] == 0
    if args == CASES[0]:
        # natural order, one variable: a slice holds 64 consecutive interior nodes of a line, whose columns are
        # base + lane, or base + lane + 2 past the two boundary nodes where the slice wraps to the next line
        assert res["c2"] > 4 * (res["c8"] + res["c16"])
    if minbits >= 16:
        assert res["c8"] == 0
    if minbits == 32:
        assert res["c16"] == 0 and res["c32"] == res["slots"]
    if not permute and r == 1 and nv == 1 and minbits == 8:
        assert res["c8"] > res["c16"]        # natural order, narrow stencil: mostly 8-bit offsets
    if args == CASES[-1]:
        assert res["c32"] > 0
