"""The internal DoF order of every level, built on the host by the library's own code
(exfemus_amd/csrc/order_host.cpp, on meshes from exfemus_amd/csrc/gencase.cpp), checked by
tests/cpp/order_check.cpp against what is recomputed from node coordinates and element geometry alone: a
permutation with its inverse, the DoFs of a node adjacent and in variable order, groups inside one slab of 8
node planes with the group's signature and lexicographic node order, element colours that share no node, and,
for coupled (Q2-Q1) systems, the rows of the groups holding colour c equal to the DoFs of colour c's elements
(the rows the coloured Vanka sweep recomputes its residual on).  2D and 3D, heat-shaped (one Q2 variable) and
Navier-Stokes-shaped (dim Q2 + one Q1 variable), cubic and elongated boxes (ties and non-ties of the slowest
axis), every level of 3 to 7-level hierarchies, up to the 96 x 24 x 24 channel."""
import json
import os
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
CHECK = os.path.join(HERE, "cpp", "_build", "order_check")

CASES = [
    # dim nx ny nz nlev nq2 nq1
    (2, 3, 2, 1, 3, 1, 0),
    (2, 3, 2, 1, 4, 2, 1),
    (2, 8, 8, 1, 6, 2, 1),
    (3, 2, 1, 1, 3, 1, 0),     # cells 8 x 4 x 4: tie of y and z
    (3, 2, 1, 1, 4, 3, 1),
    (3, 1, 2, 3, 3, 3, 1),     # slowest axis z
    (3, 4, 4, 4, 4, 1, 0),     # 64^3 heat
    (3, 12, 3, 3, 4, 3, 1),    # 96 x 24 x 24 channel, 1.45M DoFs
]


@pytest.mark.parametrize("args", CASES, ids=lambda a: "x".join(map(str, a)))
def test_level_order(args):
    if not os.path.exists(CHECK):
        pytest.fail("order checker not built: run __graft_entry__.build()")
    out = subprocess.run([CHECK, *map(str, args)], capture_output=True, text=True, timeout=300)
    res = json.loads(out.stdout.strip().splitlines()[-1])
    assert out.returncode == 0 and res["ok"], res
    dim, nx, ny, nz, nlev, nq2, nq1 = args
    cells = [n << (nlev - 1) for n in (nx, ny, nz)[:dim]]
    nodes = 1
    for c in cells:
        nodes *= 2 * c + 1
    vert = 1
    for c in cells:
        vert *= c + 1
    assert res["ndof"] == nq2 * nodes + nq1 * vert
    assert res["slabs"] == (2 * max(cells)) // 8 + 1
