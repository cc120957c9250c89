"""The element-buffer gather on the CPU: tests/cpp/egather_check builds the gather plan of a box hierarchy's top
level with the library's own host units and replays the gather kernel on it block by block (random block
order, every array and LDS access bounds-checked), comparing the slot values and row right-hand sides bit for
bit with direct element-order sums.  Heat (1 Q2 variable) and Navier-Stokes (DIM Q2 + 1 Q1) layouts, 2D and
3D, permuted and natural internal orders, slice windows of 64 to 4096 rows, and LDS budgets that force every
lane group down to 2; a budget below one row block is refused."""
import json
import os
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
CHECK = os.path.join(HERE, "cpp", "_build", "egather_check")


@pytest.fixture(scope="module")
def checker():
    if not os.path.exists(CHECK):
        pytest.fail("gather checker not built: run __graft_entry__.build()")
    return CHECK


CASES = [
    # dim nx ny nz nlev nq2 nq1 W max_lds seed permute -> lane group
    ((3, 2, 2, 2, 3, 1, 0, 4096, 49152, 1, 1), 16),
    ((3, 2, 1, 1, 3, 3, 1, 4096, 49152, 2, 1), 4),
    ((2, 3, 2, 1, 3, 2, 1, 64, 49152, 3, 1), 16),
    ((3, 2, 2, 1, 2, 1, 0, 4096, 8192, 4, 1), 2),
    ((2, 2, 2, 1, 4, 1, 0, 4096, 49152, 6, 0), 16),
    ((3, 1, 2, 1, 3, 3, 1, 512, 20000, 7, 1), 2),
    ((3, 2, 2, 2, 3, 3, 1, 4096, 57344, 5, 0), 4),    # the library's Q2-Q1 budget (56 KB)
]


@pytest.mark.parametrize("args,lg", CASES)
def test_gather_plan_replay_bitexact(checker, args, lg):
    r = subprocess.run([checker] + [str(a) for a in args], capture_output=True, text=True, timeout=300)
    out = json.loads(r.stdout.strip().splitlines()[-1])
    assert out["ok"], out
    assert out["lg"] == lg
    assert out["lds"] <= args[8]


def test_gather_plan_refuses_too_small_lds(checker):
    r = subprocess.run([checker] + [str(a) for a in (3, 1, 1, 1, 2, 3, 1, 256, 6000, 5, 0)], capture_output=True,
                       text=True, timeout=300)
    out = json.loads(r.stdout.strip().splitlines()[-1])
    assert not out["ok"] and "exceed the LDS" in out["error"]
