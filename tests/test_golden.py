"""The oracle reproduces the golden fixtures written from the reference core (bit-exact).

Fixtures: tests/golden/*.npz, made by tests/golden/make_golden.py from oracle/_ref (the
reference's own C++ sources).  They pin the oracle on machines without the reference tree.
"""
import sys
from pathlib import Path

import numpy as np
import pytest

import checkers

sys.path.insert(0, str(Path(__file__).resolve().parent / "golden"))
import make_golden  # noqa: E402

CASES = {c[0]: c for c in make_golden.cases()}
GOLDEN = Path(__file__).resolve().parent / "golden"


@pytest.mark.parametrize("name", sorted(CASES))
def test_oracle_matches_golden(name):
    _, p, rg, iid, h, params, opts = CASES[name]
    want = np.load(GOLDEN / f"{name}.npz", allow_pickle=False)
    assert np.allclose(want["input_checksum"], [np.nansum(p.X), np.sum(p.y)], rtol=0, atol=0), \
        "panel generator drifted; regenerate fixtures"
    got = make_golden.pack(make_golden.fit(checkers.oracle(), name, p, rg, iid, h, params, opts), p)
    for k in want.files:
        assert np.array_equal(want[k], got[k], equal_nan=want[k].dtype.kind == "f"), f"{name}: {k}"
