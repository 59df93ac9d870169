"""The committed Si8 fixture (tests/golden/si8_L32_B4.npz, written by
tests/golden/make_golden.py from the oracle) against the current system
builder and oracle: a regression pin of the inputs and of every stage."""
import hashlib
import os

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden", "si8_L32_B4.npz")


@pytest.fixture(scope="module")
def gold():
    return dict(np.load(GOLD, allow_pickle=False))


def test_inputs_fingerprints(gold, si8_system):
    s = si8_system
    fp = [hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()
          for a in (s.vks, s.proj, s.map)]
    assert list(gold["fingerprints"]) == fp


def test_oracle_reproduces_fixture(gold):
    import sys
    sys.path.insert(0, os.path.join(HERE, "golden"))
    import make_golden
    d = make_golden.compute()
    for k in ("chi", "f", "T1", "T2", "T4", "coeffs", "acc", "tl", "moments", "mu", "emin_emax",
              "rho_sum", "rho_norm", "rho_sub", "energies"):
        ref = gold[k]
        assert np.allclose(d[k], ref, rtol=1e-12, atol=1e-12 * np.max(np.abs(ref))), k
