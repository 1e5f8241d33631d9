"""The committed golden fixtures (tests/golden/golden.npz, written by
tests/golden/make_golden.py from the CPU oracle on seeded inputs) against the
oracle now: FP32 and integer outputs bit for bit, FP64 ones to 1e-12
relative (libm may differ by an ulp between hosts). A change to the oracle or
to the scene generators shows up here first."""
import importlib.util
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = os.path.join(HERE, "golden", "golden.npz")


def _make_golden():
    spec = importlib.util.spec_from_file_location("make_golden", os.path.join(HERE, "golden", "make_golden.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def test_oracle_reproduces_golden_fixtures():
    ref = np.load(GOLDEN)
    got = _make_golden().compute()
    assert set(got) == set(ref.files)
    for k in ref.files:
        a, b = got[k], ref[k]
        assert a.shape == b.shape and a.dtype == b.dtype, k
        if a.dtype == np.float64:
            np.testing.assert_allclose(a, b, rtol=1e-12, atol=1e-300, err_msg=k)
        elif a.dtype.kind == "U":
            assert str(a) == str(b), k
        else:
            assert np.array_equal(a.view(np.uint8), b.view(np.uint8)), k
