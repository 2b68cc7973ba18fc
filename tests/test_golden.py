"""The committed golden fixture tests/golden/tiny_chain.npz (written by tests/golden/make_fixtures.py
from the NumPy restatement of the solvers, on the reference's tiny set-up): one station's gradient
chain -- Green run, correlation run, adjoint runs mode 1 and mode 0 -- with its inputs and outputs.

  * the C oracle reproduces it to 1e-13 (the two restatements agree);
  * the generator reproduces the committed file (the fixture is what the script makes);
  * the HIP engine, through the C ABI, matches it within the parity tolerances: traces and spectra
    1e-5, kernels 1e-4 (relative L2).
"""
import os
import sys

import numpy as np
import pytest

from conftest import rel_l2

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))
import make_fixtures as MF  # noqa: E402

TOL = {"green_seis": 1e-5, "G_fft_sub": 1e-5, "correlation": 1e-5, "adj1_K": 1e-4, "adj1_seis": 1e-5,
       "stf_fft_sub": 1e-5, "adj0_K": 1e-4}


@pytest.fixture(scope="module")
def fixture():
    with np.load(os.path.join(HERE, "golden", "tiny_chain.npz"), allow_pickle=False) as f:
        return {k: f[k] for k in f.files}


def _inputs(fx):
    return {k[3:]: fx[k] for k in fx if k.startswith("in_")}


def _check(out, fx, tol):
    for k in TOL:
        ref = fx["out_" + k]
        got = np.asarray(out[k])
        if k.endswith("_K"):  # the structure kernel (the source kernel is zero without G_fft)
            got, ref = got[:, :, -1], ref[:, :, -1]
        e = rel_l2(got, ref)
        print(f"{k}: {e:.2e}")
        assert got.shape == ref.shape and np.abs(ref).max() > 0
        assert e < (tol if tol is not None else TOL[k]), (k, e)


def test_oracle_reproduces_the_fixture(orc, fixture):
    c = MF.config()
    _check(MF.chain(orc, c, _inputs(fixture), False), fixture, 1e-13)


def test_generator_reproduces_the_committed_fixture(fixture):
    c = MF.config()
    d = MF.inputs(c)
    for k, v in d.items():
        assert np.array_equal(v, fixture["in_" + k]), k
    out = MF.chain(MF.NP, c, d, True)
    for k, v in out.items():
        assert np.array_equal(v, fixture["out_" + k]), k


@pytest.mark.gpu
def test_engine_matches_the_fixture(fixture):
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import fd2d_noise_amd as fd
    c = MF.config()
    _check(MF.chain(fd, c, _inputs(fixture), False), fixture, None)
