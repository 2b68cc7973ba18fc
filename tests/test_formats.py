"""MAT-file interop (fd2d_noise_amd/formats.py): the reference's file names, variable names and
shapes round-trip through MAT v5 files that MATLAB's load() reads."""
import types

import numpy as np
import pytest

from fd2d_noise_amd import Config, formats as F


def test_array_and_data_files(tmp_path):
    rng = np.random.default_rng(0)
    array = rng.uniform(0, 2e6, (5, 2))
    ref = array[[1, 3]]
    c = Config(nt=40)
    c_data = rng.standard_normal((2 * 4, c.n_corr))
    F.save_array(str(tmp_path / "array_2_ref.mat"), array, ref)
    F.save_data(str(tmp_path / "data_2_ref_0.mat"), c_data, c.t_corr())
    with open(tmp_path / "data_2_ref_0.mat", "rb") as f:
        assert f.read(19) == b"MATLAB 5.0 MAT-file"
    net = F.load_array(str(tmp_path / "array_2_ref.mat"))
    dat = F.load_data(str(tmp_path / "data_2_ref_0.mat"))
    assert np.array_equal(net["array"], array) and np.array_equal(net["ref_stat"], ref)
    assert np.array_equal(dat["c_data"], c_data)
    # t is MATLAB's row vector -(nt-1)dt : dt : (nt-1)dt
    assert dat["t"].shape == (1, c.n_corr) and np.array_equal(dat["t"][0], c.t_corr())
    assert dat["t"][0, c.nt - 1] == 0.0


def test_g_fft_parsave_parload(tmp_path):
    rng = np.random.default_rng(1)
    G = (rng.standard_normal((6, 5, 4)) + 1j * rng.standard_normal((6, 5, 4))).astype(np.complex128)
    p = str(tmp_path / "G_fft_3.mat")
    F.save_G_fft(p, G)
    back = F.load_G_fft(p)
    assert back.dtype == np.complex128 and back.shape == (6, 5, 4) and np.array_equal(back, G)
    import scipy.io as sio
    assert set(k for k in sio.loadmat(p) if not k.startswith("__")) == {"x"}  # parsave.m:3
    two = str(tmp_path / "two.mat")
    sio.savemat(two, {"a": 1.0, "b": 2.0})
    with pytest.raises(ValueError, match="Problem with"):
        F.load_G_fft(two)


def test_model_file_fields(tmp_path):
    c = Config(nx=7, nz=6, n_basis_fct=2)
    usr = types.SimpleNamespace(config=c, sigma_source=[1e4, 2e4], sigma_structure=[3e4, 4e4], ring={"switch": "no"})
    model = {"objective": 1.5, "m": np.arange(7 * 6 * 3, dtype=np.float64), "gradient": np.ones(7 * 6 * 3),
             "normg": 2.0, "correlation": np.zeros((3, 11))}
    s = F.model_struct(model, usr)
    assert list(s)[:5] == sorted(model)  # orderfields (new_iteration.m:28)
    p = str(tmp_path / "model_4.mat")
    F.save_model(p, model, usr)
    back = F.load_model(p)
    assert float(np.squeeze(back["objective"])) == 1.5
    assert np.array_equal(np.squeeze(back["m"]), model["m"])
    assert np.array_equal(np.squeeze(back["sigma"]["structure"]), [3e4, 4e4])
    assert float(np.squeeze(back["config"]["n_basis_fct"])) == 2 and float(np.squeeze(back["config"]["nx"])) == 7
    assert "ring" not in back


def test_solution_file(tmp_path):
    p = str(tmp_path / "solution.mat")
    F.save_solution(p, 0, np.arange(4.0), {"type": "structure"})
    import scipy.io as sio
    d = sio.loadmat(p)
    assert d["flag"].item() == 0 and d["mfinal"].shape == (4, 1)
