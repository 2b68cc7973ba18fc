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


def test_solution_file_v5(tmp_path):
    p = str(tmp_path / "solution.mat")
    F.save_solution(p, 0, np.arange(4.0), {"type": "structure"}, version="5")
    import scipy.io as sio
    d = sio.loadmat(p)
    assert d["flag"].item() == 0 and d["mfinal"].shape == (4, 1)
    back = F.load_solution(p)
    assert back["usr_par"]["type"] == "structure"


def F_mat73_available():
    from fd2d_noise_amd import mat73
    return mat73.available()


def _usr_par_like():
    """usr_par-shaped contents: nested structs, strings, numbers, logicals, empties, complex and
    integer arrays, a struct array (usr_par.model(it)) and a cell."""
    return {"type": "structure", "measurement": {"source": "log_amplitude_ratio", "structure": "cc_time_shift"},
            "config": {"nx": 7.0, "nz": 5.0, "n_basis_fct": 0.0, "f_sample": np.linspace(0.01, 0.1, 4)},
            "model": [{"m": np.arange(6.0), "objective": 1.5, "name": "abcdefgh"},
                      {"m": np.ones(6), "objective": 0.25, "name": "ijklmnop"}],
            "data": {"c_data": np.arange(12.0).reshape(3, 4), "t": np.linspace(-1, 1, 4)},
            "debug": False, "nothing": None, "blank": "", "cells": ["a", np.eye(2), 3.0],
            "spectrum": np.array([[1 + 2j], [3 - 4j]]), "ids": np.arange(6, dtype=np.int32).reshape(2, 3),
            "mask": np.array([[True, False, True]]), "single": np.float32([1.5, 2.5])}


@pytest.mark.skipif(not F_mat73_available(), reason="no HDF5 C library in this environment")
def test_solution_file_v73(tmp_path):
    """MAT v7.3 (start_inversion.m:215): an HDF5 file with MATLAB's 512-byte user block and header,
    each value in MATLAB's layout (read back through the HDF5 library)."""
    from fd2d_noise_amd import mat73
    p = str(tmp_path / "solution.mat")
    usr = _usr_par_like()
    F.save_solution(p, 1, np.arange(4.0), usr)
    raw = open(p, "rb").read()
    assert raw.startswith(b"MATLAB 7.3 MAT-file, Platform: GLNXA64, Created on: ") and b"HDF5 schema 1.00 ." in raw[:116]
    assert raw[124:128] == b"\x00\x02IM"
    assert raw[512:520] == b"\x89HDF\r\n\x1a\n"  # the HDF5 superblock after the user block
    assert mat73._h5().H5Fis_hdf5(p.encode()) > 0
    b = F.load_solution(p)
    assert b["flag"].shape == (1, 1) and b["flag"][0, 0] == 1.0
    assert b["mfinal"].shape == (4, 1) and np.array_equal(b["mfinal"][:, 0], np.arange(4.0))
    u = b["usr_par"]
    assert u["type"] == "structure" and u["measurement"]["structure"] == "cc_time_shift"
    assert np.array_equal(u["config"]["f_sample"], usr["config"]["f_sample"].reshape(1, -1))
    assert [m["name"] for m in u["model"]] == ["abcdefgh", "ijklmnop"]
    assert np.array_equal(u["model"][0]["m"], np.arange(6.0).reshape(1, -1)) and u["model"][1]["objective"][0, 0] == 0.25
    assert np.array_equal(u["data"]["c_data"], usr["data"]["c_data"])
    assert u["debug"].dtype == np.bool_ and not u["debug"][0, 0]
    assert u["nothing"].shape == (0, 0) and u["blank"] == ""
    assert u["cells"][0] == "a" and np.array_equal(u["cells"][1], np.eye(2)) and u["cells"][2][0, 0] == 3.0
    assert np.array_equal(u["spectrum"], usr["spectrum"]) and u["spectrum"].dtype == np.complex128
    assert u["ids"].dtype == np.int32 and np.array_equal(u["ids"], usr["ids"])
    assert np.array_equal(u["mask"], usr["mask"]) and u["single"].dtype == np.float32


def test_mat73_layout(tmp_path):
    """The HDF5 objects carry MATLAB's attributes: MATLAB_class on every value, MATLAB_fields on
    structs, MATLAB_int_decode on char and logical, MATLAB_empty on empties, references for cells."""
    if not F_mat73_available():
        pytest.skip("no HDF5 C library in this environment")
    from fd2d_noise_amd import mat73
    import ctypes
    p = str(tmp_path / "x.mat")
    mat73.save(p, {"a": np.arange(6.0).reshape(2, 3), "s": {"x": "hi", "y": True}, "e": None, "c": ["q", 1.0]})
    L = mat73._h5()
    r = mat73._Reader(L)
    fid = L.H5Fopen(p.encode(), 0, 0)
    try:
        def attrs(path):
            o = L.H5Oopen(fid, path.encode(), 0)
            out = {k: r._attr(o, k) for k in ("MATLAB_class", "MATLAB_int_decode", "MATLAB_empty")}
            dims = r._dims(o) if L.H5Iget_type(o) == 5 else None
            L.H5Oclose(o)
            return out, dims
        a, dims = attrs("a")
        assert a["MATLAB_class"] == "double" and dims == (2, 3)
        s, _ = attrs("s")
        assert s["MATLAB_class"] == "struct"
        o = L.H5Oopen(fid, b"s", 0)
        assert r._fields(o) == ["x", "y"]
        L.H5Oclose(o)
        x, dims = attrs("s/x")
        assert x["MATLAB_class"] == "char" and x["MATLAB_int_decode"] == 2 and dims == (1, 2)
        y, _ = attrs("s/y")
        assert y["MATLAB_class"] == "logical" and y["MATLAB_int_decode"] == 1
        e, dims = attrs("e")
        assert e["MATLAB_class"] == "double" and e["MATLAB_empty"] == 1 and dims == (2,)
        c, dims = attrs("c")
        assert c["MATLAB_class"] == "cell" and dims == (1, 2)
        ra, _ = attrs("#refs#/a")
        assert ra["MATLAB_class"] == "char"
    finally:
        L.H5Fclose(fid)
