"""THC ERI file (thc.eri.h5) reader and writer, libthcgw_eri.so (no GPU).

Pins: a write/read round trip; the file as seen by h5py (an independent HDF5 reader, from
the conda Python of this image) -- dataset names, shapes, dtypes, the __complex__ tag and
the values; a file written by h5py in the same layout read back by the library; and the
consistency checks of thc_reader_t::read (thc_reader_t.hpp:519-530).  The reference ships
no THC ERI file, so the layout is pinned by its writer and reader code, not by a fixture.
"""
import json
import os
import subprocess

import numpy as np
import pytest

from coqui_amd.eri import ERI_EXPORTS, ERI_LIB_PATH, load_eri_library, read_thc_eri, write_thc_eri
from coqui_amd.system import make_system

H5PY_PYTHON = "/opt/conda/bin/python3.9"
need_h5py = pytest.mark.skipif(
    not os.path.exists(H5PY_PYTHON) or subprocess.run([H5PY_PYTHON, "-c", "import h5py"],
                                                      capture_output=True).returncode != 0,
    reason="no Python with h5py in this image")


def sample(kgrid=(2, 2, 1), no=5, nmu=12, sid=3):
    s = make_system(kgrid, no, nmu, system_id=sid)
    n1, n2, n3 = kgrid
    kc = np.array([[i / n1, j / n2, l / n3] for i in range(n1) for j in range(n2) for l in range(n3)])
    return s, kc


def test_library_exports():
    lib = load_eri_library()
    assert os.path.exists(ERI_LIB_PATH)
    for name in ERI_EXPORTS:
        assert hasattr(lib, name)


def test_round_trip(tmp_path):
    s, kc = sample()
    p = tmp_path / "thc.eri.h5"
    ip = np.arange(s.nmu, dtype=np.int64) * 7
    write_thc_eri(str(p), s.X, s.V, kpts=kc, qpts=kc, nbnd=s.no, interp_points=ip)
    r = read_thc_eri(str(p))
    assert r["info"]["Np"] == s.nmu and r["info"]["nkpts"] == s.nk and r["info"]["nqpts_ibz"] == s.nk
    assert (r["info"]["x_first"], r["info"]["x_last"]) == (0, s.no) and not r["info"]["has_Y"]
    assert np.array_equal(r["X"][0], s.X) and np.array_equal(r["V"], s.V)
    assert np.array_equal(r["kpts"], kc) and np.array_equal(r["qpts"], kc)


def test_y_collocation_round_trip(tmp_path):
    s, kc = sample(no=6)
    p = tmp_path / "thc_xy.eri.h5"
    Y = s.X[..., 2:5] * (1 + 0.5j)
    write_thc_eri(str(p), s.X, s.V, x_range=(0, 6), Y=Y, y_range=(2, 5), nbnd=6)
    r = read_thc_eri(str(p))
    assert r["info"]["has_Y"] and np.array_equal(r["Y"][0], Y)


@need_h5py
def test_layout_seen_by_h5py(tmp_path):
    s, kc = sample()
    p = tmp_path / "thc.eri.h5"
    write_thc_eri(str(p), s.X, s.V, kpts=kc, qpts=kc, nbnd=s.no, interp_points=np.arange(s.nmu))
    code = r'''
import h5py, json, sys, numpy as np
out = {}
with h5py.File(sys.argv[1], "r") as f:
    for k in f:
        d = f[k]
        a = np.asarray(d[()])
        e = {"shape": list(a.shape), "dtype": str(a.dtype)}
        if "__complex__" in d.attrs:
            v = d.attrs["__complex__"]
            e["complex"] = v.decode() if isinstance(v, bytes) else str(v)
        if a.dtype.kind == "f":
            e["sum"] = float(a.sum()); e["sumsq"] = float((a * a).sum())
        else:
            e["value"] = a.tolist()
        out[k] = e
print(json.dumps(out))
'''
    res = subprocess.run([H5PY_PYTHON, "-c", code, str(p)], capture_output=True, text=True, check=True)
    seen = json.loads(res.stdout)
    assert seen["Np"]["value"] == s.nmu and seen["number_of_bands"]["value"] == s.no
    assert seen["X_orbital_range"]["value"] == [0, s.no] and seen["Y_orbital_range"]["value"] == [0, s.no]
    assert seen["collocation_matrix"]["shape"] == [1, s.nk, s.nmu, s.no, 2]
    assert seen["collocation_matrix"]["dtype"] == "float64" and seen["collocation_matrix"]["complex"] == "1"
    assert seen["coulomb_matrix"]["shape"] == [s.nk, s.nmu, s.nmu, 2] and seen["coulomb_matrix"]["complex"] == "1"
    assert seen["kpts"]["shape"] == [s.nk, 3] and seen["interpolating_points"]["value"] == list(range(s.nmu))
    Xr = s.X.view(np.float64)
    assert np.isclose(seen["collocation_matrix"]["sum"], Xr.sum(), rtol=1e-13)
    assert np.isclose(seen["coulomb_matrix"]["sumsq"], (s.V.view(np.float64) ** 2).sum(), rtol=1e-13)


@need_h5py
def test_reads_file_written_by_h5py(tmp_path):
    """A file in the nda layout written by h5py is read back by the library."""
    s, kc = sample(kgrid=(1, 1, 2), no=4, nmu=9, sid=4)
    src = tmp_path / "arrays.npz"
    np.savez(src, X=s.X[None].view(np.float64).reshape((1,) + s.X.shape + (2,)), V=s.V.view(np.float64).reshape(s.V.shape + (2,)),
             k=kc)
    p = tmp_path / "h5py.eri.h5"
    code = r'''
import h5py, sys, numpy as np
z = np.load(sys.argv[1])
with h5py.File(sys.argv[2], "w") as f:
    f["Np"] = np.int32(z["X"].shape[2]); f["number_of_bands"] = np.int32(z["X"].shape[3])
    f["X_orbital_range"] = np.array([0, z["X"].shape[3]], np.int32)
    f["Y_orbital_range"] = np.array([0, z["X"].shape[3]], np.int32)
    f["kpts"] = z["k"]; f["qpts"] = z["k"]
    f["nkpts_ibz"] = np.int32(z["X"].shape[1]); f["nqpts_ibz"] = np.int32(z["V"].shape[0])
    for k in ("X", "V"):
        name = "collocation_matrix" if k == "X" else "coulomb_matrix"
        f[name] = z[k]; f[name].attrs["__complex__"] = "1"
    f["interpolating_points"] = np.arange(z["X"].shape[2], dtype=np.int64)
'''
    subprocess.run([H5PY_PYTHON, "-c", code, str(src), str(p)], check=True)
    r = read_thc_eri(str(p))
    assert np.array_equal(r["X"][0], s.X) and np.array_equal(r["V"], s.V) and np.array_equal(r["kpts"], kc)


@need_h5py
def test_consistency_checks(tmp_path):
    """thc_reader_t::read refuses ranges beyond nbnd and interpolating_points of the wrong length."""
    s, kc = sample(no=4, nmu=8)
    good = tmp_path / "good.eri.h5"
    write_thc_eri(str(good), s.X, s.V, nbnd=4, interp_points=np.arange(8))
    cases = {
        "range": 'f["X_orbital_range"][...] = np.array([0, 9], np.int32)',
        "rp": 'del f["interpolating_points"]; f["interpolating_points"] = np.arange(5)',
        "np": 'del f["Np"]; f["Np"] = np.int32(7)',
    }
    for tag, edit in cases.items():
        bad = tmp_path / ("bad_%s.eri.h5" % tag)
        bad.write_bytes(good.read_bytes())
        subprocess.run([H5PY_PYTHON, "-c", "import h5py, numpy as np, sys\nwith h5py.File(sys.argv[1], 'a') as f:\n    "
                        + edit, str(bad)], check=True)
        with pytest.raises(IOError):
            read_thc_eri(str(bad))
    with pytest.raises(IOError, match="cannot open"):
        read_thc_eri(str(tmp_path / "missing.h5"))


def test_qpts_rows_differ_from_kpts(tmp_path):
    """qpts with nq != nk rows round-trip into a buffer sized from qpts, not from nk."""
    s, kc = sample()
    qc = np.vstack([kc, kc[:3] + 0.5])  # nq = nk + 3
    p = tmp_path / "nq.eri.h5"
    write_thc_eri(str(p), s.X, s.V, kpts=kc, qpts=qc, nbnd=s.no)
    r = read_thc_eri(str(p))
    assert r["info"]["nqpts"] == s.nk + 3 and np.array_equal(r["qpts"], qc) and np.array_equal(r["kpts"], kc)


def test_read_refuses_wrong_capacity(tmp_path):
    """Every read function refuses an output buffer whose size differs from the dataset."""
    import ctypes

    s, kc = sample()
    p = tmp_path / "cap.eri.h5"
    write_thc_eri(str(p), s.X, s.V, kpts=kc, qpts=kc, nbnd=s.no)
    lib = load_eri_library()
    b = os.fsencode(str(p))
    big = np.zeros(2 * s.X.size + 64)
    ptr = big.ctypes.data_as(ctypes.c_void_p)
    err = ctypes.create_string_buffer(256)
    assert lib.thcgw_eri_read_X(b, 0, ptr, 2 * s.X.size - 2, err, 256) != 0 and b"output buffer" in err.value
    assert lib.thcgw_eri_read_X(b, 0, ptr, 2 * s.X.size, err, 256) == 0
    vb = np.zeros(2 * s.V.size + 2)
    assert lib.thcgw_eri_read_V(b, vb.ctypes.data_as(ctypes.c_void_p), 2 * s.V.size + 2, err, 256) != 0
    kb = np.zeros(3 * s.nk)
    assert lib.thcgw_eri_read_kpts(b, None, 0, kb.ctypes.data_as(ctypes.c_void_p), 3 * s.nk - 3, err, 256) != 0
    assert lib.thcgw_eri_read_kpts(b, None, 0, kb.ctypes.data_as(ctypes.c_void_p), 3 * s.nk, err, 256) == 0
    with pytest.raises(ValueError):
        write_thc_eri(str(tmp_path / "bad.eri.h5"), s.X, s.V, kpts=kc[:-1])


@need_h5py
def test_mis_shaped_y_is_refused(tmp_path):
    """read_info checks Y_collocation_matrix against (ns*npol, nk, Np, ny) before any read."""
    s, kc = sample(no=6)
    good = tmp_path / "y.eri.h5"
    write_thc_eri(str(good), s.X, s.V, x_range=(0, 6), Y=s.X[..., 2:5], y_range=(2, 5), nbnd=6)
    bad = tmp_path / "y_bad.eri.h5"
    bad.write_bytes(good.read_bytes())
    edit = ('y = f["Y_collocation_matrix"][()]; del f["Y_collocation_matrix"]; '
            'f["Y_collocation_matrix"] = np.concatenate([y, y], axis=2); f["Y_collocation_matrix"].attrs["__complex__"] = "1"')
    subprocess.run([H5PY_PYTHON, "-c", "import h5py, numpy as np, sys\nwith h5py.File(sys.argv[1], 'a') as f:\n    "
                    + edit, str(bad)], check=True)
    with pytest.raises(IOError, match="Y_collocation_matrix"):
        read_thc_eri(str(bad))
