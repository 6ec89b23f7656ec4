"""Convert the reference's IR sparse-sampling tables (HDF5) into .npz files.

The tables are numerical data shipped by CoQui under Apache-2.0
(src/numerics/imag_axes_ft/ir/data/*.h5, written by gen_sparse_ir.py with
sparse_ir 1.1.7).  They are stored here unchanged, in their dimensionless
form: the beta-dependent rescaling of src/numerics/imag_axes_ft/ir/ir_driver.hpp:68-73
is applied at load time by coqui_amd.iaft.

Run with a Python that has h5py (this image: /opt/conda/bin/python3.9):

    /opt/conda/bin/python3.9 coqui_amd/data/ir/convert_ir_tables.py /root/reference

Dataset names follow the HDF5 layout (ir_driver.hpp:45-65).  Complex datasets are
stored in the HDF5 files as (..., 2) float64 arrays and are converted to complex128.
"""
import os
import sys

import h5py
import numpy as np

TABLES = [
    "1e2.1e-15", "1e3.1e-15", "1e4.1e-15", "1e5.1e-15", "1e6.1e-15",
    "1e4.1e-10", "1e5.1e-06",
]
COMPLEX_F = ["Ttw", "Twt", "Ttt_bf", "T_betat", "Tct"]
COMPLEX_B = ["Ttw", "Twt", "Ttt_fb", "Tct"]


def _cplx(ds):
    a = np.asarray(ds[()], dtype=np.float64)
    if a.shape and a.shape[-1] == 2:
        return a[..., 0] + 1j * a[..., 1]
    return a.astype(np.complex128)


def convert(ref_root, out_dir):
    src_dir = os.path.join(ref_root, "src/numerics/imag_axes_ft/ir/data")
    for name in TABLES:
        with h5py.File(os.path.join(src_dir, name + ".h5"), "r") as f:
            out = {"version": np.array(f["version"][()]).astype(str)}
            fg, bg = f["fermion"], f["boson"]
            out["f_nt"] = np.int64(fg["nt"][()])
            out["f_nw"] = np.int64(fg["nw"][()])
            out["f_tau_mesh"] = np.asarray(fg["tau_mesh"][()], np.float64)
            out["f_wn_mesh"] = np.asarray(fg["wn_mesh"][()], np.int64)
            for k in COMPLEX_F:
                out["f_" + k] = _cplx(fg[k])
            # T_zerot is written as a real vector by gen_sparse_ir.py:62-63
            out["f_T_zerot"] = _cplx(fg["T_zerot"])
            out["b_nt"] = np.int64(bg["nt"][()])
            out["b_nw"] = np.int64(bg["nw"][()])
            out["b_tau_mesh"] = np.asarray(bg["tau_mesh"][()], np.float64)
            out["b_wn_mesh"] = np.asarray(bg["wn_mesh"][()], np.int64)
            for k in COMPLEX_B:
                out["b_" + k] = _cplx(bg[k])
        np.savez_compressed(os.path.join(out_dir, name + ".npz"), **out)
        print(name, int(out["f_nt"]), int(out["f_nw"]), int(out["b_nt"]), int(out["b_nw"]))


if __name__ == "__main__":
    ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    convert(ref, os.path.dirname(os.path.abspath(__file__)))
