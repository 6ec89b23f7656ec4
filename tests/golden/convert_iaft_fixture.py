"""Copy the reference's IAFT golden data into tests/golden as .npz.

Source: tests/unit_test_files/pyscf/si_kp222_krhf/gw_Gw_Gt_beta1000_wmax1.2_{high,medium}.h5
(written by gen_gw_gt.py in the same directory with the sparse_ir-based Python IAFT).
The reference checks its C++ IAFT against these arrays in
src/numerics/imag_axes_ft/tests/test_iaft.cpp:46-119 (tolerance 1e-11 high, 1e-9 medium).
These are data (inputs and expected outputs); they are copied unchanged.

Run with a Python that has h5py:
    /opt/conda/bin/python3.9 tests/golden/convert_iaft_fixture.py /root/reference
"""
import os
import sys

import h5py
import numpy as np


def _cplx(ds):
    a = np.asarray(ds[()], dtype=np.float64)
    return a[..., 0] + 1j * a[..., 1]


def main(ref_root):
    src = os.path.join(ref_root, "tests/unit_test_files/pyscf/si_kp222_krhf")
    out_dir = os.path.dirname(os.path.abspath(__file__))
    for prec in ("high", "medium"):
        fn = os.path.join(src, "gw_Gw_Gt_beta1000_wmax1.2_%s.h5" % prec)
        with h5py.File(fn, "r") as f:
            np.savez_compressed(os.path.join(out_dir, "iaft_si_kp222_%s.npz" % prec),
                                Gt=_cplx(f["Gt"]), Gw=_cplx(f["Gw"]), Dm=_cplx(f["Dm"]),
                                beta=np.float64(1000.0), wmax=np.float64(1.2))
        print(prec, "ok")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
