"""bench.py's command line and form table, without a GPU: the multi-GPU refusal, form
validation, and the flags each form sets (the plain form keeps four real products per
complex product, every optimized form uses three)."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import bench  # noqa: E402


def run(*args):
    return subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), *args], capture_output=True, text=True,
                          timeout=120, env=dict(os.environ, HIP_VISIBLE_DEVICES=""))


def test_more_than_one_gpu_is_refused():
    r = run("--gpus", "2")
    assert r.returncode == 2 and "one GPU only" in r.stderr


@pytest.mark.parametrize("args,msg", [(("--forms", "plain,nonsense"), "unknown form"),
                                      (("--forms", "plain,mixed_resolvent_trev"), "needs --trev"),
                                      (("--forms", "plain,resolvent_ibz"), "needs --c4")])
def test_bad_forms_are_refused(args, msg):
    r = run(*args)
    assert r.returncode == 2 and msg in r.stderr


def test_form_table():
    for name, f in bench.FORMS.items():
        assert len(f) == 4 and f[0] in (0, 1, 2) and f[1] in (0, 1, 2) and f[2] in (0, 1, 2, 3), name
        # flag 12: four real products in the plain forms (the reference's arithmetic) only
        assert f[3] == (0 if name.startswith("plain") and name != "plain_g3" else 1), name
        assert (f[2] == 1) == name.endswith("_trev") and (f[2] == 2) == name.endswith("_c4"), name
        assert (f[2] == 3) == name.endswith("_ibz"), name
    assert bench.FORMS["mixed_resolvent"][:2] == (2, 1)


def test_kernel_names_follow_the_product_count():
    assert bench.gemm_kernel_name("NC") == "gemm_kernel<double, c128, c128, c128, 0, 2, 64, 64, false>"
    assert bench.gemm_kernel_name("NC", True) == "gemm_kernel<double, c128, c128, c128, 0, 2, 64, 64, false, true>"
    assert bench.gemm_kernel_name("CN32") == "gemm_kernel<float, c64, c64, c64, 2, 0, 64, 64, false>"
