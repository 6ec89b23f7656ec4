"""The Python binding refuses mis-sized, mis-typed and host-resident arrays with ValueError
before any call into libthcgw (the C ABI takes raw pointers, so a wrong size would be read
or written out of bounds on the GPU).  No GPU: the context object is assembled by hand and
its library handle is None, so any call that got through to the library would fail with
an AttributeError instead of the ValueError asserted here."""
import numpy as np
import pytest
import torch

from coqui_amd.thcgw import LearnedDyson, Mlp, ThcBatch, ThcGW, check_array


class _Iaft:
    nt, nt_half, nw_half, nw_b, nw_f = 6, 3, 4, 7, 6


def fake_ctx(kgrid=(2, 1, 1), no=3, nmu=5):
    c = object.__new__(ThcGW)
    c._torch = torch
    c.lib = None
    c.h = None
    c.kgrid = kgrid
    c.nk = int(np.prod(kgrid))
    c.no, c.nmu = no, nmu
    c.iaft = _Iaft()
    return c


def cplx(*shape):
    return torch.zeros(shape, dtype=torch.complex128)


def test_check_array_rules():
    a = np.zeros((2, 3), np.complex128)
    check_array(a, (2, 3), "a")
    check_array(a, (1, 2, 3), "a")  # axes of length 1 may be added or dropped
    with pytest.raises(ValueError, match="shape"):
        check_array(a, (3, 2), "a")
    with pytest.raises(ValueError, match="dtype"):
        check_array(a.astype(np.complex64), (2, 3), "a")
    with pytest.raises(ValueError, match="contiguous"):
        check_array(np.zeros((3, 2), np.complex128).T, (2, 3), "a")
    with pytest.raises(ValueError, match="GPU"):
        check_array(cplx(2, 3), (2, 3), "a")


@pytest.mark.parametrize("call", [
    lambda c: c.pi(cplx(c.nt, 1, c.nk, c.no, c.no + 1), cplx(1, c.nk, c.nmu, c.no)),
    lambda c: c.pi(cplx(c.nt, 1, c.nk, c.no, c.no), cplx(1, c.nk, c.nmu - 1, c.no)),
    lambda c: c.tau_to_w(cplx(c.nt_half + 1, c.nk, c.nmu, c.nmu)),
    lambda c: c.dyson_w(cplx(c.nw_half, c.nk, c.nmu, c.nmu), cplx(c.nk + 1, c.nmu, c.nmu)),
    lambda c: c.w_to_tau(cplx(c.nw_half, c.nk, c.nmu, c.nmu), out=cplx(c.nk, c.nt_half, c.nmu, c.nmu - 1)),
    lambda c: c.sigma(cplx(c.nt, 1, c.nk, c.no, c.no), cplx(c.nk, c.nt_half - 1, c.nmu, c.nmu),
                      cplx(1, c.nk, c.nmu, c.no)),
    lambda c: c.gw_iteration(cplx(c.nt, 1, c.nk, c.no, c.no), cplx(1, c.nk, c.nmu, c.no), cplx(c.nk, c.nmu, c.nmu),
                             Sigma=cplx(c.nt - 1, 1, c.nk, c.no, c.no)),
    lambda c: c.gw_iteration(cplx(c.nt, 1, c.nk, c.no, c.no), cplx(1, c.nk, c.nmu, c.no), cplx(c.nk, c.nmu, c.nmu),
                             W_qt=cplx(c.nk, c.nt_half, c.nmu + 1, c.nmu)),
    lambda c: c.rpa_ec(cplx(c.nw_half, c.nk, c.nmu, c.nmu + 2), cplx(c.nk, c.nmu, c.nmu)),
    lambda c: c.update_G(cplx(c.nk, c.no, c.no), np.zeros((c.nk, c.no + 1))),
    lambda c: c.update_G(cplx(c.nk, c.no, c.no), torch.zeros((c.nk, c.no), dtype=torch.float32)),
    lambda c: c.primary_to_aux(cplx(1, c.nk, c.nmu, c.no), cplx(7)),
    lambda c: c.hermitize(cplx(2, 3, 4)),
    lambda c: c.solve_qp(cplx(c.nt, 1, c.nk, c.no, c.no), cplx(c.nk, c.no, c.no), cplx(c.nk, c.no, c.no),
                         np.zeros((c.nk, c.no - 1))),
    lambda c: c.hf(cplx(1, c.nk, c.no, c.no), cplx(1, c.nk, c.nmu, c.no), cplx(c.nk, c.nmu, c.nmu),
                   out=cplx(1, c.nk, c.no, c.no + 1)),
    lambda c: c.dyson_g(cplx(1, c.nk, c.no, c.no), cplx(c.nt, 1, c.nk, c.no, c.no), 0.0, H0=cplx(c.nk, c.no)),
    lambda c: c.eps_inv_head(cplx(c.nk, c.nt_half, c.nmu, c.nmu), cplx(c.nk, c.nmu), np.zeros((c.nk, 2)), 1.0),
    lambda c: c.gm_energy(cplx(c.nt, 1, c.nk, c.no, c.no), np.zeros((c.nt, 1, c.nk, c.no, c.no), np.complex64)),
], ids=lambda f: "")
def test_mis_sized_arrays_are_refused(call):
    with pytest.raises(ValueError):
        call(fake_ctx())


def test_host_resident_output_is_refused():
    c = fake_ctx()
    with pytest.raises(ValueError, match="GPU"):
        c.w_to_tau(cplx(c.nw_half, c.nk, c.nmu, c.nmu), out=cplx(c.nk, c.nt_half, c.nmu, c.nmu))


def test_batch_and_learned_checks():
    c = fake_ctx()
    b = object.__new__(ThcBatch)
    b.ctxs, b.lib, b._torch = [c], None, torch
    ok = [cplx(c.nt, 1, c.nk, c.no, c.no)], [cplx(1, c.nk, c.nmu, c.no)], [cplx(c.nk, c.nmu, c.nmu)]
    with pytest.raises(ValueError, match="one array per system"):
        b.gw_iteration(*ok, [], [])
    with pytest.raises(ValueError, match="system 0 G: expected a GPU tensor"):
        b.gw_iteration(*ok, [cplx(c.nt, 1, c.nk, c.no, c.no)], [cplx(c.nk, c.nt_half, c.nmu, c.nmu)])
    L = object.__new__(LearnedDyson)
    L.ctx, L.lib, L._torch, L.h = c, None, torch, None
    with pytest.raises(ValueError, match="PiW"):
        L.dyson(cplx(c.nw_half - 1, c.nk, c.nmu, c.nmu), cplx(c.nk, c.nmu, c.nmu))


def test_mlp_checks():
    m = object.__new__(Mlp)
    m._torch, m.lib, m.h = torch, None, None
    m.n_in, m.width, m.n_out, m.n_params = 4, 8, 3, 10
    with pytest.raises(ValueError):
        m.forward(torch.zeros((5, 3), dtype=torch.float32))
    with pytest.raises(ValueError):
        m.backward(np.zeros((5, 5), np.float32), np.zeros((5, 3), np.float32))
    with pytest.raises(ValueError):
        m.set_params(np.zeros(9, np.float32))
    with pytest.raises(ValueError):
        m.train_mse(torch.zeros((5, 4), dtype=torch.float64), torch.zeros((5, 3), dtype=torch.float32), 1e-3)


def test_bz_checks():
    """The Brillouin-zone symmetry binding: mis-shaped tables are refused before
    thcgw_set_bz, the wedge entry points need a symmetry set, and their arrays have the
    wedge's shapes (irreducible k for G and Sigma, irreducible q for Pi, W and V)."""
    from coqui_amd.system import make_system_bz

    _, bz = make_system_bz((3, 2, 2), 2, 4, system_id=1, kind="trev")
    c = fake_ctx(kgrid=(3, 2, 2), no=2, nmu=4)
    with pytest.raises(ValueError, match="no Brillouin-zone symmetry"):
        c.pi_ibz(cplx(c.nt, 1, bz["nk_ibz"], c.no, c.no), cplx(1, c.nk, c.nmu, c.no))
    with pytest.raises(ValueError, match="qk_to_k2"):
        c.set_bz(dict(bz, qk_to_k2=bz["qk_to_k2"][:-1]))
    with pytest.raises(ValueError, match=r"bz\[D\]"):
        c.set_bz(dict(bz, D=bz["D"][:, :-1]))
    c._bz = (bz["nk_ibz"], bz["nq_ibz"])  # as after a successful thcgw_set_bz
    nki, nqi = c._bz
    full = cplx(c.nt, 1, c.nk, c.no, c.no)  # G on every k, not on the irreducible k
    with pytest.raises(ValueError, match="G: expected shape"):
        c.pi_ibz(full, cplx(1, c.nk, c.nmu, c.no))
    with pytest.raises(ValueError, match="G: expected shape"):
        c.sigma_ibz(full, cplx(nqi, c.nt_half, c.nmu, c.nmu), cplx(1, c.nk, c.nmu, c.no))
    with pytest.raises(ValueError, match="G: expected shape"):
        c.gw_iteration_ibz(full, cplx(1, c.nk, c.nmu, c.no), cplx(nqi, c.nmu, c.nmu))
    # the wedge shapes get past the shape checks (then stop at the host-resident test)
    with pytest.raises(ValueError, match="GPU"):
        c.pi_ibz(cplx(c.nt, 1, nki, c.no, c.no), cplx(1, c.nk, c.nmu, c.no))
