"""GPU adjoint sensitivities through the C ABI: the element contraction kernel
against the FP64 CPU adjoint (oracle) on random vectors, and compliance /
tip-displacement gradients of the SIMP block (rho_e -> E_e = Emin + rho^3
(E0 - Emin)) through a torch.autograd.Function, checked by central finite
differences with h = 1e-6 in rho on 32 random elements. Each difference
quotient re-assembles K on the device: every CSR slot is summed in element
order, so K(rho +- h) differs from K(rho) only where element e contributes,
and the quotients are not disturbed by summation-order noise."""
import numpy as np
import pytest

import oracle as O
from vibestran_amd import decks
from vibestran_amd import model as M
from vibestran_amd.assembly import (assemble_model, compliance_gradients, dof_map_device,
                                    element_sensitivity)
from tests.test_gpu_assembly import mixed_mesh
from tests.test_gpu_shell import warped_mesh

pytestmark = pytest.mark.gpu


def _kernel_vs_oracle(be, m, params):
    eq, n = dof_map_device(be, m)
    assemble_model(be, m, eq, n)
    rng = np.random.default_rng(9)
    u, v = rng.standard_normal(n), rng.standard_normal(n)
    for p in params:
        g = element_sensitivity(be, m, p, u, v)
        go = O.element_sensitivity(n, m, eq, p, u, v)
        assert np.abs(g - go).max() <= 1e-11 * np.abs(go).max(), p


@pytest.mark.parametrize("mk", [lambda: decks.config4_simp(nx=8, ny=4, nz=3, seed=2),
                                mixed_mesh,
                                lambda: warped_mesh(form=0), lambda: warped_mesh(form=1),
                                lambda: decks.cquad4_tube(n_circ=12, n_axial=6, n_cases=1)])
def test_sensitivity_kernel_matches_cpu_adjoint(gpu_backend_factory, mk):
    m = mk()
    _kernel_vs_oracle(gpu_backend_factory(), m, (0, 1) if m.thickness is not None else (0,))


def test_sensitivity_needs_mesh(gpu_backend_factory):
    from vibestran_amd import _lib as L
    be = gpu_backend_factory()
    out = np.zeros(4)
    x = np.zeros(4)
    assert L.load().vbh_element_sensitivity(be._h, 0, L.dp(x), L.dp(x), L.dp(out)) == L.VBH_ERR_STATE


torch = pytest.importorskip("torch")


class SimpCompliance(torch.autograd.Function):
    """C(rho) = F^T u, K(E(rho)) u = F, E = Emin + rho^3 (E0 - Emin); backward
    through the GPU adjoint kernel (dC/dE_e = -u_e^T K_e u_e / E_e)."""

    @staticmethod
    def forward(ctx, rho, be, m, eq, n, F):
        r = rho.detach().cpu().numpy()
        mm = M.Model(**{**m.__dict__, "E": m.meta["Emin"] + r ** 3 * (m.meta["E0"] - m.meta["Emin"])})
        C, dC_dE, _, _ = compliance_gradients(be, mm, eq, n, F)
        dE = 3.0 * r ** 2 * (m.meta["E0"] - m.meta["Emin"])
        ctx.save_for_backward(torch.from_numpy(dC_dE * dE))
        return torch.tensor(C, dtype=torch.float64)

    @staticmethod
    def backward(ctx, g):
        (grad,) = ctx.saved_tensors
        return g * grad, None, None, None, None, None


def test_simp_compliance_autograd_vs_fd(gpu_backend_factory):
    m = decks.config4_simp(nx=10, ny=5, nz=3, seed=11)
    be = gpu_backend_factory(tolerance=1e-13, max_iters=200000)
    eq, n = dof_map_device(be, m)
    F = M.assemble_rhs(m, eq, n, m.loads[0])
    rho = torch.tensor(m.meta["density"], dtype=torch.float64, requires_grad=True)
    C = SimpCompliance.apply(rho, be, m, eq, n, F)
    (grad,) = torch.autograd.grad(C, rho)
    # the same gradient from the CPU adjoint
    mm = M.Model(**{**m.__dict__, "E": m.meta["Emin"] + m.meta["density"] ** 3 * (m.meta["E0"] - m.meta["Emin"])})
    u = be.solve(assemble_model(be, mm, eq, n), F)
    go = O.element_sensitivity(n, mm, eq, 0, u, u) * 3.0 * m.meta["density"] ** 2 * (m.meta["E0"] - m.meta["Emin"])
    assert np.abs(grad.numpy() - go).max() <= 1e-10 * np.abs(go).max()
    h = 1e-6

    def compliance(r):  # C(rho) on the device assembly, solved on the GPU
        mr = M.Model(**{**m.__dict__, "E": m.meta["Emin"] + r ** 3 * (m.meta["E0"] - m.meta["Emin"])})
        return float(F @ be.solve(assemble_model(be, mr, eq, n), F))

    r0 = m.meta["density"]
    assert abs(compliance(r0) - C.item()) <= 1e-9 * abs(C.item())
    for e in np.random.default_rng(5).choice(m.ne, 32, replace=False):
        rp, rm = r0.copy(), r0.copy()
        rp[e] += h
        rm[e] -= h
        fd = (compliance(rp) - compliance(rm)) / (2 * h)
        assert abs(grad[e].item() - fd) <= 1e-4 * abs(fd) + 1e-7 * grad.abs().max().item(), e


def test_simp_tip_displacement_gradient_fd(gpu_backend_factory):
    m = decks.config4_simp(nx=10, ny=5, nz=3, seed=12)
    be = gpu_backend_factory(tolerance=1e-13, max_iters=200000)
    eq, n = dof_map_device(be, m)
    F = M.assemble_rhs(m, eq, n, m.loads[0])
    tip = int(eq[m.meta["tip_node"], 2])
    C, dC, ut, dut = compliance_gradients(be, m, eq, n, F, tip_dof=tip)
    assert abs(C - (-1000.0) * ut) <= 1e-12 * abs(C)  # single point load: C = F_tip u_tip
    assert np.abs(dC - (-1000.0) * dut).max() <= 1e-8 * np.abs(dC).max()
    E0, Emin, r0 = m.meta["E0"], m.meta["Emin"], m.meta["density"]
    dut_drho = dut * 3.0 * r0 ** 2 * (E0 - Emin)
    h = 1e-6

    def u_tip(r):
        mr = M.Model(**{**m.__dict__, "E": Emin + r ** 3 * (E0 - Emin)})
        return be.solve(assemble_model(be, mr, eq, n), F)[tip]

    for e in np.random.default_rng(6).choice(m.ne, 32, replace=False):
        rp, rm = r0.copy(), r0.copy()
        rp[e] += h
        rm[e] -= h
        fd = (u_tip(rp) - u_tip(rm)) / (2 * h)
        assert abs(dut_drho[e] - fd) <= 1e-4 * abs(fd) + 1e-7 * np.abs(dut_drho).max(), e


def test_shell_thickness_gradient_fd(gpu_backend_factory):
    m = warped_mesh(nx=5, ny=3, form=0)
    be = gpu_backend_factory(tolerance=1e-13, max_iters=200000)
    eq, n = dof_map_device(be, m)
    F = np.random.default_rng(4).standard_normal(n)
    K = assemble_model(be, m, eq, n)
    u = be.solve(K, F)
    g = element_sensitivity(be, m, 1, u, u)  # dC/dt_e
    for e in range(0, m.ne, 3):
        h = 1e-3 * m.thickness[e]
        tp, tm = m.thickness.copy(), m.thickness.copy()
        tp[e] += h
        tm[e] -= h
        cp = F @ be.solve(assemble_model(be, M.Model(**{**m.__dict__, "thickness": tp}), eq, n), F)
        cm = F @ be.solve(assemble_model(be, M.Model(**{**m.__dict__, "thickness": tm}), eq, n), F)
        fd = (cp - cm) / (2 * h)
        assert abs(g[e] - fd) <= 1e-4 * abs(fd) + 2e-7 * np.abs(g).max(), e
