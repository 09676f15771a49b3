"""The device integrators (sph_integrate_rk4, sph_integrate_semi_implicit)
against the closed-form step maps of a linear system, x' = v, v' = -w^2 x,
rho_hat' = -k (rho_hat - 1): one RK4 step is the polynomial
I + hA + (hA)^2/2 + (hA)^3/6 + (hA)^4/24 (simulationBase.py:679-705), one
semi-implicit Euler step is [[1 - (hw)^2, h], [-h w^2, 1]]
(simulationBase.py:647-649).  With h w = 0.6 other stage weights differ at
1e-3 and more; fp32 arithmetic keeps the device within 2e-6 relative.
"""
import numpy as np
import pytest
import torch

from pytorchsph_amd import _lib as L

pytestmark = pytest.mark.gpu

W, K, H, R0 = 3.0, 2.0, 0.2, 1000.0


def phys_tensor(dt, dev):
    rec = np.zeros(1, dtype=L.PHYS_DTYPE)
    rec["dt"][0] = dt
    rec["rho0"][0] = R0
    return L.records_to_tensor(rec, dev)


def rhs(x, rho):
    """acc = -w^2 x, dpdt = -k rho0 (rho_hat - 1) (dpdt is in kg m^-3 s^-1,
    integrated as rho_hat += c dt dpdt / rho0)."""
    return -W * W * x, -K * R0 * (rho - 1.0)


def test_rk4_kernel_is_the_rk4_map(cuda):
    n = 64
    rng = np.random.default_rng(8)
    x0 = rng.normal(size=(n, 2)).astype(np.float32)
    v0 = rng.normal(size=(n, 2)).astype(np.float32)
    r0 = (1.0 + rng.uniform(0, 0.1, n)).astype(np.float32)
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(cuda)  # noqa: E731
    P = (t(x0), t(v0), t(r0))
    S0 = tuple(torch.empty_like(a) for a in P)
    S1 = tuple(torch.empty_like(a) for a in P)
    Kacc = torch.zeros((n, 5), dtype=torch.float32, device=cuda)
    acc_out = torch.zeros((n, 2), dtype=torch.float32, device=cuda)
    ps = torch.zeros(n, dtype=torch.int32, device=cuda)
    phys = phys_tensor(H, cuda)
    sh = L.stream_handle()
    for k, (cur, out) in enumerate([(P, S0), (S0, S1), (S1, S0), (S0, P)], start=1):
        a, dp = rhs(cur[0].cpu().numpy().astype(np.float64), cur[2].cpu().numpy())
        at, dpt = t(a.astype(np.float32)), t(dp.astype(np.float32))
        L.call("sph_integrate_rk4", k, L.ptr(ps), L.ptr(phys), n, L.ptr(P[0]), L.ptr(P[1]),
               L.ptr(P[2]), L.ptr(cur[1]), L.ptr(at), L.ptr(dpt), L.ptr(Kacc), L.ptr(out[0]),
               L.ptr(out[1]), L.ptr(out[2]), L.ptr(acc_out) if k == 4 else None, sh)
        torch.cuda.synchronize()
    A = np.array([[0.0, 1.0], [-W * W, 0.0]]) * H
    M = np.eye(2) + A + A @ A / 2 + A @ A @ A / 6 + A @ A @ A @ A / 24
    x1, v1, rho1 = (a.cpu().numpy().astype(np.float64) for a in P)
    for d in range(2):
        want = M @ np.stack([x0[:, d].astype(np.float64), v0[:, d].astype(np.float64)])
        assert np.allclose(x1[:, d], want[0], rtol=2e-6, atol=2e-6)
        assert np.allclose(v1[:, d], want[1], rtol=2e-6, atol=2e-6)
    z = -K * H
    want_r = 1 + (1 + z + z * z / 2 + z ** 3 / 6 + z ** 4 / 24) * (r0.astype(np.float64) - 1)
    assert np.allclose(rho1, want_r, rtol=1e-6)
    # acc_out is the weighted stage sum k1 + 2 k2 + 2 k3 + k4 (the
    # reference's fluidAcceleration after RK4, simulationBase.py:701)
    assert np.allclose(acc_out.cpu().numpy(), (v1 - v0) / H * 6, rtol=1e-4, atol=1e-4)


def test_semi_implicit_kernel_is_symplectic_euler(cuda):
    n = 64
    rng = np.random.default_rng(9)
    x0 = rng.normal(size=(n, 2)).astype(np.float32)
    v0 = rng.normal(size=(n, 2)).astype(np.float32)
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(cuda)  # noqa: E731
    xt, vt = t(x0), t(v0)
    at = t((-W * W * x0).astype(np.float32))
    ps = torch.zeros(n, dtype=torch.int32, device=cuda)
    h = 0.1
    phys = phys_tensor(h, cuda)
    xo, vo = torch.empty_like(xt), torch.empty_like(vt)
    L.call("sph_integrate_semi_implicit", L.ptr(ps), L.ptr(phys), n, L.ptr(xt), L.ptr(vt), None,
           L.ptr(vt), L.ptr(at), None, L.ptr(xo), L.ptr(vo), None, L.stream_handle())
    torch.cuda.synchronize()
    M = np.array([[1 - (h * W) ** 2, h], [-h * W * W, 1.0]])
    for d in range(2):
        want = M @ np.stack([x0[:, d].astype(np.float64), v0[:, d].astype(np.float64)])
        assert np.allclose(xo.cpu().numpy()[:, d], want[0], rtol=2e-6, atol=2e-6)
        assert np.allclose(vo.cpu().numpy()[:, d], want[1], rtol=2e-6, atol=2e-6)
