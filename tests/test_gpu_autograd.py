"""Differentiable degrid/grid (torch.autograd.Function) on the GPU.

Checks, as the differentiable-imaging workload asks:
  * dot-product adjoint test  Re<A x, v> = <x, A^H v>      (relative 1e-5)
  * central finite differences (eps = 1e-3) on seeded pixels of a weighted
    chi-squared loss                                        (relative 2e-2)
  * the autograd gradient against the direct-DFT gradient of the same loss
    on a visibility subset (oracle co_direct_adjoint)     (relative 3e-2 of max;
    the gridded operator approximates the DFT to the accuracy of the
    spheroidal kernel, the reference's own direct-vs-FFT bound is 3e-2 rms,
    lib.rs:89880-89887)
"""
import importlib

import numpy as np
import pytest
import torch

import synth

pytestmark = pytest.mark.gpu


def _op(gpu, nx, cell, u, v, w, planes=None):
    ag = importlib.import_module("casa_rs_amd.autograd")
    g = gpu.StandardGridder(nx, nx, cell, cell)
    if planes:
        g.attach_wproject(float(np.max(np.abs(w))), planes=planes)
    n = len(u)
    ones = torch.ones(n, dtype=torch.float32)
    b = gpu.VisibilityBatch(torch.from_numpy(u), torch.from_numpy(v), torch.from_numpy(w), ones,
                            torch.zeros(n, dtype=torch.complex64), ones)
    vs = g.visibility_set(b)
    return ag.DegridOperator(g, vs), g


CASES = [(False, 64, 20.0), (True, 96, 300.0)]


@pytest.mark.parametrize("wproj,nx,arcsec", CASES)
def test_dot_product_adjoint(gpu, wproj, nx, arcsec):
    cell = np.deg2rad(arcsec / 3600.0)
    rng = np.random.default_rng(5)
    n = 20000
    umax = 0.4 / cell
    u, v = rng.uniform(-umax, umax, n), rng.uniform(-umax, umax, n)
    w = rng.uniform(-200, 200, n) if wproj else np.zeros(n)
    op, g = _op(gpu, nx, cell, u, v, w, planes=16 if wproj else None)
    x = torch.from_numpy(rng.normal(size=(nx, nx)).astype(np.float32)).cuda()
    y = torch.from_numpy((rng.normal(size=n) + 1j * rng.normal(size=n)).astype(np.complex64)).cuda()
    Ax = op.forward_linear(x).to(torch.complex128)
    AHy = op.adjoint_linear(y).double()
    lhs = torch.sum(Ax.conj() * y.to(torch.complex128)).real.item()
    rhs = torch.sum(x.double() * AHy).item()
    assert abs(lhs - rhs) <= 1e-5 * abs(lhs)


@pytest.mark.parametrize("wproj,nx,arcsec", CASES)
def test_finite_differences(gpu, wproj, nx, arcsec):
    cell = np.deg2rad(arcsec / 3600.0)
    rng = np.random.default_rng(6)
    n = 20000
    umax = 0.4 / cell
    u, v = rng.uniform(-umax, umax, n), rng.uniform(-umax, umax, n)
    w = rng.uniform(-200, 200, n) if wproj else np.zeros(n)
    op, g = _op(gpu, nx, cell, u, v, w, planes=16 if wproj else None)
    x_true = np.zeros((nx, nx), np.float32)
    x_true[rng.integers(16, nx - 16, 6), rng.integers(16, nx - 16, 6)] = 1.0
    data = op.forward_linear(torch.from_numpy(x_true).cuda())
    data = data + 1e-3 * torch.randn_like(data)
    wts = torch.from_numpy(rng.uniform(0.5, 1.5, n).astype(np.float32)).cuda()
    # close to the truth, so the residual is small and the FP32 rounding of the
    # model visibilities stays far below the eps = 1e-3 difference
    x0 = torch.from_numpy((x_true + 1e-3 * rng.normal(size=(nx, nx))).astype(np.float32)).cuda()

    def loss64(img):
        r = (op(img) - data).to(torch.complex128)
        return torch.sum(wts.double() * (r.real ** 2 + r.imag ** 2))

    x = x0.clone().requires_grad_(True)
    L = loss64(x)
    L.backward()
    grad = x.grad.double().cpu().numpy()
    eps = 1e-3
    pix = [(int(a), int(b)) for a, b in zip(rng.integers(10, nx - 10, 12), rng.integers(10, nx - 10, 12))]
    for (a, b) in pix:
        xp, xm = x0.clone(), x0.clone()
        xp[a, b] += eps
        xm[a, b] -= eps
        fd = (loss64(xp).item() - loss64(xm).item()) / (2 * eps)
        assert abs(fd - grad[a, b]) <= 2e-2 * max(abs(grad[a, b]), 1e-3 * np.max(np.abs(grad)))


def test_gradient_against_direct_dft(gpu, oracle_mod):
    nx, cell = 64, np.deg2rad(20.0 / 3600.0)
    rng = np.random.default_rng(7)
    n = 10000
    umax = 0.3 / cell
    u, v = rng.uniform(-umax, umax, n), rng.uniform(-umax, umax, n)
    w = np.zeros(n)
    op, g = _op(gpu, nx, cell, u, v, w)
    r = (rng.normal(size=n) + 1j * rng.normal(size=n)).astype(np.complex64)
    # gradient of L = sum |A x - d|^2 at residual r is 2 Re(A^H r)
    got = 2.0 * op.adjoint_linear(torch.from_numpy(r).cuda()).double().cpu().numpy()
    # direct DFT adjoint: image += 2 Re(value e^{+i phi}) (lib.rs:73890-73910)
    ref = oracle_mod.direct_adjoint(nx, nx, cell, cell, u, v, w, r).astype(np.float64)
    # the gridded adjoint is exact only inside the well-corrected central region
    sl = slice(nx // 8, nx - nx // 8)
    err = np.max(np.abs(got[sl, sl] - ref[sl, sl])) / np.max(np.abs(ref[sl, sl]))
    assert err < 3e-2


def test_gradient_against_major_cycle_residual_image(gpu):
    """The chi^2 gradient of the autograd operator against the major cycle's
    residual image (compute_residual_standard_tiled, lib.rs:68862-68932):
    for L = sum w |A x - V|^2, dL/dx = -2 Re(A^H W (V - A x)) =
    -2 gx gy sumwt psf_peak R(x), with R the residual image (the centered
    inverse transform is F^H / (gx gy), and both paths apply the same grid
    correction).  FP32 on both sides: 2e-4 of the gradient's peak."""
    ag = importlib.import_module("casa_rs_amd.autograd")
    nx, cell = 64, np.deg2rad(20.0 / 3600.0)
    rng = np.random.default_rng(11)
    n = 30000
    umax = 0.4 / cell
    u, v = rng.uniform(-umax, umax, n), rng.uniform(-umax, umax, n)
    w = np.zeros(n)
    g = gpu.StandardGridder(nx, nx, cell, cell)
    wt = rng.uniform(0.5, 2.0, n).astype(np.float32)
    x_true = np.zeros((nx, nx), np.float32)
    x_true[rng.integers(12, nx - 12, 8), rng.integers(12, nx - 12, 8)] = rng.uniform(0.2, 1.0, 8)
    # data: the forward operator of x_true plus noise, through a set without data
    zeros = torch.zeros(n, dtype=torch.complex64)
    probe = g.visibility_set(gpu.VisibilityBatch(torch.from_numpy(u), torch.from_numpy(v), torch.from_numpy(w),
                                                 torch.from_numpy(wt), zeros, torch.ones(n)))
    V = (ag.DegridOperator(g, probe).forward_linear(torch.from_numpy(x_true).cuda()).cpu()
         + 0.01 * torch.randn(n, dtype=torch.complex64))
    vs = g.visibility_set(gpu.VisibilityBatch(torch.from_numpy(u), torch.from_numpy(v), torch.from_numpy(w),
                                              torch.from_numpy(wt), V, torch.ones(n)))
    op = ag.DegridOperator(g, vs)
    psf = g.new_grids(1)
    vs.grid(gpu.MODE_PSF, psf=psf)
    nsum = vs.sums[0]["normalization_sumwt"]
    _, _, peak = g.finish_products(psf, None, nsum)

    x0 = torch.from_numpy(0.5 * x_true).cuda()
    x = x0.clone().requires_grad_(True)
    Vd, wd = V.cuda(), torch.from_numpy(wt).cuda()
    r = op(x) - Vd
    L = torch.sum(wd * (r.real ** 2 + r.imag ** 2))
    L.backward()
    grad = x.grad.double().cpu().numpy()
    R = vs.major_cycle(x0[None], nsum, peak)[0].double().cpu().numpy()
    expect = -2.0 * g.info.gx * g.info.gy * nsum * float(peak[0]) * R
    err = float(np.max(np.abs(grad - expect)) / np.max(np.abs(expect)))
    print(f"gradient vs -2 gx gy sumwt peak R: {err:.2e} of peak")
    assert err < 2e-4
