"""Averaging functions (electrodynamics/kernel_lib.h:560-634) on the GPU.

- Ker_BasicAverFunc against the reference compiled from its sources (golden
  fixture aver.npz, 400 pairs inside and outside eps, r = 0 and r = eps).
- Ker_AverDiv and Ker_AverDiv_NearEdge cannot be instantiated in the
  reference (they read KernelParam::eps_edge, which KernelParam lacks), so
  they are pinned by the mathematics: each is the x-gradient of a closed form,
      Ker_AverDiv          = grad_x (2 - rho^2) e^{-rho^2} / (pi eps^2),
      Ker_AverDiv_NearEdge = grad_x (6 - 4 rho^2) e^{-rho^2} / (pi eps^2),
  rho = |x - y| / eps, checked by central differences (a wrong sign, power or
  constant in the kernel fails it).
"""
import numpy as np
import pytest

from _util import golden

pytestmark = pytest.mark.gpu


def test_basic_aver_vs_reference(ctx):
    g = golden("aver.npz")
    got = np.array([ctx.aver_batch(2, g["x"][i:i + 1], g["y"][i:i + 1], g["eps"][i])[0] for i in range(len(g["eps"]))])
    ref = g["basic"]
    assert np.all(got[:, 1:] == 0.0)
    assert np.array_equal(got[:, 0] == 0.0, ref == 0.0)           # same support (r > eps -> 0)
    nz = ref != 0.0
    assert np.max(np.abs(got[nz, 0] - ref[nz]) / np.abs(ref[nz])) <= 1e-14


@pytest.mark.parametrize("which,c0,c2", [(0, 2.0, -1.0), (1, 6.0, -4.0)])
def test_divergence_kernels_are_gradients(ctx, which, c0, c2):
    rng = np.random.default_rng(81 + which)
    eps = 0.37
    x = rng.normal(size=(300, 3))
    y = x + rng.normal(size=(300, 3)) * eps * rng.uniform(0.05, 2.5, (300, 1))
    got = ctx.aver_batch(which, x, y, eps)

    def g(xx):
        rho2 = np.sum((xx - y) ** 2, axis=1) / eps ** 2
        return (c0 + c2 * rho2) * np.exp(-rho2) / (np.pi * eps ** 2)
    h = 1e-6 * eps
    fd = np.stack([(g(x + h * e) - g(x - h * e)) / (2 * h) for e in np.eye(3)], axis=1)
    scale = np.abs(fd).max()
    assert np.abs(got - fd).max() <= 1e-7 * scale
    # and they vanish at r = 0 (odd in x - y)
    assert np.all(ctx.aver_batch(which, x[:4], x[:4], eps) == 0.0)


def test_aver_argument_errors(ctx):
    import bielcc_amd as B
    with pytest.raises(B.BielccError):
        ctx.aver_batch(3, np.zeros((1, 3)), np.zeros((1, 3)), 1.0)
    with pytest.raises(B.BielccError):
        ctx.aver_batch(0, np.zeros((1, 3)), np.zeros((1, 3)), 0.0)
