"""Pins of the oracle's SPH kernel functions (density/device.cuh:10-88,
pbf.cu:26-32) by normalisation, known values, and finite differences."""
import numpy as np
import pytest

import oracle

H = float(np.float32(0.1))  # the oracle takes h as FP32


def radial(kind, r, precision="f64"):
    d = np.stack([r, np.zeros_like(r), np.zeros_like(r)])
    return oracle.kernel_values(kind, H, d, precision)


def simpson(y, x):
    n = len(x) - 1
    assert n % 2 == 0
    dx = (x[-1] - x[0]) / n
    return dx / 3.0 * (y[0] + y[-1] + 4.0 * y[1:-1:2].sum() + 2.0 * y[2:-1:2].sum())


@pytest.mark.parametrize("kind", [oracle.POLY6, oracle.CUBIC])
def test_normalised(kind):
    # 4 pi int_0^h W(r) r^2 dr = 1
    r = np.linspace(0.0, H, 20001)
    w, _ = radial(kind, r)
    assert abs(4.0 * np.pi * simpson(w * r * r, r) - 1.0) < 1e-9


def test_spiky_gradient_integrates_to_normalised_kernel():
    # |grad W| integrated from r to h rebuilds W(r); W must then integrate to 1
    r = np.linspace(0.0, H, 4001)[1:]
    _, g = radial(oracle.SPIKY, r)
    assert np.all(g[0] <= 0.0)  # points from the neighbour towards the particle: W decreases outwards
    grad = -g[0]
    w = np.array([simpson_tail(grad, r, k) for k in range(len(r))])
    assert abs(4.0 * np.pi * np.trapezoid(w * r * r, r) - 1.0) < 1e-3


def simpson_tail(y, x, k):
    return np.trapezoid(y[k:], x[k:])


def test_cubic_known_values():
    sigma = 8.0 / (np.pi * H**3)
    w, _ = radial(oracle.CUBIC, np.array([0.0, 0.5 * H, H, 1.5 * H]))
    np.testing.assert_allclose(w[:2], [sigma, sigma / 4.0], rtol=1e-14)
    assert w[2] == 0.0 and w[3] == 0.0


def test_poly6_known_values():
    c = 315.0 / (64.0 * np.pi * H**9)
    w, _ = radial(oracle.POLY6, np.array([0.0, 0.5 * H, H]))
    np.testing.assert_allclose(w[:2], [c * H**6, c * (0.75 * H * H) ** 3], rtol=1e-14)
    assert w[2] == 0.0


def test_viscosity_laplacian_mueller():
    # Mueller et al. 2003: lap W_visc(r) = 45 / (pi h^6) (h - r)
    r = np.array([0.0, 0.25 * H, 0.9 * H, H])
    w, _ = radial(oracle.VISCOSITY_LAPLACIAN, r)
    np.testing.assert_allclose(w[:3], 45.0 / (np.pi * H**6) * (H - r[:3]), rtol=1e-14)
    assert w[3] == 0.0


@pytest.mark.parametrize("kind", [oracle.POLY6, oracle.CUBIC])
def test_gradient_is_derivative_of_value(kind):
    rng = np.random.default_rng(3)
    d = rng.uniform(-0.06, 0.06, size=(3, 200))
    d = d[:, (np.linalg.norm(d, axis=0) > 0.01) & (np.linalg.norm(d, axis=0) < 0.095)]
    _, g = oracle.kernel_values(kind, H, d, "f64")
    eps = 1e-7
    for a in range(3):
        e = np.zeros((3, 1))
        e[a] = eps
        wp, _ = oracle.kernel_values(kind, H, d + e, "f64")
        wm, _ = oracle.kernel_values(kind, H, d - e, "f64")
        fd = (wp - wm) / (2 * eps)
        np.testing.assert_allclose(g[a], fd, rtol=1e-5, atol=1e-6 * np.abs(fd).max())


def test_f32_matches_f64():
    rng = np.random.default_rng(5)
    d = rng.uniform(-0.1, 0.1, size=(3, 1000)).astype(np.float32)
    for kind in (oracle.POLY6, oracle.CUBIC, oracle.SPIKY, oracle.VISCOSITY_LAPLACIAN):
        w32, g32 = oracle.kernel_values(kind, H, d, "f32")
        w64, g64 = oracle.kernel_values(kind, H, d.astype(np.float64), "f64")
        assert np.abs(w32 - w64).max() <= 2e-5 * max(np.abs(w64).max(), 1e-30)
        assert np.abs(g32 - g64).max() <= 2e-5 * max(np.abs(g64).max(), 1e-30)
