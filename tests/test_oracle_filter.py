"""Pins of the oracle's pixel filters (render_target.h:70-184, create_pixel_filter render_target.cpp:28-44),
CPU only.

- eval: the Blackman-Harris window is (a0 + a1 + a2 + a3)^2 = 1 at the centre and (a0 - a1 + a2 - a3)^2 at
  the corners, and equals the published four-term window (Harris 1978) in FP64; the gaussian filter is the
  separable product of the normal density minus its value at the radius, zero at the edge.
- sample: offsets stay in [-r, r]^2; the mass of each sample-table cell is proportional to the filter at
  the cell centre (chi^2 against FP64 cell weights from the formulas), and inside a cell the offset is
  uniform (sample_linear); the box filter of any width is uniform over its square.
"""

import numpy as np
import pytest
from scipy import stats

import ks_oracle as ko

BH = (0.35875, 0.48829, 0.14128, 0.01168)


def _bh(x, r):
    t = (np.asarray(x, np.float64) + r) / (2 * r)
    a0, a1, a2, a3 = BH
    return a0 - a1 * np.cos(2 * np.pi * t) + a2 * np.cos(4 * np.pi * t) - a3 * np.cos(6 * np.pi * t)


def _gauss(x, sigma, r):
    g = lambda v: np.exp(-np.asarray(v, np.float64) ** 2 / (2 * sigma**2)) / np.sqrt(2 * np.pi * sigma**2)  # noqa: E731
    return g(x) - g(r)


def _prm(kind, width=0.0, sigma=0.0):
    return ko.params(pixel_filter=kind, filter_width=width, filter_sigma=sigma, spp=4)


@pytest.mark.parametrize("width", [3.0, 1.5, 4.0])
def test_blackman_harris_eval(width):
    r = width / 2
    rng = np.random.default_rng(1)
    xy = rng.uniform(-r, r, (5000, 2)).astype(np.float32)
    _, val, n = ko.pixel_filter(_prm(2, width), xy=xy)
    assert n == int(np.ceil(32 * r))
    want = _bh(xy[:, 0], r) * _bh(xy[:, 1], r)
    assert np.allclose(val, want, rtol=2e-5, atol=2e-6)
    _, c, _ = ko.pixel_filter(_prm(2, width), xy=np.array([[0, 0], [-r, -r], [r, r], [r * 1.01, 0]], np.float32))
    assert abs(c[0] - sum(BH) ** 2) < 1e-6 and sum(BH) == pytest.approx(1.0)
    assert abs(c[1] - (BH[0] - BH[1] + BH[2] - BH[3]) ** 2) < 1e-7 and c[3] == 0.0


@pytest.mark.parametrize("width,sigma", [(3.0, 0.5), (2.0, 0.3), (4.0, 1.0)])
def test_gaussian_eval(width, sigma):
    r = width / 2
    rng = np.random.default_rng(2)
    xy = rng.uniform(-r, r, (5000, 2)).astype(np.float32)
    _, val, _ = ko.pixel_filter(_prm(1, width, sigma), xy=xy)
    want = _gauss(xy[:, 0], sigma, r) * _gauss(xy[:, 1], sigma, r)
    assert np.allclose(val, want, rtol=1e-4, atol=1e-6)
    _, c, _ = ko.pixel_filter(_prm(1, width, sigma), xy=np.array([[0, 0], [r, 0]], np.float32))
    assert abs(c[0] - _gauss(0.0, sigma, r) ** 2) < 1e-5 * c[0]  # (1 / sqrt(2 pi) sigma - g(r))^2
    assert abs(c[1]) < 1e-7


@pytest.mark.parametrize("kind,width,sigma", [(1, 3.0, 0.5), (2, 3.0, 0.0), (2, 1.5, 0.0), (1, 2.0, 0.3)])
def test_tabulated_sampling(kind, width, sigma):
    r = width / 2
    rng = np.random.default_rng(3)
    m = 400_000
    u = rng.uniform(0, 1, (m, 2)).astype(np.float32)
    off, _, n = ko.pixel_filter(_prm(kind, width, sigma), u=u)
    assert np.all(np.abs(off) <= r)
    # cell masses: the filter at the n x n cell centres (FP64), normalised
    c = -r + (np.arange(n) + 0.5) * (2 * r / n)
    f1 = _gauss(c, sigma or 0.5, r) if kind == 1 else _bh(c, r)
    w = np.outer(f1, f1)  # separable: [y, x]
    ix = np.clip(((off[:, 0].astype(np.float64) + r) / (2 * r) * n).astype(int), 0, n - 1)
    iy = np.clip(((off[:, 1].astype(np.float64) + r) / (2 * r) * n).astype(int), 0, n - 1)
    obs = np.bincount(iy * n + ix, minlength=n * n).astype(np.float64)
    exp = (w / w.sum()).ravel() * m
    keep = exp > 10
    chi2 = (((obs - exp) ** 2)[keep] / exp[keep]).sum()
    assert stats.chi2.sf(chi2, keep.sum() - 1) > 1e-4, chi2
    assert obs[~keep].sum() <= 3 * exp[~keep].sum() + 20
    # uniform inside the cells: the fractional position in the cell is U(0, 1)
    fx = (off[:, 0].astype(np.float64) + r) / (2 * r) * n - ix
    assert stats.kstest(np.clip(fx, 0, 1), "uniform").pvalue > 1e-4


@pytest.mark.parametrize("width", [1.0, 2.0, 0.5])
def test_box_sampling(width):
    rng = np.random.default_rng(4)
    u = rng.uniform(0, 1, (100_000, 2)).astype(np.float32)
    off, _, n = ko.pixel_filter(_prm(0, width), u=u)
    assert n == 0
    # lerp(-r, r, u) = (1 - u)(-r) + u r = (2u - 1) r
    assert np.allclose(off, (2 * u.astype(np.float64) - 1) * width / 2, atol=1e-6)
