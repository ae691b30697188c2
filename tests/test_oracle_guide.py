"""Pins the CPU oracle of the path guide (oracle/ks_guide_oracle.py) against facts it does not restate:
torch's BF16 rounding, exact reproduction of linear fields by trilinear interpolation, vMF normalisation
by quadrature, the closed-form mean cosine (VMF::mean_cosine, vmf.h:58-65), finite differences of the
loss and torch.optim.Adam. No GPU is needed."""

import math

import numpy as np
import pytest
import torch

import ks_guide_oracle as go


def test_bf16_matches_torch():
    rng = np.random.default_rng(1)
    x = np.concatenate([rng.standard_normal(20000) * 10.0 ** rng.integers(-8, 8, 20000),
                        [0.0, -0.0, 1.0, 1 + 2**-8, 1 + 3 * 2**-8, 65504.0, 1e-40, -3.0e38]]).astype(np.float32)
    ref = torch.from_numpy(x).to(torch.bfloat16).to(torch.float64).numpy()
    assert np.array_equal(go.bf16(x), ref)


def test_resolutions_closed_form():
    r = go.resolutions(16.0, 2048.0)
    assert r[0] == 16 and r[-1] == 2048 and all(b > a for a, b in zip(r, r[1:]))
    b = 2 ** (7 / 15)  # 128^(1/15)
    assert r == [int(math.floor(16 * b**l + 1e-9)) for l in range(16)]
    assert go.resolutions(4.0, 4.0) == [4] * 16


def test_dense_index_is_a_bijection():
    for res in (3, 4, 15):
        side = res + 1
        g = np.stack(np.meshgrid(*[np.arange(side)] * 3, indexing="ij"), -1).reshape(-1, 3)
        idx = go.grid_index(g[:, 0], g[:, 1], g[:, 2], res, 12)
        assert sorted(idx.tolist()) == list(range(side**3))


def test_hash_index_uses_wrapping_products():
    rng = np.random.default_rng(2)
    p = rng.integers(0, 3000, size=(500, 3))
    idx = go.grid_index(p[:, 0], p[:, 1], p[:, 2], 2048, 19)
    ref = [(int(x) ^ ((int(y) * 2654435761) % 2**32) ^ ((int(z) * 805459861) % 2**32)) & (2**19 - 1) for x, y, z in p]
    assert idx.tolist() == ref


def test_trilinear_reproduces_linear_fields():
    """A grid holding a.v/res + b at every vertex v interpolates to a.(u res)/res + b (dense and hashed
    levels); samples with a corner whose table entry collides with another vertex are skipped."""
    lay = go.Layout(log2_t=16, base=4.0, max_res=512.0, bound_lo=(-1, 0, 2), bound_hi=(3, 1, 5))
    rng = np.random.default_rng(3)
    x = rng.uniform([-1, 0, 2], [3, 1, 5], size=(60, 3)).astype(np.float32)
    u = np.clip((x - lay.lo) * lay.inv_ext, np.float32(0), np.float32(1)).astype(np.float32)
    a, b = np.array([0.3, -0.7, 0.2]), 0.1
    for l, (idx, w) in enumerate(lay.corners(x)):
        assert np.allclose(w.sum(axis=1), 1.0, atol=1e-6)
        res = lay.res[l]
        xs = (u * np.float32(res)).astype(np.float32)
        cell = np.minimum(np.floor(xs), res - 1).astype(np.float64)
        owner = {}
        vals = np.zeros((len(x), 8))
        for k in range(8):
            v = cell + np.array([k & 1, (k >> 1) & 1, (k >> 2) & 1])
            vals[:, k] = (v / res) @ a + b
            for i in range(len(x)):
                owner.setdefault(int(idx[i, k]), set()).add(tuple(v[i]))
        ok = np.array([all(len(owner[int(j)]) == 1 for j in idx[i]) for i in range(len(x))])
        f = (w * vals).sum(axis=1)
        want = (xs.astype(np.float64) / res) @ a + b
        assert ok.mean() > 0.8, l
        assert np.allclose(f[ok], want[ok], atol=2e-6), l
        if l == 0:  # dense level: no collisions at all
            assert ok.all()


def _random_mixture(rng, n, kmax=1e3):
    w = rng.uniform(0.0, 1.0, (n, 8))
    w[:, rng.integers(0, 8)] = 0.0  # some empty lobes
    w /= w.sum(axis=1, keepdims=True)
    mu = rng.standard_normal((n, 8, 3))
    mu /= np.linalg.norm(mu, axis=2, keepdims=True)
    kappa = np.exp(rng.uniform(np.log(1e-2), np.log(kmax), (n, 8)))
    return go.mixture_array(w, mu, kappa)


def _sphere_quadrature(n_theta=1500, n_phi=600):
    th = (np.arange(n_theta) + 0.5) * np.pi / n_theta
    ph = (np.arange(n_phi) + 0.5) * 2 * np.pi / n_phi
    T, P = np.meshgrid(th, ph, indexing="ij")
    d = np.stack([np.sin(T) * np.cos(P), np.sin(T) * np.sin(P), np.cos(T)], -1).reshape(-1, 3)
    dA = (np.sin(T) * (np.pi / n_theta) * (2 * np.pi / n_phi)).ravel()
    return d, dA


@pytest.mark.parametrize("kappa", [0.0, 1e-5, 1e-3, 0.5, 5.0, 50.0, 400.0])
def test_vmf_normalised(kappa):
    d, dA = _sphere_quadrature()
    mu = np.array([0.3, -0.5, 0.8]) / np.linalg.norm([0.3, -0.5, 0.8])
    p = go.vmf_eval(mu[None], np.full(len(d), kappa), d)
    assert (p * dA).sum() == pytest.approx(1.0, rel=2e-3)


def test_mixture_pdf_normalised_and_matches_sum():
    rng = np.random.default_rng(4)
    mix = _random_mixture(rng, 3, kmax=200.0)
    d, dA = _sphere_quadrature()
    for m in mix:
        p = go.mixture_pdf(np.repeat(m[None], len(d), 0), d)
        assert (p * dA).sum() == pytest.approx(1.0, rel=3e-3)
        lob = m[8:].reshape(8, 4)
        direct = sum(m[k] * go.vmf_eval(lob[k, :3][None], np.full(len(d), lob[k, 3]), d) for k in range(8))
        assert np.allclose(p, direct, rtol=1e-10)


@pytest.mark.parametrize("kappa", [1e-3, 0.7, 8.0, 300.0, 1e4])
def test_vmf_sample_mean_cosine(kappa):
    """E[mu . w] = coth(kappa) - 1/kappa (VMF::mean_cosine, vmf.h:58-65), stratified samples."""
    n = 200_000
    u0 = (np.arange(n) + 0.5) / n
    u1 = np.random.default_rng(5).uniform(size=n)
    mu = np.array([0.0, 0.6, -0.8])
    w = go.vmf_sample(np.repeat(mu[None], n, 0), np.full(n, kappa), u0, u1)
    assert np.allclose(np.linalg.norm(w, axis=1), 1.0, atol=1e-9)
    want = 1.0 / math.tanh(kappa) - 1.0 / kappa
    assert (w @ mu).mean() == pytest.approx(want, abs=2e-5 + 1e-4 * (1 - want))


def test_vmf_sample_uniform_at_zero_kappa():
    n = 100_000
    rng = np.random.default_rng(6)
    w = go.vmf_sample(np.tile([0.0, 0.0, 1.0], (n, 1)), np.zeros(n), rng.uniform(size=n), rng.uniform(size=n))
    assert np.abs(w.mean(axis=0)).max() < 0.01
    assert np.allclose((w**2).mean(axis=0), 1 / 3, atol=0.01)


def test_mixture_sample_histogram_follows_pdf():
    """Chi-square of mixture samples in 24 x 12 equal-area bins against the quadrature of the pdf."""
    rng = np.random.default_rng(7)
    m = _random_mixture(rng, 1, kmax=30.0)[0]
    n = 400_000
    u = rng.uniform(size=(n, 2))
    w, _ = go.mixture_sample(np.repeat(m[None], n, 0), u)
    nz, nph = 24, 12

    def bins(d):
        iz = np.minimum(((d[:, 2] + 1) / 2 * nz).astype(int), nz - 1)
        ip = np.minimum(((np.arctan2(d[:, 1], d[:, 0]) + np.pi) / (2 * np.pi) * nph).astype(int), nph - 1)
        return iz * nph + ip

    counts = np.bincount(bins(w), minlength=nz * nph)
    d, dA = _sphere_quadrature(1200, 720)
    expect = np.bincount(bins(d), weights=go.mixture_pdf(np.repeat(m[None], len(d), 0), d) * dA,
                         minlength=nz * nph) * n
    keep = expect > 20
    chi2 = ((counts[keep] - expect[keep]) ** 2 / expect[keep]).sum()
    dof = keep.sum() - 1
    assert chi2 < dof + 5 * math.sqrt(2 * dof)


def test_head_gradient_finite_differences():
    rng = np.random.default_rng(8)
    o = rng.standard_normal((6, 48)) * 0.8
    o[:, 4:40:5] = rng.uniform(-3, 6, (6, 8))
    o[0, 4] = -20.0  # clamped kappa: zero gradient
    wi = rng.standard_normal((6, 3))
    c = rng.uniform(0.2, 2.0, 6)
    _, g = go.head_backward(o, wi, c)
    h = 1e-6
    for i in range(6):
        for j in range(40):
            op, om = o.copy(), o.copy()
            op[i, j] += h
            om[i, j] -= h
            fd = (go.head_backward(op, wi, c)[0][i] - go.head_backward(om, wi, c)[0][i]) / (2 * h)
            assert g[i, j] == pytest.approx(fd, rel=1e-5, abs=1e-7), (i, j)
    assert g[0, 4] == 0.0
    assert np.all(g[:, 40:] == 0.0)


def _random_params(lay, rng, grid_scale=0.3):
    p = np.zeros(lay.n_grid + go.MLP_SIZE, np.float32)
    p[: lay.n_grid] = rng.uniform(-grid_scale, grid_scale, lay.n_grid)
    lim = math.sqrt(6 / 64)
    p[lay.n_grid:lay.n_grid + go.W1_SIZE + go.W2_SIZE] = rng.uniform(-lim, lim, go.W1_SIZE + go.W2_SIZE)
    p[lay.n_grid + go.W1_SIZE + go.W2_SIZE:] = rng.uniform(-lim, lim, go.W3_SIZE) * 0.5
    return p


def _random_queries(rng, n, lo, hi):
    x = rng.uniform(lo, hi, (n, 3))
    nrm = rng.standard_normal((n, 3))
    nrm /= np.linalg.norm(nrm, axis=1, keepdims=True)
    wo = rng.standard_normal((n, 3))
    wo /= np.linalg.norm(wo, axis=1, keepdims=True)
    return np.concatenate([x, nrm, wo], axis=1).astype(np.float32)


def test_network_gradient_finite_differences():
    lay = go.Layout(log2_t=10, base=2.0, max_res=64.0, bound_lo=(0, 0, 0), bound_hi=(2, 1, 1))
    rng = np.random.default_rng(9)
    p = _random_params(lay, rng).astype(np.float64)
    q = _random_queries(rng, 48, [0, 0, 0], [2, 1, 1])
    wi = rng.standard_normal((48, 3))
    c = rng.uniform(0.1, 1.0, 48)
    c[3] = 0.0  # inactive sample
    _, g = go.gradient(p, lay, q, wi, c)
    touched = np.flatnonzero(g[: lay.n_grid])
    picks = np.concatenate([rng.choice(touched, 30, replace=False),
                            lay.n_grid + rng.choice(go.MLP_SIZE, 60, replace=False)])
    h = 1e-5
    for i in picks:
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        fd = (go.loss(pp, lay, q, wi, c) - go.loss(pm, lay, q, wi, c)) / (2 * h)
        assert g[i] == pytest.approx(fd, rel=2e-4, abs=2e-6), i


def test_quantised_gradient_close_to_exact():
    """BF16 operands change the gradient by about the BF16 resolution, not more."""
    lay = go.Layout(log2_t=12, base=4.0, max_res=256.0)
    rng = np.random.default_rng(10)
    p = _random_params(lay, rng)
    q = _random_queries(rng, 256, [0, 0, 0], [1, 1, 1])
    wi = rng.standard_normal((256, 3))
    c = rng.uniform(0.1, 1.0, 256)
    l0, g0 = go.gradient(p, lay, q, wi, c, quantised=False)
    l1, g1 = go.gradient(p, lay, q, wi, c, quantised=True)
    assert l1 == pytest.approx(l0, rel=2e-2)
    o = lay.n_grid
    for a, b in [(0, o), (o, o + go.W1_SIZE), (o + go.W1_SIZE, o + go.W1_SIZE + go.W2_SIZE),
                 (o + go.W1_SIZE + go.W2_SIZE, o + go.MLP_SIZE)]:
        err = np.linalg.norm(g1[a:b] - g0[a:b]) / np.linalg.norm(g0[a:b])
        assert err < 0.05, (a, err)


def test_forward_bound_covers_operand_noise():
    """The interval bound of the quantised forward covers the outputs obtained when the FP32 sums are
    perturbed inside their error radius (a stand-in for a different summation order)."""
    lay = go.Layout(log2_t=12, base=4.0, max_res=256.0)
    rng = np.random.default_rng(11)
    p = _random_params(lay, rng)
    net = go.Net(p, lay, quantised=True)
    q = _random_queries(rng, 512, [0, 0, 0], [1, 1, 1])
    o, _, bound = go.forward(net, q, with_bound=True)
    assert np.all(bound >= 0) and np.median(bound / (np.abs(o) + 1e-3)) < 1e-2
    # reference forward with FP32 arithmetic in a different order (torch float32, reversed K)
    x, _ = go.encode(net, q)
    h = torch.from_numpy(go.bf16(x.astype(np.float32))).float()
    for layer, w in enumerate(net.w):
        wt = torch.from_numpy(w).float()
        a = (h.flip(1) @ wt.flip(0)).double().numpy()
        if layer < 2:
            h = torch.from_numpy(go.bf16(np.maximum(a, 0.0))).float()
    assert np.all(np.abs(a - o) <= bound + 1e-12)


def test_adam_matches_torch():
    rng = np.random.default_rng(12)
    p0 = rng.standard_normal(1000).astype(np.float32)
    grads = [rng.standard_normal(1000).astype(np.float32) * 10.0 ** rng.uniform(-6, 1) for _ in range(5)]
    t = torch.nn.Parameter(torch.from_numpy(p0.astype(np.float64)))
    # adam.slang holds beta1/beta2 as float: torch gets the same FP32 values, whose rounding matters in
    # 1 - beta2^t at small t
    f = lambda v: float(np.float32(v))  # noqa: E731
    opt = torch.optim.Adam([t], lr=f(1e-2), betas=(f(0.9), f(0.999)), eps=f(1e-8))
    p, m1, m2 = p0.copy(), np.zeros(1000, np.float32), np.zeros(1000, np.float32)
    for it, g in enumerate(grads, start=1):
        t.grad = torch.from_numpy(g.astype(np.float64))
        opt.step()
        p, m1, m2 = go.adam_step(p, g, m1, m2, it)
        # FP32 parameters (|p| < 4) round by up to one ulp = 2^-22 per step against torch's float64
        assert np.abs(p - t.detach().numpy()).max() <= it * 2.0**-22 * 1.5


def test_adam_first_step_closed_form():
    """With bias correction the first step is -lr * g / (|g| + eps)."""
    g = np.array([3.0, -1e-3, 0.0, 2e-9], np.float32)
    p, _, _ = go.adam_step(np.zeros(4, np.float32), g, np.zeros(4, np.float32), np.zeros(4, np.float32), 1, lr=0.1)
    # v / (1 - beta2) divides two FP32 values carrying the cancellation of 1 - beta2 (relative 2^-24 /
    # 1e-3 each): the closed form holds to about 1e-4
    assert np.allclose(p, -0.1 * g / (np.abs(g) + 1e-8), rtol=1e-4)


def test_fp8_matches_torch_e4m3fn():
    """The oracle's e4m3 rounding against torch.float8_e4m3fn (OCP e4m3, nearest even) inside the
    saturation range, ties and subnormals included, and saturation at 448 beyond it."""
    rng = np.random.default_rng(5)
    x = np.concatenate([rng.standard_normal(40000) * np.exp(rng.uniform(-14, 7, 40000)),
                        [0.0, -0.0, 448.0, -448.0, 2.0**-9, 2.0**-10, 3 * 2.0**-11, 1.0625, 1.1875, 2.0**-6,
                         2.0**-6 - 2.0**-10, 17.0, 464.0 - 1e-3]]).astype(np.float32)
    inside = np.abs(x) <= 448
    ref = torch.from_numpy(x[inside]).to(torch.float8_e4m3fn).to(torch.float64).numpy()
    assert np.array_equal(go.fp8_e4m3(x[inside]), ref)
    assert np.array_equal(go.fp8_e4m3(np.array([500.0, -1e9, 449.0])), [448.0, -448.0, 448.0])


def test_fp8_pow2_scale_closed_forms():
    assert go.fp8_pow2_scale(448.0) == 1.0
    assert go.fp8_pow2_scale(448.0 * 2**-5) == 2**-5
    assert go.fp8_pow2_scale(449.0) == 2.0
    assert go.fp8_pow2_scale(1.0) == 2.0**-8  # 1 / 2^-8 = 256 <= 448 < 512
    assert go.fp8_pow2_scale(0.0) == 1.0
    rng = np.random.default_rng(6)
    for v in np.exp(rng.uniform(-20, 20, 200)).astype(np.float32):
        s = go.fp8_pow2_scale(v)
        assert v / s <= 448.0 * (1 + 2**-23) and v / (s / 2) > 448.0 * (1 - 2**-23)


def test_fp8_forward_bound_covers_reordered_sums_and_is_close_to_bf16():
    """The FP8 forward's interval covers a torch float32 forward of the same FP8 operands summed in
    reversed order; FP8 outputs stay within a few percent of the FP32 network; a forward that forgets
    the input scale falls outside the bound."""
    lay = go.Layout(log2_t=12, base=4.0, max_res=256.0)
    rng = np.random.default_rng(12)
    p = _random_params(lay, rng)
    net8 = go.Fp8Net(p, lay)
    q = _random_queries(rng, 512, [0, 0, 0], [1, 1, 1])
    exact = go.Net(p, lay, quantised=False)
    o32, acts = go.forward(exact, q)
    s_in = go.fp8_input_scales([np.abs(a).max() for a in acts])
    o, bound = go.forward_fp8(net8, q, s_in)
    assert np.all(bound >= 0)
    x, _ = go.encode(net8, q)
    h = torch.from_numpy(go.fp8_e4m3(x.astype(np.float32) / s_in[0])).float()
    for layer, w in enumerate(net8.w):
        a = (h.flip(1) @ torch.from_numpy(w).float().flip(0)).double().numpy() * (s_in[layer] * net8.sw[layer])
        if layer < 2:
            h = torch.from_numpy(go.fp8_e4m3(np.maximum(a, 0.0) / s_in[layer + 1])).float()
    assert np.all(np.abs(a - o) <= bound + 1e-12)
    rel = np.linalg.norm(o - o32) / np.linalg.norm(o32)
    assert rel < 0.1
    # an FP8 conversion that truncates instead of rounding to nearest even leaves the bound
    rnd = go.fp8_e4m3

    def truncating(a):
        v = np.clip(np.asarray(a, np.float64), -448.0, 448.0)
        m = np.abs(v)
        _, ex = np.frexp(np.where(m > 0, m, 1.0))
        qq = np.ldexp(1.0, np.maximum(ex - 1, -6) - 3)
        return np.sign(v) * np.floor(m / qq) * qq

    try:
        go.fp8_e4m3 = truncating
        wrong = go.forward_fp8(net8, q, s_in)[0]
    finally:
        go.fp8_e4m3 = rnd
    assert np.mean(np.abs(wrong - o) > bound) > 0.5


def _mfma_fp8_group_emulation(x, w, layer, rng):
    """The measured FP8 MFMA accumulation: products of one lane group aligned to the group's largest
    (LSB 2^(E - 14)) and truncated toward zero or down at random, group sums added in float32."""
    acc = np.zeros((x.shape[0], w.shape[1]), np.float32)
    for grp in go.fp8_groups(layer):
        p = x[:, grp][:, :, None] * w[grp, :][None, :, :]
        m = np.abs(p).max(axis=1, keepdims=True)
        _, ex = np.frexp(np.where(m > 0, m, 1.0))
        lsb = np.ldexp(1.0, ex - 1 - 14)
        t = np.where(rng.uniform(size=p.shape) < 0.5, np.trunc(p / lsb), np.floor(p / lsb)) * lsb
        t = np.where(np.abs(p) == m, p, t)  # the largest product is exact
        acc = (acc + t.sum(axis=1).astype(np.float32)).astype(np.float32)
    return acc.astype(np.float64)


def test_fp8_bound_covers_group_aligned_accumulation():
    """The FP8 bound covers a forward whose dot products follow the measured MFMA model (group-aligned
    truncation, float32 group sums) with random truncation direction."""
    lay = go.Layout(log2_t=12, base=4.0, max_res=256.0)
    rng = np.random.default_rng(13)
    p = _random_params(lay, rng)
    net8 = go.Fp8Net(p, lay)
    q = _random_queries(rng, 400, [0, 0, 0], [1, 1, 1])
    _, acts = go.forward(go.Net(p, lay, quantised=False), q)
    s_in = go.fp8_input_scales([np.abs(a).max() for a in acts])
    o, bound = go.forward_fp8(net8, q, s_in)
    x, _ = go.encode(net8, q)
    h = go.fp8_e4m3(x.astype(np.float32) / s_in[0])
    for layer, w in enumerate(net8.w):
        a = _mfma_fp8_group_emulation(h, w, layer, rng) * (s_in[layer] * net8.sw[layer])
        if layer < 2:
            h = go.fp8_e4m3(np.maximum(a, 0.0) / s_in[layer + 1])
    assert np.all(np.abs(a - o) <= bound + 1e-12)
