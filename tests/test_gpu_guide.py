"""GPU parity of the path guide against the CPU oracle (oracle/ks_guide_oracle.py).

Forward: every mixture parameter lies inside the oracle's bound, an interval carried through the FP32
sums and each BF16 rounding (ulp flips included). It is tight enough that a transposed operand or a
wrong lane layout of the MFMA tiles fails: the test also requires the median bound to be small.
Gradient: the relative L2 error per parameter block is within 5e-3 of the oracle's emulation of the same
BF16 operands; rounding flips (rate ~2^-10, size 2^-8) predict ~1e-4. Adam: the first step of
adam.slang with bias correction and a second step from the oracle's moments."""

import math

import numpy as np
import pytest

import ks_guide_oracle as go

pytestmark = pytest.mark.gpu


def _queries(rng, n, lo, hi):
    x = rng.uniform(lo, hi, (n, 3))
    nrm = rng.standard_normal((n, 3))
    nrm /= np.linalg.norm(nrm, axis=1, keepdims=True)
    wo = rng.standard_normal((n, 3))
    wo /= np.linalg.norm(wo, axis=1, keepdims=True)
    return np.concatenate([x, nrm, wo], axis=1).astype(np.float32)


def _params(n_grid, rng, grid_scale=0.3):
    p = np.zeros(n_grid + go.MLP_SIZE, np.float32)
    p[:n_grid] = rng.uniform(-grid_scale, grid_scale, n_grid)
    lim = math.sqrt(6 / 64)
    p[n_grid:n_grid + go.W1_SIZE + go.W2_SIZE] = rng.uniform(-lim, lim, go.W1_SIZE + go.W2_SIZE)
    p[n_grid + go.W1_SIZE + go.W2_SIZE:] = rng.uniform(-lim, lim, go.W3_SIZE) * 0.5
    return p


CONFIGS = {
    # dense coarse levels ((res+1)^3 <= T) and hashed fine ones
    "small": dict(log2_table_size=12, base_resolution=4.0, max_resolution=256.0),
    "default": dict(),
    "fine": dict(max_resolution=2048.0),
}
BOUNDS = ((-1.0, 0.0, 2.0), (3.0, 1.5, 5.0))


@pytest.fixture(params=list(CONFIGS))
def guide(request, ks, device):
    cfg = ks.guide_config(seed=7, **CONFIGS[request.param])
    g = ks.Guide(device, BOUNDS[0], BOUNDS[1], cfg)
    c = CONFIGS[request.param]
    lay = go.Layout(c.get("log2_table_size", 19), c.get("base_resolution", 16.0), c.get("max_resolution", 256.0),
                    BOUNDS[0], BOUNDS[1])
    yield g, lay
    g.close()


def test_layout_matches_oracle(guide):
    g, lay = guide
    res, lo, inv = g.grid_info()
    assert res.tolist() == lay.res
    assert np.array_equal(lo, lay.lo) and np.array_equal(inv, lay.inv_ext)
    assert g.n_grid == lay.n_grid and g.n_mlp == go.MLP_SIZE


def test_params_round_trip_and_init(guide):
    g, lay = guide
    p = g.get_params()
    assert np.all(np.isfinite(p))
    assert np.abs(p[: g.n_grid]).max() <= 1e-4 and np.abs(p[: g.n_grid]).max() > 0
    assert np.abs(p[g.n_grid:]).max() <= math.sqrt(6 / 64)
    q = _params(g.n_grid, np.random.default_rng(1))
    g.set_params(q)
    assert np.array_equal(g.get_params(), q)


def test_forward_within_oracle_bound(guide):
    g, lay = guide
    rng = np.random.default_rng(2)
    p = _params(g.n_grid, rng)
    g.set_params(p)
    n = 1000  # not a multiple of the 128-query block
    q = _queries(rng, n, BOUNDS[0], BOUNDS[1])
    q[:5, 0:3] = [[-5, -5, -5], [9, 9, 9], BOUNDS[0], BOUNDS[1], [1.0, 0.75, 3.5]]  # clamped and edge positions
    mix = g.forward(q)
    net = go.Net(p, lay, quantised=True)
    o, _, bound = go.forward(net, q, with_bound=True)
    w, mu, kappa = go.head(o)
    dw, dmu, dk = go.head_bound(o, bound)
    gw = mix[:, :8]
    lob = mix[:, 8:].reshape(n, 8, 4)
    print("median output bound", float(np.median(bound)), "median |o|", float(np.median(np.abs(o))), flush=True)
    assert np.median(bound) < 1e-3 * max(1.0, float(np.median(np.abs(o))))
    assert np.all(np.abs(gw - w) <= dw)
    assert np.all(np.linalg.norm(lob[:, :, :3] - mu, axis=2) <= dmu)
    assert np.all(np.abs(lob[:, :, 3] - kappa) <= dk)
    assert np.allclose(gw.sum(axis=1), 1.0, atol=1e-5)


def test_fp8_forward_within_oracle_bound(guide):
    """FP8 (e4m3) inference: the weight scales follow the power-of-two rule on the FP32 master weights,
    the input scales follow it on the calibration maxima, and every mixture parameter lies inside the
    oracle's FP8 interval bound (FP8 rounding of both interval ends; the MFMA's group-aligned
    accumulation, which keeps 15 bits below each lane group's largest product)."""
    g, lay = guide
    rng = np.random.default_rng(21)
    p = _params(g.n_grid, rng)
    g.set_params(p)
    cal = _queries(rng, 3000, BOUNDS[0], BOUNDS[1])
    g.calibrate_fp8(cal)
    g.set_inference(True)
    sc = g.fp8_scales()
    net8 = go.Fp8Net(p, lay)
    assert [sc["sw1"], sc["sw2"], sc["sw3"]] == net8.sw
    _, acts = go.forward(go.Net(p, lay, quantised=True), cal)
    amax = [float(np.abs(a).max()) for a in acts]
    want = go.fp8_input_scales(amax)
    got = [sc["sx"], sc["sh1"], sc["sh2"]]
    print("fp8 input scales", got, "oracle", want, "amax", amax, "weight scales", net8.sw, flush=True)
    for a, w_, g_ in zip(amax, want, got):
        # an activation maximum within a BF16 step of a scale boundary may land on either side
        boundary = abs(math.log2(a / 448.0) - round(math.log2(a / 448.0))) < 2**-7
        assert g_ == w_ or (boundary and g_ in (w_ / 2, w_ * 2))
    assert sc["inv_sx"] * sc["sx"] == 1.0 and sc["c1"] == sc["sx"] * sc["sw1"] and sc["c3"] == sc["sh2"] * sc["sw3"]
    n = 1000
    q = _queries(rng, n, BOUNDS[0], BOUNDS[1])
    q[:5, 0:3] = [[-5, -5, -5], [9, 9, 9], BOUNDS[0], BOUNDS[1], [1.0, 0.75, 3.5]]
    mix = g.forward(q)
    o, bound = go.forward_fp8(net8, q, got)
    w, mu, kappa = go.head(o)
    dw, dmu, dk = go.head_bound(o, bound)
    gw = mix[:, :8]
    lob = mix[:, 8:].reshape(n, 8, 4)
    print("fp8 median output bound", float(np.median(bound)), "median |o|", float(np.median(np.abs(o))), flush=True)
    assert np.median(bound) < 0.05 * float(np.median(np.abs(o)))
    assert np.all(np.abs(gw - w) <= dw)
    assert np.all(np.linalg.norm(lob[:, :, :3] - mu, axis=2) <= dmu)
    assert np.all(np.abs(lob[:, :, 3] - kappa) <= dk)
    assert np.allclose(gw.sum(axis=1), 1.0, atol=1e-5)
    # against the BF16 network: the FP8 mixture weights stay close
    g.set_inference(False)
    mix_bf = g.forward(q)
    err = np.abs(mix_bf[:, :8] - gw).max()
    print("max |w_fp8 - w_bf16|", float(err), flush=True)
    assert err < 0.1


def test_gradient_matches_oracle(guide):
    g, lay = guide
    rng = np.random.default_rng(3)
    p = _params(g.n_grid, rng)
    g.set_params(p)
    n = 700
    q = _queries(rng, n, BOUNDS[0], BOUNDS[1])
    wi = rng.standard_normal((n, 3)).astype(np.float32)
    c = rng.uniform(0.05, 2.0, n).astype(np.float32)
    c[::17] = 0.0  # inactive samples contribute nothing
    loss, grad = g.gradient(q, wi, c)
    l_ref, g_ref = go.gradient(p, lay, q, wi, c, quantised=True)
    assert loss == pytest.approx(l_ref, rel=1e-4)
    o = g.n_grid
    blocks = {"grid": (0, o), "W1": (o, o + go.W1_SIZE), "W2": (o + go.W1_SIZE, o + go.W1_SIZE + go.W2_SIZE),
              "W3": (o + go.W1_SIZE + go.W2_SIZE, o + go.MLP_SIZE)}
    for name, (a, b) in blocks.items():
        err = np.linalg.norm(grad[a:b] - g_ref[a:b]) / np.linalg.norm(g_ref[a:b])
        print(name, "relative L2 error", err, flush=True)
        assert err < 5e-3, name
    # the grid gradient touches exactly the entries the oracle touches
    assert np.array_equal(grad[:o] != 0, g_ref[:o] != 0)
    # a second call returns the same gradient: the accumulators were cleared
    loss2, grad2 = g.gradient(q, wi, c)
    assert loss2 == pytest.approx(loss, rel=1e-6)
    assert np.linalg.norm(grad2 - grad) <= 1e-5 * np.linalg.norm(grad)


def test_adam_steps(ks, device):
    cfg = ks.guide_config(seed=3, learning_rate=5e-3, **CONFIGS["small"])
    g = ks.Guide(device, BOUNDS[0], BOUNDS[1], cfg)
    try:
        rng = np.random.default_rng(4)
        p0 = _params(g.n_grid, rng)
        g.set_params(p0)
        q = _queries(rng, 300, BOUNDS[0], BOUNDS[1])
        wi = rng.standard_normal((300, 3)).astype(np.float32)
        c = np.ones(300, np.float32)
        _, g1 = g.gradient(q, wi, c)
        loss1 = g.train_step(q, wi, c)
        p1 = g.get_params()
        z = np.zeros_like(p0)
        want1, m1, v1 = go.adam_step(p0, g1, z, z, 1, lr=5e-3)
        tol = 4 * np.spacing(np.abs(p0)) + 1e-4 * 5e-3  # parameter rounding + the 1 - beta2 cancellation
        assert np.all(np.abs(p1 - want1) <= tol)
        _, g2 = g.gradient(q, wi, c)
        loss2 = g.train_step(q, wi, c)
        want2, _, _ = go.adam_step(p1, g2, m1, v1, 2, lr=5e-3)
        assert np.all(np.abs(g.get_params() - want2) <= 4 * np.spacing(np.abs(p1)) + 1e-4 * 5e-3)
        assert loss2 < loss1
    finally:
        g.close()


def test_learns_a_position_dependent_lobe(ks, device):
    """Samples wi ~ vMF(mu(x), 20) with mu depending on the octant of x: the mean -log q on held-out
    samples falls from log(4 pi) = 2.53 towards the entropy of that vMF (-0.16)."""
    cfg = ks.guide_config(seed=11, learning_rate=1e-2)
    g = ks.Guide(device, (0, 0, 0), (1, 1, 1), cfg)
    try:
        rng = np.random.default_rng(5)
        dirs = rng.standard_normal((8, 3))
        dirs /= np.linalg.norm(dirs, axis=1, keepdims=True)

        def batch(n):
            q = _queries(rng, n, (0, 0, 0), (1, 1, 1))
            octant = (q[:, 0] > 0.5) * 1 + (q[:, 1] > 0.5) * 2 + (q[:, 2] > 0.5) * 4
            wi = go.vmf_sample(dirs[octant], np.full(n, 20.0), rng.uniform(size=n), rng.uniform(size=n))
            return q, wi.astype(np.float32)

        qt, wt = batch(4096)
        ones = np.ones(4096, np.float32)
        nll0 = g.gradient(qt, wt, ones)[0] / 4096
        for _ in range(300):
            q, wi = batch(4096)
            g.train_step(q, wi, ones)
        nll = g.gradient(qt, wt, ones)[0] / 4096
        print("held-out NLL", nll0, "->", nll, flush=True)
        assert nll0 == pytest.approx(math.log(4 * math.pi), abs=0.3)
        assert nll < 0.3
    finally:
        g.close()


def _mixtures(rng, n, kmax, w_min=0.0):
    w = rng.uniform(w_min, 1.0, (n, 8))
    w[rng.uniform(size=(n, 8)) < 0.2] = 0.0
    w[:, 0] += 1e-3
    w /= w.sum(axis=1, keepdims=True)
    mu = rng.standard_normal((n, 8, 3))
    mu /= np.linalg.norm(mu, axis=2, keepdims=True)
    kappa = np.exp(rng.uniform(np.log(1e-2), np.log(kmax), (n, 8)))
    kappa[rng.uniform(size=(n, 8)) < 0.05] = 0.0
    return go.mixture_array(w, mu, kappa).astype(np.float32)


def test_vmm_pdf(ks, device):
    rng = np.random.default_rng(6)
    n = 5000
    mix = _mixtures(rng, n, 1e4)
    d = rng.standard_normal((n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    lob = mix[:, 8:].reshape(n, 8, 4)
    d[:200] = lob[:200, 0, :3]  # at the mean of a lobe
    d = d.astype(np.float32)
    p = ks.vmm_pdf(device, mix, d)
    want = go.mixture_pdf(mix, d)
    # FP32 evaluation: log-sum-exp terms of size up to 2 kappa + 10 and a dot product of ~3 ulp scaled by
    # kappa; 16 ulp of kappa_max bounds both
    tol = 1e-5 + 16 * 2.0**-24 * lob[:, :, 3].max(axis=1)
    rel = np.abs(p - want) / np.maximum(want, 1e-30)
    print("max relative error / tolerance", float((rel / tol)[want > 1e-30].max()), flush=True)
    assert np.all((rel <= tol) | (want < 1e-30))


def test_vmm_sample(ks, device):
    """Directions agree with sample_vmm for the same (u0, u1) up to the FP32 error of the selection remap
    (carried through the oracle's derivative of the direction in u0) and of the lobe mapping."""
    rng = np.random.default_rng(7)
    n = 5000
    mix = _mixtures(rng, n, 1e3, w_min=0.02)
    u = rng.uniform(size=(n, 2)).astype(np.float32)
    d, p = ks.vmm_sample(device, mix, u)
    want, sel = go.mixture_sample(mix, u)
    w = mix[:, :8].astype(np.float64)
    wn = w / w.sum(axis=1, keepdims=True)
    cdf = np.cumsum(wn, axis=1)
    # a u0 within FP32 rounding of a CDF step may select the neighbouring lobe: skipped
    clear = np.abs(cdf - u[:, :1]).min(axis=1) > 1e-5
    assert clear.mean() > 0.99
    rows = np.arange(n)
    lob = mix[:, 8:].reshape(n, 8, 4)[rows, sel]
    _, ur = go.mixture_select(w, u[:, 0])
    # remapped u0 = (u0 - cdf) / w_sel: FP32 cdf sums of 8 terms carry 16 ulp, divided by w_sel
    dur = 16 * 2.0**-24 / wn[rows, sel]
    h = 1e-7
    up = go.vmf_sample(lob[:, :3], lob[:, 3], np.clip(ur + h, 0, 1), u[:, 1])
    dn = go.vmf_sample(lob[:, :3], lob[:, 3], np.clip(ur - h, 0, 1), u[:, 1])
    jac = np.linalg.norm(up - dn, axis=1) / (2 * h)
    ct = np.einsum("nc,nc->n", want, lob[:, :3])
    st = np.sqrt(np.maximum(1e-14, 1 - ct**2))
    # lobe mapping in FP32: cos(theta) to ~2^-23 (1 + |log(.)| / kappa), sin(theta) via 1 - ct^2, basis and
    # trig to a few ulp
    k = np.maximum(lob[:, 3], 1e-2)
    dct = 2.0**-23 * (1 + np.abs(np.log(np.maximum(ur, 1e-30))) / k)
    tol = 2 * (jac * dur + dct / st + 2.0**-22 / st + 8 * 2.0**-24)
    err = np.linalg.norm(d - want, axis=1)
    print("max error / tolerance", float((err / tol)[clear].max()), flush=True)
    assert np.all(err[clear] <= tol[clear])
    assert np.allclose(np.linalg.norm(d, axis=1), 1.0, atol=1e-5)
    assert np.allclose(p, ks.vmm_pdf(device, mix, d), rtol=1e-6)


def test_invalid_config_rejected(ks, device):
    with pytest.raises(ks.KsError, match="log2_table_size"):
        ks.Guide(device, (0, 0, 0), (1, 1, 1), ks.guide_config(log2_table_size=30))
    with pytest.raises(ks.KsError, match="guide_fraction"):
        ks.Guide(device, (0, 0, 0), (1, 1, 1), ks.guide_config(guide_fraction=1.5))
    with pytest.raises(ks.KsError, match="bound"):
        ks.Guide(device, (0, 0, 0), (1, -1, 1))
