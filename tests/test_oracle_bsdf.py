"""Pins of the oracle's principled BSDF (principled_bsdf.cpp:246-705), its GGX microfacet model
(microfacet.h:20-153) and Fresnel terms (fresnel.h:9-37), CPU only.

Each check is independent of the oracle's own code: FP64 closed forms written in a different algebraic
form (Walter et al. 2007 / Heitz 2014 angle forms of D and Lambda, pbrt's two-index Fresnel), integrals
by quadrature, scipy's Beta function, and chi^2 tests of the samplers against the densities they claim.
A dropped term, a swapped operand or a wrong sign in the oracle fails at least one of them.
"""

import numpy as np
import pytest
from scipy import special, stats

import ks_oracle as ko
from keepsake_amd.api import MATERIAL_DTYPE


def _dirs(theta, phi):
    st = np.sin(theta)
    return np.stack([st * np.cos(phi), st * np.sin(phi), np.cos(theta)], -1)


def _unit(rng, n, upper=None):
    v = rng.normal(size=(n, 3))
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    if upper is True:
        v[:, 2] = np.abs(v[:, 2])
    elif upper is False:
        v[:, 2] = -np.abs(v[:, 2])
    return v


def ggx_D_walter(m, ax, ay):
    """GGX normal distribution in angle form (Heitz 2014, eq. 85): D = 1 / (pi ax ay cos^4 t (1 + tan^2 t
    (cos^2 p / ax^2 + sin^2 p / ay^2))^2), zero on the lower hemisphere."""
    m = np.asarray(m, np.float64)
    c2 = m[:, 2] ** 2
    s2 = np.maximum(1.0 - c2, 0.0)
    rxy = np.maximum(m[:, 0] ** 2 + m[:, 1] ** 2, 1e-300)
    cp2, sp2 = m[:, 0] ** 2 / rxy, m[:, 1] ** 2 / rxy
    tan2 = s2 / np.maximum(c2, 1e-300)
    with np.errstate(over="ignore", invalid="ignore"):  # grazing m: D -> 0
        d = 1.0 / (np.pi * ax * ay * c2 * c2 * (1.0 + tan2 * (cp2 / ax**2 + sp2 / ay**2)) ** 2)
    return np.where((m[:, 2] > 0) & np.isfinite(d), d, 0.0)


def ggx_lambda_closed(w, ax, ay):
    """Smith Lambda of GGX (Heitz 2014, eq. 86) with the projected roughness of direction w."""
    w = np.asarray(w, np.float64)
    c2 = w[:, 2] ** 2
    s2 = np.maximum(1.0 - c2, 0.0)
    rxy = np.maximum(w[:, 0] ** 2 + w[:, 1] ** 2, 1e-300)
    a2 = (w[:, 0] ** 2 * ax**2 + w[:, 1] ** 2 * ay**2) / rxy
    return 0.5 * (-1.0 + np.sqrt(1.0 + a2 * s2 / c2))


def fresnel_pbrt(cos_i, eta_i, eta_t):
    """Unpolarised dielectric Fresnel reflectance with two indices (pbrt-v3 FrDielectric), FP64."""
    cos_i = np.asarray(cos_i, np.float64)
    sin_t = eta_i / eta_t * np.sqrt(np.maximum(0.0, 1.0 - cos_i**2))
    tir = sin_t >= 1.0
    cos_t = np.sqrt(np.maximum(0.0, 1.0 - sin_t**2))
    r_par = (eta_t * cos_i - eta_i * cos_t) / (eta_t * cos_i + eta_i * cos_t)
    r_perp = (eta_i * cos_i - eta_t * cos_t) / (eta_i * cos_i + eta_t * cos_t)
    return np.where(tir, 1.0, 0.5 * (r_par**2 + r_perp**2))


# ------------------------------------------------------------------------------------------------
# GGX
# ------------------------------------------------------------------------------------------------
@pytest.mark.parametrize("alpha", [0.01, 0.05, 0.2, 0.5, 1.0])
def test_ggx_projected_area_is_one(alpha):
    """int D(m) cos(theta_m) dm = 1 (the microfacets' projected area is the macro-surface's). With w =
    cos^2(theta_m) the integral is pi * int_0^1 D dw; the grid is graded towards the peak at w = 1."""
    t = np.concatenate([[0.0], np.geomspace(1e-9, 1.0, 40001)])
    z = np.sqrt(1.0 - t).astype(np.float32)  # the oracle sees float32 directions
    z = np.unique(z)
    m = np.stack([np.sqrt(np.maximum(0.0, 1.0 - z.astype(np.float64) ** 2)), np.zeros_like(z, np.float64),
                  z.astype(np.float64)], -1).astype(np.float32)
    d = ko.ggx("D", alpha, alpha, m).astype(np.float64)
    w = z.astype(np.float64) ** 2
    integral = np.pi * np.trapezoid(d, w)
    assert abs(integral - 1.0) < 2e-3, integral


@pytest.mark.parametrize("ax,ay", [(0.3, 0.3), (0.05, 0.05), (0.2, 0.5), (0.6, 0.15), (1e-4, 1e-4)])
def test_ggx_D_matches_angle_form(ax, ay):
    rng = np.random.default_rng(1)
    m = _unit(rng, 4000, upper=True)
    m[:10] = [[0, 0, 1]] * 10
    got = ko.ggx("D", ax, ay, m.astype(np.float32)).astype(np.float64)
    # the oracle's float alphas, clamped to min_alpha (microfacet.h:152)
    axc, ayc = (float(np.clip(np.float32(v), np.float32(1e-3), np.float32(1.0))) for v in (ax, ay))
    want = ggx_D_walter(m.astype(np.float32), axc, ayc)
    # float conditioning of D: the ulp of z (|d ln D / d z| ulp(z)) and the rounding of (a^2 - 1) next to 1
    # inside t = 1 + (a^2 - 1) z^2, which is relative to t itself and dominates at the peak of small alphas
    z = m[:, 2].astype(np.float64)
    a2 = axc * ayc
    t = 1.0 + (a2 - 1.0) * z * z
    cond = 8 * 2.0**-24 * (1.0 + 4.0 * z * z * (1.0 - a2)) / t
    assert np.all(ko.ggx("D", ax, ay, -m.astype(np.float32)) == 0.0)
    peak = ko.ggx("D", ax, ax, np.array([[0, 0, 1.0]], np.float32))[0] * np.pi * axc**2
    assert abs(peak - 1.0) < 4e-6 + 8 * 2.0**-24 / axc**2


@pytest.mark.parametrize("ax,ay", [(0.1, 0.1), (0.5, 0.5), (1.0, 1.0), (0.2, 0.6)])
def test_ggx_lambda_g1_g2(ax, ay):
    rng = np.random.default_rng(2)
    wo = _unit(rng, 3000, upper=True).astype(np.float32)
    wi = _unit(rng, 3000).astype(np.float32)
    lo = ko.ggx("lambda", ax, ay, wo).astype(np.float64)
    want = ggx_lambda_closed(wo, ax, ay)
    assert np.allclose(lo, want, rtol=2e-5, atol=2e-6)
    g1o = ko.ggx("G1", ax, ay, wo).astype(np.float64)
    assert np.allclose(g1o, 1.0 / (1.0 + want), rtol=1e-5)
    g1i = ko.ggx("G1", ax, ay, wi).astype(np.float64)
    g2 = ko.ggx("G2", ax, ay, wo, wi).astype(np.float64)
    same = wo[:, 2] * wi[:, 2] >= 0
    # height-correlated G2 = 1 / (1 + L(o) + L(i)), i.e. 1/G2 = 1/G1(o) + 1/G1(i) - 1 (Heitz 2014, eq. 99)
    assert np.allclose(1.0 / g2[same], 1.0 / g1o[same] + 1.0 / g1i[same] - 1.0, rtol=2e-5)
    # transmission: G2 = B(1 + L(o), 1 + L(i)) (Heitz 2014, eq. 101)
    li = ggx_lambda_closed(wi, ax, ay)
    a, b = 1.0 + want[~same], 1.0 + li[~same]
    beta = special.beta(a, b)
    # std::beta on floats goes through float lgamma sums: an absolute error of a few ulp of the largest
    # lgamma term becomes a relative error of the result at grazing angles (large Lambda)
    tol = 5e-5 + 4 * 2.0**-24 * (np.abs(special.gammaln(a)) + np.abs(special.gammaln(b)) + np.abs(special.gammaln(a + b)))
    # (below float's smallest normal the float result flushes towards zero)
    assert np.all(np.abs(g2[~same] - beta) <= tol * beta + 1.2e-38), np.max(np.abs(g2[~same] - beta) / beta / tol)
    assert np.all(g2 <= np.minimum(g1o, g1i) + 1e-6)
    assert np.all(ko.ggx("lambda", ax, ay, np.array([[0, 0, 1.0]], np.float32)) == 0.0)


@pytest.mark.parametrize("alpha", [0.05, 0.3, 0.8])
@pytest.mark.parametrize("cos_o", [0.1, 0.4, 0.95])
def test_ggx_g1_projected_area_identity(alpha, cos_o):
    """int G1(wo) <wo, m>+ D(m) dm = cos(theta_o): G1 is the visible fraction of the projected
    microfacet area (Heitz 2014, eq. 16). Pins Lambda against D."""
    wo = np.array([[np.sqrt(1 - cos_o**2), 0.0, cos_o]], np.float32)
    g1 = float(ko.ggx("G1", alpha, alpha, wo)[0])
    # quadrature over the hemisphere: graded theta (peak of D at theta = 0), uniform phi
    u = np.concatenate([np.geomspace(1e-7, 1.0, 3000)])
    theta = np.arctan(alpha * np.sqrt(u / (1.0 - u + 1e-12)) * 4.0)
    theta = np.unique(np.clip(np.concatenate([[0.0], theta, np.linspace(0, np.pi / 2, 400)]), 0, np.pi / 2))
    phi = (np.arange(256) + 0.5) * (2 * np.pi / 256)
    T, P = np.meshgrid(theta, phi, indexing="ij")
    m = _dirs(T.ravel(), P.ravel())
    d = ko.ggx("D", alpha, alpha, m.astype(np.float32)).astype(np.float64).reshape(T.shape)
    proj = np.maximum(0.0, m @ wo[0].astype(np.float64)).reshape(T.shape)
    inner = (d * proj).mean(axis=1) * 2 * np.pi  # over phi
    integral = g1 * np.trapezoid(inner * np.sin(theta), theta)
    assert abs(integral - cos_o) < 3e-3 * max(cos_o, 0.1), (integral, cos_o)


@pytest.mark.parametrize("alpha", [0.1, 0.4, 0.9])
@pytest.mark.parametrize("cos_o", [0.2, 0.7, 1.0])
def test_ggx_vndf_sampling_chi2(alpha, cos_o):
    """Sampled microfacet normals follow D_wo(m) = G1(wo) <wo, m>+ D(m) / cos(theta_o) (microfacet.h:
    98-118), by chi^2 over (cos theta_m, phi_m) bins whose probabilities are integrated in FP64 from the
    closed forms."""
    rng = np.random.default_rng(3)
    n = 200_000
    so = np.sqrt(max(0.0, 1 - cos_o**2))
    wo = np.array([so, 0.0, cos_o], np.float32)
    m = ko.ggx_sample(alpha, alpha, wo, rng.uniform(0, 1, (n, 2)).astype(np.float32)).astype(np.float64)
    assert np.all(m @ wo.astype(np.float64) >= -1e-6), "VNDF produced a back-facing normal"
    # bins: equal-probability in cos theta under D alone (z^2 = 1 - t, t = a^2 s/(1 - s + a^2 s)), 8 in phi
    nz, nphi = 12, 8
    s_edges = np.linspace(0, 1, nz + 1)
    t_edges = alpha**2 * s_edges / np.maximum(1 - s_edges + alpha**2 * s_edges, 1e-300)
    z_edges = np.sqrt(np.clip(1 - t_edges, 0, 1))  # decreasing from 1 to 0
    phi_edges = np.linspace(-np.pi, np.pi, nphi + 1)
    zi = np.clip(np.searchsorted(-z_edges, -m[:, 2], side="right") - 1, 0, nz - 1)
    pi = np.clip(np.searchsorted(phi_edges, np.arctan2(m[:, 1], m[:, 0]), side="right") - 1, 0, nphi - 1)
    obs = np.bincount(zi * nphi + pi, minlength=nz * nphi).astype(np.float64)
    g1 = 1.0 / (1.0 + ggx_lambda_closed(wo[None].astype(np.float64), alpha, alpha)[0])
    exp = np.zeros(nz * nphi)
    k = 64
    for a in range(nz):
        zz = z_edges[a] + (z_edges[a + 1] - z_edges[a]) * (np.arange(k) + 0.5) / k  # midpoint in z
        dz = abs(z_edges[a + 1] - z_edges[a]) / k
        for b in range(nphi):
            pp = phi_edges[b] + (phi_edges[b + 1] - phi_edges[b]) * (np.arange(k) + 0.5) / k
            Z, PH = np.meshgrid(zz, pp, indexing="ij")
            st = np.sqrt(np.maximum(0, 1 - Z**2))
            mm = np.stack([st * np.cos(PH), st * np.sin(PH), Z], -1).reshape(-1, 3)
            dens = g1 * np.maximum(0.0, mm @ wo.astype(np.float64)) * ggx_D_walter(mm, alpha, alpha) / cos_o
            exp[a * nphi + b] = dens.sum() * dz * (phi_edges[1] - phi_edges[0]) / k
    assert abs(exp.sum() - 1.0) < 5e-3, exp.sum()
    exp = exp / exp.sum() * n
    keep = exp > 5
    chi2 = float((((obs - exp) ** 2)[keep] / exp[keep]).sum())
    p = stats.chi2.sf(chi2, keep.sum() - 1)
    assert p > 1e-4, (chi2, keep.sum(), p)
    assert obs[~keep].sum() <= max(10.0, 3 * exp[~keep].sum() + 10)


# ------------------------------------------------------------------------------------------------
# Fresnel, Schlick, Henyey-Greenstein
# ------------------------------------------------------------------------------------------------
@pytest.mark.parametrize("eta", [1 / 1.5, 1.5, 1 / 1.33, 1.33, 1 / 2.4, 2.4, 1.0])
def test_fresnel_dielectric(eta):
    """fresnel_dielectric(cos_i, eta) with eta = n_i / n_t (fresnel.h:9-22)."""
    # c = 0 with eta = 1 is 0/0 in the reference formula (never reached: eval and sample return early at
    # wo.z == 0); next to the critical cosine of total internal reflection the slope is infinite
    c = np.linspace(0.0, 1.0, 2001, dtype=np.float32)[1:]
    got = ko.fresnel_dielectric(c, eta).astype(np.float64)
    want = fresnel_pbrt(c, eta, 1.0)
    crit = np.sqrt(max(0.0, 1 - 1 / eta**2))
    away = np.abs(c - crit) > 2e-3
    if eta == 1.0:  # cos_t = sqrt(1 - (1 - c^2)) cancels at grazing: float error ~ ulp(1) / c^2
        away &= c > 0.02
    assert np.allclose(got[away], want[away], atol=3e-6, rtol=3e-6)
    # normal incidence ((eta - 1) / (eta + 1))^2
    assert abs(got[-1] - ((eta - 1) / (eta + 1)) ** 2) < 1e-6
    if eta == 1.0:
        assert np.all(got[c > 0.02] < 1e-6)
    if eta > 1.0:  # total internal reflection below the critical cosine
        crit = np.sqrt(1 - 1 / eta**2)
        assert np.all(got[c < crit - 1e-4] == 1.0) and np.all(got[c > crit + 1e-4] < 1.0)
    # Brewster's angle: the p-polarised part vanishes, R = rs^2 / 2
    tb = np.arctan(1.0 / eta)
    cb = np.float32(np.cos(tb))
    st = eta * np.sin(tb)
    if st < 1:
        ct = np.sqrt(1 - st**2)
        rs = (eta * cb - ct) / (eta * cb + ct)  # s-polarised amplitude, n_i cos_i vs n_t cos_t
        assert abs(ko.fresnel_dielectric(np.array([cb]), eta)[0] - 0.5 * rs**2) < 1e-5


def test_fresnel_schlick_and_hg():
    c = np.linspace(-0.5, 1.5, 4001, dtype=np.float32)
    got = ko.fresnel_schlick(c).astype(np.float64)
    want = np.clip(1.0 - c.astype(np.float64), 0.0, 1.0) ** 5
    assert np.allclose(got, want, atol=1e-6)
    # Henyey-Greenstein (bsdf.h:45-48): normalised over the sphere, mean cosine g, isotropic at g = 0
    mu = np.linspace(-1, 1, 200001)
    for g in (0.0, 0.3, 0.8, -0.5):
        p = ko.hg_phase(mu.astype(np.float32), g).astype(np.float64)
        assert abs(2 * np.pi * np.trapezoid(p, mu) - 1.0) < 1e-4
        assert abs(2 * np.pi * np.trapezoid(p * mu, mu) - g) < 1e-4
    assert np.allclose(ko.hg_phase(mu[::1000].astype(np.float32), 0.0), 1 / (4 * np.pi), rtol=1e-6)


# ------------------------------------------------------------------------------------------------
# Principled BSDF: lobes in FP64 closed form, reciprocity, sampling consistency
# ------------------------------------------------------------------------------------------------
def _materials():
    """The five materials of the device parity test, plus a white diffuse and a white metal."""
    m = np.zeros(7, MATERIAL_DTYPE)
    m["basecolor"] = [[0.73, 0.73, 0.73], [0.8, 0.3, 0.2], [0.9, 0.9, 0.9], [0.5, 0.6, 0.7], [0.2, 0.8, 0.4],
                      [1, 1, 1], [1, 1, 1]]
    m["roughness"] = [1.0, 0.3, 0.05, 0.6, 0.4, 0.5, 0.25]
    m["metallic"] = [0.0, 1.0, 0.0, 0.4, 0.0, 0.0, 1.0]
    m["ior"] = [1.0, 1.5, 1.5, 1.45, 1.33, 1.5, 1.5]
    m["specular_r0_mul"] = 0.5
    m["specular_trans"] = [0.0, 0.0, 0.0, 0.3, 0.8, 0.0, 0.0]
    m["diffuse_trans"] = [0.0, 0.0, 0.0, 0.2, 0.0, 0.0, 0.0]
    m["diffuse_trans_fwd"] = [0.0, 0.0, 0.0, 0.5, 0.0, 0.0, 0.0]
    return m


def _closure(mat):
    """eval_closure's clamps (principled_bsdf.cpp:246-275)."""
    c = dict(base=np.clip(np.asarray(mat["basecolor"], np.float64), 0, 1),
             r=float(np.clip(mat["roughness"], 0, 1)), metal=float(np.clip(mat["metallic"], 0, 1)),
             ior=float(np.clip(mat["ior"], 1, 4)), r0mul=max(float(mat["specular_r0_mul"]) * 2.0, 0.0),
             st=float(np.clip(mat["specular_trans"], 0, 1)), dt=float(np.clip(mat["diffuse_trans"], 0, 1)),
             dtf=float(np.clip(mat["diffuse_trans_fwd"], 0, 1)))
    c["a"] = float(np.clip(c["r"] ** 2, 1e-3, 1.0))
    return c


def _schlick(x):
    return np.clip(1.0 - x, 0.0, 1.0) ** 5


def _hg(mu, g):
    return (1 - g * g) / (1 + g * g - 2 * g * mu) ** 1.5 / (4 * np.pi)


def lobes_fp64(mat, wo, wi):
    """f*cos of each lobe (principled_bsdf.cpp:423-508) in FP64, written with the angle forms of D and
    Lambda and the two-index Fresnel above; returns (n, 4, 3)."""
    c = _closure(mat)
    wo = np.asarray(wo, np.float64)
    wi = np.asarray(wi, np.float64)
    n = wo.shape[0]
    a = c["a"]
    out = np.zeros((n, 4, 3))
    # diffuse with the Disney retro-reflection factor
    wh = wo + wi
    wh /= np.linalg.norm(wh, axis=1, keepdims=True)
    fd90 = 0.5 + 2.0 * c["r"] * wh[:, 2]
    retro = (1 + (fd90 - 1) * _schlick(wi[:, 2])) * (1 + (fd90 - 1) * _schlick(wo[:, 2]))
    w_d = (1 - c["metal"]) * (1 - c["st"]) * (1 - c["dt"])
    out[:, 0] = w_d * c["base"][None] / np.pi * (retro * wi[:, 2])[:, None]
    # diffuse transmission with the HG lobe around -wo
    w_dt = (1 - c["metal"]) * (1 - c["st"]) * c["dt"]
    out[:, 1] = w_dt * c["base"][None] / np.pi * (-wi[:, 2] * _hg(np.sum(wo * -wi, 1), c["dtf"]))[:, None]
    # metallic: D G2 F / (4 cos_o), F = Schlick towards white
    lo, li = ggx_lambda_closed(wo, a, a), ggx_lambda_closed(wi, a, a)
    g2_same = 1.0 / (1 + lo + li)
    D = ggx_D_walter(wh, a, a)
    fs = _schlick(np.sum(wo * wh, 1))
    F = c["base"][None] + (1 - c["base"][None]) * fs[:, None]
    out[:, 2] = c["metal"] * (D * g2_same / (4 * wo[:, 2]))[:, None] * F
    # dielectric: reflection and rough transmission (Walter 2007, eq. 21 without the eta^2 radiance factor)
    refl = wo[:, 2] * wi[:, 2] >= 0
    eta = np.where(wo[:, 2] >= 0, c["ior"], 1 / c["ior"])
    ht = wo + wi * eta[:, None]
    wh2 = np.where(refl[:, None], wo + wi, ht)
    wh2 /= np.linalg.norm(wh2, axis=1, keepdims=True)
    wh2 = np.where(wh2[:, 2:3] < 0, -wh2, wh2)
    D2 = ggx_D_walter(wh2, a, a)
    G2 = np.where(refl, g2_same, special.beta(1 + lo, 1 + li))
    odh, idh = np.sum(wo * wh2, 1), np.sum(wi * wh2, 1)
    # fresnel_dielectric(|wo.wh|, 1/eta): incident index 1/eta relative to 1
    Fr = fresnel_pbrt(np.abs(odh), 1.0 / eta, 1.0)
    R0 = ((1 - c["ior"]) / (1 + c["ior"])) ** 2
    sch = _schlick(np.abs(odh))
    den = R0 + sch * (1 - R0)
    adj = np.where(den > 0, (c["r0mul"] * R0 + sch * (1 - c["r0mul"] * R0)) / np.where(den > 0, den, 1), 1.0)
    f_r = D2 * G2 * Fr * adj / (4 * np.abs(wo[:, 2]))
    f_t = np.abs(D2 * G2 * (1 - Fr) * odh * idh / ((odh + eta * idh) ** 2 * wo[:, 2]))
    f = (1 - c["metal"]) * np.where(refl, f_r, c["st"] * f_t)
    out[:, 3] = f[:, None]
    return out


@pytest.mark.parametrize("mi", range(7))
def test_lobes_match_fp64_closed_forms(mi):
    mat = _materials()[mi]
    rng = np.random.default_rng(10 + mi)
    n = 6000
    wo = _unit(rng, n)
    wi = _unit(rng, n)
    keep = (np.abs(wo[:, 2]) > 0.05) & (np.abs(wi[:, 2]) > 0.05)
    wo, wi = wo[keep].astype(np.float32), wi[keep].astype(np.float32)
    want = lobes_fp64(mat, wo, wi)
    a = _closure(mat)["a"]
    for lobe in range(4):
        got = ko.bsdf_lobe(mat, lobe, wo, wi).astype(np.float64)
        if lobe == 0 or lobe == 2:  # the full eval only adds these on the upper/upper side
            sel = (wo[:, 2] > 0) & (wi[:, 2] > 0)
        elif lobe == 1:
            sel = (wo[:, 2] > 0) & (wi[:, 2] < 0)
        else:
            sel = np.ones(len(wo), bool)
        w = want[sel, lobe]
        g = got[sel]
        # GGX near its peak: relative float error ~ |d ln D / d z| ulp(z)
        ht = wo[sel].astype(np.float64) + wi[sel].astype(np.float64)
        z = np.abs(ht[:, 2]) / np.linalg.norm(ht, axis=1)
        cond = 1e-4 + 64 * 2.0**-24 * 4 * z * (1 - a * a) / np.maximum(1 + (a * a - 1) * z * z, 1e-12)
        err = np.abs(g - w)
        ok = err <= cond[:, None] * np.abs(w) + 1e-7
        assert ok.mean() > 0.9995, (lobe, ok.mean(), np.max(err / (np.abs(w) + 1e-7)))


@pytest.mark.parametrize("mi", [0, 1, 2, 5, 6])
def test_reflection_is_reciprocal(mi):
    """For opaque materials f(wo, wi) / cos(wi) = f(wi, wo) / cos(wo) (Helmholtz reciprocity of every
    reflection lobe); a transposed operand breaks it."""
    mat = _materials()[mi]
    rng = np.random.default_rng(20 + mi)
    wo = _unit(rng, 5000, upper=True)
    wi = _unit(rng, 5000, upper=True)
    keep = (wo[:, 2] > 0.05) & (wi[:, 2] > 0.05)
    wo, wi = wo[keep].astype(np.float32), wi[keep].astype(np.float32)
    f1, _ = ko.bsdf_eval(mat, wo, wi)
    f2, _ = ko.bsdf_eval(mat, wi, wo)
    a = f1.astype(np.float64) / wi[:, 2:3]
    b = f2.astype(np.float64) / wo[:, 2:3]
    assert np.allclose(a, b, rtol=2e-3, atol=1e-6), np.max(np.abs(a - b) / (np.abs(b) + 1e-6))


def test_metal_albedo_closed_forms():
    """White metal (F = 1) seen head-on: the albedo is int D(m) cos(theta_m) G2 / G1 dm with wi the mirror
    of wo about m (G1(wo) = 1 at normal incidence); the FP64 quadrature of the closed forms must equal
    the mean sample weight f / pdf of the oracle's sampler."""
    rng = np.random.default_rng(5)
    for rough in (0.2, 0.3, 0.5, 1.0):
        a = max(rough**2, 1e-3)
        u = np.geomspace(1e-9, 1.0, 20001)
        th = np.unique(np.concatenate([[0.0], np.arctan(a * np.sqrt(u / (1 - u + 1e-15)) * 3), np.linspace(0, np.pi / 2, 2000)]))
        th = th[th <= np.pi / 2]
        m = _dirs(th, np.zeros_like(th))
        wi_z = 2 * m[:, 2] ** 2 - 1
        wi = np.stack([np.sqrt(np.maximum(0, 1 - wi_z**2)), np.zeros_like(wi_z), wi_z], -1)
        g2_over_g1 = np.where(wi_z > 0, 1.0 / (1.0 + ggx_lambda_closed(wi, a, a)), 0.0)
        albedo = 2 * np.pi * np.trapezoid(ggx_D_walter(m, a, a) * m[:, 2] * g2_over_g1 * np.sin(th), th)
        mat = _materials()[6].copy()
        mat["roughness"] = rough
        n = 100_000
        wo = np.tile(np.array([[0.0, 0.0, 1.0]], np.float32), (n, 1))
        _, beta, _ = ko.bsdf_sample(mat, wo, rng.uniform(0, 1, (n, 3)).astype(np.float32))
        b = beta[:, 0].astype(np.float64)
        # the float GGX is ill-conditioned at the peak of small alphas (relative ~ 4 ulp(1) / a^2 per
        # sample, roughly symmetric), hence the extra term
        tol = 5 * b.std() / np.sqrt(n) + 2.0**-22 / a**2
        assert abs(b.mean() - albedo) < tol, (rough, b.mean(), albedo, tol)
        assert albedo <= 1.0


def _sphere_grid(nz=64, nphi=128, sub=6):
    """Equal-area bins (uniform in z and phi) with sub x sub midpoint nodes each."""
    zk = -1 + (np.arange(nz * sub) + 0.5) * (2.0 / (nz * sub))
    pk = -np.pi + (np.arange(nphi * sub) + 0.5) * (2 * np.pi / (nphi * sub))
    Z, P = np.meshgrid(zk, pk, indexing="ij")
    s = np.sqrt(1 - Z**2)
    dirs = np.stack([s * np.cos(P), s * np.sin(P), Z], -1)
    dw = (2.0 / (nz * sub)) * (2 * np.pi / (nphi * sub))
    return dirs.reshape(-1, 3), dw, (nz, nphi, sub)


def lobe_weights_fp64(mat, wo):
    """lobe_sample_weights (principled_bsdf.cpp:518-540) in FP64."""
    c = _closure(mat)
    wo = np.asarray(wo, np.float64)
    L = float(c["base"] @ np.array([0.212671, 0.715160, 0.072169]))
    up = wo[:, 2] > 0
    wd = np.where(up, (1 - c["metal"]) * (1 - c["st"]) * L / np.pi, 0.0)
    eta = np.where(wo[:, 2] >= 0, c["ior"], 1 / c["ior"])
    w = np.stack([wd * (1 - c["dt"]), wd * c["dt"],
                  np.where(up, c["metal"] * (L + (1 - L) * _schlick(wo[:, 2])), 0.0),
                  (1 - c["metal"]) * fresnel_pbrt(np.abs(wo[:, 2]), 1 / eta, 1.0)], 1)
    s = w.sum(1, keepdims=True)
    return np.where(s > 1e-20, w / np.where(s > 1e-20, s, 1), 0)  # the float sum is exactly 0 there


def sampling_densities(mat, wo, wi):
    """Solid-angle density of the directions each lobe's sampler produces (principled_bsdf.cpp:543-643),
    derived in FP64 from the samplers themselves: cosine hemisphere; Henyey-Greenstein around -wo; VNDF
    microfacet normals m around wo_s = sgn(wo.z) wo, reflected (Jacobian 1 / (4 |wo.m|)) or refracted
    (Jacobian eta^2 |wi.m| / (wo.m + eta wi.m)^2) with the Fresnel-weighted choice. Only directions a
    sampler can reach get density: reflected/refracted directions need wo_s.m > 0 for their m, and a
    refracted direction lies on the other side of its microfacet. (pdf() of the reference drops the last
    condition, so it also gives density to transmitted directions no sampler produces.)
    Returns (n, 4)."""
    c = _closure(mat)
    wo = np.asarray(wo, np.float64)
    wi = np.asarray(wi, np.float64)
    a = c["a"]
    n = wi.shape[0]
    q = np.zeros((n, 4))
    q[:, 0] = np.maximum(wi[:, 2], 0.0) / np.pi
    if c["dtf"] > 0.05:
        q[:, 1] = _hg(np.sum(-wo * wi, 1), c["dtf"])
    else:
        q[:, 1] = np.maximum(-wi[:, 2], 0.0) / np.pi
    sg = np.where(wo[:, 2] >= 0, 1.0, -1.0)
    wo_s = wo * sg[:, None]
    g1 = 1.0 / (1.0 + ggx_lambda_closed(wo_s, a, a))

    def vndf(m):  # density of m: G1(wo_s) <wo_s, m>+ D(m) / wo_s.z
        return g1 * np.maximum(0.0, np.sum(wo_s * m, 1)) * ggx_D_walter(m, a, a) / wo_s[:, 2]

    # reflection through m = +-normalize(wo + wi), m.z > 0
    same = wo[:, 2] * wi[:, 2] > 0
    mr = wo + wi
    mr /= np.maximum(np.linalg.norm(mr, axis=1, keepdims=True), 1e-300)
    mr = np.where(mr[:, 2:3] < 0, -mr, mr)
    jr = 1.0 / (4.0 * np.maximum(np.abs(np.sum(wo * mr, 1)), 1e-300))
    q_refl = np.where(same, vndf(mr) * jr, 0.0)
    q[:, 2] = q_refl
    # dielectric: Fresnel-weighted choice between reflection and refraction at m
    eta = np.where(wo[:, 2] >= 0, c["ior"], 1 / c["ior"])

    def choice(m):
        odm = np.abs(np.sum(wo * m, 1))
        Fr = fresnel_pbrt(odm, 1.0 / eta, 1.0)
        R0 = ((1 - c["ior"]) / (1 + c["ior"])) ** 2
        sch = _schlick(odm)
        den = R0 + sch * (1 - R0)
        adj = np.where(den > 0, (c["r0mul"] * R0 + sch * (1 - c["r0mul"] * R0)) / np.where(den > 0, den, 1), 1.0)
        fr, ft = Fr * adj, (1 - Fr) * c["st"]
        fr = np.where(ft == 0, 1.0, fr)
        return fr / (fr + ft), ft / (fr + ft)

    pr_r, _ = choice(mr)
    mt = wo + wi * eta[:, None]
    mt /= np.maximum(np.linalg.norm(mt, axis=1, keepdims=True), 1e-300)
    mt = np.where(mt[:, 2:3] < 0, -mt, mt)
    _, pr_t = choice(mt)
    odt, idt = np.sum(wo * mt, 1), np.sum(wi * mt, 1)
    jt = eta**2 * np.abs(idt) / np.maximum((odt + eta * idt) ** 2, 1e-300)
    # a refraction leaves wo and wi on opposite sides of its microfacet: (wo.m)(wi.m) < 0
    q_refr = np.where((wo[:, 2] * wi[:, 2] < 0) & (odt * idt < 0), vndf(mt) * jt * pr_t, 0.0)
    q[:, 3] = q_refl * pr_r + q_refr
    return q


SINGLE_LOBE = {
    # (lobe, material fields): lobes other than `lobe` get zero weight
    "diffuse": (0, dict(basecolor=[1, 1, 1], roughness=0.5, ior=1.0)),
    "diffuse_trans_hg": (1, dict(basecolor=[1, 1, 1], diffuse_trans=1.0, diffuse_trans_fwd=0.5, ior=1.0)),
    "diffuse_trans_cos": (1, dict(basecolor=[1, 1, 1], diffuse_trans=1.0, diffuse_trans_fwd=0.0, ior=1.0)),
    "metal_0.3": (2, dict(basecolor=[1, 1, 1], roughness=0.3, metallic=1.0)),
    "metal_0.7": (2, dict(basecolor=[0.5, 0.5, 0.5], roughness=0.7, metallic=1.0)),
    "glass_0.3": (3, dict(basecolor=[0, 0, 0], roughness=0.3, ior=1.5, specular_trans=1.0)),
    "glass_0.6_half": (3, dict(basecolor=[0, 0, 0], roughness=0.6, ior=1.33, specular_trans=0.5)),
}


def _make(fields):
    m = np.zeros((), MATERIAL_DTYPE)
    m["specular_r0_mul"] = 0.5
    for k, v in fields.items():
        m[k] = v
    return m


def _hist(dirs, nz, nphi):
    zi = np.clip(((dirs[:, 2].astype(np.float64) + 1) / 2 * nz).astype(int), 0, nz - 1)
    phi = np.arctan2(dirs[:, 1].astype(np.float64), dirs[:, 0].astype(np.float64))
    pj = np.clip(((phi + np.pi) / (2 * np.pi) * nphi).astype(int), 0, nphi - 1)
    return np.bincount(zi * nphi + pj, minlength=nz * nphi).astype(np.float64)


def _chi2_ok(obs, expected_prob, n_ok):
    exp = expected_prob / expected_prob.sum() * n_ok
    keep = exp > 20
    chi2 = float((((obs - exp) ** 2)[keep] / exp[keep]).sum())
    dof = int(keep.sum()) - 1
    # the bin probabilities come from a midpoint quadrature: allow its small bias on top of the 1e-4 quantile
    return chi2 < stats.chi2.isf(1e-4, dof) * 1.03 and obs[~keep].sum() <= 3 * exp[~keep].sum() + 50, (chi2, dof)


@pytest.mark.parametrize("case", list(SINGLE_LOBE))
@pytest.mark.parametrize("wo_z", [0.9, 0.35, -0.6])
def test_single_lobe_sampler_follows_its_density(case, wo_z):
    """Each lobe's sampler on a material where only that lobe has weight:
    - the histogram of sampled directions follows the density derived from the sampler (chi^2);
    - int q dw = 1 - P(sample rejected) (quadrature against the sampler's rejection rate);
    - the returned pdf equals that density, and the weight equals eval / pdf;
    - E[weight] equals int f dw over the directions the sampler reaches."""
    lobe, fields = SINGLE_LOBE[case]
    mat = _make(fields)
    so = np.sqrt(1 - wo_z**2)
    wo1 = np.array([so * 0.6, so * 0.8, wo_z], np.float32)
    w = lobe_weights_fp64(mat, wo1[None])[0]
    if w.sum() == 0:  # e.g. reflection lobes seen from below: the sampler must reject everything
        _, _, pdf = ko.bsdf_sample(mat, np.tile(wo1, (1000, 1)), np.random.default_rng(0).uniform(0, 1, (1000, 3)).astype(np.float32))
        assert np.all(pdf == 0)
        return
    assert w[lobe] == 1.0 and ko.bsdf_lobe_weights(mat, wo1[None])[0, lobe] == 1.0
    rng = np.random.default_rng(list(SINGLE_LOBE).index(case) * 100 + int(10 * wo_z) + 50)
    n = 200_000
    wo = np.tile(wo1, (n, 1))
    wi, beta, pdf = ko.bsdf_sample(mat, wo, rng.uniform(0, 1, (n, 3)).astype(np.float32))
    ok = pdf > 0
    q_s = sampling_densities(mat, wo[ok], wi[ok])[:, lobe]
    assert np.allclose(pdf[ok], q_s, rtol=2e-3, atol=1e-6), np.max(np.abs(pdf[ok] - q_s) / q_s)
    f, _ = ko.bsdf_eval(mat, wo[ok], wi[ok])
    assert np.allclose(beta[ok], f / pdf[ok, None], rtol=2e-4, atol=1e-6)
    dirs, dw, (nz, nphi, sub) = _sphere_grid()
    q = sampling_densities(mat, np.tile(wo1, (len(dirs), 1)), dirs.astype(np.float32))[:, lobe]
    p_fail = 1.0 - ok.mean()
    assert abs(q.sum() * dw - (1.0 - p_fail)) < 5e-3 + 4 * np.sqrt(p_fail * (1 - p_fail) / n), (q.sum() * dw, p_fail)
    good, info = _chi2_ok(_hist(wi[ok], nz, nphi), q.reshape(nz, sub, nphi, sub).sum((1, 3)).ravel(), ok.sum())
    assert good, info
    fg, _ = ko.bsdf_eval(mat, np.tile(wo1, (len(dirs), 1)), dirs.astype(np.float32))
    f_int = (fg.astype(np.float64) * (q > 0)[:, None]).sum(0) * dw
    b = beta.astype(np.float64)
    assert np.all(np.abs(b.mean(0) - f_int) < 5 * b.std(0) / np.sqrt(n) + 5e-3 * np.abs(f_int) + 1e-4), (b.mean(0), f_int)


@pytest.mark.parametrize("mi", [0, 1, 2, 3, 4, 5])
@pytest.mark.parametrize("wo_z", [0.9, 0.35, -0.6])
def test_mixture_sampling(mi, wo_z):
    """The full sample() (principled_bsdf.cpp:326-393) on the mixed materials: directions follow
    sum_l w_l q_l with the FP64 lobe weights (chi^2); the weight is eval / pdf; the returned pdf equals
    pdf() (:395-421) except for the reference's own mismatch: a diffuse-transmission (Henyey-Greenstein)
    sample that lands above the surface gets pdf_lobe[1] + pdf_lobe[3] in sample() but the reflection
    lobes' pdfs in pdf(), which gates pdf_lobe[1] to wi below the surface."""
    mat = _materials()[mi]
    so = np.sqrt(1 - wo_z**2)
    wo1 = np.array([so * 0.6, so * 0.8, wo_z], np.float32)
    w = lobe_weights_fp64(mat, wo1[None])[0]
    rng = np.random.default_rng(200 + mi)
    n = 200_000
    wo = np.tile(wo1, (n, 1))
    u = rng.uniform(0, 1, (n, 3)).astype(np.float32)
    wi, beta, pdf = ko.bsdf_sample(mat, wo, u)
    ok = pdf > 0
    if w.sum() == 0:
        assert not ok.any()
        return
    f, p_eval = ko.bsdf_eval(mat, wo[ok], wi[ok])
    assert np.allclose(beta[ok], f / pdf[ok, None], rtol=2e-4, atol=1e-6)
    cdf = np.cumsum(w)
    lobe = np.searchsorted(cdf, u[ok, 0].astype(np.float64), side="right")
    near_edge = np.min(np.abs(u[ok, 0, None].astype(np.float64) - cdf[None]), axis=1) < 1e-5
    quirk = (lobe == 1) & (wi[ok, 2] > 0)
    sel = ~quirk & ~near_edge
    # GGX conditioning at small roughness (material 2: alpha = 0.0025)
    a = _closure(mat)["a"]
    rtol = 1e-4 + 2.0**-20 / a**2
    rel = np.abs(pdf[ok][sel] - p_eval[sel]) / p_eval[sel]
    # sample() keeps its microfacet normal, pdf() recomputes it from wo and wi: a few small refraction
    # pdfs move by more than the bulk
    assert (rel <= rtol).mean() > 0.9999 and rel.max() < 1e-2 + 2.0**-20 / a**2, (np.sort(rel)[-5:], rtol)
    if _closure(mat)["r"] < 0.1:
        return  # too peaked for the sphere quadrature below
    dirs, dw, (nz, nphi, sub) = _sphere_grid()
    q = sampling_densities(mat, np.tile(wo1, (len(dirs), 1)), dirs.astype(np.float32)) @ w
    p_fail = 1.0 - ok.mean()
    assert abs(q.sum() * dw - (1.0 - p_fail)) < 5e-3 + 4 * np.sqrt(p_fail * (1 - p_fail) / n), (q.sum() * dw, p_fail)
    good, info = _chi2_ok(_hist(wi[ok], nz, nphi), q.reshape(nz, sub, nphi, sub).sum((1, 3)).ravel(), ok.sum())
    assert good, info


def test_lobe_selection_weights():
    """lobe_sample_weights (principled_bsdf.cpp:518-540) against its FP64 form: diffuse ~ (1-m)(1-st)
    lum / pi, metallic ~ m lerp(lum, 1, schlick(cos_o)), dielectric ~ (1-m) F(cos_o), normalised; the
    reflection lobes vanish below the surface."""
    rng = np.random.default_rng(7)
    for mat in _materials():
        wo = _unit(rng, 500).astype(np.float32)
        if mat["ior"] == 1.0:
            # below the surface only the dielectric lobe is left, whose Fresnel weight at eta = 1 is zero
            # up to float cancellation (a 1e-11 weight normalises to 1, FP64's 1e-33 falls to 0)
            wo = wo[wo[:, 2] > 0.02]
        got = ko.bsdf_lobe_weights(mat, wo).astype(np.float64)
        w = lobe_weights_fp64(mat, wo)
        assert np.allclose(got, w, atol=2e-6), np.max(np.abs(got - w))
