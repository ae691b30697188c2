"""Test infrastructure only: a numpy restatement of the reference's image output path, the checker for
keepsake_amd/host/image_io.cpp. Only tests import it.

- tone mappers of tonemap.h: ACES fit (tonemap.h:27-76), AgX with its looks (tonemap.h:84-210),
  Khronos PBR Neutral (tonemap.h:212-237), create_tone_mapper (tonemap.h:239-264);
- linear_to_srgb (maths.h:651-662);
- the PNG composite of RenderTarget2::composite_and_save_to_png (render_target.cpp:127-142) and the EXR
  composite of composite_and_save_to_exr (render_target.cpp:181-192).

All arithmetic is float32 in the reference's order of operations; vectors are (..., 3) arrays.

Parity pins (tests/test_image_io.py): the ACES and AgX matrices keep grey grey (rows sum to one), so a
grey input has the closed forms below; Khronos PBR Neutral is the identity shift v - f below its knee
and tends to one; half conversion is checked against numpy's IEEE float16; the composites' two formulas
agree with each other wherever both are defined. No reference golden vectors exist for these functions
(the reference's test data submodule is not part of the snapshot), so beyond these checks the
tone-mapper curves are "parity unpinned"."""

import numpy as np

F = np.float32

ACES_IN = np.array([[0.59719, 0.35458, 0.04823], [0.07600, 0.90834, 0.01566], [0.02840, 0.13383, 0.83777]], F)
ACES_OUT = np.array([[1.60475, -0.53108, -0.07367], [-0.10208, 1.10813, -0.00605], [-0.00327, -0.07276, 1.07602]], F)
AGX_INSET = np.array([[0.856627153315983, 0.0951212405381588, 0.0482516061458583],
                      [0.137318972929847, 0.761241990602591, 0.101439036467562],
                      [0.11189821299995, 0.0767994186031903, 0.811302368396859]], F)
AGX_OUTSET_INV = np.array([[0.899796955911611, 0.0871996192028351, 0.013003424885555],
                           [0.11142098895748, 0.875575586156966, 0.0130034248855548],
                           [0.11142098895748, 0.0871996192028349, 0.801379391839686]], F)
REC2020_TO_SRGB = np.array([[1.6605, -0.5876, -0.0728], [-0.1246, 1.1329, -0.0083], [-0.0182, -0.1006, 1.1187]], F)
SRGB_TO_REC2020 = np.array([[0.6274, 0.3293, 0.0433], [0.0691, 0.9195, 0.0113], [0.0164, 0.0880, 0.8956]], F)
AGX_MIN_EV = F(-12.47393)
AGX_MAX_EV = F(4.026069)


def _mat(m, v):
    """mat3 * vec3 per row, summed left to right in float32."""
    v = v.astype(F)
    return np.stack([m[i, 0] * v[..., 0] + m[i, 1] * v[..., 1] + m[i, 2] * v[..., 2] for i in range(3)], axis=-1)


def aces(v, exposure=1.0):
    v = v.astype(F) * F(exposure / 0.6)
    v = _mat(ACES_IN, v)
    a = v * (v + F(0.0245786)) - F(0.000090537)
    b = v * (F(0.983729) * v + F(0.4329510)) + F(0.238081)
    v = a / b
    return np.clip(_mat(ACES_OUT, v), F(0), F(1))


def agx(v, exposure=1.0, look="none"):
    v = np.maximum(v.astype(F) * F(exposure), F(0))
    v = _mat(SRGB_TO_REC2020, v)
    v = _mat(AGX_INSET, v)
    v = (np.log2(np.maximum(v, F(1e-10))) - AGX_MIN_EV) / (AGX_MAX_EV - AGX_MIN_EV)
    x = np.clip(v, F(0), F(1))
    x2 = x * x
    x4 = x2 * x2
    v = (F(15.5) * x4 * x2 - F(40.14) * x4 * x + F(31.96) * x4 - F(6.868) * x2 * x + F(0.4298) * x2
         + F(0.1191) * x - F(0.00232))
    if look != "none":
        luma = v[..., 0] * F(0.2126) + v[..., 1] * F(0.7152) + v[..., 2] * F(0.0722)
        if look == "golden":
            slope, power, sat = np.array([1.0, 0.9, 0.5], F), F(0.8), F(1.3)
        elif look == "punchy":
            slope, power, sat = np.ones(3, F), F(1.35), F(1.4)
        else:
            raise ValueError(look)
        c = np.power(v * slope + F(0), power)
        v = luma[..., None] + sat * (c - luma[..., None])
    v = _mat(np.linalg.inv(AGX_OUTSET_INV.astype(np.float64)).astype(F), v)
    v = np.power(np.maximum(v, F(0)), F(2.2))
    return np.clip(_mat(REC2020_TO_SRGB, v), F(0), F(1))


def khronos_pbr_neutral(v, exposure=1.0):
    v = v.astype(F) * F(exposure)
    f90, ks, kd = F(0.04), F(0.8) - F(0.04), F(0.15)
    x = v.min(axis=-1)
    f = np.where(x <= F(2) * f90, x - x * x / (F(4) * f90), f90)
    p = (v - f[..., None]).max(axis=-1)
    pn = F(1) - (F(1) - ks) * (F(1) - ks) / (p + F(1) - F(2) * ks)
    g = F(1) / (kd * (p - pn) + F(1))
    low = v - f[..., None]
    high = (v - f[..., None]) * (pn / p)[..., None] * g[..., None] + (pn * (F(1) - g))[..., None]
    return np.where((p <= ks)[..., None], low, high).astype(F)


def tone_map(v, type, exposure=1.0, look="none"):
    with np.errstate(invalid="ignore", divide="ignore"):  # the reference's NaN / inf cases pass through
        if type == "aces":
            return aces(v, exposure)
        if type == "agx":
            return agx(v, exposure, look)
        if type == "khronos_pbr_neutral":
            return khronos_pbr_neutral(v, exposure)
    raise ValueError(type)


def linear_to_srgb(x):
    x = x.astype(F)
    with np.errstate(invalid="ignore"):
        y = np.where(x < F(0.0031308), x * F(12.92), np.power(np.maximum(x, F(0)), F(1.0) / F(2.4)) * F(1.055) - F(0.055))
    return np.where(x <= 0, F(0), np.where(x >= 1, F(1), y)).astype(F)


def srgb_bytes(c):
    q = np.floor(linear_to_srgb(c) * F(255))
    return np.clip(np.nan_to_num(q, nan=0.0), 0, 255).astype(np.uint8)


def composite_png(sums, weights, backdrop, tone=None):
    """(sum + w_miss * backdrop) / (w_hit + w_miss) in FP64, cast to float32, tone-mapped, sRGB bytes."""
    ws = weights[..., 0] + weights[..., 1]
    c = ((sums + weights[..., 1:2] * np.asarray(backdrop, np.float64)) / ws[..., None]).astype(F)
    if tone is not None:
        c = tone_map(c, **tone)
    return srgb_bytes(c)


def composite_exr(sums, weights, backdrop):
    """lerp(backdrop, sum / w_hit, t) with t = w_hit / (w_hit + w_miss) rounded to float32 (lerp's t is a
    float, maths.h:686-690): (1 - t) * backdrop + t * fg in FP64, cast to float32."""
    ws = np.floor(weights[..., 0] + weights[..., 1])  # uint32_t w_sum
    with np.errstate(invalid="ignore", divide="ignore"):
        t = np.where(ws == 0, 0.0, weights[..., 0] / ws).astype(F)
        fg = np.where(weights[..., :1] == 0, 0.0, sums / weights[..., :1])
    b = np.asarray(backdrop, F).astype(np.float64)
    return ((F(1) - t).astype(np.float64)[..., None] * b + t.astype(np.float64)[..., None] * fg).astype(F)
