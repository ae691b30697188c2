"""Image files and tone mappers of the host tools (keepsake_amd/host/image_io.cpp), checked without a
GPU through the `tonemap` task (tonemap.cpp:9-32): EXR inputs written by an independent Python writer
(NONE / RLE / ZIP, FLOAT / HALF, any channel order), Radiance .hdr inputs (flat and run-length), EXR
output read back by an independent reader, HALF output against numpy's IEEE float16, PNG output
against the sRGB encoding of the oracle's tone mappers (oracle/ks_image_oracle.py)."""

import subprocess
from pathlib import Path

import numpy as np
import pytest

import ks_image_oracle as io
from image_files import read_exr, read_exr_rgb, read_png, rgbe_to_float, to_rgbe, write_exr, write_hdr

ROOT = Path(__file__).resolve().parents[1]
TOOL = ROOT / "keepsake_amd" / "_lib" / "ks_render"


def run(tmp_path, text):
    cfg = tmp_path / "c.toml"
    cfg.write_text(f'output_dir = "{tmp_path / "out"}"\n' + text)
    return subprocess.run([str(TOOL), "-c", str(cfg), "-r", str(tmp_path)], capture_output=True, text=True)


def tonemap(tmp_path, inp, tm, out="o.exr", extra=""):
    r = run(tmp_path, f"""
[[task]]
type = "tonemap"
input_path = "{inp}"
output_path = "{out}"
tone_mapper = {tm}
{extra}""")
    assert r.returncode == 0, r.stderr
    return r


def hdr_image(h=37, w=53, seed=0):
    """Colours over many stops: darks, mid-tones, highlights, saturated and grey pixels."""
    rng = np.random.default_rng(seed)
    img = np.exp(rng.uniform(-9, 5, (h, w, 3))).astype(np.float32)
    img[0, :, :] = np.exp(np.linspace(-9, 5, w))[:, None]  # grey ramp
    special = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1], [1e-3, 2e-3, 4e-3]], np.float32)
    img[1, :min(5, w)] = special[:min(5, w)]
    return img


TONE_MAPPERS = [
    ('{ type = "aces" }', dict(type="aces")),
    ('{ type = "aces", exposure = 2.5 }', dict(type="aces", exposure=2.5)),
    ('{ type = "agx" }', dict(type="agx")),
    ('{ type = "agx", look = "golden", exposure = 0.7 }', dict(type="agx", look="golden", exposure=0.7)),
    ('{ type = "agx", look = "punchy" }', dict(type="agx", look="punchy")),
    ('{ type = "khronos_pbr_neutral" }', dict(type="khronos_pbr_neutral")),
    ('{ type = "khronos_pbr_neutral", exposure = 4.0 }', dict(type="khronos_pbr_neutral", exposure=4.0)),
]


@pytest.mark.parametrize("toml,tone", TONE_MAPPERS)
def test_tone_mappers_match_oracle(tmp_path, toml, tone):
    img = hdr_image()
    write_exr(tmp_path / "in.exr", img)
    tonemap(tmp_path, "in.exr", toml, extra='png_path = "o.png"')
    got, info = read_exr_rgb(tmp_path / "o.exr")
    assert info["compression"] == 3 and [c[1] for c in info["channels"]] == [2, 2, 2]
    assert [c[0] for c in info["channels"]] == ["B", "G", "R"]
    want = io.tone_map(img, **tone)
    np.testing.assert_allclose(got, want, rtol=3e-5, atol=3e-6)
    # the 8-bit sRGB copy: floor(255 * srgb) may move by one where the float curve sits on a step
    png = read_png(tmp_path / "o.png")
    diff = np.abs(png.astype(int) - io.srgb_bytes(want).astype(int))
    assert diff.max() <= 1 and (diff > 0).mean() < 2e-3


def test_tone_mapper_closed_forms():
    """Grey in, grey out (the ACES / AgX matrix rows sum to one); Khronos PBR Neutral is v - f below its
    knee with f = x - x^2 / 0.16 (x <= 0.08) or 0.04, and tends to one."""
    g = np.array([1e-3, 0.01, 0.18, 0.5, 1.0, 4.0], np.float32)
    grey = np.repeat(g[:, None], 3, axis=1)
    for m in (io.ACES_IN, io.ACES_OUT, io.AGX_INSET, io.AGX_OUTSET_INV):
        np.testing.assert_allclose(m.astype(np.float64).sum(axis=1), 1.0, atol=2e-5)
    a = io.aces(grey)
    v = g.astype(np.float64) / 0.6
    fit = (v * (v + 0.0245786) - 0.000090537) / (v * (0.983729 * v + 0.4329510) + 0.238081)
    np.testing.assert_allclose(a[:, 0], np.clip(fit, 0, 1), rtol=1e-5, atol=1e-6)
    assert np.all(np.abs(a - a[:, :1]) < 1e-5)
    assert np.all(np.diff(a[:, 0]) > 0)
    x = io.agx(grey)
    assert np.all(np.abs(x - x[:, :1]) < 2e-3) and np.all(np.diff(x[:, 0]) > 0)
    k = io.khronos_pbr_neutral(np.array([[0.1, 0.2, 0.3], [0.02, 0.5, 0.5], [1e4, 1e4, 1e4]], np.float32))
    np.testing.assert_allclose(k[0], [0.06, 0.16, 0.26], atol=1e-7)
    f = 0.02 - 0.02**2 / 0.16
    np.testing.assert_allclose(k[1], [0.02 - f, 0.5 - f, 0.5 - f], atol=1e-7)
    assert abs(k[2, 0] - 1.0) < 1e-3


@pytest.mark.parametrize("compression", [0, 1, 3])
@pytest.mark.parametrize("half", [False, True])
def test_exr_inputs(tmp_path, compression, half):
    img = hdr_image(h=40, w=19, seed=1)
    img[5:9] = 0.25  # flat rows give RLE runs
    write_exr(tmp_path / "in.exr", img[..., ::-1], channels="BGR", half=half, compression=compression)
    tonemap(tmp_path, "in.exr", '{ type = "khronos_pbr_neutral", exposure = 1e-3 }')
    got, _ = read_exr_rgb(tmp_path / "o.exr")
    src = img.astype(np.float16).astype(np.float32) if half else img
    np.testing.assert_allclose(got, io.tone_map(src, 'khronos_pbr_neutral', 1e-3), rtol=1e-6, atol=1e-9)


def test_exr_single_channel_and_alpha(tmp_path):
    img = hdr_image(h=9, w=7, seed=2)
    write_exr(tmp_path / "y.exr", img[..., :1], channels="Y")
    tonemap(tmp_path, "y.exr", '{ type = "aces" }')
    got, _ = read_exr_rgb(tmp_path / "o.exr")
    np.testing.assert_allclose(got, io.aces(np.repeat(img[..., :1], 3, axis=-1)), rtol=3e-5, atol=3e-6)
    rgba = np.concatenate([img, np.ones_like(img[..., :1])], axis=-1)
    write_exr(tmp_path / "a.exr", rgba, channels="RGBA")
    tonemap(tmp_path, "a.exr", '{ type = "aces" }')
    got, _ = read_exr_rgb(tmp_path / "o.exr")
    np.testing.assert_allclose(got, io.aces(img), rtol=3e-5, atol=3e-6)


def test_half_output_matches_ieee_float16(tmp_path):
    """HALF channels round to nearest even, subnormals included; overflow goes to infinity."""
    rng = np.random.default_rng(4)
    vals = np.concatenate([
        rng.uniform(-70000, 70000, 400), np.exp(rng.uniform(-20, 11, 400)), -np.exp(rng.uniform(-20, 11, 200)),
        np.array([0.0, -0.0, 65504, 65519.99, 65520, 1e6, 2.0**-24, 2.0**-25, 1.5 * 2.0**-25, 2.0**-26, 2.0**-14,
                  2.0**-14 - 2.0**-25, 1 + 2.0**-11, 1 + 3 * 2.0**-11, 2048 + 1, 2048 + 3, 0.1, 1 / 3]),
    ]).astype(np.float32)
    # plus values exactly halfway between neighbouring halves
    h = rng.integers(1, 0x7BFF, 300).astype(np.uint16).view(np.float16)
    nxt = (h.view(np.uint16) + 1).view(np.float16)
    mid = ((h.astype(np.float64) + nxt.astype(np.float64)) / 2).astype(np.float32)
    vals = np.concatenate([vals, mid, -mid])
    n = vals.size
    w = 16
    vals = np.pad(vals, (0, (-n) % w))
    img = np.repeat(vals.reshape(-1, w)[..., None], 3, axis=-1)
    write_exr(tmp_path / "in.exr", img, compression=0)
    # without a tone mapper the task copies the image: the HALF file holds float_to_half of each value
    r = run(tmp_path, """
[[task]]
type = "tonemap"
input_path = "in.exr"
output_path = "h.exr"
half = true
""")
    assert r.returncode == 0, r.stderr
    ch, info = read_exr(tmp_path / "h.exr")
    assert [c[1] for c in info["channels"]] == [1, 1, 1]
    got = ch["R"].astype(np.float16)  # widened from float16 exactly, so this narrowing is exact
    with np.errstate(over="ignore"):
        want = img[..., 0].astype(np.float16)
    assert np.array_equal(got.view(np.uint16), want.view(np.uint16))


@pytest.mark.parametrize("rle", [True, False])
def test_hdr_input(tmp_path, rle):
    img = hdr_image(h=12, w=40, seed=3)
    img[3:6] = img[3, 0]  # flat rows: long runs
    rgbe = to_rgbe(img)
    write_hdr(tmp_path / "in.hdr", rgbe, rle=rle)
    tonemap(tmp_path, "in.hdr", '{ type = "aces" }')
    got, _ = read_exr_rgb(tmp_path / "o.exr")
    np.testing.assert_allclose(got, io.aces(rgbe_to_float(rgbe)), rtol=3e-5, atol=3e-6)


def test_exr_output_stores_large_blocks_compressed(tmp_path):
    """A smooth 300 x 70 image: ZIP blocks of 16 lines, the last one short, each smaller than raw."""
    y, x = np.mgrid[0:70, 0:300].astype(np.float32)
    img = np.stack([x / 300, y / 70, 0.5 * np.ones_like(x)], axis=-1)
    write_exr(tmp_path / "in.exr", img, compression=0)
    tonemap(tmp_path, "in.exr", '{ type = "khronos_pbr_neutral" }')
    got, info = read_exr_rgb(tmp_path / "o.exr")
    assert info["width"] == 300 and info["height"] == 70
    np.testing.assert_allclose(got, io.tone_map(img, 'khronos_pbr_neutral'), rtol=1e-6, atol=1e-7)
    assert (tmp_path / "o.exr").stat().st_size < 300 * 70 * 12 * 0.5


@pytest.mark.parametrize(
    "setup,msg",
    [
        ("piz", "compression 4 is not supported"),
        ("ext", "Invalid image extension .png"),
        ("type", "Invalid tone mapper type [filmic]"),
        ("look", "Invalid AgX tone mapper look [sepia]"),
        ("notexr", "not an OpenEXR file"),
    ],
)
def test_tonemap_errors(tmp_path, setup, msg):
    img = hdr_image(h=4, w=4)
    write_exr(tmp_path / "in.exr", img)
    inp, tm = "in.exr", '{ type = "aces" }'
    if setup == "piz":
        b = bytearray((tmp_path / "in.exr").read_bytes())
        i = b.index(b"compression\0compression\0") + len("compression\0compression\0") + 4
        b[i] = 4
        (tmp_path / "in.exr").write_bytes(bytes(b))
    elif setup == "ext":
        (tmp_path / "in.png").write_bytes(b"")
        inp = "in.png"
    elif setup == "type":
        tm = '{ type = "filmic" }'
    elif setup == "look":
        tm = '{ type = "agx", look = "sepia" }'
    elif setup == "notexr":
        (tmp_path / "in.exr").write_bytes(b"\0" * 64)
    r = run(tmp_path, f"""
[[task]]
type = "tonemap"
input_path = "{inp}"
output_path = "o.exr"
tone_mapper = {tm}
""")
    assert r.returncode != 0
    assert msg in r.stderr


def test_composite_formulas_agree():
    """The EXR lerp form and the PNG sum form of the film composite are the same value wherever the hit
    weight is positive (up to the float rounding of the lerp parameter)."""
    rng = np.random.default_rng(6)
    n = 1000
    wh = rng.integers(1, 64, n).astype(np.float64)
    wm = rng.integers(0, 64, n).astype(np.float64)
    sums = rng.uniform(0, 5, (n, 3)) * wh[:, None]
    bd = np.array([0.25, 0.5, 0.75], np.float32)
    wts = np.stack([wh, wm], axis=-1)
    e = io.composite_exr(sums, wts, bd)
    p = ((sums + wm[:, None] * bd.astype(np.float64)) / (wh + wm)[:, None])
    np.testing.assert_allclose(e, p, rtol=1e-6)
    # no samples at all: the EXR composite shows the backdrop
    np.testing.assert_array_equal(io.composite_exr(np.zeros((1, 3)), np.zeros((1, 2)), bd)[0], bd)
