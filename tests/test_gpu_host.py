"""The C++ task path (ks_render -> ks::SmallPT::run -> C ABI) renders the same film as a direct call, and
writes it as the reference's small_pt does: an EXR of the film composite per progress interval and,
with a tone mapper, a PNG of the tone-mapped sRGB composite (small_pt.cpp:278-294)."""

import subprocess
from pathlib import Path

import numpy as np
import pytest

import ks_image_oracle as io
from image_files import read_exr_rgb, read_png

pytestmark = pytest.mark.gpu

ROOT = Path(__file__).resolve().parents[1]
TOOL = ROOT / "keepsake_amd" / "_lib" / "ks_render"


def read_pfm(path):
    data = Path(path).read_bytes()
    parts = data.split(b"\n", 3)
    assert parts[0] == b"PF"
    w, h = (int(x) for x in parts[1].split())
    scale = float(parts[2])
    img = np.frombuffer(parts[3], dtype="<f4" if scale < 0 else ">f4").reshape(h, w, 3)
    return img[::-1]


def test_ks_render_matches_api(tmp_path, ks, cornell, device):
    cfg = tmp_path / "c.toml"
    cfg.write_text(
        f"""
output_dir = "{tmp_path / 'out'}"
[[task]]
type = "small_pt"
task_dir = "cornell"
procedural = "cornell"
bounces = 8
render_width = 64
render_height = 64
spp = 8
spp_prog_interval = 4
output_pfm = true
tone_mapper = {{ type = "agx", look = "punchy", exposure = 1.5 }}
backdrop = [0.1, 0.2, 0.3]
"""
    )
    r = subprocess.run([str(TOOL), "-c", str(cfg)], capture_output=True, text=True, timeout=120)
    print(r.stdout[-2000:], r.stderr[-2000:], flush=True)
    assert r.returncode == 0
    out = tmp_path / "out" / "cornell"
    for f in ("small_pt_spp4.exr", "small_pt_spp4.png", "small_pt_spp4.pfm", "render_stats.json"):
        assert (out / f).exists(), f
    exr, info = read_exr_rgb(out / "small_pt_spp8.exr")
    assert [c[1] for c in info["channels"]] == [2, 2, 2]  # FLOAT, as composite_and_save_to_exr writes
    sc = ks.Scene(device, cornell)
    film = ks.Film(device, 64, 64, backdrop=(0.1, 0.2, 0.3))
    ks.small_pt_render(sc, cornell.camera, ks.small_pt_params(bounces=8, render_width=64, render_height=64, spp=8,
                                                               spp_prog_interval=4, include_background=0,
                                                               backdrop=(0.1, 0.2, 0.3)), film)
    ref = film.composite()
    assert np.array_equal(read_pfm(out / "small_pt_spp8.pfm"), ref)
    assert np.array_equal(exr, ref)
    # the composite is the oracle's lerp form of the downloaded sums
    sums, wts = film.download()
    np.testing.assert_array_equal(ref, io.composite_exr(sums, wts, (0.1, 0.2, 0.3)))
    png = read_png(out / "small_pt_spp8.png")
    want = io.composite_png(sums, wts, (0.1, 0.2, 0.3), dict(type="agx", look="punchy", exposure=1.5))
    diff = np.abs(png.astype(int) - want.astype(int))
    print("png bytes off by one:", int((diff > 0).sum()), "of", diff.size, flush=True)
    assert diff.max() <= 1 and (diff > 0).mean() < 2e-3
    film.close()
    sc.close()


def test_oversized_film_is_a_clean_error(ks, device):
    """A film larger than the device (2^20 x 2^20 pixels, 40 TiB) fails with KS_ERR_OOM and a message; the
    device keeps working."""
    with pytest.raises(ks.KsError) as e:
        ks.Film(device, 1 << 20, 1 << 20)
    assert e.value.status == ks.KS_ERR_OOM and "alloc" in str(e.value)
    f = ks.Film(device, 8, 8)
    f.close()
