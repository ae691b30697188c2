"""The C++ task path (ks_render -> ks::SmallPT::run -> C ABI) renders the same film as a direct call."""

import subprocess
from pathlib import Path

import numpy as np
import pytest

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
"""
    )
    r = subprocess.run([str(TOOL), "-c", str(cfg)], capture_output=True, text=True, timeout=120)
    print(r.stdout[-2000:], r.stderr[-2000:], flush=True)
    assert r.returncode == 0
    out = tmp_path / "out" / "cornell"
    assert (out / "small_pt_spp4.pfm").exists() and (out / "render_stats.json").exists()
    img = read_pfm(out / "small_pt_spp8.pfm")
    sc = ks.Scene(device, cornell)
    film = ks.Film(device, 64, 64)
    ks.small_pt_render(sc, cornell.camera, ks.small_pt_params(bounces=8, render_width=64, render_height=64, spp=8,
                                                               spp_prog_interval=4), film)
    ref = film.composite()
    assert np.array_equal(img, ref)
    film.close()
    sc.close()
