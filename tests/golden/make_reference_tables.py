"""Writes tests/golden/reference_tables.json from the reference's source text.

The reference cannot be built here (Eigen, Embree, TBB are absent), so its sampler tables and hash
constants are read as data from its sources and committed as a fixture; the tests then check the
oracle's and the GPU library's copies against them:
  - sobol_matrices_32[0..103]      (sobol.cpp:154-172): generator matrices of dimensions 0 and 1
  - the 24 digit permutations       (sobol.h:171-197)
  - FastOwenScrambler constants     (sobol.h:17-33)
  - mix_bits / murmur constants     (hash.h:140-214)
  - the sRGB u8 -> linear table     (texture.cpp:376-401) and the EWA weight table (texture.cpp:720-856)
Run from the repository root: python tests/golden/make_reference_tables.py [/root/reference]
"""

from __future__ import annotations

import json
import re
import sys
from pathlib import Path

OUT = Path(__file__).resolve().parent / "reference_tables.json"


def main(ref: Path) -> None:
    sobol_cpp = (ref / "sobol.cpp").read_text()
    start = sobol_cpp.index("uint32_t sobol_matrices_32[")
    body = sobol_cpp[start : start + 20000]
    words = [int(w, 16) for w in re.findall(r"0x[0-9a-fA-F]{8}", body)[:104]]
    assert len(words) == 104

    sobol_h = (ref / "sobol.h").read_text()
    pstart = sobol_h.index("permutations[24][4]")
    pend = sobol_h.index("};", pstart)
    perms = [[int(x) for x in m] for m in re.findall(r"\{(\d), (\d), (\d), (\d)\}", sobol_h[pstart:pend])]
    assert len(perms) == 24

    scr = sobol_h[sobol_h.index("struct FastOwenScrambler") :]
    scr = scr[: scr.index("uint32_t seed;")]
    scrambler_consts = [int(c, 16) for c in re.findall(r"0x[0-9a-fA-F]+", scr)]

    hash_h = (ref / "hash.h").read_text()
    mix = hash_h[hash_h.index("inline uint64_t mix_bits(uint64_t v)") :]
    mix = mix[: mix.index("}")]
    mix_consts = [int(c, 16) for c in re.findall(r"0x[0-9a-fA-F]+", mix)]
    mix_shifts = [int(s) for s in re.findall(r">> (\d+)", mix)]
    mur = hash_h[hash_h.index("inline uint64_t murmur_hash_64A") :]
    mur = mur[: mur.index("return h;")]
    murmur_m = int(re.search(r"const uint64_t m = (0x[0-9a-fA-F]+)", mur).group(1), 16)
    murmur_r = int(re.search(r"const int r = (\d+);", mur).group(1))

    tex_cpp = (ref / "texture.cpp").read_text()
    lut = tex_cpp[tex_cpp.index("srgb_u8_to_linear_f32_lut[256] = {") :]
    lut = lut[lut.index("{") + 1 : lut.index("};")]
    srgb_lut = [float(x) for x in re.findall(r"[0-9]+\.[0-9]+", lut)]
    assert len(srgb_lut) == 256
    ewa = tex_cpp[tex_cpp.index("ewa_lut[ewa_lut_size] = {") :]
    ewa = ewa[ewa.index("{") + 1 : ewa.index("};")]
    ewa_lut = [float(x.rstrip("f")) for x in re.findall(r"[0-9]+\.[0-9]*(?:e-?[0-9]+)?f", ewa)]
    assert len(ewa_lut) == 128, len(ewa_lut)

    data = {
        "source": "keepsake reference sources (sobol.cpp, sobol.h, hash.h, texture.cpp), read as text",
        "sobol_matrices_dim01": words,
        "perm4": perms,
        "owen_scrambler_constants": scrambler_consts,
        "mix_bits_constants": mix_consts,
        "mix_bits_shifts": mix_shifts,
        "murmur64a_m": murmur_m,
        "murmur64a_r": murmur_r,
        "srgb_u8_to_linear": srgb_lut,
        "ewa_weights": ewa_lut,
    }
    OUT.write_text(json.dumps(data, indent=1) + "\n")
    print(f"wrote {OUT}")


if __name__ == "__main__":
    main(Path(sys.argv[1] if len(sys.argv) > 1 else "/root/reference"))
