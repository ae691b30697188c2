"""Independent Python readers/writers of the file formats the host tools produce and consume, written
from the published format descriptions (OpenEXR scanline files with NONE / RLE / ZIP compression,
PNG, Radiance RGBE), for checking keepsake_amd/host/image_io.cpp."""

import struct
import zlib

import numpy as np

# ---- OpenEXR -------------------------------------------------------------------------------------

_LINES = {0: 1, 1: 1, 2: 1, 3: 16}


def _predict(raw):
    """Even bytes, then odd bytes, then deltas + 128 (the ZIP / RLE pre-pass)."""
    b = np.frombuffer(raw, np.uint8)
    t = np.concatenate([b[0::2], b[1::2]])
    d = t.astype(np.int32)
    d[1:] = (d[1:] - d[:-1] + 128) % 256
    return d.astype(np.uint8).tobytes()


def _unpredict(data):
    d = np.frombuffer(data, np.uint8).astype(np.int64)
    t = (np.cumsum(d - 128) + 128) % 256  # t[i] = t[i-1] + d[i] - 128, t[0] = d[0]
    t = t.astype(np.uint8)
    n = t.size
    h = (n + 1) // 2
    out = np.empty(n, np.uint8)
    out[0::2] = t[:h]
    out[1::2] = t[h:]
    return out.tobytes()


def _rle(data):
    out = bytearray()
    i = 0
    while i < len(data):
        j = i
        while j + 1 < len(data) and data[j + 1] == data[i] and j - i < 127:
            j += 1
        if j > i + 1:  # run of j - i + 1 >= 3
            out += struct.pack("b", j - i) + data[i:i + 1]
            i = j + 1
        else:
            k = i
            while k < len(data) and k - i < 127 and not (k + 2 < len(data) and data[k] == data[k + 1] == data[k + 2]):
                k += 1
            k = max(k, i + 1)
            out += struct.pack("b", -(k - i)) + data[i:k]
            i = k
    return bytes(out)


def write_exr(path, img, channels="RGB", half=False, compression=3):
    """img: (H, W, C) float array with channels named by `channels` (any order); written sorted."""
    h, w, c = img.shape
    names = list(channels)
    order = sorted(range(c), key=lambda k: names[k])
    ptype = 1 if half else 2
    hdr = bytearray(struct.pack("<II", 20000630, 2))

    def attr(name, typ, value):
        hdr.extend(name.encode() + b"\0" + typ.encode() + b"\0" + struct.pack("<i", len(value)) + value)

    cl = b"".join(names[k].encode() + b"\0" + struct.pack("<iBBBBii", ptype, 0, 0, 0, 0, 1, 1) for k in order) + b"\0"
    attr("channels", "chlist", cl)
    attr("compression", "compression", bytes([compression]))
    attr("dataWindow", "box2i", struct.pack("<4i", 0, 0, w - 1, h - 1))
    attr("displayWindow", "box2i", struct.pack("<4i", 0, 0, w - 1, h - 1))
    attr("lineOrder", "lineOrder", b"\0")
    attr("pixelAspectRatio", "float", struct.pack("<f", 1.0))
    attr("screenWindowCenter", "v2f", struct.pack("<2f", 0, 0))
    attr("screenWindowWidth", "float", struct.pack("<f", 1.0))
    hdr.append(0)
    lines = _LINES[compression]
    blocks = (h + lines - 1) // lines
    table_pos = len(hdr)
    body = bytearray(hdr + b"\0" * (8 * blocks))
    dt = np.float16 if half else np.float32
    for b in range(blocks):
        y0 = b * lines
        raw = b"".join(img[y, :, k].astype(dt).astype("<" + np.dtype(dt).str[1:]).tobytes()
                       for y in range(y0, min(h, y0 + lines)) for k in order)
        if compression == 3:
            data = zlib.compress(_predict(raw))
        elif compression == 1:
            data = _rle(_predict(raw))
        else:
            data = raw
        if len(data) >= len(raw):
            data = raw
        struct.pack_into("<Q", body, table_pos + 8 * b, len(body))
        body += struct.pack("<ii", y0, len(data)) + data
    with open(path, "wb") as f:
        f.write(body)


def read_exr(path):
    """-> (dict channel name -> (H, W) float32 array, header dict)."""
    b = open(path, "rb").read()
    magic, ver = struct.unpack_from("<II", b, 0)
    assert magic == 20000630 and ver & 0xFF == 2 and not ver & 0x200
    p = 8
    attrs = {}

    def cstr():
        nonlocal p
        e = b.index(b"\0", p)
        s = b[p:e].decode()
        p = e + 1
        return s

    while True:
        name = cstr()
        if not name:
            break
        typ = cstr()
        (size,) = struct.unpack_from("<i", b, p)
        p += 4
        attrs[name] = (typ, b[p:p + size])
        p += size
    chans = []
    cl = attrs["channels"][1]
    q = 0
    while cl[q] != 0:
        e = cl.index(b"\0", q)
        nm = cl[q:e].decode()
        ptype, _, _, _, _, xs, ys = struct.unpack_from("<iBBBBii", cl, e + 1)
        assert xs == 1 and ys == 1
        chans.append((nm, ptype))
        q = e + 1 + 16
    comp = attrs["compression"][1][0]
    x0, y0, x1, y1 = struct.unpack("<4i", attrs["dataWindow"][1])
    w, h = x1 - x0 + 1, y1 - y0 + 1
    lines = _LINES[comp]
    blocks = (h + lines - 1) // lines
    offsets = struct.unpack_from(f"<{blocks}Q", b, p)
    sizes = {1: 2, 2: 4, 0: 4}
    line_bytes = sum(w * sizes[t] for _, t in chans)
    out = {nm: np.zeros((h, w), np.float32) for nm, _ in chans}
    for off in offsets:
        yb, n = struct.unpack_from("<ii", b, off)
        data = b[off + 8:off + 8 + n]
        nl = min(lines, h - (yb - y0))
        if len(data) != line_bytes * nl:
            if comp == 3:
                data = _unpredict(zlib.decompress(data))
            elif comp == 1:
                raw = bytearray()
                i = 0
                while i < len(data):
                    (cnt,) = struct.unpack_from("b", data, i)
                    i += 1
                    if cnt < 0:
                        raw += data[i:i - cnt]
                        i -= cnt
                    else:
                        raw += data[i:i + 1] * (cnt + 1)
                        i += 1
                data = _unpredict(bytes(raw))
        o = 0
        for ln in range(nl):
            for nm, t in chans:
                dt = {1: "<f2", 2: "<f4", 0: "<u4"}[t]
                out[nm][yb - y0 + ln] = np.frombuffer(data, dt, w, o).astype(np.float32)
                o += w * sizes[t]
    return out, {"compression": comp, "channels": chans, "width": w, "height": h}


def read_exr_rgb(path):
    ch, info = read_exr(path)
    return np.stack([ch["R"], ch["G"], ch["B"]], axis=-1), info


# ---- PNG -----------------------------------------------------------------------------------------


def read_png(path):
    """8-bit, non-interlaced PNG -> (H, W, C) uint8, with every CRC checked."""
    b = open(path, "rb").read()
    assert b[:8] == b"\x89PNG\r\n\x1a\n"
    p = 8
    idat = b""
    while True:
        (n,) = struct.unpack_from(">I", b, p)
        typ = b[p + 4:p + 8]
        body = b[p + 8:p + 8 + n]
        (crc,) = struct.unpack_from(">I", b, p + 8 + n)
        assert zlib.crc32(typ + body) == crc, typ
        p += 12 + n
        if typ == b"IHDR":
            w, h, depth, ctype, _, _, lace = struct.unpack(">IIBBBBB", body)
            assert depth == 8 and lace == 0
        elif typ == b"IDAT":
            idat += body
        elif typ == b"IEND":
            break
    c = {0: 1, 4: 2, 2: 3, 6: 4}[ctype]
    raw = zlib.decompress(idat)
    stride = w * c
    img = np.zeros((h, stride), np.int32)
    prev = np.zeros(stride, np.int32)
    for y in range(h):
        f = raw[y * (stride + 1)]
        row = np.frombuffer(raw, np.uint8, stride, y * (stride + 1) + 1).astype(np.int32)
        if f == 0:
            cur = row
        elif f == 2:
            cur = (row + prev) % 256
        else:
            cur = np.zeros(stride, np.int32)  # Sub / Average / Paeth need the running left value
            for i in range(stride):
                a = cur[i - c] if i >= c else 0
                up = prev[i]
                ul = prev[i - c] if i >= c else 0
                if f == 1:
                    pr = a
                elif f == 3:
                    pr = (a + up) // 2
                else:
                    pa, pb, pc = abs(up - ul), abs(a - ul), abs(a + up - 2 * ul)
                    pr = a if pa <= pb and pa <= pc else (up if pb <= pc else ul)
                cur[i] = (row[i] + pr) % 256
        img[y] = cur
        prev = cur
    return img.astype(np.uint8).reshape(h, w, c)


# ---- Radiance RGBE -------------------------------------------------------------------------------


def to_rgbe(img):
    """float RGB -> RGBE bytes (mantissas from frexp of the largest channel)."""
    m = img.max(axis=-1)
    e = np.zeros(m.shape, np.int32)
    mant = np.zeros(m.shape)
    nz = m > 1e-32
    mant[nz], e[nz] = np.frexp(m[nz])
    out = np.zeros(img.shape[:-1] + (4,), np.uint8)
    scale = np.where(nz, mant * 256.0 / np.where(nz, m, 1), 0)
    out[..., :3] = np.clip(img * scale[..., None], 0, 255).astype(np.uint8)
    out[..., 3] = np.where(nz, e + 128, 0)
    return out


def write_hdr(path, rgbe, rle=True):
    h, w, _ = rgbe.shape
    out = bytearray(b"#?RADIANCE\nFORMAT=32-bit_rle_rgbe\n\n" + f"-Y {h} +X {w}\n".encode())
    for y in range(h):
        if not rle:
            out += rgbe[y].tobytes()
            continue
        out += bytes([2, 2, w >> 8, w & 255])
        for k in range(4):
            ch = rgbe[y, :, k]
            i = 0
            while i < w:
                j = i
                while j < w and ch[j] == ch[i] and j - i < 127:
                    j += 1
                if j - i >= 3:
                    out += bytes([128 + j - i, ch[i]])
                    i = j
                else:
                    k2 = min(w, i + 128)
                    n = k2 - i
                    out += bytes([n]) + ch[i:k2].tobytes()
                    i = k2
    with open(path, "wb") as f:
        f.write(out)


def rgbe_to_float(rgbe):
    e = rgbe[..., 3].astype(np.int32)
    s = np.where(e > 0, np.ldexp(np.float32(1), e - 136), 0).astype(np.float32)
    return rgbe[..., :3].astype(np.float32) * s[..., None]


def write_png_any(path, samples, ctype, depth, palette=None, trns=None, interlace=False, seed=0):
    """PNG writer for every colour type (0 gray, 2 RGB, 3 palette, 4 gray+alpha, 6 RGBA) and bit depth,
    with a random row filter per row (all five kinds) and optional Adam7 interlacing.
    samples: (H, W, C) integer array of raw samples (palette indices for type 3)."""
    rng = np.random.default_rng(seed)
    h, w, c = samples.shape

    def pack_rows(img):
        rows = []
        hh, ww, _ = img.shape
        for y in range(hh):
            flat = img[y].reshape(-1).astype(np.int64)
            if depth == 16:
                raw = b"".join(struct.pack(">H", int(v)) for v in flat)
            elif depth == 8:
                raw = bytes(int(v) for v in flat)
            else:
                bits = "".join(format(int(v), f"0{depth}b") for v in flat)
                bits += "0" * (-len(bits) % 8)
                raw = bytes(int(bits[i:i + 8], 2) for i in range(0, len(bits), 8))
            rows.append(raw)
        return rows

    bpp = max(1, c * depth // 8)

    def filter_rows(rows):
        out = bytearray()
        prev = bytes(len(rows[0])) if rows else b""
        for r in rows:
            f = int(rng.integers(0, 5))
            fr = bytearray(len(r))
            for i in range(len(r)):
                a = r[i - bpp] if i >= bpp else 0
                b = prev[i]
                cc = prev[i - bpp] if i >= bpp else 0
                if f == 0:
                    pr = 0
                elif f == 1:
                    pr = a
                elif f == 2:
                    pr = b
                elif f == 3:
                    pr = (a + b) // 2
                else:
                    p = a + b - cc
                    pa, pb, pc = abs(p - a), abs(p - b), abs(p - cc)
                    pr = a if pa <= pb and pa <= pc else (b if pb <= pc else cc)
                fr[i] = (r[i] - pr) % 256
            out += bytes([f]) + bytes(fr)
            prev = r
        return bytes(out)

    if interlace:
        data = b""
        for x0, y0, dx, dy in [(0, 0, 8, 8), (4, 0, 8, 8), (0, 4, 4, 8), (2, 0, 4, 4), (0, 2, 2, 4), (1, 0, 2, 2), (0, 1, 1, 2)]:
            sub = samples[y0::dy, x0::dx]
            if sub.size:
                data += filter_rows(pack_rows(sub))
    else:
        data = filter_rows(pack_rows(samples))

    def chunk(t, d):
        return struct.pack(">I", len(d)) + t + d + struct.pack(">I", zlib.crc32(t + d))

    out = b"\x89PNG\r\n\x1a\n" + chunk(b"IHDR", struct.pack(">IIBBBBB", w, h, depth, ctype, 0, 0, 1 if interlace else 0))
    if palette is not None:
        out += chunk(b"PLTE", bytes(np.asarray(palette, np.uint8).reshape(-1)))
    if trns is not None:
        out += chunk(b"tRNS", bytes(trns))
    out += chunk(b"IDAT", zlib.compress(data)) + chunk(b"IEND", b"")
    with open(path, "wb") as f:
        f.write(out)


def lz4_decode(src, size):
    """LZ4 block decoder from the published block format."""
    out = bytearray()
    i = 0
    while i < len(src):
        tok = src[i]
        i += 1
        lit = tok >> 4
        if lit == 15:
            while True:
                b = src[i]
                i += 1
                lit += b
                if b != 255:
                    break
        out += src[i:i + lit]
        i += lit
        if i >= len(src):
            break
        off = src[i] | src[i + 1] << 8
        i += 2
        ml = tok & 15
        if ml == 15:
            while True:
                b = src[i]
                i += 1
                ml += b
                if b != 255:
                    break
        ml += 4
        for _ in range(ml):
            out.append(out[-off])
    assert len(out) == size
    return bytes(out)
