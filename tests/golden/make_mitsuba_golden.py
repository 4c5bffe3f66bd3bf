"""Writes tests/golden/mitsuba_golden.npz from the reference's golden-image test data
(tests/data/golden/<scene>.photon, <scene>_mitsuba.exr, and the environment map the HDRI scene names;
test_golden_reference.cpp): per scene the scene document (the keys the renderer reads, as JSON text) and the
Mitsuba 3 image as float32 RGB, plus the environment map as its raw Radiance RGBE bytes.

The reference tree exists only where the fixtures are made; the tests read the .npz. Run from the repository
root:  python tests/golden/make_mitsuba_golden.py
"""
from __future__ import annotations

import json
import struct
import zlib
from pathlib import Path

import numpy as np

REF = Path("/root/reference")
GOLDEN = REF / "tests/data/golden"
SCENES = ("cornell_diffuse", "cornell_specular", "hdri_materials")
OUT = Path(__file__).resolve().parent / "mitsuba_golden.npz"


def read_exr(path: Path) -> np.ndarray:
    """Scanline OpenEXR with FLOAT channels and ZIP (16-line) or no compression, as the golden images are
    (OpenEXR file layout: magic, version, attribute list, line-offset table, chunks of y, size, data; ZIP
    chunks hold the zlib stream of the byte-delta-coded, even/odd-split block)."""
    b = path.read_bytes()
    if b[:4] != b"\x76\x2f\x31\x01":
        raise ValueError("not an OpenEXR file")
    p, attrs = 8, {}
    while b[p] != 0:
        e = b.index(b"\0", p)
        name = b[p:e].decode()
        p = e + 1
        e = b.index(b"\0", p)
        p = e + 1
        size = struct.unpack_from("<i", b, p)[0]
        p += 4
        attrs[name] = b[p:p + size]
        p += size
    p += 1
    chans, q = [], 0
    cl = attrs["channels"]
    while cl[q] != 0:
        e = cl.index(b"\0", q)
        name = cl[q:e].decode()
        ptype = struct.unpack_from("<i", cl, e + 1)[0]
        if ptype != 2:
            raise ValueError("only FLOAT channels are read")
        chans.append(name)
        q = e + 1 + 16
    x0, y0, x1, y1 = struct.unpack("<4i", attrs["dataWindow"])
    w, h = x1 - x0 + 1, y1 - y0 + 1
    comp = attrs["compression"][0]
    lines = {0: 1, 3: 16}[comp]
    nchunks = (h + lines - 1) // lines
    offsets = struct.unpack_from("<%dQ" % nchunks, b, p)
    img = np.zeros((h, len(chans), w), np.float32)
    for off in offsets:
        y, size = struct.unpack_from("<ii", b, off)
        data = b[off + 8:off + 8 + size]
        n = min(lines, y1 + 1 - y)
        raw = n * len(chans) * w * 4
        if comp == 3 and size < raw:
            d = np.frombuffer(zlib.decompress(data), np.uint8).astype(np.int64)
            d[1:] -= 128  # undo the byte predictor: t[i] = t[i-1] + d[i] - 128 (mod 256)
            t = (np.cumsum(d) % 256).astype(np.uint8)
            half = (t.size + 1) // 2
            u = np.empty_like(t)
            u[0::2] = t[:half]
            u[1::2] = t[half:]
            data = u.tobytes()
        arr = np.frombuffer(data, "<f4").reshape(n, len(chans), w)
        img[y - y0:y - y0 + n] = arr
    rgb = [chans.index(c) for c in "RGB"]
    return np.ascontiguousarray(img[:, rgb, :].transpose(0, 2, 1))


def read_rgbe(path: Path) -> np.ndarray:
    """Radiance .hdr (-Y H +X W) with new-style run-length scanlines or flat pixels -> (H, W, 4) uint8 RGBE."""
    b = path.read_bytes()
    p = 0
    while True:
        e = b.index(b"\n", p)
        line = b[p:e]
        p = e + 1
        if line.startswith(b"-Y"):
            _, hs, _, ws = line.split()
            h, w = int(hs), int(ws)
            break
    out = np.zeros((h, w, 4), np.uint8)
    for y in range(h):
        if b[p] == 2 and b[p + 1] == 2 and (b[p + 2] << 8 | b[p + 3]) == w:
            p += 4
            for c in range(4):
                x = 0
                while x < w:
                    count = b[p]
                    p += 1
                    if count > 128:
                        out[y, x:x + count - 128, c] = b[p]
                        p += 1
                        x += count - 128
                    else:
                        out[y, x:x + count, c] = np.frombuffer(b, np.uint8, count, p)
                        p += count
                        x += count
        else:
            out[y] = np.frombuffer(b, np.uint8, 4 * w, p).reshape(w, 4)
            p += 4 * w
    return out


def main():
    arrays = {}
    for name in SCENES:
        doc = json.loads((GOLDEN / f"{name}.photon").read_text())
        keep = {k: doc[k] for k in ("camera", "environment", "lights", "materials", "nodes", "render")}
        hdr = keep["environment"].get("hdr", "")
        if hdr:
            keep["environment"]["hdr"] = "env_rgbe"
            arrays["env_rgbe"] = read_rgbe((GOLDEN / hdr).resolve())
        arrays[f"{name}_doc"] = np.frombuffer(json.dumps(keep).encode(), np.uint8)
        arrays[f"{name}_image"] = read_exr(GOLDEN / f"{name}_mitsuba.exr")
    np.savez_compressed(OUT, **arrays)
    for k, v in arrays.items():
        print(k, v.shape, v.dtype)


if __name__ == "__main__":
    main()
