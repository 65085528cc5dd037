"""Extract the reference's own test fixtures into compressed numpy files.

Reads the data files the reference's tests hold (test/files/Apr/<case>/*.tif
and the parameter attributes of <case>/sphere.apr) and writes
tests/golden/<case>.npz.  Run in the container that has the reference tree:

    python tests/golden/make_fixtures.py [/root/reference/test/files/Apr]

Only data is stored (input image, saved gradient / local intensity scale, and
the per-pixel particle level / x / y / z / intensity images the reference
uses to check get_apr, APRTest.cpp:2425-2490, 2771-2841).
"""
import os
import struct
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def read_tiff(path):
    """Minimal reader for the uncompressed multi-page grayscale TIFFs of the
    reference fixtures.  Returns an array [pages, rows, cols]."""
    d = open(path, "rb").read()
    bo = {b"II": "<", b"MM": ">"}[d[:2]]
    assert struct.unpack(bo + "H", d[2:4])[0] == 42
    off = struct.unpack(bo + "I", d[4:8])[0]
    tsz = {1: 1, 2: 1, 3: 2, 4: 4, 5: 8, 6: 1, 7: 1, 8: 2, 9: 4, 10: 8, 11: 4, 12: 8}
    pages = []
    while off:
        n = struct.unpack(bo + "H", d[off:off + 2])[0]
        tags = {}
        for i in range(n):
            e = d[off + 2 + 12 * i: off + 14 + 12 * i]
            tag, typ, cnt = struct.unpack(bo + "HHI", e[:8])
            size = tsz[typ] * cnt
            raw = e[8:12] if size <= 4 else d[struct.unpack(bo + "I", e[8:12])[0]:][:size]
            fmt = {1: "B", 3: "H", 4: "I", 8: "h", 9: "i", 11: "f", 12: "d", 16: "Q"}.get(typ, "B")
            tags[tag] = struct.unpack(bo + fmt * cnt, raw[:tsz[typ] * cnt])
        w, h = tags[256][0], tags[257][0]
        bits = tags[258][0]
        assert tags.get(259, (1,))[0] == 1, "compressed TIFF not supported"
        sf = tags.get(339, (1,))[0]
        dt = {(16, 1): "u2", (8, 1): "u1", (32, 3): "f4", (32, 1): "u4", (16, 2): "i2"}[(bits, sf)]
        buf = b"".join(d[o:o + c] for o, c in zip(tags[273], tags[279]))
        pages.append(np.frombuffer(buf, np.dtype(bo + dt), count=w * h).reshape(h, w))
        off = struct.unpack(bo + "I", d[off + 2 + 12 * n: off + 6 + 12 * n])[0]
    return np.stack(pages)


def apr_attributes(path, names):
    """Scalar attributes of an .apr (HDF5) file: version-1 attribute messages
    (name, datatype, dataspace, data; each 8-byte padded)."""
    d = open(path, "rb").read()
    out = {}
    for name in names:
        i = d.find(name.encode() + b"\x00")
        ver, _, nsz, dtsz, dssz = struct.unpack("<BBHHH", d[i - 8:i])
        assert ver == 1
        pad = lambda n: (n + 7) // 8 * 8
        dt = d[i + pad(nsz): i + pad(nsz) + dtsz]
        cls = dt[0] & 0x0F
        size = struct.unpack("<I", dt[4:8])[0]
        o = i + pad(nsz) + pad(dtsz) + pad(dssz)
        if cls == 1:
            out[name] = struct.unpack("<f" if size == 4 else "<d", d[o:o + size])[0]
        else:
            out[name] = int.from_bytes(d[o:o + 4], "little")
    return out


def as_zxy(tiff_pages):
    """TIFF pages are z; each page is rows=y? The reference reads page z, row y
    is stored along x... TiffUtils stores x as rows, y as columns (y fastest)."""
    return np.ascontiguousarray(tiff_pages)


def main(base):
    cases = {"sphere_120": "sphere", "sphere_diff_dims": "sphere"}
    for case, stem in cases.items():
        src = os.path.join(base, case)
        arrs = {}
        for key, fn in [("original", f"{stem}_original.tif"), ("gradient_saved", "gradient_saved.tif"),
                        ("scale_saved", "scale_saved.tif"), ("level", f"{stem}_level.tif"), ("x", f"{stem}_x.tif"),
                        ("y", f"{stem}_y.tif"), ("z", f"{stem}_z.tif"), ("pc", f"{stem}_pc.tif")]:
            arrs[key] = as_zxy(read_tiff(os.path.join(src, fn)))
        attrs = apr_attributes(os.path.join(src, f"{stem}.apr"),
                               ["lambda", "sigma_th", "sigma_th_max", "I_th", "rel_error", "grad_th", "psfx", "psfy",
                                "psfz", "total_number_particles", "level_max", "level_min"])
        for k, v in attrs.items():
            arrs["attr_" + k] = np.array(v)
        out = os.path.join(HERE, f"{case}.npz")
        np.savez_compressed(out, **arrs)
        print(case, {k: v.shape for k, v in arrs.items() if v.ndim}, attrs, os.path.getsize(out))


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference/test/files/Apr")
