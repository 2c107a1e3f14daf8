"""Host asset readers of the scene loader (orf_hdr_read, orf_obj_read), checked
without a GPU against inputs written here from the file formats' published
descriptions (tests/rgbe.py for Radiance RGBE; OBJ text below):

- RGBE: flat and run-length-encoded scanlines decode to (m + 1/2) 2^(e - 136)
  (0 for e = 0), divided by the exposure of the last header line holding
  EXPOSURE=, with row 0 = the file's bottom row (the row flip of
  loadHDRTexture); malformed files raise OrfError.
- OBJ: v / vn / f in the v, v/t, v//n and v/t/n forms, polygons as fans,
  negative (relative) indices, comments; nindex is dropped when a face has no
  normals; out-of-range indices raise OrfError.
- The reference's own assets, where its tree is present (this container):
  data/CedarCity.hdr (1600 x 800, run-length encoded) decodes pixel for pixel as
  tests/rgbe.py's reader + decoder read it, and data/Plane.obj (one quad, v/vt/vn
  faces, normals +y) loads as two fan triangles of the unit square at y = 0.
"""
import numpy as np
import pytest

from open_rendering_framework_amd import rt
from tests import rgbe


def _image(h, w, seed):
    rng = np.random.default_rng(seed)
    img = rng.lognormal(0.0, 3.0, size=(h, w, 3))
    img[h // 3, :, :] = 0.0          # black row (e = 0)
    img[:, w // 2, :] = 7.25         # constant column (runs)
    return img


@pytest.mark.parametrize("rle,w", [(True, 37), (False, 37), (True, 5)])
def test_hdr_decode(tmp_path, rle, w):
    h = 11
    enc = rgbe.to_rgbe(_image(h, w, 3))
    p = tmp_path / "img.hdr"
    rgbe.write(p, enc, rle=rle)
    got = rt.hdr_read(p)
    assert got.shape == (h, w, 4)
    want = rgbe.decode(enc)[::-1]
    np.testing.assert_array_equal(got[..., :3], want)
    assert np.all(got[..., 3] == 1.0)


def test_hdr_exposure_and_quantisation(tmp_path):
    img = _image(6, 16, 5)
    enc = rgbe.to_rgbe(img)
    p = tmp_path / "e.hdr"
    rgbe.write(p, enc, exposure=2.0)
    got = rt.hdr_read(p)[::-1, :, :3]
    np.testing.assert_allclose(got, rgbe.decode(enc) / 2.0, rtol=1e-7)
    # RGBE keeps 8 mantissa bits of the largest channel: relative error < 2^-8 there
    mx = img.max(axis=-1)
    rel = np.abs(got.max(axis=-1) * 2.0 - mx) / np.maximum(mx, 1e-30)
    assert np.all(rel[mx > 0] < 2.0 ** -7)


def test_hdr_last_exposure_line_wins(tmp_path):
    """Two EXPOSURE lines: the later sets the exposure (ORF/HDRLoader.cpp:157-160)."""
    enc = rgbe.to_rgbe(_image(4, 8, 9))
    p = tmp_path / "x.hdr"
    rgbe.write(p, enc, exposure=2.0)
    data = p.read_bytes().replace(b"EXPOSURE=2\n", b"EXPOSURE=2\nPRIMARIES EXPOSURE=4\n")
    p.write_bytes(data)
    np.testing.assert_allclose(rt.hdr_read(p)[::-1, :, :3], rgbe.decode(enc) / 4.0, rtol=1e-7)


def test_hdr_rejects_bad_files(tmp_path):
    bad = tmp_path / "bad.hdr"
    bad.write_bytes(b"P6\n1 1\n255\n\x00\x00\x00")
    with pytest.raises(rt.OrfError):
        rt.hdr_read(bad)
    flipped = tmp_path / "flipped.hdr"
    flipped.write_bytes(b"#?RADIANCE\nFORMAT=32-bit_rle_rgbe\n\n+Y 1 +X 1\n\x80\x80\x80\x81")
    with pytest.raises(rt.OrfError):
        rt.hdr_read(flipped)
    short = tmp_path / "short.hdr"
    short.write_bytes(b"#?RADIANCE\nFORMAT=32-bit_rle_rgbe\n\n-Y 2 +X 2\n\x80\x80")
    with pytest.raises(rt.OrfError):
        rt.hdr_read(short)
    with pytest.raises(rt.OrfError):
        rt.hdr_read(tmp_path / "missing.hdr")


OBJ = """# a quad, a triangle with relative indices, texture coordinates ignored
v 0 0 0
v 1 0 0
v 1 1 0
v 0 1 0
vt 0 0
vn 0 0 1
vn 0 0 -1
f 1/1/1 2/1/1 3/1/1 4/1/1
v 2 0 0
f -1//2 -4//2 -3//2
"""


def test_obj_read(tmp_path):
    p = tmp_path / "m.obj"
    p.write_text(OBJ)
    o = rt.obj_read(p)
    np.testing.assert_array_equal(o["verts"], [[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0], [2, 0, 0]])
    np.testing.assert_array_equal(o["tris"], [[0, 1, 2], [0, 2, 3], [4, 1, 2]])
    np.testing.assert_array_equal(o["normals"], [[0, 0, 1], [0, 0, -1]])
    np.testing.assert_array_equal(o["nindex"], [[0, 0, 0], [0, 0, 0], [1, 1, 1]])


def test_obj_without_normals_and_errors(tmp_path):
    p = tmp_path / "n.obj"
    p.write_text(OBJ + "f 1 2 5\n")
    assert rt.obj_read(p)["nindex"] is None
    bad = tmp_path / "bad.obj"
    bad.write_text("v 0 0 0\nv 1 0 0\nf 1 2 3\n")
    with pytest.raises(rt.OrfError):
        rt.obj_read(bad)


REF_DATA = "/root/reference/data"


@pytest.mark.skipif(not __import__("os").path.exists(REF_DATA + "/CedarCity.hdr"), reason="reference tree absent")
def test_reference_cedarcity_hdr():
    enc, header = rgbe.read(REF_DATA + "/CedarCity.hdr")
    assert enc.shape == (800, 1600, 4)
    assert any(h.startswith("FORMAT=32-bit_rle_rgbe") for h in header)
    exposure = [float(h.split("=", 1)[1]) for h in header if "EXPOSURE=" in h][-1]
    got = rt.hdr_read(REF_DATA + "/CedarCity.hdr")
    assert got.shape == (800, 1600, 4)
    want = rgbe.decode(enc)[::-1] / np.float32(exposure)
    np.testing.assert_array_equal(got[..., :3], want)
    assert np.all(got[..., 3] == 1.0)
    assert np.isfinite(got).all() and got[..., :3].max() > 1.0  # a real HDR sky


@pytest.mark.skipif(not __import__("os").path.exists(REF_DATA + "/Plane.obj"), reason="reference tree absent")
def test_reference_plane_obj():
    o = rt.obj_read(REF_DATA + "/Plane.obj")
    np.testing.assert_array_equal(o["verts"], [[-1, 0, 1], [1, 0, 1], [-1, 0, -1], [1, 0, -1]])
    # f 1/1/1 2/2/2 4/4/3 3/3/4 as a fan from its first corner
    np.testing.assert_array_equal(o["tris"], [[0, 1, 3], [0, 3, 2]])
    np.testing.assert_array_equal(o["normals"], [[0, 1, 0]] * 4)
    np.testing.assert_array_equal(o["nindex"], [[0, 1, 2], [0, 2, 3]])
    v = o["verts"].astype(np.float64)
    for t in o["tris"]:
        n = np.cross(v[t[1]] - v[t[0]], v[t[2]] - v[t[0]])
        assert n[1] > 0  # the file's corner order faces +y, the side of its vn normals
