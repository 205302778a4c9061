"""Host-side exports of libsdtfe_amd.so that carry the reference's
prototypes and touch no GPU: convert_to_qhdata, rotate3d, shuffle_buff
(src/utils.c:8-79), the bitmap (src/bitmap.c:8-31) and the map's file output
(src/io.c:13-95)."""
import ctypes as C
import math
import struct

import numpy as np
import pytest

from sdtfe_amd import capi


def test_convert_and_rotate_match_reference(golden):
    """The C exports give the reference's own outputs (G5, written by the
    compiled reference) bit for bit."""
    g = golden("g5_utils.npz")
    pos32 = np.asarray(g["pos32"], dtype=np.float32)
    qh = capi.convert_to_qhdata(pos32.T.ravel())  # planar: x..., y..., z...
    np.testing.assert_array_equal(qh[:, :3], g["qh"])
    assert (qh[:, 3] == 0).all()
    capi.rotate3d(qh, qh.shape[0], g["theta"], g["center"])
    np.testing.assert_array_equal(qh[:, :3], g["rotated"])


def test_rotate_partial_rows_and_identity():
    d = np.random.default_rng(1).random((20, 4))
    ref = d.copy()
    capi.rotate3d(d, 5, [0.0, 0.0, 0.0], [0.5, 0.5, 0.5])
    np.testing.assert_allclose(d, ref, rtol=0, atol=1e-15)
    capi.rotate3d(d, 5, [0.0, 0.0, math.pi / 2], [0.0, 0.0, 0.0])
    # a quarter turn about z: (x, y) -> (-y, x); rows 5.. untouched
    np.testing.assert_allclose(d[:5, 0], -ref[:5, 1], atol=1e-15)
    np.testing.assert_allclose(d[:5, 1], ref[:5, 0], atol=1e-15)
    np.testing.assert_array_equal(d[5:], ref[5:])


def test_shuffle_follows_libc_rand():
    """shuffle_buff swaps row i-1 with row rand() % (n - i) + i for
    i = 1 .. k-1 (src/utils.c:18-30), with the C library's rand()."""
    libc = C.CDLL(None)
    libc.rand.restype = C.c_int
    n, k, seed = 50, 30, 1234
    libc.srand(seed)
    draws = [libc.rand() for _ in range(k - 1)]
    want = np.arange(4 * n, dtype=np.float64).reshape(n, 4)
    for i in range(1, k):
        r = draws[i - 1] % (n - i) + i
        want[[i - 1, r]] = want[[r, i - 1]]
    got = np.arange(4 * n, dtype=np.float64).reshape(n, 4)
    libc.srand(seed)
    capi.shuffle_buff(got, n, k)
    np.testing.assert_array_equal(got, want)


def test_bitmap_layout():
    """Four bits per 32-bit word, as the reference's BITS_PER_WORD gives:
    bit 4 t + f is bit f of word t."""
    L = capi.lib()
    ptr = L.create_bitmap(4 * 10)
    words = np.ctypeslib.as_array(C.cast(ptr, C.POINTER(C.c_uint32)), shape=(11,))
    assert (words == 0).all()
    bm = np.zeros(11, dtype=np.uint32)
    for n in (0, 5, 6, 39):
        L.set_bit(bm, n)
    assert bm[0] == 1 and bm[1] == 0b0110 and bm[9] == 0b1000
    assert [L.get_bit(bm, n) for n in range(8)] == [1, 0, 0, 0, 0, 1, 1, 0]
    L.clear_bit(bm, 5)
    assert bm[1] == 0b0100
    L.delete_bitmap(ptr)


def test_write_rho_raw(tmp_path):
    rho = np.random.default_rng(2).random(64 * 64)
    path = tmp_path / "m.rho.bin"
    assert capi.write_rho(str(path), rho) == rho.size
    np.testing.assert_array_equal(np.fromfile(path, dtype=np.float64), rho)


def jet8(i):
    """The Jet palette from its definition: 64-step ramps of 255/64, each
    rounded to nearest."""
    def ramp(k):
        return int(math.floor(255 * k / 64 + 0.5))
    x = i
    r = 0 if x < 96 else ramp(x - 95) if x < 160 else 255 if x < 224 else ramp(287 - x)
    g = 0 if x < 32 else ramp(x - 31) if x < 96 else 255 if x < 160 else \
        ramp(223 - x) if x < 224 else 0
    b = ramp(x + 33) if x < 32 else 255 if x < 96 else ramp(159 - x) if x < 160 else 0
    return r, g, b


def read_tiff(path):
    """Minimal little-endian TIFF reader: tags and the strips' bytes."""
    data = open(path, "rb").read()
    assert data[:4] == b"II*\x00"
    (ifd,) = struct.unpack_from("<I", data, 4)
    (n,) = struct.unpack_from("<H", data, ifd)
    tags = {}
    for k in range(n):
        tag, typ, cnt, val = struct.unpack_from("<HHII", data, ifd + 2 + 12 * k)
        if typ == 3 and cnt == 1:
            val &= 0xFFFF
        if typ == 3 and cnt > 2:
            val = list(struct.unpack_from(f"<{cnt}H", data, val))
        tags[tag] = val
    assert list(tags) == sorted(tags)
    off, nbytes = tags[273], tags[279]
    return tags, np.frombuffer(data[off:off + nbytes], dtype=np.uint8)


def test_write_image_indices_and_palette(tmp_path):
    G = 32
    rho = np.random.default_rng(3).lognormal(0, 2, G * G)
    rho[::7] = 0.0
    path = tmp_path / "m.rho.tiff"
    capi.write_image(str(path), rho)
    tags, pix = read_tiff(path)
    assert tags[256] == G and tags[257] == G and tags[258] == 8 and tags[259] == 1
    assert tags[262] == 3 and tags[277] == 1 and tags[278] == G and tags[279] == G * G
    # indices: 128 (log(rho + 1) - log(min + 1)) / (log(max + 1) - log(min + 1)), truncated
    pos = rho[rho > 0]
    lo, hi = math.log(pos.min() + 1.0), math.log(rho.max() + 1.0)
    want = np.zeros(G * G, dtype=np.uint8)
    for i, v in enumerate(rho):
        if v > 0:
            want[i] = int(128.0 * (math.log(v + 1.0) - lo) / (hi - lo))
    np.testing.assert_array_equal(pix, want)
    assert pix.max() == 128 and pix[rho == 0].max() == 0
    cmap = np.array(tags[320]).reshape(3, 256)
    for i in range(256):
        assert tuple(cmap[:, i] // 257) == jet8(i) and (cmap[:, i] % 257 == 0).all()
    try:
        from PIL import Image
    except ImportError:
        return
    im = Image.open(path)
    assert im.mode == "P" and im.size == (G, G)
    np.testing.assert_array_equal(np.asarray(im).ravel(), want)
    pal = np.array(im.getpalette()[:3 * 256]).reshape(256, 3)
    assert [tuple(p) for p in pal[:129]] == [jet8(i) for i in range(129)]


def test_write_image_refuses_negative_and_nan(tmp_path, capfd):
    for bad in (-1.0, float("nan")):
        rho = np.ones(16)
        rho[3] = bad
        path = tmp_path / "bad.tiff"
        capi.write_image(str(path), rho)
        assert not path.exists()
    out = capfd.readouterr().out
    assert "Negative density" in out and "NaN density" in out


def test_write_image_constant_map(tmp_path):
    """One positive value everywhere: the reference divides 0 by 0; here
    every index is 0."""
    path = tmp_path / "c.tiff"
    capi.write_image(str(path), np.full(64, 2.5))
    _, pix = read_tiff(path)
    assert (pix == 0).all()


@pytest.mark.parametrize("n_tri", [3])
def test_copy_fh_triangles_rejects_overfull_bitmap(n_tri):
    """A bitmap that marks more faces than n_tri is refused on the host
    before any device work (the reference would write past its array)."""
    pos = np.random.default_rng(0).random((8, 4))
    tet = np.zeros((2, 8), dtype=np.int32)
    tet[:, 4:] = -1
    bm = np.array([0b1111, 0b0011, 0], dtype=np.uint32)
    with pytest.raises(capi.SdtfeError, match="more than n_tri"):
        capi.copy_fh_triangles(tet, pos, n_tri, bm)
