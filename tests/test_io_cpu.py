"""CPU: HEALPix map files and NEST <-> RING (the data formats either side of
the likelihood; no GPU compute).

- tests/golden/mask1.fits is the reference's own test mask
  (slow_test_files/mask1.fits, nside 32, NESTED), read by the library and by an
  independent NumPy parse of its bytes;
- nest2ring / ring2nest: bijection and inverse pair at several nside, identity
  at nside 1; geometry fixes the numbering: the four NEST children of a pixel
  surround it, and the reference mask (a galactic-plane cut) keeps its shape
  only under the right permutation;
- FITS write -> read round trips (RING and NESTED), float32 / multi-column
  tables written by hand, and error paths.
"""
import math
import os

import numpy as np
import pytest

import cosmopp_amd as cm
import oracle

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
MASK = os.path.join(ROOT, "tests", "golden", "mask1.fits")


def numpy_read_single_column(path):
    """Independent reader for a 2-HDU file: primary header, BINTABLE header,
    then rows of '<repeat>D' big-endian doubles."""
    raw = open(path, "rb").read()
    pos, cards = 0, []
    for hdu in range(2):
        while True:
            blk = raw[pos:pos + 2880].decode("ascii")
            pos += 2880
            cs = [blk[i:i + 80] for i in range(0, 2880, 80)]
            cards += cs
            if any(c.startswith("END ") or c.rstrip() == "END" for c in cs):
                break
    kv = {}
    for c in cards:
        if len(c) > 9 and c[8] == "=":
            kv[c[:8].strip()] = c[10:].split("/")[0].strip().strip("'").strip()
    rows, width = int(kv["NAXIS2"]), int(kv["NAXIS1"])
    data = np.frombuffer(raw[pos:pos + rows * width], dtype=">f8").astype(np.float64)
    return data, kv


def test_reference_mask_reads_like_an_independent_parse():
    m, nside, nested = cm.read_map(MASK)
    ref, kv = numpy_read_single_column(MASK)
    assert (nside, nested) == (32, True) and kv["ORDERING"] == "NESTED" and int(kv["NSIDE"]) == 32
    assert np.array_equal(m, ref)
    assert set(np.unique(m)) == {0.0, 1.0}
    assert int(m.sum()) == 9096


@pytest.mark.parametrize("nside", [1, 2, 4, 8, 32, 128])
def test_nest_ring_bijection(nside):
    npix = 12 * nside * nside
    r = np.array([cm.nest2ring(nside, i) for i in range(npix)])
    assert np.array_equal(np.sort(r), np.arange(npix))
    assert all(cm.ring2nest(nside, int(r[i])) == i for i in range(npix))
    if nside == 1:
        assert np.array_equal(r, np.arange(12))  # base pixels carry the same numbers


def unit(theta, phi):
    return np.array([math.sin(theta) * math.cos(phi), math.sin(theta) * math.sin(phi), math.cos(theta)])


@pytest.mark.parametrize("nside", [4, 16, 64])
def test_nest_children_surround_their_parent(nside):
    """NEST is hierarchical: pixels 4n..4n+3 at 2 nside tile pixel n.  Their
    centres (through RING pix2ang, pinned separately) lie within one parent
    pixel size of the parent's centre and their mean is close to it."""
    rng = np.random.default_rng(nside)
    size = math.sqrt(4 * math.pi / (12 * nside * nside))
    for n in rng.integers(0, 12 * nside * nside, 200):
        c = unit(*oracle.pix2ang_ring(nside, cm.nest2ring(nside, int(n))))
        kids = [unit(*oracle.pix2ang_ring(2 * nside, cm.nest2ring(2 * nside, 4 * int(n) + k))) for k in range(4)]
        assert max(np.linalg.norm(k - c) for k in kids) < 0.75 * size
        mean = np.mean(kids, axis=0)
        assert np.linalg.norm(mean / np.linalg.norm(mean) - c) < 0.1 * size


@pytest.mark.parametrize("nside", [2, 4, 32])
def test_nest_face_axes_point_north_east_and_north_west(nside):
    """Handedness of the NEST numbering (Gorski et al. 2005, sect. 4): inside a
    face, bit 0 of the index steps x towards the north-east and bit 1 steps y
    towards the north-west.  (Ring membership is symmetric in x and y, so the
    tests above would not see a mirrored numbering.)"""
    fp = nside * nside
    for f in range(12):
        z, phi = [], []
        for k in (0, 1, 2):  # (x, y) = (0, 0), (1, 0), (0, 1)
            t, ph = oracle.pix2ang_ring(nside, cm.nest2ring(nside, f * fp + k))
            z.append(math.cos(t))
            phi.append(ph)
        d1 = (phi[1] - phi[0] + math.pi) % (2 * math.pi) - math.pi
        d2 = (phi[2] - phi[0] + math.pi) % (2 * math.pi) - math.pi
        assert z[1] > z[0] and z[2] > z[0]
        assert d1 > 0 > d2


def test_reference_mask_keeps_its_shape_in_ring_order():
    m, nside, _ = cm.read_map(MASK)
    perm = np.array([cm.ring2nest(nside, r) for r in range(m.size)])
    ring = m[perm]
    theta = np.array([oracle.pix2ang_ring(nside, p)[0] for p in range(m.size)])
    lat = 90.0 - np.degrees(theta)
    assert ring[np.abs(lat) < 15].mean() < 0.3       # galactic-plane cut
    assert ring[np.abs(lat) > 45].mean() > 0.99
    same = np.mean(ring[:-1] == ring[1:])           # neighbours along rings agree
    assert same > 0.95                               # a scrambled order gives ~0.62


@pytest.mark.parametrize("nside,nested", [(1, False), (2, True), (16, False), (64, True)])
def test_fits_round_trip(tmp_path, nside, nested):
    rng = np.random.default_rng(nside)
    m = rng.standard_normal(12 * nside * nside)
    m[::7] = -1.6375e30  # HEALPix UNSEEN values pass through unchanged
    path = str(tmp_path / "m.fits")
    cm.write_map(path, m, nside, nested)
    back, ns, nest = cm.read_map(path)
    assert (ns, nest) == (nside, nested) and np.array_equal(back, m)
    ref, kv = numpy_read_single_column(path)
    assert np.array_equal(ref, m) and kv["PIXTYPE"] == "HEALPIX"
    assert os.path.getsize(path) % 2880 == 0
    with pytest.raises(cm.CmblError):
        cm.write_map(path, m, nside, nested, overwrite=False)


def card(key, value, comment=""):
    s = "%-8s= %20s" % (key, value) if not value.startswith("'") else "%-8s= %s" % (key, value)
    if comment:
        s += " / " + comment
    return s.ljust(80)


def write_table(path, header_cards, payload):
    prim = "".join([card("SIMPLE", "T"), card("BITPIX", "8"), card("NAXIS", "0"), card("EXTEND", "T"), "END".ljust(80)])
    prim = prim.ljust((len(prim) + 2879) // 2880 * 2880)
    ext = "".join(header_cards) + "END".ljust(80)
    ext = ext.ljust((len(ext) + 2879) // 2880 * 2880)
    pad = (-len(payload)) % 2880
    open(path, "wb").write(prim.encode() + ext.encode() + payload + b"\0" * pad)


def test_fits_float32_three_columns_with_scaling(tmp_path):
    """Polarisation-style table: I, Q, U as '1024E' columns; column 2 scaled."""
    nside = 16
    npix = 12 * nside * nside
    rng = np.random.default_rng(3)
    cols = [rng.standard_normal(npix).astype(">f4") for _ in range(3)]
    rows = npix // 1024
    payload = b"".join(b"".join(c[r * 1024:(r + 1) * 1024].tobytes() for c in cols) for r in range(rows))
    cards = [card("XTENSION", "'BINTABLE'"), card("BITPIX", "8"), card("NAXIS", "2"),
             card("NAXIS1", str(3 * 4096)), card("NAXIS2", str(rows)), card("PCOUNT", "0"), card("GCOUNT", "1"),
             card("TFIELDS", "3"), card("TFORM1", "'1024E   '"), card("TFORM2", "'1024E   '"),
             card("TFORM3", "'1024E   '"), card("TSCAL2", "2.0D0"), card("TZERO2", "0.5"),
             card("PIXTYPE", "'HEALPIX '"), card("ORDERING", "'RING    '"), card("NSIDE", str(nside)),
             card("INDXSCHM", "'IMPLICIT'")]
    path = str(tmp_path / "iqu.fits")
    write_table(path, cards, payload)
    for c in (1, 3):
        m, ns, nested = cm.read_map(path, column=c)
        assert ns == nside and not nested
        assert np.array_equal(m, cols[c - 1].astype(np.float64))
    m2, _, _ = cm.read_map(path, column=2)
    assert np.array_equal(m2, cols[1].astype(np.float64) * 2.0 + 0.5)
    with pytest.raises(cm.CmblError):
        cm.read_map(path, column=4)


def test_fits_errors(tmp_path):
    with pytest.raises(cm.CmblError):
        cm.read_map(str(tmp_path / "missing.fits"))
    bad = tmp_path / "bad.fits"
    bad.write_bytes(b"not a fits file".ljust(2880))
    with pytest.raises(cm.CmblError):
        cm.read_map(str(bad))
    with pytest.raises(cm.CmblError):
        cm.nest2ring(3, 0)
    with pytest.raises(cm.CmblError):
        cm.ring2nest(4, 12 * 16)
