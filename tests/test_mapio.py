"""Map files and initial maps through the C ABI (host code, no GPU): the formats of the
reference's crates/stl-mpi/src/io.rs and the initial maps of crates/stl-cli/src/lib.rs.

Pinned by numpy's own `.npy` reader and writer, a FITS file laid out here by hand (primary
HDU with data to skip, a RING table of 64-pixel rows as healpy writes them) decoded through
the oracle's RING->NESTED map (itself pinned by the reference's ring2nest fixtures), the
published first output of SplitMix64 and the moments of the normal law; the reference's
own round-trip test (io.rs:523-569) is restated case by case.
"""
import math
import os

import numpy as np
import pytest

import oracle as O
from stl_port import maps as M
from stl_port import stats as S
from stl_port import synth as Y

from conftest import GOLDEN

VALUES = np.sin(np.arange(12 * 16, dtype=np.float64))
SINGLE = VALUES.astype(np.float32).astype(np.float64)


def test_fits_and_npy_roundtrip(tmp_path):
    # io.rs:523-569, case by case.
    p = tmp_path / "m.fits"
    M.write_map(p, VALUES, "VALUE", double=True)
    assert np.array_equal(M.read_map(p), VALUES)
    M.write_map(p, VALUES, "VALUE", double=False)
    assert np.array_equal(M.read_map(p), SINGLE)
    assert os.path.getsize(p) % 2880 == 0
    q = tmp_path / "m.npy"
    M.write_map(q, VALUES)
    assert np.array_equal(M.read_map(q), VALUES)
    assert np.array_equal(np.load(q), VALUES)
    be = tmp_path / "be.npy"
    np.save(be, VALUES.astype(">f4"))
    assert M.map_len(be) == VALUES.size
    assert np.array_equal(M.read_map_range(be, 5, 77), SINGLE[5:77])
    for path in (p, q):
        full = M.read_map(path)
        assert M.map_len(path) == full.size
        assert np.array_equal(M.read_map_range(path, 37, 101), full[37:101])
    for name, double in (("s.npy", False), ("s.fits", False), ("d.fits", True)):
        path = tmp_path / name
        out = M.SlicedMap(path, double)
        out.create(VALUES.size, "VALUE")
        for k in range(0, VALUES.size, 50):
            out.write(k, VALUES[k:k + 50])
        expect = SINGLE if name == "s.fits" else VALUES
        assert np.array_equal(M.read_map(path), expect), name
        if name.endswith(".npy"):
            assert np.array_equal(np.load(path), VALUES)


@pytest.mark.parametrize("dtype", ["<f8", ">f8", "<f4", ">f4", "<i8", ">i8", "<i4", ">i4"])
def test_npy_dtypes_as_numpy_writes_them(tmp_path, dtype):
    a = (VALUES * 1000).astype(dtype)
    path = tmp_path / "x.npy"
    np.save(path, a)
    assert np.array_equal(M.read_map(path), a.astype(np.float64))
    assert np.array_equal(M.read_map_range(path, 11, 150), a.astype(np.float64)[11:150])


def test_npy_shapes(tmp_path):
    a = np.arange(2 * 3 * 4, dtype=np.float64).reshape(2, 3, 4)
    path = tmp_path / "a.npy"
    M.write_npy(path, a)
    b = np.load(path)
    assert b.shape == a.shape and b.dtype == np.float64 and np.array_equal(a, b)
    c = M.read_npy(path)
    assert c.shape == a.shape and np.array_equal(a, c)
    np.save(path, a.astype(np.float32).T.copy())
    assert np.array_equal(M.read_npy(path), a.T.astype(np.float32).astype(np.float64))
    # The header is padded so that the data starts on a 64-byte boundary (npy format 1.0).
    with open(path, "rb") as f:
        head = f.read(10)
    assert head[:8] == b"\x93NUMPY\x01\x00"


def test_npy_fortran_order_is_refused(tmp_path):
    path = tmp_path / "f.npy"
    np.save(path, np.asfortranarray(np.arange(12.0).reshape(3, 4)))
    with pytest.raises(RuntimeError, match="Fortran"):
        M.read_map(path)


def _card(key, value=None):
    c = key.ljust(8)
    if value is not None:
        c += "= " + value
    return c.ljust(80).encode()


def _block(cards):
    b = b"".join(cards)
    return b + b" " * (-len(b) % 2880)


def _ring_fits(path, values_ring, nside, repeat=64, ordering="RING"):
    """A HEALPix table as healpy lays it out: rows of `repeat` float32 values, comments after
    '/', and a primary HDU that holds 100 bytes of data to skip."""
    n = values_ring.size
    primary = _block([_card("SIMPLE", "T".rjust(20)), _card("BITPIX", "8".rjust(20)),
                      _card("NAXIS", "1".rjust(20)), _card("NAXIS1", "100".rjust(20) + " / bytes"),
                      _card("END")])
    pdata = bytes(range(100)) + b"\0" * (2880 - 100)
    table = _block([_card("XTENSION", "'BINTABLE'           / binary table extension"),
                    _card("BITPIX", "8".rjust(20)), _card("NAXIS", "2".rjust(20)),
                    _card("NAXIS1", str(4 * repeat).rjust(20)), _card("NAXIS2", str(n // repeat).rjust(20)),
                    _card("PCOUNT", "0".rjust(20)), _card("GCOUNT", "1".rjust(20)),
                    _card("TFIELDS", "1".rjust(20)), _card("TTYPE1", "'TEMPERATURE'"),
                    _card("TFORM1", f"'{repeat}E'".ljust(20) + " / data format of field"),
                    _card("PIXTYPE", "'HEALPIX '"), _card("ORDERING", f"'{ordering}'".ljust(10)),
                    _card("NSIDE", str(nside).rjust(20)), _card("END")])
    data = values_ring.astype(">f4").tobytes()
    data += b"\0" * (-len(data) % 2880)
    path.write_bytes(primary + pdata + table + data)


@pytest.mark.parametrize("nside", [4, 16])
def test_ring_fits_is_reordered_to_nested(tmp_path, nside):
    order = int(np.log2(nside))
    n = 12 * nside * nside
    ring = np.arange(n, dtype=np.float64) * 0.5 - 7.0
    path = tmp_path / "ring.fits"
    _ring_fits(path, ring, nside)
    nest = M.read_map(path)
    expect = np.empty(n)
    for p in range(n):
        expect[O.ring2nest(p, order)] = ring[p]
    assert np.array_equal(nest, expect)
    assert M.map_len(path) == n
    assert np.array_equal(M.read_map_range(path, 13, 99), expect[13:99])
    # The same layout marked NESTED is taken as is, read by ranges across rows.
    _ring_fits(path, ring, nside, ordering="NESTED")
    assert np.array_equal(M.read_map(path), ring)
    assert np.array_equal(M.read_map_range(path, 63, 130), ring[63:130])


def test_fits_header_cards(tmp_path):
    path = tmp_path / "h.fits"
    M.write_map(path, VALUES, "SIGNAL", double=False)
    raw = path.read_bytes()
    assert len(raw) == 3 * 2880
    header = raw[2880:5760].decode("ascii")
    cards = {header[i:i + 8].strip(): header[i + 10:i + 80].strip() for i in range(0, 2880, 80)
             if header[i + 8:i + 10] == "= "}
    assert cards["XTENSION"] == "'BINTABLE'"
    assert cards["TTYPE1"] == "'SIGNAL  '"
    assert cards["TFORM1"] == "'E       '"
    assert cards["ORDERING"] == "'NESTED  '"
    assert cards["NSIDE"] == "4" and cards["NAXIS2"] == "192" and cards["LASTPIX"] == "191"
    assert np.array_equal(np.frombuffer(raw[5760:5760 + 4 * 192], dtype=">f4"), VALUES.astype(np.float32))
    # The primary header: SIMPLE, BITPIX, NAXIS 0, EXTEND, END.
    assert raw[:80].decode().startswith("SIMPLE  =                    T")


def test_errors_name_the_file(tmp_path):
    with pytest.raises(RuntimeError, match="missing.npy"):
        M.read_map(tmp_path / "missing.npy")
    with pytest.raises(RuntimeError, match="not a full-sky"):
        M.write_map(tmp_path / "x.fits", np.zeros(100))
    with pytest.raises(ValueError, match="not a full-sky"):
        M.nside_of(12 * 12 * 12)
    path = tmp_path / "m.npy"
    M.write_map(path, VALUES)
    with pytest.raises(RuntimeError, match="beyond"):
        M.read_map_range(path, 100, 300)
    bad = tmp_path / "bad.fits"
    _ring_fits(bad, np.zeros(192), 4)
    raw = bytearray(bad.read_bytes())
    at = raw.find(b"'64E'")
    raw[at:at + 5] = b"'64J'"
    bad.write_bytes(bytes(raw))
    with pytest.raises(RuntimeError, match="not a float map"):
        M.read_map(bad)


@pytest.mark.parametrize("name, double", [("s.npy", True), ("s.fits", False), ("s.fits", True)])
def test_sliced_writes_stay_inside_the_map(tmp_path, name, double):
    path = tmp_path / name
    out = M.SlicedMap(path, double)
    out.create(192, "VALUE")
    size = path.stat().st_size
    out.write(100, VALUES[100:192])
    with pytest.raises(RuntimeError, match="beyond the map"):
        out.write(100, VALUES[:93])
    with pytest.raises(RuntimeError, match="beyond the map"):
        out.write(193, VALUES[:0])
    assert path.stat().st_size == size
    out.write(0, VALUES[:100])
    expect = VALUES if double else SINGLE
    np.testing.assert_array_equal(M.read_map(path), expect)


def test_npy_shape_beyond_the_file_is_refused(tmp_path):
    # a header whose shape announces more data than the file holds (or overflows the
    # element count) is refused before anything is read
    for shape in ("(1000,)", "(4294967296, 4294967296, 16)"):
        path = tmp_path / "h.npy"
        header = "{'descr': '<f8', 'fortran_order': False, 'shape': " + shape + ", }"
        header += " " * ((64 - (10 + len(header) + 1) % 64) % 64) + "\n"
        path.write_bytes(b"\x93NUMPY\x01\x00" + len(header).to_bytes(2, "little") + header.encode() + bytes(64))
        with pytest.raises(RuntimeError, match="exceeds the file|overflows"):
            M.read_npy(path)
        with pytest.raises(RuntimeError, match="exceeds the file|overflows"):
            M.map_len(path)


def test_takahashi_fixture_reads_as_numpy_does():
    path = GOLDEN / "takahashi" / "takahashi_r000_zs1_nside64_nested.npy"
    a = np.load(path)
    assert np.array_equal(M.read_map(path), a)
    assert M.map_len(path) == 12 * 64 * 64


# ---------------------------------------------------------------------------
# Initial maps (lib.rs:523-540, 646-760)
# ---------------------------------------------------------------------------

MASK = (1 << 64) - 1


def splitmix64(x):
    x = (x + 0x9E3779B97F4A7C15) & MASK
    x = ((x ^ (x >> 30)) * 0xBF58476D1CE4E5B9) & MASK
    x = ((x ^ (x >> 27)) * 0x94D049BB133111EB) & MASK
    return x ^ (x >> 31)


def test_splitmix64_known_answer():
    # First output of SplitMix64 from state 0 (Steele, Lea and Flood 2014; Vigna's reference code).
    assert splitmix64(0) == 0xE220A8397B1DCDAF


@pytest.mark.parametrize("seed", [0, 11, 12, 2**63 + 5])
def test_pixel_normal_matches_its_definition(seed):
    got = M.pixel_normal(seed, 0, 512)
    for p in range(512):
        a = splitmix64(seed ^ splitmix64(p))
        b = splitmix64(a)
        u1 = ((a >> 11) + 0.5) / 2.0 ** 53
        u2 = ((b >> 11) + 0.5) / 2.0 ** 53
        ref = math.sqrt(-2.0 * math.log(u1)) * math.cos(2.0 * math.pi * u2)
        assert abs(got[p] - ref) <= 1e-14 * max(1.0, abs(ref)), p


def test_pixel_normal_is_counter_based_and_normal():
    full = M.pixel_normal(7, 0, 400_000)
    assert np.array_equal(M.pixel_normal(7, 1234, 5678), full[1234:5678])
    n = full.size
    assert abs(full.mean()) < 5 / math.sqrt(n)
    assert abs(full.var() - 1.0) < 5 * math.sqrt(2.0 / n)
    assert abs((full ** 4).mean() / full.var() ** 2 - 3.0) < 0.05
    assert abs(np.corrcoef(full, M.pixel_normal(8, 0, n))[0, 1]) < 5 / math.sqrt(n)


def test_white_noise_initial_maps():
    target = np.zeros(12 * 8 * 8)
    x, info = M.initial_maps(target, members=3, seed=11)
    assert info is None and x.shape == (3, target.size)
    for m in range(3):
        assert np.array_equal(x[m], M.pixel_normal(11 + m, 0, target.size))


def test_coarse_to_fine_init(tmp_path):
    rng = np.random.default_rng(3)
    nside, nside_init = 16, 4
    target = rng.standard_normal(12 * nside ** 2)
    coarse = rng.standard_normal(12 * nside_init ** 2) * 3.0 + 2.0
    path = tmp_path / "init.npy"
    M.write_map(path, coarse)
    mean, std = 2.0, 3.0
    factor = (nside // nside_init) ** 2
    # auto noise: RMS of the target's scales below nside 4 (deviations from the means of its
    # groups of `factor` NESTED pixels).
    rms = 0.0
    for k in range(coarse.size):
        g = target[k * factor:(k + 1) * factor]
        rms += ((g - g.mean()) ** 2).sum()
    rms = math.sqrt(rms / target.size)
    x, info = M.initial_maps(target, 2, 5, path, "target", "auto", mean, std)
    assert info == {"file": str(path), "per_member": False, "nside": nside_init, "noise_std": pytest.approx(rms)}
    for m in range(2):
        base = np.array([(coarse[p // factor] - mean) / std for p in range(target.size)])
        assert np.allclose(x[m], base + rms * M.pixel_normal(5 + m, 0, target.size), rtol=0, atol=1e-12)
    # Standardised units and a fixed noise of 0 (one member).
    x, info = M.initial_maps(target, 1, 5, path, "standardized", 0.0)
    assert np.array_equal(x[0], np.repeat(coarse, factor)) and info["noise_std"] == 0.0
    # Several members from one init need a noise.
    with pytest.raises(ValueError, match="identical members"):
        M.initial_maps(target, 2, 5, path, "standardized", 0.0)
    # One init per member: "{m}" becomes the member index on three digits.
    for m in range(2):
        M.write_map(tmp_path / f"run_m{m:03d}.npy", coarse + m)
    x, info = M.initial_maps(target, 2, 5, str(tmp_path / "run_m{m}.npy"), "standardized", 0.0)
    assert info["per_member"] and np.array_equal(x[1], np.repeat(coarse + 1, factor))
    # An init finer than the target is refused.
    with pytest.raises(ValueError, match="incompatible"):
        M.initial_maps(np.zeros(12 * 2 * 2), 1, 5, path)


# ---------------------------------------------------------------------------
# Command lines (no GPU work)
# ---------------------------------------------------------------------------

def test_stats_command_line():
    assert S.main(["--out", "x.json"]) == 2  # no "--"
    assert S.main(["--out", "x.json", "--"]) == 2  # no map
    assert S.main(["--bogus", "--out", "x.json", "--", "m.npy"]) == 2
    assert S.parse_engine_options(["--J", "5", "--gpu", "--ranks", "4", "--precision", "fp64"]) == {
        "J": 5, "precision": "fp64"}
    assert S.default_j(64) == 5 and S.default_j(1024) == 9 and S.default_j(2) == 1


def test_synth_command_line():
    a = Y.parse_args(["--target", "t.npy", "--ensemble", "4", "--family-weights", "S1=10, S2=2.5",
                      "--init-noise", "0.3", "--weights", "ergodic", "--gpu", "--L", "4"])
    assert a["ensemble"] == 4 and a["init_noise"] == 0.3 and a["weights"] == "ergodic"
    assert a["family_weights"] == [("S1", 10.0), ("S2", 2.5)]
    for bad, msg in ((["--target", "t.npy", "--scatter-weight", "1"], "not supported by this library"),
                     (["--target", "t.npy", "--L", "2"], "not supported by this library"),
                     (["--target", "t.npy", "--family-weights", "S5=1"], "unknown family"),
                     (["--target", "t.npy", "--init-noise", "conditional"], "conditional"),
                     (["--niter", "3"], "--target is required")):
        with pytest.raises(ValueError, match=msg):
            Y.parse_args(bad)
    assert Y.parse_args(["--target", "t.npy", "--scatter-weight", "0"])["target"] == "t.npy"
