"""NcmMSetCatalog FITS layout (ncm_mset_catalog.c:1640-1735,
ncm_mset_catalog.h:234-257) written without cfitsio: structure against the
FITS standard (NASA FITS 4.0) and round trips.  The reference ships no
catalog file to read, so byte-level compatibility with cfitsio's output is
parity unpinned; the keys and column order follow the reference's writer."""
import struct

import numpy as np
import pytest

from numcosmo_amd import catalog as K


def _cards(raw, start):
    out = []
    pos = start
    while True:
        block = raw[pos:pos + 2880].decode("ascii")
        pos += 2880
        for k in range(0, 2880, 80):
            c = block[k:k + 80]
            out.append(c)
            if c.startswith("END     ") or c.rstrip() == "END":
                return out, pos


def _cat(nrows=37, ndim=3, nchains=4):
    g = np.random.default_rng(1)
    cat = K.MSetCatalog([f"NcmModelMVND:mu_{i}" for i in range(ndim)], [f"\\mu_{i}" for i in range(ndim)],
                        nchains=nchains)
    cat.add_rows(g.normal(size=(nrows, ndim + 1)))
    return cat


def test_round_trip_is_bit_exact(tmp_path):
    cat = _cat()
    p = str(tmp_path / "chain.fits")
    cat.write(p)
    back = K.MSetCatalog.read(p)
    assert np.array_equal(back.as_array(), cat.as_array())
    assert back.param_names == cat.param_names and back.param_symbols == cat.param_symbols
    assert back.add_names == ["NcmFit:m2lnL"] and back.add_symbols == ["-2\\ln(L)"]
    assert back.nchains == 4 and back.weighted is False and back.get_m2lnp_var() == 0
    assert len(back) == 37 and np.array_equal(back.peek_row(5), cat.peek_row(5))


def test_fits_structure(tmp_path):
    cat = _cat(nrows=100, ndim=5)
    p = str(tmp_path / "chain.fits")
    cat.write(p)
    raw = open(p, "rb").read()
    assert len(raw) % 2880 == 0
    prim, pos = _cards(raw, 0)
    assert prim[0] == "SIMPLE  =                    T / conforms to FITS standard".ljust(80)
    assert [c[:8].strip() for c in prim[:4]] == ["SIMPLE", "BITPIX", "NAXIS", "EXTEND"]
    ext, data0 = _cards(raw, pos)
    keys = [c[:8].strip() for c in ext]
    assert keys[:8] == ["XTENSION", "BITPIX", "NAXIS", "NAXIS1", "NAXIS2", "PCOUNT", "GCOUNT", "TFIELDS"]
    assert ext[0][10:20] == "'BINTABLE'"
    assert int(ext[3][10:30]) == 8 * 6 and int(ext[4][10:30]) == 100
    for lab in ("EXTNAME", "RTYPE", "NCHAINS", "NADDVAL", "WEIGHTED", "ASYMB1", "FSYMB1", "FSYMB5", "FIRST_ID",
                "M2LNP_ID", "TTYPE1", "TFORM6"):
        assert lab in keys
    assert all(len(c) == 80 and all(32 <= ord(ch) <= 126 for ch in c) for c in prim + ext)
    # big-endian float64 rows, column order (m2lnL, parameters)
    first = cat.peek_row(0)
    assert raw[data0:data0 + 8] == struct.pack(">d", first[0])
    assert raw[data0 + 8:data0 + 16] == struct.pack(">d", first[1])
    # zero padding after the table
    end = data0 + 100 * 6 * 8
    assert set(raw[end:]) <= {0}


def test_card_strings_and_errors(tmp_path):
    assert K._parse_value("'it''s    '") == "it's"
    assert K._parse_value("                   T / c") is True
    assert K._parse_value("                  42") == 42
    with pytest.raises(ValueError):
        K._card("TTYPE1", "x" * 80)
    cat = _cat()
    with pytest.raises(ValueError):
        cat.add_from_vector(np.zeros(3))
    bad = tmp_path / "bad.fits"
    bad.write_bytes(b"NOT A FITS FILE".ljust(2880))
    with pytest.raises(ValueError):
        K.MSetCatalog.read(str(bad))


def test_trim_mean_covar():
    cat = _cat(nrows=40, ndim=2, nchains=4)    # 10 steps of 4 chains
    A = cat.as_array()
    np.testing.assert_allclose(cat.get_mean(), A[:, 1:].mean(0), rtol=1e-14)
    np.testing.assert_allclose(cat.get_covar(), np.cov(A[:, 1:], rowvar=False), rtol=1e-13)
    cat.trim(3, 2)
    assert len(cat) == 4 * 4 and cat.first_id == 12
    np.testing.assert_array_equal(cat.as_array(), A.reshape(10, 4, 3)[3::2].reshape(-1, 3))
    with pytest.raises(ValueError):
        cat.trim(0, 0)
