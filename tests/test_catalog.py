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


def _id_cat(nrows, ndim=2, nchains=6, first_id=0):
    """Rows whose values encode their sample id: m2lnL = 1000 + id, theta =
    (id, -id, ...)."""
    cat = K.MSetCatalog([f"p{i}" for i in range(ndim)], nchains=nchains, first_id=first_id)
    for r in range(nrows):
        i = first_id + r
        cat.add_from_vector([1000.0 + i] + [i * (-1.0) ** j for j in range(ndim)])
    return cat


def test_ids_and_rows():
    cat = _id_cat(20, first_id=12)
    assert cat.get_first_id() == 12 and cat.get_cur_id() == 31
    assert cat.get_row_from_time(12) == 0 and cat.get_row_from_time(31) == 19
    for t in (11, 32):
        with pytest.raises(IndexError):
            cat.get_row_from_time(t)
    assert K.MSetCatalog(["a"], nchains=3).get_cur_id() == -1


@pytest.mark.parametrize("nrows,left", [(24, 18), (22, 18), (19, 18), (18, 12), (5, 0), (6, 0)])
def test_remove_last_ensemble(nrows, left):
    # ncm_mset_catalog.c:5390-5446: a whole last ensemble goes (6 rows), a
    # trailing partial one goes alone (len % 6 rows)
    cat = _id_cat(nrows)
    before = cat.as_array()
    cat.remove_last_ensemble()
    assert len(cat) == left
    assert np.array_equal(cat.as_array(), before[:left].reshape(left, 3))
    assert cat.get_cur_id() == left - 1


def test_remove_last_ensemble_backs_up_the_file(tmp_path):
    p = str(tmp_path / "run.fits")
    cat = _id_cat(30)
    cat.write(p)
    open(p + ".0.bak", "wb").write(b"older backup")
    cat.remove_last_ensemble(p)
    assert open(p + ".0.bak", "rb").read() == b"older backup"
    assert np.array_equal(K.MSetCatalog.read(p + ".1.bak").as_array(), _id_cat(30).as_array())
    assert np.array_equal(K.MSetCatalog.read(p).as_array(), _id_cat(24).as_array())
    empty = K.MSetCatalog(["a"], nchains=2)
    with pytest.warns(UserWarning):
        empty.remove_last_ensemble()


def _latest_rows(cat, nw):
    """Brute force: each walker's last row in the catalog (walker = id mod
    nw), or None when the catalog holds none for it."""
    rows = cat.as_array()
    out = [None] * nw
    for r in range(len(cat)):
        out[(cat.get_first_id() + r) % nw] = rows[r]
    return out


@pytest.mark.parametrize("nrows,first_id", [(36, 0), (39, 0), (41, 0), (30, 6), (33, 12), (4, 0), (6, 0)])
def test_catalog_state_is_each_walkers_latest_row(nrows, first_id):
    from numcosmo_amd.apes import catalog_state
    nw = 6
    cat = _id_cat(nrows, nchains=nw, first_id=first_id)
    th, m2 = catalog_state(cat, nw)
    want = [r for r in _latest_rows(cat, nw) if r is not None]
    assert th.shape == (len(want), 2)
    assert np.array_equal(m2, [r[0] for r in want])
    assert np.array_equal(th, [r[1:] for r in want])


def test_catalog_state_trimmed_without_a_first_ensemble():
    from numcosmo_amd.apes import catalog_state
    # ids 2..4 of a 6-walker run: no complete first ensemble, and rows before
    # id 2 are gone (ncm_fit_esmcmc.c:1830-1831)
    with pytest.raises(ValueError, match="first_id > 0"):
        catalog_state(_id_cat(3, first_id=2), 6)
    # ids 8..13: walker 0..1 at step 2 need ids 12, 13 and walkers 2..5 ids 8..11
    th, _ = catalog_state(_id_cat(6, first_id=8), 6)
    assert np.array_equal(th[:, 0], [12, 13, 8, 9, 10, 11])
    # ids 8..11 only: walkers 0..5 of step 1 would need ids 6, 7 (trimmed)
    with pytest.raises(IndexError):
        catalog_state(_id_cat(4, first_id=8), 6)


def test_validate_rows():
    from numcosmo_amd.apes import validate_rows
    cat = _id_cat(18)
    like = lambda th: 1000.0 + th[:, 0]   # noqa: E731 - recomputes the rows' own m2lnL
    assert validate_rows(cat, like, 12, 18)
    rows = cat.as_array()
    for rel, ok in [(0.9e-3, True), (1.1e-3, False)]:
        c = K.MSetCatalog(["p0", "p1"], nchains=6)
        r = rows.copy()
        r[15, 0] *= 1.0 + rel
        c.add_rows(r)
        assert validate_rows(c, like, 12, 18) is ok
        assert validate_rows(c, like, 0, 12)   # the planted row is outside
    # a NaN difference does not fail (the comparison is false), as in the reference
    c = K.MSetCatalog(["p0", "p1"], nchains=6)
    r = rows.copy()
    r[14, 0] = np.nan
    c.add_rows(r)
    assert validate_rows(c, like, 12, 18)
