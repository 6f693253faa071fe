"""MCMC chains in NcmMSetCatalog's FITS layout, without cfitsio.

NcmMSetCatalog (ncm_mset_catalog.c:1640-1735, labels in
ncm_mset_catalog.h:234-257) keeps an ensemble run as a FITS file with:
- a primary HDU (the reference also stores its model-set description there;
  this build has no NcmMSet and leaves it out);
- one binary table, EXTNAME 'NcmMSetCatalog:DATA', with one 1D (float64)
  column per additional value first (for ESMCMC: 'NcmFit:m2lnL'), then one
  per free parameter;
- one row per walker and step, walkers of a step consecutive (NCHAINS per
  step);
- header keys RTYPE, NCHAINS, NADDVAL, WEIGHTED, ASYMB<i> and FSYMB<i>
  (symbols of the additional values and parameters), FIRST_ID and M2LNP_ID.

The file is written and read here by a restatement of the FITS standard
(NASA FITS 4.0): 2880-byte blocks, 80-character header cards, big-endian
binary table data.  ``APES.write_catalog`` stores a run's chain this way.
"""
from __future__ import annotations

import os
import warnings
from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np

EXTNAME = "NcmMSetCatalog:DATA"
M2LNL_COLNAME = "NcmFit:m2lnL"
M2LNL_SYMBOL = "-2\\ln(L)"
BLOCK = 2880
CARD = 80


# ---- FITS cards ---------------------------------------------------------------
def _card(key: str, value=None, comment: Optional[str] = None) -> str:
    """One 80-character header card (fixed format for logical, integer, real
    and string values)."""
    if value is None:
        text = key.ljust(8)
    else:
        if isinstance(value, bool):
            v = ("T" if value else "F").rjust(20)
        elif isinstance(value, (int, np.integer)):
            v = str(int(value)).rjust(20)
        elif isinstance(value, (float, np.floating)):
            v = repr(float(value)).upper().rjust(20)
        else:
            s = str(value).replace("'", "''")
            v = ("'" + s.ljust(8) + "'").ljust(20)
        text = key.ljust(8) + "= " + v
        if comment:
            text += " / " + comment
    if len(text) > CARD or any(ord(ch) < 32 or ord(ch) > 126 for ch in text):
        raise ValueError(f"FITS card does not fit 80 ASCII characters: {text!r}")
    return text.ljust(CARD)


def _header(cards: Sequence[str]) -> bytes:
    h = "".join(cards) + "END".ljust(CARD)
    pad = (-len(h)) % BLOCK
    return (h + " " * pad).encode("ascii")


def _parse_value(raw: str):
    raw = raw.strip()
    if raw.startswith("'"):
        # string: up to the closing quote ('' is an escaped quote)
        out, i = [], 1
        while i < len(raw):
            if raw[i] == "'":
                if i + 1 < len(raw) and raw[i + 1] == "'":
                    out.append("'")
                    i += 2
                    continue
                break
            out.append(raw[i])
            i += 1
        return "".join(out).rstrip()
    val = raw.split("/", 1)[0].strip()
    if val == "T":
        return True
    if val == "F":
        return False
    try:
        return int(val)
    except ValueError:
        return float(val.replace("D", "E"))


def _read_header(fh) -> Tuple[List[Tuple[str, object]], int]:
    cards: List[Tuple[str, object]] = []
    nbytes = 0
    while True:
        block = fh.read(BLOCK)
        if len(block) != BLOCK:
            raise ValueError("truncated FITS header")
        nbytes += BLOCK
        text = block.decode("ascii")
        for k in range(0, BLOCK, CARD):
            c = text[k:k + CARD]
            key = c[:8].strip()
            if key == "END":
                return cards, nbytes
            if c[8:10] == "= ":
                cards.append((key, _parse_value(c[10:])))
            elif key:
                cards.append((key, None))


# ---- the catalog ------------------------------------------------------------------
class MSetCatalog:
    """Rows of (additional values, free parameters) of an ensemble run.

    Mirrors the parts of NcmMSetCatalog that carry an ESMCMC chain:
    ``ncm_mset_catalog_new`` (nadd_vals, nchains, weighted, names and
    symbols), ``ncm_mset_catalog_add_from_vector``, ``ncm_mset_catalog_peek_row``,
    ``ncm_mset_catalog_len``, ``ncm_mset_catalog_get_m2lnp_var``,
    ``ncm_mset_catalog_set_file`` + ``ncm_mset_catalog_sync`` (``write``) and
    ``ncm_mset_catalog_new_from_file_ro`` (``read``)."""

    def __init__(self, param_names: Sequence[str], param_symbols: Optional[Sequence[str]] = None,
                 nchains: int = 1, weighted: bool = False,
                 add_vals: Sequence[Tuple[str, str]] = ((M2LNL_COLNAME, M2LNL_SYMBOL),),
                 rtype: str = "NcmFitESMCMC:APES", first_id: int = 0, m2lnp_var: int = 0):
        self.param_names = list(param_names)
        self.param_symbols = list(param_symbols) if param_symbols is not None else list(param_names)
        if len(self.param_symbols) != len(self.param_names):
            raise ValueError("one symbol per parameter")
        self.add_names = [a for a, _ in add_vals]
        self.add_symbols = [s for _, s in add_vals]
        self.nchains, self.weighted, self.rtype = int(nchains), bool(weighted), str(rtype)
        self.first_id, self.m2lnp_var = int(first_id), int(m2lnp_var)
        self._rows: List[np.ndarray] = []

    @property
    def ncols(self) -> int:
        return len(self.add_names) + len(self.param_names)

    def add_from_vector(self, row) -> None:
        row = np.asarray(row, dtype=np.float64).reshape(-1)
        if row.size != self.ncols:
            raise ValueError(f"row of {row.size} values, the catalog has {self.ncols} columns")
        self._rows.append(row.copy())

    def add_rows(self, rows) -> None:
        rows = np.asarray(rows, dtype=np.float64).reshape(-1, self.ncols)
        self._rows.extend(r.copy() for r in rows)

    def __len__(self) -> int:
        return len(self._rows)

    def peek_row(self, i: int) -> np.ndarray:
        return self._rows[i]

    def as_array(self) -> np.ndarray:
        return np.array(self._rows).reshape(len(self._rows), self.ncols)

    def get_m2lnp_var(self) -> int:
        return self.m2lnp_var

    def get_first_id(self) -> int:
        return self.first_id

    def get_cur_id(self) -> int:
        """ncm_mset_catalog_get_cur_id (ncm_mset_catalog.c:2731): id of the
        last row, first_id - 1 when the catalog is empty."""
        return self.first_id + len(self._rows) - 1

    def get_row_from_time(self, t: int) -> int:
        """ncm_mset_catalog_get_row_from_time (ncm_mset_catalog.c:2699): the
        row holding sample id t."""
        row = t - self.first_id
        if t < self.first_id or row >= len(self._rows):
            raise IndexError(f"sample id {t} is not in the catalog (ids {self.first_id}..{self.get_cur_id()})")
        return row

    def remove_last_ensemble(self, path: Optional[str] = None) -> None:
        """ncm_mset_catalog_remove_last_ensemble (ncm_mset_catalog.c:5390-5446):
        drop the last NCHAINS rows, or the trailing partial ensemble
        (len % NCHAINS rows) when there is one.  With ``path``, the file
        there is first renamed to ``path.<k>.bak`` (the first k not in use)
        and the shortened catalog is written to ``path``."""
        n = len(self._rows)
        if n == 0:
            warnings.warn("MSetCatalog.remove_last_ensemble: called on an empty catalog.")
            return
        drop = self.nchains if n % self.nchains == 0 else n % self.nchains
        del self._rows[n - drop:]
        if path is not None:
            k = 0
            while os.path.exists(f"{path}.{k}.bak"):
                k += 1
            if os.path.exists(path):
                os.rename(path, f"{path}.{k}.bak")
            self.write(path)

    def trim(self, tc: int, thin: int = 1) -> None:
        """ncm_mset_catalog_trim (ncm_mset_catalog.c:5255-5315): drop the
        first tc steps (tc x nchains rows) and keep every thin-th step after
        them; FIRST_ID advances by tc x nchains."""
        if thin < 1 or tc < 0:
            raise ValueError("trim: thin >= 1 and tc >= 0")
        if len(self._rows) % self.nchains != 0:
            raise ValueError("trim: the catalog does not hold whole steps")
        if thin == 1 and tc == 0:
            return
        steps = len(self._rows) // self.nchains
        rows = self._rows
        self._rows = [rows[t * self.nchains + i] for t in range(tc, steps, thin) for i in range(self.nchains)]
        self.first_id += tc * self.nchains

    def get_mean(self) -> np.ndarray:
        """ncm_mset_catalog_get_mean: mean of the parameter columns."""
        return self.as_array()[:, len(self.add_names):].mean(axis=0)

    def get_covar(self) -> np.ndarray:
        """ncm_mset_catalog_get_covar: covariance of the parameter columns
        (bias 1/(n-1), as NcmStatsVec)."""
        P = self.as_array()[:, len(self.add_names):]
        return np.atleast_2d(np.cov(P, rowvar=False, ddof=1))

    # ---- FITS -------------------------------------------------------------------
    def write(self, path: str) -> None:
        data = self.as_array()
        nrows, ncols = data.shape
        prim = _header([_card("SIMPLE", True, "conforms to FITS standard"), _card("BITPIX", 8), _card("NAXIS", 0),
                        _card("EXTEND", True)])
        cards = [_card("XTENSION", "BINTABLE", "binary table extension"), _card("BITPIX", 8), _card("NAXIS", 2),
                 _card("NAXIS1", 8 * ncols, "width of table in bytes"), _card("NAXIS2", nrows, "number of rows"),
                 _card("PCOUNT", 0), _card("GCOUNT", 1), _card("TFIELDS", ncols, "number of columns")]
        for k, name in enumerate(self.add_names + self.param_names):
            cards.append(_card(f"TTYPE{k + 1}", name))
            cards.append(_card(f"TFORM{k + 1}", "1D"))
        cards += [_card("EXTNAME", EXTNAME), _card("RTYPE", self.rtype, "Run type string."),
                  _card("NCHAINS", self.nchains, "Number of chains."),
                  _card("NADDVAL", len(self.add_names), "Number of additional values."),
                  _card("WEIGHTED", self.weighted, "Whether the catalog is weighted.")]
        for k, s in enumerate(self.add_symbols):
            cards.append(_card(f"ASYMB{k + 1}", s))
        for k, s in enumerate(self.param_symbols):
            cards.append(_card(f"FSYMB{k + 1}", s))
        cards += [_card("FIRST_ID", self.first_id, "Id of the first element."),
                  _card("M2LNP_ID", self.m2lnp_var, "Id of the m2lnp variable.")]
        body = data.astype(">f8").tobytes()
        body += b"\0" * ((-len(body)) % BLOCK)
        tmp = path + ".tmp"
        with open(tmp, "wb") as fh:
            fh.write(prim)
            fh.write(_header(cards))
            fh.write(body)
        os.replace(tmp, path)

    @classmethod
    def read(cls, path: str) -> "MSetCatalog":
        with open(path, "rb") as fh:
            prim, _ = _read_header(fh)
            p = dict(prim)
            if p.get("SIMPLE") is not True or p.get("NAXIS", 0) != 0:
                raise ValueError("not a FITS file with an empty primary HDU")
            cards, _ = _read_header(fh)
            h: Dict[str, object] = dict(cards)
            if h.get("XTENSION") != "BINTABLE" or h.get("EXTNAME") != EXTNAME:
                raise ValueError(f"no {EXTNAME} binary table")
            ncols, nrows = int(h["TFIELDS"]), int(h["NAXIS2"])
            if int(h["NAXIS1"]) != 8 * ncols or any(h[f"TFORM{k + 1}"] != "1D" for k in range(ncols)):
                raise ValueError("only 1D (float64) columns are supported")
            nadd = int(h["NADDVAL"])
            names = [str(h[f"TTYPE{k + 1}"]) for k in range(ncols)]
            add_vals = [(names[k], str(h.get(f"ASYMB{k + 1}", names[k]))) for k in range(nadd)]
            pnames = names[nadd:]
            psymbs = [str(h.get(f"FSYMB{k + 1}", pnames[k])) for k in range(len(pnames))]
            raw = fh.read(8 * ncols * nrows)
            if len(raw) != 8 * ncols * nrows:
                raise ValueError("truncated FITS table")
        cat = cls(pnames, psymbs, nchains=int(h["NCHAINS"]), weighted=bool(h["WEIGHTED"]), add_vals=add_vals,
                  rtype=str(h["RTYPE"]), first_id=int(h.get("FIRST_ID", 0)), m2lnp_var=int(h.get("M2LNP_ID", 0)))
        cat.add_rows(np.frombuffer(raw, dtype=">f8").astype(np.float64).reshape(nrows, ncols))
        return cat
