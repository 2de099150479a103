"""Marks files and the precomputed-marks estimator.

The reference keeps per-lens marks between runs in a FITS file
(create_marks_fits_file / load_marks_fits_file, src/input_formats.f90:605-806)
and re-forms the correlation from them with new lens values and weights
(calculate_tpcf_from_marks, src/main.f90:1017-1139).  This module reads and
writes the same layout with numpy alone (no cfitsio):

  primary HDU  (BITPIX 8, NAXIS 0) with SURVEY, NFIELD, NBIN, BINTYPE,
               THETAMIN, THETAMAX, [IHSTHETA], BINCE1.., BINED1.., [BINTH1..]
  one BINTABLE per field, EXTNAME 'FIELD  i', LENSFILE, columns
               x, y, w            (E, or D in double precision)
               xi_E_MARKS, xi_B_MARKS, K_MARKS   (n_bin E or n_bin D)

with w the lens' marks_weight.  Rows are lenses; a marks column holds the
lens' n_bin values, bin fastest (the reference's (n_bin, n_A) arrays).
"""
from __future__ import annotations

import os
from dataclasses import dataclass, field
from typing import Optional

import numpy as np

BLOCK = 2880
CARD = 80


# ------------------------------------------------------------- headers ----
def _fmt_value(v) -> str:
    if isinstance(v, bool):
        return f"{'T' if v else 'F':>20}"
    if isinstance(v, (int, np.integer)):
        return f"{int(v):>20d}"
    if isinstance(v, (float, np.floating)):
        s = f"{float(v):.7E}"
        return f"{s:>20}"
    raise TypeError(type(v))


def _card(key: str, value=None, comment: str = "") -> list[str]:
    """One header keyword as 80-character cards (long strings continue with
    the OGIP 1.0 CONTINUE convention, as cfitsio's FTPKLS writes them)."""
    if value is None:
        return [f"{key:<80}"[:80]]
    if isinstance(value, str):
        text = value.replace("'", "''")
        if len(text) <= 68:
            body = f"{key:<8}= '{text:<8}'"
            if comment:
                body += f" / {comment}"
            return [f"{body:<80}"[:80]]
        cards = []
        first, rest = text[:67], text[67:]
        cards.append(f"{key:<8}= '{first}&'"[:80].ljust(80))
        while rest:
            chunk, rest = rest[:67], rest[67:]
            amp = "&" if rest else ""
            cards.append(f"CONTINUE  '{chunk}{amp}'"[:80].ljust(80))
        return cards
    body = f"{key:<8}= {_fmt_value(value)}"
    if comment:
        body += f" / {comment}"
    return [f"{body:<80}"[:80]]


def _header_bytes(cards: list[str]) -> bytes:
    cards = cards + [f"{'END':<80}"]
    raw = "".join(cards).encode("ascii")
    pad = (-len(raw)) % BLOCK
    return raw + b" " * pad


def _parse_header(buf: bytes, pos: int) -> tuple[dict, int]:
    """Header cards from `pos` to END; returns (keywords, data offset)."""
    hdr: dict = {}
    last_str_key = None
    while True:
        block = buf[pos:pos + BLOCK]
        if len(block) < BLOCK:
            raise ValueError("truncated FITS header")
        pos += BLOCK
        done = False
        for i in range(0, BLOCK, CARD):
            card = block[i:i + CARD].decode("ascii")
            key = card[:8].strip()
            if key == "END":
                done = True
                break
            if key == "CONTINUE" and last_str_key is not None:
                s = _parse_string(card[8:].strip())
                prev = hdr[last_str_key]
                hdr[last_str_key] = (prev[:-1] if prev.endswith("&") else prev) + s
                continue
            if card[8:10] != "= ":
                continue
            val = card[10:].strip()
            if val.startswith("'"):
                hdr[key] = _parse_string(val)
                last_str_key = key
            else:
                last_str_key = None
                v = val.split("/")[0].strip()
                if v in ("T", "F"):
                    hdr[key] = v == "T"
                else:
                    try:
                        hdr[key] = int(v)
                    except ValueError:
                        hdr[key] = float(v.replace("D", "E"))
        if done:
            return hdr, pos


def _parse_string(v: str) -> str:
    out, i = [], 1
    while i < len(v):
        c = v[i]
        if c == "'":
            if i + 1 < len(v) and v[i + 1] == "'":
                out.append("'")
                i += 2
                continue
            break
        out.append(c)
        i += 1
    return "".join(out).rstrip()


# -------------------------------------------------------------- tables ----
_TCODE = {"E": ">f4", "D": ">f8", "J": ">i4", "K": ">i8", "I": ">i2", "B": "u1"}


def _tform_dtype(tform: str):
    t = tform.strip()
    i = 0
    while i < len(t) and t[i].isdigit():
        i += 1
    rep = int(t[:i]) if i else 1
    code = t[i]
    if code not in _TCODE:
        raise ValueError(f"unsupported TFORM {tform}")
    return np.dtype(_TCODE[code]), rep


@dataclass
class MarksField:
    """One field of a marks file: lens positions, marks_weight and the
    per-lens marks [n][n_bin]."""
    x: np.ndarray
    y: np.ndarray
    w: np.ndarray
    xi_E_marks: np.ndarray
    xi_B_marks: np.ndarray
    K_marks: np.ndarray
    lensfile: str = ""

    @property
    def n(self) -> int:
        return len(self.x)


@dataclass
class MarksFile:
    n_bin: int
    spacing: int
    theta_min: float
    theta_max: float
    ihs_theta: float = 1.0
    bin_centers: Optional[np.ndarray] = None
    bin_edges: Optional[np.ndarray] = None
    bin_centers_eff: Optional[np.ndarray] = None
    fields: list = field(default_factory=list)


def write_marks_fits(filename: str, mf: MarksFile, double_precision: bool = False) -> None:
    """create_marks_fits_file (src/input_formats.f90:605-694)."""
    nb = mf.n_bin
    cards = []
    cards += _card("SIMPLE", True, "file does conform to FITS standard")
    cards += _card("BITPIX", 8, "number of bits per data pixel")
    cards += _card("NAXIS", 0, "number of data axes")
    cards += _card("EXTEND", True, "FITS dataset may contain extensions")
    cards += _card("SURVEY", "")
    cards += _card("NFIELD", len(mf.fields), "Number of fields")
    cards += _card("NBIN", nb, "Number of angular bins")
    cards += _card("BINTYPE", int(mf.spacing), "Bin spacing. lin = 0, log = 1, ihs = 2")
    cards += _card("THETAMIN", float(np.float32(mf.theta_min)), "Innermost bin edge")
    cards += _card("THETAMAX", float(np.float32(mf.theta_max)), "Outermost bin edge")
    if mf.spacing == 2:
        cards += _card("IHSTHETA", float(np.float32(mf.ihs_theta)), "IHS binning theta parameter")
    if mf.bin_centers is not None:
        for i, v in enumerate(np.asarray(mf.bin_centers, dtype=np.float32)):
            cards += _card(f"BINCE{i + 1}", float(v), "Bin center")
    if mf.bin_edges is not None:
        for i, v in enumerate(np.asarray(mf.bin_edges, dtype=np.float32)):
            cards += _card(f"BINED{i + 1}", float(v), "Bin edge")
    if mf.bin_centers_eff is not None:
        for i, v in enumerate(np.asarray(mf.bin_centers_eff, dtype=np.float32)):
            cards += _card(f"BINTH{i + 1}", float(v), "Mean theta")
    out = [_header_bytes(cards)]
    code = "D" if double_precision else "E"
    dt = np.dtype(">f8" if double_precision else ">f4")
    for i, f in enumerate(mf.fields):
        n = f.n
        rec = np.dtype([("x", dt), ("y", dt), ("w", dt), ("xi_E_MARKS", dt, (nb,)), ("xi_B_MARKS", dt, (nb,)),
                        ("K_MARKS", dt, (nb,))])
        tab = np.zeros(n, dtype=rec)
        tab["x"], tab["y"], tab["w"] = f.x, f.y, f.w
        tab["xi_E_MARKS"] = np.asarray(f.xi_E_marks).reshape(n, nb)
        tab["xi_B_MARKS"] = np.asarray(f.xi_B_marks).reshape(n, nb)
        tab["K_MARKS"] = np.asarray(f.K_marks).reshape(n, nb)
        h = []
        h += _card("XTENSION", "BINTABLE", "binary table extension")
        h += _card("BITPIX", 8, "8-bit bytes")
        h += _card("NAXIS", 2, "2-dimensional binary table")
        h += _card("NAXIS1", rec.itemsize, "width of table in bytes")
        h += _card("NAXIS2", n, "number of rows in table")
        h += _card("PCOUNT", 0, "size of special data area")
        h += _card("GCOUNT", 1, "one data group (required keyword)")
        h += _card("TFIELDS", 6, "number of fields in each row")
        for j, (name, form) in enumerate((("x", code), ("y", code), ("w", code), ("xi_E_MARKS", f"{nb}{code}"),
                                          ("xi_B_MARKS", f"{nb}{code}"), ("K_MARKS", f"{nb}{code}")), start=1):
            h += _card(f"TTYPE{j}", name)
            h += _card(f"TFORM{j}", form)
        h += _card("EXTNAME", f"FIELD{i + 1:3d}", "name of this binary table extension")
        if len(f.lensfile) > 68:
            h += _card("LONGSTRN", "OGIP 1.0", "The OGIP Long String Convention may be used.")
        h += _card("LENSFILE", f.lensfile)
        data = tab.tobytes()
        out += [_header_bytes(h), data + b"\0" * ((-len(data)) % BLOCK)]
    tmp = filename + ".tmp"
    with open(tmp, "wb") as fh:
        for part in out:
            fh.write(part)
    os.replace(tmp, filename)


def read_marks_fits(filename: str) -> MarksFile:
    """load_marks_fits_file (src/input_formats.f90:696-790): NFIELD must
    equal the number of table HDUs; marks columns are returned as [n][n_bin]
    FP64 arrays, the other columns as FP64 vectors."""
    with open(filename, "rb") as fh:
        buf = fh.read()
    hdr, pos = _parse_header(buf, 0)
    nb = int(hdr["NBIN"])
    mf = MarksFile(n_bin=nb, spacing=int(hdr["BINTYPE"]), theta_min=float(hdr["THETAMIN"]),
                   theta_max=float(hdr["THETAMAX"]), ihs_theta=float(hdr.get("IHSTHETA", 1.0)))
    if "BINCE1" in hdr:
        mf.bin_centers = np.array([hdr[f"BINCE{i + 1}"] for i in range(nb)])
    if "BINED1" in hdr:
        mf.bin_edges = np.array([hdr[f"BINED{i + 1}"] for i in range(nb + 1)])
    if "BINTH1" in hdr:
        mf.bin_centers_eff = np.array([hdr[f"BINTH{i + 1}"] for i in range(nb)])
    # primary data (NAXIS 0: none)
    while pos < len(buf):
        h, pos = _parse_header(buf, pos)
        if h.get("XTENSION") != "BINTABLE":
            raise ValueError("expected a BINTABLE extension")
        nrow, width = int(h["NAXIS2"]), int(h["NAXIS1"])
        names, fmts, shapes = [], [], []
        for j in range(1, int(h["TFIELDS"]) + 1):
            dt, rep = _tform_dtype(h[f"TFORM{j}"])
            names.append(h[f"TTYPE{j}"].upper())
            fmts.append(dt)
            shapes.append((rep,) if rep > 1 else ())
        rec = np.dtype([(n, d, s) for n, d, s in zip(names, fmts, shapes)])
        if rec.itemsize != width:
            raise ValueError("row width does not match the columns")
        nbytes = nrow * width
        tab = np.frombuffer(buf, dtype=rec, count=nrow, offset=pos)
        pos += nbytes + ((-nbytes) % BLOCK)

        def col(name):
            return np.asarray(tab[name], dtype=np.float64)

        mf.fields.append(MarksField(x=col("X"), y=col("Y"), w=col("W"),
                                    xi_E_marks=col("XI_E_MARKS").reshape(nrow, nb),
                                    xi_B_marks=col("XI_B_MARKS").reshape(nrow, nb),
                                    K_marks=col("K_MARKS").reshape(nrow, nb), lensfile=h.get("LENSFILE", "")))
    if len(mf.fields) != int(hdr["NFIELD"]):
        raise ValueError("number of HDUs does not match NFIELD")
    return mf


# ------------------------------------------------------------ estimator ----
def tpcf_from_marks(mf: MarksFile, lenses: list, tol: float = 1e-5) -> dict:
    """calculate_tpcf_from_marks (src/main.f90:1017-1139): marks made with
    unit lens values and weights, re-weighted with each field's lens values
    and weights: xi marks x val w, K marks x w, marks weight x w; lenses with
    weight 0 dropped; per field g = sum xi / sum K, weights = sum w, and the
    stack over fields with the summed xi and K.  ``lenses``: per field
    objects with x, y, val, w (the positions must match the file within
    `tol`, as the reference checks)."""
    if len(lenses) != len(mf.fields):
        raise ValueError("number of fields does not match")
    nb = mf.n_bin
    per_field = []
    st_e, st_b, st_k, st_w = np.zeros(nb), np.zeros(nb), np.zeros(nb), 0.0
    for f, L in zip(mf.fields, lenses):
        x, y = np.asarray(L.x, dtype=np.float64), np.asarray(L.y, dtype=np.float64)
        if len(x) != f.n:
            raise ValueError("number of lenses does not match")
        if np.any(np.abs(f.x - x) > tol) or np.any(np.abs(f.y - y) > tol):
            raise ValueError("lens coordinates do not match the marks file")
        val, w = np.asarray(L.val, dtype=np.float64), np.asarray(L.w, dtype=np.float64)
        xe = f.xi_E_marks * (val * w)[:, None]
        xb = f.xi_B_marks * (val * w)[:, None]
        km = f.K_marks * w[:, None]
        ww = f.w * w
        keep = ww > 0.0
        se, sb, sk = xe[keep].sum(axis=0), xb[keep].sum(axis=0), km[keep].sum(axis=0)
        with np.errstate(invalid="ignore", divide="ignore"):
            per_field.append(dict(g_t=se / sk, g_x=sb / sk, weights=ww[keep].sum(), n=int(keep.sum())))
        st_e += se
        st_b += sb
        st_k += sk
        st_w += ww[keep].sum()
    with np.errstate(invalid="ignore", divide="ignore"):
        stacked = dict(g_t=st_e / st_k, g_x=st_b / st_k, weights=st_w)
    return dict(fields=per_field, stacked=stacked)
