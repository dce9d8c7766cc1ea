"""Obit's uv FITS tables (.uvtab): read into the record layout the device
path uses, write back.

Reading follows ObitIOUVFITS (src/ObitIOUVFITS.c:1540-1640 for the
descriptor, :680 for the data): the 'AIPS UV' binary table's rows are
taken as raw bytes, so the random parameters are the stored values (u, v, w
in wavelengths at the reference frequency, the date as days from the
column's TZERO, which becomes JDObs) and the VISIBILITIES column is the
record [IF][chan][Stokes][re, im, wt] in the order of its TDIM.  The
channel frequencies follow ObitUVDescGetFreq (src/ObitUVDesc.c:520-610):
the 'AIPS FQ' table's IF offsets, channel widths and sidebands about the
FREQ axis.  Writing replaces the rows of the 'AIPS UV' table and copies
every other HDU byte for byte; a file can also be written from a UVData
alone, with the scalings ObitIOUVFITS writes (TSCAL = 1/freq for u, v, w
and TZERO = JDObs for the date, :1916-1944).
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from .synth import UVData

BLOCK = 2880


# ------------------------------------------------------------------ FITS
@dataclass
class HDU:
    cards: list            # [(key, value string, comment)], value as written
    data: bytes            # data bytes, unpadded
    raw_header: bytes = b""  # the header blocks as read (written back unchanged)

    def get(self, key, default=None):
        for k, v, _ in self.cards:
            if k == key:
                return v
        return default

    def num(self, key, default=None):
        v = self.get(key)
        return default if v is None else float(v.replace("D", "E"))

    def text(self, key, default=""):
        v = self.get(key)
        if v is None:
            return default
        return v.strip().strip("'").rstrip()


def _parse_card(card):
    key = card[:8].strip()
    if card[8:10] != "= ":
        return key, None, card[8:].rstrip()
    rest = card[10:]
    if rest.lstrip().startswith("'"):
        s = rest.lstrip()
        end = 1
        while True:
            end = s.find("'", end)
            if end < 0:
                end = len(s)
                break
            if end + 1 < len(s) and s[end + 1] == "'":
                end += 2
                continue
            break
        val = s[:end + 1]
        comment = s[end + 1:].split("/", 1)[1].strip() if "/" in s[end + 1:] else ""
        return key, val, comment
    val, _, comment = rest.partition("/")
    return key, val.strip(), comment.strip()


def read_hdus(raw: bytes):
    """Every HDU of a FITS byte string, in order."""
    out, off = [], 0
    while off + BLOCK <= len(raw):
        cards, done, start = [], False, off
        while not done:
            block = raw[off:off + BLOCK]
            off += BLOCK
            for i in range(36):
                card = block[80 * i:80 * (i + 1)].decode("ascii", "replace")
                if card[:8].strip() == "END":
                    done = True
                    break
                if card.strip():
                    cards.append(_parse_card(card))
        h = HDU(cards, b"", raw[start:off])
        naxis = int(h.num("NAXIS", 0))
        size = 0
        if naxis and not (h.get("XTENSION") is None and int(h.num("NAXIS1", 0)) == 777777701):
            size = abs(int(h.num("BITPIX", 8))) // 8
            for i in range(1, naxis + 1):
                size *= int(h.num("NAXIS%d" % i))
            if h.get("XTENSION") is not None:
                size += int(h.num("PCOUNT", 0))
        h.data = raw[off:off + size]
        off += ((size + BLOCK - 1) // BLOCK) * BLOCK
        out.append(h)
    return out


def _fmt_card(key, val, comment=""):
    if val is None:
        card = f"{key:<8}{comment}"
    elif val.startswith("'"):
        card = f"{key:<8}= {val:<20}"  # strings start in column 11
        if comment:
            card += f" / {comment}"
    else:
        card = f"{key:<8}= {val:>20}"  # numbers and logicals end in column 30
        if comment:
            card += f" / {comment}"
    return card[:80].ljust(80)


def _pad(b: bytes, fill: bytes) -> bytes:
    return b + fill * ((-len(b)) % BLOCK)


def write_hdus(hdus) -> bytes:
    out = []
    for h in hdus:
        if h.raw_header:
            out.append(h.raw_header)
        else:
            hdr = "".join(_fmt_card(k, v, c) for k, v, c in h.cards) + "END".ljust(80)
            out.append(_pad(hdr.encode("ascii"), b" "))
        out.append(_pad(h.data, b"\0"))
    return b"".join(out)


_FMT = {"E": (">f4", 4), "D": (">f8", 8), "J": (">i4", 4), "I": (">i2", 2), "A": ("S1", 1)}


def table_columns(h: HDU):
    """{name: numpy array} of a binary table (E, D, J, I, A fields)."""
    n1, n2 = int(h.num("NAXIS1")), int(h.num("NAXIS2"))
    dt = []
    for i in range(1, int(h.num("TFIELDS")) + 1):
        tf = h.text("TFORM%d" % i)
        cnt = int(tf[:-1] or 1)
        code, _ = _FMT[tf[-1]]
        dt.append((h.text("TTYPE%d" % i), code if code != "S1" else f"S{cnt}", (cnt,) if code != "S1" else ()))
    arr = np.frombuffer(h.data[:n1 * n2], dtype=np.dtype(dt))
    return {name: np.array(arr[name]) for name, _, _ in dt}


# ------------------------------------------------------------- uv tables
@dataclass
class UVTable:
    """What a .uvtab holds besides the records."""
    ptype: list                 # random parameter names, in record order
    ctype: list                 # visibility axes, fastest first (COMPLEX, STOKES, FREQ, IF, RA, DEC)
    naxis: list
    crval: list
    cdelt: list
    crpix: list
    crota: list
    jd_obs: float = 0.0         # TZERO of the date column
    hdus: list = field(default_factory=list)   # the file's HDUs (the UV table's are replaced on write)
    extra: list = field(default_factory=list)  # other keywords of the UV table header

    def axis(self, name):
        return self.ctype.index(name) if name in self.ctype else -1


def _freq_arr(t: UVTable, fq):
    """ObitUVDescGetFreq: channel frequencies, frequency axis first."""
    jf, ji = t.axis("FREQ"), t.axis("IF")
    nfreq = t.naxis[jf]
    nif = t.naxis[ji] if ji >= 0 else 1
    f0 = t.crval[jf]
    off = np.zeros(nif)
    bw = np.full(nif, t.cdelt[jf])
    sb = np.zeros(nif, np.int64)
    if fq is not None and len(fq.get("FRQSEL", [])):
        sel = np.nonzero(fq["FRQSEL"].reshape(-1) == 1)[0]
        row = int(sel[0]) if sel.size else 0
        off = fq["IF FREQ"][row].astype(np.float64).reshape(-1)[:nif]
        bw = fq["CH WIDTH"][row].astype(np.float64).reshape(-1)[:nif]
        sb = fq["SIDEBAND"][row].astype(np.int64).reshape(-1)[:nif]
    out = np.empty(nif * nfreq)
    for j in range(nif):
        d = -abs(bw[j]) if sb[j] < 0 else bw[j]
        for i in range(nfreq):
            out[j * nfreq + i] = off[j] + f0 + (i + 1.0 - t.crpix[jf]) * d
    return out


def read(path):
    """(UVData, UVTable) of a .uvtab file."""
    hdus = read_hdus(open(path, "rb").read())
    uvh = next(h for h in hdus if h.text("EXTNAME") == "AIPS UV")
    fqh = next((h for h in hdus if h.text("EXTNAME") == "AIPS FQ"), None)
    ncol = int(uvh.num("TFIELDS"))
    ptype = [uvh.text("TTYPE%d" % i) for i in range(1, ncol)]
    for i in range(1, ncol):
        if uvh.text("TFORM%d" % i) not in ("1E", "E"):
            raise ValueError(f"UV table column {i} is not a scalar float")
    if uvh.text("TTYPE%d" % ncol)[:12] != "VISIBILITIES":
        raise ValueError("last UV table column is not VISIBILITIES")
    dims = [int(x) for x in uvh.text("TDIM%d" % ncol).strip("()").split(",")]
    jd = 0.0
    for i, p in enumerate(ptype, start=1):
        if p.startswith(("DATE", "TIME")):
            jd = uvh.num("TZERO%d" % i, 0.0)
    v = ncol
    t = UVTable(ptype=ptype, naxis=dims,
                ctype=[uvh.text("%dCTYP%d" % (k, v)) for k in range(1, len(dims) + 1)],
                crval=[uvh.num("%dCRVL%d" % (k, v), 0.0) for k in range(1, len(dims) + 1)],
                cdelt=[uvh.num("%dCDLT%d" % (k, v), 1.0) for k in range(1, len(dims) + 1)],
                crpix=[uvh.num("%dCRPX%d" % (k, v), 1.0) for k in range(1, len(dims) + 1)],
                crota=[uvh.num("%dCROT%d" % (k, v), 0.0) for k in range(1, len(dims) + 1)],
                jd_obs=jd, hdus=hdus)
    own = {"XTENSION", "BITPIX", "NAXIS", "NAXIS1", "NAXIS2", "PCOUNT", "GCOUNT", "TFIELDS", "EXTNAME",
           "EXTVER"}
    for k, val, c in uvh.cards:
        if k in own or k[:5] in ("TFORM", "TTYPE", "TUNIT", "TSCAL", "TZERO") or k[:4] == "TDIM" or \
                (len(k) > 5 and k[0].isdigit() and k[1:5] in ("CTYP", "CRVL", "CDLT", "CRPX", "CROT")):
            continue
        t.extra.append((k, val, c))
    if t.naxis[0] != 3:
        raise ValueError("compressed uv data are not supported")
    n1, n2 = int(uvh.num("NAXIS1")), int(uvh.num("NAXIS2"))
    nfl = n1 // 4
    if nfl * 4 != n1 or nfl != len(ptype) + int(np.prod(dims)):
        raise ValueError("UV table row is not all 4-byte floats")
    rows = np.frombuffer(uvh.data[:n1 * n2], dtype=">f4").reshape(n2, nfl).astype(np.float32)
    fq = table_columns(fqh) if fqh is not None else None
    js, jf, ji = t.axis("STOKES"), t.axis("FREQ"), t.axis("IF")
    if js != 1 or jf != 2 or (ji >= 0 and ji != 3):
        raise ValueError(f"axis order {t.ctype} is not COMPLEX, STOKES, FREQ, IF")
    loc = {name: next((i for i, p in enumerate(ptype) if p.startswith(pre)), -1)
           for name, pre in (("u", "UU-"), ("v", "VV-"), ("w", "WW-"), ("b", "BASELINE"))}
    lt = next((i for i, p in enumerate(ptype) if p.startswith(("DATE", "TIME"))), -1)
    uv = UVData(buffer=rows, nchan=dims[jf], nif=dims[ji] if ji >= 0 else 1, nstok=dims[js],
                freq=t.crval[jf], freqArr=_freq_arr(t, fq), crval_stokes=t.crval[js], nrparm=len(ptype),
                ilocu=loc["u"], ilocv=loc["v"], ilocw=loc["w"], ilocb=loc["b"], iloct=lt)
    return uv, t


def _uv_hdu(uv: UVData, t: UVTable) -> HDU:
    nfl = uv.lrec
    cards = [("XTENSION", "'BINTABLE'", ""), ("BITPIX", "8", ""), ("NAXIS", "2", ""),
             ("NAXIS1", str(4 * nfl), ""), ("NAXIS2", str(uv.nrow), ""), ("PCOUNT", "0", ""),
             ("GCOUNT", "1", ""), ("TFIELDS", str(uv.nrparm + 1), ""), ("EXTNAME", "'AIPS UV '", ""),
             ("EXTVER", "1", "")]
    for i, p in enumerate(t.ptype, start=1):
        cards += [("TFORM%d" % i, "'1E      '", ""), ("TTYPE%d" % i, "'%-8s'" % p, "")]
        if p.startswith(("UU-", "VV-", "WW-")):
            cards.append(("TSCAL%d" % i, "%.13E" % (1.0 / max(1.0, uv.freq)), "Scaling for %s to time units" % p[:1]))
        if p.startswith(("DATE", "TIME")):
            cards.append(("TZERO%d" % i, "%.13E" % t.jd_obs, "Offset of Date from JD"))
    v = uv.nrparm + 1
    nvis = int(np.prod(t.naxis))
    cards += [("TFORM%d" % v, "'%dE'" % nvis, ""), ("TTYPE%d" % v, "'VISIBILITIES'", ""),
              ("TDIM%d" % v, "'(%s)'" % ",".join(str(n) for n in t.naxis), "")]
    for k in range(len(t.naxis)):
        cards += [("%dCTYP%d" % (k + 1, v), "'%-8s'" % t.ctype[k], ""),
                  ("%dCRVL%d" % (k + 1, v), "%.11E" % t.crval[k], ""),
                  ("%dCDLT%d" % (k + 1, v), "%.9E" % t.cdelt[k], ""),
                  ("%dCRPX%d" % (k + 1, v), "%.9E" % t.crpix[k], ""),
                  ("%dCROT%d" % (k + 1, v), "%.9E" % t.crota[k], "")]
    cards += list(t.extra)
    data = np.ascontiguousarray(uv.buffer, np.float32).astype(">f4").tobytes()
    return HDU(cards, data)


def table_for(uv: UVData, ptype=None, jd_obs=0.0, ra=0.0, dec=0.0) -> UVTable:
    """A UVTable describing uv (COMPLEX, STOKES, FREQ, IF, RA, DEC axes;
    uniform channels about the first frequency of freqArr)."""
    ptype = ptype or ["UU-L-SIN", "VV-L-SIN", "WW-L-SIN", "BASELINE", "TIME1"][:uv.nrparm]
    fa = np.asarray(uv.freqArr, np.float64).reshape(uv.nif, uv.nchan)
    df = float(fa[0, 1] - fa[0, 0]) if uv.nchan > 1 else 1.0
    return UVTable(ptype=list(ptype), ctype=["COMPLEX", "STOKES", "FREQ", "IF", "RA", "DEC"],
                   naxis=[3, uv.nstok, uv.nchan, uv.nif, 1, 1],
                   crval=[1.0, float(uv.crval_stokes), float(uv.freq), 1.0, ra, dec],
                   cdelt=[1.0, -1.0 if uv.crval_stokes < 0 else 1.0, df, 1.0, 1.0, 1.0],
                   crpix=[1.0, 1.0, 1.0 + (uv.freq - fa[0, 0]) / df if uv.nchan > 1 else 1.0, 1.0, 1.0, 1.0],
                   crota=[0.0] * 6, jd_obs=jd_obs)


def _fq_hdu(uv: UVData, t: UVTable) -> HDU:
    fa = np.asarray(uv.freqArr, np.float64).reshape(uv.nif, uv.nchan)
    jf = t.axis("FREQ")
    off = fa[:, 0] - (uv.freq + (1.0 - t.crpix[jf]) * t.cdelt[jf])
    bw = np.full(uv.nif, t.cdelt[jf], np.float32)
    nif = uv.nif
    dt = np.dtype([("FRQSEL", ">i4", (1,)), ("IF FREQ", ">f8", (nif,)), ("CH WIDTH", ">f4", (nif,)),
                   ("TOTAL BANDWIDTH", ">f4", (nif,)), ("SIDEBAND", ">i4", (nif,))])
    row = np.zeros(1, dt)
    row["FRQSEL"] = 1
    row["IF FREQ"] = off
    row["CH WIDTH"] = bw
    row["TOTAL BANDWIDTH"] = bw * uv.nchan
    row["SIDEBAND"] = 1
    cards = [("XTENSION", "'BINTABLE'", ""), ("BITPIX", "8", ""), ("NAXIS", "2", ""),
             ("NAXIS1", str(dt.itemsize), ""), ("NAXIS2", "1", ""), ("PCOUNT", "0", ""), ("GCOUNT", "1", ""),
             ("TFIELDS", "5", ""), ("EXTNAME", "'AIPS FQ '", ""), ("EXTVER", "1", "")]
    for i, (name, form) in enumerate((("FRQSEL", "1J"), ("IF FREQ", "%dD" % nif), ("CH WIDTH", "%dE" % nif),
                                      ("TOTAL BANDWIDTH", "%dE" % nif), ("SIDEBAND", "%dJ" % nif)), start=1):
        cards += [("TFORM%d" % i, "'%s'" % form, ""), ("TTYPE%d" % i, "'%s'" % name, "")]
    cards.append(("NO_IF", str(nif), ""))
    return HDU(cards, row.tobytes())


def write(path, uv: UVData, t: UVTable | None = None):
    """Write uv as a .uvtab: the other HDUs of t's file copied, the UV table
    rewritten (or a primary HDU, an FQ table and the UV table when t has no
    file behind it)."""
    if t is None:
        t = table_for(uv)
    if list(t.naxis[1:4]) != [uv.nstok, uv.nchan, uv.nif] or len(t.ptype) != uv.nrparm:
        raise ValueError("uv data do not match the table description")
    uvh = _uv_hdu(uv, t)
    if t.hdus:
        hdus = [uvh if h.text("EXTNAME") == "AIPS UV" else h for h in t.hdus]
    else:
        prim = HDU([("SIMPLE", "T", ""), ("BITPIX", "8", ""), ("NAXIS", "0", ""), ("EXTEND", "T", "")], b"")
        hdus = [prim, _fq_hdu(uv, t), uvh]
    open(path, "wb").write(write_hdus(hdus))
