"""Binary head and cell-by-cell budget files in the reference's layout.

MODFLOW 6 opens both files as unformatted stream files (OpenSpec.f90:18,
ACCESS = 'STREAM'): a file is the bytes of its Fortran writes back to back,
with no record markers.  Integers are int32, reals float64, text fields 16
characters; byte order is the machine's (little-endian here).

Head file (Dis%record_array with a positive unit, Dis.f90:1463-1477): one
ulasav record per layer (InputOutput.f90:924-940)::

    kstp, kper (i4)  pertim, totim (f8)  text (16)  ncol, nrow, ilay (i4)
    ncol * nrow values (f8), row after row

Budget file, in gwf_ot_flow order (gwf.f90:911-936):

* STO-SS, then STO-SY when any cell is convertible (sto_save_model_flows,
  gwf-sto.f90:602-640): ubdsv1 of the whole grid (record_array with a
  negative unit, Dis.f90:1478-1482);
* FLOW-JA-FACE (record_connection_array, DiscretizationBase.f90:955-969):
  ubdsv1 of the nja face flows as ncol = nja, nrow = 1, nlay = 1;
* each boundary package in bndlist order (WEL, DRN, RCH, CHD) as a list
  (save_print_model_flows, BoundaryPackage.f90:1753-1902): a ubdsv06
  header (InputOutput.f90:979-1027) and one ubdsvd entry per active bound
  (:1053-1071) holding the user node number, the bound's 1-based index and
  its rate.

ubdsv1 (InputOutput.f90:944-975)::

    kstp, kper (i4)  text (16)  ncol, nrow, -nlay (i4)
    imeth = 1 (i4)  delt, pertim, totim (f8)
    values (f8)

ubdsv06::

    kstp, kper (i4)  text (16)  ncol, nrow, -nlay (i4)
    imeth = 6 (i4)  delt, pertim, totim (f8)
    model, package, model, package (16 each)
    naux + 1 (i4)  [aux names (16 each)]  nlist (i4)
    nlist x (node (i4), node2 (i4), q (f8) [, aux (f8) ...])
"""
import struct

import numpy as np

TEXT_HEAD = "HEAD"
TEXT_STO_SS = "          STO-SS"      # gwf-sto.f90:28-29
TEXT_STO_SY = "          STO-SY"
TEXT_FLOWJA = "    FLOW-JA-FACE"      # DiscretizationBase.f90:964
TEXT_PACKAGE = {                      # the packages' text (gwf-wel.f90:37,
    "WEL": "             WEL",        # gwf-drn.f90:22, gwf-rch.f90:24-25,
    "DRN": "             DRN",        # gwf-chd.f90:23)
    "RCH": "             RCH",
    "RCHA": "            RCHA",       # RCH with READASARRAYS
    "CHD": "             CHD",
}
BNDLIST_ORDER = ("WEL", "DRN", "RCH", "CHD")

_ENTRY = np.dtype([("node", "<i4"), ("node2", "<i4"), ("q", "<f8")])


def text16(s):
    """A character(len=16) value: padded with blanks on the right, or cut
    to 16 characters, as Fortran assignment does."""
    b = s.encode("ascii")
    return b[:16].ljust(16, b" ")


def ulasav(buf, text, kstp, kper, pertim, totim, ilay):
    """Bytes of one ulasav record; buf is (nrow, ncol)."""
    buf = np.ascontiguousarray(buf, dtype="<f8")
    nrow, ncol = buf.shape
    return (struct.pack("<iidd", kstp, kper, pertim, totim) + text16(text)
            + struct.pack("<iii", ncol, nrow, ilay) + buf.tobytes())


def ubdsv1(kstp, kper, text, buff, ncol, nrow, nlay, delt, pertim, totim):
    """Bytes of one ubdsv1 record (imeth 1: a full array)."""
    buff = np.ascontiguousarray(buff, dtype="<f8").ravel()
    if buff.size != ncol * nrow * nlay:
        raise ValueError(f"ubdsv1: {buff.size} values for {ncol} x {nrow} x {nlay}")
    return (struct.pack("<ii", kstp, kper) + text16(text)
            + struct.pack("<iii", ncol, nrow, -nlay)
            + struct.pack("<iddd", 1, delt, pertim, totim) + buff.tobytes())


def ubdsv06(kstp, kper, text, model1, package1, model2, package2, ncol, nrow, nlay,
            nodes, nodes2, q, delt, pertim, totim, aux=None, auxnames=()):
    """Bytes of one ubdsv06 header and its ubdsvd entries (imeth 6: a
    list).  nodes / nodes2 are written as given (1-based)."""
    nodes = np.asarray(nodes)
    nlist = nodes.size
    naux = len(auxnames)
    out = [struct.pack("<ii", kstp, kper), text16(text),
           struct.pack("<iii", ncol, nrow, -nlay),
           struct.pack("<iddd", 6, delt, pertim, totim),
           text16(model1), text16(package1), text16(model2), text16(package2),
           struct.pack("<i", naux + 1)]
    out += [text16(a) for a in auxnames]
    out.append(struct.pack("<i", nlist))
    if naux == 0:
        e = np.empty(nlist, dtype=_ENTRY)
        e["node"], e["node2"], e["q"] = nodes, nodes2, q
        out.append(e.tobytes())
    else:
        dt = np.dtype([("node", "<i4"), ("node2", "<i4"), ("q", "<f8"),
                       ("aux", "<f8", (naux,))])
        e = np.empty(nlist, dtype=dt)
        e["node"], e["node2"], e["q"] = nodes, nodes2, q
        e["aux"] = np.asarray(aux, dtype=np.float64).reshape(nlist, naux)
        out.append(e.tobytes())
    return b"".join(out)


def head_records(heads, nlay, nrow, ncol, kstp, kper, pertim, totim, text=TEXT_HEAD):
    """The head file records of one save: a ulasav record per layer."""
    h = np.asarray(heads, dtype=np.float64).reshape(nlay, nrow, ncol)
    return b"".join(ulasav(h[k], text, kstp, kper, pertim, totim, k + 1)
                    for k in range(nlay))


def budget_records(kstp, kper, delt, pertim, totim, nlay, nrow, ncol, flowja,
                   strgss=None, strgsy=None, packages=(), model_name="GWF"):
    """The budget file records of one save, in gwf_ot_flow order.

    strgss / strgsy: storage rates per cell (None: no STO package / no
    convertible cell).  packages: (kind, package name, 0-based nodes,
    simvals) tuples; kind WEL, DRN, RCH, RCHA or CHD; written in bndlist
    order."""
    out = []
    if strgss is not None:
        out.append(ubdsv1(kstp, kper, TEXT_STO_SS, strgss, ncol, nrow, nlay, delt,
                          pertim, totim))
    if strgsy is not None:
        out.append(ubdsv1(kstp, kper, TEXT_STO_SY, strgsy, ncol, nrow, nlay, delt,
                          pertim, totim))
    fj = np.asarray(flowja, dtype=np.float64)
    out.append(ubdsv1(kstp, kper, TEXT_FLOWJA, fj, fj.size, 1, 1, delt, pertim, totim))

    def order(p):
        kind = "RCH" if p[0] == "RCHA" else p[0]
        return BNDLIST_ORDER.index(kind)

    for kind, pname, nodes, sim in sorted(packages, key=order):
        nodes = np.asarray(nodes, dtype=np.int64)
        out.append(ubdsv06(kstp, kper, TEXT_PACKAGE[kind], model_name, model_name,
                           model_name, pname, ncol, nrow, nlay, nodes + 1,
                           np.arange(1, nodes.size + 1), sim, delt, pertim, totim))
    return b"".join(out)


# --- readers ---------------------------------------------------------------

def read_heads(path):
    """[(kstp, kper, pertim, totim, text, ilay, array (nrow, ncol))]"""
    b = open(path, "rb").read()
    recs, o = [], 0
    while o < len(b):
        kstp, kper, pertim, totim = struct.unpack_from("<iidd", b, o)
        text = b[o + 24:o + 40].decode("ascii")
        ncol, nrow, ilay = struct.unpack_from("<iii", b, o + 40)
        o += 52
        a = np.frombuffer(b, dtype="<f8", count=ncol * nrow, offset=o).reshape(nrow, ncol)
        o += 8 * ncol * nrow
        recs.append((kstp, kper, pertim, totim, text, ilay, a.copy()))
    return recs


def read_budget(path):
    """[dict(kstp, kper, text, ncol, nrow, nlay, imeth, delt, pertim, totim,
    data | (model1, package1, model2, package2, auxnames, entries))]"""
    b = open(path, "rb").read()
    recs, o = [], 0
    while o < len(b):
        kstp, kper = struct.unpack_from("<ii", b, o)
        text = b[o + 8:o + 24].decode("ascii")
        ncol, nrow, nlay = struct.unpack_from("<iii", b, o + 24)
        imeth, delt, pertim, totim = struct.unpack_from("<iddd", b, o + 36)
        o += 64
        r = dict(kstp=kstp, kper=kper, text=text, ncol=ncol, nrow=nrow, nlay=nlay,
                 imeth=imeth, delt=delt, pertim=pertim, totim=totim)
        if imeth == 1:
            n = ncol * nrow * abs(nlay)
            r["data"] = np.frombuffer(b, dtype="<f8", count=n, offset=o).copy()
            o += 8 * n
        elif imeth == 6:
            names = [b[o + 16 * i:o + 16 * (i + 1)].decode("ascii") for i in range(4)]
            o += 64
            ndat, = struct.unpack_from("<i", b, o)
            o += 4
            aux = [b[o + 16 * i:o + 16 * (i + 1)].decode("ascii") for i in range(ndat - 1)]
            o += 16 * (ndat - 1)
            nlist, = struct.unpack_from("<i", b, o)
            o += 4
            dt = np.dtype([("node", "<i4"), ("node2", "<i4"), ("q", "<f8")]
                          + ([("aux", "<f8", (ndat - 1,))] if ndat > 1 else []))
            r["model1"], r["package1"], r["model2"], r["package2"] = names
            r["auxnames"] = aux
            r["entries"] = np.frombuffer(b, dtype=dt, count=nlist, offset=o).copy()
            o += dt.itemsize * nlist
        else:
            raise ValueError(f"unsupported imeth {imeth} at byte {o - 64}")
        recs.append(r)
    return recs
