"""Writes tests/golden/dcdata.csv from the reference's data/dcData.RData.

The file is a gzip'd R save() image in XDR serialization format version 2 ("RDX2\\nX\\n"): a
pairlist of (name, object); dcData is a data.frame (a list of REALSXP columns X, Y, Time with
names / class / row.names attributes). This script reads that format directly (only the item
types the file uses) and writes the three columns as CSV with full double precision. No R code is
run. Columns (R/data.R): X, Y in km from the western / southern boundary of Washington D.C.;
Time in minutes from the beginning of 2010.

Usage: python tests/golden/make_dcdata.py [path/to/dcData.RData]
"""
import gzip
import os
import struct
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
DEFAULT = "/root/reference/data/dcData.RData"

NILVALUE, REFSXP = 254, 255
SYMSXP, LISTSXP, CHARSXP, LGLSXP, INTSXP, REALSXP, STRSXP, VECSXP = 1, 2, 9, 10, 13, 14, 16, 19
GLOBALENV, EMPTYENV, BASEENV, MISSINGARG = 253, 242, 241, 251


class Reader:
    def __init__(self, data):
        self.d = data
        self.p = 0
        self.refs = []

    def int(self):
        v = struct.unpack(">i", self.d[self.p:self.p + 4])[0]
        self.p += 4
        return v

    def doubles(self, n):
        v = struct.unpack(">%dd" % n, self.d[self.p:self.p + 8 * n])
        self.p += 8 * n
        return list(v)

    def item(self):
        flags = self.int()
        typ = flags & 0xFF
        has_attr = bool(flags & (1 << 9))
        has_tag = bool(flags & (1 << 10))
        if typ == NILVALUE:
            return None
        if typ in (GLOBALENV, EMPTYENV, BASEENV, MISSINGARG):
            return ("env", typ)
        if typ == REFSXP:
            idx = flags >> 8
            if idx == 0:
                idx = self.int()
            return self.refs[idx - 1]
        if typ == SYMSXP:
            name = self.item()
            sym = ("sym", name)
            self.refs.append(sym)
            return sym
        if typ == CHARSXP:
            n = self.int()
            if n == -1:
                return None
            s = self.d[self.p:self.p + n].decode("utf-8")
            self.p += n
            return s
        if typ == LISTSXP:
            attr = self.item() if has_attr else None
            tag = self.item() if has_tag else None
            car = self.item()
            cdr = self.item()
            return ("list", tag, car, cdr, attr)
        if typ in (LGLSXP, INTSXP):
            n = self.int()
            v = [self.int() for _ in range(n)]
        elif typ == REALSXP:
            n = self.int()
            v = self.doubles(n)
        elif typ == STRSXP:
            n = self.int()
            v = [self.item() for _ in range(n)]
        elif typ == VECSXP:
            n = self.int()
            v = [self.item() for _ in range(n)]
        else:
            raise ValueError("unsupported SEXP type %d at byte %d" % (typ, self.p))
        attr = self.item() if has_attr else None
        return ("vec", typ, v, attr)


def pairlist_items(pl):
    out = []
    while pl is not None:
        _, tag, car, cdr, _ = pl
        out.append((tag[1] if tag else None, car))
        pl = cdr
    return out


def main():
    path = sys.argv[1] if len(sys.argv) > 1 else DEFAULT
    raw = gzip.open(path).read()
    assert raw[:7] == b"RDX2\nX\n", raw[:7]
    r = Reader(raw[7:])
    assert r.int() == 2  # serialization version
    r.int()
    r.int()
    objs = dict(pairlist_items(r.item()))
    df = objs["dcData"]
    _, _, cols, attr = df
    names = dict(pairlist_items(attr))["names"][2]
    data = {nm: col[2] for nm, col in zip(names, cols)}
    out = os.path.join(HERE, "dcdata.csv")
    with open(out, "w") as f:
        f.write("X,Y,Time\n")
        for x, y, t in zip(data["X"], data["Y"], data["Time"]):
            f.write("%r,%r,%r\n" % (x, y, t))
    print("wrote", out, len(data["X"]), "rows; columns", names)


if __name__ == "__main__":
    main()
