"""Polygon file formats of the reference's workflow.

* Count-first vertex files, the input of bin/seg (src/clip.cpp:172-311): the
  vertex count n on the first line, then n lines "x,y", the closing vertex
  not repeated.
* Result files, the output of bin/seg (foster/polygon.h:576-585): "x y," per
  vertex and ";" after each component.  The reference prints 6 significant
  digits; write_results() defaults to 17, which round-trips every double.
* Shapefile-derived text (WKT): POLYGON ((x y, x y, ...), (hole ...)) and
  MULTIPOLYGON (((...)), ((...))), the layers the reference's data
  preparation reads (foster/readShapefile.cpp loadPolygonFromShapeFile /
  loadPolygonFromShapeFile2).  read_wkt() returns every polygon as a list of
  rings (outer ring first, then holes) without the repeated closing vertex.

Command line (converting a layer into bin/seg inputs, one file per polygon's
outer ring, as the run scripts expect):

    python -m segclip.io wkt2seg layer.wkt out_dir [--limit N] [--prefix name]
"""
import argparse
import os
import re

import numpy as np

_NUM = r"[-+]?(?:\d+\.?\d*|\.\d+)(?:[eE][-+]?\d+)?"


def read_count_first(path):
    """One ring from a count-first file as an (n, 2) float64 array."""
    with open(path) as f:
        tokens = f.read().replace(",", " ").split()
    if not tokens:
        raise ValueError("%s: empty file" % path)
    n = int(tokens[0])
    vals = np.array([float(t) for t in tokens[1:1 + 2 * n]], dtype=np.float64)
    if n < 3 or len(vals) != 2 * n:
        raise ValueError("%s: %d vertices announced, %d coordinates read" % (path, n, len(vals)))
    return vals.reshape(n, 2)


def write_count_first(path, ring, precision=17):
    ring = np.asarray(ring, dtype=np.float64).reshape(-1, 2)
    with open(path, "w") as f:
        f.write("%d\n" % len(ring))
        for x, y in ring:
            f.write("%.*g,%.*g\n" % (precision, x, precision, y))


def write_results(path, rings, precision=17):
    """Rings in bin/seg's output format ('x y,' per vertex, ';' after a component)."""
    with open(path, "w") as f:
        for r in rings:
            r = np.asarray(r, dtype=np.float64).reshape(-1, 2)
            for i, (x, y) in enumerate(r):
                f.write("%.*g %.*g%s\n" % (precision, x, precision, y, ";" if i == len(r) - 1 else ","))


def read_results(path):
    """Rings of a result file as a list of (n, 2) arrays."""
    rings, cur = [], []
    with open(path) as f:
        for line in f:
            line = line.strip()
            if not line:
                continue
            end = line.endswith(";")
            x, y = line.rstrip(",;").split()
            cur.append((float(x), float(y)))
            if end:
                rings.append(np.array(cur, dtype=np.float64))
                cur = []
    if cur:
        raise ValueError("%s: a component is not closed by ';'" % path)
    return rings


def _ring(text):
    vals = np.array([float(v) for v in re.findall(_NUM, text)], dtype=np.float64)
    if len(vals) % 2:
        raise ValueError("odd number of coordinates in a ring")
    r = vals.reshape(-1, 2)
    if len(r) >= 2 and np.array_equal(r[0], r[-1]):
        r = r[:-1]  # WKT repeats the first vertex; the clip's rings do not
    return r


def parse_wkt(text):
    """Every POLYGON / MULTIPOLYGON in text as a list of polygons, each a list
    of rings (outer first, then holes)."""
    out = []
    for m in re.finditer(r"(MULTIPOLYGON|POLYGON)\s*\(", text, flags=re.I):
        # the balanced parenthesised body that starts at the keyword's '('
        start = m.end() - 1
        depth, i = 0, start
        while i < len(text):
            if text[i] == "(":
                depth += 1
            elif text[i] == ")":
                depth -= 1
                if depth == 0:
                    break
            i += 1
        body = text[start + 1:i]
        if m.group(1).upper() == "POLYGON":
            out.append([_ring(r) for r in re.findall(r"\(([^()]*)\)", body)])
        else:
            for poly in re.findall(r"\(((?:\s*\([^()]*\)\s*,?)*)\)", body):
                out.append([_ring(r) for r in re.findall(r"\(([^()]*)\)", poly)])
    return out


def read_wkt(path, limit=-1):
    with open(path) as f:
        polys = parse_wkt(f.read())
    return polys if limit < 0 else polys[:limit]


def wkt2seg(src, out_dir, limit=-1, prefix=None):
    """Write the outer ring of every polygon of a WKT layer as a count-first
    file out_dir/<prefix>_<i>.txt; returns the paths."""
    os.makedirs(out_dir, exist_ok=True)
    prefix = prefix or os.path.splitext(os.path.basename(src))[0]
    paths = []
    for i, poly in enumerate(read_wkt(src, limit)):
        if not poly or len(poly[0]) < 3:
            continue
        p = os.path.join(out_dir, "%s_%d.txt" % (prefix, i))
        write_count_first(p, poly[0])
        paths.append(p)
    return paths


def main(argv=None):
    ap = argparse.ArgumentParser(prog="python -m segclip.io")
    sub = ap.add_subparsers(dest="cmd", required=True)
    w = sub.add_parser("wkt2seg", help="WKT layer -> count-first files for bin/seg (outer rings)")
    w.add_argument("src")
    w.add_argument("out_dir")
    w.add_argument("--limit", type=int, default=-1)
    w.add_argument("--prefix", default=None)
    a = ap.parse_args(argv)
    if a.cmd == "wkt2seg":
        for p in wkt2seg(a.src, a.out_dir, a.limit, a.prefix):
            print(p)


if __name__ == "__main__":
    main()
