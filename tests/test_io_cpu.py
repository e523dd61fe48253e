"""The file formats of the reference's workflow (segclip.io): shapefile-derived
WKT layers (polygons with holes, multipolygons, the repeated closing vertex
dropped), bin/seg's count-first input files and its result files, both
round-tripping every double at 17 significant digits."""
import os

import numpy as np

import segclip.io as IO
import segclip.synth as S

WKT = """1,POLYGON ((0 0, 4 0, 4 3, 0 3, 0 0), (1 1, 2 1, 2 2, 1 1))
2,MULTIPOLYGON (((10 10, 11 10, 10.5 11.5e0, 10 10)), ((-1.5e-3 -2, -1 -2, -1 -1, -1.5e-3 -2)))
"""


def test_parse_wkt():
    polys = IO.parse_wkt(WKT)
    assert len(polys) == 3
    assert [len(p) for p in polys] == [2, 1, 1]
    assert np.array_equal(polys[0][0], [[0, 0], [4, 0], [4, 3], [0, 3]])
    assert np.array_equal(polys[0][1], [[1, 1], [2, 1], [2, 2]])
    assert np.array_equal(polys[1][0], [[10, 10], [11, 10], [10.5, 11.5]])
    assert np.array_equal(polys[2][0], [[-1.5e-3, -2], [-1, -2], [-1, -1]])


def test_wkt2seg_and_count_first_round_trip(tmp_path):
    src = tmp_path / "layer.wkt"
    src.write_text(WKT)
    paths = IO.wkt2seg(str(src), str(tmp_path / "out"))
    assert [os.path.basename(p) for p in paths] == ["layer_0.txt", "layer_1.txt", "layer_2.txt"]
    assert np.array_equal(IO.read_count_first(paths[0]), [[0, 0], [4, 0], [4, 3], [0, 3]])
    ring = S.harmonic(1000, 7)
    p = str(tmp_path / "h.txt")
    IO.write_count_first(p, ring)
    assert np.array_equal(IO.read_count_first(p).view(np.uint64), ring.view(np.uint64))


def test_results_round_trip(tmp_path):
    rings = [S.star(50, 3), S.star(7, 4, 2.0, 2.0)]
    p = str(tmp_path / "r.txt")
    IO.write_results(p, rings)
    back = IO.read_results(p)
    assert len(back) == 2
    for a, b in zip(rings, back):
        assert np.array_equal(a.view(np.uint64), b.view(np.uint64))
    # the reference's default precision
    IO.write_results(p, rings, precision=6)
    back = IO.read_results(p)
    assert np.array_equal(back[0], [[float("%.6g" % x), float("%.6g" % y)] for x, y in rings[0]])
