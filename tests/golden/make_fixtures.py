"""Writes the golden fixtures in this directory from the reference's own tests.

Run in a container where the reference checkout is present:

    python tests/golden/make_fixtures.py /root/reference

It reads data only: the input vectors and expected outputs the reference's
gtests hold, at the cited lines, and writes them as JSON. No reference code is
run or imported. The JSON files are committed; the tests read only the JSON.
"""

import json
import os
import re
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
NUM = r"-?\d+(?:\.\d*)?(?:[eE][-+]?\d+)?"


def lines(ref, rel, a, b):
    with open(os.path.join(ref, rel)) as f:
        return "".join(f.readlines()[a - 1 : b])


def pairs(text):
    return [[float(u), float(v)] for u, v in re.findall(r"\{\s*(%s)\s*,\s*(%s)\s*\}" % (NUM, NUM), text)]


def brace_lists(text):
    """Every innermost {...} list of numbers or true/false, in order."""
    out = []
    for body in re.findall(r"\{([^{}]*)\}", text):
        toks = [t.strip() for t in body.split(",") if t.strip()]
        if not toks:
            out.append([])
            continue
        vals = []
        for t in toks:
            if t in ("true", "false"):
                vals.append(t == "true")
            elif re.fullmatch(NUM, t):
                vals.append(float(t) if ("." in t or "e" in t.lower()) else int(t))
            else:
                vals = None
                break
        if vals is not None:
            out.append(vals)
    return out


def quadtree_cases(ref):
    src = "cpp/tests/index/point_quadtree_test.cu"
    cases = [
        # test_empty, :69-80
        dict(name="test_empty", source=src + ":69-80", points=[], vertex_1=[0.0, 0.0],
             vertex_2=[1.0, 1.0], scale=1.0, max_depth=1, max_size=1,
             key=[], level=[], is_internal_node=[], length=[], offset=[]),
        # test_single, :82-103
        dict(name="test_single", source=src + ":82-103", points=[[0.45, 0.45]],
             vertex_1=[0.0, 0.0], vertex_2=[1.0, 1.0], scale=1.0, max_depth=1, max_size=1,
             key=[0], level=[0], is_internal_node=[False], length=[1], offset=[0]),
        # test_two, :105-126
        dict(name="test_two", source=src + ":105-126", points=[[0.45, 0.45], [1.45, 1.45]],
             vertex_1=[0.0, 0.0], vertex_2=[2.0, 2.0], scale=1.0, max_depth=1, max_size=1,
             key=[0, 3], level=[0, 0], is_internal_node=[False, False], length=[1, 1],
             offset=[0, 1]),
        # test_all_lowest_level_quads, :128-148 (scale -1 is clamped)
        dict(name="test_all_lowest_level_quads", source=src + ":128-148",
             points=[[-100.0, -100.0], [100.0, 100.0]], vertex_1=[-1000.0, -1000.0],
             vertex_2=[1000.0, 1000.0], scale=-1.0, max_depth=2, max_size=1,
             key=[3, 12, 15], level=[0, 1, 1], is_internal_node=[True, False, False],
             length=[2, 1, 1], offset=[1, 0, 1]),
    ]
    text = lines(ref, src, 159, 221)
    pts = pairs(text)[:71]
    assert len(pts) == 71, len(pts)
    tail = brace_lists(text[text.index("vertex_1,") :])
    key, level, internal, length, offset = tail[-5:]
    assert len(key) == 14 and len(internal) == 14
    cases.append(dict(name="test_small", source=src + ":150-221", points=pts,
                      vertex_1=[0.0, 0.0], vertex_2=[8.0, 8.0], scale=1.0, max_depth=3,
                      max_size=12, key=key, level=level, is_internal_node=internal,
                      length=length, offset=offset))
    return cases


def join_pip_small(ref):
    src = "cpp/tests/join/quadtree_point_in_polygon_test_small.cu"
    pts = pairs(lines(ref, src, 56, 94))
    assert len(pts) == 71, len(pts)
    poly_text = lines(ref, src, 98, 123)
    lists = brace_lists(poly_text)
    part_offsets, ring_offsets = lists[1], lists[2]
    verts = pairs(poly_text)
    assert len(verts) == 19, len(verts)
    exp_join = brace_lists(lines(ref, src, 143, 144))
    exp_pip = brace_lists(lines(ref, src, 160, 163))
    return dict(
        source=src + ":44-166",
        points=pts, v_min=[0.0, 0.0], v_max=[8.0, 8.0], scale=1.0, max_depth=3, max_size=12,
        geometry_offsets=lists[0], part_offsets=part_offsets, ring_offsets=ring_offsets,
        vertices=verts,
        expected_join_bbox_offset=exp_join[0], expected_join_quad_offset=exp_join[1],
        expected_polygon_index=exp_pip[0], expected_point_index=exp_pip[1])


def pip_kats():
    """Planar cases of cpp/tests/point_in_polygon/point_in_polygon_test.cu.

    Each case: points, part_offsets (rings per polygon), ring_offsets, vertices,
    expected per-point bitmask (bit j = inside polygon j)."""
    src = "cpp/tests/point_in_polygon/point_in_polygon_test.cu"
    sq8 = [[-2.0, 0.0], [2.0, 0.0], [0.0, -2.0], [0.0, 2.0],
           [-0.5, 0.0], [0.5, 0.0], [0.0, -0.5], [0.0, 0.5]]
    cases = [
        dict(name="OnePolygonOneRing", source=src + ":108-123", points=sq8,
             part_offsets=[0, 1], ring_offsets=[0, 5],
             vertices=[[-1.0, -1.0], [1.0, -1.0], [1.0, 1.0], [-1.0, 1.0], [-1.0, -1.0]],
             expected=[0, 0, 0, 0, 1, 1, 1, 1]),
        dict(name="OnePolygonOneRingUnclosed", source=src + ":148-164", points=sq8,
             part_offsets=[0, 1], ring_offsets=[0, 4],
             vertices=[[-1.0, -1.0], [1.0, -1.0], [1.0, 0.0], [1.0, 1.0]],
             expected=[0, 0, 0, 0, 0, 1, 1, 0]),
        dict(name="TwoPolygonsOneRingEach", source=src + ":182-210", points=sq8,
             part_offsets=[0, 1, 2], ring_offsets=[0, 5, 10],
             vertices=[[-1.0, -1.0], [-1.0, 1.0], [1.0, 1.0], [1.0, -1.0], [-1.0, -1.0],
                       [0.0, 1.0], [1.0, 0.0], [0.0, -1.0], [-1.0, 0.0], [0.0, 1.0]],
             expected=[0, 0, 0, 0, 3, 3, 3, 3]),
        dict(name="OnePolygonTwoRings", source=src + ":234-253",
             points=[[0.0, 0.0], [-0.4, 0.0], [-0.6, 0.0], [0.0, 0.4], [0.0, -0.6]],
             part_offsets=[0, 2], ring_offsets=[0, 5, 10],
             vertices=[[-1.0, -1.0], [1.0, -1.0], [1.0, 1.0], [-1.0, 1.0], [-1.0, -1.0],
                       [-0.5, -0.5], [-0.5, 0.5], [0.5, 0.5], [0.5, -0.5], [-0.5, -0.5]],
             expected=[0, 0, 1, 0, 1]),
        dict(name="EdgesOfSquare", source=src + ":279-296", points=[[0.0, 0.0]],
             part_offsets=[0, 1, 2, 3, 4], ring_offsets=[0, 5, 10, 15, 20],
             vertices=[[-1.0, -1.0], [0.0, -1.0], [0.0, 1.0], [-1.0, 1.0], [-1.0, -1.0],
                       [0.0, -1.0], [1.0, -1.0], [1.0, 1.0], [0.0, 1.0], [0.0, -1.0],
                       [-1.0, -1.0], [-1.0, 0.0], [1.0, 0.0], [1.0, -1.0], [-1.0, 1.0],
                       [-1.0, 0.0], [-1.0, 1.0], [1.0, 1.0], [1.0, 0.0], [-1.0, 0.0]],
             expected=[0]),
        dict(name="CornersOfSquare", source=src + ":298-315", points=[[0.0, 0.0]],
             part_offsets=[0, 1, 2, 3, 4], ring_offsets=[0, 5, 10, 15, 20],
             vertices=[[-1.0, -1.0], [-1.0, 0.0], [0.0, 0.0], [0.0, -1.0], [-1.0, -1.0],
                       [-1.0, 0.0], [-1.0, 1.0], [0.0, 1.0], [-1.0, 0.0], [-1.0, 0.0],
                       [0.0, -1.0], [0.0, 0.0], [1.0, 0.0], [1.0, -1.0], [0.0, -1.0],
                       [0.0, 0.0], [0.0, 1.0], [1.0, 1.0], [1.0, 0.0], [0.0, 0.0]],
             expected=[0]),
        dict(name="SelfClosingLoopLeftEdgeMissing", source=src + ":407-418",
             points=[[-2.0, 0.0], [0.0, 0.0], [2.0, 0.0]], part_offsets=[0, 1],
             ring_offsets=[0, 4], vertices=[[-1.0, 1.0], [1.0, 1.0], [1.0, -1.0], [-1.0, -1.0]],
             expected=[0, 1, 0]),
        dict(name="SelfClosingLoopRightEdgeMissing", source=src + ":420-429",
             points=[[-2.0, 0.0], [0.0, 0.0], [2.0, 0.0]], part_offsets=[0, 1],
             ring_offsets=[0, 4], vertices=[[1.0, -1.0], [-1.0, -1.0], [-1.0, 1.0], [1.0, 1.0]],
             expected=[0, 1, 0]),
        dict(name="ContainsButCollinearWithBoundary", source=src + ":431-441",
             points=[[0.5, 0.5]], part_offsets=[0, 1], ring_offsets=[0, 9],
             vertices=[[0.0, 0.0], [0.0, 1.0], [1.0, 1.0], [1.0, 0.5], [1.5, 0.5], [1.5, 1.0],
                       [2.0, 1.0], [2.0, 0.0], [0.0, 0.0]],
             expected=[1]),
    ]
    # 31PolygonSupport (:366-405): 31 squares generated by index functors, points
    # (0,0) inside all and (2,0) inside none.
    xs = [(-1.0, -1.0, 1.0, 1.0, -1.0)[i % 5] for i in range(31 * 5)]
    ys = [(-1.0, 1.0, 1.0, -1.0, -1.0)[i % 5] for i in range(31 * 5)]
    cases.append(dict(name="31PolygonSupport", source=src + ":341-405",
                      points=[[0.0, 0.0], [2.0, 0.0]], part_offsets=list(range(32)),
                      ring_offsets=[5 * i for i in range(32)],
                      vertices=[[a, b] for a, b in zip(xs, ys)],
                      expected=[(1 << 31) - 1, 0]))
    return cases


def pairwise_pip_kats():
    """cpp/tests/point_in_polygon/pairwise_point_in_polygon_test.cu: point i is
    tested against polygon i. Cases that the reference runs as several calls
    against one polygon range are written as one call with that polygon repeated.

    Each case: points, part_offsets, ring_offsets, vertices, expected (0/1 per pair)."""
    src = "cpp/tests/point_in_polygon/pairwise_point_in_polygon_test.cu"
    sq = [[-1.0, -1.0], [1.0, -1.0], [1.0, 1.0], [-1.0, 1.0], [-1.0, -1.0]]
    sq8 = [[-2.0, 0.0], [2.0, 0.0], [0.0, -2.0], [0.0, 2.0],
           [-0.5, 0.0], [0.5, 0.0], [0.0, -0.5], [0.0, 0.5]]

    def repeat(vertex_lists, counts):
        """polygons given as lists of rings (lists of vertices), in pair order"""
        parts, rings, verts = [0], [0], []
        for poly in vertex_lists:
            for ring in poly:
                verts += ring
                rings.append(len(verts))
            parts.append(len(rings) - 1)
        return parts, rings, verts

    cases = []
    p, r, v = repeat([[sq]] * 8, None)
    cases.append(dict(name="OnePolygonOneRing", source=src + ":75-114", points=sq8,
                      part_offsets=p, ring_offsets=r, vertices=v,
                      expected=[0, 0, 0, 0, 1, 1, 1, 1]))
    a = [[-1.0, -1.0], [-1.0, 1.0], [1.0, 1.0], [1.0, -1.0], [-1.0, -1.0]]
    b = [[0.0, 1.0], [1.0, 0.0], [0.0, -1.0], [-1.0, 0.0], [0.0, 1.0]]
    p, r, v = repeat([[a], [b], [a], [b]], None)
    cases.append(dict(name="TwoPolygonsOneRingEach", source=src + ":116-164",
                      points=sq8[:4], part_offsets=p, ring_offsets=r, vertices=v,
                      expected=[0, 0, 0, 0]))
    hole = [[-0.5, -0.5], [-0.5, 0.5], [0.5, 0.5], [0.5, -0.5], [-0.5, -0.5]]
    p, r, v = repeat([[sq, hole]] * 5, None)
    cases.append(dict(name="OnePolygonTwoRings", source=src + ":166-209",
                      points=[[0.0, 0.0], [-0.4, 0.0], [-0.6, 0.0], [0.0, 0.4], [0.0, -0.6]],
                      part_offsets=p, ring_offsets=r, vertices=v, expected=[0, 0, 1, 0, 1]))
    cases.append(dict(name="EdgesOfSquare", source=src + ":211-226", points=[[0.0, 0.0]] * 4,
                      part_offsets=[0, 1, 2, 3, 4], ring_offsets=[0, 5, 10, 15, 20],
                      vertices=[[-1.0, -1.0], [0.0, -1.0], [0.0, 1.0], [-1.0, 1.0],
                                [-1.0, -1.0], [0.0, -1.0], [1.0, -1.0], [1.0, 1.0], [0.0, 1.0],
                                [0.0, -1.0], [-1.0, -1.0], [-1.0, 0.0], [1.0, 0.0], [1.0, -1.0],
                                [-1.0, 1.0], [-1.0, 0.0], [-1.0, 1.0], [1.0, 1.0], [1.0, 0.0],
                                [-1.0, 0.0]],
                      expected=[0, 0, 0, 0]))
    cases.append(dict(name="CornersOfSquare", source=src + ":228-244", points=[[0.0, 0.0]] * 4,
                      part_offsets=[0, 1, 2, 3, 4], ring_offsets=[0, 5, 10, 15, 20],
                      vertices=[[-1.0, -1.0], [-1.0, 0.0], [0.0, 0.0], [0.0, -1.0],
                                [-1.0, -1.0], [-1.0, 0.0], [-1.0, 1.0], [0.0, 1.0], [-1.0, 0.0],
                                [-1.0, 0.0], [0.0, -1.0], [0.0, 0.0], [1.0, 0.0], [1.0, -1.0],
                                [0.0, -1.0], [0.0, 0.0], [0.0, 1.0], [1.0, 1.0], [1.0, 0.0],
                                [0.0, 0.0]],
                      expected=[0, 0, 0, 0]))
    # 32PolygonSupport (:246-323): 32 index-generated squares, points alternating
    # (0,0) (inside) and (2,0) (outside)
    xs = [(-1.0, -1.0, 1.0, 1.0, -1.0)[i % 5] for i in range(32 * 5)]
    ys = [(-1.0, 1.0, 1.0, -1.0, -1.0)[i % 5] for i in range(32 * 5)]
    cases.append(dict(name="32PolygonSupport", source=src + ":246-323",
                      points=[[0.0, 0.0], [2.0, 0.0]] * 16, part_offsets=list(range(33)),
                      ring_offsets=[5 * i for i in range(33)],
                      vertices=[[a, b] for a, b in zip(xs, ys)], expected=[1, 0] * 16))
    return cases


def contains_properly_kats(ref):
    """python/cuspatial/cuspatial/tests/binpreds/test_contains_properly.py: one
    point against one polygon (a closed single ring) and the value
    GeoSeries.contains_properly returns for it. The points and polygons are read
    from the cited parametrize lists; `expected` is what each test asserts of
    the result (test_float_precision_limits_failures asserts that none holds).
    Multi-row cases list their rows and the expected per-point values."""
    src = "python/cuspatial/cuspatial/tests/binpreds/test_contains_properly.py"
    pt_re = r"Point\(\[\s*(%s),\s*(%s)\]\)" % (NUM, NUM)
    poly_re = r"Polygon\(\[((?:\[\s*%s,\s*%s\],?\s*)+)\]\)" % (NUM, NUM)

    def ring(text):
        return [[float(a), float(b)] for a, b in re.findall(r"\[\s*(%s),\s*(%s)\]" % (NUM, NUM),
                                                             text)]

    cases = []

    def single(a, b, forced=None):
        text = lines(ref, src, a, b)
        named = {"clockwiseTriangle": [[0, 0], [0, 1], [1, 1], [0, 0]],
                 "clockwiseSquare": [[-0.5, -0.5], [-0.5, 0.5], [0.5, 0.5], [0.5, -0.5],
                                     [-0.5, -0.5]]}
        for m in re.finditer(r"\[\s*%s,\s*(%s|clockwiseTriangle|clockwiseSquare),\s*(True|False),?\s*\]"
                             % (pt_re, poly_re), text):
            x, y, poly_text, verts, expect = m.group(1), m.group(2), m.group(3), m.group(4), m.group(5)
            poly = named[poly_text] if poly_text in named else ring(verts)
            e = (expect == "True") if forced is None else forced
            cases.append(dict(source="%s:%d-%d" % (src, a, b), points=[[float(x), float(y)]],
                              polygons=[[[float(u), float(v)] for u, v in poly]],
                              expected=[e]))

    single(85, 100, forced=False)  # test_float_precision_limits_failures: `not np.any(got)`
    single(113, 127)  # test_float_precision_limits
    single(152, 184)  # test_point_in_polygon
    tri_ccw = [[0.0, 0.0], [1.0, 0.0], [1.0, 1.0], [0.0, 0.0]]
    tri_cw = [[0.0, 0.0], [0.0, 1.0], [1.0, 1.0], [0.0, 0.0]]
    cases.append(dict(source=src + ":196-203", points=[[0.0, 0.0], [0.0, 0.0]],
                      polygons=[tri_ccw], expected=[False, False]))
    cases.append(dict(source=src + ":206-218", points=[[0.0, 0.0]],
                      polygons=[tri_ccw, [[-2.0, -2.0], [-2.0, 2.0], [2.0, 2.0], [-2.0, -2.0]]],
                      expected=[False]))
    # test_ten_fixed_points (:221-254): rows pair up; (0,0) is a vertex of every
    # triangle, (0.5, 0.25) lies inside the counter-clockwise one only
    cases.append(dict(source=src + ":221-254",
                      points=[[0.0, 0.0]] * 2 + [[0.5, 0.25]] * 6 + [[0.0, 0.0]] * 2,
                      polygons=[tri_ccw] * 5 + [tri_cw] * 5,
                      expected=[False, False, True, True, True, False, False, False, False,
                                False]))
    return cases


def nearest_linestring_small(ref):
    """cpp/tests/join/quadtree_point_to_nearest_linestring_test_small.cu:36-209: 71
    points in [0, 8)^2, a depth-3 quadtree (scale 1, max_size 12), 4 closed
    linestrings with boxes expanded by 2; the expected join pairs and, per point
    position, the nearest linestring and its distance (float64 branch)."""
    src = "cpp/tests/join/quadtree_point_to_nearest_linestring_test_small.cu"
    points = pairs(lines(ref, src, 48, 84))
    verts = pairs(lines(ref, src, 96, 117))

    def ints(a, b):
        return [int(v) for v in re.findall(r"\d+", lines(ref, src, a, b).split("(", 1)[1])]

    dist_text = lines(ref, src, 173, 192)
    return {
        "source": src + ":36-209",
        "points": points,
        "bbox": [0.0, 8.0, 0.0, 8.0], "scale": 1.0, "max_depth": 3, "max_size": 12,
        "linestring_offsets": [0, 4, 10, 14, 19],
        "vertices": verts,
        "expansion_radius": 2.0,
        "expected_join_linestring_index": ints(138, 139),
        "expected_join_quad_index": ints(140, 141),
        "expected_point_index": ints(196, 199),
        "expected_linestring_index": ints(201, 204),
        "expected_distance": [float(v) for v in re.findall(NUM, dist_text.split("make_device_vector<T>(", 1)[1])],
    }


def main():
    ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    with open(os.path.join(HERE, "quadtree_cases.json"), "w") as f:
        json.dump(quadtree_cases(ref), f, indent=1)
    with open(os.path.join(HERE, "join_pip_small.json"), "w") as f:
        json.dump(join_pip_small(ref), f, indent=1)
    with open(os.path.join(HERE, "pip_kats.json"), "w") as f:
        json.dump(pip_kats(), f, indent=1)
    with open(os.path.join(HERE, "pairwise_pip_kats.json"), "w") as f:
        json.dump(pairwise_pip_kats(), f, indent=1)
    with open(os.path.join(HERE, "nearest_linestring_small.json"), "w") as f:
        json.dump(nearest_linestring_small(ref), f, indent=1)
    with open(os.path.join(HERE, "contains_properly_kats.json"), "w") as f:
        json.dump(contains_properly_kats(ref), f, indent=1)
    print("fixtures written to", HERE)


if __name__ == "__main__":
    main()
