"""Writes the hand-derived known-answer fixtures of tests/golden/.

Nothing here runs the oracle or the library: every expected pair list below is
worked out by hand from the pair rule of the reference broad phase
(closed overlap on all three axes, src/collision.cu:38-42 and
src/cpu/sweep.cpp:32-37; is_valid_pair, src/aabb.cu:12-40; covertex,
src/collision.cu:44-48), with the reasoning in the comment of each case.
Ids are input positions.  Run:  python tests/golden/make_fixtures.py
"""
import json
import math
import os
import struct

HERE = os.path.dirname(os.path.abspath(__file__))


def V(v):
    return [v, -v - 1, -v - 1]


def E(a, b):
    return [a, b, -a - 1]


def F(a, b, c):
    return [a, b, c]


def box(lo, hi, vids):
    return {"min": list(lo), "max": list(hi), "vids": vids}


U0, U1 = (0.0, 0.0, 0.0), (1.0, 1.0, 1.0)
NEXT1 = math.nextafter(1.0, 2.0)

CASES = {
    # x intervals [0,1] and [1,2] share x = 1 (closed intervals) -> pair
    "touch_x": ([box(U0, U1, V(0)), box((1, 0, 0), (2, 1, 1), F(1, 2, 3))], [[0, 1]]),
    "touch_y": ([box(U0, U1, V(0)), box((0, 1, 0), (1, 2, 1), F(1, 2, 3))], [[0, 1]]),
    "touch_z": ([box(U0, U1, V(0)), box((0, 0, 1), (1, 1, 2), F(1, 2, 3))], [[0, 1]]),
    # one ulp above 1.0 on each axis: disjoint -> no pair
    "gap_x": ([box(U0, U1, V(0)), box((NEXT1, 0, 0), (2, 1, 1), F(1, 2, 3))], []),
    "gap_y": ([box(U0, U1, V(0)), box((0, NEXT1, 0), (1, 2, 1), F(1, 2, 3))], []),
    "gap_z": ([box(U0, U1, V(0)), box((0, 0, NEXT1), (1, 1, 2), F(1, 2, 3))], []),
    # identical boxes of a vertex and a face -> pair
    "identical": ([box(U0, U1, V(0)), box(U0, U1, F(1, 2, 3))], [[0, 1]]),
    # vertex box strictly inside a face box -> pair
    "nested": ([box((0, 0, 0), (10, 10, 10), F(1, 2, 3)), box((4, 4, 4), (5, 5, 5), V(0))], [[0, 1]]),
    # -0.0 >= +0.0 holds in IEEE, so max.x = -0.0 touches min.x = +0.0 -> pair
    "signed_zero": ([box((-1, -1, -1), (-0.0, 1, 1), V(0)), box((0.0, 0, 0), (1, 1, 1), F(1, 2, 3))],
                    [[0, 1]]),
    # six identical boxes: V0 V1 E0 E1 F0 F1 with no shared ids; valid pairs are
    # V-F (4 of them) and E-E (1); VV, VE, EF, FF are filtered out
    "types": ([box(U0, U1, V(10)), box(U0, U1, V(11)), box(U0, U1, E(20, 21)), box(U0, U1, E(22, 23)),
               box(U0, U1, F(30, 31, 32)), box(U0, U1, F(33, 34, 35))],
              [[0, 4], [0, 5], [1, 4], [1, 5], [2, 3]]),
    # same box for all; V{7} - F{7,8,9} share 7 (excluded); V{7} - F{1,2,3} pair;
    # E{7,5} - E{5,6} share 5; E{5,6} - E{4,6} share 6; E{7,5} - E{4,6} pair
    "shared_vertex": ([box(U0, U1, V(7)), box(U0, U1, F(7, 8, 9)), box(U0, U1, F(1, 2, 3)),
                       box(U0, U1, E(7, 5)), box(U0, U1, E(5, 6)), box(U0, U1, E(4, 6))],
                      [[0, 2], [3, 5]]),
    # padding convention: E{1,2,-2} and E{3,4,-4} have different pads -> pair
    # (a pad of -1 for both would wrongly look like a shared id); V{5,-6,-6}
    # and F{6,7,8} share no real id -> pair; E{3,9,-4} shares 3 with E{3,4,-4}
    "padding": ([box(U0, U1, E(1, 2)), box(U0, U1, E(3, 4)), box(U0, U1, V(5)), box(U0, U1, F(6, 7, 8)),
                 box(U0, U1, E(3, 9))],
                [[0, 1], [0, 4], [2, 3]]),
    # equal sweep-axis minima, zero-width x extents included: every V-F pair
    # overlaps (all x intervals contain 0.5, y/z identical)
    "equal_minima": ([box((0.5, 0, 0), (0.6, 1, 1), V(0)), box((0.5, 0, 0), (0.5, 1, 1), F(1, 2, 3)),
                      box((0.5, 0, 0), (0.9, 1, 1), F(4, 5, 6)), box((0.5, 0, 0), (0.5, 1, 1), V(7))],
                     [[0, 1], [0, 2], [1, 3], [2, 3]]),
    # x: A=[0,1], B=[2,3], C=[0.5,0.6]; A-C overlap, B is past A's chain end
    "chain_break": ([box((0, 0, 0), (1, 1, 1), V(0)), box((2, 0, 0), (3, 1, 1), F(1, 2, 3)),
                     box((0.5, 0, 0), (0.6, 1, 1), F(4, 5, 6))],
                    [[0, 2]]),
    # overlap on x and z only (y disjoint) -> no pair
    "yz_miss": ([box((0, 0, 0), (1, 1, 1), V(0)), box((0, 2, 0), (1, 3, 1), F(1, 2, 3))], []),
    # a long box over a run of short ones: the long vertex overlaps the three
    # faces that lie inside its x range and not the one beyond it
    "long_chain": ([box((0, 0, 0), (10, 1, 1), V(0))]
                   + [box((float(k), 0, 0), (k + 0.5, 1, 1), F(3 * k + 1, 3 * k + 2, 3 * k + 3)) for k in (1, 4, 9, 11)],
                   [[0, 1], [0, 2], [0, 3]]),
    "empty": ([], []),
    "single": ([box(U0, U1, V(0))], []),
}


def f32_bits(x):
    return struct.unpack("<I", struct.pack("<f", x))[0]


# FP32 outward rounding rule of the reference (src/aabb.cu:56-61): round to
# nearest float f, then lo = nextafterf(f, f - FLT_MAX), hi = nextafterf(f, f + FLT_MAX).
# Expected bit patterns derived by hand from IEEE-754 binary32.
ROUNDING = [
    # value, lo bits, hi bits
    (1.0, 0x3F7FFFFF, 0x3F800001),
    (-1.0, 0xBF800001, 0xBF7FFFFF),
    (0.1, 0x3DCCCCCC, 0x3DCCCCCE),        # (float)0.1 = 0x3DCCCCCD
    (0.0, 0x80000001, 0x00000001),        # +-denorm_min around zero
    (-0.0, 0x80000001, 0x00000001),
    (1e-45, 0x00000000, 0x00000002),      # rounds to denorm_min 0x00000001
    (3.4028234663852886e38, 0x7F7FFFFE, 0x7F800000),  # FLT_MAX: up is +inf
    (0.5 + 2.0 ** -25, 0x3EFFFFFF, 0x3F000001),  # tie rounds to even 0x3F000000
]


# A tetrahedron, written in the three mesh formats the reference reads
# (src/io.cu:36-45).  Vertex coordinates are exact in binary32 so that the
# float-typed PLY holds them unchanged.  Its edges, hand-derived from the four
# faces: {0,1} {0,2} {0,3} {1,2} {1,3} {2,3}, listed as igl::edges reads them
# off the column-major adjacency matrix (by larger index, then smaller).
TETRA_V = [[0.0, 0.0, 0.0], [1.5, 0.0, 0.25], [0.0, 2.0, -0.5], [0.75, 0.5, 3.0]]
TETRA_F = [[0, 1, 2], [0, 1, 3], [0, 2, 3], [1, 2, 3]]
TETRA_E = [[0, 1], [0, 2], [1, 2], [0, 3], [1, 3], [2, 3]]
# A ground-truth list in the reference's JSON form (src/groundtruth.cu:43-47)
# and pairs to check it against: (1, 3) and (2, 7) are found as given, (5, 0)
# only when both sides are taken as (min, max); (4, 4) is not in the list.
TRUTH = [[0, 5], [1, 3], [2, 7]]
FOUND = [[1, 3], [2, 7], [5, 0], [4, 4]]


def write_meshes():
    with open(os.path.join(HERE, "tetra.obj"), "w") as fh:
        fh.write("# tetrahedron\n")
        for v in TETRA_V:
            fh.write("v %r %r %r\n" % tuple(v))
        fh.write("vn 0 0 1\n")
        fh.write("f 1 2 3\n")
        fh.write("f 1/1/1 2/2/2 4/4/4\n")   # v/vt/vn form
        fh.write("f 1//1 3//1 4//1\n")      # v//vn form
        fh.write("f -3 -2 -1\n")            # relative indices: 2 3 4
    with open(os.path.join(HERE, "tetra_ascii.ply"), "w") as fh:
        fh.write("ply\nformat ascii 1.0\ncomment tetrahedron\nelement vertex 4\n"
                 "property float x\nproperty float y\nproperty float z\n"
                 "element face 4\nproperty list uchar int vertex_indices\nend_header\n")
        for v in TETRA_V:
            fh.write("%r %r %r\n" % tuple(v))
        for f in TETRA_F:
            fh.write("3 %d %d %d\n" % tuple(f))
    with open(os.path.join(HERE, "tetra_binary.ply"), "wb") as fh:
        fh.write(b"ply\nformat binary_little_endian 1.0\nelement vertex 4\n"
                 b"property double x\nproperty double y\nproperty double z\nproperty uchar red\n"
                 b"element face 4\nproperty list uchar uint vertex_indices\nend_header\n")
        for k, v in enumerate(TETRA_V):
            fh.write(struct.pack("<dddB", v[0], v[1], v[2], 10 * k))
        for f in TETRA_F:
            fh.write(struct.pack("<BIII", 3, *f))
    with open(os.path.join(HERE, "mesh_expected.json"), "w") as fh:
        json.dump({"V": TETRA_V, "F": TETRA_F, "E": TETRA_E}, fh, indent=1)
    with open(os.path.join(HERE, "truth_tiny.json"), "w") as fh:
        json.dump(TRUTH, fh)
    with open(os.path.join(HERE, "recall_expected.json"), "w") as fh:
        json.dump({"pairs": FOUND, "as_given": [2, 3], "canonical": [3, 3]}, fh)


def main():
    write_meshes()
    out = {name: {"boxes": b, "pairs": p} for name, (b, p) in CASES.items()}
    with open(os.path.join(HERE, "known_answers.json"), "w") as fh:
        json.dump(out, fh, indent=1)
    with open(os.path.join(HERE, "rounding.json"), "w") as fh:
        json.dump([{"value": v, "lo_bits": lo, "hi_bits": hi} for v, lo, hi in ROUNDING], fh, indent=1)


if __name__ == "__main__":
    main()
