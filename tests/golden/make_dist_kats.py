"""Writes dist_kats.json: the known-answer cases of the reference's segment
distance tests (ABM-3d/ABM-3d-cpp/test/src/testDistBetweenCells.cpp:143-506),
as inputs and hand-derived targets (tolerance 1e-8 in the reference).

Each case: r1, n1, half_l1, r2, n2, half_l2 and the target d (cell 1 -> 2), s, t.
nearest cases: r, n, half_l, q and the target cell-body coordinate.
"""
import json
import math
import os


def lines(r1, n1, r2, n2):
    """Closest points of two infinite lines (the reference test's distBetweenLines)."""
    r12 = [b - a for a, b in zip(r1, r2)]
    dot = lambda u, v: sum(x * y for x, y in zip(u, v))
    c = dot(n1, n2)
    s = (dot(r12, n1) - c * dot(r12, n2)) / (1.0 - c * c)
    t = (-dot(r12, n2) + c * dot(r12, n1)) / (1.0 - c * c)
    d = [r12[k] + t * n2[k] - s * n1[k] for k in range(3)]
    return d, s, t


def main():
    nearest = [
        dict(r=[2, 1, 0], n=[0.6, 0.8, 0], half_l=1, q=[0, 3, 0], target=0.4),
        dict(r=[6, 2, 2], n=[12 / 13, 5 / 13, 0], half_l=0.75, q=[7, 1, 2], target=0.5384615385),
        dict(r=[6, 2, 2], n=[12 / 13, 5 / 13, 0], half_l=0.5, q=[7, 1, 2], target=0.5),
    ]
    cases = []

    def add(c1, c2, d, s, t, name):
        cases.append(dict(name=name, r1=c1[0:3], n1=c1[3:6], half_l1=c1[6], r2=c2[0:3],
                          n2=c2[3:6], half_l2=c2[6], d=list(d), s=s, t=t))

    add([2, 1, 0, 0.6, 0.8, 0, 1], [0, 3, 0, 0, 1, 0, 0.5], [-2, 1.5, 0], 0, -0.5, "skew-h1")
    s, t = -0.07692307692, 0.5
    add([4, 4, 2, 5 / 13, 12 / 13, 0, 2.7], [5, 3, 2, 0, 1, 0, 0.5],
        [(5 + t * 0) - (4 + s * 5 / 13), (3 + t) - (4 + s * 12 / 13), 0], s, t, "skew-h2")
    add([4, 4, 2, 5 / 13, 12 / 13, 0, 2.7], [5, 3, 2, 0, 1, 0, 5.0], [0, 0, 0], 2.6, 3.4,
        "skew-h3-intersect")
    s, t = 1.0, 1.9230769231
    add([4, 4, 2, 5 / 13, 12 / 13, 0, 1.0], [5, 3, 2, 0, 1, 0, 5.0],
        [(5 + t * 0) - (4 + s * 5 / 13), (3 + t) - (4 + s * 12 / 13), 0], s, t, "skew-h4")
    s, t = -0.6, 0.5
    add([4, 4, 5, 5 / 13, 12 / 13, 0, 0.6], [6, 2, 5, 0, 1, 0, 0.5],
        [(6 + t * 0) - (4 + s * 5 / 13), (2 + t) - (4 + s * 12 / 13), 0], s, t, "skew-h5")
    add([0, 0, 0, 1, 0, 0, 0.5], [0, 1.4, 0, 0, 0, 1, 0.5], [0, 1.4, 0], 0, 0, "skew-3d1")
    th = math.pi / 6
    hl1 = hl2 = 0.5
    l1, l2 = 1.0, 1.0
    x2 = 0.75 * l1 / math.cos(th) - hl1 * math.cos(th)
    dd = [0.75 * math.tan(th) * math.sin(th), 0, -0.75 * math.sin(th)]
    add([0, 0, hl1 * math.sin(th), math.cos(th), 0, math.sin(th), 0.5], [x2, 0, 0, 0, 1, 0, 0.5],
        dd, 0.25 * l1, 0, "skew-3d2")
    add([0, 0, hl1 * math.sin(th), math.cos(th), 0, math.sin(th), 0.5],
        [x2, l2 / 4, 0, 0, 1, 0, 0.5], dd, 0.25 * l1, -0.25 * l2, "skew-3d3")
    eps = 5 * math.pi / 180
    r1 = [0, 0, hl1 * math.sin(th)]
    n1 = [math.cos(th), 0, math.sin(th)]
    r2 = [x2, 0, 0]
    n2 = [-math.sin(eps), math.cos(eps), 0]
    d, s, t = lines(r1, n1, r2, n2)
    add(r1 + n1 + [0.5], r2 + n2 + [0.5], d, s, t, "skew-3d4-lines")
    # parallel and nearly parallel (testDistBetweenCells.cpp:346-506)
    s, t = -1, 0.75
    add([0, 0, 0, 1, 0, 0, 1.0], [-2, 2, 0, 1, 0, 0, 0.75], [(-2 + t) - s, 2, 0], s, t, "par-1")
    add([0, 0, 2, 1, 0, 0, 1.0], [-0.8, 2, 2, 1, 0, 0, 1.5], [0, 2, 0], -0.15, 0.65, "par-2")
    add([0, 0, 2, 1, 0, 0, 1.0], [-0.8, 2, 2, -1, 0, 0, 1.5], [0, 2, 0], -0.15, -0.65, "par-3")
    add([0, 0, 5, 1, 0, 0, 1.0], [0.1, 2, 5, 1, 0, 0, 1.5], [0, 2, 0], 0, -0.1, "par-4")
    add([0, 0, 5, 1, 0, 0, 1.0], [0.1, 2, 5, 1, 0, 0, 0.6], [0, 2, 0], 0.1, 0, "par-5")
    add([0, 0, 10, 1, 0, 0, 1.0], [0.5, 2, 10, 1, 0, 0, 1.1], [0, 2, 0], 0.2, -0.3, "par-6")
    add([0, 0, 10, 1, 0, 0, 1.0], [0.5, 2, 10, -1, 0, 0, 1.1], [0, 2, 0], 0.2, 0.3, "par-7")
    add([0.5, 2, 10, 1, 0, 0, 1.1], [0, 0, 10, 1, 0, 0, 1.0], [0, -2, 0], -0.3, 0.2, "par-8-swap")
    s, t = 1, -1.5
    add([0, 0, 7, 1, 0, 0, 1.0], [5, 2, 7, 1, 0, 0, 1.5], [(5 + t) - s, 2, 0], s, t, "par-9")
    add([0, 0, 7, -1, 0, 0, 1.0], [5, 2, 7, 1, 0, 0, 1.5], [(5 + t) - s, 2, 0], -s, t, "par-10")
    add([0, 0, 0.5, 1, 0, 0, 1.0], [5, 0, 0.5, 1, 0, 0, 1.5], [(5 + t) - s, 0, 0], s, t, "par-11")
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "dist_kats.json")
    with open(out, "w") as f:
        json.dump(dict(nearest=nearest, cases=cases), f, indent=0)
    print(f"wrote {len(cases)} distance cases and {len(nearest)} nearest cases to {out}")


if __name__ == "__main__":
    main()
