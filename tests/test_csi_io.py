"""Result files and clustering of the CSI application (CPU only).

Pins: published Java formatting behaviour (``Double.toString`` examples from
its specification; ``%.f`` rounds the shortest decimal digits half-up, e.g.
``String.format("%.1f", 0.15)`` is "0.2" where C prints "0.1"), a 4-image
example worked by hand through ``Clustering.java`` (including its stale
average for merged rows), ``java.util.HashMap`` iteration order, and
round trips of the binary matrix."""
import os

import numpy as np
import pytest

from rocket_amd import csi_io as io


@pytest.mark.parametrize("x,s", [
    (1.0, "1.0"), (100.0, "100.0"), (0.001, "0.001"), (1e-4, "1.0E-4"), (1e7, "1.0E7"),
    (9999999.0, "9999999.0"), (123456.789, "123456.789"), (-0.5, "-0.5"), (0.0, "0.0"), (-0.0, "-0.0"),
    (0.1 + 0.2, "0.30000000000000004"), (1.5e-5, "1.5E-5"), (1e22, "1.0E22"), (2.5e-3, "0.0025"),
    (float("nan"), "NaN"), (float("inf"), "Infinity"), (1.0 / 3.0, "0.3333333333333333"),
])
def test_java_double_to_string(x, s):
    assert io.java_double(x) == s


@pytest.mark.parametrize("x,s", [
    (0.0, "0.000000"), (1.0, "1.000000"), (0.1234565, "0.123457"), (1.0000005, "1.000001"),
    (-1e-9, "-0.000000"), (0.9999995, "1.000000"), (-2.5, "-2.500000"), (12345.6789, "12345.678900"),
])
def test_java_fixed6(x, s):
    assert io.java_fixed6(x) == s


def hand_matrix():
    return np.array([[0, .9, .1, .2], [.9, 0, .3, .4], [.1, .3, 0, .8], [.2, .4, .8, 0]], dtype=np.float64)


def test_hierarchical_clustering_hand_example():
    """Merge (0,1) at .9, then (2,3) at .8; the third merge links clusters 4
    and 5 at .2 -- the stale value left in row 0 after the first merge (the
    true average linkage would be .25), because the reference refreshes only
    singleton rows."""
    link = io.hierarchical_clustering(hand_matrix(), ["a", "b", "c", "d"])
    assert [repr(l) for l in link] == ["[0,1,0.9,2]", "[2,3,0.8,2]", "[4,5,0.2,4]"]


def test_clustering_matches_brute_force_restatement():
    """Random symmetric matrices against a literal, loop-for-loop restatement
    of Clustering.java."""
    rng = np.random.default_rng(3)

    def literal(c):
        n = len(c)
        m = [row[:] for row in c]
        nid, ids, mem, out = n, list(range(n)), {i: [i] for i in range(n)}, []
        for _ in range(n - 1):
            hi, idx = -1e100, (0, 0)
            for i in range(n):
                for j in range(n):
                    if m[i][j] > hi:
                        hi, idx = m[i][j], (i, j)
            n1, n2 = idx
            if n1 == n2:
                break
            c1, c2 = ids[n1], ids[n2]
            a = mem.pop(c1)
            b = mem.pop(c2)
            a.extend(b)
            mem[nid] = a
            ids[n1] = nid
            out.append((c1, c2, m[n1][n2], len(a)))
            if len(a) >= n:
                break
            for i in range(n):
                if i in mem:
                    other = mem[ids[i]]
                    s = 0.0
                    for x in mem[nid]:
                        for y in other:
                            s += c[x][y]
                    avg = s / (len(mem[nid]) * len(other))
                    m[n1][i] = avg
                    m[i][n1] = avg
            for i in range(n):
                m[n2][i] = -1e200
                m[i][n2] = -1e200
            nid += 1
        return out

    for n in (2, 5, 9, 17):
        a = rng.normal(size=(n, n))
        c = (a + a.T) / 2
        np.fill_diagonal(c, 0.0)
        got = [(l.n1, l.n2, l.dist, l.size) for l in io.hierarchical_clustering(c)]
        assert got == literal(c.tolist())


def test_flat_clustering_text():
    link = io.hierarchical_clustering(hand_matrix())
    assert io.flat_clustering_text(link, 4) == (
        "flat clustering:\n0=[0]\n1=[1]\n2=[2]\n3=[3]\n\nlabels:\n[1 2 3 4 ]\n\n")
    assert io.flat_clustering_text(link, 4, threshold=0.5) == (
        "flat clustering:\n4=[0, 1]\n5=[2, 3]\n\nlabels:\n[1 1 2 2 ]\n\n")


def test_java_hashmap_iteration_order():
    """Capacity 16 until the 13th key; keys k and k + capacity share a bucket
    and iterate in insertion order; the table never shrinks."""
    m = io._JavaIntHashMap()
    for k in (17, 1, 3, 16, 0):
        m.put(k, k)
    assert [k for k, _ in m.entries()] == [16, 0, 17, 1, 3]
    for k in range(5, 12):
        m.put(k, k)
    assert m.cap == 16 and len(m.items) == 12
    m.put(40, 40)                                  # 13 entries > 12: capacity 32, 40 -> bucket 8 after 8
    assert m.cap == 32
    assert [k for k, _ in m.entries()] == [0, 1, 3, 5, 6, 7, 8, 40, 9, 10, 11, 16, 17]
    for k in (5, 6, 7, 8, 9, 10, 11, 40):
        m.remove(k)
    assert m.cap == 32 and [k for k, _ in m.entries()] == [0, 1, 3, 16, 17]


def test_result_files_round_trip(tmp_path):
    rng = np.random.default_rng(1)
    n = 6
    pairs = [(i, j) for i in range(n) for j in range(i + 1, n)]
    scores = rng.uniform(-0.1, 0.9, size=len(pairs))
    c = io.correlation_matrix(n, pairs, scores)
    assert np.array_equal(c, c.T) and np.all(np.diag(c) == 0)
    names = [f"img{i}.jpg" for i in range(n)]
    paths = io.write_all("case1", c, names, str(tmp_path))
    back = np.fromfile(paths["matrix_bin"], dtype=">f8").reshape(n, n)
    assert np.array_equal(back, c)
    assert os.path.getsize(paths["matrix_bin"]) == 8 * n * n
    lines = open(paths["edgelist"]).read().splitlines()
    assert len(lines) == n * (n - 1) // 2
    assert lines[0] == f"img1.jpg img0.jpg {io.java_double(c[1, 0])}"
    rows = open(paths["matrix_txt"]).read().split("\n")
    assert rows[0] == "".join(io.java_fixed6(v) + ", " for v in c[0]) and rows[n] == "" and rows[n + 1] == ""
    link = open(paths["linkage"]).read().splitlines()
    assert len(link) == n and link[-1] == ""          # n-1 links, then an empty line


def test_jpeg_directory_listing_and_testcase_name(tmp_path):
    """Main.java:69-73, 88: the test-case name check and the sorted,
    case-insensitive ".jpg" listing (Util.java:91-106), before any GPU work."""
    from rocket_amd import app
    for nm in ("b.jpg", "A.JPG", "c.jpeg", "a.Jpg", "notes.txt", "z.jpg"):
        (tmp_path / nm).write_bytes(b"")
    assert app.scan_jpeg_directory(str(tmp_path)) == ["A.JPG", "a.Jpg", "b.jpg", "z.jpg"]
    with pytest.raises(ValueError, match="special characters"):
        app.common_source_identification_dir("bad name!", str(tmp_path))
    with pytest.raises(FileNotFoundError):
        app.scan_jpeg_directory(str(tmp_path / "missing"))
