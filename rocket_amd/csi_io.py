"""Result files and clustering of the CSI application, host side.

Restates ``CSI/util/IO.java`` and ``CSI/util/Clustering.java`` (paths under
``src/main/java/nl/esciencecenter/common_source_identification/``) so a run of
this framework writes the same files as the reference's ``Main``
(``CSI/Main.java:75-139``): ``matrix-<case>.txt`` / ``.dat``,
``edgelist-<case>.txt``, ``linkage-<case>.txt``, ``clustering-<case>.txt``.

Behaviour kept from the reference, as it is:

* the matrix is symmetric with a zero diagonal (``Main.java:121-126``);
* text numbers follow Java: ``Double.toString`` (shortest round-trip digits,
  scientific notation outside [1e-3, 1e7)) and ``%.6f`` (half-up rounding of
  those shortest digits);
* ``hierarchical_clustering`` is average linkage on similarities, but after a
  merge only rows that are still singletons get the new average
  (``Clustering.java:76-93`` tests ``cluster_members.containsKey(i)`` with a row
  index against cluster ids), so rows of earlier merged clusters keep stale
  similarities; the merged-away row is set to -1e200;
* ``write_flat_clustering`` cuts the linkage at the first link below 60 and
  prints clusters in ``java.util.HashMap`` iteration order (buckets of
  ``key & (capacity - 1)``, insertion order inside a bucket).
"""
from __future__ import annotations

import math
from decimal import ROUND_HALF_UP, Decimal

import numpy as np


# ------------------------------------------------------------ Java numbers
def java_double(x: float) -> str:
    """``Double.toString(x)``: the shortest digits that round-trip, written
    as plain decimal for 1e-3 <= |x| < 1e7 and as d.dddE<exp> otherwise."""
    x = float(x)
    if math.isnan(x):
        return "NaN"
    if math.isinf(x):
        return "Infinity" if x > 0 else "-Infinity"
    if x == 0.0:
        return "-0.0" if math.copysign(1.0, x) < 0 else "0.0"
    sign = "-" if x < 0 else ""
    digits, exp = _shortest_digits(abs(x))      # value = 0.d1d2... * 10^exp
    ax = abs(x)
    if 1e-3 <= ax < 1e7:
        if exp <= 0:
            return sign + "0." + "0" * (-exp) + digits
        if exp >= len(digits):
            return sign + digits + "0" * (exp - len(digits)) + ".0"
        return sign + digits[:exp] + "." + digits[exp:]
    mant = digits[0] + "." + (digits[1:] if len(digits) > 1 else "0")
    return f"{sign}{mant}E{exp - 1}"


def _shortest_digits(ax: float) -> tuple[str, int]:
    """(digits, exp) with ax = 0.digits * 10**exp, digits the shortest
    round-trip representation (Python's repr algorithm)."""
    r = repr(ax)
    if "e" in r:
        m, e = r.split("e")
        e = int(e)
    else:
        m, e = r, 0
    if "." in m:
        ip, fp = m.split(".")
    else:
        ip, fp = m, ""
    digits = (ip + fp).lstrip("0")
    # position of the decimal point relative to the first significant digit
    lead_zeros = len(ip + fp) - len((ip + fp).lstrip("0"))
    exp = len(ip) - lead_zeros + e
    digits = digits.rstrip("0") or "0"
    return digits, exp


def java_fixed6(x: float) -> str:
    """``String.format("%.6f", x)``: half-up rounding of the shortest decimal
    digits of x to six places."""
    x = float(x)
    if math.isnan(x):
        return "NaN"
    if math.isinf(x):
        return "Infinity" if x > 0 else "-Infinity"
    d = Decimal(repr(x)).quantize(Decimal("0.000001"), rounding=ROUND_HALF_UP)
    s = format(d, "f")
    if s.startswith("-") and d == 0:
        s = s  # Java keeps "-0.000000" for negative values that round to zero
    return s


# ------------------------------------------------------------ matrix files
def write_matrix_text(cortable, path: str) -> None:
    """``IO.write_matrix_text`` (``IO.java:24-40``): rows of "%.6f, ", then an
    empty line."""
    c = np.asarray(cortable, dtype=np.float64)
    n = c.shape[1]
    with open(path, "w", newline="\n") as f:
        for i in range(n):
            f.write("".join(java_fixed6(c[i, j]) + ", " for j in range(n)))
            f.write("\n")
        f.write("\n")


def write_matrix_binary(cortable, path: str) -> None:
    """``IO.write_matrix_binary`` (``IO.java:218-235``): row-major big-endian FP64."""
    np.asarray(cortable, dtype=">f8").tofile(path)


def write_edgelist(cortable, filenames, path: str) -> None:
    """``IO.write_edgelist`` (``IO.java:195-210``): "name_i name_j value" for j < i."""
    c = np.asarray(cortable, dtype=np.float64)
    n = c.shape[1]
    with open(path, "w", newline="\n") as f:
        for i in range(n):
            for j in range(i):
                f.write(f"{filenames[i]} {filenames[j]} {java_double(c[i, j])}\n")


def correlation_matrix(n: int, pairs, scores) -> np.ndarray:
    """``Main.java:121-126``: corrs[i][j] = corrs[j][i] = score, zero diagonal."""
    m = np.zeros((n, n), dtype=np.float64)
    for (i, j), v in zip(pairs, scores):
        m[i, j] = m[j, i] = float(v)
    return m


# -------------------------------------------------------------- clustering
class Link:
    """``Clustering.Link``: merged cluster ids, similarity, new size."""

    __slots__ = ("n1", "n2", "dist", "size")

    def __init__(self, n1: int, n2: int, dist: float, size: int):
        self.n1, self.n2, self.dist, self.size = int(n1), int(n2), float(dist), int(size)

    def __repr__(self) -> str:
        return f"[{self.n1},{self.n2},{java_double(self.dist)},{self.size}]"

    def __eq__(self, other) -> bool:
        return (self.n1, self.n2, self.dist, self.size) == (other.n1, other.n2, other.dist, other.size)


def _find_max(matrix: np.ndarray) -> tuple[int, int]:
    """``Clustering.findMax``: first (row-major) entry strictly above -1e100
    that beats every earlier one."""
    flat = int(np.argmax(matrix))                  # first occurrence of the maximum
    if not matrix.flat[flat] > -1e100:
        return 0, 0
    n = matrix.shape[1]
    return flat // n, flat % n


def hierarchical_clustering(cortable, filenames=None) -> list:
    """``Clustering.hierarchical_clustering`` (``Clustering.java:10-103``)."""
    cort = np.array(cortable, dtype=np.float64)
    n = cort.shape[0]
    matrix = cort.copy()
    next_id = n
    cluster_ids = list(range(n))
    members = {i: [i] for i in range(n)}
    linkage = []
    for _ in range(n - 1):
        n1, n2 = _find_max(matrix)
        if n1 == n2:
            break
        c1, c2 = cluster_ids[n1], cluster_ids[n2]
        m1 = members.pop(c1)
        m2 = members.pop(c2)
        m1.extend(m2)
        members[next_id] = m1
        cluster_ids[n1] = next_id
        new_size = len(m1)
        linkage.append(Link(c1, c2, matrix[n1, n2], new_size))
        if new_size >= n:
            break
        # only rows whose index is still a cluster id (singletons) are updated;
        # sums run over the members in list order, as the reference's loops
        acc = np.zeros(n, dtype=np.float64)
        for a in m1:
            acc += cort[a]
        for i in range(n):
            if i in members:
                other = members[cluster_ids[i]]
                if len(other) == 1:
                    avg = acc[other[0]] / (len(m1) * 1)
                else:                               # not reachable: keys < n are singletons
                    s = 0.0
                    for a in m1:
                        for b in other:
                            s += cort[a, b]
                    avg = s / (len(m1) * len(other))
                matrix[n1, i] = avg
                matrix[i, n1] = avg
        matrix[n2, :] = -1e200
        matrix[:, n2] = -1e200
        next_id += 1
    return linkage


def write_linkage(linkage, path: str) -> None:
    """``IO.write_linkage`` (``IO.java:45-56``)."""
    with open(path, "w", newline="\n") as f:
        for l in linkage:
            f.write(repr(l) + "\n")
        f.write("\n")


class _JavaIntHashMap:
    """Iteration order of ``java.util.HashMap<Integer, V>`` for small keys:
    capacity doubles from 16 while size > 0.75 * capacity and never shrinks;
    entries iterate by bucket (key & (capacity - 1)), insertion order within a
    bucket (no tree bins: every key here is below 2 * capacity)."""

    def __init__(self):
        self.cap = 16
        self.items: dict = {}
        self.seq: dict = {}
        self.t = 0

    def put(self, k: int, v) -> None:
        if k not in self.items:
            self.seq[k] = self.t
            self.t += 1
        self.items[k] = v
        while len(self.items) > 0.75 * self.cap:
            self.cap *= 2

    def remove(self, k: int):
        self.seq.pop(k)
        return self.items.pop(k)

    def get(self, k: int):
        return self.items[k]

    def entries(self):
        keys = sorted(self.items, key=lambda k: ((k ^ (k >> 16)) & (self.cap - 1), self.seq[k]))
        return [(k, self.items[k]) for k in keys]


def flat_clustering_text(linkage, n: int, threshold: float = 60.0) -> str:
    """``IO.write_flat_clustering`` (``IO.java:58-131``) as text."""
    out = ["flat clustering:\n"]
    members = _JavaIntHashMap()
    for i in range(n):
        members.put(i, [i])
    next_id = n
    terminated = False
    it = iter(linkage)
    for _ in range(n - 1):
        if terminated:
            break
        try:
            link = next(it)
        except StopIteration:
            return ""          # the reference throws here before anything is flushed
        if link.dist < threshold:
            for key, lst in members.entries():
                lst.sort()
                out.append(f"{key}=[{', '.join(str(v) for v in lst)}]\n")
            terminated = True
        if not terminated:
            m1 = members.remove(link.n1)
            m2 = members.remove(link.n2)
            m1.extend(m2)
            members.put(next_id, m1)
            next_id += 1
    out.append("\n")
    out.append("labels:\n")
    labeling = [0] * n
    label = 1
    for _, lst in members.entries():
        for m in lst:
            labeling[m] = label
        label += 1
    digits = int(math.log10(label)) + 1
    out.append("[" + "".join(f"{l:{digits}d} " for l in labeling) + "]\n")
    out.append("\n")
    return "".join(out)


def write_flat_clustering(linkage, filenames, path: str, threshold: float = 60.0) -> None:
    with open(path, "w", newline="\n") as f:
        f.write(flat_clustering_text(linkage, len(filenames), threshold))


def write_all(testcase: str, cortable, filenames, directory: str = ".") -> dict:
    """Everything ``Main.java:127-139`` writes, with its file names."""
    import os
    names = {
        "linkage": f"linkage-{testcase}.txt",
        "clustering": f"clustering-{testcase}.txt",
        "edgelist": f"edgelist-{testcase}.txt",
        "matrix_txt": f"matrix-{testcase}.txt",
        "matrix_bin": f"matrix-{testcase}.dat",
    }
    paths = {k: os.path.join(directory, v) for k, v in names.items()}
    linkage = hierarchical_clustering(cortable, filenames)
    write_linkage(linkage, paths["linkage"])
    write_flat_clustering(linkage, filenames, paths["clustering"])
    write_edgelist(cortable, filenames, paths["edgelist"])
    write_matrix_text(cortable, paths["matrix_txt"])
    write_matrix_binary(cortable, paths["matrix_bin"])
    return paths
