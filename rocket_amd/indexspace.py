"""Upper-triangle pair scheduling: mirror of Rocket's index spaces over the C ABI.

* :class:`HilbertCurveIndexSpace` -- ``R/indexspace/HilbertCurveIndexSpace.java:7-123``
* :class:`TilingIndexSpace` -- ``R/indexspace/TilingIndexSpace.java:15-91``
* :class:`IndexSpaceDivision` -- ``R/indexspace/IndexSpaceDivision.java``
* :func:`calculate_triangle_size` -- ``R/util/Util.java:51-77``

(``R/`` = ``src/main/java/nl/esciencecenter/rocket/``.)  The recursion runs in
host C++ inside libprnu_hip.so (no device call), so these work without a GPU.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib


class _Node(ctypes.Structure):
    _fields_ = [("left_offset", ctypes.c_int32), ("left_length", ctypes.c_int32),
                ("right_offset", ctypes.c_int32), ("right_length", ctypes.c_int32),
                ("rule", ctypes.c_int32)]


def calculate_triangle_size(row_begin: int, row_end: int, col_begin: int, col_end: int,
                            include_diagonal: bool) -> int:
    return int(_lib.load().prnu_triangle_size(row_begin, row_end, col_begin, col_end, int(include_diagonal)))


class IndexSpaceDivision:
    """Children to spawn plus the entries (key pairs) a divide step emits."""

    def __init__(self):
        self.sub_index_spaces: list = []
        self.leaf_tasks: list = []

    def add_subspace(self, space) -> None:
        self.sub_index_spaces.append(space)

    def add_entry(self, i, j) -> None:
        self.leaf_tasks.append((i, j))

    def get_children(self):
        return self.sub_index_spaces

    def get_entries(self):
        return self.leaf_tasks


class _IndexSpace:
    KIND = None

    def __init__(self, keys, param: int, include_diagonal: bool, node: _Node):
        self.keys = list(keys)
        self.param = int(param)
        self.include_diagonal = bool(include_diagonal)
        self.node = node

    def _make(self, node: _Node):
        return type(self)._from_node(self.keys, self.param, self.include_diagonal, node)

    def size(self) -> int:
        n = self.node
        return calculate_triangle_size(n.left_offset, n.left_offset + n.left_length, n.right_offset,
                                       n.right_offset + n.right_length, self.include_diagonal)

    def divide(self, result: IndexSpaceDivision) -> None:
        lib = _lib.load()
        cap_children = 4 * max(1, (self.node.left_length // 1 + 1) * (self.node.right_length // 1 + 1))
        cap_children = min(cap_children, 1 << 20)
        children = (_Node * cap_children)()
        n_children = ctypes.c_int()
        n_pairs = ctypes.c_int64()
        # first call sizes the pair list
        status = lib.prnu_index_divide(self.KIND, ctypes.byref(self.node), self.param, int(self.include_diagonal),
                                       children, cap_children, ctypes.byref(n_children), None, 0,
                                       ctypes.byref(n_pairs))
        if status not in (_lib.OK, _lib.EINVAL):
            _lib.check(status, "prnu_index_divide")
        pairs = np.zeros((max(1, n_pairs.value), 2), dtype=np.int32)
        _lib.check(lib.prnu_index_divide(self.KIND, ctypes.byref(self.node), self.param, int(self.include_diagonal),
                                         children, cap_children, ctypes.byref(n_children),
                                         pairs.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), len(pairs),
                                         ctypes.byref(n_pairs)), "prnu_index_divide")
        for k in range(n_children.value):
            c = children[k]
            result.add_subspace(self._make(_Node(c.left_offset, c.left_length, c.right_offset, c.right_length,
                                                 c.rule)))
        for i, j in pairs[: n_pairs.value]:
            result.add_entry(self.keys[i], self.keys[j])


class HilbertCurveIndexSpace(_IndexSpace):
    KIND = _lib.INDEX_HILBERT

    def __init__(self, keys, min_split_size: int, include_diagonal: bool):
        n = len(keys)
        super().__init__(keys, min_split_size, include_diagonal, _Node(0, n, 0, n, 0))

    @classmethod
    def _from_node(cls, keys, param, include_diagonal, node):
        obj = cls.__new__(cls)
        _IndexSpace.__init__(obj, keys, param, include_diagonal, node)
        return obj


class TilingIndexSpace(_IndexSpace):
    KIND = _lib.INDEX_TILING

    def __init__(self, tile_size: int, keys, include_diagonal: bool):
        n = len(keys)
        super().__init__(keys, tile_size, include_diagonal, _Node(0, n, 0, n, 0))

    @classmethod
    def _from_node(cls, keys, param, include_diagonal, node):
        obj = cls.__new__(cls)
        _IndexSpace.__init__(obj, keys, param, include_diagonal, node)
        return obj


def all_pairs(count: int, kind: str = "hilbert", param: int = 1, include_diagonal: bool = False) -> np.ndarray:
    """Every (i, j) pair of a `count`-key space in curve order, int32 (P, 2)."""
    lib = _lib.load()
    k = _lib.INDEX_HILBERT if kind == "hilbert" else _lib.INDEX_TILING
    n = ctypes.c_int64()
    _lib.check(lib.prnu_index_pairs(k, count, param, int(include_diagonal), None, 0, ctypes.byref(n)))
    out = np.zeros((max(1, n.value), 2), dtype=np.int32)
    _lib.check(lib.prnu_index_pairs(k, count, param, int(include_diagonal),
                                    out.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), len(out), ctypes.byref(n)))
    return out[: n.value]


def tile_blocks(count: int, tile: int = 256, kind: str = "hilbert", param: int = 1) -> list:
    """Rocket's pair scheduling lifted to GPU tiles: the `count` keys are cut
    into ceil(count/tile) tiles, the index space enumerates the tile pairs
    (I, J), I <= J (diagonal included), in its curve order, and each pair
    becomes a block (row_offset, row_length, col_offset, col_length) of the
    key triangle; I == J blocks are triangles, the rest rectangles."""
    nt = (count + tile - 1) // tile
    out = []
    for I, J in all_pairs(nt, kind, param, include_diagonal=True):
        r0, c0 = int(I) * tile, int(J) * tile
        out.append((r0, min(tile, count - r0), c0, min(tile, count - c0)))
    return out


def ncc_by_index_space(gram, store, count: int, first: int = 0, kind: str = "hilbert", param: int = 1,
                       tile: int = 256, slab_values: int = 0, use_graph: bool = True):
    """All-pairs NCC of slots [first, first+count) of a PatternStore, scheduled
    the way a Rocket host divides the pair triangle (index space over tiles of
    `tile` keys, curve order) but run as ONE launch: every tile pair becomes a
    block of a grouped plan writing straight into a (count, count) FP64 CUDA
    matrix.  Returns (matrix, blocks); the upper triangle (and the diagonal
    blocks' lower halves) are filled."""
    import torch
    from .csi import GramPlan
    blocks = tile_blocks(count, tile, kind, param)
    full = torch.zeros((count, count), dtype=torch.float64, device="cuda")
    spec = [(store, first + r0, rn, first + c0, cn, full[r0:r0 + rn, c0:c0 + cn]) for (r0, rn, c0, cn) in blocks]
    plan = GramPlan(gram, blocks=spec, slab_values=slab_values)
    plan.run(use_graph=use_graph)
    plan.close()
    return full, blocks
