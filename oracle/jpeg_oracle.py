"""CPU restatement of the reference's JPEG reader (TEST INFRASTRUCTURE ONLY).

The reference decodes frames with ``readJPEG`` (``readjpeg/src/main/cpp/
native.cpp:59-134``): libjpeg with its default decompression parameters and
RGB output, anything that does not decode to 3-component RGB rejected
(``:99-103``).  libjpeg is a third-party library absent from the reference
(the Makefile links the system ``-ljpeg``); its defaults are the IJG
algorithms of libjpeg 6b, which libjpeg-turbo reproduces bit for bit:

* baseline / extended sequential Huffman decoding (ITU T.81 F.2.2), DC
  prediction per component reset at restart markers;
* the "ISLOW" integer inverse DCT (the Loeffler-Ligtenberg-Moschytz
  factorisation with 13-bit fixed-point constants, 2 extra bits kept between
  the column and the row pass, output level-shifted and clamped through a
  range table indexed by the low 10 bits);
* "fancy" upsampling of subsampled chroma: a triangle filter, 3/4 of the
  nearer and 1/4 of the further sample, with libjpeg's alternating rounding
  biases; edge rows and columns replicated;
* YCbCr -> RGB with 16-bit fixed-point tables (R = Y + 1.402 Cr, G = Y -
  0.34414 Cb - 0.71414 Cr, B = Y + 1.772 Cb), clamped to [0, 255].

This module is the checker of the GPU decoder, never the product: only
``tests/`` and ``tests/golden`` scripts import it.  It is pinned against
Pillow's decoder (Pillow bundles libjpeg-turbo; ``tests/test_jpeg_oracle.py``
compares every pixel of encoder outputs at several sizes, qualities,
subsamplings and restart intervals), and its fixed-point constants against
the cosine expressions they stand for.
"""
from __future__ import annotations

import math

import numpy as np

# zigzag position k -> natural (row-major) index
ZIGZAG = np.array([
    0, 1, 8, 16, 9, 2, 3, 10, 17, 24, 32, 25, 18, 11, 4, 5, 12, 19, 26, 33, 40, 48, 41, 34, 27, 20, 13, 6, 7, 14,
    21, 28, 35, 42, 49, 56, 57, 50, 43, 36, 29, 22, 15, 23, 30, 37, 44, 51, 58, 59, 52, 45, 38, 31, 39, 46, 53, 60,
    61, 54, 47, 55, 62, 63], dtype=np.int64)

CONST_BITS, PASS1_BITS = 13, 2


def _fix13(x: float) -> int:
    return int(round(x * (1 << CONST_BITS)))


# The ISLOW constants: |sqrt(2) x (a combination of c_k = cos(k pi / 16))|
# in 13-bit fixed point; the names are the values to 9 decimals.
_C = [math.cos(k * math.pi / 16) for k in range(8)]
_R2 = math.sqrt(2.0)
FIX = {name: _fix13(abs(_R2 * v)) for name, v in {
    "0_298631336": -_C[1] + _C[3] + _C[5] - _C[7],
    "0_390180644": _C[5] - _C[3],
    "0_541196100": _C[6],
    "0_765366865": _C[2] - _C[6],
    "0_899976223": _C[7] - _C[3],
    "1_175875602": _C[3],
    "1_501321110": _C[1] + _C[3] - _C[5] - _C[7],
    "1_847759065": _C[2] + _C[6],
    "1_961570560": -_C[3] - _C[5],
    "2_053119869": _C[1] + _C[3] - _C[5] + _C[7],
    "2_562915447": -_C[1] - _C[3],
    "3_072711026": _C[1] + _C[3] + _C[5] - _C[7],
}.items()}


def _descale(x, n):
    return (x + (1 << (n - 1))) >> n


def _idct_1d(d0, d1, d2, d3, d4, d5, d6, d7):
    """One 8-point ISLOW pass on integer arrays (columns or rows); returns
    the 8 outputs before descaling."""
    z2, z3 = d2, d6
    z1 = (z2 + z3) * FIX["0_541196100"]
    tmp2 = z1 + z3 * (-FIX["1_847759065"])
    tmp3 = z1 + z2 * FIX["0_765366865"]
    tmp0 = (d0 + d4) << CONST_BITS
    tmp1 = (d0 - d4) << CONST_BITS
    tmp10, tmp13 = tmp0 + tmp3, tmp0 - tmp3
    tmp11, tmp12 = tmp1 + tmp2, tmp1 - tmp2
    t0, t1, t2, t3 = d7, d5, d3, d1
    z1, z2, z3, z4 = t0 + t3, t1 + t2, t0 + t2, t1 + t3
    z5 = (z3 + z4) * FIX["1_175875602"]
    t0 = t0 * FIX["0_298631336"]
    t1 = t1 * FIX["2_053119869"]
    t2 = t2 * FIX["3_072711026"]
    t3 = t3 * FIX["1_501321110"]
    z1 = z1 * (-FIX["0_899976223"])
    z2 = z2 * (-FIX["2_562915447"])
    z3 = z3 * (-FIX["1_961570560"]) + z5
    z4 = z4 * (-FIX["0_390180644"]) + z5
    t0 = t0 + z1 + z3
    t1 = t1 + z2 + z4
    t2 = t2 + z2 + z3
    t3 = t3 + z1 + z4
    return (tmp10 + t3, tmp11 + t2, tmp12 + t1, tmp13 + t0, tmp13 - t0, tmp12 - t1, tmp11 - t2, tmp10 - t3)


def _range_limit(v):
    """libjpeg's post-IDCT range table indexed by v & 1023: v + 128 clamped
    to [0, 255] for v in [-512, 511] (wrapping outside)."""
    k = v & 1023
    return np.where(k < 128, k + 128, np.where(k < 512, 255, np.where(k < 896, 0, k - 896))).astype(np.uint8)


def idct_islow(coef, q):
    """coef (..., 64) natural-order quantized coefficients (int), q (64,)
    natural-order quantization table -> (..., 8, 8) uint8 samples."""
    c = (np.asarray(coef, dtype=np.int64) * np.asarray(q, dtype=np.int64)).reshape(coef.shape[:-1] + (8, 8))
    # pass 1: columns (index [..., row, col]; a column is fixed col)
    d = [c[..., r, :] for r in range(8)]
    out = _idct_1d(*d)
    ws = np.stack([_descale(o, CONST_BITS - PASS1_BITS) for o in out], axis=-2)  # (..., 8 rows, 8 cols)
    # pass 2: rows
    d = [ws[..., :, k] for k in range(8)]
    out = _idct_1d(*d)
    res = np.stack([_descale(o, CONST_BITS + PASS1_BITS + 3) for o in out], axis=-1)
    return _range_limit(res)


# ----------------------------------------------------------------- bitstream
class JpegError(ValueError):
    pass


def _u16(b, i):
    return (b[i] << 8) | b[i + 1]


def parse(data: bytes) -> dict:
    """Markers of a baseline / extended sequential Huffman JPEG."""
    b = memoryview(data)
    if len(b) < 4 or b[0] != 0xFF or b[1] != 0xD8:
        raise JpegError("not a JPEG (no SOI)")
    i = 2
    info = {"q": {}, "dc": {}, "ac": {}, "restart": 0, "scans": [], "adobe": None}
    while i < len(b):
        if b[i] != 0xFF:
            raise JpegError(f"marker expected at {i}")
        while i < len(b) and b[i] == 0xFF:
            i += 1
        m = b[i]
        i += 1
        if m == 0xD9:
            info["eoi"] = True
            break
        if m in (0x01,) or 0xD0 <= m <= 0xD7:
            continue
        ln = _u16(b, i)
        seg = bytes(b[i + 2:i + ln])
        i += ln
        if m == 0xDB:
            j = 0
            while j < len(seg):
                pq, tq = seg[j] >> 4, seg[j] & 15
                if pq:
                    q = [_u16(seg, j + 1 + 2 * k) for k in range(64)]
                    j += 129
                else:
                    q = list(seg[j + 1:j + 65])
                    j += 65
                nat = np.zeros(64, dtype=np.int64)
                nat[ZIGZAG] = q
                info["q"][tq] = nat
        elif m == 0xC4:
            j = 0
            while j < len(seg):
                tc, th = seg[j] >> 4, seg[j] & 15
                counts = list(seg[j + 1:j + 17])
                syms = list(seg[j + 17:j + 17 + sum(counts)])
                j += 17 + sum(counts)
                table, code, k = {}, 0, 0
                for ln_ in range(1, 17):
                    for _ in range(counts[ln_ - 1]):
                        table[(ln_, code)] = syms[k]
                        code += 1
                        k += 1
                    code <<= 1
                (info["ac"] if tc else info["dc"])[th] = table
        elif m in (0xC0, 0xC1):
            if seg[0] != 8:
                raise JpegError("only 8-bit samples")
            info["height"], info["width"] = _u16(seg, 1), _u16(seg, 3)
            nc = seg[5]
            info["comps"] = [(seg[6 + 3 * k], seg[7 + 3 * k] >> 4, seg[7 + 3 * k] & 15, seg[8 + 3 * k])
                             for k in range(nc)]
        elif m in (0xC2, 0xC3) or 0xC5 <= m <= 0xCF and m not in (0xC8, 0xCC):
            raise JpegError(f"unsupported frame type SOF{m - 0xC0} (progressive, lossless or arithmetic)")
        elif m == 0xDD:
            info["restart"] = _u16(seg, 0)
        elif m == 0xEE and seg[:5] == b"Adobe":
            info["adobe"] = seg[11] if len(seg) > 11 else 0
        elif m == 0xDA:
            ns = seg[0]
            sc = [(seg[1 + 2 * k], seg[2 + 2 * k] >> 4, seg[2 + 2 * k] & 15) for k in range(ns)]
            # entropy-coded data runs to the next marker that is not RST or stuffing
            j = i
            while j < len(b) - 1:
                if b[j] == 0xFF and b[j + 1] != 0x00 and not (0xD0 <= b[j + 1] <= 0xD7):
                    break
                j += 1
            info["scans"].append((sc, bytes(b[i:j])))
            i = j
        # APPn, COM and others: skipped
    if "comps" not in info:
        raise JpegError("no SOF0/SOF1 frame")
    if not info.get("eoi"):
        # readJPEG's source errors when libjpeg asks for bytes past the end
        # (native.cpp:26-29), which it does for a file without EOI
        raise JpegError("truncated file (no EOI marker)")
    if len(info["comps"]) != 3:
        raise JpegError("image not in RGB color space (not 3 components)")
    if info["adobe"] not in (None, 1):
        raise JpegError("image not in RGB color space (Adobe transform)")
    return info


class _Bits:
    def __init__(self, data: bytes):
        self.d = data
        self.p = 0
        self.acc = 0
        self.n = 0

    def _fill(self):
        while self.n <= 24:
            if self.p < len(self.d):
                c = self.d[self.p]
                if c == 0xFF:
                    nxt = self.d[self.p + 1] if self.p + 1 < len(self.d) else 0
                    if nxt == 0x00:
                        self.p += 2
                    else:           # a marker: feed zeros, as libjpeg does
                        c = 0
                else:
                    self.p += 1
            else:
                c = 0
            self.acc = (self.acc << 8) | c
            self.n += 8

    def bits(self, k):
        if k == 0:
            return 0
        self._fill()
        v = (self.acc >> (self.n - k)) & ((1 << k) - 1)
        self.n -= k
        return v

    def huff(self, table):
        code = 0
        for ln in range(1, 17):
            code = (code << 1) | self.bits(1)
            s = table.get((ln, code))
            if s is not None:
                return s
        raise JpegError("bad Huffman code")

    def restart(self):
        """Drop the remaining bits of the byte and the RSTn marker."""
        self.acc = 0
        self.n = 0
        while self.p + 1 < len(self.d) and not (self.d[self.p] == 0xFF and 0xD0 <= self.d[self.p + 1] <= 0xD7):
            self.p += 1
        self.p += 2


def _extend(v, s):
    return v - (1 << s) + 1 if s and v < (1 << (s - 1)) else v


def decode_coefficients(info):
    """Quantized coefficients per component, (blocks_h, blocks_w, 64) natural
    order, over the whole-MCU padded block grid."""
    comps = info["comps"]
    hmax = max(c[1] for c in comps)
    vmax = max(c[2] for c in comps)
    W, H = info["width"], info["height"]
    mcux = -(-W // (8 * hmax))
    mcuy = -(-H // (8 * vmax))
    coef = {}
    for cid, hs, vs, tq in comps:
        coef[cid] = np.zeros((mcuy * vs, mcux * hs, 64), dtype=np.int64)
    geom = {c[0]: c for c in comps}
    for sc, ecs in info["scans"]:
        bits = _Bits(ecs)
        pred = {cid: 0 for cid, _, _ in sc}
        if len(sc) == 1:
            cid = sc[0][0]
            _, hs, vs, _ = geom[cid]
            bw = -(-(-(-W * hs // hmax)) // 8)
            bh = -(-(-(-H * vs // vmax)) // 8)
            units = [[(cid, by, bx)] for by in range(bh) for bx in range(bw)]
        else:
            units = []
            for my in range(mcuy):
                for mx in range(mcux):
                    u = []
                    for cid, _, _ in sc:
                        _, hs, vs, _ = geom[cid]
                        for v in range(vs):
                            for h in range(hs):
                                u.append((cid, my * vs + v, mx * hs + h))
                    units.append(u)
        tabs = {cid: (info["dc"][td], info["ac"][ta]) for cid, td, ta in sc}
        ri = info["restart"]
        for n, unit in enumerate(units):
            if ri and n and n % ri == 0:
                bits.restart()
                pred = {cid: 0 for cid in pred}
            for cid, by, bx in unit:
                dct, act = tabs[cid]
                blk = coef[cid][by, bx]
                s = bits.huff(dct)
                pred[cid] += _extend(bits.bits(s), s)
                blk[0] = pred[cid]
                k = 1
                while k < 64:
                    rs = bits.huff(act)
                    r, s = rs >> 4, rs & 15
                    if s:
                        k += r
                        blk[ZIGZAG[k]] = _extend(bits.bits(s), s)
                        k += 1
                    else:
                        if r != 15:
                            break
                        k += 16
    return coef


def _fancy_h2v1(p):
    """(rows, dw) -> (rows, 2 dw), libjpeg h2v1 triangle filter (dw > 2)."""
    p = p.astype(np.int64)
    out = np.empty((p.shape[0], 2 * p.shape[1]), dtype=np.int64)
    left = np.concatenate([p[:, :1], p[:, :-1]], axis=1)
    right = np.concatenate([p[:, 1:], p[:, -1:]], axis=1)
    out[:, 0::2] = (3 * p + left + 1) >> 2
    out[:, 1::2] = (3 * p + right + 2) >> 2
    out[:, 0] = p[:, 0]
    out[:, -1] = p[:, -1]
    return out


def _fancy_h2v2(p):
    """(dh, dw) -> (2 dh, 2 dw), libjpeg h2v2 triangle filter (dw > 2); the
    row above row 0 and below row dh - 1 replicate the edge row."""
    p = p.astype(np.int64)
    above = np.concatenate([p[:1], p[:-1]], axis=0)
    below = np.concatenate([p[1:], p[-1:]], axis=0)
    out = np.empty((2 * p.shape[0], 2 * p.shape[1]), dtype=np.int64)
    for v, nb in ((0, above), (1, below)):
        cs = 3 * p + nb                                  # column sums
        last = np.concatenate([cs[:, :1], cs[:, :-1]], axis=1)
        nxt = np.concatenate([cs[:, 1:], cs[:, -1:]], axis=1)
        even = (3 * cs + last + 8) >> 4
        odd = (3 * cs + nxt + 7) >> 4
        even[:, 0] = (4 * cs[:, 0] + 8) >> 4
        odd[:, -1] = (4 * cs[:, -1] + 7) >> 4
        out[v::2, 0::2] = even
        out[v::2, 1::2] = odd
    return out


def upsample(plane, dw, dh, hf, vf):
    """Chroma plane (IDCT grid) -> (vf dh, hf dw): libjpeg's choice (jdsample
    jinit_upsampler): the triangle filter for 2x1 and 2x2 when the component
    is more than 2 samples wide, plain replication otherwise."""
    p = plane[:dh, :dw]
    if hf == 1 and vf == 1:
        return p.astype(np.int64)
    if dw > 2 and hf == 2 and vf == 1:
        return _fancy_h2v1(p)
    if dw > 2 and hf == 2 and vf == 2:
        return _fancy_h2v2(p)
    return np.repeat(np.repeat(p.astype(np.int64), vf, axis=0), hf, axis=1)


def _ycc_tables():
    SB, HALF = 16, 1 << 15

    def fix(x):
        return int(x * (1 << SB) + 0.5)
    x = np.arange(256) - 128
    cr_r = (fix(1.40200) * x + HALF) >> SB
    cb_b = (fix(1.77200) * x + HALF) >> SB
    cr_g = -fix(0.71414) * x
    cb_g = -fix(0.34414) * x + HALF
    return cr_r, cb_b, cr_g, cb_g


def decode(data: bytes) -> np.ndarray:
    """(H, W, 3) uint8 RGB, as readJPEG returns it."""
    info = parse(data)
    coef = decode_coefficients(info)
    comps = info["comps"]
    hmax = max(c[1] for c in comps)
    vmax = max(c[2] for c in comps)
    W, H = info["width"], info["height"]
    planes = []
    for cid, hs, vs, tq in comps:
        c = coef[cid]
        smp = idct_islow(c, info["q"][tq])                       # (bh, bw, 8, 8)
        plane = smp.transpose(0, 2, 1, 3).reshape(c.shape[0] * 8, c.shape[1] * 8)
        dw = -(-W * hs // hmax)
        dh = -(-H * vs // vmax)
        if hmax % hs or vmax % vs or hmax // hs > 2 or vmax // vs > 2 or (hmax // hs, vmax // vs) == (1, 2):
            raise JpegError(f"unsupported sampling {hs}x{vs} of {hmax}x{vmax}")
        up = upsample(plane, dw, dh, hmax // hs, vmax // vs)[:H, :W]
        planes.append(up)
    y, cb, cr = planes
    cr_r, cb_b, cr_g, cb_g = _ycc_tables()
    r = np.clip(y + cr_r[cr], 0, 255)
    g = np.clip(y + ((cb_g[cb] + cr_g[cr]) >> 16), 0, 255)
    b = np.clip(y + cb_b[cb], 0, 255)
    return np.stack([r, g, b], axis=-1).astype(np.uint8)
