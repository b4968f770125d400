"""A small sequential-Huffman JPEG writer for tests (ITU-T T.81).

It writes given quantised DCT coefficients as they are, so a decoder's output
can be checked against the inverse DCT of exactly those coefficients.  It
covers what Pillow's encoder does not produce: components coded in separate
(non-interleaved) scans, arbitrary sampling factors, 16-bit quantisation
tables under an extended (SOF1) frame header, restart intervals on any MCU
count, and an Adobe marker selecting RGB colour.

Huffman tables are built from the symbol counts of the data (code lengths at
most 16, the all-ones code left unused as T.81 annex C requires).
"""
from __future__ import annotations

import heapq
import struct

import numpy as np


def zigzag() -> list[int]:
    """Natural (row-major) index of each zig-zag position (T.81 figure A.6):
    anti-diagonals s = row + col in turn, odd ones walked down-left (row
    rising), even ones up-right."""
    order = []
    for s in range(15):
        rows = range(max(0, s - 7), min(s, 7) + 1)
        cells = [(r, s - r) for r in rows]
        if s % 2 == 0:
            cells.reverse()
        order += [r * 8 + c for r, c in cells]
    return order


ZIGZAG = zigzag()


def category(v: int) -> int:
    return int(abs(int(v))).bit_length()


def amplitude_bits(v: int, s: int) -> int:
    return int(v) if v >= 0 else int(v) + (1 << s) - 1


class BitWriter:
    def __init__(self):
        self.out = bytearray()
        self.acc = 0
        self.n = 0

    def put(self, code: int, length: int):
        for i in range(length - 1, -1, -1):
            self.acc = (self.acc << 1) | ((code >> i) & 1)
            self.n += 1
            if self.n == 8:
                self.out.append(self.acc)
                if self.acc == 0xFF:
                    self.out.append(0x00)   # byte stuffing
                self.acc = 0
                self.n = 0

    def flush(self):
        if self.n:
            self.put((1 << (8 - self.n)) - 1, 8 - self.n)   # pad with one bits


def huffman_lengths(counts: dict[int, int]) -> dict[int, int]:
    """Code length per symbol: a Huffman code over the symbols plus a
    reserved dummy that takes the all-ones code; if it would need more than
    16 bits, a flat code instead."""
    syms = sorted(counts)
    heap = [(counts[s], 0, [s]) for s in syms] + [(0, 0, [256])]
    heapq.heapify(heap)
    depth = {s: 0 for s in syms + [256]}
    tick = 1
    while len(heap) > 1:
        a = heapq.heappop(heap)
        b = heapq.heappop(heap)
        for s in a[2] + b[2]:
            depth[s] += 1
        heapq.heappush(heap, (a[0] + b[0], tick, a[2] + b[2]))
        tick += 1
    if max(depth.values()) > 16:
        L = max(1, (len(syms) + 1 - 1).bit_length())
        while len(syms) > (1 << L) - 1:
            L += 1
        depth = {s: L for s in syms}
        depth[256] = L
    return depth


def canonical(depth: dict[int, int]):
    """(BITS, HUFFVAL, {symbol: (code, length)}) of a canonical code; the
    dummy symbol 256, sorted last among the longest codes, is dropped."""
    order = sorted(depth, key=lambda s: (depth[s], s))
    code, prev = 0, depth[order[0]]
    table = {}
    for s in order:
        code <<= depth[s] - prev
        prev = depth[s]
        table[s] = (code, depth[s])
        code += 1
    del table[256]
    bits = [0] * 16
    for s, (_, l) in table.items():
        bits[l - 1] += 1
    vals = [s for s in order if s != 256]
    return bits, vals, table


def _segment(marker: int, body: bytes) -> bytes:
    return struct.pack(">BBH", 0xFF, marker, len(body) + 2) + body


def write_jpeg(width: int, height: int, comps: list[dict], qtables: dict[int, np.ndarray], scans: list[list[int]],
               restart: int = 0, extended: bool = False, adobe_transform: int | None = None,
               table_of=lambda c: c % 2) -> bytes:
    """comps: [{"id", "h", "v", "tq", "coef"}], coef an int array of shape
    (rows, cols, 64) in natural order covering the component's MCU-padded
    block grid; qtables: {slot: 64 values in natural order}; scans: lists
    of component indices; table_of(c): Huffman slot of component c."""
    hmax = max(c["h"] for c in comps)
    vmax = max(c["v"] for c in comps)
    mcux = -(-width // (8 * hmax))
    mcuy = -(-height // (8 * vmax))
    for c in comps:
        assert c["coef"].shape == (mcuy * c["v"], mcux * c["h"], 64)

    def scan_blocks(scan):
        """(component index, block row, block col) in coding order, and the
        number of blocks per MCU."""
        if len(scan) == 1:
            c = comps[scan[0]]
            cw = -(-width * c["h"] // hmax)
            ch = -(-height * c["v"] // vmax)
            seq = [(scan[0], by, bx) for by in range(-(-ch // 8)) for bx in range(-(-cw // 8))]
            return seq, 1
        seq = []
        for my in range(mcuy):
            for mx in range(mcux):
                for ci in scan:
                    c = comps[ci]
                    for v in range(c["v"]):
                        for h in range(c["h"]):
                            seq.append((ci, my * c["v"] + v, mx * c["h"] + h))
        return seq, sum(comps[ci]["h"] * comps[ci]["v"] for ci in scan)

    def symbols(ci, by, bx, pred):
        z = [int(comps[ci]["coef"][by, bx, ZIGZAG[k]]) for k in range(64)]
        diff = z[0] - pred[ci]
        pred[ci] = z[0]
        s = category(diff)
        dc = (s, amplitude_bits(diff, s), s)
        ac = []
        run = 0
        last = max([k for k in range(1, 64) if z[k] != 0], default=0)
        for k in range(1, last + 1):
            if z[k] == 0:
                run += 1
                continue
            while run > 15:
                ac.append((0xF0, 0, 0))
                run -= 16
            s = category(z[k])
            ac.append(((run << 4) | s, amplitude_bits(z[k], s), s))
            run = 0
        if last < 63:
            ac.append((0x00, 0, 0))
        return dc, ac

    # first pass: symbol counts per Huffman slot
    dc_counts, ac_counts = {}, {}
    coded = []
    for scan in scans:
        seq, per_mcu = scan_blocks(scan)
        pred = {ci: 0 for ci in scan}
        blocks = []
        for i, (ci, by, bx) in enumerate(seq):
            if restart and i % per_mcu == 0 and i and (i // per_mcu) % restart == 0:
                pred = {cj: 0 for cj in scan}
            dc, ac = symbols(ci, by, bx, pred)
            t = table_of(ci)
            dc_counts.setdefault(t, {}).setdefault(dc[0], 0)
            dc_counts[t][dc[0]] += 1
            for a in ac:
                ac_counts.setdefault(t, {}).setdefault(a[0], 0)
                ac_counts[t][a[0]] += 1
            blocks.append((i, ci, dc, ac))
        coded.append((scan, per_mcu, blocks))
    dc_tab = {t: canonical(huffman_lengths(c)) for t, c in dc_counts.items()}
    ac_tab = {t: canonical(huffman_lengths(c)) for t, c in ac_counts.items()}

    out = bytearray(b"\xff\xd8")
    if adobe_transform is not None:
        out += _segment(0xEE, b"Adobe" + struct.pack(">HHHB", 100, 0, 0, adobe_transform))
    for slot, q in qtables.items():
        q = np.asarray(q).reshape(64)
        wide = bool(np.any(q > 255))
        body = bytes([(1 if wide else 0) << 4 | slot])
        for k in range(64):
            v = int(q[ZIGZAG[k]])
            body += struct.pack(">H", v) if wide else bytes([v])
        out += _segment(0xDB, body)
    sof = struct.pack(">BHHB", 8, height, width, len(comps))
    for c in comps:
        sof += bytes([c["id"], c["h"] << 4 | c["v"], c["tq"]])
    out += _segment(0xC1 if extended else 0xC0, sof)
    for cls, tabs in ((0, dc_tab), (1, ac_tab)):
        for t, (bits, vals, _) in tabs.items():
            out += _segment(0xC4, bytes([cls << 4 | t] + bits + vals))
    if restart:
        out += _segment(0xDD, struct.pack(">H", restart))
    for scan, per_mcu, blocks in coded:
        sos = bytes([len(scan)])
        for ci in scan:
            t = table_of(ci)
            sos += bytes([comps[ci]["id"], t << 4 | t])
        sos += bytes([0, 63, 0])
        out += _segment(0xDA, sos)
        bw = BitWriter()
        rst = 0
        for i, ci, dc, ac in blocks:
            if restart and i % per_mcu == 0 and i and (i // per_mcu) % restart == 0:
                bw.flush()
                bw.out += bytes([0xFF, 0xD0 + rst % 8])
                rst += 1
            t = table_of(ci)
            code, length = dc_tab[t][2][dc[0]]
            bw.put(code, length)
            bw.put(dc[1], dc[2])
            for sym, amp, s in ac:
                code, length = ac_tab[t][2][sym]
                bw.put(code, length)
                bw.put(amp, s)
        bw.flush()
        out += bw.out
    out += b"\xff\xd9"
    return bytes(out)


def idct_blocks(coef: np.ndarray, q: np.ndarray) -> np.ndarray:
    """Samples of a component: the orthonormal 2-D inverse DCT-II of the
    dequantised blocks (which is T.81's IDCT), +128, rounded and clamped;
    coef (rows, cols, 64) natural order -> (rows * 8, cols * 8) uint8."""
    from scipy.fft import idctn

    rows, cols, _ = coef.shape
    f = coef.reshape(rows, cols, 8, 8).astype(np.float64) * np.asarray(q, np.float64).reshape(8, 8)
    s = idctn(f, type=2, norm="ortho", axes=(2, 3)) + 128.0
    s = np.clip(np.floor(s + 0.5), 0, 255).astype(np.uint8)
    return s.transpose(0, 2, 1, 3).reshape(rows * 8, cols * 8)
