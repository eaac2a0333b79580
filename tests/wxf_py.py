"""A small WXF encoder/decoder for the tests (the published Wolfram Exchange Format layout:
"8:" header, varint lengths, f/A/C/j/i/L/r/S/s tokens). Test infrastructure only."""
from __future__ import annotations

import struct


class Symbol(str):
    pass


class Rule:
    def __init__(self, lhs, rhs):
        self.lhs, self.rhs = lhs, rhs


def _varint(n: int) -> bytes:
    out = bytearray()
    while n >= 0x80:
        out.append((n & 0x7F) | 0x80)
        n >>= 7
    out.append(n)
    return bytes(out)


def _enc(x, out: bytearray):
    if isinstance(x, bool):
        _enc(Symbol("True" if x else "False"), out)
    elif isinstance(x, int):
        for tag, fmt, lo, hi in ((b"C", "<b", -128, 127), (b"j", "<h", -2**15, 2**15 - 1),
                                 (b"i", "<i", -2**31, 2**31 - 1), (b"L", "<q", -2**63, 2**63 - 1)):
            if lo <= x <= hi:
                out += tag + struct.pack(fmt, x)
                return
        raise ValueError("integer out of range")
    elif isinstance(x, float):
        out += b"r" + struct.pack("<d", x)
    elif isinstance(x, Symbol):
        b = x.encode()
        out += b"s" + _varint(len(b)) + b
    elif isinstance(x, str):
        b = x.encode()
        out += b"S" + _varint(len(b)) + b
    elif isinstance(x, Rule):
        out += b"f" + _varint(2)
        _enc(Symbol("Rule"), out)
        _enc(x.lhs, out)
        _enc(x.rhs, out)
    elif isinstance(x, (list, tuple)):
        out += b"f" + _varint(len(x))
        _enc(Symbol("List"), out)
        for y in x:
            _enc(y, out)
    elif isinstance(x, dict):
        out += b"A" + _varint(len(x))
        for k, v in x.items():
            out += b"-"
            _enc(k, out)
            _enc(v, out)
    else:
        raise TypeError(type(x))


def dumps(x) -> bytes:
    out = bytearray(b"8:")
    _enc(x, out)
    return bytes(out)


def loads(b: bytes):
    assert b[:2] == b"8:", "missing header"
    pos = [2]

    def byte():
        v = b[pos[0]]
        pos[0] += 1
        return v

    def varint():
        v, sh = 0, 0
        while True:
            c = byte()
            v |= (c & 0x7F) << sh
            if not c & 0x80:
                return v
            sh += 7

    def take(n):
        v = b[pos[0]:pos[0] + n]
        pos[0] += n
        return v

    def expr():
        t = chr(byte())
        if t == "C":
            return struct.unpack("<b", take(1))[0]
        if t == "j":
            return struct.unpack("<h", take(2))[0]
        if t == "i":
            return struct.unpack("<i", take(4))[0]
        if t == "L":
            return struct.unpack("<q", take(8))[0]
        if t == "r":
            return struct.unpack("<d", take(8))[0]
        if t == "S":
            return take(varint()).decode()
        if t == "s":
            return Symbol(take(varint()).decode())
        if t == "f":
            n = varint()
            head = expr()
            args = [expr() for _ in range(n)]
            if head == "List":
                return args
            if head == "Rule":
                return Rule(args[0], args[1])
            return (head, args)
        if t == "A":
            n = varint()
            d = {}
            for _ in range(n):
                assert chr(byte()) in "-:"
                k = expr()
                d[k] = expr()
            return d
        raise ValueError(f"token {t!r}")

    v = expr()
    assert pos[0] == len(b), "trailing bytes"
    return v
