"""formats.py — CPU restatement of the reference's on-disk input readers.

TEST INFRASTRUCTURE ONLY. This module is a checker: only tests/ import it. The product's
readers are libptmi's C++ (pt_obj_parse, pt_inst_parse); nothing in the package imports this.

Pure Python, statement by statement after the reference's JavaScript, whose numbers are IEEE
doubles (Python floats) rounded to float32 wherever it stores into a Float32Array (`f32`):
- parse_obj: src/authoring/loadOBJ.ts:62-163 (parseOBJ), JavaScript string semantics by
  regular expression (parseFloat's StrDecimalLiteral prefix, parseInt's digits);
- vertex_normals: src/authoring/loadOBJ.ts:35-60 (computeVertexNormals), Math.hypot as V8
  evaluates it (scaled, Kahan-summed squares);
- parse_instances: src/authoring/loadInstances.ts:30-98 (parseInstances), over struct.
"""
import math
import re
import struct

import numpy as np

_F = re.compile(r"[\t\n\v\f\r ]*([+-]?(?:Infinity|\d+\.?\d*(?:[eE][+-]?\d+)?|\.\d+(?:[eE][+-]?\d+)?))")
_I = re.compile(r"[\t\n\v\f\r ]*([+-]?)(\d+)")


def f32(x):
    """A Float32Array store: round to nearest float32 (overflow to infinity)."""
    with np.errstate(over="ignore"):
        return float(np.float32(x))


def js_parse_float(s):
    """parseFloat (ECMA-262 21.1.2.12): the longest StrDecimalLiteral prefix, else NaN."""
    if s is None:
        return math.nan
    m = _F.match(s)
    if not m:
        return math.nan
    lit = m.group(1)
    if lit.lstrip("+-") == "Infinity":
        return -math.inf if lit.startswith("-") else math.inf
    return float(lit)


def js_parse_int(s):
    """parseInt(s, 10) (ECMA-262 21.1.2.13): sign and leading decimal digits, else NaN."""
    m = _I.match(s)
    if not m:
        return math.nan
    v = float(int(m.group(2)))
    return -v if m.group(1) == "-" else v


def js_hypot(*xs):
    """Math.hypot as V8 computes it: NaN / Infinity rules, then sqrt of Kahan-summed squares
    of the values scaled by the largest magnitude, times that magnitude."""
    ab = [abs(x) for x in xs]
    nan = any(math.isnan(a) for a in ab)
    mx = max([a for a in ab if not math.isnan(a)] + [0.0])
    if mx == math.inf:
        return math.inf
    if nan:
        return math.nan
    if mx == 0.0:
        return 0.0
    s = c = 0.0
    for a in ab:
        q = a / mx
        y = q * q - c
        t = s + y
        c = (t - s) - y
        s = t
    return math.sqrt(s) * mx


def vertex_normals(P, I):
    """computeVertexNormals over float32 positions P (flat list) and indices I (flat list)."""
    n = [0.0] * len(P)
    largest = [0.0] * len(P)
    largest_len = [0.0] * (len(P) // 3)
    for t in range(0, len(I), 3):
        ia, ib, ic = I[t] * 3, I[t + 1] * 3, I[t + 2] * 3
        e1 = [P[ib + k] - P[ia + k] for k in range(3)]
        e2 = [P[ic + k] - P[ia + k] for k in range(3)]
        cx = e1[1] * e2[2] - e1[2] * e2[1]
        cy = e1[2] * e2[0] - e1[0] * e2[2]
        cz = e1[0] * e2[1] - e1[1] * e2[0]
        cl = js_hypot(cx, cy, cz)
        for i in (ia, ib, ic):
            n[i] = f32(n[i] + cx)
            n[i + 1] = f32(n[i + 1] + cy)
            n[i + 2] = f32(n[i + 2] + cz)
            if cl > largest_len[i // 3]:
                largest_len[i // 3] = f32(cl)
                largest[i], largest[i + 1], largest[i + 2] = f32(cx), f32(cy), f32(cz)
    for i in range(0, len(n), 3):
        ln = js_hypot(n[i], n[i + 1], n[i + 2])
        if ln == 0:
            n[i], n[i + 1], n[i + 2] = largest[i], largest[i + 1], largest[i + 2]
            ln = largest_len[i // 3]
        if ln == 0:
            n[i], n[i + 1], n[i + 2] = 0.0, 0.0, 1.0
            continue
        n[i], n[i + 1], n[i + 2] = f32(n[i] / ln), f32(n[i + 1] / ln), f32(n[i + 2] / ln)
    return n


def _at(stream, i, per, k):
    """A JavaScript array read stream[i * per + k]: undefined (NaN when stored) off the end."""
    if not (i >= 0) or i != math.floor(i) or i * per + k >= len(stream):
        return math.nan
    return stream[int(i) * per + k]


def parse_obj(text, smooth_normals=False):
    """parseOBJ: returns (positions, indices, normals or None, uvs or None) as numpy arrays
    (float32 [V, 3], uint32 [T, 3], float32 [V, 3], float32 [V, 2]); raises ValueError when
    there are no faces."""
    v, vt, vn = [], [], []
    pos, nrm, uv, idx, has_n = [], [], [], [], []
    corners = {}
    used_n = used_uv = False

    def resolve(raw, count):
        return count + raw if raw < 0 else raw - 1

    def key(x):
        return "NaN" if math.isnan(x) else x

    def corner(tok):
        nonlocal used_n, used_uv
        f = tok.split("/")
        vi = resolve(js_parse_int(f[0]), len(v) // 3)
        ti = resolve(js_parse_int(f[1]), len(vt) // 2) if len(f) > 1 and f[1] != "" else -1
        ni = resolve(js_parse_int(f[2]), len(vn) // 3) if len(f) > 2 and f[2] != "" else -1
        k = (key(vi), key(ti), key(ni))
        if k in corners:
            return corners[k]
        out = len(pos) // 3
        pos.extend(_at(v, vi, 3, c) for c in range(3))
        if ti >= 0:
            uv.extend((_at(vt, ti, 2, 0), _at(vt, ti, 2, 1)))
            used_uv = True
        else:
            uv.extend((0.0, 0.0))
        if ni >= 0:
            nrm.extend(_at(vn, ni, 3, c) for c in range(3))
            used_n = True
        else:
            nrm.extend((0.0, 0.0, 0.0))
        has_n.append(ni >= 0)
        corners[k] = out
        return out

    if text.startswith("\ufeff"):
        text = text[1:]
    for line in text.split("\n"):
        s = line.strip(" \t\r\v\f")
        if s == "" or s.startswith("#"):
            continue
        parts = re.split(r"[ \t\r\v\f]+", s)
        tag = parts[0]
        get = lambda k: parts[k] if k < len(parts) else None
        if tag == "v":
            v.extend(js_parse_float(get(k)) for k in (1, 2, 3))
        elif tag == "vt":
            vt.extend((js_parse_float(get(1)), js_parse_float(get(2) if get(2) is not None else "0")))
        elif tag == "vn":
            vn.extend(js_parse_float(get(k)) for k in (1, 2, 3))
        elif tag == "f":
            c = [corner(t) for t in parts[1:]]
            for i in range(1, len(c) - 1):
                idx.extend((c[0], c[i], c[i + 1]))
    if not pos or not idx:
        raise ValueError("parseOBJ: no faces found (empty or unsupported OBJ)")
    P = [f32(x) for x in pos]
    normals = None
    if used_n:
        normals = [f32(x) for x in nrm]
        if not all(has_n):
            syn = vertex_normals(P, idx)
            for i, h in enumerate(has_n):
                if not h:
                    normals[3 * i: 3 * i + 3] = syn[3 * i: 3 * i + 3]
    elif smooth_normals:
        normals = vertex_normals(P, idx)
    arr = lambda a, w: np.asarray(a, np.float32).reshape(-1, w)
    return (arr(P, 3), np.asarray(idx, np.uint32).reshape(-1, 3), None if normals is None else arr(normals, 3),
            arr([f32(x) for x in uv], 2) if used_uv else None)


def parse_instances(buf):
    """parseInstances: returns a dict of count, flags, aabb, provenance, names and float32
    column arrays; raises ValueError('inst: ...') as the reference throws."""
    L = len(buf)
    if L < 48:
        raise ValueError(f"inst: buffer too small for a header ({L} bytes)")
    u32 = lambda o: struct.unpack_from("<I", buf, o)[0]
    if u32(0) != 0x54534E49:
        raise ValueError("inst: bad magic — not a '.inst' file")
    if u32(4) != 1:
        raise ValueError(f"inst: unsupported version {u32(4)} (loader speaks 1)")
    n = u32(8)
    if n == 0:
        raise ValueError("inst: count is 0 — nothing to place")
    flags = u32(12)
    if flags & ~7:
        raise ValueError(f"inst: unknown flag bits 0x{flags & ~7:x} — file written by a newer format version")
    box = struct.unpack_from("<6f", buf, 16)
    if not all(math.isfinite(x) for x in box) or any(box[a] > box[3 + a] for a in range(3)):
        raise ValueError("inst: header AABB is non-finite or empty")
    k = u32(40)
    ne = 44 + 32 * k
    if L < ne + 4:
        raise ValueError("inst: buffer truncated in the column-name table")
    names = []
    for c in range(k):
        raw = bytes(buf[44 + 32 * c: 76 + 32 * c])
        nm = raw.split(b"\0", 1)[0].decode("utf-8", errors="replace")
        if not nm:
            raise ValueError(f"inst: scalar column {c} has an empty name")
        names.append(nm)
    if len(set(names)) != len(names):
        raise ValueError("inst: duplicate scalar column names")
    pl = u32(ne)
    head = ne + 4 + pl + (4 - pl % 4) % 4
    if L < head:
        raise ValueError("inst: buffer truncated in the provenance string")
    prov = bytes(buf[ne + 4: ne + 4 + pl]).decode("utf-8", errors="replace")
    nf = 3 * n + (n if flags & 1 else 0) + (3 * n if flags & 2 else 0) + (4 * n if flags & 4 else 0) + k * n
    if L != head + 4 * nf:
        raise ValueError(f"inst: byte length {L} does not match header arithmetic ({head} header + {nf} floats) "
                         "— truncated or corrupt")
    off = head
    out = {"count": n, "flags": flags, "aabb": box, "provenance": prov, "names": names}

    def block(m):
        nonlocal off
        a = np.frombuffer(bytes(buf[off: off + 4 * m]), "<f4")
        off += 4 * m
        return a

    out["positions"] = block(3 * n).reshape(n, 3)
    out["sizes"] = block(n) if flags & 1 else None
    out["colors"] = block(3 * n).reshape(n, 3) if flags & 2 else None
    out["orientations"] = block(4 * n).reshape(n, 4) if flags & 4 else None
    out["scalars"] = {nm: block(n) for nm in names}
    return out
