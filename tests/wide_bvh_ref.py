"""Python model of the wide-BVH node format and of the kernel's walk, for CPU-side checks of
the builder (structure, containment, and a scalar replay of the traversal order logic)."""
import ctypes as C

import numpy as np

from pathtracerglsl_amd import _lib


def build(tri_vertices):
    lib = _lib.load()
    f = lib.pt_debug_wide_bvh
    f.restype = C.c_longlong
    f.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_longlong, C.c_void_p, C.c_longlong,
                  C.POINTER(C.c_longlong), C.POINTER(C.c_uint32)]
    tv = np.ascontiguousarray(tri_vertices, np.float32)
    nt, depth = C.c_longlong(), C.c_uint32()
    nn = f(tv.ctypes.data, len(tv), None, 0, None, 0, C.byref(nt), C.byref(depth))
    nodes = np.zeros((nn, 20), np.uint32)
    tris = np.zeros((nt.value, 12), np.float32)
    f(tv.ctypes.data, len(tv), nodes.ctypes.data, nn, tris.ctypes.data, nt.value, C.byref(nt), C.byref(depth))
    return nodes, tris, depth.value


def decode(node):
    w = [int(x) for x in node]
    p = np.array(w[0:3], np.uint32).view(np.float32)
    e = [(w[3] >> (8 * a)) & 0xFF for a in range(3)]
    s = np.array([e_ << 23 for e_ in e], np.uint32).view(np.float32)
    imask = w[3] >> 24
    meta = [(w[6 + (i >> 2)] >> (8 * (i & 3))) & 0xFF for i in range(8)]

    def q(word0):
        return [(w[word0 + (i >> 2)] >> (8 * (i & 3))) & 0xFF for i in range(8)]
    qlo = [q(8), q(10), q(12)]
    qhi = [q(14), q(16), q(18)]
    children = []
    for i in range(8):
        if meta[i] == 0:
            continue
        lo = np.array([np.float32(p[a]) + np.float32(qlo[a][i]) * s[a] for a in range(3)], np.float32)
        hi = np.array([np.float32(p[a]) + np.float32(qhi[a][i]) * s[a] for a in range(3)], np.float32)
        internal = bool(imask >> i & 1)
        children.append({"slot": i, "internal": internal, "meta": meta[i], "lo": lo, "hi": hi,
                         "child": w[4] + bin(imask & ((1 << i) - 1)).count("1") if internal else None,
                         "tris": None if internal else list(range(w[5] + (meta[i] & 31),
                                                                 w[5] + (meta[i] & 31) + bin(meta[i] >> 5).count("1")))})
    return children
