"""On-disk scene inputs: Wavefront OBJ meshes and `.inst` instance clouds.

The parsing runs in libptmi's C++ readers (pt_obj_parse, pt_inst_parse); this module is the
host-side mirror of the reference's authoring entry points around them:
- parse_obj / load_obj: src/authoring/loadOBJ.ts:62-170 (parseOBJ, loadOBJ);
- parse_instances / load_instances: src/authoring/loadInstances.ts:30-109;
- encode_instances: the write side, tools/inst-format.mjs:39-120 (encodeInstances);
- instance_cloud: src/authoring/instance.ts:86-135 (instanceCloud) over a mesh prototype.

A parsed OBJ is a MeshObject-shaped dict (compiler/types.ts:134-158) that compile_scene takes
as it is. A parsed `.inst` is an InstanceTable of numpy views over the file's bytes.
"""
import ctypes as C
import struct
from dataclasses import dataclass, field

import numpy as np

from . import _lib, abi
from .compiler import CompileError


def parse_obj(text, material, transform=None, name=None, smooth_normals=False, closed=None):
    """parseOBJ(text, opts) (authoring/loadOBJ.ts:62-163) as a MeshObject dict: float32
    positions [V, 3] in object-local space, uint32 indices [T, 3], normals [V, 3] and uvs
    [V, 2] when the file has them (normals also when smooth_normals asks for synthesized
    ones). Raises _lib.PtError when the text has no faces."""
    if isinstance(text, str):
        text = text.encode("utf-8")
    lib = _lib.load()
    h = C.c_void_p()
    _lib.check(lib.pt_obj_parse(text, len(text), abi.OBJ_SMOOTH_NORMALS if smooth_normals else 0, C.byref(h)))
    try:
        nv, ni, hn, hu = C.c_uint32(), C.c_uint32(), C.c_int32(), C.c_int32()
        _lib.check(lib.pt_mesh_data_info(h, C.byref(nv), C.byref(ni), C.byref(hn), C.byref(hu)))
        pos = np.zeros((nv.value, 3), np.float32)
        idx = np.zeros((ni.value // 3, 3), np.uint32)
        nrm = np.zeros((nv.value, 3), np.float32) if hn.value else None
        uvs = np.zeros((nv.value, 2), np.float32) if hu.value else None
        ptr = lambda a: None if a is None else a.ctypes.data
        _lib.check(lib.pt_mesh_data_copy(h, ptr(pos), ptr(idx), ptr(nrm), ptr(uvs)))
    finally:
        lib.pt_mesh_data_free(h)
    mesh = {"type": "mesh", "positions": pos, "indices": idx, "material": material}
    if nrm is not None:
        mesh["normals"] = nrm
    if uvs is not None:
        mesh["uvs"] = uvs
    if transform is not None:
        mesh["transform"] = transform
    if name is not None:
        mesh["name"] = name
    if closed is not None:
        mesh["closed"] = bool(closed)
    return mesh


def load_obj(path, material, **opts):
    """loadOBJ(url, opts) (authoring/loadOBJ.ts:166-170) for a file path."""
    with open(path, "rb") as f:
        return parse_obj(f.read(), material, **opts)


@dataclass
class InstanceTable:
    """InstanceTable (authoring/loadInstances.ts:9-22): typed views over the parsed buffer."""
    count: int
    aabb_min: tuple
    aabb_max: tuple
    provenance: str
    positions: np.ndarray                       # [N, 3] f32
    sizes: np.ndarray = None                    # [N]
    colors: np.ndarray = None                   # [N, 3] linear RGB
    orientations: np.ndarray = None             # [N, 4] unit quaternions x y z w
    scalars: dict = field(default_factory=dict)  # name -> [N] f32, in file order


def parse_instances(buffer):
    """parseInstances(buffer) (authoring/loadInstances.ts:30-98). `buffer` is bytes-like; the
    returned arrays are read-only views of it. Raises _lib.PtError ("inst: ...") on a corrupt
    or unsupported file."""
    buf = bytes(buffer)
    tab = abi.PtInstTable()
    _lib.check(_lib.load().pt_inst_parse(buf, len(buf), C.byref(tab)))
    n = tab.count

    def block(off, per):
        a = np.frombuffer(buf, np.float32, count=per * n, offset=off)
        return a.reshape(n, per) if per > 1 else a

    names = []
    for c in range(tab.n_scalars):
        raw = buf[tab.names_offset + 32 * c: tab.names_offset + 32 * (c + 1)]
        names.append(raw.split(b"\0", 1)[0].decode("utf-8", errors="replace"))
    prov = buf[tab.provenance_offset: tab.provenance_offset + tab.provenance_len].decode("utf-8", errors="replace")
    scal = {nm: np.frombuffer(buf, np.float32, count=n, offset=tab.scalars_offset + 4 * n * c)
            for c, nm in enumerate(names)}
    return InstanceTable(
        count=n, aabb_min=tuple(tab.aabb[:3]), aabb_max=tuple(tab.aabb[3:]), provenance=prov,
        positions=block(tab.positions_offset, 3),
        sizes=block(tab.sizes_offset, 1) if tab.flags & 1 else None,
        colors=block(tab.colors_offset, 3) if tab.flags & 2 else None,
        orientations=block(tab.orientations_offset, 4) if tab.flags & 4 else None,
        scalars=scal)


def load_instances(path):
    """loadInstances(src) (authoring/loadInstances.ts:101-109) for a file path."""
    with open(path, "rb") as f:
        return parse_instances(f.read())


def encode_instances(positions, sizes=None, colors=None, orientations=None, scalars=None, provenance=""):
    """encodeInstances(table) (tools/inst-format.mjs:39-120): the version-1 `.inst` bytes of
    an instance table (float32 columns; the header box is the positions' box). Raises
    ValueError for the inputs the reference's encoder rejects."""
    P = np.asarray(positions, np.float32).reshape(-1)
    if P.size == 0 or P.size % 3:
        raise ValueError("inst encode: positions must be a nonempty Float32Array of length 3N")
    n = P.size // 3
    cols = []
    for nm, arr, per in (("sizes", sizes, 1), ("colors", colors, 3), ("orientations", orientations, 4)):
        if arr is not None:
            a = np.asarray(arr, np.float32).reshape(-1)
            if a.size != per * n:
                raise ValueError(f"inst encode: {nm} must be a Float32Array of length {per}N ({per * n}), got {a.size}")
            cols.append(a)
        else:
            cols.append(None)
    scalars = dict(scalars or {})
    name_bytes = []
    for nm in scalars:
        b = nm.encode("utf-8")
        if not 0 < len(b) < 32:
            raise ValueError(f"inst encode: scalar column name '{nm}' must be 1..31 UTF-8 bytes")
        name_bytes.append(b)
    sc = []
    for nm, arr in scalars.items():
        a = np.asarray(arr, np.float32).reshape(-1)
        if a.size != n:
            raise ValueError(f"inst encode: scalars.{nm} must be a Float32Array of length 1N ({n}), got {a.size}")
        sc.append(a)
    bad = np.flatnonzero(~np.isfinite(P))
    if bad.size:
        k = int(bad[0])
        raise ValueError(f"inst encode: positions[{k}] (instance {k // 3}) is not finite")
    pv = P.reshape(n, 3)
    flags = (1 if cols[0] is not None else 0) | (2 if cols[1] is not None else 0) | (4 if cols[2] is not None else 0)
    prov = provenance.encode("utf-8")
    head = struct.pack("<IIII", 0x54534E49, 1, n, flags) + pv.min(axis=0).tobytes() + pv.max(axis=0).tobytes()
    head += struct.pack("<I", len(name_bytes)) + b"".join(b.ljust(32, b"\0") for b in name_bytes)
    head += struct.pack("<I", len(prov)) + prov + b"\0" * ((4 - len(prov) % 4) % 4)
    body = b"".join(a.astype("<f4").tobytes() for a in [P] + [c for c in cols if c is not None] + sc)
    return head + body


def instance_cloud(table, prototype, material=None, name=None, size_scale=None, size=None, use_colors=False):
    """instanceCloud(table, opts) (authoring/instance.ts:86-135) as one instance batch with
    packed placements: positions, orientations where the file has them, and sizes (the file's
    column, or the `size(scalars, i)` hook, times size_scale) as uniform scales of
    `prototype`, a mesh object (triangle or indexed form) whose material the batch takes
    unless `material` is given. The reference's unit analytic prototypes ('sphere', 'cube')
    and per-instance colour attributes are not provided by this backend: a table with colours
    needs use_colors=False, which renders the batch material."""
    if isinstance(prototype, str):
        raise CompileError(f"instanceCloud: analytic prototype '{prototype}' is not supported by this backend; "
                           "pass a mesh prototype")
    if table.colors is not None and use_colors:
        raise CompileError("instanceCloud: per-instance colour attributes are not supported by this backend")
    from .compiler import indexed_mesh_object
    proto = indexed_mesh_object(dict(prototype, transform=None) if "positions" in prototype else prototype)
    n = table.count
    sizes = table.sizes
    if size is not None:
        sizes = np.array([size(table.scalars, i) for i in range(n)], np.float32)
    if size_scale is not None and size_scale != 1:       # f64 product, stored as float32
        base = sizes.astype(np.float64) if sizes is not None else np.ones(n)
        sizes = (float(size_scale) * base).astype(np.float32)
    obj = {"type": "instances", "triangles": proto["triangles"], "material": material or prototype["material"],
           "closed": bool(proto.get("closed", False)),
           "placements": {"count": n, "positions": table.positions, "orientations": table.orientations,
                          "sizes": sizes}}
    if proto.get("normals") is not None:
        obj["normals"] = proto["normals"]
    if name is not None:
        obj["name"] = name
    return obj


__all__ = ["parse_obj", "load_obj", "InstanceTable", "parse_instances", "load_instances", "encode_instances",
           "instance_cloud"]
