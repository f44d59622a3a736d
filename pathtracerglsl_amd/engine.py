"""Engine — the host-side executor over libptmi, mirroring the reference Engine's surface
(src/engine/Engine.ts:83-474): loadRenderer, renderFrame(s), readExport, clearAccumulation,
pinResetSalt, sampleCount; plus the traversal parity entry points."""
import ctypes as C

import numpy as np

from . import _lib, abi


class Engine:
    def __init__(self, device=0):
        self.lib = _lib.load()
        h = C.c_void_p()
        _lib.check(self.lib.pt_ctx_create(device, C.byref(h)))
        self.ctx = h
        self._scene = None
        self._renderers = {}
        self._selected = None
        self._salt = None

    # -- scene data (registerDataTexture / registerEnvironmentTable) --
    def upload_scene(self, arrays):
        self._scene = arrays
        self._desc = arrays.desc()
        _lib.check(self.lib.pt_scene_upload(self.ctx, C.byref(self._desc)))

    # -- renderers --
    def load_renderer(self, compiled):
        """Engine.loadRenderer: uploads the compiled scene and selects the kernel instantiation."""
        if self._scene is not compiled.scene:
            self.upload_scene(compiled.scene)
        h = C.c_void_p()
        prog = compiled.program
        _lib.check(self.lib.pt_renderer_load(self.ctx, C.byref(prog), C.byref(h)))
        self._renderers[compiled.id] = (h, compiled)
        self._selected = compiled.id
        return compiled.id

    def load_renderer_description(self, description, compiled):
        """Engine.loadRenderer from the reference's plan record: the ProgramDescription (a dict or
        its JSON text) is lowered by the library (pt_renderer_load_description) over compiled's
        scene, film size and frames in flight; compiled also supplies the per-frame params."""
        import json
        if self._scene is not compiled.scene:
            self.upload_scene(compiled.scene)
        text = (description if isinstance(description, str) else json.dumps(description)).encode()
        h = C.c_void_p()
        _lib.check(self.lib.pt_renderer_load_description(self.ctx, text, len(text), compiled.width, compiled.height,
                                                         compiled.program.frames_in_flight, C.byref(h)))
        rid = f"{compiled.id}#description"
        if rid in self._renderers:
            self.lib.pt_renderer_unload(self._renderers[rid][0])
        self._renderers[rid] = (h, compiled)
        self._selected = rid
        return rid

    def select_renderer(self, rid):
        """Engine.selectRenderer: the renderer's own scene becomes resident again when another
        load replaced it (every render call checks the program against the resident scene)."""
        if rid not in self._renderers:
            raise KeyError(f"unknown renderer '{rid}'")
        _, compiled = self._renderers[rid]
        if self._scene is not compiled.scene:
            self.upload_scene(compiled.scene)
        self._selected = rid

    def _cur(self):
        if self._selected is None:
            raise RuntimeError("no renderer loaded")
        return self._renderers[self._selected]

    def pin_reset_salt(self, salt):
        self._salt = int(salt)

    def render_frames(self, n, params=None):
        h, comp = self._cur()
        prm = params if params is not None else comp.params(self._salt)
        _lib.check(self.lib.pt_render_frames(h, C.byref(prm), int(n)))

    def render_frame(self, params=None):
        self.render_frames(1, params)

    # -- differentiation (the adjoint pair) --
    def forward(self, n, params=None):
        """pt_forward: clear, then render n frames from sampleCount 0 (the primal image)."""
        h, comp = self._cur()
        prm = params if params is not None else comp.params(self._salt)
        _lib.check(self.lib.pt_forward(h, C.byref(prm), int(n)))

    def backward(self, n, adj, dL_dimage, params=None):
        """pt_backward: dL/dparams (abi.ADJ_N_PARAMS floats) of the replayed n-frame forward pass
        for the adjoint image dL_dimage [H][W][4] (row 0 = bottom)."""
        h, comp = self._cur()
        prm = params if params is not None else comp.params(self._salt)
        dL = np.ascontiguousarray(dL_dimage, np.float32)
        if dL.shape != (comp.height, comp.width, 4):
            raise ValueError(f"dL_dimage must be [{comp.height}][{comp.width}][4]")
        out = np.zeros(abi.ADJ_N_PARAMS, np.float32)
        _lib.check(self.lib.pt_backward(h, C.byref(prm), int(n), C.byref(adj), dL.ctypes.data, out.ctypes.data))
        return out

    def update_materials(self, materials):
        """pt_scene_update_materials: new values for the uploaded scene's material records."""
        m = np.ascontiguousarray(materials, dtype=abi.MATERIAL_DTYPE)
        _lib.check(self.lib.pt_scene_update_materials(self.ctx, m.ctypes.data, len(m)))
        if self._scene is not None:
            self._scene.materials = m

    @property
    def sample_count(self):
        h, _ = self._cur()
        return int(self.lib.pt_sample_count(h))

    def clear_accumulation(self):
        h, _ = self._cur()
        _lib.check(self.lib.pt_clear_accumulation(h))

    def read_export(self, name="radiance"):
        h, comp = self._cur()
        out = np.empty((comp.height, comp.width, 4), np.float32)
        _lib.check(self.lib.pt_read_export(h, name.encode(), out.ctypes.data, out.size))
        return out

    def ray_counts(self):
        h, _ = self._cur()
        a, b = C.c_uint64(), C.c_uint64()
        _lib.check(self.lib.pt_last_ray_counts(h, C.byref(a), C.byref(b)))
        return int(a.value), int(b.value)

    def set_traversal_counting(self, on):
        h, _ = self._cur()
        _lib.check(self.lib.pt_renderer_set_traversal_counting(h, int(bool(on))))

    def traversal_counts(self):
        h, _ = self._cur()
        a, b = C.c_uint64(), C.c_uint64()
        _lib.check(self.lib.pt_last_traversal_counts(h, C.byref(a), C.byref(b)))
        return int(a.value), int(b.value)

    def finish(self):
        _lib.check(self.lib.pt_finish(self.ctx))

    # -- timing --
    def set_timing(self, on):
        _lib.check(self.lib.pt_set_timing(self.ctx, int(bool(on))))

    def timings(self):
        buf = (abi.PtKernelTiming * 32)()
        n = self.lib.pt_get_timings(self.ctx, buf, 32)
        return {buf[i].name.decode(): (buf[i].total_ms, buf[i].launches, buf[i].items) for i in range(min(n, 32))}

    # -- traversal parity entry points --
    def trace_closest(self, rays):
        rays = np.ascontiguousarray(rays, np.float32)
        n = rays.shape[1]
        soa = abi.PtRaysSoa(*[rays[k].ctypes.data for k in range(7)])
        t = np.empty(n, np.float32)
        p = np.empty(n, np.int32)
        _lib.check(self.lib.pt_trace_closest(self.ctx, C.byref(soa), n, t.ctypes.data, p.ctypes.data))
        return t, p

    def trace_any(self, rays):
        rays = np.ascontiguousarray(rays, np.float32)
        n = rays.shape[1]
        soa = abi.PtRaysSoa(*[rays[k].ctypes.data for k in range(7)])
        h = np.empty(n, np.uint8)
        _lib.check(self.lib.pt_trace_any(self.ctx, C.byref(soa), n, h.ctypes.data))
        return h

    def scene_stats(self):
        a, b, c = C.c_uint32(), C.c_uint32(), C.c_uint32()
        _lib.check(self.lib.pt_scene_stats(self.ctx, C.byref(a), C.byref(b), C.byref(c)))
        return {"wide_nodes": a.value, "wide_depth": b.value, "stack": c.value}

    def close(self):
        if self.ctx:
            for h, _ in self._renderers.values():
                self.lib.pt_renderer_unload(h)
            self._renderers.clear()
            self.lib.pt_ctx_destroy(self.ctx)
            self.ctx = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()
