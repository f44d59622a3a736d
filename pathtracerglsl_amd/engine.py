"""Engine — the host-side executor over libptmi, mirroring the reference Engine's surface
(src/engine/Engine.ts:83-474): loadRenderer, renderFrame(s), readExport, clearAccumulation,
pinResetSalt, sampleCount; plus the traversal parity entry points."""
import ctypes as C

import numpy as np

from . import _lib, abi

# Live parameter roots (the dotted paths of Engine.setParameter, Engine.ts:258): the authored
# field each one drives. Geometry is not a parameter: light geometry rows are constant by the
# reference's contract (resolveLightValues, features/lighting.ts:337-348).
ENV_PARAMS = {"intensity": "intensity", "rotation": "rotation", "selectProb": "select_prob", "color": "color",
              "totalWeight": "total_weight"}
LIGHT_EMISSION_ROWS = ("radiance", "intensity")


def with_parameter(source, name, value):
    """The authored scene with parameter `name` set to value: 'camera.<field>', 'env.intensity' |
    'env.rotation' | 'env.selectProb' | 'env.color' | 'env.totalWeight', 'lights.<i>.radiance' |
    'lights.<i>.intensity' (the emission row) and 'materials.<name>.<field>' (a medium's field as
    'materials.<name>.medium.<field>'). Only the branch the path names is copied."""
    parts = name.split(".")
    out = dict(source)
    if parts[0] == "camera" and len(parts) == 2:
        out["camera"] = dict(source["camera"], **{parts[1]: value})
    elif parts[0] == "env" and len(parts) == 2 and parts[1] in ENV_PARAMS:
        out["environment"] = dict(source.get("environment", {"type": "none"}), **{ENV_PARAMS[parts[1]]: value})
    elif parts[0] == "lights" and len(parts) == 3:
        i = int(parts[1])
        lights = list(source.get("lights", []))
        if not 0 <= i < len(lights):
            raise KeyError(f"parameter '{name}': no light {i}")
        if parts[2] not in LIGHT_EMISSION_ROWS or parts[2] not in lights[i]:
            raise KeyError(f"parameter '{name}': a light's live row is its emission ({lights[i].get('type')} "
                           f"light: '{'radiance' if 'radiance' in lights[i] else 'intensity'}')")
        lights[i] = dict(lights[i], **{parts[2]: value})
        out["lights"] = lights
    elif parts[0] == "materials" and len(parts) in (3, 4):
        mats = dict(source.get("materials", {}))
        if parts[1] not in mats:
            raise KeyError(f"parameter '{name}': no material '{parts[1]}'")
        m = dict(mats[parts[1]])
        if len(parts) == 4:
            if parts[2] != "medium" or m.get("medium") is None:
                raise KeyError(f"parameter '{name}': unknown material field")
            m["medium"] = dict(m["medium"], **{parts[3]: value})
        elif parts[2] == "model":
            raise KeyError(f"parameter '{name}': the material model selects a kernel; compile a new renderer")
        else:
            m[parts[2]] = value
        mats[parts[1]] = m
        out["materials"] = mats
    else:
        raise KeyError(f"unknown parameter '{name}'")
    return out


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
        self._params = {}
        self._env_table = None
        self._descriptions = {}   # renderers loaded from a ProgramDescription: its JSON text
        self.state = "ready"
        self._lost_selected = None

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
        if compiled.id in self._renderers:           # a reload replaces the renderer of that id
            self.unload_renderer(compiled.id)
        self._renderers[compiled.id] = (h, compiled)
        self._selected = compiled.id
        return compiled.id

    def unload_renderer(self, rid):
        """Release a loaded renderer (its film and path state)."""
        h, _ = self._renderers.pop(rid)
        self._descriptions.pop(rid, None)
        if h:
            self.lib.pt_renderer_unload(h)
        if self._selected == rid:
            self._selected = None

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
            self.unload_renderer(rid)
        self._renderers[rid] = (h, compiled)
        self._descriptions[rid] = text
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

    # -- live parameters (setParameter / getParameter / getAllParameters, Engine.ts:258-268) --
    def set_parameter(self, name, value):
        """Engine.setParameter: the selected renderer's scene with one parameter changed. The
        host recompiles the authored scene with the new value (the reference's CPU-side compute
        of derived uniforms: light selection table, env selection probability, majorants) and
        sends the changed records to the resident scene (pt_scene_update_materials / _lights /
        _env); the camera rides the per-frame params. Geometry never changes. Accumulation is
        not cleared (the caller clears, as the reference's triggersReset does)."""
        from . import compiler
        _, comp = self._cur()
        src = with_parameter(comp.source, name, value)
        new = compiler.compile_scene(src, comp.strategy, comp.width, comp.height, comp.reset_salt)
        if self._scene is not comp.scene:
            self.upload_scene(comp.scene)
        self._apply_records(comp.scene, new.scene)
        comp.source = src
        comp.camera = new.camera
        self._params[name] = value

    def get_parameter(self, name):
        return self._params.get(name)

    def get_all_parameters(self):
        return dict(self._params)

    def _apply_records(self, sc, ns):
        if not (np.array_equal(sc.tri_vertices, ns.tri_vertices) and sc.objects.tobytes() == ns.objects.tobytes()):
            raise ValueError("a live parameter may not change the scene's geometry")
        if sc.materials.tobytes() != ns.materials.tobytes():
            self.update_materials(ns.materials)
        if sc.lights.tobytes() != ns.lights.tobytes():
            _lib.check(self.lib.pt_scene_update_lights(self.ctx, ns.lights.ctypes.data, len(ns.lights)))
            sc.lights[...] = ns.lights
        if bytes(sc.env) != bytes(ns.env):
            _lib.check(self.lib.pt_scene_update_env(self.ctx, C.byref(ns.env)))
            C.memmove(C.byref(sc.env), C.byref(ns.env), C.sizeof(abi.PtEnv))

    # -- environment table (registerEnvironmentTable, Engine.ts:450-460) --
    def register_environment_table(self, rgb, width, height, chart="equirect", compensation=False):
        """The resident sky's importance table rebuilt from an RGB picture [height][width][3] on
        `chart` (pt_env_table_upload); returns {'totalWeight': mass}. The table's size, chart and
        mass become the env.size / env.totalWeight the derived env selection probability reads,
        and the picture is kept so a context restore registers it again."""
        from .compiler import ENV_CHARTS
        a = np.ascontiguousarray(rgb, np.float32)
        if a.size != width * height * 3:
            raise ValueError(f"rgb must hold {height} x {width} x 3 floats")
        if chart not in ENV_CHARTS:
            raise ValueError(f"chart {chart!r}: expected one of {sorted(ENV_CHARTS)}")
        tw = C.c_double()
        _lib.check(self.lib.pt_env_table_upload(self.ctx, a.ctypes.data, int(width), int(height), ENV_CHARTS[chart],
                                                int(bool(compensation)), C.byref(tw)))
        self._env_table = (a.copy(), int(width), int(height), chart, bool(compensation))
        if self._selected is not None:
            _, comp = self._cur()
            env = dict(comp.source.get("environment", {}), table_size=(int(width), int(height)), chart=chart,
                       compensated=bool(compensation), total_weight=tw.value)
            comp.source = dict(comp.source, environment=env)
            from . import compiler
            new = compiler.compile_scene(comp.source, comp.strategy, comp.width, comp.height, comp.reset_salt)
            self._apply_records(comp.scene, new.scene)
        return {"totalWeight": tw.value}

    # -- context loss and restore (Engine._handleContextLoss / _handleContextRestored, Engine.ts:476-520) --
    def lose_context(self):
        """Every device resource goes (pt_ctx_reset); the compiled renderers are kept as data.
        Rendering no-ops while lost, as renderFrame does (Engine.ts:193)."""
        _lib.check(self.lib.pt_ctx_reset(self.ctx))
        for h, _ in self._renderers.values():
            self.lib.pt_renderer_unload(h)
        self._renderers = {rid: (None, comp) for rid, (_, comp) in self._renderers.items()}
        self._lost_selected = self._selected
        self._selected = None
        self._scene = None
        self.state = "context-lost"

    def restore_context(self):
        """Rebuild from the retained compiled renderers: upload the scene, register the last
        environment table again, load every renderer (fresh, zeroed accumulation) and select the
        one that was active."""
        kept = self._renderers
        self._renderers = {}
        self.state = "ready"
        for rid, (_, compiled) in kept.items():
            if self._scene is not compiled.scene:
                self.upload_scene(compiled.scene)
            h = C.c_void_p()
            text = self._descriptions.get(rid)
            if text is not None:
                _lib.check(self.lib.pt_renderer_load_description(self.ctx, text, len(text), compiled.width, compiled.height,
                                                                 compiled.program.frames_in_flight, C.byref(h)))
            else:
                _lib.check(self.lib.pt_renderer_load(self.ctx, C.byref(compiled.program), C.byref(h)))
            self._renderers[rid] = (h, compiled)
        if self._lost_selected in self._renderers:
            self.select_renderer(self._lost_selected)
        if self._env_table is not None:
            rgb, w, h2, chart, comp = self._env_table
            tw = C.c_double()
            _lib.check(self.lib.pt_env_table_upload(self.ctx, rgb.ctypes.data, w, h2, abi.CHART_OCTAHEDRAL
                                                    if chart == "octahedral" else abi.CHART_EQUIRECT, int(comp),
                                                    C.byref(tw)))

    def pin_reset_salt(self, salt):
        self._salt = int(salt)

    def render_frames(self, n, params=None):
        if self.state == "context-lost":
            return
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
                if h:
                    self.lib.pt_renderer_unload(h)
            self._renderers.clear()
            self.lib.pt_ctx_destroy(self.ctx)
            self.ctx = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()
