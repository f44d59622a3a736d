"""Batched persistent wavefront over libptmi's pt_batch_* entry points: many compiled scenes of
uneven size rendered as one hipGraph-captured sequence of mixed-scene waves, with the
autotuner that picks wave size, traversal block size, grid and stack depth."""
import ctypes as C

import numpy as np

from . import _lib, abi


class BatchRenderer:
    def __init__(self, engine, compiled):
        """engine: an Engine (its context and stream); compiled: CompiledRenderers."""
        self.engine = engine
        self.lib = engine.lib
        self.compiled = list(compiled)
        n = len(self.compiled)
        self._descs = (abi.PtSceneDesc * n)(*[c.scene.desc() for c in self.compiled])
        self._progs = (abi.PtProgram * n)(*[c.program for c in self.compiled])
        self._params = (abi.PtParams * n)(*[c.params() for c in self.compiled])
        h = C.c_void_p()
        _lib.check(self.lib.pt_batch_create(engine.ctx, n, C.addressof(self._descs), C.addressof(self._progs),
                                            C.addressof(self._params), C.byref(h)))
        self.h = h

    @property
    def config(self):
        c = abi.PtBatchConfig()
        _lib.check(self.lib.pt_batch_get_config(self.h, C.byref(c)))
        return c

    @config.setter
    def config(self, c):
        _lib.check(self.lib.pt_batch_set_config(self.h, C.byref(c)))

    def render(self, spp):
        """Clear and render spp[s] frames of every scene s (asynchronous)."""
        s = np.ascontiguousarray(spp, np.uint32)
        assert s.size == len(self.compiled)
        self._spp = s
        _lib.check(self.lib.pt_batch_render(self.h, s.ctypes.data))

    def read_export(self, scene):
        c = self.compiled[scene]
        out = np.empty((c.height, c.width, 4), np.float32)
        _lib.check(self.lib.pt_batch_read_export(self.h, int(scene), out.ctypes.data, out.size))
        return out

    def autotune(self, spp, probe_div=8, max_results=64):
        """Sweep the launch configuration on a reduced-spp copy of the batch; keeps the fastest."""
        s = np.ascontiguousarray(spp, np.uint32)
        best = abi.PtBatchConfig()
        res = (abi.PtBatchConfig * max_results)()
        ms = np.zeros(max_results, np.float32)
        n = C.c_int32()
        _lib.check(self.lib.pt_batch_autotune(self.h, s.ctypes.data, int(probe_div), C.byref(best), C.addressof(res),
                                              ms.ctypes.data, max_results, C.byref(n)))
        sweep = [(res[i].as_dict(), float(ms[i])) for i in range(min(n.value, max_results))]
        return best, sweep

    def close(self):
        if self.h:
            self.lib.pt_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
