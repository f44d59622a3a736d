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


def main():
    """Render the 64-scene batch workload (scenes.batch_workload) at a given launch shape, several
    times, printing each render's wall time: the command the batch's kernel-trace profile runs."""
    import argparse
    import time

    from . import compiler, scenes
    from .engine import Engine
    ap = argparse.ArgumentParser()
    ap.add_argument("--renders", type=int, default=3)
    ap.add_argument("--wave-paths", type=int, default=1 << 28)
    ap.add_argument("--trace-block", type=int, default=64)
    ap.add_argument("--trace-blocks-per-cu", type=int, default=32)
    ap.add_argument("--shade-blocks-per-cu", type=int, default=8)
    ap.add_argument("--stack", type=int, default=16)
    ap.add_argument("--no-graph", action="store_true", help="launch each kernel (so a kernel trace sees every one)")
    a = ap.parse_args()
    items = scenes.batch_workload()
    strategy = compiler.RenderStrategy("mis", 12, 3, 0.95, 64)
    comps = [compiler.compile_scene(sc, strategy, w, h) for sc, w, h, _ in items]
    spp = [n for _, _, _, n in items]
    eng = Engine(0)
    br = BatchRenderer(eng, comps)
    c = abi.PtBatchConfig()
    c.wave_paths, c.trace_block, c.trace_blocks_per_cu = a.wave_paths, a.trace_block, a.trace_blocks_per_cu
    c.shade_blocks_per_cu, c.stack, c.use_graph = a.shade_blocks_per_cu, a.stack, 0 if a.no_graph else 1
    br.config = c
    paths = sum(cc.width * cc.height * n for cc, n in zip(comps, spp))
    for r in range(a.renders):
        eng.finish()
        t0 = time.perf_counter()
        br.render(spp)
        eng.finish()
        dt = time.perf_counter() - t0
        print(f"batch render {r}: {dt:.4f} s, {paths / dt / 1e6:.1f} Mpaths/s", flush=True)
    br.close()
    eng.close()


if __name__ == "__main__":
    main()
