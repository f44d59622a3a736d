"""Session-level cache of compiled workloads shared by the GPU tests."""
import functools

from pathtracerglsl_amd import compiler, scenes  # noqa: F401


@functools.lru_cache(maxsize=None)
def cornell(est="mis", w=64, h=64, bounces=8, rr=3, fif=1):
    return compiler.compile_scene(scenes.cornell_box(), compiler.RenderStrategy(est, bounces, rr, 0.95, fif), w, h)


@functools.lru_cache(maxsize=None)
def cloud_scene():
    return scenes.icosphere_cloud()


@functools.lru_cache(maxsize=None)
def cloud(est="mis", w=64, h=36, bounces=12, rr=3, fif=1):
    return compiler.compile_scene(cloud_scene(), compiler.RenderStrategy(est, bounces, rr, 0.95, fif), w, h)
