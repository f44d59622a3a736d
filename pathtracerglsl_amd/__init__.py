"""MI355X-native transport-loop backend for PathTracerGLSL scenes (libptmi + host compiler)."""
from . import abi, compiler, scenes  # noqa: F401
