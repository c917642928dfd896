"""ngap_amd: MI355X-native non-blocking automata processing (ngAP's worklist
path) — HIP kernels for gfx950 behind a C ABI (include/ngap_amd.h)."""
from ._lib import ALGORITHMS, NgapError  # noqa: F401
from .engine import Engine, decode_reports  # noqa: F401
from .groups import Frontend, Group, load_anml, make_group, parse_symbol_set, write_anml  # noqa: F401
