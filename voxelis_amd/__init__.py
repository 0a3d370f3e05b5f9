"""voxelis_amd: MI355X-native batched SVO-DAG edit path (apply_batch, hash-consing, get).

The compute path is libvoxelis_hip.so (hand-written gfx950 HIP kernels behind the
C ABI in include/voxelis_hip.h).  Importing the package does not touch the GPU;
the first interner loads the library and fails loudly if it is missing.
"""
from ._lib import LIB_PATH, STATUS_NAMES, VxError, lib  # noqa: F401
from .morton import encode_path  # noqa: F401


def __getattr__(name):
    # torch-backed classes load lazily so `import voxelis_amd` stays cheap
    if name in ("VoxInterner", "VoxTree", "Batch", "Forest", "EMPTY", "INVALID"):
        from . import api
        return getattr(api, name)
    raise AttributeError(name)
