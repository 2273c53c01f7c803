"""MI355X-native ByteCRT: exact negacyclic products mod (X^N+1, q) on gfx950.

Load from the repository root with ``__graft_entry__.load_package()`` (the
directory name is not a Python identifier).
"""
from . import bytecrt, datagen, probes  # noqa: F401
from .bytecrt import (Engine, Inputs, Mix, Outputs, Point, Server, Tensor, fherma_free, fherma_init,  # noqa: F401
                      fherma_run, load_library)
