"""MI355X-native evaluator for fmmgen's octree FMM path.

* ``generate_code`` -- the operator generator (host C/C++ and HIP device
  printers), same interface as ``fmmgen.generate_code``;
* ``build_tree`` / ``Tree`` -- the example's tree / evaluate API on one GPU,
  backed by libfmm_hip.so (HIP kernels for gfx950 behind a C ABI).
"""
from .codegen import generate_code, Nterms, Msize, Lsize, generate_mappings

__all__ = ["generate_code", "Nterms", "Msize", "Lsize", "generate_mappings", "build_tree", "Tree",
           "direct_sum"]


def __getattr__(name):
    # torch is imported lazily so the generator works without it
    if name in ("build_tree", "Tree", "direct_sum", "op_batch"):
        from . import tree
        return getattr(tree, name)
    raise AttributeError(name)
