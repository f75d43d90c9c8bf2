"""naivedynamics.jl_amd -- MI355X (gfx950) implementation of NaiveDynamics.jl's
per-step hot path behind the reference's entry points.

The directory name is not a Python identifier; load the package with
`load_package()` from __graft_entry__ or import it under an alias:

    spec = importlib.util.spec_from_file_location(
        "naivedynamics_amd", ".../naivedynamics.jl_amd/__init__.py",
        submodule_search_locations=[".../naivedynamics.jl_amd"])
"""
from . import _lib, adjoint, batched, bvh, cells, forces, simulator, synth
from ._lib import (MORTON_HILBERT, MORTON_INTERLEAVE, MORTON_MASKED, NODE_DTYPE, PAIR_DTYPE, PRED_SCALAR,
                   PRED_SIMD, PREC_FP32, PREC_FP64, PREC_MIXED, QUERY_OWN_LEAF, QUERY_ROOT,
                   CapacityError, Context, DeviceArray, DomainError, NDError, SpecError,
                   default_context)
from .bvh import (SpheresBVHSpecs, TreeData, binwidth, bounding_volume_hierarchy,
                  build_traverse_bvh, canonical, exptsimdbuild_traverse_bvh, gpubvh_neighborlist,
                  leafbuild_traverse_bvh, leafneighbor_traverse, leafTreeData, mortoncodes,
                  neighbor_traverse, neighborlist_device, quantized_positions,
                  simdbuild_traverse_bvh, sortperm, threshold_unique_pairs)
from .cells import neighborlist as cell_neighborlist
from .forces import force_coulomb_, force_lennardjones_, sum_forces_
from .simulator import ObjectCollection, SimSpec, Simulation, simulate_, simulate_bvh_
from .batched import BatchedSimulation, autotune
from .adjoint import Adjoint

LIB_PATH = _lib.LIB_PATH


def library():
    """Load libnaivehip.so (raises ImportError if it was not built)."""
    return _lib.load()
