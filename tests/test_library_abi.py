"""CPU: the C-ABI library loads and exports every entry point declared in
include/*.h; host-only logic (specification checks) behaves like the
reference's error paths.  No device work is issued here."""
import ctypes
import glob
import os
import re
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names |= set(re.findall(r"\b(nd_[a-z0-9_]+)\s*\(", text))
    return sorted(names)


def test_library_exports_every_declared_symbol(pkg):
    lib = pkg.library()
    missing = [n for n in declared_symbols() if not hasattr(lib, n)]
    assert not missing, missing
    assert len(declared_symbols()) >= 30


def test_bindings_cover_every_declared_symbol(pkg):
    assert set(declared_symbols()) == set(pkg._lib.SIGNATURES)


def test_struct_layouts(pkg, tmp_path):
    """The ctypes mirrors agree with the C header field by field: a probe
    compiled against include/naivedynamics.h prints every offset and size."""
    assert pkg.NODE_DTYPE.itemsize == 32 and pkg.PAIR_DTYPE.itemsize == 12
    structs = {"nd_sim_params": pkg._lib.SimParams, "nd_adj_params": pkg._lib.AdjParams,
               "nd_bvh_spec": pkg._lib.BvhSpec, "nd_sim_stats": pkg._lib.SimStats}
    lines = ['#include <stddef.h>', '#include <stdio.h>', '#include "naivedynamics.h"', "int main(void) {"]
    for cname, cls in structs.items():
        lines.append('printf("%s size %%zu\\n", sizeof(%s));' % (cname, cname))
        for f, _ in cls._fields_:
            lines.append('printf("%s.%s %%zu\\n", offsetof(%s, %s));' % (cname, f, cname, f))
    lines.append("return 0; }")
    src = tmp_path / "probe.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "probe"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    out = dict(line.rsplit(" ", 1) for line in subprocess.run([str(exe)], check=True, capture_output=True,
                                                              text=True).stdout.splitlines())
    for cname, cls in structs.items():
        assert int(out[cname + " size"]) == ctypes.sizeof(cls), cname
        for f, _ in cls._fields_:
            assert int(out[cname + "." + f]) == getattr(cls, f).offset, (cname, f)


def test_version(pkg):
    assert pkg.library().nd_version() == 1


@pytest.mark.parametrize("n,apl,ok", [(8, 1, True), (8, 4, True), (8, 3, False), (4, 4, False),
                                      (5000, 1000, True), (262144, 4, True), (0, 1, False)])
def test_spec_validation_matches_reference(pkg, n, apl, ok):
    # SpheresBVHSpecs (BVHTraverse.jl:74-80): apl must divide N and nL >= 2
    if ok:
        s = pkg.SpheresBVHSpecs(neighbor_distance=0.1, atom_count=n, atomsperleaf=apl)
        assert s.leaves_count == n // apl and s.branches_count == n // apl - 1
    else:
        with pytest.raises(pkg.SpecError):
            pkg.SpheresBVHSpecs(neighbor_distance=0.1, atom_count=n, atomsperleaf=apl)


def test_c_spec_init_rejects_bad_specs(pkg):
    lib = pkg.library()
    s = pkg._lib.BvhSpec()
    assert lib.nd_bvh_spec_init(ctypes.byref(s), np.float32(0.5), 12, 5) == pkg._lib.ND_ERR_ARG
    assert lib.nd_bvh_spec_init(ctypes.byref(s), np.float32(0.5), 12, 12) == pkg._lib.ND_ERR_ARG
    assert lib.nd_bvh_spec_init(ctypes.byref(s), np.float32(0.5), 12, 4) == 0
    assert (s.atom_count, s.leaves_count, s.branches_count, s.atomsperleaf) == (12, 3, 2, 4)


def test_package_fails_loudly_without_library(pkg, tmp_path):
    with pytest.raises(ImportError):
        pkg._lib.load.__wrapped__(str(tmp_path / "missing.so")) if hasattr(pkg._lib.load, "__wrapped__") \
            else _load_fresh(pkg, str(tmp_path / "missing.so"))


def _load_fresh(pkg, path):
    saved = pkg._lib._lib
    pkg._lib._lib = None
    try:
        pkg._lib.load(path)
    finally:
        pkg._lib._lib = saved
