"""CPU-side checks of the product boundary: the HIP library loads and exports every
entry point include/yarnstep.h declares (no device call is made), the host mirror
builds valid descriptors, and the workloads are well formed."""
import ctypes
import os
import re

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    txt = open(os.path.join(ROOT, "include", "yarnstep.h")).read()
    return sorted(set(re.findall(r"\b(yc_[a-z_]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol(ys):
    lib = ctypes.CDLL(ys.lib_path())
    names = declared_symbols()
    assert len(names) >= 18
    for n in names:
        assert hasattr(lib, n), n
    assert sorted(ys._abi.EXPORTS) == names


def test_descriptor_layout(ys, cs):
    spec = cs.heter_spec(9, "satin", 3)
    d = ys._abi.make_desc(spec)
    assert d.width == 9 and d.n_yarns == 2 and d.pattern == 2 and d.n_handles == 2
    assert d.handles[2] == 8 and d.handles[3] == 0
    # the C struct and the ctypes mirror agree on size and field offsets
    import shutil
    import subprocess
    import tempfile
    if shutil.which("gcc") is None:
        return
    src = (
        '#include <stdio.h>\n#include <stddef.h>\n#include "yarnstep.h"\n'
        'int main(void){printf("%zu %zu %zu %zu", sizeof(yc_desc), offsetof(yc_desc, kc),'
        ' offsetof(yc_desc, G), offsetof(yc_desc, handles)); return 0;}\n')
    with tempfile.TemporaryDirectory() as td:
        with open(os.path.join(td, "t.c"), "w") as f:
            f.write(src)
        subprocess.check_call(["gcc", "-I", os.path.join(ROOT, "include"), "-o", os.path.join(td, "t"), os.path.join(td, "t.c")])
        out = subprocess.check_output([os.path.join(td, "t")]).decode().split()
    D = ys._abi.YcDesc
    assert [int(x) for x in out] == [ctypes.sizeof(D), D.kc.offset, D.G.offset, D.handles.offset]


def test_workloads_are_well_formed(cs):
    for name in ("homo5", "heter64_plain", "heter64_twill", "heter64_satin"):
        spec, st, ext, steps = cs.workload(name)
        N = spec["width"] * spec["length"]
        assert st["LPos"].shape == (N, 3) and st["EPos"].shape == (N, 2)
        assert ext is None and steps in (100, 200)
    f = cs.pressure_force(cs.homo_spec(4), cs.initial_state(cs.homo_spec(4)), 2.0)
    # total pressure force = p * area of the 3x3-cell sheet, all along +z
    assert abs(f[2::3][:16].sum() - 2.0 * (3 * 1e-3) ** 2) < 1e-15
    assert np.all(f[0:48:3] == 0)


def test_compute_raises_without_device(ys, cs):
    """There is no CPU fallback: creating a context on a machine without a HIP device
    fails loudly."""
    import pytest
    import torch
    if torch.cuda.is_available():
        pytest.skip("a device is present")
    with pytest.raises(ys.YcError):
        ys.Cloth(cs.homo_spec(5))


def test_cpp_adapter_driver_builds(tmp_path):
    """include/yarnstep.hpp (the C++ mirror of Cloth / PhysicsCloth / EigenSolver /
    SimStep over the C ABI) compiles and links into tests/abi_driver.cpp, a driver
    written as the reference's inline step (src/Simulate.h:86-96); the GPU tests run it."""
    import __graft_entry__
    exe = __graft_entry__.build_abi_driver(out=str(tmp_path / "abi_driver"))
    assert os.path.exists(exe)
