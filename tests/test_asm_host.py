"""The first assembly kernel's work (assemble1_block, csrc/assemble1.hpp) built for the
host by tests/asm_host.cpp under AddressSanitizer and UndefinedBehaviorSanitizer (g++), every
thread of a block an OS thread and the block's LDS an exact-size array: no sanitizer
report on the parity cases (one to four tiles, partial tiles, handles, wind), and the
rows it writes -- every term but the shear, which the second kernel adds -- equal the
oracle's with the shear switched off (A, F, b within 1e-12 relative; the compression
within 1e-12).  This is the check for undefined behaviour (an index out of range, a read
of an uninitialised slot) in the code the GPU kernels call."""
import os
import subprocess

import numpy as np
import pytest
import scipy.sparse as sp

from oracle.oracle import OracleCloth
from tests.golden_cases import cases, excited_state

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "fine-grained-differentiable-physics-a-yarn-level-model-for-fabrics_amd")
SAN_ENV = dict(os.environ, ASAN_OPTIONS="detect_leaks=0:abort_on_error=1",
               UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1")


@pytest.fixture(scope="module")
def harness(tmp_path_factory):
    exe = str(tmp_path_factory.mktemp("asm_host") / "asm_host")
    subprocess.run(["g++", "-std=c++20", "-O1", "-g", "-fsanitize=address,undefined", "-fno-sanitize-recover=undefined",
                    "-fno-omit-frame-pointer", "-I", os.path.join(ROOT, "tests", "host_shim"),
                    "-I", os.path.join(ROOT, "include"), "-I", os.path.join(PKG, "csrc"),
                    os.path.join(ROOT, "tests", "asm_host.cpp"), "-o", exe, "-lpthread"], check=True)
    return exe


def _planes(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.float64).T).ravel()


def _run(harness, tmp_path, ys, spec, state, normals, h):
    W, Lg = spec["width"], spec["length"]
    N = W * Lg
    nrm = np.array(normals, dtype=np.float64)
    inter = np.zeros(N, bool)
    for v in range(1, Lg - 1):
        inter[v * W + 1:v * W + W - 1] = True
    nrm[~inter] = 0.0   # the device keeps boundary normals at zero
    inp, out = tmp_path / "in.bin", tmp_path / "out.bin"
    with open(inp, "wb") as f:
        f.write(bytes(ys._abi.make_desc(spec)))
        f.write(np.array([h], dtype=np.float64).tobytes())
        for k in ("LPos", "LVel", "LPosFix", "EPos", "EVel", "EPosBar"):
            f.write(_planes(state[k]).tobytes())
        f.write(_planes(nrm).tobytes())
    r = subprocess.run([harness, str(inp), str(out)], env=SAN_ENV, capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stderr[-4000:]
    assert "runtime error" not in r.stderr and "AddressSanitizer" not in r.stderr, r.stderr[-4000:]
    raw = open(out, "rb").read()
    nnz = int(np.frombuffer(raw, np.int64, 1)[0])
    o = 8
    rows = np.frombuffer(raw, np.int32, nnz, o); o += 4 * nnz
    cols = np.frombuffer(raw, np.int32, nnz, o); o += 4 * nnz
    vals = np.frombuffer(raw, np.float64, nnz, o); o += 8 * nnz
    rest = np.frombuffer(raw, np.float64, offset=o)
    Wp, Hp = W + 4, Lg + 4
    P = ((Wp * Hp + 63) // 64) * 64
    sizes = [5 * N, 5 * P, N, 2 * N, 2 * N, 2]
    parts, i = [], 0
    for s in sizes:
        parts.append(rest[i:i + s])
        i += s
    F5, b5, comp, _, _, flags = parts
    n = 3 * N + 2 * (W - 2) * (Lg - 2)
    A = sp.coo_matrix((vals, (rows, cols)), shape=(n, n)).tocsr()
    # F and b in the global layout
    F = np.zeros(n)
    b = np.zeros(n)
    for v in range(Lg):
        for u in range(W):
            k = u + v * W
            pk = (u + 2) + (v + 2) * Wp
            for c in range(3):
                F[3 * k + c] = F5[c * N + k]
                b[3 * k + c] = b5[c * P + pk]
            if inter[k]:
                e = (u - 1) + (v - 1) * (W - 2)
                for c in range(2):
                    F[3 * N + 2 * e + c] = F5[(3 + c) * N + k]
                    b[3 * N + 2 * e + c] = b5[(3 + c) * P + pk]
    return A, F, b, comp, flags


def _rel(a, b):
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (nb if nb > 0 else 1.0)


def _frel(A, B):
    D = (A - B).tocsr()
    nb = np.sqrt((B.tocsr().data ** 2).sum())
    return np.sqrt((D.data ** 2).sum()) / nb


def _all_cases(cs):
    out = [(name, spec, 400 + i) for i, (name, spec) in enumerate(cases(cs))]
    s = cs.heter_spec(13, "twill", 5)   # 2 x 2 tiles, partial
    s["wind_vel"] = [0.0, 3.0, 0.0]
    out.append(("heter13_twill_wind", s, 450))
    s = cs.heter_spec(19, "satin", 6)   # 3 x 3 tiles, partial
    s["length"] = 10
    out.append(("heter19x10_satin", s, 451))
    return out


@pytest.mark.parametrize("idx", range(6))
def test_assembly_body_on_host_under_sanitizers(ys, cs, harness, tmp_path, idx):
    name, spec, seed = _all_cases(cs)[idx]
    state = excited_state(cs, spec, seed)
    h = spec["h"]
    r = OracleCloth(spec)
    r.set_state(state)
    normals = r.normals()
    r.set_terms([t for t in OracleCloth.TERMS if t != "shear"])
    r.assemble(h)
    A, F, b, comp, flags = _run(harness, tmp_path, ys, spec, state, normals, h)
    assert flags[0] == 0 and flags[1] == 0, (name, flags)
    eA, eF, eb = _frel(A, r.csr(0)), _rel(F, r.vec("F")), _rel(b, r.vec("b"))
    _, cp, _ = r.crossing()
    W, Lg = spec["width"], spec["length"]
    inter = [u + v * W for v in range(1, Lg - 1) for u in range(1, W - 1)]
    ec = _rel(comp[inter], cp)
    print("%s: |dA| %.2e |dF| %.2e |db| %.2e |dcomp| %.2e" % (name, eA, eF, eb, ec))
    assert eA < 1e-12 and eF < 1e-12 and eb < 1e-12 and ec < 1e-12
