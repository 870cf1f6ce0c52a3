"""The .obj frame writer (src/Render.cpp:3-131).  CPU only: closed forms on the flat rest
pose of a non-square cloth."""
import numpy as np
import pytest


def _read(path):
    v, vt, vn, f, l = [], [], [], [], []
    for line in open(path):
        t = line.split()
        if not t:
            continue
        if t[0] == "v":
            v.append([float(a) for a in t[1:4]])
        elif t[0] == "vt":
            vt.append([float(a) for a in t[1:3]])
        elif t[0] == "vn":
            vn.append([float(a) for a in t[1:4]])
        elif t[0] == "f":
            f.append([int(a.split("/")[0]) for a in t[1:]])
        elif t[0] == "l":
            l.append([int(a) for a in t[1:3]])
    return np.array(v), np.array(vt), np.array(vn), f, l


@pytest.fixture
def cloth(ys):
    cs = ys.cloth_spec
    spec = cs.homo_spec(6)
    spec["length"] = 4
    return spec, cs.initial_state(spec)["LPos"]


def test_triangle_form(ys, cloth, tmp_path):
    spec, x = cloth
    W, Lg = 6, 4
    p = tmp_path / "t.obj"
    ys.render.render_cloth(p, spec, x, "triangle")
    v, vt, vn, f, _ = _read(p)
    assert v.shape == (W * Lg, 3) and np.allclose(v, x, atol=5e-6)
    assert np.allclose(vt[W + 2], [2 / 5, 1 / 5], atol=1e-6)
    # the rest pose lies in the xz-plane with v growing downwards: every face normal is +y
    assert np.allclose(vn, [0.0, 1.0, 0.0])
    assert len(f) == 2 * (W - 1) * (Lg - 1)
    # alternating diagonals: cell (0,0) splits along tl-br, cell (1,0) along tr-bl
    assert f[0] == [1, 2, W + 2] and f[1] == [1, W + 2, W + 1]
    assert f[2] == [2, 3, W + 2] and f[3] == [3, W + 3, W + 2]


def test_rectangle_form(ys, cloth, tmp_path):
    spec, x = cloth
    p = tmp_path / "r.obj"
    ys.render.render_cloth(p, spec, x, "rectangle")
    v, _, _, f, _ = _read(p)
    assert len(f) == 5 * 3 and f[0] == [1, 2, 8, 7]


def test_yarn_form(ys, cloth, tmp_path):
    spec, x = cloth
    W, Lg = 6, 4
    nrm = np.zeros((W * Lg, 3))
    for v in range(1, Lg - 1):
        for u in range(1, W - 1):
            nrm[u + v * W] = (0.0, 1.0, 0.0)
    p = tmp_path / "y.obj"
    ys.render.render_cloth(p, spec, x, "yarn", normals=nrm)
    v, _, _, _, l = _read(p)
    assert len(v) == (Lg - 2) * W + (W - 2) * Lg
    assert len(l) == (Lg - 2) * (W - 1) + (W - 2) * (Lg - 1)
    R = spec["R"]
    # warp copy of interior node (2, 1): x - R n; of boundary node (0, 1): x
    assert np.allclose(v[(1 - 1) * W + 2], x[2 + W] - R * np.array([0, 1, 0]), atol=5e-6)
    assert np.allclose(v[0], x[W], atol=5e-6)
    # weft copy of interior node (1, 2): x + R n
    assert np.allclose(v[(Lg - 2) * W + 2], x[1 + 2 * W] + R * np.array([0, 1, 0]), atol=5e-6)
    with pytest.raises(ValueError):
        ys.render.render_cloth(p, spec, x, "yarn")
