"""Wavefront .obj frames of a cloth (src/Render.cpp:3-131).

Three forms, as the reference's RenderType:
  "triangle"   vertices, texture coordinates, vertex normals (mean of the unit normals of
               the faces around the vertex) and the two triangles of every cell with the
               alternating diagonal of the cloth's face list (src/Cloth.cpp:272-286);
  "rectangle"  vertices and one quad per cell;
  "yarn"       the warp and weft copies of the nodes (x - R n on warp yarns, x + R n on weft
               yarns at interior crossings, src/Cloth.cpp:321-325) as polylines.
Coordinates are written with 5 decimals and texture coordinates with 6, as the
reference's streams do.  The reference keeps its texture step in a function-level
static (fixed by the first cloth it renders); here it is computed per cloth.
"""
import numpy as np

FORMS = ("triangle", "rectangle", "yarn")


def _cell_faces(W, Lg):
    """The cloth's triangles (vertex indices, 0-based) in face-list order."""
    out = []
    for i in range(Lg - 1):
        for j in range(W - 1):
            tl, tr, bl, br = j + i * W, j + 1 + i * W, j + (i + 1) * W, j + 1 + (i + 1) * W
            if (i + j) % 2 == 0:
                out += [(tl, tr, br), (tl, br, bl)]
            else:
                out += [(tl, tr, bl), (tr, br, bl)]
    return out


def vertex_normals(x, W, Lg):
    """Mean of the unit normals of the faces around each vertex (Face::update,
    src/ClothUnit.h:268-273: cross(x2 - x1, x3 - x1), unit unless zero)."""
    acc = np.zeros((W * Lg, 3))
    cnt = np.zeros(W * Lg)
    for a, b, c in _cell_faces(W, Lg):
        n = np.cross(x[b] - x[a], x[c] - x[a])
        nn = np.linalg.norm(n)
        if nn != 0.0:
            n = n / nn
        for k in (a, b, c):
            acc[k] += n
            cnt[k] += 1
    return acc / np.maximum(cnt, 1)[:, None]


def render_cloth(path, spec, lpos, form="triangle", normals=None):
    """Writes one frame.  lpos: (N, 3) node positions, index u + v*width; normals: (N, 3)
    crossing normals (``Cloth.normals()``), needed by the "yarn" form."""
    if form not in FORMS:
        raise ValueError("form must be one of %s" % (FORMS,))
    W, Lg = int(spec["width"]), int(spec["length"])
    x = np.asarray(lpos, dtype=np.float64).reshape(W * Lg, 3)
    lines = ["o cloth"]
    fx = lambda p: " ".join("%.5f" % c for c in p) + " "   # noqa: E731
    if form == "triangle":
        unit = 1.0 / (max(W, Lg) - 1)
        lines += ["v " + fx(x[j + i * W]) for i in range(Lg) for j in range(W)]
        lines += ["vt %f %f" % (j * unit, i * unit) for i in range(Lg) for j in range(W)]
        vn = vertex_normals(x, W, Lg)
        lines += ["vn " + fx(vn[j + i * W]) for i in range(Lg) for j in range(W)]
        for a, b, c in _cell_faces(W, Lg):
            lines.append("f " + " ".join("%d/%d/%d" % (k + 1, k + 1, k + 1) for k in (a, b, c)))
    elif form == "rectangle":
        lines += ["v " + fx(x[j + i * W]) for i in range(Lg) for j in range(W)]
        for i in range(Lg - 1):
            for j in range(W - 1):
                tl, tr, bl, br = j + i * W + 1, j + 2 + i * W, j + (i + 1) * W + 1, j + 2 + (i + 1) * W
                lines.append("f %d %d %d %d" % (tl, tr, br, bl))
    else:
        if normals is None:
            raise ValueError("the yarn form needs the crossing normals")
        nrm = np.asarray(normals, dtype=np.float64).reshape(W * Lg, 3)
        R = spec["R"]

        def copy(u, v, sign):
            k = u + v * W
            inter = 0 < u < W - 1 and 0 < v < Lg - 1
            return x[k] + sign * R * nrm[k] if inter else x[k]
        lines += ["v " + fx(copy(u, v, -1.0)) for v in range(1, Lg - 1) for u in range(W)]   # warp copies
        lines += ["v " + fx(copy(u, v, 1.0)) for u in range(1, W - 1) for v in range(Lg)]    # weft copies
        lines += ["l %d %d " % ((v - 1) * W + u + 1, (v - 1) * W + u + 2)
                  for v in range(1, Lg - 1) for u in range(W - 1)]
        lines += ["l %d %d " % ((Lg - 2) * W + (u - 1) * Lg + v + 1, (Lg - 2) * W + (u - 1) * Lg + v + 2)
                  for u in range(1, W - 1) for v in range(Lg - 1)]
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")
