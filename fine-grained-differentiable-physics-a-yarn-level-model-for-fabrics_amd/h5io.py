"""Trajectories in the reference's HDF5 format.

Layout (writer src/Saver.cpp:184-426, structs src/Saver.h:7-92, reader
src/Loader.cpp:102-267):

  /Properties/Size                   compound {int width, int length}, one element
  /Properties/Physical properties    compound of doubles (Gravity, rho | Yarn1_rho Yarn2_rho,
                                     R, L, Y | Yarn1_Y Yarn2_Y, B | Yarn1_B Yarn2_B, kc, kf, df, mu, S,
                                     handle_stiff, wind_velocity_x/y/z, wind_density,
                                     wind_drag) and int woven_pattern, one element
  /Steps/Position/Node u v/LPOS, LPOSFix  (steps, 3) doubles
  /Steps/Position/Node u v/EPOS, EPOSBar  (steps, 2) doubles
  /Steps/Velocity/Node u v/LVEL           (steps, 3) doubles
  /Steps/Velocity/Node u v/EVEL           (steps, 2) doubles

Frames here are arrays (steps, N, 3|2) indexed u + v * width, the order of the
reference loader's tensors.  The file is read and written through the HDF5 C library
(ctypes); it is located from YS_HDF5_LIB, the conda prefix or the loader path, and its
absence raises.
"""
import ctypes
import ctypes.util
import os

import numpy as np

_hid = ctypes.c_int64
_hsize = ctypes.c_ulonglong
H5P_DEFAULT = 0
H5S_ALL = 0
H5F_ACC_RDONLY = 0x0000
H5F_ACC_TRUNC = 0x0002
H5T_INTEGER = 0
H5T_FLOAT = 1
H5T_COMPOUND = 6

# frame key -> (group, dataset, width)
DATASETS = {"LPos": ("Position", "LPOS", 3), "LPosFix": ("Position", "LPOSFix", 3),
            "EPos": ("Position", "EPOS", 2), "EPosBar": ("Position", "EPOSBar", 2),
            "LVel": ("Velocity", "LVEL", 3), "EVel": ("Velocity", "EVEL", 2)}
HOMO_FIELDS = ("Gravity", "rho", "R", "L", "Y", "B", "kc", "kf", "df", "mu", "S", "handle_stiff",
               "wind_velocity_x", "wind_velocity_y", "wind_velocity_z", "wind_density", "wind_drag")
HETER_FIELDS = ("Gravity", "Yarn1_rho", "Yarn2_rho", "R", "L", "Yarn1_Y", "Yarn2_Y", "Yarn1_B", "Yarn2_B", "kc", "kf",
                "df", "mu", "S", "handle_stiff", "wind_velocity_x", "wind_velocity_y", "wind_velocity_z",
                "wind_density", "wind_drag")   # member names of src/Saver.cpp:466-486
PATTERNS = ("plain", "twill", "satin")   # WovenPattern, src/Cloth.h:16-20

_lib = None


def _find():
    cands = [os.environ.get("YS_HDF5_LIB"), "/opt/conda/lib/libhdf5.so"]
    if os.environ.get("CONDA_PREFIX"):
        cands.append(os.path.join(os.environ["CONDA_PREFIX"], "lib", "libhdf5.so"))
    cands.append(ctypes.util.find_library("hdf5"))
    for c in cands:
        if c and (os.path.exists(c) or not os.path.isabs(c)):
            try:
                return ctypes.CDLL(c)
            except OSError:
                continue
    raise OSError("the HDF5 C library (libhdf5.so) was not found; set YS_HDF5_LIB")


def lib():
    global _lib
    if _lib is not None:
        return _lib
    L = _find()
    if L.H5open() < 0:
        raise OSError("H5open failed")
    for name, res, args in [
            ("H5Fcreate", _hid, [ctypes.c_char_p, ctypes.c_uint, _hid, _hid]),
            ("H5Fopen", _hid, [ctypes.c_char_p, ctypes.c_uint, _hid]),
            ("H5Fclose", ctypes.c_int, [_hid]),
            ("H5Gcreate2", _hid, [_hid, ctypes.c_char_p, _hid, _hid, _hid]),
            ("H5Gclose", ctypes.c_int, [_hid]),
            ("H5Screate_simple", _hid, [ctypes.c_int, ctypes.POINTER(_hsize), ctypes.POINTER(_hsize)]),
            ("H5Sget_simple_extent_dims", ctypes.c_int, [_hid, ctypes.POINTER(_hsize), ctypes.POINTER(_hsize)]),
            ("H5Sget_simple_extent_ndims", ctypes.c_int, [_hid]),
            ("H5Sclose", ctypes.c_int, [_hid]),
            ("H5Tcreate", _hid, [ctypes.c_int, ctypes.c_size_t]),
            ("H5Tinsert", ctypes.c_int, [_hid, ctypes.c_char_p, ctypes.c_size_t, _hid]),
            ("H5Tget_nmembers", ctypes.c_int, [_hid]),
            ("H5Tget_member_name", ctypes.c_void_p, [_hid, ctypes.c_uint]),
            ("H5Tget_member_class", ctypes.c_int, [_hid, ctypes.c_uint]),
            ("H5Tclose", ctypes.c_int, [_hid]),
            ("H5Dcreate2", _hid, [_hid, ctypes.c_char_p, _hid, _hid, _hid, _hid, _hid]),
            ("H5Dopen2", _hid, [_hid, ctypes.c_char_p, _hid]),
            ("H5Dwrite", ctypes.c_int, [_hid, _hid, _hid, _hid, _hid, ctypes.c_void_p]),
            ("H5Dread", ctypes.c_int, [_hid, _hid, _hid, _hid, _hid, ctypes.c_void_p]),
            ("H5Dget_space", _hid, [_hid]),
            ("H5Dget_type", _hid, [_hid]),
            ("H5Dclose", ctypes.c_int, [_hid]),
            ("H5free_memory", ctypes.c_int, [ctypes.c_void_p])]:
        f = getattr(L, name)
        f.restype, f.argtypes = res, args
    L.T_DOUBLE = _hid.in_dll(L, "H5T_NATIVE_DOUBLE_g").value
    L.T_INT = _hid.in_dll(L, "H5T_NATIVE_INT_g").value
    _lib = L
    return L


def _ok(v, what):
    if v < 0:
        raise OSError("HDF5: %s failed" % what)
    return v


def _compound(fields):
    """A C-layout structured dtype (doubles, then an int) and its HDF5 compound type."""
    L = lib()
    dt = np.dtype([(f, np.float64) for f in fields] + [("woven_pattern", np.int32)], align=True)
    t = _ok(L.H5Tcreate(H5T_COMPOUND, dt.itemsize), "H5Tcreate")
    for f in dt.names:
        _ok(L.H5Tinsert(t, f.encode(), dt.fields[f][1], L.T_INT if f == "woven_pattern" else L.T_DOUBLE), "H5Tinsert")
    return dt, t


def _write(loc, name, mtype, arr):
    L = lib()
    arr = np.ascontiguousarray(arr)
    dims = (_hsize * max(arr.ndim, 1))(*(arr.shape if arr.ndim else (1,)))
    sp = _ok(L.H5Screate_simple(max(arr.ndim, 1), dims, None), "H5Screate_simple")
    ds = _ok(L.H5Dcreate2(loc, name.encode(), mtype, sp, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), "H5Dcreate2 " + name)
    _ok(L.H5Dwrite(ds, mtype, H5S_ALL, H5S_ALL, H5P_DEFAULT, arr.ctypes.data), "H5Dwrite " + name)
    L.H5Dclose(ds)
    L.H5Sclose(sp)


def props_from_spec(spec):
    """The Physical properties record of a spec (homogeneous or two-yarn field set)."""
    y = spec["yarns"]
    p = {"Gravity": spec["G"][2], "R": spec["R"], "L": spec["L"], "kc": spec["kc"], "kf": spec["kf"],
         "df": spec["df"], "mu": spec["mu"], "S": spec["S"], "handle_stiff": spec["handle_k"],
         "wind_velocity_x": spec["wind_vel"][0], "wind_velocity_y": spec["wind_vel"][1],
         "wind_velocity_z": spec["wind_vel"][2], "wind_density": spec["wind_density"],
         "wind_drag": spec["wind_drag"], "woven_pattern": PATTERNS.index(spec["pattern"])}
    if len(y) == 1:
        p.update({"rho": y[0]["rho"], "Y": y[0]["Y"], "B": y[0]["B"]})
    else:
        p.update({"Yarn1_rho": y[0]["rho"], "Yarn2_rho": y[1]["rho"], "Yarn1_Y": y[0]["Y"], "Yarn2_Y": y[1]["Y"],
                  "Yarn1_B": y[0]["B"], "Yarn2_B": y[1]["B"]})
    return p


def save_trajectory(path, spec, frames):
    """Writes `frames` (get_ClothState keys -> (steps, N, 3|2)) of the cloth `spec` in
    the reference's layout (src/Saver.cpp:184-426)."""
    L = lib()
    W, Lg = int(spec["width"]), int(spec["length"])
    steps = int(frames["LPos"].shape[0])
    f = _ok(L.H5Fcreate(str(path).encode(), H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT), "H5Fcreate")
    try:
        g = _ok(L.H5Gcreate2(f, b"/Properties", H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), "H5Gcreate2")
        sdt = np.dtype([("width", np.int32), ("length", np.int32)], align=True)
        st = _ok(L.H5Tcreate(H5T_COMPOUND, sdt.itemsize), "H5Tcreate")
        for n in sdt.names:
            _ok(L.H5Tinsert(st, n.encode(), sdt.fields[n][1], L.T_INT), "H5Tinsert")
        _write(g, "Size", st, np.array([(W, Lg)], dtype=sdt))
        L.H5Tclose(st)
        p = props_from_spec(spec)
        fields = HOMO_FIELDS if "rho" in p else HETER_FIELDS
        pdt, pt = _compound(fields)
        rec = np.zeros(1, dtype=pdt)
        for n in pdt.names:
            rec[n] = p[n]
        _write(g, "Physical properties", pt, rec)
        L.H5Tclose(pt)
        L.H5Gclose(g)
        root = _ok(L.H5Gcreate2(f, b"/Steps", H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), "H5Gcreate2")
        grp = {k: _ok(L.H5Gcreate2(root, k.encode(), H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), "H5Gcreate2")
               for k in ("Position", "Velocity")}
        for v in range(Lg):
            for u in range(W):
                k = u + v * W
                node = {gk: _ok(L.H5Gcreate2(gid, ("Node %d %d" % (u, v)).encode(), H5P_DEFAULT, H5P_DEFAULT,
                                             H5P_DEFAULT), "H5Gcreate2") for gk, gid in grp.items()}
                for key, (gk, dname, dim) in DATASETS.items():
                    a = np.ascontiguousarray(np.asarray(frames[key], dtype=np.float64)[:, k, :dim].reshape(steps, dim))
                    _write(node[gk], dname, L.T_DOUBLE, a)
                for gid in node.values():
                    L.H5Gclose(gid)
        for gid in grp.values():
            L.H5Gclose(gid)
        L.H5Gclose(root)
    finally:
        L.H5Fclose(f)


def _read_ds(f, name, mtype, dtype):
    L = lib()
    ds = _ok(L.H5Dopen2(f, name.encode(), H5P_DEFAULT), "H5Dopen2 " + name)
    try:
        sp = L.H5Dget_space(ds)
        nd = L.H5Sget_simple_extent_ndims(sp)
        dims = (_hsize * nd)()
        L.H5Sget_simple_extent_dims(sp, dims, None)
        L.H5Sclose(sp)
        out = np.zeros(tuple(dims[i] for i in range(nd)), dtype=dtype)
        _ok(L.H5Dread(ds, mtype, H5S_ALL, H5S_ALL, H5P_DEFAULT, out.ctypes.data), "H5Dread " + name)
        return out
    finally:
        L.H5Dclose(ds)


def _read_compound(f, name):
    """Every member of a one-element compound dataset, by name (each read through a
    one-member memory type, so the file's member order and padding do not matter)."""
    L = lib()
    ds = _ok(L.H5Dopen2(f, name.encode(), H5P_DEFAULT), "H5Dopen2 " + name)
    try:
        ft = L.H5Dget_type(ds)
        members = []
        for i in range(L.H5Tget_nmembers(ft)):
            ptr = L.H5Tget_member_name(ft, i)
            members.append((ctypes.string_at(ptr).decode(), L.H5Tget_member_class(ft, i)))
            L.H5free_memory(ptr)
        L.H5Tclose(ft)
        out = {}
        for m, cls in members:
            is_int = cls == H5T_INTEGER
            mt = _ok(L.H5Tcreate(H5T_COMPOUND, 4 if is_int else 8), "H5Tcreate")
            _ok(L.H5Tinsert(mt, m.encode(), 0, L.T_INT if is_int else L.T_DOUBLE), "H5Tinsert")
            buf = np.zeros(1, dtype=np.int32 if is_int else np.float64)
            _ok(L.H5Dread(ds, mt, H5S_ALL, H5S_ALL, H5P_DEFAULT, buf.ctypes.data), "H5Dread " + m)
            L.H5Tclose(mt)
            out[m] = int(buf[0]) if is_int else float(buf[0])
        return out
    finally:
        L.H5Dclose(ds)


def load_trajectory(path, start=0, end=None):
    """Reads rows [start, end) of a trajectory file (src/Loader.cpp:102-267).  Returns
    (size (width, length), properties dict, frames dict (steps, N, 3|2))."""
    L = lib()
    f = _ok(L.H5Fopen(str(path).encode(), H5F_ACC_RDONLY, H5P_DEFAULT), "H5Fopen " + str(path))
    try:
        size = _read_compound(f, "/Properties/Size")
        W, Lg = size["width"], size["length"]
        props = _read_compound(f, "/Properties/Physical properties")
        frames = {}
        for key, (gk, dname, dim) in DATASETS.items():
            arr = None
            for v in range(Lg):
                for u in range(W):
                    d = _read_ds(f, "/Steps/%s/Node %d %d/%s" % (gk, u, v, dname), L.T_DOUBLE, np.float64)
                    d = d.reshape(d.shape[0], dim)[start:end]
                    if arr is None:
                        arr = np.zeros((d.shape[0], W * Lg, dim))
                    arr[:, u + v * W, :] = d
            frames[key] = arr
        return (W, Lg), props, frames
    finally:
        L.H5Fclose(f)


def spec_from_props(size, props, handles=None, h=1e-3, hight=5.0, pose="upright"):
    """A cloth spec from a file's size and Physical properties.  The file does not hold
    the handles, the time step, the hanging height or the pose; they are arguments
    (default: the two top corners, 1 ms, 5 m, upright, as the reference's drivers)."""
    W, Lg = size
    if "rho" in props:
        yarns = [{"rho": props["rho"], "Y": props["Y"], "B": props["B"]}]
    else:
        yarns = [{"rho": props["Yarn1_rho"], "Y": props["Yarn1_Y"], "B": props["Yarn1_B"]},
                 {"rho": props["Yarn2_rho"], "Y": props["Yarn2_Y"], "B": props["Yarn2_B"]}]
    return {"width": W, "length": Lg, "hight": hight, "L": props["L"], "R": props["R"], "yarns": yarns,
            "kc": props["kc"], "kf": props["kf"], "df": props["df"], "mu": props["mu"], "S": props["S"],
            "handle_k": props["handle_stiff"], "pattern": PATTERNS[props["woven_pattern"]], "pose": pose,
            "G": [0.0, 0.0, props["Gravity"]],
            "wind_vel": [props["wind_velocity_x"], props["wind_velocity_y"], props["wind_velocity_z"]],
            "wind_density": props["wind_density"], "wind_drag": props["wind_drag"],
            "handles": list(handles) if handles is not None else [(0, 0), (W - 1, 0)], "h": h}
