"""Writes tests/golden/wind_yarn12_plain5.npz from the reference's only shipped
trajectory, TrainingData/Table_2/Wind_Yarn_1_2_Plain_Size5.h5 (5x5 crossings, two
yarns, plain weave, wind (0, 5, 0)), using the HDF5 command-line dumper.

Run in the build container (it reads /root/reference); the committed .npz is data:
the physical properties and the first ROWS rows of every node's LPOS, LPOSFix, EPOS,
EPOSBar, LVEL and EVEL datasets (layout of src/Saver.cpp:184-426).
"""
import os
import subprocess
import tempfile

import numpy as np

SRC = "/root/reference/TrainingData/Table_2/Wind_Yarn_1_2_Plain_Size5.h5"
H5DUMP = "/opt/conda/bin/h5dump"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "wind_yarn12_plain5.npz")
ROWS = 121
PROPS = ["Gravity", "Yarn1_rho", "Yarn2_rho", "R", "L", "Yarn1_Y", "Yarn2_Y", "Yarn1_B", "Yarn2_B", "kc", "kf",
         "df", "mu", "S", "handle_stiff", "wind_velocity_x", "wind_velocity_y", "wind_velocity_z", "wind_density",
         "wind_drag", "woven_pattern"]


def dump(ds):
    with tempfile.NamedTemporaryFile("r", suffix=".txt") as f:
        subprocess.check_call([H5DUMP, "-y", "-w", "0", "-m", "%.17g", "-o", f.name, "-d", ds, SRC],
                              stdout=subprocess.DEVNULL)
        txt = f.read()
    return np.array([float(t) for t in txt.replace("{", " ").replace("}", " ").replace(",", " ").split()])


def main():
    props = dump("/Properties/Physical properties")
    out = {"props": props, "prop_names": np.array(PROPS)}
    W = 5
    for name, grp, dim in [("LPOS", "Position", 3), ("LPOSFix", "Position", 3), ("EPOS", "Position", 2),
                           ("EPOSBar", "Position", 2), ("LVEL", "Velocity", 3), ("EVEL", "Velocity", 2)]:
        arr = np.zeros((ROWS, W * W, dim))
        for v in range(W):
            for u in range(W):
                d = dump("/Steps/%s/Node %d %d/%s" % (grp, u, v, name)).reshape(-1, dim)
                arr[:, u + v * W, :] = d[:ROWS]
        out[name] = arr
    np.savez_compressed(OUT, **out)
    print("wrote", OUT)


if __name__ == "__main__":
    main()
