"""Convert the reference's FEniCSx gold files into compact .npz fixtures.

Run in a container that holds the reference checkout (it is read as text and
NumPy data only; nothing from it is imported or executed):

    python tests/golden/make_fixtures.py /root/reference

Inputs (reference paths):
  tests/benchmarks/<case>/fenicsx/sol_p0_000000.vtu   ASCII VTU, cell type 72 (HEX8)
  tests/benchmarks/hyperelasticity/fenicsx/traction.npy
  tests/benchmarks/linear_elasticity_cylinder/fenicsx/surface_area.npy
  tests/benchmarks/plasticity/fenicsx/{disps,avg_stresses}.npy

Output: tests/golden/fenicsx_<case>.npz with arrays points (n,3) f64,
cells (m,8) i32, sol (n,vec) f64 and the scalar gold values where present.
The data is GPL-3.0 (same licence as the reference) and only the numbers are kept.
"""
import os
import re
import sys

import numpy as np

CASES = {
    "linear_poisson": 1,
    "linear_elasticity_cube": 3,
    "linear_elasticity_cylinder": 3,
    "hyperelasticity": 3,
    "plasticity": 3,
}


def _data_array(text, name=None, section=None):
    if section is not None:
        m = re.search(r"<%s[^>]*>(.*?)</%s>" % (section, section), text, re.S)
        text = m.group(1)
    if name is None:
        m = re.search(r"<DataArray[^>]*>(.*?)</DataArray>", text, re.S)
    else:
        m = re.search(r'<DataArray[^>]*Name="%s"[^>]*>(.*?)</DataArray>' % name, text, re.S)
    return np.array(m.group(1).split(), dtype=np.float64)


def read_vtu(path, vec):
    text = open(path).read()
    points = _data_array(text, section="Points").reshape(-1, 3)
    conn = _data_array(text, "connectivity").astype(np.int64)
    offsets = _data_array(text, "offsets").astype(np.int64)
    types = _data_array(text, "types").astype(np.int64)
    assert np.all(types == 72) and np.all(np.diff(offsets) == 8)
    cells = conn.reshape(-1, 8).astype(np.int32)
    sol = _data_array(text, "sol", section="PointData").reshape(len(points), vec)
    return points, cells, sol


def main(ref_root):
    out_dir = os.path.dirname(os.path.abspath(__file__))
    bench = os.path.join(ref_root, "tests", "benchmarks")
    for case, vec in CASES.items():
        d = os.path.join(bench, case, "fenicsx")
        points, cells, sol = read_vtu(os.path.join(d, "sol_p0_000000.vtu"), vec)
        extra = {}
        for fn in ("traction.npy", "surface_area.npy", "disps.npy", "avg_stresses.npy"):
            p = os.path.join(d, fn)
            if os.path.exists(p):
                extra[fn[:-4]] = np.load(p, allow_pickle=False)
        np.savez_compressed(os.path.join(out_dir, "fenicsx_%s.npz" % case),
                            points=points, cells=cells, sol=sol, **extra)
        print(case, points.shape, cells.shape, sol.shape, sorted(extra))


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
