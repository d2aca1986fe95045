"""Writes golden box-mesh fixtures by running the reference's own generator.

python/sfem/mesh/box_mesh.py:479-700 (`create`) fixes node numbering (np.meshgrid
'xy' flattening) and the 6-tet split; no closed form pins that numbering, so the
fixtures are its actual output on small boxes.  Run in a container that holds the
reference checkout; the fixtures (.npz, data only) are committed.

    python tests/golden/make_box_mesh_golden.py /root/reference
"""
import os
import sys

import numpy as np

ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
sys.path.insert(0, os.path.join(ref, "python", "sfem", "mesh"))
import box_mesh  # noqa: E402  (reference module, read from the checkout)

out = os.path.dirname(os.path.abspath(__file__))
for cell in ("hex8", "tet4"):
    for (nx, ny, nz, w, h, t) in ((3, 4, 5, 1.0, 1.0, 1.0), (4, 4, 4, 1.0, 2.0, 0.5)):
        idx, points = box_mesh.create(w, h, t, nx, ny, nz, cell)
        name = f"box_{cell}_{nx}x{ny}x{nz}_{w:g}_{h:g}_{t:g}.npz"
        np.savez(os.path.join(out, name),
                 elements=np.stack([np.asarray(i, dtype=np.int32) for i in idx]),
                 points=np.stack([np.asarray(p, dtype=np.float32) for p in points]),
                 dims=np.array([nx, ny, nz]), extent=np.array([w, h, t]))
        print("wrote", name)
