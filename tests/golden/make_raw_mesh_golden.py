"""Writes a golden raw-mesh folder by running the reference's own box-mesh writer.

The SFEM raw mesh format (meta.yaml with the element type, point and element file lists; one
float32 file per coordinate, one int32 file per element node; boundary node sets) is defined
by the script body of python/sfem/mesh/box_mesh.py:697-900; no closed form pins the file names,
key names or dtypes, so the fixture is that script's actual output on a 2x2x2 hex8 box.  Run in
a container that holds the reference checkout; the folder (data only) is committed.

    python tests/golden/make_raw_mesh_golden.py /root/reference
"""
import os
import shutil
import subprocess
import sys

ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
script = os.path.join(ref, "python", "sfem", "mesh", "box_mesh.py")
out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "raw_hex8_2x2x2")
if os.path.exists(out):
    shutil.rmtree(out)
subprocess.run([sys.executable, script, out, "-c", "hex8", "-x", "2", "-y", "2", "-z", "2", "--width=1",
                "--height=2", "--depth=3"], check=True)
print("wrote", out)
