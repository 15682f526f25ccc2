"""Copies the glTF assets the loader tests read from the reference's
examples (data files: Blender exports with vertex colours stored as 16-bit
and float, a skin, meshes of several primitives).  Run from the repository
root where /root/reference exists; the copies are committed."""
import os
import shutil

SRC = "/root/reference/examples"
FILES = ["paths/paths.glb", "properties/properties.glb", "skeletal2d/skeletal2d.glb", "text/text.glb"]

here = os.path.dirname(os.path.abspath(__file__))
for f in FILES:
    shutil.copyfile(os.path.join(SRC, f), os.path.join(here, os.path.basename(f)))
    print("copied", f)
