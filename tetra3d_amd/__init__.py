"""MI355X geometry pipeline for Tetra3D's Camera.Render (see include/t3d.h)."""
