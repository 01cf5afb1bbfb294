"""MI355X (gfx950) implementation of the Neural 3D Strokes stroke-field hot path.

Subpackages / modules:

* ``strokelib``   drop-in for the reference's ``source.strokelib`` (get_stroke,
                  compose_strokes, _stroke_fn, _compositing_fn)
* ``gridencoder`` drop-in for the reference's ``source.gridencoder`` (GridEncoder, grid_encode)
* ``render``      fused stroke-field renderer (stroke SDF -> compositing -> ray march)
* ``plain``       plain per-op PyTorch formulation of the same math (the comparison baseline)
* ``scenes``      seeded synthetic stroke scenes and camera rays
* ``build``       in-tree hipcc build of ``libn3ds.so``

All operators call hand-written HIP kernels through the C-ABI in
``include/n3ds.h``; there is no CPU fallback.
"""
__version__ = '0.1.0'
