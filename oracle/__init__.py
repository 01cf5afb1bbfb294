"""CPU oracle for the Neural 3D Strokes stroke-field hot path.

TEST INFRASTRUCTURE ONLY.  This package restates, in plain PyTorch on the CPU,
the arithmetic of the reference's stroke kernels (``source/strokelib/src``),
compositing kernels, ray-march helpers (``source/utils/render.py``) and hash-grid
encoder (``source/gridencoder``).  It is the checker the HIP path is compared
against.  Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline``
leg of ``bench.py`` may import it; the product path (``neural3dstrokes_amd``)
never imports or calls it.

Every function is dtype-generic: run it in float64 for the accuracy oracle
("exact" adjoints come from torch.autograd of the float64 forward), or in
float32 to emulate the reference's single-precision operation order.

Pinning (what checks each function against something other than itself) is
listed per function in the module docstrings and in the tests under
``tests/test_oracle_*.py``.
"""
