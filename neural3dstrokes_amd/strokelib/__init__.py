"""Drop-in for the reference's ``source.strokelib`` package (``__init__.py``: get_stroke, compose_strokes)."""
from .strokes import get_stroke, compose_strokes, set_adjoint, get_adjoint, _stroke_fn, _compositing_fn  # noqa: F401
