"""luce-render on MI355X: the reference's host API over hand-written gfx950
HIP kernels (liblrhip.so). `from luce_render import render` mirrors the
reference's `from luce_render import render`."""
from . import render  # noqa: F401
