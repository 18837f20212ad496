"""MI355X (gfx950) range-limited MD force path: libmdrl.so and its Python host mirror."""
from . import mdrl  # noqa: F401
