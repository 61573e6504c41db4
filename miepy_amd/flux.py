"""Cross-section container (src/miepy/flux.py:9)."""

from collections import namedtuple

cross_sections = namedtuple("cross_sections", ["scattering", "absorption", "extinction"])
