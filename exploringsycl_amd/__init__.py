"""MI355X-native TeaLeaf CG backend.

The product is native: ``lib/libteahip.so`` (HIP kernels for gfx950 behind
the C ABI declared in ``include/tea_hip.h``) and ``bin/tealeaf`` (the host
driver, linked to it through the kernel_interface.h adapter).  This Python
package is a thin mirror of that C ABI for tests and benchmarks:

- :mod:`.native` loads the library (and fails loudly if it is not built);
- :mod:`.chunk` wraps one chunk handle with the reference's entry-point names;
- :mod:`.deck` writes tea.in decks, including the seeded random-rectangle
  problems;
- :mod:`.runner` runs ``bin/tealeaf`` on a deck and returns its run record.
"""

from .deck import Deck, seeded_rectangles, splitmix64  # noqa: F401

__all__ = ["Deck", "seeded_rectangles", "splitmix64"]
