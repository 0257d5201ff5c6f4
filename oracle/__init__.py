"""FP64 CPU oracle for the SHORE fit (test infrastructure only).

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s CPU-baseline
leg import this package, as the checker / CPU baseline. The product library
(``cs_dmri_amd``) never imports, links or calls it.
"""

from .oracle import Oracle, load  # noqa: F401
