"""FP64 CPU oracle of the LMM association scan -- test infrastructure only.

Importable by tests/, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py``, as the
checker or the timed CPU baseline; the HIP product path never imports, links or executes it.
"""
