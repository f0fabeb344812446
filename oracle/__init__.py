"""ORACLE — test infrastructure only: the CPU FP64 restatement of the reference's SIMPLEC
outer iteration (``simplec_oracle``). Imported only by ``tests/``,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py``."""
