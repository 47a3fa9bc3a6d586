"""Test infrastructure only: the CPU oracle for the LightFEM Q4 + PCG path.

``oracle.lfo`` wraps the C restatement (``lfo.c``, built into ``build/liblfo.so``)
and ``oracle/_ref`` holds the compiled reference itself (this container only,
recipe in ``Makefile``). Only tests/, ``__graft_entry__.smoke()`` and bench.py's
CPU-baseline leg may use anything here, and only as the checker.
"""
