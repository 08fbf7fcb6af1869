"""CPU oracle for the 2D TM FDTD + DAS hot path.

TEST INFRASTRUCTURE ONLY. This package is a CPU restatement of the
reference algorithm and is used as the checker: only ``tests/``,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py``
may import it. The product (``waveforge_amd``) never imports, links or
runs anything from here, and must fail loudly if its HIP library is
missing instead of falling back to this code.

Modules
-------
fdtd2d_ref
    Torch (CPU) restatement of ``FDTD2D.step`` with the reference's exact
    fp32 operation order (``src/core/fdtd2d.py:187-266``), Mur ABC
    (``src/core/boundaries.py:230-274``) and soft sources
    (``src/core/sources.py:863-887, 1017-1035, 1060-1071``); runs in fp32
    or fp64.
waveforms_ref
    Restatement of the waveform tables (``src/core/sources.py:112-138`` and
    the ``_compute`` methods).
das_ref
    NumPy restatement of the DAS backprojection variants
    (``examples/brain_mimo_imaging.py:211-266``,
    ``examples/09_brain_clot_dataset_sample.py:152-193``,
    ``examples/10_breast_tumor_mimo.py:148-165``).
brute (C, ``fdtd2d_brute.c``)
    Independent per-cell loop restatement of one step written from the
    update equations, built by ``oracle/Makefile`` into ``oracle/_ref``-free
    ``oracle/build/libwf_brute.so``; used to pin ``fdtd2d_ref``.
"""
