"""FP64 CPU oracle for the fdtdz() Yee FDTD hot path -- TEST INFRASTRUCTURE ONLY.

Nothing in the product package (``pjz_amd``) imports, links or runs anything
under this directory.  Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` use it, and only as the checker.

Contents
--------
``helpers``     numpy restatements of pjz's boundary / waveform / frequency /
                source helpers (``src/pjz/_field.py:56-167,271-279``).
``fdtd_spec``   vectorised FP64 stepper of the update equations (numpy, and
                torch-CPU-float64 for autograd gradients).
``brute``       per-cell triple-loop stepper for tiny grids; pins the indexing
                of ``fdtd_spec``.
``fdtd_ref.c``  the same stepper in plain C + OpenMP (the CPU baseline).

Pinning status (what each function is checked against) is listed in the
header of each module.
"""
