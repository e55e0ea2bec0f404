"""CPU oracle of the FastXC hot path -- TEST INFRASTRUCTURE ONLY.

Each module restates one reference stage step by step in FP64 NumPy (or plain
C for the multi-threaded timing baseline), citing the reference file:line it
follows.  Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import anything here, and only as
the checker.  The product path (``fastxc_amd``) never imports it.

Pinning status of every function is listed in each module header: a function
is "pinned" when a test checks it against something other than its own
formula (a closed form, brute force, a library routine or the reference's own
C code compiled into ``oracle/_ref``).
"""
