"""Oracle package: test infrastructure only.

A CPU restatement (plain C, ``gzoracle.c``) of gzippy's parallel gzip decode
path, loaded through ctypes.  Only ``tests/``, ``__graft_entry__.smoke()`` and
the ``cpu_baseline`` leg of ``bench.py`` may import it, and only as the
checker; the product (``gzippy_amd``) never imports anything from here.
"""
