"""CPU checker for the HIP LOITS path.  TEST INFRASTRUCTURE ONLY.

Only tests/, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import this package.  The product (``quantom_loits_amd``)
never imports it and fails loudly when its HIP library is missing.
"""
