"""CPU oracle for the MPS construction + sampling path -- TEST INFRASTRUCTURE ONLY.

This package restates, in float64/complex128 NumPy and plain C, the algorithm of the
reference's CPU scripts (kron_cpu.py, MPS_cpu.py, sampling_cpu.py, and make_cov.py for
instance generation).  It exists to check the HIP product path:

* only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s CPU-baseline leg may
  import, call, link or execute anything in here, and only as the checker or the timed
  CPU baseline -- never as the thing measured or shipped;
* the product package ``bosonsupremacy_amd`` never imports it.

Pinning status of each restated function is recorded in ``gbs_oracle.py``'s header.
"""
