"""CPU oracle of the search-node fast path -- TEST INFRASTRUCTURE ONLY.

This package restates the reference algorithm (dsa110/dsa110-rt) in float64 /
int64 numpy plus one plain-C file for the dedispersion inner loop, so that the
HIP path can be checked against it.  It is imported only by ``tests/``,
``__graft_entry__.smoke()`` (as the checker) and ``bench.py``'s
``cpu_baseline`` leg; the product path (``dsx``) never imports, links or
calls it.

Modules and the reference code each follows:

* ``oracle.corr``   -- nibble unpack and the int64 correlator
  (``src/dsart/services/slow_corr_kernel.py:161-178,220-403,717-744``,
  ``src/dsart/x2/ref/corr2x_ref.py:76-130``).
* ``oracle.image``  -- geometric ``np.add.at`` gridder, textbook
  ``fftshift(Re(ifft2))`` imager and the signed edge mask
  (``src/dsart/x2/ref/corr2x_ref.py:138-166``, ``src/dsart/x2/ref/search2x_ref.py:120-170``,
  ``src/dsart/image/imager.py:57-175``).
* ``oracle.search`` -- subband shift table, shift-and-add dedispersion,
  centred boxcars, float64 running mean/variance, chunk maxima and the
  top-k + merge-radius decoder (``src/dsart/fine_dm/combiner.py:150-655``,
  ``src/dsart/detector/forward.py:163-504``, ``src/dsart/detector/decoder.py:388-533``,
  ``src/dsart/x2/ref/search2x_ref.py:231-313``).

Pinning status of each function is recorded in its docstring; every function
here is pinned by a known-answer, closed-form or brute-force check in
``tests/test_oracle_*.py``.
"""
