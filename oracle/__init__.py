"""FP64 CPU oracle for the EUVlitho mask -> EM -> M3D -> image path.

TEST INFRASTRUCTURE ONLY.  This package is a plain NumPy/LAPACK restatement of
the reference algorithm, written step by step from the reference sources
(cnn/include/header.h, cnn/intensity/abbe/intensity.cpp, cnn/intensity/m3d/m3d.cpp,
cnn/intensity/linear/linear.py, cnn/intensity/stccsocs/socs.py,
cnn/intensity/stccsocs/eigenfunction.py and the elitho helpers they import).
Every function cites the reference file:line it follows.

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import it, and only as the checker or the timed CPU baseline.
The product package ``euvlitho_amd`` never imports it.

How each part is pinned (checked against something other than itself):

* geometry (orders, pupil points, source lattice): exact match with the
  reference fixture ``inputlm.csv`` (1901 rows) and the ``inputxx.csv`` row
  lengths (tests/test_oracle_pins.py).
* mask I/O: bit-exact round trip and the fixture density of ``mask.bin``.
* mask spectrum: ``numpy.fft.fft2`` of the pattern (textbook DFT) and the
  affine identities between the eps, 1/eps and log-eps tables.
* multilayer cascade: an independent Airy/Parratt recursion of the same stack
  (plane-wave reflectance) and the zero-thickness composition identity.
* EM core (ampS): blank-mask closed forms, invariance of the amplitudes to
  eigenvector scaling/permutation, mirror symmetry and the integer-shift
  theorem.  End to end, the reference fixtures ``inputxx.csv``/``emint.csv``
  at the reference-default optics (N = 8705, blank solves N = 1009) are
  reproduced to their 6 significant digits on the masks as the revision that
  wrote them read them (every mask array one row off in y, the out-of-range
  row from the adjacent memory: tests/golden/fixture_reads.py), at the
  fixtures' 2e-5 bar (tests/golden/ref_default_e2e.py, the slow
  test_reference_default_end_to_end).
* Abbe image: blank-mask uniformity, grid refinement sub-sampling identity
  and ``emint.csv`` as above.
* M3D fit: exact recovery of synthetic linear data in both branches (Cramer
  solve where snum >= 8, mean elsewhere) and equality with
  ``numpy.linalg.lstsq`` on the design matrix [1, ls, ms].
* mask generators (maskgen.py): the C++ line/space program with the C
  library's own rand() reproduces the reference fixture ``mask.bin`` bit for
  bit; LinePattern fed the same rand() draws reproduces it too (y-flipped);
  RandomLineSpacePattern by closed forms under constant draws and by its band
  width bounds.
* linear.py / socs.py images: parity unpinned by reference fixtures (the
  reference's intft/intlinear/intsocs CSVs are not in the repository); pinned
  by invariants only (a0=ax=ay=0 => intlinear == intft; SOCS with all kernels
  and an alias-free grid == Abbe with the same source).
"""
