"""CPU oracle for the semi-numerical reionization hot path.

TEST INFRASTRUCTURE ONLY.  This package restates, in NumPy on the CPU, the
algorithm of the reference (yunfanz/21cmpy) for the four stages
initialize -> perturb_field -> find_HII_bubbles -> delta_T.  It is the checker:
only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import it.  The product package never imports it and has no
CPU fallback.

Each function cites the reference file:line it follows.  Two flavours exist:

* ``*_f32`` functions replay the reference kernels' FP32 operation order
  (the parity target for the GPU path);
* ``*_f64`` functions compute the same quantities in FP64 (the error yardstick).

Pinning status of each module (what checks it besides the GPU):

* ``philox``   - pinned by the Random123 known-answer vectors and by the
  host-compiled rocRAND Philox4x32-10 generator.
* ``schedule`` - pinned by the radius counts derived from the reference rule
  (37 / 40 / 41 at (64,100) / (256,300) / (512,500)) and the float32 list.
* ``sigma``    - ``sig0_trapz`` pinned by the reference-held ``sig0.npz`` table
  (quad integration) evaluated on the ``Del2k.npz`` spectrum; the discrete
  mode sum pinned by brute force on small grids and by its R->0 limit.
* ``bubbles``  - filter windows pinned by closed forms and a brute-force DFT;
  the loop by constant-box closed forms, OR-monotonicity and a pure-Python
  loop restatement on tiny boxes; the nonlinear-input branch (growth scaling
  and box-mean normalisation) by the constant-box value zeta*fc_mean, the
  per-radius box mean fc_mean, growth-only-through-the-input, one scalar per
  radius and zeta-monotonicity, each shown to catch a planted mistake.
* ``grf``      - Hermitian symmetry (irfftn round trip), variances, Philox.
* ``perturb``  - NGP indices by a pure-Python loop restatement, mass
  conservation, zero-displacement block sums, integer-shift rolls.
* ``deltat``   - plane-wave and white-noise closed forms, exact mode counts.
"""
