"""FP64 NumPy oracle for the NNEDI3 / RAVU prescalers -- TEST INFRASTRUCTURE ONLY.

This package is a CPU restatement of the shader math of bjin/mpv-prescalers, written
from the `.hook` files (each function cites the file:line it follows).  It exists to
check the HIP kernels: only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import it.  The product path
(``mpv-prescalers_amd``) never calls it and fails loudly when its HIP library is
missing.

Conventions
-----------
* Planes are ``float64`` arrays ``(H, W)`` (luma) or ``(3, H, W)`` (yuv / rgb, planar).
* Sampling is clamp-to-edge on the texture being read, as mpv's sampler does
  (``HOOKED_texOff``; intermediates such as ``ravu_int11`` clamp on their own size).
* RAVU lookup tables are used FP16-rounded (round-to-nearest-even), which is what the
  reference's ``FORMAT rgba16f`` texture upload stores (``ravu-lite-r3.hook:132-136``).
* NNEDI3 weights are the exact float32 values of the ``intBitsToFloat`` literals.

Pinning status (what checks each function; see tests/test_oracle_*.py):
* weights: bit-exact against the reference text (extractor asserts, CPU test);
* every variant's main pass: against golden vectors produced by executing the
  reference's own hook() text (tests/golden/hookexec.py) on small frames;
* closed forms / invariants: flat input, pass-through positions, AR over constant taps,
  NNEDI3 constant window = mean, transposition symmetry.
"""
