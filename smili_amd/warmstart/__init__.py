"""Learned warm start: a U-Net that proposes the starting image of the
L-BFGS-B imager from the dirty image and dirty beam (no counterpart in the
reference, which starts from a prior image, smili/imaging/lbfgs.py:202-489).

* ``unet``: ctypes binding of libsmili_unet.so (forward, loss and gradient,
  device-resident training with Adam);
* ``data``: seeded synthetic skies, uv coverage, dirty image / beam and
  plain-solve targets;
* ``solve``: warm-started imaging and the time-to-image comparison.
"""
