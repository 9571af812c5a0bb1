"""CPU restatement of the azimuth-crossing test of the position detectors --
TEST INFRASTRUCTURE ONLY.

Only ``tests/`` uses this module, as the checker of the GPU detection kernel
(ffa_detect_crossing) and of the detector classes; the product package never
imports it.

``crossing_mask`` follows detect_azimuth_crossing
(src/FFAG_MathTools.py:1633-1695): both azimuths reduced to [0, 2 pi) by
``x - 2 pi floor(x / 2 pi)`` (fast_mod_parallel, src/FFAG_MathTools.py:
1553-1557); a particle whose pre-step survive flag is set crosses when

* ``post >= pre``: ``pre <= target <= post``;
* ``post <  pre`` (the step wrapped): ``target >= pre`` (and ``< 2 pi``) or
  ``0 <= target <= post``.

``crossing_mask_portion`` is the detect_azimuth_crossing_portion form
(src/FFAG_MathTools.py:1697-1757): the same test over a list of rows.

Pinning: for forward motion by less than a turn the rule is the continuous
statement "some ``target + 2 pi k`` lies in ``[fi_pre, fi_post]``", which the
tests evaluate by brute force over k; the boundary cases (target equal to
either end, a wrap landing exactly on 0, NaN azimuths, dead particles) are
fixed by hand from the rule above.
"""
from __future__ import annotations

import numpy as np

TWO_PI = 2 * np.pi


def reduce_azimuth(x):
    x = np.asarray(x, dtype=np.float64)
    return x - TWO_PI * np.floor(x / TWO_PI)


def crossing_mask(fi_pre, fi_post, alive, target):
    target = target % TWO_PI
    pre, post = reduce_azimuth(fi_pre), reduce_azimuth(fi_post)
    alive = np.asarray(alive, dtype=bool)
    out = np.zeros(pre.shape, dtype=bool)
    for i in range(pre.size):
        if not alive[i]:
            continue
        a, b = pre[i], post[i]
        if b >= a:
            out[i] = a <= target <= b
        else:
            out[i] = (target >= a and target < TWO_PI) or (target >= 0 and target <= b)
    return out


def crossing_mask_portion(fi_pre, fi_post, alive, target, rows):
    full = crossing_mask(np.asarray(fi_pre)[rows], np.asarray(fi_post)[rows], np.asarray(alive)[rows], target)
    return full
