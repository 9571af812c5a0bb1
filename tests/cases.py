"""Shared test inputs: one lattice, its Taylor maps, SEO tables, RF file and
seeded bunches, packaged both for the oracle (``oracle.Machine``) and for the
GPU path (``FFAG_GlobalParameters`` with the product's field objects).

Lattices: ``prescribed`` is the 12-sector scaling FFAG exactly as the
workload specifies it (D polarity -0.728), which is vertically unstable under
the reference field model, so it is only used for short (single-step /
few-step) parity checks; ``stable`` changes only the D polarity to -1.4
(tunes measured with the oracle: Qx, Qz fractional parts ~0.30, ~0.37;
vertical dynamic aperture ~2 cm) and is used for long runs."""
from __future__ import annotations

import numpy as np

from oracle import adapters as A
from oracle import oracle as O
from pyffa_amd import workloads as W
from pyffa_amd.workloads import APERTURE, D_STABLE, bunch, emap_file, lattice, maps, seo, set_bunch_seo  # noqa: F401


def bmap_config(lat):
    return lat.to_config()


def machines(kind="stable", max_order=2, rf=False, aperture=None, harmonics=None, tmpdir=None):
    """Returns (oracle Machine, product FFAG_GlobalParameters) holding the
    same fields."""
    gp = W.global_parameters(kind, max_order=max_order, rf=rf, aperture=aperture, harmonics=harmonics, tmpdir=tmpdir)
    return oracle_machine(gp), gp


def oracle_machine(gp):
    """Oracle Machine holding the same fields as a product
    FFAG_GlobalParameters (oracle.adapters.machine_from)."""
    return A.machine_from(gp)


def oracle_seo(kind="stable"):
    return A.seo_from(seo(kind))


def sync_tracker_from_bunch(tr, B):
    """Make an oracle Tracker continue exactly from a GPU bunch's state:
    state columns, flags, SC endpoint history and masks, injection head."""
    tr.M[:] = B.LocalBunch
    fl = B.ordered("flags").cpu().numpy()
    tr.E_pre[:] = B.ordered("esc_pre").cpu().numpy().astype(np.float64).T
    tr.E_prepre[:] = B.ordered("esc_prepre").cpu().numpy().astype(np.float64).T
    tr.m_pre[:] = (fl & 4) != 0
    tr.m_prepre[:] = (fl & 8) != 0
    tr.head = B._uninject_head
    tr.step_survive[:] = (fl & 2) != 0


STATE_COLS = (0, 1, 2, 3, 4, 5, 9)


def rel_err(a, b, cols=STATE_COLS):
    """Per-column max |a-b| / (|b| + rms(b)) over the state columns."""
    out = {}
    for c in cols:
        x, y = a[:, c], b[:, c]
        scale = np.abs(y) + np.sqrt(np.mean(y * y)) + 1e-300
        out[c] = float(np.max(np.abs(x - y) / scale))
    return out
