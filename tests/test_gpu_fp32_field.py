"""The FP32 tier's static field (its own evaluation: position rounded to
FP32, the push's atan2, FP32 order-0 field and FP32 Taylor map; ffa_bfield_eval
mode 3) against the FP64 path's field (mode 0) over the map: the tier's
field error, stated against the 1e-4 relative bound.  Order 6 runs the push's
kernel instance for order-6 maps (the gather written out for that order)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("max_order", [2, 4, 6])
def test_fp32_field_error_below_1e4(torch_cuda, max_order):
    from pyffa_amd import workloads as W
    from pyffa_amd.FFAG_Field import field_at_points
    print(f"[case] fp32 field error order {max_order}", flush=True)
    gp = W.global_parameters("stable", max_order=max_order)
    rng = np.random.default_rng(4)
    n = 400_000
    r = rng.uniform(8.0, 9.3, n)
    fi = rng.uniform(-40.0, 40.0, n)               # unwrapped azimuths, several turns either way
    z = rng.uniform(-0.02, 0.02, n)
    b64 = field_at_points(gp, r, fi, z, "fp64")
    b32 = field_at_points(gp, r, fi, z, "fp32")
    rms = np.sqrt((b64 ** 2).sum(0).mean())
    err = np.sqrt(((b32 - b64) ** 2).sum(0)) / rms
    print("rel rms", np.sqrt(np.mean(err ** 2)), "p99", np.quantile(err, 0.99), "max", err.max(), flush=True)
    assert np.sqrt(np.mean(err ** 2)) < 1e-5
    assert err.max() < 1e-4
    # the FP32 evaluation is not the FP64 one by accident: it differs
    assert np.abs(b32 - b64).max() > 0.0
