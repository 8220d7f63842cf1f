"""Host logic of the learned warm start (no GPU): the density compensation
of the radial trajectory follows the trajectory's own sample layout."""
import numpy as np

from bart_amd import inputs as I
from bart_amd import warmstart as WS


def test_radial_dcf_is_the_ramp_of_the_trajectory():
    # w[m] = |k_m| (in units of the readout step) for sample m = r + R spoke
    # of inputs.radial_traj; a transposed layout (spoke-fastest) would fail
    for R, S in ((512, 201), (64, 7)):
        traj = I.radial_traj(R, S)
        k = np.hypot(traj[:, 0].real, traj[:, 1].real) * 2.0  # readout oversampling 2
        w = WS.radial_dcf(R, S)
        assert w.shape == (R * S,) and w.dtype == np.float32
        assert np.allclose(w, k, rtol=1e-5, atol=1e-5)
        assert w.min() == 0.5 and w.max() == R / 2 - 0.5
