"""Host-side pieces of the learned prior, without a GPU: the closed-form twist maps against scipy, the input
scaling, and the parameter layout the library reports."""

import numpy as np
from scipy.spatial.transform import Rotation


def test_twist_maps_match_scipy(lw):
    rng = np.random.default_rng(0)
    for scale in (0.0, 1e-9, 1e-4, 0.05, 1.0, 2.9, 3.1):
        w = rng.normal(size=3)
        w = w / np.linalg.norm(w) * scale
        R = Rotation.from_rotvec(w).as_matrix()
        assert np.abs(lw.prior.rotmat_of(w) - R).max() < 1e-14
        assert np.abs(lw.prior.rotvec_of(R) - w).max() < 1e-12
    xi = np.array([0.1, -0.2, 0.05, 0.01, 0.02, -0.03])
    assert np.allclose(lw.prior.body_twist(lw.prior.twist_matrix(xi)), xi, atol=1e-15)


def test_motion_inputs(lw):
    norm = lw.prior.Normalizer(np.arange(6), np.full(6, 2.0), np.zeros(6), np.ones(6))
    hist = [np.arange(6) + 2.0, np.arange(6) + 4.0]
    m = norm.motion(hist)
    assert np.allclose(m[:6], 1.0) and np.allclose(m[6:12], 2.0)
    assert np.allclose(m[12:], (-np.arange(6) / 2.0).tolist() * 2)  # missing lags read as zero motion


def test_parameter_layout(lw):
    P = lw._abi.lib().lw_mlp_param_count()
    assert P == 160 * 256 + 256 + 2 * (256 * 256 + 256) + 256 * 16 + 16
    p = lw.prior.init_params(np.random.default_rng(1))
    W1 = p[:160 * 256].reshape(160, 256)
    o4 = 160 * 256 + 256 + 2 * (256 * 256 + 256)
    W4 = p[o4:o4 + 256 * 16].reshape(256, 16)
    assert np.all(W1[152:] == 0) and np.all(W4[:, 6:] == 0) and np.all(p[o4 + 256 * 16:] == 0)
