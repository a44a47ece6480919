"""Host side of the learned coarse-level correction, on the CPU: the input
features (right-hand side over its RMS, per-node log10 viscosity as the
average over the adjacent elements) that k_cnet and k_cnet_logeta form on
the device, against brute force."""
import numpy as np

from citcom_amd import learn


def test_nodal_average_brute_force():
    dims = (4, 3, 5)   # nox, noy, noz
    nox, noy, noz = dims
    rng = np.random.default_rng(2)
    el = rng.standard_normal((noy - 1) * (nox - 1) * (noz - 1))
    got = learn._nodal_from_elements(el, dims)
    E = el.reshape(noy - 1, nox - 1, noz - 1)
    for j in range(noy):
        for i in range(nox):
            for k in range(noz):
                vals = [E[ey, ex, ez] for ey in (j - 1, j) for ex in (i - 1, i) for ez in (k - 1, k)
                        if 0 <= ey < noy - 1 and 0 <= ex < nox - 1 and 0 <= ez < noz - 1]
                # which elements are averaged is the point, not the order of the sum
                assert abs(got[k + noz * (i + nox * j)] - np.mean(vals)) <= 1e-15 * max(1.0, np.abs(vals).max())


def test_coarse_features_layout_and_scale():
    rng = np.random.default_rng(3)
    nno = 18
    F = rng.standard_normal(3 * nno)
    leta = rng.uniform(-2, 0, nno)
    x, s = learn.coarse_features(F, leta)
    assert x.shape == (nno, learn.CH) and x.dtype == np.float32
    assert s == np.sqrt(np.mean(F * F))
    # node n's three components in channels 0..2, scaled to unit RMS
    assert np.allclose(x[:, :3].reshape(-1), F / s, rtol=1e-7)
    assert abs(np.sqrt(np.mean(x[:, :3].astype(np.float64) ** 2)) - 1.0) < 1e-6
    assert np.array_equal(x[:, 3], leta.astype(np.float32))
    assert not x[:, 4:].any()
    # a zero right-hand side gives zero inputs (and a zero correction)
    x0, s0 = learn.coarse_features(np.zeros(3 * nno), leta)
    assert s0 == 0.0 and not x0[:, :3].any()


def test_random_coarse_rhs_shape_and_seed():
    dims = (5, 5, 3)
    a = learn.random_coarse_rhs(dims, np.random.default_rng(1))
    b = learn.random_coarse_rhs(dims, np.random.default_rng(1))
    assert a.shape == (3 * 5 * 5 * 3,) and np.array_equal(a, b)
