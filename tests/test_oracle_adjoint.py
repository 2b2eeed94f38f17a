"""The oracle's path-replay adjoint for Lambert albedos (or_render_adjoint)
against central finite differences of the oracle's own renders with common
random numbers: a Lambert-only material's albedo changes no sampling
decision, so the frame is a polynomial in the albedos and the difference
quotient with eps = 1e-3 matches the derivative to float noise."""
import numpy as np

import oracle
from ymrt_amd import abi, scenes


def _loss(sa, rs, adj):
    img = oracle.Scene(sa).render(rs, 1, threads=8)
    return float((img[..., :3].astype(np.float64) * adj).sum())


def test_adjoint_matches_finite_differences():
    rs = abi.render_settings(24, 24, 8, 5)
    rng = np.random.default_rng(0)
    adj = rng.uniform(0.5, 1.5, (24, 24, 3)).astype(np.float32)
    sa = scenes.cornell_box(24, 24)
    lam = [oracle.lambert_only(m) for m in sa.materials]
    assert lam == [True, True, True, False, False, False]
    img, gd = oracle.Scene(sa).render_adjoint(rs, adj, 1, threads=8)
    g = gd["albedo"]          # the Lambert-only materials 0-2, in order
    assert g.shape == (3, 3)
    # the adjoint's frame is the plain frame
    assert np.array_equal(img, oracle.Scene(sa).render(rs, 1, threads=8))
    eps = 1e-3
    for mi in range(3):
        for c in range(3):
            up, dn = scenes.cornell_box(24, 24), scenes.cornell_box(24, 24)
            up.materials[mi].diffuse_albedo[c] += eps
            dn.materials[mi].diffuse_albedo[c] -= eps
            fd = (_loss(up, rs, adj) - _loss(dn, rs, adj)) / (2 * eps)
            assert abs(fd - g[mi, c]) <= 1e-3 * abs(fd), (mi, c, fd, g[mi, c])


def test_adjoint_zero_albedo_channel():
    """A channel with albedo 0 gets gradient 0 (the estimator divides by the
    albedo); the other channels are unaffected."""
    sa = scenes.cornell_box(16, 16)
    sa.materials[1].diffuse_albedo[2] = 0.0
    rs = abi.render_settings(16, 16, 4, 4)
    adj = np.ones((16, 16, 3), np.float32)
    g = oracle.Scene(sa).render_adjoint(rs, adj, 1, threads=4)[1]["albedo"]
    assert g[1, 2] == 0.0 and g[1, 0] > 0.0
