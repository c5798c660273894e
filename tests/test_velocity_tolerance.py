"""The velocity comparison rule used by the GPU parity tests, checked on the
CPU between two independent FP64 LAPACK eigensolvers (zheevx: bisection and
inverse iteration, as pzheevx; zheevd: divide and conquer)."""
import numpy as np
import pytest

from oracle import phonon as O
from parphom_amd import synth
from tests import _util as U


@pytest.mark.parametrize("iq", [0, 5, 26, 39])
def test_velocity_rule_between_lapack_drivers(iq):
    S = U.oracle_system(3, 4, 1003)
    q = synth.q_path_gkmg()[iq]
    D = S.D(q)
    w1, Z1, _ = O.eig(D)
    w2, Z2 = np.linalg.eigh(D)
    dD = [S.D(q, a) for a in (1, 2, 3)]
    om = O.to_cm(w1)
    v1 = O.velocity(Z1, dD, om)
    v2 = O.velocity(Z2, dD, O.to_cm(w2))
    groups = U.clusters(w1, 1e-6)
    noise = U.vel_noise(dD, om, w1, groups)
    U.assert_vel_close(v2, v1, groups, noise)
    vp1 = U.vel_projected(Z1, dD, om, groups)
    vp2 = U.vel_projected(Z2, dD, om, groups)
    assert np.all(np.abs(vp1 - vp2) <= 1e-6 * np.abs(vp1) + 1e-9 + noise[:, None])
