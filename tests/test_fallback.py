"""The learned-correction fallback switch of the host driver
(mdforce.mlp_train.LearnedForce) on the oracle path: engines backed by the
oracle's force_pot (yukawa_p3m_ord6.py:6-29) and a stand-in network whose
correction is far outside any budget.  When the judge refuses the learned
form, the forces the driver returns are bitwise the plain engine's; within
the budget the network is attached to the reduced engine."""
import numpy as np
import pytest
import torch

from oracle import p3m_oracle as O


class OracleEngine:
    def __init__(self, N, L, kappa, alpha, rc, M):
        self.P = O.P3MParams(N, L, kappa, alpha, rc, M, af=1.0)
        self._mlp = None

    def force(self, pos, Z=None, af=1.0 / 3.0):
        U, acc, _ = self.P.force(pos.numpy())
        out = torch.from_numpy(acc) * af
        if self._mlp is not None:
            out = out + af * self._mlp.correction(self, pos)
        return out, torch.tensor([U, 0.0, 0.0, 0.0], dtype=torch.float64)


class StandInNet:
    """Adds 0.05 to every force component; records attach / detach."""

    def __init__(self):
        self.attached = None

    def set_meta(self, **kw):
        pass

    def correction(self, engine, pos):
        return torch.full(pos.shape, 0.05, dtype=torch.float64)

    def attach(self, engine):
        engine._mlp = self
        self.attached = engine

    def detach(self, engine):
        engine._mlp = None
        self.attached = None


@pytest.fixture(scope="module")
def setup():
    N = 300
    L = O.box_length(N)
    pos, _ = O.initial_split(3, 4, N, [L, L, L], 0.01)
    red = OracleEngine(N, L, 2.0, 1.0, 2.5, 16)
    plain = OracleEngine(N, L, 2.0, 0.9, 3.5, 16)
    return torch.from_numpy(pos), red, plain


def test_fallback_forces_are_bitwise_the_plain_path(setup):
    from mdforce.mlp_train import LearnedForce
    pos, red, plain = setup
    net = StandInNet()
    drv = LearnedForce(net, red, plain, 2.0, 1.0, budget=1e-5)
    with pytest.raises(RuntimeError, match="decide"):
        drv.force(pos)
    form, est = drv.decide(pos)
    assert form == "plain" and est > 1e-5
    assert net.attached is None and red._mlp is None
    a_drv, u_drv = drv.force(pos, None, af=1.0)
    a_plain, u_plain = plain.force(pos, None, af=1.0)
    assert torch.equal(a_drv, a_plain) and torch.equal(u_drv, u_plain)
    # a budget of zero refuses even a perfect correction
    drv0 = LearnedForce(net, red, plain, 2.0, 1.0, budget=0.0)
    assert drv0.decide(pos)[0] == "plain"
    assert torch.equal(drv0.force(pos, None, af=1.0)[0], a_plain)


def test_learned_form_attaches_the_network(setup):
    from mdforce.mlp_train import LearnedForce
    pos, red, plain = setup
    net = StandInNet()
    drv = LearnedForce(net, red, plain, 2.0, 1.0, budget=1e9)
    form, est = drv.decide(pos)
    assert form == "learned" and net.attached is red
    a_drv, _ = drv.force(pos, None, af=1.0)
    red._mlp = None
    a_red, _ = red.force(pos, None, af=1.0)
    np.testing.assert_array_equal(a_drv.numpy(), (a_red + 0.05).numpy())
    # deciding again with a tight budget detaches it
    drv.budget = 1e-5
    assert drv.decide(pos)[0] == "plain" and net.attached is None
