"""Host logic of the learned-guess gate: when the network is consulted."""
import types

from mantle_convection_lab_amd.model import Model


def _consults(outcomes, steps):
    """Steps on which the network is consulted, given its win/loss outcomes
    in consultation order (True = the network's guess was accepted)."""
    m = types.SimpleNamespace(_net_losses=0, _net_idle=0, NET_LOSSES=Model.NET_LOSSES,
                              NET_RECHECK=Model.NET_RECHECK)
    it = iter(outcomes)
    used = []
    for step in range(steps):
        if Model._consult_net(m):
            used.append(step)
            won = next(it)
            m._net_losses = 0 if won else m._net_losses + 1
    return used


def test_consulted_every_step_while_winning():
    assert _consults([True] * 30, 30) == list(range(30))


def test_backs_off_after_two_losses_and_rechecks():
    # wins at steps 0 and 1, then always loses: consulted at 2 and 3, then
    # every NET_RECHECK steps
    used = _consults([True, True] + [False] * 30, 40)
    r = Model.NET_RECHECK
    assert used[:4] == [0, 1, 2, 3]
    assert used[4:] == [3 + r * k for k in range(1, len(used) - 3)]


def test_a_win_on_recheck_restores_every_step():
    r = Model.NET_RECHECK
    used = _consults([False, False, True] + [True] * 40, 2 + r + 10)
    assert used[:3] == [0, 1, 1 + r]
    assert used[3:] == list(range(2 + r, 2 + r + 10))
