"""The restart-length search of fem_codes_amd.autotune (host logic): the
fastest converged candidate wins, unconverged candidates are never chosen,
and a search where nothing converges fails loudly."""
import pytest

from fem_codes_amd import FemGpuError, autotune


def test_fastest_converged_wins():
    table = {20: (3.0, True), 30: (1.0, False), 45: (2.0, True), 64: (2.5, True)}
    km, times = autotune.tune_restart(lambda k: table[k], restarts=(20, 30, 45, 64))
    assert km == 45
    assert times == {20: 3.0, 30: None, 45: 2.0, 64: 2.5}


def test_nothing_converged_raises():
    with pytest.raises(FemGpuError):
        autotune.tune_restart(lambda k: (1.0, False), restarts=(10, 20))


def test_key_names():
    assert autotune.key_str((512, autotune.ORDER_MESH, 4)) == "512/mesh/4"
    assert autotune.key_str((256, autotune.ORDER_BFS, 0)) == "256/bfs/0"
