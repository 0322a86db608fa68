"""Phase timing of the fused cycle (pennant_time_cycles' five entries): K1,
K2, K3, K4 and the learned pass, timed by events between the launches.  The
learned pass has time only in the learned form; the pipelined cycle reports
its zone pass first and no separate corrector."""
import pytest

import pennant_amd as pa
from pennant_amd import decks as wd

pytestmark = pytest.mark.gpu


def closure():
    from pennant_amd.closure import Mlp
    from pennant_amd.train_closure import DEFAULT_OUT
    return Mlp.load(DEFAULT_OUT)


def phases(mode, pipeline=0, mlp=None):
    with pa.Problem(wd.sedov_rect(96, cstop=10 ** 6)) as pb:
        pb.set_tuning("pipeline", pipeline)
        if mlp is not None:
            pb.set_closure(mlp)
        pb.run(4, mode)
        total, ph = pb.time_cycles(6, mode, phases=True)
        assert pb.scalars()["cycle"] == 10
        return total, ph


@pytest.mark.parametrize("mode", [pa.MODE_FUSED, pa.MODE_MIXED])
def test_phases_without_learned_pass(mode):
    total, ph = phases(mode)
    print(f"[phases] mode {mode}: total {total:.3f} ms, {ph}", flush=True)
    assert len(ph) == 5
    assert all(p > 0 for p in ph[:4])
    assert ph[4] == 0.0
    assert sum(ph) <= total * 1.001


def test_phases_learned_and_pipelined():
    m = closure()
    total, ph = phases(pa.MODE_LEARNED, mlp=m)
    print(f"[phases] learned: total {total:.3f} ms, {ph}", flush=True)
    assert len(ph) == 5 and all(p > 0 for p in ph)
    assert sum(ph) <= total * 1.001
    total, ph = phases(pa.MODE_LEARNED, pipeline=1, mlp=m)
    print(f"[phases] learned pipelined: total {total:.3f} ms, {ph}", flush=True)
    assert ph[2] == 0.0
    assert ph[0] > 0 and ph[1] > 0 and ph[3] > 0 and ph[4] > 0
    m.close()
