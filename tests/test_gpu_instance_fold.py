"""Instance folding on the device: the child at N x L lanes on the word engine
(generated kernel and interpreter) against the flattened top at L lanes on
the same engine, cycle by cycle through the top-level ports, and a
device-resident run of a 64-instance top folded into 16,384 lanes against the
packed oracle on sampled lane blocks."""
import numpy as np
import pytest

import oracle
from conftest import flush_print
from rrtl_amd.backend import IN_CONST, IN_RANDOM, HipSimulator
from rrtl_amd.instance_fold import InstanceFoldSimulator
from test_instance_fold import run_fold_vs_flat, top

pytestmark = pytest.mark.gpu


class HipEngine(HipSimulator):
    def __init__(self, program, lanes, codegen="on"):
        super().__init__(program, lanes, engine="word", codegen=codegen)


@pytest.mark.parametrize("codegen", ["on", "off"])
def test_fold_matches_flattened_on_the_word_engine(codegen):
    flush_print(f"case: instance fold 8 x 100 lanes vs flattened, codegen={codegen}")
    run_fold_vs_flat(8, 100, 20, lambda p, lanes: HipEngine(p, lanes, codegen))


def test_folded_run_of_64_instances_against_the_oracle():
    """64 instances x 256 stimulus lanes: one device-resident run of the
    child at 16,384 lanes; lane blocks of three instances replayed on the
    packed oracle of the child."""
    from harness import packed_run
    flush_print("case: instance fold 64 x 256 lanes, device run vs packed oracle")
    fold = InstanceFoldSimulator(top(64), "Top", 256, engine=HipEngine)
    p = fold.plan.child
    ins = [(p.signal_index("clk"), IN_CONST, 1), (p.signal_index("din"), IN_RANDOM, 0),
           (p.signal_index("en"), IN_RANDOM, 0)]
    obs = [p.signal_index("o"), p.signal_index("odd")]
    fold.sim.run(300, seed=5, inputs=ins, observe=obs)
    _, _, dg = packed_run(p, fold.total_lanes, 5, 300, ins, obs)
    assert np.array_equal(fold.sim.digest(), dg)
    for inst in (0, 31, 63):
        lo = fold.instance_lane(inst, 0)
        assert np.array_equal(fold.get_top_limbs(f"o{inst}"), fold.sim.get_signal_limbs(p.signal_index("o"))[:, lo:lo + 256])
