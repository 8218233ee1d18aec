"""The launch-configuration interface of the autotuner (include/cloudsc_tune.h), decided on the host: which
register bounds and compiled-in level counts exist per variant, the round trip through get/set, the legacy
work-group-size entry, and apply() of a recorded cache.  No GPU needed."""
import pytest

from openifs_amd import autotune as A
from openifs_amd import cloudsc as C

KLEV = 211  # a level count no recorded cache uses, so the tests leave the process-wide table as they found it


@pytest.fixture(autouse=True)
def _clear():
    yield
    for v in A.VARIANTS:
        A.set_launch(v, KLEV, 0)


@pytest.mark.parametrize("variant,waves", [(v, w) for v, ws in A.WAVES.items() for w in ws])
@pytest.mark.parametrize("block", [128, 256])
@pytest.mark.parametrize("kspec", [0, 1])
def test_instantiated_launches_round_trip(variant, waves, block, kspec):
    A.set_launch(variant, KLEV, block, waves, kspec)
    assert A.get_launch(variant, KLEV) == (block, waves, kspec)
    assert A.get_block(variant, KLEV) == block


@pytest.mark.parametrize("variant,block,waves,kspec", [
    ("sp", 256, 3, 0),      # sp has the 4-wave bound only
    ("dp", 256, 4, 0),      # dp has the 2-wave bound only
    ("mixed", 256, 2, 0),   # mixed: 3 or 4
    ("sp", 64, 4, 0),       # register bounds only for 128/256-column groups
    ("mixed", 512, 0, 1),   # compiled-in KLEV likewise
    ("sp", 96, 0, 0),       # not a work-group size
    ("sp", 256, 0, 2),      # klev_compiled is 0 or 1
])
def test_launches_that_are_not_instantiated_are_rejected(variant, block, waves, kspec):
    with pytest.raises(C.CloudscError) as e:
        A.set_launch(variant, KLEV, block, waves, kspec)
    assert e.value.code == -1
    assert A.get_launch(variant, KLEV) == (256, 0, 0)  # unchanged: the default


def test_block_zero_removes_the_setting():
    A.set_launch("mixed", KLEV, 128, 3, 1)
    A.set_launch("mixed", KLEV, 0)
    assert A.get_launch("mixed", KLEV) == (256, 0, 0)
    A.set_block("sp", KLEV, 512)
    assert A.get_launch("sp", KLEV) == (512, 0, 0)
    A.set_block("sp", KLEV, 0)
    assert A.get_block("sp", KLEV) == 256


def test_apply_installs_recorded_launches():
    cache = {"gfx950": {"mixed": {str(KLEV): {"block": 128, "waves": 3, "klev_compiled": 1, "nproma": 128}},
                        "sp": {str(KLEV): {"block": 64, "nproma": 64}}}}  # a record without the register fields
    applied = A.apply(cache, arch="gfx950")
    assert applied[("mixed", KLEV)] == (128, 3, 1) and applied[("sp", KLEV)] == (64, 0, 0)
    assert A.get_launch("mixed", KLEV) == (128, 3, 1)
    assert A.get_launch("sp", KLEV) == (64, 0, 0)
    assert A.recommended_nproma("mixed", KLEV, cache=cache) == 128
