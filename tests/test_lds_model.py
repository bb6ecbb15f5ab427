"""The LDS pitch rule of the pass tiles (fft_device.hpp lds_pitch_for, mirrored
in build.py) against the bank model it was chosen with, and the mirror's LDS
sizes against those the device compiler reports for the same tiles."""

import pytest

from reikna_amd.csrc.build import lds_pitch, tile_lds_bytes, tile_pitch
from tests.lds_bank_model import score, tile_cycles


@pytest.mark.parametrize("length,c", [(1024, 4), (1024, 16), (512, 8), (256, 16), (2048, 8)])
def test_rule_beats_odd_pitch_in_the_model(length, c):
    assert tile_pitch(length, c) != lds_pitch(length)
    assert score(length, c, tile_pitch(length, c)) < 0.8 * score(length, c, lds_pitch(length))


@pytest.mark.parametrize("length,c", [(128, 32), (64, 64), (160, 16), (4096, 1)])
def test_odd_pitch_kept(length, c):
    assert tile_pitch(length, c) == lds_pitch(length)


def test_column_exchanges_conflict_free_at_rule_pitch():
    # 1024-point column tile of 4 instances, direct load and store: only the
    # two stage exchanges touch LDS, at the ideal rate
    t, ideal = tile_cycles(1024, 4, tile_pitch(1024, 4), True, "direct", "direct")
    assert t == ideal
    t0, _ = tile_cycles(1024, 4, lds_pitch(1024), True, "direct", "direct")
    assert t0 > 3 * ideal


@pytest.mark.parametrize("length,c,lds", [(1024, 16, 69888), (1024, 8, 69888), (1024, 4, 35072),
                                          (128, 32, 34560), (64, 64, 34304), (256, 16, 37120)])
def test_lds_bytes_match_the_device_compiler(length, c, lds):
    # LDS Size [bytes/block] that hipcc -Rpass-analysis=kernel-resource-usage
    # reports for fft_pass_body<L, C, V> on gfx950
    assert tile_lds_bytes(length, c) == lds
