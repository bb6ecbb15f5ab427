"""Batch planning without a GPU (dry-run plans on a VirtualDevice)."""

import pytest

from reikna_amd.fft import BatchFFT, BatchItem, BatchOptions
from reikna_amd.fft.batch import packed_items, sample_lengths
from reikna_amd.runtime import VirtualDevice


@pytest.fixture(scope="module")
def vdev():
    return VirtualDevice()


def info(vdev, lengths, counts=None, **opts):
    return BatchFFT(vdev, packed_items(lengths, counts), BatchOptions(**opts)).info


def test_single_pass_lengths_one_level(vdev):
    i = info(vdev, [64, 96, 160, 1024, 3072, 4096])
    assert i["levels"] == 1 and i["passes"] == 6 and i["launches"] == 1
    assert i["workspace_bytes"] == 0


def test_chain_lengths(vdev):
    # 8192 -> two passes; 65521 (prime) -> Bluestein over 2^17 = 512 * 256:
    # two chains of two passes (three each with chain passes of at most 256)
    assert info(vdev, [8192])["levels"] == 2
    assert info(vdev, [49152])["levels"] == 2          # 3 * 2^14 = 192 * 256
    assert info(vdev, [65521])["levels"] == 4
    assert info(vdev, [65521], factor_max=256)["levels"] == 6
    assert info(vdev, [997])["levels"] == 1            # single-kernel Bluestein over 2048
    assert info(vdev, [5003])["levels"] == 4           # Bluestein over 2^14 = 128 * 128, twice
    i = info(vdev, [8192, 64, 65521])
    assert i["levels"] == 4 and i["passes"] == 2 + 1 + 4 and i["launches"] == 1
    assert i["workspace_bytes"] >= 8 * (8192 + 2 * 131072)   # one region per item, two for Bluestein chains


def test_lengths_without_a_case_use_bluestein(vdev):
    # 1000 = 2^3 5^3 and 7 * 2^9 have no single-pass case in the batch kernel
    assert info(vdev, [1000])["levels"] == 1
    assert info(vdev, [7 * 2**9])["levels"] == 4


def test_knobs_change_the_chains(vdev):
    assert info(vdev, [4096], single_max=1024)["levels"] == 2
    assert info(vdev, [2003], single_max=2048)["levels"] == 4   # Bluestein over 4096 in chains of two passes
    assert info(vdev, [65536], factor_max=64)["levels"] == 3


def test_tiles_cover_every_instance(vdev):
    # 4096 x 3 copies: one instance per tile; 64 x 1: one tile of 64 slots
    i = info(vdev, [4096, 64], [3, 1])
    assert i["tiles"] == 3 + 1
    assert i["items"] == 4


def test_length_one_items_are_copies(vdev):
    i = info(vdev, [1, 64])
    assert i["launches"] == 2 and i["items"] == 2


def test_grouped_mode(vdev):
    i = info(vdev, [64, 65536, 997], mode="grouped")   # single-shape plans: 65536 = 256 * 256
    assert i["launches"] == 1 + 2 + 1 and i["mode"] == 1


def test_c4_workload_plans(vdev):
    lengths = sample_lengths(512)
    assert len(lengths) == 512 and min(lengths) >= 64 and max(lengths) <= 65536
    i = info(vdev, lengths)
    assert i["launches"] == 1 and i["items"] == 512
    assert i["lds_bytes"] <= 64 * 1024


def test_invalid_items(vdev):
    with pytest.raises(ValueError):
        BatchFFT(vdev, [BatchItem(0)])
    with pytest.raises(ValueError):
        BatchFFT(vdev, [BatchItem(64, count=2, in_dist=32, out_dist=64)])
    with pytest.raises(ValueError):
        BatchFFT(vdev, [BatchItem(64)], BatchOptions(mode="bogus"))


def test_dry_batch_cannot_execute(vdev):
    b = BatchFFT(vdev, [BatchItem(64)])
    with pytest.raises(ValueError, match="VirtualDevice"):
        b(None, None, None)


def test_sample_lengths_families():
    from reikna_amd.fft.batch import _is_prime

    ls = sample_lengths(512, seed=5)
    fam = {"pow2": 0, "3": 0, "5": 0, "prime": 0}
    for n in ls:
        if n & (n - 1) == 0:
            fam["pow2"] += 1
        elif n % 3 == 0 and (n // 3) & (n // 3 - 1) == 0:
            fam["3"] += 1
        elif n % 5 == 0 and (n // 5) & (n // 5 - 1) == 0:
            fam["5"] += 1
        else:
            assert _is_prime(n)
            fam["prime"] += 1
    assert all(v > 64 for v in fam.values())
    assert ls == sample_lengths(512, seed=5)


def test_kernel_flavour(vdev):
    # the light kernel unless a pass needs a single-kernel Bluestein case
    assert info(vdev, [64, 8192, 3 * 2**10])["flavour"] == 0
    assert info(vdev, [64, 997])["flavour"] == 1
    assert info(vdev, [65521])["flavour"] == 0          # Bluestein over 2^17 runs as chains
    assert info(vdev, [2003])["flavour"] == 0           # Bluestein over 4096: two single-pass chains


def test_order_modes_keep_dependencies(vdev):
    lengths = sample_lengths(64, seed=3)
    for order in (0, 1, 2):
        i = info(vdev, lengths, order=order)
        assert i["launches"] == 1 and i["levels"] >= 4


def test_r2c_items(vdev):
    from reikna_amd.fft.batch import packed_items as pk

    # half-length chain + one combine pass
    assert info(vdev, [1024])["levels"] == 1
    i = BatchFFT(vdev, pk([1024], r2c=True)).info
    assert i["levels"] == 2 and i["passes"] == 2
    i = BatchFFT(vdev, pk([131072], r2c=True)).info      # 65536 = 256 * 256, then the combine
    assert i["levels"] == 3 and i["passes"] == 3
    i = BatchFFT(vdev, pk([1994], r2c=True)).info        # 997 (prime): single-kernel Bluestein, then the combine
    assert i["levels"] == 2
    # the half-length spectrum lives in the workspace
    assert BatchFFT(vdev, pk([1024], [3], r2c=True)).workspace_bytes >= 3 * 512 * 8
    # mixed with complex items: still one launch
    i = BatchFFT(vdev, pk([1024, 997, 8192], r2c=[True, False, True])).info
    assert i["launches"] == 1


@pytest.mark.parametrize("n", [2, 5, 1001])
def test_r2c_bad_lengths(vdev, n):
    from reikna_amd.fft.batch import packed_items as pk

    with pytest.raises(Exception, match="even"):
        BatchFFT(vdev, pk([n], r2c=True))


def test_r2c_grouped_refused(vdev):
    from reikna_amd.fft.batch import packed_items as pk

    with pytest.raises(Exception, match="queue mode"):
        BatchFFT(vdev, pk([64], r2c=True), BatchOptions(mode="grouped"))
