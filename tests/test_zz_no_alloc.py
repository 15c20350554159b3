"""tsf_hp_compact_noalloc: the compaction with the library's static scratch
(no device allocation) against tsf_hp_compact with its own workspace, called
twice in a row, the second time on a smaller volume, so stale scratch
contents from the first call would show."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _mask(shape, p, seed):
    return (np.random.default_rng(seed).uniform(0, 1, shape) < p).astype(np.uint8)


def _compact(tsf, fn, mask):
    m = tsf.DeviceArray.from_host(mask)
    n = int(mask.sum())
    pos = tsf.DeviceArray((max(n, 1), 3), np.int32)
    cnt = fn(m, pos)
    return cnt, pos.download()[:cnt]


@pytest.mark.parametrize("first,second", [(((200, 160, 256), 0.01), ((20, 12, 28), 0.3)),
                                          (((512, 512, 64), 0.001), ((64, 64, 64), 0.05)),
                                          (((100, 100, 100), 0.02), ((5, 9, 7), 0.5))])
def test_noalloc_matches_workspace_entry_point(tsf, first, second):
    for k, (shape, p) in enumerate((first, second)):
        mask = _mask(shape, p, 40 + k)
        n0, p0 = _compact(tsf, tsf.hp_compact, mask)
        n1, p1 = _compact(tsf, tsf.hp_compact_noalloc, mask)
        print(f"call {k}: {shape} p={p}: {n1} positions")
        assert n1 == n0 == int(mask.sum())
        assert np.array_equal(p1, p0)


def test_noalloc_count_only_and_capacity(tsf):
    mask = _mask((48, 40, 32), 0.2, 7)
    m = tsf.DeviceArray.from_host(mask)
    assert tsf.hp_compact_noalloc(m, None) == int(mask.sum())
    pos = tsf.DeviceArray((10, 3), np.int32)
    with pytest.raises(tsf.TSFError) as e:
        tsf.hp_compact_noalloc(m, pos)
    assert e.value.status == 5
    ref = tsf.DeviceArray((int(mask.sum()), 3), np.int32)
    tsf.hp_compact(m, ref)
    assert np.array_equal(pos.download(), ref.download()[:10])
