"""Host logic of the query server that runs without a GPU: the shared
gc.freeze hold of running servers."""
import gc

from rocket_amd import serving


def test_gc_hold_is_shared_by_running_servers():
    assert gc.get_freeze_count() == 0
    junk = [[i] for i in range(1000)]
    serving._gc_hold()
    frozen = gc.get_freeze_count()
    assert frozen >= len(junk)
    serving._gc_hold()                              # a second server: no second collection
    assert gc.get_freeze_count() == frozen
    serving._gc_release()
    assert gc.get_freeze_count() == frozen          # one server still running
    serving._gc_release()
    assert gc.get_freeze_count() == 0
    assert serving._gc_servers == 0


def test_device_arena_first_fit_and_merge():
    import pytest
    a = serving.DeviceArena(4096, align=256)
    x = a.alloc(100)                     # rounded to 256
    y = a.alloc(1000)                    # 1024
    z = a.alloc(256)
    assert (x, y, z) == (0, 256, 1280) and a.used == 1536
    assert a.alloc(4096) is None
    a.free(y, 1000)
    assert a.largest_free() == 4096 - 1536      # the tail is larger than the hole
    assert a.alloc(512) == 256                  # first fit: the hole
    with pytest.raises(ValueError):
        a.free(x + 1, 10)                       # not an allocation boundary
    a.free(x, 100)
    with pytest.raises(ValueError):
        a.free(x, 100)                          # double free
    a.free(256, 512)
    a.free(z, 256)
    assert a.free_ranges == [(0, 4096)] and a.used == 0 and a.high_water == 1536


def test_batch_bytes_grows_with_the_batch():
    one = serving.batch_bytes(1024 * 1024, 20000, 10, 1)
    assert one >= 5 * 1024 * 1024 + 8 * 20000
    assert serving.batch_bytes(1024 * 1024, 20000, 10, 64) <= 64 * one
