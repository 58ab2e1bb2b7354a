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
