"""Host checks of the spill tier's slot-image files (no GPU): a file is
accepted only if its header matches the tier and the block and it holds
exactly the bytes that header describes."""
import numpy as np
import pytest

from rocket_amd import _lib, spill


def image(n, fmt, count, extra=0):
    per = spill.SLOT_BLOCK_VALUES[fmt]
    nkb = (n + per - 1) // per
    size = spill.SLOT_HEADER_BYTES + count * (nkb * spill.SLOT_BLOCK_BYTES + spill.SLOT_STATS_BYTES) + extra
    data = np.zeros(max(size, 0), dtype=np.uint8)
    data[:8] = np.frombuffer(b"PRNUSLOT", dtype=np.uint8)
    data[8:16] = np.array([n], dtype=np.int64).view(np.uint8)
    data[16:24] = np.array([_lib.PRECISIONS[fmt], count], dtype=np.int32).view(np.uint8)
    data[24:32] = np.array([nkb], dtype=np.int64).view(np.uint8)
    return data


def test_check_image_accepts_exact_and_refuses_others():
    tier = spill.HostPatternTier(1000, 300, 128, "fp16")
    tier._check_image(image(1000, "fp16", 128), 128, "ok")
    for bad, count in [(image(1000, "fp16", 128, extra=-1), 128),       # truncated
                       (image(1000, "fp16", 128, extra=64), 128),       # padded
                       (image(1000, "fp16", 44), 128),                   # count of another block
                       (image(1001, "fp16", 128), 128),                  # another pattern length
                       (image(1000, "split_bf16", 128), 128),            # another format
                       (np.zeros(10, dtype=np.uint8), 128)]:             # no header
        with pytest.raises(ValueError):
            tier._check_image(bad, count, "bad")


def test_load_files_refuses_truncated_block(tmp_path):
    tier = spill.HostPatternTier(1000, 200, 128, "split_bf16")
    image(1000, "split_bf16", 128).tofile(tmp_path / "block-00000.slots")
    image(1000, "split_bf16", 72, extra=-24).tofile(tmp_path / "block-00001.slots")
    with pytest.raises(ValueError):
        tier.load_files(str(tmp_path))
