"""Measurement aids: the streaming copy and the random-gather ceiling run,
return positive rates, and reject malformed arguments."""
import math

import pytest

import knightking_amd as kk

pytestmark = pytest.mark.gpu


def test_stream_copy_and_random_gather(engine):
    print("[gpu] diag: stream copy 64 MiB, random gather 256 MiB", flush=True)
    ms = engine.stream_copy_ms(64 << 20, 3)
    assert ms > 0 and math.isfinite(ms)
    lps = engine.random_gather_lines_per_s(256 << 20, reads=8, iters=2)
    assert lps > 0 and math.isfinite(lps)


def test_diag_rejects_bad_arguments(engine):
    with pytest.raises(kk.KkwError) as e:
        engine.random_gather_lines_per_s(64)
    assert e.value.code == kk.KKW_E_INVALID
    with pytest.raises(kk.KkwError) as e:
        engine.random_gather_lines_per_s(1 << 20, reads=0)
    assert e.value.code == kk.KKW_E_INVALID


def test_kept_device_memory_goes_back_when_needed(engine):
    """Device memory the library frees is kept for reuse; an allocation the
    runtime cannot serve beside the kept blocks gets them back first.  A
    100 GiB buffer is freed (and kept), then a 200 GiB one is needed: more
    than an MI355X's 288 GB together, so the second succeeds only if the kept
    block is returned to the runtime first."""
    print("[gpu] device pool: 100 GiB kept, then 200 GiB", flush=True)
    assert engine.random_gather_lines_per_s(100 << 30, reads=1, iters=1) > 0
    assert engine.random_gather_lines_per_s(200 << 30, reads=1, iters=1) > 0
