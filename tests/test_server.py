"""The serving loop's scheduler on the CPU (dry-run mode: admission, FIFO,
budget accounting, batching, cancellation and shutdown; no device work and no
outputs), and its C-ABI error paths."""
import pytest

REQ_BYTES = 64 * 64 * 32  # one 64x64 request without shadow rays


@pytest.fixture
def cam(wx):
    return wx.camera_basis((400, 300, -200), (128, 128, 128))


def dry(wx, budget, streams=1, max_batch=8):
    s = wx.Server(budget, streams=streams, max_batch_requests=max_batch, dryrun=True)
    s.add_grid(None)
    return s


def test_request_larger_than_budget_is_rejected(wx, cam):
    s = dry(wx, REQ_BYTES - 1)
    with pytest.raises(wx.WoxelError) as e:
        s.submit(0, cam, 64, 64)
    assert e.value.code == wx.WX_ERR_BUDGET
    assert s.stats().rejected == 1
    s.close()


def test_admission_waits_for_released_memory(wx, cam):
    s = dry(wx, 2 * REQ_BYTES)
    t = [s.submit(0, cam, 64, 64) for _ in range(3)]
    s.wait(t[0], 5.0)
    s.wait(t[1], 5.0)
    with pytest.raises(wx.WoxelError) as e:  # the third one does not fit yet
        s.wait(t[2], 0.2)
    assert e.value.code == wx.WX_ERR_STATE
    st = s.stats()
    assert st.completed == 2 and st.budget_in_use == 2 * REQ_BYTES and st.budget_peak <= 2 * REQ_BYTES
    s.release(t[0])
    r = s.wait(t[2], 5.0)
    assert r.status == wx.WX_OK and r.width == 64
    s.release(t[1])
    s.release(t[2])
    assert s.stats().budget_in_use == 0
    s.close()


def test_shadow_requests_take_twice_the_memory(wx, cam):
    s = dry(wx, 2 * REQ_BYTES)
    a = s.submit(0, cam, 64, 64, shadow=True)
    b = s.submit(0, cam, 64, 64)
    s.wait(a, 5.0)
    with pytest.raises(wx.WoxelError):
        s.wait(b, 0.2)
    s.release(a)
    s.wait(b, 5.0)
    s.close()


def test_queued_requests_are_batched_when_memory_frees(wx, cam):
    s = dry(wx, 20 * REQ_BYTES, max_batch=8)
    big = s.submit(0, cam, 64 * 20, 64)  # fills the whole budget
    s.wait(big, 5.0)
    small = [s.submit(0, cam, 64, 64) for _ in range(20)]
    before = s.stats().batches
    s.release(big)
    res = [s.wait(t, 5.0) for t in small]
    st = s.stats()
    assert st.batches - before == 3  # 20 queued requests, at most 8 per launch
    assert sorted(r.batch_requests for r in res) == [4] * 4 + [8] * 16
    assert all(r.submit_s <= r.launch_s <= r.done_s for r in res)
    s.close()


def test_fifo_order_is_kept(wx, cam):
    s = dry(wx, 4 * REQ_BYTES, max_batch=1)
    t = [s.submit(0, cam, 64, 64) for _ in range(4)]
    res = [s.wait(x, 5.0) for x in t]
    assert [r.launch_s for r in res] == sorted(r.launch_s for r in res)
    s.close()


def test_cancel_queued_request(wx, cam):
    s = dry(wx, REQ_BYTES)
    a = s.submit(0, cam, 64, 64)
    b = s.submit(0, cam, 64, 64)
    s.wait(a, 5.0)
    s.cancel(b)
    with pytest.raises(wx.WoxelError):
        s.wait(b, 1.0)
    with pytest.raises(wx.WoxelError) as e:
        s.cancel(a)
    assert e.value.code == wx.WX_ERR_STATE
    assert s.stats().cancelled == 1
    s.close()


def test_shutdown_cancels_queued_and_does_not_hang(wx, cam):
    s = dry(wx, REQ_BYTES)
    a = s.submit(0, cam, 64, 64)
    for _ in range(5):
        s.submit(0, cam, 64, 64)  # cannot be admitted while `a` holds the budget
    s.wait(a, 5.0)
    s.close()


def test_bad_arguments(wx, cam):
    with pytest.raises(wx.WoxelError):
        wx.Server(1 << 20, streams=0, dryrun=True)
    with pytest.raises(wx.WoxelError):
        wx.Server(1 << 20, max_batch_requests=5000, dryrun=True)
    s = dry(wx, 1 << 20)
    with pytest.raises(wx.WoxelError):
        s.submit(7, cam, 64, 64)  # unknown grid
    with pytest.raises(wx.WoxelError):
        s.submit(0, cam, 0, 64)
    with pytest.raises(wx.WoxelError):
        s.release(12345)
    s.close()


def test_batches_hold_at_most_64_grids(wx, cam):
    """A server with more than 64 grids: the packer closes a batch before a 65th
    distinct grid (the limit of one batched launch), so every request is served."""
    s = wx.Server(200 * REQ_BYTES, streams=1, max_batch_requests=1024, dryrun=True)
    ids = [s.add_grid(None) for _ in range(100)]
    blocker = s.submit(ids[0], cam, 64 * 200, 64)  # the whole budget: the rest queues
    s.wait(blocker, 5.0)
    tks = [s.submit(ids[k % 100], cam, 64, 64) for k in range(150)]
    before = s.stats().batches
    s.release(blocker)
    res = [s.wait(t, 5.0) for t in tks]
    assert all(r.status == wx.WX_OK for r in res)
    # 150 requests over 100 grids: 64 distinct grids per batch -> batches of 64, 64, 22
    assert sorted(r.batch_requests for r in res) == [22] * 22 + [64] * 128
    assert s.stats().batches - before == 3
    s.close()


def test_queue_limit_returns_busy(wx, cam):
    s = wx.Server(REQ_BYTES, streams=1, max_batch_requests=8, dryrun=True, max_queued=3)
    s.add_grid(None)
    a = s.submit(0, cam, 64, 64)
    s.wait(a, 5.0)  # holds the whole budget: later requests stay queued
    q = [s.submit(0, cam, 64, 64) for _ in range(3)]
    with pytest.raises(wx.WoxelError) as e:
        s.submit(0, cam, 64, 64)
    assert e.value.code == wx.WX_ERR_BUSY
    assert s.stats().rejected == 1
    s.release(a)
    s.wait(q[0], 5.0)
    s.submit(0, cam, 64, 64)  # room again once a queued request has launched
    s.close()


def test_latency_percentiles_from_histogram(wx, cam):
    """p50 / p99 come from the fixed-size histogram: within a bin (1 %) of the
    exact percentiles of the served latencies."""
    import numpy as np

    s = dry(wx, 64 * REQ_BYTES, max_batch=1)
    lat = []
    for _ in range(60):
        t = s.submit(0, cam, 64, 64)
        r = s.wait(t, 5.0)
        lat.append((r.done_s - r.submit_s) * 1e3)
        s.release(t)
    st = s.stats()
    assert st.completed == 60
    lo, hi = np.percentile(lat, 50, method="lower"), np.percentile(lat, 50, method="higher")
    assert lo / 1.011 <= st.p50_ms <= hi * 1.011, (st.p50_ms, lo, hi)
    assert st.p99_ms >= st.p50_ms and abs(st.mean_ms - np.mean(lat)) <= 1e-6 + 1e-6 * np.mean(lat)
    s.close()
