"""Epochs streamed to a GPU consumer (kkw_run_epochs): every epoch's paths
handed over in HBM with a stream ordered after its walk while the next epoch
walks; rows equal to the oracle's, buffers reused only after the consumer's
work, a consumer that stops the run."""
import ctypes as C

import numpy as np
import pytest

from helpers import HipRuntime, O, assert_rows_match_oracle, kk, oracle_walk

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def graph12(engine):
    print("[gpu] epochs: R-MAT scale 12 weighted", flush=True)
    engine.graph_rmat(12, 16, seed=4, weighted=True)
    engine.prepare()
    rp, col, w, _ = engine.download()
    return O.Graph(rp, col, w)


def rows_of(parts):
    return (np.concatenate([p.vertices for p in parts]), np.concatenate([p.lengths for p in parts]),
            np.concatenate([p.truncated for p in parts]))


@pytest.mark.parametrize("mode", [kk.MODE_PLAIN, kk.MODE_PERSISTENT], ids=["plain", "persistent"])
@pytest.mark.parametrize("kind", ["node2vec", "ppr"])
def test_epochs_match_the_oracle(engine, graph12, mode, kind):
    if kind == "node2vec":
        job = kk.node2vec_job(50000, 40, 2.0, 0.5, True, seed=17, walker_begin=3)
    else:
        job = kk.ppr_job(50000, 0.1, True, seed=18, start_vertices=np.array([5, 9, 100], np.uint32))
    seen = []

    def consumer(paths, stream):
        assert paths.on_device == 1 and stream
        seen.append((paths.walker_begin, paths.walker_num))
        parts.append(kk.Engine.download_paths(paths))

    parts = []
    print(f"[gpu] run_epochs {kind} mode {mode}", flush=True)
    engine.run_epochs(job, 8192, consumer, mode=mode)
    assert seen == [(job.walker_begin + 8192 * e, min(8192, 50000 - 8192 * e)) for e in range(7)]
    v, l, t = rows_of(parts)
    op, ol, ot, _ = oracle_walk(graph12, job)

    class R:
        vertices, lengths, truncated = v, l, t
    assert_rows_match_oracle(R, op, ol, ot, what=f"{kind} epochs")


def test_consumer_work_on_its_stream(engine, graph12):
    """The consumer enqueues an asynchronous copy of each epoch into pinned
    host memory on the stream it is given and returns at once; the run ends
    when that work is done, and every copied row is the oracle's (the engine
    never overwrote an epoch's buffer before the copy had read it)."""
    rt = HipRuntime()
    job = kk.deepwalk_job(60000, 30, True, seed=19)
    W = 6000
    stride = 32
    host = rt.host_alloc(60000 * stride * 4)
    hlen = rt.host_alloc(60000 * 4)

    def consumer(paths, stream):
        assert paths.stride == stride
        off = paths.walker_begin - job.walker_begin
        rt.copy_async_d2h(host + off * stride * 4, C.cast(paths.vertices, C.c_void_p).value,
                          paths.walker_num * stride * 4, stream)
        rt.copy_async_d2h(hlen + off * 4, C.cast(paths.lengths, C.c_void_p).value, paths.walker_num * 4, stream)

    engine.run_epochs(job, W, consumer, mode=kk.MODE_PERSISTENT)
    v = np.ctypeslib.as_array((C.c_uint32 * (60000 * stride)).from_address(host)).reshape(60000, stride).copy()
    raw = np.ctypeslib.as_array((C.c_uint32 * 60000).from_address(hlen)).copy()
    op, ol, ot, _ = oracle_walk(graph12, job)

    class R:
        vertices, lengths, truncated = v, raw & 0x7FFFFFFF, (raw >> 31).astype(bool)
    assert_rows_match_oracle(R, op, ol, ot, what="stream-ordered copies")
    rt.lib.hipHostFree(host)
    rt.lib.hipHostFree(hlen)


def test_consumer_can_stop_the_run(engine, graph12):
    calls = []

    def consumer(paths, stream):
        calls.append(paths.walker_begin)
        return len(calls) < 3

    with pytest.raises(kk.KkwError) as e:
        engine.run_epochs(kk.deepwalk_job(50000, 10, True, seed=1), 5000, consumer)
    assert e.value.code == kk.KKW_E_CANCELLED and calls == [0, 5000, 10000]
    # an exception in the consumer stops the run and comes back to the caller
    with pytest.raises(ZeroDivisionError):
        engine.run_epochs(kk.deepwalk_job(50000, 10, True, seed=1), 5000, lambda p, s: 1 / 0)


SKIPGRAM = r"""
import sys
import numpy as np
import torch
torch.cuda.init()                     # torch's HIP runtime first: libkkw binds to it
sys.path.insert(0, sys.argv[1])
import knightking_amd as kk
from knightking_amd.skipgram import SkipGram
from oracle import oracle as O

# four communities of 200 vertices: dense inside, sparse across
rng = np.random.default_rng(3)
B, n = 4, 200
V = B * n
blk = np.repeat(np.arange(B), n)
iu = np.triu_indices(V, 1)
p = np.where(blk[iu[0]] == blk[iu[1]], 0.08, 0.002)
keep = rng.random(len(p)) < p
src, dst = iu[0][keep].astype(np.uint32), iu[1][keep].astype(np.uint32)
eng = kk.Engine()
eng.graph_from_edges(V, src, dst, (1 + 4 * rng.random(len(src))).astype(np.float32), make_undirected=True,
                     sort_adj=True)
eng.prepare()
rp, col, w, _ = eng.download()
job = kk.deepwalk_job(40 * V, 40, True, seed=5)
sg = SkipGram(V, dim=32, window=5, negatives=5, seed=1)
train = sg.consumer()
epochs = []
def consume(paths, stream):
    if not epochs:   # the first epoch's rows, as the trainer sees them, against the oracle
        host = kk.Engine.download_paths(paths)
        op, ol, _, _ = O.walk(O.Graph(rp, col, w), O.DEEPWALK, paths.walker_num, 40, True, seed=5)
        assert np.array_equal(host.lengths, ol)
        m = np.arange(41)[None, :] < ol[:, None]
        assert np.all((host.vertices[:, :41] == op) | ~m)
    epochs.append((paths.walker_begin, paths.walker_num))
    train(paths, stream)
eng.run_epochs(job, 4000, consume)
torch.cuda.synchronize()
assert [e[0] for e in epochs] == list(range(0, 40 * V, 4000))
losses = [float(l) for l in sg.losses]
a = torch.randint(0, V, (20000,), device="cuda")
b = torch.randint(0, V, (20000,), device="cuda")
bt = torch.as_tensor(blk, device="cuda")
s = sg.similarity(a, b)
same = s[bt[a] == bt[b]].mean().item()
diff = s[bt[a] != bt[b]].mean().item()
nn_same = (bt[sg.nearest()] == bt).float().mean().item()
print(f"losses {losses[0]:.3f} -> {losses[-1]:.3f}; cosine same community {same:.3f}, across {diff:.3f}; "
      f"nearest neighbour in the same community {nn_same:.3f}")
assert losses[-1] < 0.9 * losses[0] and same - diff > 0.5 and nn_same > 0.9
eng.close()
print("skipgram ok")
"""


def test_skipgram_trained_on_streamed_epochs():
    """DeepWalk epochs streamed straight into a skip-gram trainer on the GPU
    (knightking_amd.skipgram on torch, in a child process where torch's GPU
    comes up first): the trainer's rows are the oracle's, its loss falls, and
    vertices of one community end closer than vertices of two."""
    import os
    import subprocess
    import sys
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    print("[gpu] skip-gram on streamed epochs in a child process", flush=True)
    r = subprocess.run([sys.executable, "-c", SKIPGRAM, root], capture_output=True, text=True, timeout=150)
    print(r.stdout[-500:], flush=True)
    assert r.returncode == 0 and "skipgram ok" in r.stdout, r.stdout[-2000:] + r.stderr[-3000:]
