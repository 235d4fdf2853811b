"""Skip-gram with negative sampling (word2vec; Mikolov et al., "Distributed
representations of words and phrases and their compositionality", NIPS 2013)
trained on walks in HBM: the embedding step that follows DeepWalk and
node2vec walks, fed epoch by epoch by Engine.run_epochs while the next epoch
walks (the consumer WalkConfig::set_output_consumer hands PathSets to in
KnightKing, include/walk.hpp:133-138, without the copy to the host).

torch is plumbing here: an epoch's rows and lengths become tensors over the
engine's buffers without a copy, and the trainer's work runs on torch's
stream, ordered after the walk that wrote the epoch and before the engine
reuses its buffers.  torch must initialise its GPU before the first Engine
(one HIP runtime per process).
"""
from __future__ import annotations

import ctypes as C

import torch


def epoch_tensors(paths, device):
    """(vertices int32 [n, stride], lengths int32 [n]) over an epoch's device
    buffers (a kkw_paths with on_device = 1), without a copy."""
    n, s = int(paths.walker_num), int(paths.stride)

    class _CAI:
        def __init__(self, ptr, shape):
            self.__cuda_array_interface__ = {"shape": shape, "typestr": "<i4", "data": (ptr, False),
                                             "version": 2, "strides": None}

    v = _CAI(C.cast(paths.vertices, C.c_void_p).value, (n, s))
    l = _CAI(C.cast(paths.lengths, C.c_void_p).value, (n,))
    return torch.as_tensor(v, device=device), torch.as_tensor(l, device=device)


class SkipGram:
    """Vertex embeddings (input vectors `emb`, output vectors `ctx`) trained by
    minibatch SGD on (center, context) pairs within `window` positions of each
    other on a path, each with `negatives` contexts drawn from the visit
    counts raised to 3/4 (word2vec's unigram^0.75); a minibatch moves each
    vertex by the mean gradient of its pairs."""

    def __init__(self, num_vertices: int, dim: int = 64, window: int = 5, negatives: int = 5, lr: float = 0.5,
                 seed: int = 0, device="cuda", batch_pairs: int = 1 << 18):
        self.V, self.dim, self.window, self.k, self.lr = num_vertices, dim, window, negatives, lr
        self.device = torch.device(device)
        self.gen = torch.Generator(device=self.device)
        self.gen.manual_seed(seed)
        self.emb = (torch.rand(num_vertices, dim, generator=self.gen, device=self.device) - 0.5) / dim
        self.ctx = torch.zeros(num_vertices, dim, device=self.device)
        self.counts = torch.zeros(num_vertices, device=self.device)
        self.batch_pairs = batch_pairs
        self.losses = []

    def pairs(self, verts: torch.Tensor, lens: torch.Tensor):
        """(center, context) pairs of every path, both directions, as int64."""
        L = (lens & 0x7FFFFFFF).long()
        pos = torch.arange(verts.shape[1], device=self.device)
        cs, xs = [], []
        for d in range(1, self.window + 1):
            ok = (pos[None, : verts.shape[1] - d] + d) < L[:, None]
            a, b = verts[:, :-d][ok].long(), verts[:, d:][ok].long()
            cs += [a, b]
            xs += [b, a]
        if not cs:
            z = torch.zeros(0, dtype=torch.long, device=self.device)
            return z, z
        return torch.cat(cs), torch.cat(xs)

    def step(self, verts: torch.Tensor, lens: torch.Tensor) -> torch.Tensor:
        """One pass over an epoch's paths; returns the mean loss as a tensor."""
        L = (lens & 0x7FFFFFFF).long()
        mask = torch.arange(verts.shape[1], device=self.device)[None, :] < L[:, None]
        self.counts += torch.bincount(verts[mask].long(), minlength=self.V).float()
        noise = self.counts.pow(0.75)
        noise = noise / noise.sum()
        c_all, x_all = self.pairs(verts, lens)
        total, n = torch.zeros((), device=self.device), 0
        for b in range(0, c_all.numel(), self.batch_pairs):
            c, x = c_all[b: b + self.batch_pairs], x_all[b: b + self.batch_pairs]
            neg = torch.multinomial(noise, c.numel() * self.k, replacement=True, generator=self.gen).view(-1, self.k)
            ec, ex, en = self.emb[c], self.ctx[x], self.ctx[neg]
            sp = (ec * ex).sum(-1)
            sn = (en * ec[:, None, :]).sum(-1)
            total = total + -(torch.nn.functional.logsigmoid(sp) +
                              torch.nn.functional.logsigmoid(-sn).sum(-1)).sum()
            n += c.numel()
            gp = torch.sigmoid(sp) - 1.0              # d loss / d sp
            gn = torch.sigmoid(sn)                    # d loss / d sn
            g_c = gp[:, None] * ex + (gn[..., None] * en).sum(1)
            # one minibatch updates a vertex once, by the mean of its pairs'
            # gradients: summing them (a hub appears thousands of times in a
            # batch) would take steps far beyond what sequential word2vec
            # takes, and diverge
            nc = torch.bincount(c, minlength=self.V).clamp_(min=1).float()
            xo = torch.cat([x, neg.reshape(-1)])
            nx = torch.bincount(xo, minlength=self.V).clamp_(min=1).float()
            self.emb.index_add_(0, c, g_c / nc[c][:, None], alpha=-self.lr)
            g_o = torch.cat([gp[:, None] * ec, (gn[..., None] * ec[:, None, :]).reshape(-1, self.dim)])
            self.ctx.index_add_(0, xo, g_o / nx[xo][:, None], alpha=-self.lr)
        loss = total / max(n, 1)
        self.losses.append(loss)
        return loss

    def consumer(self):
        """A Engine.run_epochs consumer: trains on each epoch on torch's
        stream, ordered after the epoch's walk and before the engine reuses
        the epoch's buffers."""
        def consume(paths, stream):
            ext = torch.cuda.ExternalStream(stream, device=self.device)
            cur = torch.cuda.current_stream(self.device)
            cur.wait_stream(ext)                      # after the walk that wrote the epoch
            verts, lens = epoch_tensors(paths, self.device)
            self.step(verts, lens)
            ext.wait_stream(cur)                      # the engine reuses the buffers after this
        return consume

    def vectors(self) -> torch.Tensor:
        """Unit input vectors after removing their mean (the direction every
        vector shares under negative sampling says nothing about a vertex)."""
        e = self.emb - self.emb.mean(0, keepdim=True)
        return torch.nn.functional.normalize(e, dim=-1)

    def similarity(self, a: torch.Tensor, b: torch.Tensor) -> torch.Tensor:
        """Cosine similarity of vertex pairs (a[i], b[i]) over vectors()."""
        e = self.vectors()
        return (e[a.long()] * e[b.long()]).sum(-1)

    def nearest(self) -> torch.Tensor:
        """Each vertex's most similar other vertex."""
        e = self.vectors()
        s = e @ e.T
        s.fill_diagonal_(-2.0)
        return s.argmax(1)
