"""Read batches: concatenated read bytes plus an offsets table.

A batch of n reads is ``seq`` (uint8) and ``offsets`` (int64, n + 1 entries,
offsets[0] = 0); read r is seq[offsets[r]:offsets[r+1]].  Per-k-mer results
are packed per read, read r owning max(0, len_r - K + 1) entries.
"""
import numpy as np


class ReadBatch:
    """Host (numpy) read batch."""

    def __init__(self, seq, offsets):
        self.seq = np.ascontiguousarray(seq, dtype=np.uint8)
        self.offsets = np.ascontiguousarray(offsets, dtype=np.int64)
        if self.offsets.ndim != 1 or len(self.offsets) < 1:
            raise ValueError('offsets must be a 1-d array of n_reads + 1 entries')
        if self.offsets[0] != 0 or np.any(np.diff(self.offsets) < 0):
            raise ValueError('offsets must start at 0 and be non-decreasing')
        if int(self.offsets[-1]) > len(self.seq):
            raise ValueError('offsets run past the end of seq')

    @classmethod
    def from_strings(cls, reads):
        enc = [r.encode() if isinstance(r, str) else bytes(r) for r in reads]
        offsets = np.zeros(len(enc) + 1, np.int64)
        offsets[1:] = np.cumsum([len(e) for e in enc]) if enc else []
        seq = np.frombuffer(b''.join(enc), dtype=np.uint8) if enc else np.zeros(0, np.uint8)
        return cls(seq.copy(), offsets)

    @property
    def n_reads(self):
        return len(self.offsets) - 1

    @property
    def n_bases(self):
        return int(self.offsets[-1])

    def lengths(self):
        return np.diff(self.offsets)

    def kmer_counts(self, K):
        return np.maximum(self.lengths() - K + 1, 0)

    def kmer_offsets(self, K):
        ko = np.zeros(self.n_reads + 1, np.int64)
        np.cumsum(self.kmer_counts(K), out=ko[1:])
        return ko

    def read(self, r):
        return bytes(self.seq[self.offsets[r]:self.offsets[r + 1]])

    def __len__(self):
        return self.n_reads

    def to_device(self, device='cuda'):
        return DeviceReads.from_host(self, device)


class DeviceReads:
    """Read batch resident in device memory (torch tensors)."""

    def __init__(self, seq, offsets, n_bases=None):
        self.seq = seq
        self.offsets = offsets
        self.n_reads = int(offsets.numel()) - 1
        self.n_bases = int(n_bases if n_bases is not None else offsets[-1].item())

    @classmethod
    def from_host(cls, batch, device='cuda'):
        import torch
        seq = torch.from_numpy(batch.seq).to(device)
        offsets = torch.from_numpy(batch.offsets).to(device)
        return cls(seq, offsets, batch.n_bases)
