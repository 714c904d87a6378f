"""Matriplex layout helpers (byte layout of the reference, propagate-tor-test_tbb.cpp:68-125)."""
from dataclasses import dataclass

import numpy as np

TRK_FIELDS = 28  # par 6 + cov 21 + q 1 (int32)
HIT_FIELDS = 9   # pos 3 + cov 6


@dataclass(frozen=True)
class Shape:
    nevts: int
    ntrks: int
    nlayer: int
    bsize: int

    @property
    def nb(self):
        return self.ntrks // self.bsize

    @property
    def nblk(self):
        return self.nevts * self.nb

    @property
    def ntrk_total(self):
        return self.nevts * self.ntrks

    @property
    def trk_floats(self):
        return self.nblk * TRK_FIELDS * self.bsize

    @property
    def hit_floats(self):
        return self.nblk * self.nlayer * HIT_FIELDS * self.bsize

    def tuple(self):
        return (self.nevts, self.ntrks, self.nlayer, self.bsize)


def tracks_to_arrays(trk, bsize):
    """Flat Matriplex track buffer -> (par (n,6), cov (n,21), q (n,)) by flat track index."""
    t = np.asarray(trk, np.float32).reshape(-1, TRK_FIELDS, bsize)
    par = t[:, :6, :].transpose(0, 2, 1).reshape(-1, 6)
    cov = t[:, 6:27, :].transpose(0, 2, 1).reshape(-1, 21)
    q = np.ascontiguousarray(t[:, 27, :]).view(np.int32).reshape(-1)
    return par, cov, q


def hits_block_view(hits, nlayer, bsize):
    """Flat hit buffer -> view (nblk, nlayer, 9, bsize)."""
    return np.asarray(hits, np.float32).reshape(-1, nlayer, HIT_FIELDS, bsize)
