"""Result digests: how the benchmark checks the rows its timed runs
produced against fixtures computed elsewhere (tests/golden/).

A result's digest is the SHA-256 of the concatenated SHA-256 digests of
its columns, each over the column's bytes in result order; columns can be
fed in pieces (a streamed result), in order."""
from __future__ import annotations

import hashlib

import numpy as np


class Digest:
    def __init__(self, ncols: int):
        self.h = [hashlib.sha256() for _ in range(ncols)]
        self.rows = 0

    def update(self, cols: list[np.ndarray]) -> None:
        assert len(cols) == len(self.h)
        n = {len(c) for c in cols}
        assert len(n) <= 1, "columns of one result piece differ in length"
        for h, c in zip(self.h, cols):
            h.update(np.ascontiguousarray(c).view(np.uint8).tobytes())
        self.rows += n.pop() if n else 0

    def hexdigest(self) -> str:
        return hashlib.sha256(b"".join(h.digest() for h in self.h)).hexdigest()


def digest(cols: list[np.ndarray]) -> str:
    d = Digest(len(cols))
    d.update(cols)
    return d.hexdigest()
