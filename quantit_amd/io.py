"""btensor files: the block structure and data of a tensor in one ``.npz``
(plain arrays, no pickled objects), for fixtures and for replaying inputs.

The reference has no serialisation beyond printing
(include/blockTensor/btensor.h:1892-1923); the fields stored here are the
ones that define a btensor there (btensor.h:820-837): sections per
dimension, their sizes and charges, the selection rule and the block list in
its sorted order, each block row-major.
"""
from __future__ import annotations

import numpy as np


def save_parts(path, groups, secs, cvals, sel, blocks):
    """Writes btensor parts (the arguments of ``BTensor.from_parts``)."""
    keys = sorted(blocks)
    rank = len(secs)
    np.savez(path,
             groups=np.asarray(groups, dtype=np.int64),
             nsec=np.array([len(x) for x in secs], dtype=np.int64),
             sizes=np.array([v for x in secs for v in x], dtype=np.int64),
             cvals=np.array([c for d in cvals for q in d for c in q], dtype=np.int64),
             sel=np.asarray(sel, dtype=np.int64),
             index=np.array(keys, dtype=np.int64).reshape(len(keys), rank),
             data=np.concatenate([np.ascontiguousarray(blocks[k], dtype=np.float64).ravel() for k in keys])
             if keys else np.zeros(0))


def load_parts(path):
    """Reads parts written by ``save_parts`` (no pickle is ever loaded)."""
    with np.load(path, allow_pickle=False) as f:
        groups = tuple(int(x) for x in f["groups"])
        ng = len(groups)
        nsec = [int(x) for x in f["nsec"]]
        sizes, cv = f["sizes"].tolist(), f["cvals"].tolist()
        secs, cvals, o = [], [], 0
        for n in nsec:
            secs.append(sizes[o:o + n])
            cvals.append([tuple(cv[(o + i) * ng:(o + i + 1) * ng]) for i in range(n)])
            o += n
        sel = tuple(int(x) for x in f["sel"])
        index, data = f["index"], f["data"]
    blocks, off = {}, 0
    for row in index:
        key = tuple(int(x) for x in row)
        shape = tuple(secs[d][key[d]] for d in range(len(secs)))
        n = int(np.prod(shape)) if shape else 1
        blocks[key] = data[off:off + n].reshape(shape).copy()
        off += n
    if off != data.size:
        raise ValueError("block data does not match the block list")
    return dict(groups=groups, secs=secs, cvals=cvals, sel=sel, blocks=blocks)
