"""Host-side layout of the device operator (pure numpy, no GPU needed).

Given the coupling pattern and the values folded at the first refresh
window, derive:

* the off-diagonal pattern: off-diagonal entries that are non-zero at the
  first window, exactly the entries the reference keeps in ``M_off`` after
  ``split_operator`` eliminates zeros (``solvers/etd2.py:234-268``);
* the block-band layout: one block per (daughter species, mother species)
  pair with entries, stored as ``nslot`` diagonals o = j - i of ``d``
  values each (``o_lo <= o < o_lo + nslot``);
* the selection map ``sel`` from block-band slots (and the diagonal) to rows
  of the fold kernel B -- the device gathers those rows once, so every
  later fold writes straight into the split layout (the reference's
  ``build_split_fold``, ``interaction_rates.py:780-822``);
* the work items of the stage kernels: each species row split into parts
  of bounded cost, heaviest rows first.
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np


@dataclass
class BlockLayout:
    nspec: int
    d: int
    blocks: np.ndarray  # int32 [nblocks, 4]: mo, o_lo, nslot, voff
    brow_ptr: np.ndarray  # int64 [nspec+1]
    zslots: int
    sel: np.ndarray  # int64 [zslots + n]
    off_to_M: np.ndarray  # int64 [Z] M.data index of each off-diagonal entry
    off_slot: np.ndarray  # int64 [Z] block-band slot of each off-diagonal entry
    diag_to_M: np.ndarray  # int64 [n], -1 where the diagonal is absent

    @property
    def n(self):
        return self.nspec * self.d

    @property
    def nnz_off(self):
        return int(self.off_to_M.size)


def build_block_layout(indptr, indices, m0, nspec, d, keep_off=True):
    """Block-band layout from a CSR pattern and first-window values ``m0``.

    ``keep_off=False`` drops every off-diagonal entry (photo-hadronic
    coupling switched off).
    """
    n = nspec * d
    indptr = np.asarray(indptr, dtype=np.int64)
    indices = np.asarray(indices, dtype=np.int64)
    m0 = np.asarray(m0, dtype=np.float64)
    rows = np.repeat(np.arange(n, dtype=np.int64), np.diff(indptr))
    diag_mask = indices == rows
    diag_to_M = np.full(n, -1, dtype=np.int64)
    diag_to_M[rows[diag_mask]] = np.flatnonzero(diag_mask)
    off = (~diag_mask) & (m0 != 0.0)
    if not keep_off:
        off[:] = False
    k_off = np.flatnonzero(off).astype(np.int64)
    r = rows[off]
    c = indices[off]
    da, i = r // d, r % d
    mo, j = c // d, c % d
    o = j - i
    key = da * nspec + mo
    if k_off.size:
        ukeys, inv = np.unique(key, return_inverse=True)
        nb = ukeys.size
        olo = np.full(nb, np.iinfo(np.int64).max, dtype=np.int64)
        ohi = np.full(nb, np.iinfo(np.int64).min, dtype=np.int64)
        np.minimum.at(olo, inv, o)
        np.maximum.at(ohi, inv, o)
        nslot = ohi - olo + 1
        size = nslot * d
        voff = np.concatenate([[0], np.cumsum(size)[:-1]]).astype(np.int64)
        zslots = int(size.sum())
        slot = voff[inv] + (o - olo[inv]) * d + i
        blk_da = ukeys // nspec
        blk_mo = ukeys % nspec
    else:
        nb = 0
        olo = nslot = voff = blk_da = blk_mo = np.zeros(0, dtype=np.int64)
        zslots = 0
        slot = np.zeros(0, dtype=np.int64)
    if zslots >= 2**31:
        raise ValueError("coupling matrix too large for 32-bit block offsets")
    sel = np.full(zslots + n, -1, dtype=np.int64)
    sel[slot] = k_off
    sel[zslots:] = diag_to_M
    blocks = np.stack([blk_mo, olo, nslot, voff], axis=1).astype(np.int32).reshape(nb, 4)
    brow_ptr = np.searchsorted(blk_da, np.arange(nspec + 1), side="left").astype(np.int64)
    return BlockLayout(
        nspec=nspec,
        d=d,
        blocks=blocks,
        brow_ptr=brow_ptr,
        zslots=zslots,
        sel=sel,
        off_to_M=k_off,
        off_slot=slot.astype(np.int64),
        diag_to_M=diag_to_M,
    )


def block_cost(blocks):
    """Relative cost of a block in a stage kernel: a load plus one FMA per diagonal."""
    if blocks.size == 0:
        return np.zeros(0, dtype=np.int64)
    return 2 + blocks[:, 2].astype(np.int64)


def build_items(layout: BlockLayout, part_cost=96):
    """Work items [da, b_begin, b_end, part, nparts, part_base] and the part count.

    Rows whose total block cost exceeds ``part_cost`` are split into
    contiguous parts of at most ``part_cost`` (at least one block each);
    split rows come first so their long waves start early.
    """
    cost = block_cost(layout.blocks)
    split_items, plain_items = [], []
    part_base = 0
    for da in range(layout.nspec):
        b0, b1 = int(layout.brow_ptr[da]), int(layout.brow_ptr[da + 1])
        total = int(cost[b0:b1].sum()) if b1 > b0 else 0
        if total <= part_cost or b1 - b0 <= 1:
            plain_items.append([da, b0, b1, 0, 1, 0])
            continue
        bounds = [b0]
        acc = 0
        for b in range(b0, b1):
            if acc > 0 and acc + int(cost[b]) > part_cost:
                bounds.append(b)
                acc = 0
            acc += int(cost[b])
        bounds.append(b1)
        nparts = len(bounds) - 1
        if nparts == 1:
            plain_items.append([da, b0, b1, 0, 1, 0])
            continue
        for p in range(nparts):
            split_items.append([da, bounds[p], bounds[p + 1], p, nparts, part_base])
        part_base += nparts
    items = np.asarray(split_items + plain_items, dtype=np.int32).reshape(-1, 6)
    return items, part_base


def fd_tables(block):
    """FD coefficient table [d, 7] for offsets -3..3 (diagonal excluded) and the diagonal."""
    d = block.shape[0]
    dcoef = np.zeros((d, 7))
    for i in range(d):
        for o in range(-3, 4):
            j = i + o
            if o != 0 and 0 <= j < d:
                dcoef[i, o + 3] = block[i, j]
    return dcoef, np.ascontiguousarray(np.diag(block))
