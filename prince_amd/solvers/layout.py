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
* the tasks of the stage kernels: each species row cut into parts of
  bounded cost, the parts of a row packed into one 16-wave workgroup.
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


def build_block_layout(indptr, indices, m0, nspec, d, keep_off=True, data_index=None):
    """Block-band layout from a CSR pattern and first-window values ``m0``.

    ``keep_off=False`` drops every off-diagonal entry (photo-hadronic
    coupling switched off).  ``data_index`` maps the positions of this CSR to
    the rows of the fold kernel (identity when None); it is how the
    transposed layout points back into the forward kernel.
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
    src = k_off if data_index is None else np.asarray(data_index, dtype=np.int64)[k_off]
    dsrc = diag_to_M.copy()
    if data_index is not None:
        ok = dsrc >= 0
        dsrc[ok] = np.asarray(data_index, dtype=np.int64)[dsrc[ok]]
    sel = np.full(zslots + n, -1, dtype=np.int64)
    sel[slot] = src
    sel[zslots:] = dsrc
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


def split_row(cost, b0, b1, nparts):
    """Cut blocks [b0, b1) into nparts contiguous non-empty ranges of similar cost."""
    if nparts <= 1 or b1 - b0 <= 1:
        return [(b0, b1)]
    c = np.cumsum(cost[b0:b1])
    total = c[-1]
    cuts = [b0]
    for p in range(1, nparts):
        pos = b0 + int(np.searchsorted(c, total * p / nparts, side="left")) + 1
        pos = max(pos, cuts[-1] + 1)
        pos = min(pos, b1 - (nparts - p))
        cuts.append(pos)
    cuts.append(b1)
    return [(cuts[i], cuts[i + 1]) for i in range(nparts)]


PART_DIAG, PART_BAND, PART_GENERIC = 0, 1, 2
BAND_MAX_SLOTS = 8


def part_kind(blocks, b0, b1):
    """0: only diagonal blocks; 1: only blocks of <= 8 diagonals; 2: otherwise."""
    if b1 <= b0:
        return PART_DIAG
    blk = blocks[b0:b1]
    if np.all((blk[:, 2] == 1) & (blk[:, 1] == 0)):
        return PART_DIAG
    if np.all(blk[:, 2] <= BAND_MAX_SLOTS):
        return PART_BAND
    return PART_GENERIC


def build_tasks(layout: BlockLayout, part_cost=96, wg_waves=16, pair_rows=False):
    """Stage-kernel tasks [da, b_begin, b_end, lead, nparts, kind], wg_waves per workgroup.

    A row whose blocks cost more than ``part_cost`` is cut into up to
    ``wg_waves`` parts; all parts of a row are packed into consecutive waves
    of one workgroup (first fit, longest rows first) so they can be summed
    through LDS.  Unused waves carry da = -1.  ``pair_rows`` joins two
    consecutive rows that are one part of diagonal blocks each into one
    row-pair task (``TASK_PAIR``).
    """
    cost = block_cost(layout.blocks)
    groups = []
    for da in range(layout.nspec):
        b0, b1 = int(layout.brow_ptr[da]), int(layout.brow_ptr[da + 1])
        total = int(cost[b0:b1].sum()) if b1 > b0 else 0
        nparts = max(1, min(wg_waves, -(-total // part_cost), b1 - b0))
        parts = split_row(cost, b0, b1, nparts)
        groups.append([[da, p0, p1, 0, len(parts), part_kind(layout.blocks, p0, p1)]
                       for p0, p1 in parts])
    if pair_rows:
        groups = pair_row_tasks(groups)
    groups.sort(key=lambda g: -len(g))
    return _pack_first_fit(groups, wg_waves)


TASK_PAIR = 8  # kind flag of a row-pair task
TASK_PAIR_SHIFT = 8  # row da's block count sits above this bit


def pair_row_tasks(groups):
    """Join consecutive one-part diagonal rows (groups in species order) two at a time."""

    def single_diag(g):
        return len(g) == 1 and g[0][5] == PART_DIAG

    out = []
    i = 0
    while i < len(groups):
        g = groups[i]
        if i + 1 < len(groups) and single_diag(g) and single_diag(groups[i + 1]):
            a, b = g[0], groups[i + 1][0]
            assert b[0] == a[0] + 1 and b[1] == a[2]
            out.append([[a[0], a[1], b[2], 0, 1,
                         PART_DIAG | TASK_PAIR | ((a[2] - a[1]) << TASK_PAIR_SHIFT)]])
            i += 2
        else:
            out.append(g)
            i += 1
    return out


TASK_TO_DPART = 4  # kind flag of a dense-row chunk task


def _pack_first_fit(groups, wg_waves):
    """Pack task groups (each into one workgroup) first fit, in the given order."""
    wgs = []
    for grp in groups:
        target = None
        for g in wgs:
            if wg_waves - len(g) >= len(grp):
                target = g
                break
        if target is None:
            target = []
            wgs.append(target)
        lead = len(target)
        for t in grp:
            t = list(t)
            if not t[5] & TASK_TO_DPART:
                t[3] = lead
            target.append(t)
    tasks = []
    for g in wgs:
        tasks.extend(g)
        tasks.extend([[-1, 0, 0, 0, 0, 0]] * (wg_waves - len(g)))
    return np.asarray(tasks, dtype=np.int32).reshape(-1, 6)


def build_chunked_tasks(layout: BlockLayout, part_cost=96, wg_waves=16, min_blocks=100,
                        max_rows=4, chunk_cost=96, place="mother"):
    """Stage tasks with the dense rows cut into mother chunks.

    Rows with at least ``min_blocks`` blocks (at most ``max_rows``) are cut
    into ``nchunk`` contiguous block ranges of similar cost (about
    ``chunk_cost`` each; every dense row gets the same number).  Each range is
    a one-wave task flagged TASK_TO_DPART whose lead field holds the chunk
    index.  ``place="mother"`` puts it among the ordinary rows next to row
    ``first mother of the range`` (it reads mother panels that the
    neighbouring rows pull into L2 as well); ``place="first"`` puts all chunk
    tasks ahead of the ordinary rows.  The ordinary rows are cut and packed as
    in ``build_tasks`` (split rows first fit, then single rows in species
    order).

    Returns None (no dense row) or ``(tasks, fin_tasks, rows, row_of,
    nchunk)``: ``fin_tasks`` holds the dense rows' epilogues, ``min(wg_waves,
    nchunk)`` waves each with an empty block range.
    """
    blocks, brow_ptr = layout.blocks, layout.brow_ptr
    nspec = layout.nspec
    counts = np.diff(np.asarray(brow_ptr, dtype=np.int64))
    order = np.argsort(-counts, kind="stable")[:max_rows]
    rows = sorted(int(r) for r in order if counts[r] >= min_blocks)
    if not rows:
        return None
    cost = block_cost(blocks)
    row_of = np.full(nspec, -1, dtype=np.int32)
    nchunk = 1
    for ri, da in enumerate(rows):
        row_of[da] = ri
        b0, b1 = int(brow_ptr[da]), int(brow_ptr[da + 1])
        nchunk = max(nchunk, -(-int(cost[b0:b1].sum()) // chunk_cost))
    nchunk = min(nchunk, int(counts[rows].min()))
    chunks_at = {}
    for da in rows:
        b0, b1 = int(brow_ptr[da]), int(brow_ptr[da + 1])
        for c, (c0, c1) in enumerate(split_row(cost, b0, b1, nchunk)):
            kind = part_kind(blocks, c0, c1) | TASK_TO_DPART
            at = 0 if place == "first" else int(blocks[c0, 0])
            chunks_at.setdefault(at, []).append([da, c0, c1, c, 1, kind])
    if place not in ("mother", "first"):
        raise ValueError(f"unknown chunk placement {place!r}")
    multi, singles = [], []
    for da in range(nspec):
        for t in chunks_at.get(da, []):
            singles.append([t])
        if row_of[da] >= 0:
            continue
        b0, b1 = int(brow_ptr[da]), int(brow_ptr[da + 1])
        total = int(cost[b0:b1].sum()) if b1 > b0 else 0
        nparts = max(1, min(wg_waves, -(-total // part_cost), b1 - b0))
        parts = split_row(cost, b0, b1, nparts)
        grp = [[da, p0, p1, 0, len(parts), part_kind(blocks, p0, p1)] for p0, p1 in parts]
        (multi if len(parts) > 1 else singles).append(grp)
    multi.sort(key=lambda g: -len(g))
    tasks = _pack_first_fit(multi + singles, wg_waves)
    nf = min(wg_waves, nchunk)
    fin = _pack_first_fit([[[da, 0, 0, 0, nf, 0]] * nf for da in rows], wg_waves)
    return tasks, fin, rows, row_of, nchunk


def dense_split(blocks, brow_ptr, nspec, min_blocks=100, max_rows=2, max_slots=8):
    """Rows with at least ``min_blocks`` blocks (at most ``max_rows`` of them):
    their blocks of <= ``max_slots`` diagonals move to the mother-major dense
    pass.

    Returns None, or ``(rows, row_of, tab, blocks_rest, brow_rest)``: the dense
    rows, ``row_of[da]`` = dense index or -1, ``tab[r][mo] = (o_lo, nslot,
    value offset)`` (nslot 0: no block) and the block list / row pointer of
    everything the ordinary tasks still carry (same order, same offsets).
    """
    counts = np.diff(np.asarray(brow_ptr, dtype=np.int64))
    order = np.argsort(-counts, kind="stable")[:max_rows]
    rows = sorted(int(r) for r in order if counts[r] >= min_blocks)
    if not rows:
        return None
    blocks = np.asarray(blocks)
    keep = np.ones(blocks.shape[0], dtype=bool)
    tab = np.zeros((len(rows), nspec, 3), dtype=np.int32)
    row_of = np.full(nspec, -1, dtype=np.int32)
    for ri, da in enumerate(rows):
        row_of[da] = ri
        for b in range(int(brow_ptr[da]), int(brow_ptr[da + 1])):
            mo, olo, ns, voff = (int(v) for v in blocks[b])
            if ns <= max_slots:
                tab[ri, mo] = (olo, ns, voff)
                keep[b] = False
    per_row = np.array([keep[int(brow_ptr[i]):int(brow_ptr[i + 1])].sum() for i in range(nspec)])
    brow_rest = np.concatenate([[0], np.cumsum(per_row)]).astype(np.int64)
    return rows, row_of, tab, blocks[keep], brow_rest


def transposed_layout(indptr, indices, m0, nspec, d, keep_off=True):
    """Block-band layout of M^T whose selection map points at rows of the
    forward fold kernel (the entry (r, c) of M feeds slot (c, r) of M^T)."""
    import scipy.sparse as sp

    n = nspec * d
    nnz = len(indices)
    carrier = sp.csr_matrix(
        (np.arange(nnz, dtype=np.float64) + 1.0, np.asarray(indices), np.asarray(indptr)), shape=(n, n)
    )
    mt = carrier.T.tocsr()
    mt.sort_indices()
    idx = (mt.data - 1.0).astype(np.int64)
    return build_block_layout(mt.indptr, mt.indices, np.asarray(m0)[idx], nspec, d, keep_off,
                              data_index=idx)


class WindowLayout:
    """Per-window value layout: [diagonal n][forward band slots][transposed band slots].

    ``blocks`` / ``blocks_t`` carry absolute value offsets inside a window;
    ``sel`` is the row map of the fold kernel for the whole window.
    """

    def __init__(self, fwd: BlockLayout, tr: BlockLayout = None):
        n = fwd.n
        self.fwd = fwd
        self.tr = tr
        self.n = n
        self.nspec = fwd.nspec
        self.d = fwd.d
        self.diag_off = 0
        self.fwd_off = n
        self.tr_off = n + fwd.zslots
        self.stride = n + fwd.zslots + (tr.zslots if tr is not None else 0)
        self.blocks = fwd.blocks.copy()
        if self.blocks.size:
            self.blocks[:, 3] += self.fwd_off
        parts = [fwd.sel[fwd.zslots:], fwd.sel[: fwd.zslots]]
        if tr is not None:
            self.blocks_t = tr.blocks.copy()
            if self.blocks_t.size:
                self.blocks_t[:, 3] += self.tr_off
            parts.append(tr.sel[: tr.zslots])
        else:
            self.blocks_t = None
        self.sel = np.concatenate(parts)
        if self.stride >= 2**31:
            raise ValueError("window too large for 32-bit block offsets")

    @property
    def nnz_off(self):
        return self.fwd.nnz_off


def fd_transposed_tables(block):
    """Transposed FD table [d, 7]: entry (i, o) = D[i + o, i] for o != 0."""
    d = block.shape[0]
    dct = np.zeros((d, 7))
    for i in range(d):
        for o in range(-3, 4):
            j = i + o
            if o != 0 and 0 <= j < d:
                dct[i, o + 3] = block[j, i]
    return dct


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
