"""Wide stage-kernel schedule (lanes = right-hand sides), checked on the CPU.

``build_wide_tasks`` cuts heavy rows into parts whose row sums land in a
partial buffer; ``etd2_wide`` walks the blocks of a task with the index
arithmetic restated below and ``etd2_wide_combine`` sums the parts in order.
The emulation uses exactly the kernel's flat indexing (values
``m[s * 64 + i]`` against state ``x[(mo * 64 + i + olo + s) * K + k]``) and
must reproduce the off-diagonal product (M - diag M) X of the CSR operator;
every index is checked against its buffer.
"""

import numpy as np
import scipy.sparse as sp

from prince_amd import workloads
from prince_amd.solvers.layout import (
    WindowLayout,
    build_block_layout,
    build_wide_tasks,
    wide_block_cost,
)


def _layout(run, z=1.0):
    ir = run.int_rates
    m0 = ir._batch_matrix @ ir.photon_vector(z)
    d, nspec = run.cr_grid.d, run.spec_man.nspec
    cm = ir.coupling_mat
    lay = build_block_layout(cm.indptr, cm.indices, m0, nspec, d)
    M = sp.csr_matrix((m0, cm.indices, cm.indptr), shape=(nspec * d, nspec * d))
    return lay, m0, M


def _window_values(wl, m0):
    """One window of the value layout [diag n][band slots] as the fold writes it."""
    sel = wl.sel
    return np.where(sel >= 0, m0[np.maximum(sel, 0)], 0.0)


def _emulate_wide(wl, mv, X, wtasks, wheavy, nslots):
    nspec, d = wl.fwd.nspec, wl.fwd.d
    assert mv.size == wl.stride
    K = X.shape[1]
    xf = X.reshape(-1)  # [nspec][d][K] flattened
    part = np.full((max(1, nslots), d, K), np.nan)
    rows = np.full((nspec, d, K), np.nan)
    for da, b0, b1, slot, _kind in wtasks:
        acc = np.zeros((d, K))
        for b in range(b0, b1):
            mo, olo, ns, voff = (int(v) for v in wl.blocks[b])
            assert 0 <= mo < nspec and 1 <= ns and voff + ns * d <= mv.size
            for s in range(ns):
                for i in range(d):
                    j = i + olo + s
                    if 0 <= j < d:
                        idx = (mo * d + j) * K + np.arange(K)
                        assert idx.max() < xf.size
                        acc[i] += mv[voff + s * d + i] * xf[idx]
        if slot >= 0:
            assert slot < nslots
            part[slot] = acc
        else:
            assert np.isnan(rows[da, 0, 0])
            rows[da] = acc
    for da, s0, np_ in wheavy:
        assert s0 + np_ <= nslots
        acc = part[s0].copy()
        for p in range(1, np_):
            acc += part[s0 + p]
        assert np.isnan(rows[da, 0, 0])
        rows[da] = acc
    return rows


def test_wide_tasks_cover_every_block_once():
    run = workloads.nuclear_run()
    lay, _, _ = _layout(run)
    wtasks, wheavy, nslots = build_wide_tasks(lay, part_cost=24)
    cost = wide_block_cost(lay.blocks)
    seen = np.zeros(lay.blocks.shape[0], dtype=int)
    for da, b0, b1, slot, _ in wtasks:
        assert lay.brow_ptr[da] <= b0 <= b1 <= lay.brow_ptr[da + 1]
        seen[b0:b1] += 1
    assert np.all(seen == 1)
    # every row appears once: either as a whole task or as a heavy row
    whole = set(int(t[0]) for t in wtasks if t[3] < 0)
    heavy = set(int(h[0]) for h in wheavy)
    assert whole.isdisjoint(heavy) and whole | heavy == set(range(lay.nspec))
    # parts of a heavy row take consecutive slots in block order
    for da, s0, np_ in wheavy:
        parts = sorted((int(t[3]), int(t[1]), int(t[2])) for t in wtasks if t[3] >= 0 and t[0] == da)
        assert [p[0] for p in parts] == list(range(s0, s0 + np_))
        assert parts[0][1] == lay.brow_ptr[da] and parts[-1][2] == lay.brow_ptr[da + 1]
        assert all(parts[q][2] == parts[q + 1][1] for q in range(np_ - 1))
    # tasks come longest first; a part never exceeds the budget unless it is one block
    tc = [int(cost[b0:b1].sum()) for _, b0, b1, _, _ in wtasks]
    assert tc == sorted(tc, reverse=True)
    assert all(c <= 2 * 24 or b1 - b0 == 1 for c, (_, b0, b1, _, _) in zip(tc, wtasks))
    # the proton and He-4 rows (about 300 blocks each) are split
    assert len(wheavy) >= 2 and nslots >= 20


def test_wide_emulation_matches_offdiagonal_product():
    run = workloads.nuclear_run()
    lay, m0, M = _layout(run)
    wl = WindowLayout(lay, None)
    mv = _window_values(wl, m0)
    K = 3
    rng = np.random.default_rng(5)
    X = rng.standard_normal((lay.nspec * lay.d, K))
    wtasks, wheavy, nslots = build_wide_tasks(lay, part_cost=24)
    rows = _emulate_wide(wl, mv, X, wtasks, wheavy, nslots)
    Moff = M - sp.diags(M.diagonal())
    ref = (Moff @ X).reshape(lay.nspec, lay.d, K)
    scale = np.abs(Moff) @ np.abs(X)
    err = np.abs(rows - ref) / np.maximum(scale.reshape(ref.shape), 1e-300)
    assert np.all(np.isfinite(rows)) and err.max() < 1e-13
    # the diagonal region of the window holds diag(M) (used by the exponential)
    assert np.array_equal(mv[wl.diag_off : wl.diag_off + lay.n], M.diagonal())


def test_nuclear_layout_has_only_wide_shapes():
    from prince_amd.solvers.layout import wide_shapes_ok

    run = workloads.nuclear_run()
    lay, _, _ = _layout(run)
    assert wide_shapes_ok(lay.blocks)
    bad = lay.blocks.copy()
    bad[0, 1:3] = (2, 3)
    assert not wide_shapes_ok(bad)


def test_dense_split_keeps_the_product():
    """Dense rows (protons, He-4) moved to the mother-major pass: the blocks left
    in the tasks plus the dense table give the same off-diagonal product."""
    from prince_amd.solvers.layout import dense_split

    run = workloads.nuclear_run()
    lay, m0, M = _layout(run)
    wl = WindowLayout(lay, None)
    mv = _window_values(wl, m0)
    nspec, d = lay.nspec, lay.d
    out = dense_split(wl.blocks, lay.brow_ptr, nspec)
    assert out is not None
    rows, row_of, tab, rest, brow_rest = out
    pdg = [run.spec_man.species_refs[r].pdgid for r in rows]
    assert 2212 in pdg and 1000020040 in pdg
    assert rest.shape[0] + int((tab[:, :, 1] > 0).sum()) == wl.blocks.shape[0]
    rng = np.random.default_rng(3)
    X = rng.standard_normal((nspec, d))

    def apply(b, acc_row):
        mo, olo, ns, voff = (int(v) for v in b)
        for s2 in range(ns):
            for i in range(d):
                j = i + olo + s2
                if 0 <= j < d:
                    acc_row[i] += mv[voff + s2 * d + i] * X[mo, j]

    got = np.zeros((nspec, d))
    for da in range(nspec):
        for b in rest[brow_rest[da]:brow_rest[da + 1]]:
            apply(b, got[da])
    for ri, da in enumerate(rows):
        for mo in range(nspec):
            olo, ns, voff = tab[ri, mo]
            if ns > 0:
                apply((mo, olo, ns, voff), got[da])
    Moff = M - sp.diags(M.diagonal())
    ref = (Moff @ X.reshape(-1)).reshape(nspec, d)
    scale = (np.abs(Moff) @ np.abs(X.reshape(-1))).reshape(nspec, d)
    assert np.all(np.abs(got - ref) <= 1e-13 * np.maximum(scale, 1e-300))
    assert row_of[rows[0]] == 0 and (row_of >= 0).sum() == len(rows)


def test_chunked_dense_tasks_keep_the_product():
    """Dense rows (protons, He-4) cut into mother-chunk tasks: emulating the
    narrow stage kernel's task semantics (chunk tasks store dpart[chunk][row],
    the epilogue launch adds each part's share of chunks, then the parts in
    order) gives the off-diagonal product; every block, chunk slot and
    workgroup position is checked."""
    from prince_amd.solvers.layout import TASK_TO_DPART, build_chunked_tasks

    run = workloads.nuclear_run()
    lay, m0, M = _layout(run)
    wl = WindowLayout(lay, None)
    mv = _window_values(wl, m0)
    nspec, d = lay.nspec, lay.d
    wgw = 8
    tasks, fin, rows, row_of, nch = build_chunked_tasks(lay, 96, wgw)
    pdg = [run.spec_man.species_refs[r].pdgid for r in rows]
    assert 2212 in pdg and 1000020040 in pdg
    assert tasks.shape[0] % wgw == 0 and fin.shape[0] % wgw == 0
    assert (row_of >= 0).sum() == len(rows) and nch >= 2
    rng = np.random.default_rng(4)
    X = rng.standard_normal((nspec, d))

    def part_sum(b0, b1):
        acc = np.zeros(d)
        for b in range(b0, b1):
            mo, olo, ns, voff = (int(v) for v in wl.blocks[b])
            assert voff + ns * d <= mv.size
            for s2 in range(ns):
                for i in range(d):
                    j = i + olo + s2
                    if 0 <= j < d:
                        acc[i] += mv[voff + s2 * d + i] * X[mo, j]
        return acc

    covered = np.zeros(wl.blocks.shape[0], dtype=np.int64)
    dpart = np.full((nch, len(rows), d), np.nan)
    got = np.full((nspec, d), np.nan)
    red = {}
    for t, (da, b0, b1, lead, nparts, kind) in enumerate(tasks):
        if da < 0:
            continue
        covered[b0:b1] += 1
        acc = part_sum(b0, b1)
        if kind & TASK_TO_DPART:
            r = row_of[da]
            assert r >= 0 and nparts == 1 and 0 <= lead < nch
            assert np.isnan(dpart[lead, r, 0])
            dpart[lead, r] = acc
            continue
        assert row_of[da] < 0
        wg, wv = divmod(t, wgw)
        assert lead <= wv < lead + nparts <= wgw
        red.setdefault((wg, lead), []).append(acc)
        if wv == lead + nparts - 1:
            got[da] = np.sum(red[(wg, lead)], axis=0)
    assert np.all(covered == 1)
    assert not np.isnan(dpart).any()
    for t, (da, b0, b1, lead, nparts, kind) in enumerate(fin):
        if da < 0:
            continue
        assert b0 == b1 and kind == 0 and row_of[da] >= 0
        wg, wv = divmod(t, wgw)
        assert lead <= wv < lead + nparts <= wgw
        p = wv - lead
        share = dpart[p * nch // nparts:(p + 1) * nch // nparts, row_of[da]].sum(axis=0)
        red.setdefault(("fin", wg, lead), []).append(share)
        if p == nparts - 1:
            got[da] = np.sum(red[("fin", wg, lead)], axis=0)
    Moff = M - sp.diags(M.diagonal())
    ref = (Moff @ X.reshape(-1)).reshape(nspec, d)
    scale = (np.abs(Moff) @ np.abs(X.reshape(-1))).reshape(nspec, d)
    assert np.all(np.abs(got - ref) <= 1e-13 * np.maximum(scale, 1e-300))
