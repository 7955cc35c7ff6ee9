"""GPU: batches of independent scenes (mfs_batch) against one-scene solves.

The batch runs the same per-scene arithmetic as a single context -- the same
factor kernels (merged launches only regroup the blocks), the same PCG kernels
(launched in rounds on the scenes' streams) -- so each scene's strengths must
equal the single-scene result bit for bit.

Replaying the rounds from one multi-stream captured hipGraph (use_graph 1)
is switched off in the library: one full GPU test run ended in a host
segmentation fault inside mfs_batch_solve_points in that mode
(test_batch_solutions_meet_tolerance, graph mode), with no cause found in this
code.  Its cases are kept below, skipped; the library refuses use_graph = 1
(checked without a solve).  use_graph 2 replays per-scene graphs (each
scene's chunk of iterations captured from its own stream, no events inside a
capture); it must give the direct launches' results bit for bit, over many
replays with eager work and synchronisations between them."""
import numpy as np
import pytest
import torch

import oracle
from fastmbcui_amd import _lib, scenes

pytestmark = pytest.mark.gpu

DEV = torch.device("cuda", 0)


def _scenes(count=6):
    return [scenes.batch_scene(s, bodies=(2, 6), points=(96, 384)) for s in range(count)]


def _dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).to(DEV)


def _run_batch(scs, use_graph, chunk=3, tol=1e-6):
    b = _lib.Batch(len(scs))
    b.plan([sc.y for sc in scs], 3.0)
    b.factor_points([_dev(sc.x) for sc in scs], [_dev(sc.nx) for sc in scs], [_dev(sc.y) for sc in scs])
    Xs = b.solve_points([_dev(sc.b.astype(np.float32)) for sc in scs], tol=tol, chunk=chunk,
                        use_graph=use_graph)
    return b, Xs


GRAPH_OFF = pytest.mark.skip(reason="graph replay switched off: a host segmentation fault inside "
                                     "mfs_batch_solve_points in graph mode, cause not found")


@pytest.mark.parametrize("use_graph", [pytest.param(True, marks=GRAPH_OFF), False])
def test_batch_equals_single_scene_solves(use_graph):
    scs = _scenes()
    print("scene sizes", [sc.n for sc in scs], flush=True)
    b, Xs = _run_batch(scs, use_graph)
    st = b.stats()
    print(f"batch: factor launches {st.factor_launches} items {st.factor_items}, graph nodes "
          f"{st.graph_nodes}, replays {st.graph_replays}, sweeps {st.sweeps}", flush=True)
    for i, sc in enumerate(scs):
        ctx = _lib.Context(0)
        ctx.plan(sc.y, 3.0)
        ctx.factor_points(_dev(sc.x), _dev(sc.nx), _dev(sc.y), _lib.MFS_FP32)
        X = ctx.solve_points(_dev(sc.b.astype(np.float32)), tol=1e-6)
        assert torch.equal(X, Xs[i]), f"scene {i}"
        assert ctx.stats().iters[0] == b.scene_stats(i).iters[0]


@pytest.mark.parametrize("use_graph", [pytest.param(True, marks=GRAPH_OFF), False])
def test_batch_solutions_meet_tolerance(use_graph):
    scs = _scenes(8)
    b, Xs = _run_batch(scs, use_graph, chunk=2)
    for i, sc in enumerate(scs):
        M = scenes.kernel_matrix(sc.x, sc.nx, sc.y)
        res = oracle.normal_residual(M, sc.b.astype(np.float32), Xs[i].cpu().numpy())
        assert res <= 2e-6, f"scene {i}: {res}"


def test_batch_refuses_graph_replay():
    scs = _scenes(2)
    b = _lib.Batch(2)
    b.plan([sc.y for sc in scs], 3.0)
    with pytest.raises(_lib.MfsError) as e:    # refused before any solve work
        b.solve_points([_dev(sc.b.astype(np.float32)) for sc in scs], use_graph=True)
    assert e.value.code == _lib.MFS_ERR_ARG
    assert "graph replay is switched off" in str(e.value)


def test_batch_rejects_unplanned_scene():
    scs = _scenes(2)
    b = _lib.Batch(2)
    with pytest.raises(_lib.MfsError) as e:
        b.factor_points([_dev(sc.x) for sc in scs], [_dev(sc.nx) for sc in scs], [_dev(sc.y) for sc in scs])
    assert e.value.code == _lib.MFS_ERR_STATE


@pytest.mark.parametrize("chunk", [1, 3])
def test_batch_per_scene_graphs_equal_direct_launches(chunk):
    """Every round replays each running scene's graph once, with eager
    kernels (fork, join, the active-count gather) and a host poll between
    rounds; two solves in a row on the same batch."""
    scs = _scenes(6)
    b = _lib.Batch(len(scs))
    b.plan([sc.y for sc in scs], 3.0)
    b.factor_points([_dev(sc.x) for sc in scs], [_dev(sc.nx) for sc in scs], [_dev(sc.y) for sc in scs])
    Bs = [_dev(sc.b.astype(np.float32)) for sc in scs]
    Xd = b.solve_points(Bs, tol=1e-6, chunk=chunk, use_graph=0)
    it_d = [b.scene_stats(i).iters[0] for i in range(len(scs))]
    for rep in range(2):
        Xg = b.solve_points(Bs, tol=1e-6, chunk=chunk, use_graph=2)
        st = b.stats()
        print(f"chunk {chunk} solve {rep}: {st.graph_captures} graphs of up to {st.graph_nodes} nodes, "
              f"{st.graph_replays} rounds", flush=True)
        assert st.graph_captures >= len(scs) and st.graph_nodes > 0
        for i in range(len(scs)):
            assert torch.equal(Xg[i], Xd[i]), (chunk, rep, i)
            assert b.scene_stats(i).iters[0] == it_d[i]


def test_batch_rejects_unknown_launch_mode():
    scs = _scenes(2)
    b = _lib.Batch(2)
    b.plan([sc.y for sc in scs], 3.0)
    with pytest.raises(_lib.MfsError) as e:
        b.solve_points([_dev(sc.b.astype(np.float32)) for sc in scs], use_graph=3)
    assert e.value.code == _lib.MFS_ERR_ARG


def test_batch_with_row_splits_equals_single_scene_solves():
    """Row-split factor launches (MFS_PARAM_FACTOR_SPLITS): a scene's
    supernodes get the same split count in the batch's merged launches as in
    its own factorisation, so the results stay bit-identical."""
    scs = _scenes(5)
    b = _lib.Batch(len(scs))
    b.set_param(_lib.MFS_PARAM_FACTOR_SPLITS, 8)
    b.plan([sc.y for sc in scs], 3.0)
    b.factor_points([_dev(sc.x) for sc in scs], [_dev(sc.nx) for sc in scs], [_dev(sc.y) for sc in scs])
    Xs = b.solve_points([_dev(sc.b.astype(np.float32)) for sc in scs], tol=1e-6, chunk=3)
    print(f"batch factor launches with splits: {b.stats().factor_launches}", flush=True)
    for i, sc in enumerate(scs):
        ctx = _lib.Context(0)
        ctx.set_param(_lib.MFS_PARAM_FACTOR_SPLITS, 8)
        ctx.plan(sc.y, 3.0)
        ctx.factor_points(_dev(sc.x), _dev(sc.nx), _dev(sc.y), _lib.MFS_FP32)
        X = ctx.solve_points(_dev(sc.b.astype(np.float32)), tol=1e-6)
        assert torch.equal(X, Xs[i]), f"scene {i}"


@pytest.mark.parametrize("precision,products", [(_lib.MFS_FP32, 2), (_lib.MFS_FP32, 1), (_lib.MFS_FP64, 0)])
def test_batch_factor_in_one_launch_equals_launch_per_bucket(precision, products):
    """Every scene's supernodes in one launch over a device queue (mode 2)
    against one launch per tile size and split count (mode 1): the same units
    run the same code on the same scratch, so the factors are equal bit for
    bit; 24 scenes drawn from the uneven batch's ranges (2-32 bodies, 256-2048
    points per body, capped at 16 bodies x 512 points here) with row splits
    on.  The automatic mode (0) gives the same factor again.  FP32 with the
    Gram's products split into BF16 (2, the default) and on FP32 MFMA (1)."""
    scs = [scenes.batch_scene(s, bodies=(2, 16), points=(256, 512)) for s in range(24)]
    vals = {}
    for mode in (2, 1, 0):
        b = _lib.Batch(len(scs))
        b.set_factor_mode(mode)
        b.set_param(_lib.MFS_PARAM_FACTOR_SPLITS, 4)
        b.set_param(_lib.MFS_PARAM_GRAM_SPLIT, products)
        b.plan([sc.y for sc in scs], 3.0)
        b.factor_points([_dev(sc.x) for sc in scs], [_dev(sc.nx) for sc in scs], [_dev(sc.y) for sc in scs],
                        precision)
        st = b.stats()
        print(f"factor mode {mode}: launches {st.factor_launches}, items {st.factor_items}, "
              f"{st.factor_ms:.1f} ms", flush=True)
        if mode != 0:
            assert (st.factor_launches == 1) == (mode == 2)
        vals[mode] = [b.scene_factor_export(i) for i in range(len(scs))]
    for i in range(len(scs)):
        assert np.array_equal(vals[2][i], vals[1][i]), f"scene {i}"
        assert np.array_equal(vals[0][i], vals[1][i]), f"scene {i}"


def test_batch_factor_mode_is_checked():
    b = _lib.Batch(1)
    with pytest.raises(_lib.MfsError):
        b.set_factor_mode(3)
