"""Batched kernel-0 problems: one persistent work-queue launch per time step.

* every problem of a batch ends bit-identical to the same problem stepped alone
  with gr_step (same GR_FAST arithmetic, bf16 or FP32 roc2), whatever the tile
  variant, z chunk, step parity or source;
* two-step blocks (gr_batch_set_time_block) give the same bits as single steps,
  for both two-step kernels, any z chunk, odd and even step counts, sources that
  end inside a block, problems at mixed parities and problems whose fields were
  rotated by their own fused passes; a problem whose U1 and U2 halos differ makes
  the call fall back to single steps;
* graph replays of the step loop equal eager launches on every replay;
* one cooperative launch per call (mode 2, grid barrier between steps) equals a
  launch per step, for single steps and two-step blocks;
* receiver planes recorded by the batch equal a single shot's traces
  (gr_shot_forward on the same model and source) bit for bit, in every launch mode;
* against the FP64 CPU oracle, with the roc2 the batch actually holds, the FP32
  batch stays within a relative L2 error of 1e-5;
* the layered velocity model has the stated layer speeds and bf16 rounding.
"""
import ctypes

import numpy as np
import pytest

import oracle as orc
from girih_amd import gpu, physics

pytestmark = pytest.mark.gpu

SIZES = [(70, 40, 24), (64, 32, 16), (40, 72, 30), (130, 9, 11)]
DT_OVER_H = 0.4 / 4725.0  # CFL 0.4 at the fastest perturbed speed


def _model(seed, round_bf16=False, **kw):
    return gpu.vmodel(seed, DT_OVER_H, round_bf16=round_bf16, **kw)


def _init(batch, seed, src=None):
    for i, p in enumerate(batch.problems):
        p.init_fields(1, seed + i)
        p.init_roc2_model(_model(seed + 100 + i))
        if src is not None:
            nx, ny, nz = batch.sizes[i]
            p.set_source(4 + nx // 2, 4 + ny // 2, 4 + nz // 2, src)


def _alone(size, flags, seed, i, napply, src=None, pre=0):
    nx, ny, nz = size
    with gpu.Stepper(0, nx, ny, nz, dtype=gpu.GR_F32, flags=flags, coef=physics.fd8_coef()) as st:
        st.init_fields(1, seed + i)
        st.init_roc2_model(_model(seed + 100 + i))
        if src is not None:
            st.set_source(4 + nx // 2, 4 + ny // 2, 4 + nz // 2, src)
        if pre:
            st.step(pre)
        st.step(napply)
        return st.download()


@pytest.mark.parametrize("bf16", [True, False])
@pytest.mark.parametrize("variant,zchunk", [(0, 0), (1, 8), (2, 0), (3, 16), (4, 7), (5, 9), (6, 0)])
def test_batch_equals_single_problem_steps(bf16, variant, zchunk):
    flags = gpu.GR_FAST | (gpu.GR_COEF_BF16 if bf16 else 0)
    src = np.linspace(0.5, -0.2, 7)
    print("batch", bf16, variant, zchunk, flags, flush=True)
    with gpu.Batch(SIZES, flags=flags, coef=physics.fd8_coef()) as b:
        _init(b, 11, src)
        b.set_config(variant, zchunk)
        b.step(5)
        b.step(4)
        st = b.stats()
        assert st["variant"] == variant and st["zchunk"] == zchunk and st["last_launches"] == 4
        got = [p.download() for p in b.problems]
    for i, size in enumerate(SIZES):
        U1, U2 = _alone(size, flags, 11, i, 9, src)
        np.testing.assert_array_equal(got[i][0], U1, err_msg="problem %d U1" % i)
        np.testing.assert_array_equal(got[i][1], U2, err_msg="problem %d U2" % i)


def test_batch_mixed_parities_and_graph_replays():
    """Problems at different step parities advance correctly, and a replayed step
    graph equals eager launches on every replay (eager work in between)."""
    flags = gpu.GR_FAST | gpu.GR_COEF_BF16
    with gpu.Batch(SIZES, flags=flags, coef=physics.fd8_coef()) as eager, \
            gpu.Batch(SIZES, flags=flags, coef=physics.fd8_coef()) as graph:
        for b in (eager, graph):
            _init(b, 5)
            b.problems[1].step(1)  # problem 1 one application ahead (odd parity)
            b.problems[3].step(3)
        for rep in range(3):
            print("replay", rep, flush=True)
            eager.step(6)
            graph.step(6, use_graph=True)
            for i in range(len(SIZES)):
                E = eager.problems[i].download()
                G = graph.problems[i].download()
                np.testing.assert_array_equal(E[0], G[0])
                np.testing.assert_array_equal(E[1], G[1])
        for i, pre in ((1, 1), (3, 3)):
            U1, U2 = _alone(SIZES[i], flags, 5, i, 18, pre=pre)
            E = eager.problems[i].download()
            np.testing.assert_array_equal(E[0], U1)
            np.testing.assert_array_equal(E[1], U2)


@pytest.mark.parametrize("bf16", [True, False])
@pytest.mark.parametrize("block_variant,zchunk", [(0, 0), (1, 0), (0, 5), (1, 8), (0, 1)])
def test_batch_two_step_blocks_equal_single_problem_steps(bf16, block_variant, zchunk):
    flags = gpu.GR_FAST | (gpu.GR_COEF_BF16 if bf16 else 0)
    src = np.linspace(0.5, -0.2, 8)  # the third call's first block holds the last value and a zero
    print("batch two-step", bf16, block_variant, zchunk, flush=True)
    with gpu.Batch(SIZES, flags=flags, coef=physics.fd8_coef()) as b:
        _init(b, 13, src)
        b.set_time_block(2, block_variant)
        b.set_config(-1, zchunk)
        for n, launches in ((3, 2), (4, 2), (5, 3)):
            b.step(n)
            st = b.stats()
            assert st["last_launches"] == launches, (n, st)
            assert st["time_block"] == 2 and st["block_variant"] == block_variant
            assert st["zchunk"] == zchunk and (st["tile_x"], st["tile_y"]) == (64, 16)
        got = [p.download() for p in b.problems]
    for i, size in enumerate(SIZES):
        U1, U2 = _alone(size, flags, 13, i, 12, src)
        np.testing.assert_array_equal(got[i][0], U1, err_msg="problem %d U1" % i)
        np.testing.assert_array_equal(got[i][1], U2, err_msg="problem %d U2" % i)


@pytest.mark.parametrize("bf16", [True, False])
def test_batch_two_step_mixed_states_and_graph_replays(bf16):
    """Problems one and three steps ahead, one whose own fused passes rotated its
    buffers, odd step counts (each call ends with a single step and flips the
    parity), replayed from a graph: every replay equals eager launches, and the
    end state equals each problem stepped alone."""
    flags = gpu.GR_FAST | (gpu.GR_COEF_BF16 if bf16 else 0)
    pre = {1: 1, 2: 2, 3: 3}
    with gpu.Batch(SIZES, flags=flags, coef=physics.fd8_coef()) as eager, \
            gpu.Batch(SIZES, flags=flags, coef=physics.fd8_coef()) as graph:
        for b in (eager, graph):
            _init(b, 9)
            b.problems[1].step(1)
            b.problems[2].set_config(t_block=2)  # two-step pass on its own handle: spares become the fields
            b.problems[2].step(2)
            assert b.problems[2].stats()["last_launches"] == 1
            b.problems[3].step(3)
            b.set_time_block(2, 0)
        for rep in range(3):
            print("two-step replay", rep, flush=True)
            eager.step(7)
            graph.step(7, use_graph=True)
            assert graph.stats()["last_launches"] == 4
            for i in range(len(SIZES)):
                E = eager.problems[i].download()
                G = graph.problems[i].download()
                np.testing.assert_array_equal(E[0], G[0])
                np.testing.assert_array_equal(E[1], G[1])
        for i in range(len(SIZES)):
            U1, U2 = _alone(SIZES[i], flags, 9, i, 21, pre=pre.get(i, 0))
            E = eager.problems[i].download()
            np.testing.assert_array_equal(E[0], U1, err_msg="problem %d U1" % i)
            np.testing.assert_array_equal(E[1], U2, err_msg="problem %d U2" % i)


def test_batch_two_step_falls_back_on_unequal_halos():
    flags = gpu.GR_FAST
    outs = []
    for tblock in (2, 1):
        with gpu.Batch(SIZES, flags=flags, coef=physics.fd8_coef()) as b:
            _init(b, 17)
            U1, U2 = b.problems[2].download()
            U1[1, 2, 3] = 0.25  # a halo cell of U1 only
            b.problems[2].upload(U1, U2)
            b.set_time_block(tblock)
            b.step(4)
            assert b.stats()["last_launches"] == 4
            outs.append([p.download() for p in b.problems])
    for (a1, a2), (b1, b2) in zip(*outs):
        np.testing.assert_array_equal(a1, b1)
        np.testing.assert_array_equal(a2, b2)


@pytest.mark.parametrize("bf16", [True, False])
@pytest.mark.parametrize("tblock", [1, 2])
def test_batch_one_launch_equals_launch_per_step(bf16, tblock):
    """Mode 2 runs a call's steps (or two-step blocks) in one cooperative launch;
    with problems at mixed parities and a source, it gives the bits of a launch per
    step, call after call, and each problem's bits stepped alone."""
    flags = gpu.GR_FAST | (gpu.GR_COEF_BF16 if bf16 else 0)
    src = np.linspace(0.3, -0.4, 9)
    print("one launch", bf16, tblock, flush=True)
    with gpu.Batch(SIZES, flags=flags, coef=physics.fd8_coef()) as per, \
            gpu.Batch(SIZES, flags=flags, coef=physics.fd8_coef()) as one:
        for b in (per, one):
            _init(b, 23, src)
            b.problems[1].step(1)
            b.set_config(1, 8)
            b.set_time_block(tblock, 1 if tblock == 2 else -1)
        for n in (5, 4, 3):
            per.step(n)
            one.step(n, mode=2)
            st = one.stats()
            assert st["last_mode"] == 2
            expect = 1 if tblock == 1 else 1 + (n & 1)
            assert st["last_launches"] == expect, (n, st)
            for i in range(len(SIZES)):
                P = per.problems[i].download()
                O = one.problems[i].download()
                np.testing.assert_array_equal(P[0], O[0], err_msg="problem %d U1 after %d" % (i, n))
                np.testing.assert_array_equal(P[1], O[1], err_msg="problem %d U2 after %d" % (i, n))
        got = [p.download() for p in one.problems]
    for i in (0, 1):
        U1, U2 = _alone(SIZES[i], flags, 23, i, 12, src, pre=1 if i == 1 else 0)
        np.testing.assert_array_equal(got[i][0], U1)
        np.testing.assert_array_equal(got[i][1], U2)


def test_batch_one_launch_autotune_keeps_fields():
    flags = gpu.GR_FAST | gpu.GR_COEF_BF16
    with gpu.Batch(SIZES[:3], flags=flags, coef=physics.fd8_coef()) as b:
        _init(b, 29)
        before = [p.download() for p in b.problems]
        b.autotune(2, mode=2)
        assert b.stats()["last_mode"] == 2
        after = [p.download() for p in b.problems]
        for x, y in zip(before, after):
            np.testing.assert_array_equal(x[0], y[0])
            np.testing.assert_array_equal(x[1], y[1])
        b.step(3, mode=2)
        got = [p.download() for p in b.problems]
    for i, size in enumerate(SIZES[:3]):
        U1, U2 = _alone(size, flags, 29, i, 3)
        np.testing.assert_array_equal(got[i][0], U1)
        np.testing.assert_array_equal(got[i][1], U2)


@pytest.mark.parametrize("mode", [0, 1, 2])
def test_batch_receivers_equal_shot_traces(mode):
    """Shots in a batch: each problem's receiver plane, recorded after every
    application over two calls, equals the traces of gr_shot_forward for the same
    squared slowness and source; a problem without receivers and a receiver that
    stops at max_steps leave the rest untouched."""
    import torch
    from girih_amd import adjoint
    sizes = [(40, 24, 20), (32, 40, 28), (56, 16, 24)]
    nt, rec_k = 12, 4 + 6
    src = np.linspace(0.8, -0.6, nt)
    maxs = [nt, 0, 8]
    print("batch receivers, mode", mode, flush=True)
    ms, traces = [], []
    with gpu.Batch(sizes, flags=gpu.GR_FAST, coef=physics.fd8_coef()) as b:
        for i, (nx, ny, nz) in enumerate(sizes):
            p = b.problems[i]
            p.init_fields(2)
            rng = np.random.default_rng(40 + i)
            m = torch.tensor((1.0 / rng.uniform(1500.0, 4500.0, (nz, ny, nx)) ** 2).astype(np.float32), device="cuda")
            gpu.check(gpu.lib().gr_set_roc2_from_slowness2(p._h, ctypes.c_void_p(m.data_ptr()), DT_OVER_H))
            p.set_source(4 + nx // 2, 4 + ny // 2, 4 + nz // 2, src)
            t = torch.full((nt, ny, nx), -7.0, dtype=torch.float32, device="cuda")
            if maxs[i]:
                b.set_receivers(i, rec_k, t, maxs[i])
            ms.append(m)
            traces.append(t)
        b.set_time_block(2)  # receivers keep the batch on single steps
        b.step(5, mode=mode)
        b.step(7, mode=mode)
        assert b.stats()["last_launches"] == (1 if mode == 2 else 7)
        got = [t.cpu().numpy() for t in traces]
    for i, (nx, ny, nz) in enumerate(sizes):
        if not maxs[i]:
            assert np.all(got[i] == -7.0)
            continue
        shot = adjoint.Shot(nx, ny, nz, nt, DT_OVER_H, (4 + nx // 2, 4 + ny // 2, 4 + nz // 2), src, rec_k, ckpt=nt,
                            tune=False)
        try:
            ref = shot.forward(ms[i]).cpu().numpy()
        finally:
            shot.close()
        k = maxs[i]
        np.testing.assert_array_equal(got[i][:k], ref[:k], err_msg="problem %d" % i)
        assert np.all(got[i][k:] == -7.0)


def test_batch_time_block_errors():
    with gpu.Batch(SIZES[:2], flags=gpu.GR_FAST, coef=physics.fd8_coef()) as b:
        for bad in (0, 3, 4):
            with pytest.raises(gpu.GirihError):
                b.set_time_block(bad)
        with pytest.raises(gpu.GirihError):
            b.set_time_block(2, 2)
        b.set_time_block(1)
        assert b.stats()["time_block"] == 1
        for bad in (-1, 3):
            with pytest.raises(gpu.GirihError):
                b.step(1, mode=bad)


def test_batch_autotune_keeps_fields():
    flags = gpu.GR_FAST | gpu.GR_COEF_BF16
    with gpu.Batch(SIZES[:3], flags=flags, coef=physics.fd8_coef()) as b:
        _init(b, 21)
        before = [p.download() for p in b.problems]
        v, z = b.autotune(2)
        st = b.stats()
        assert 0 <= v < st["nvariants"] and st["time_block"] in (1, 2)
        print("autotune", v, z, st, flush=True)
        after = [p.download() for p in b.problems]
        for x, y in zip(before, after):
            np.testing.assert_array_equal(x[0], y[0])
            np.testing.assert_array_equal(x[1], y[1])
        b.step(3)
        got = [p.download() for p in b.problems]
    for i, size in enumerate(SIZES[:3]):
        U1, U2 = _alone(size, flags, 21, i, 3)
        np.testing.assert_array_equal(got[i][0], U1)
        np.testing.assert_array_equal(got[i][1], U2)


def test_batch_fp32_bf16_vs_fp64_oracle():
    """Zero fields, a Ricker source, the layered model: every FP32 problem against
    the FP64 oracle run on the roc2 values the batch holds (bf16-exact)."""
    sizes = [(48, 40, 36), (32, 56, 40)]
    nt = 40
    dt = 0.4 * 10 / 4725.0
    src = (0.16 * physics.ricker(30.0, dt, nt + 2, t0=0.012)).astype(np.float32).astype(np.float64)
    coef = np.zeros(11)
    coef[:5] = physics.fd8_coef()
    with gpu.Batch(sizes, flags=gpu.GR_FAST | gpu.GR_COEF_BF16, coef=physics.fd8_coef()) as b:
        for i, p in enumerate(b.problems):
            p.init_fields(2)
            p.init_roc2_model(_model(70 + i))
            nx, ny, nz = sizes[i]
            p.set_source(4 + nx // 2, 4 + ny // 2, 4 + nz // 2, src)
        b.step(nt)
        outs = [(p.download(), p.download_roc2()) for p in b.problems]
    for i, ((U1, U2), r2) in enumerate(outs):
        nx, ny, nz = sizes[i]
        shp = orc.shape_of(0, nx, ny, nz)
        u = np.zeros(shp)
        v = np.zeros(shp)
        roc2 = np.ascontiguousarray(r2[:, :, :shp[2]].astype(np.float64))
        pt = (4 + nx // 2, 4 + ny // 2, 4 + nz // 2)
        orc.run(0, coef, u, v, roc2, nt=nt, src_idx=pt, src=src)
        for got, ref in ((U1, u), (U2, v)):
            g = got[:, :, :shp[2]].astype(np.float64)
            rel = np.linalg.norm(g - ref) / np.linalg.norm(ref)
            print("problem", i, "rel L2", rel, flush=True)
            assert np.linalg.norm(ref) > 0
            assert rel < 1e-5


def test_layered_model_layers_and_bf16():
    nx, ny, nz = 24, 20, 64
    with gpu.Stepper(0, nx, ny, nz, dtype=gpu.GR_F64, coef=physics.fd8_coef()) as st:
        st.init_roc2_model(_model(3, nblobs=0, pert=0.0))
        r2 = st.download_roc2()
        st.init_roc2_model(_model(3, round_bf16=True))
        rb = st.download_roc2()
    inner = r2[4:-4, 4:-4, 4:4 + nx]
    v = np.sqrt(inner) / DT_OVER_H
    speeds = np.unique(np.round(v, 6))
    expect = 1500.0 + 3000.0 * np.arange(8) / 7
    assert set(np.round(speeds, 3)) <= set(np.round(expect, 3))
    assert len(speeds) >= 2
    for k in range(nz):  # layers are horizontal and speeds rise with depth
        assert np.all(v[k] == v[k, 0, 0])
        if k:
            assert v[k, 0, 0] >= v[k - 1, 0, 0]
    # bf16 values: the low 16 bits of the FP32 pattern are zero
    bits = rb.astype(np.float32).view(np.uint32)
    assert np.all(bits & 0xFFFF == 0)
    assert np.all(rb[4:-4, 4:-4, 4:4 + nx] > 0)
    with gpu.Stepper(0, nx, ny, nz, dtype=gpu.GR_F32, flags=gpu.GR_COEF_BF16 | gpu.GR_FAST,
                     coef=physics.fd8_coef()) as st:
        st.init_roc2_model(_model(3))
        rbf = st.download_roc2()
    np.testing.assert_array_equal(rbf[:, :, :nx + 8].astype(np.float64), rb[:, :, :nx + 8])
