"""Graph mode of the single-problem stepper (gr_set_graph).

The launch sequence of gr_step is captured once and replayed while the step
count, parity, configuration and source window repeat.  Every replay must equal
the eager launches bit for bit, with eager work and synchronisations between the
replays, including after a recapture (new step count, new source, new config).
"""
import numpy as np
import pytest

from girih_amd import gpu, physics

from test_gpu_parity import _consts, _pad, _pad_coef, _random_problem

pytestmark = pytest.mark.gpu


def _pair(kernel, nx, ny, nz, dtype, flags=0, roc2=0.0):
    dt = np.float64 if dtype == gpu.GR_F64 else np.float32
    u, v, r2, coef = _random_problem(kernel, nx, ny, nz, dt, 4242 + kernel)
    out = []
    for _ in range(2):
        st = gpu.Stepper(kernel, nx, ny, nz, dtype=dtype, flags=flags, roc2=roc2,
                         coef=_consts(coef, kernel) if kernel != 0 or flags & gpu.GR_ROC2_SCALAR else list(coef[:5]))
        st.upload(_pad(u), _pad(v), _pad(r2) if (r2 is not None and not flags & gpu.GR_ROC2_SCALAR) else None,
                  _pad_coef(coef, kernel))
        out.append(st)
    return out, _pad(u).shape


def _same(a, b, shape, what):
    A = a.download(shape)
    B = b.download(shape)
    np.testing.assert_array_equal(A[0], B[0], err_msg=what)
    np.testing.assert_array_equal(A[1], B[1], err_msg=what)


@pytest.mark.parametrize("case", ["k0_fast_fused_src", "k1_f64_fused", "k5_f32"])
def test_graph_replays_equal_eager(case):
    if case == "k0_fast_fused_src":
        (eager, graph), shape = _pair(0, 70, 36, 30, gpu.GR_F32, gpu.GR_FAST | gpu.GR_ROC2_SCALAR, 0.12)
        for st in (eager, graph):
            st.set_config(t_block=2)
            st.set_source(40, 20, 15, np.linspace(1.0, -1.0, 9))  # active for the first replays only
    elif case == "k1_f64_fused":
        (eager, graph), shape = _pair(1, 90, 40, 26, gpu.GR_F64)
        for st in (eager, graph):
            st.set_config(zchunk=8, t_block=4)
    else:
        (eager, graph), shape = _pair(5, 66, 35, 20, gpu.GR_F32)
    graph.set_graph(True)
    try:
        for rep, n in enumerate((5, 5, 5, 4, 4, 5)):
            print(case, "replay", rep, "steps", n, flush=True)
            eager.step(n)
            graph.step(n)
            assert graph.stats()["last_launches"] == eager.stats()["last_launches"]
            _same(eager, graph, shape, "%s replay %d" % (case, rep))
        for st in (eager, graph):  # a new source and a new configuration force a recapture
            st.set_source(30, 10, 12, [0.5, 0.25, 0.125])
            st.set_config(zchunk=4)
        for rep in range(3):
            eager.step(3)
            graph.step(3)
            _same(eager, graph, shape, "%s after recapture %d" % (case, rep))
    finally:
        eager.close()
        graph.close()
