"""gr_compare_fields_static, the comparison entry point that allocates nothing (its
partial sums live in a static device array): through the C ABI it returns exactly
what gr_compare_fields returns, on two calls in a row, the second on a smaller grid
(the array keeps the first call's contents, which must not leak into the second)."""
import pytest

from girih_amd import gpu

pytestmark = pytest.mark.gpu


def _pair(kernel, nx, ny, nz, seed):
    a = gpu.Stepper(kernel, nx, ny, nz, dtype=gpu.GR_F32, coef=[0.4, 0.1] if kernel == 1 else None)
    b = gpu.Stepper(kernel, nx, ny, nz, dtype=gpu.GR_F64, coef=[0.4, 0.1] if kernel == 1 else None)
    for st in (a, b):
        if gpu.NUM_COEF_ARRAYS[kernel]:
            st.init_coef_uniform(seed, 0.3, 0.5, 0.0, 0.1)
        st.init_fields(1, seed)
        st.step(7)
    return a, b


def test_static_compare_equals_allocating_compare_twice():
    cases = [(5, 150, 70, 40, 11), (1, 33, 9, 7, 12)]  # the second call is on a smaller grid
    for kernel, nx, ny, nz, seed in cases:
        print("no-alloc compare", kernel, (nx, ny, nz), flush=True)
        a, b = _pair(kernel, nx, ny, nz, seed)
        for fa, fb in ((0, 0), (1, 1), (0, 1)):
            want = a.compare(fa, b, fb)
            got = a.compare_static(fa, b, fb)
            assert want["sq_ref"] > 0 and want["sq_diff"] > 0
            for k in ("sq_diff", "sq_ref", "max_diff", "max_ref"):
                assert got[k] == want[k], (kernel, fa, fb, k, got[k], want[k])
        # one handle against itself (U1 vs U2), FP64 against FP64
        assert b.compare_static(0, b, 1) == b.compare(0, b, 1)
        a.close()
        b.close()


def test_static_compare_rejects_bad_arguments():
    with gpu.Stepper(1, 8, 8, 8) as a, gpu.Stepper(1, 9, 8, 8) as b:
        with pytest.raises(gpu.GirihError) as ei:
            a.compare_static(0, b, 0)
        assert ei.value.code == gpu.GR_EINVAL
        with pytest.raises(gpu.GirihError):
            a.compare_static(2, a, 0)
