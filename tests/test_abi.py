"""C-ABI checks that need no GPU: the library loads, exports every entry point
declared in include/girih_gpu.h, and rejects bad arguments with a status code
(the reference exits the process instead, data_structures.h:299-314)."""
import ctypes
import os
import re

import pytest

from girih_amd import gpu

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    hdr = open(os.path.join(REPO, "include", "girih_gpu.h")).read()
    return sorted(set(re.findall(r"^(?:int|const char \*)\s*(gr_\w+)\(", hdr, re.M)))


def test_header_declarations_match_binding():
    decl = _declared()
    assert set(decl) == set(gpu.EXPORTS), decl


def test_library_exports_every_declared_symbol():
    L = gpu.lib()
    for name in _declared():
        assert hasattr(L, name), name


def test_mwd_round_nt_host_only():
    # diamond_utils.c:1042-1056
    assert gpu.mwd_round_nt(100, 3) == 106
    assert gpu.mwd_round_nt(9, 1) == 10
    assert gpu.mwd_round_nt(12, 3) == 18
    assert gpu.mwd_round_nt(1000, 1) == 1002


@pytest.mark.parametrize("kw,code", [
    (dict(kernel=7, nx=8, ny=8, nz=8), gpu.GR_EINVAL),
    (dict(kernel=-1, nx=8, ny=8, nz=8), gpu.GR_EINVAL),
    (dict(kernel=1, nx=0, ny=8, nz=8), gpu.GR_EINVAL),
    (dict(kernel=1, nx=8, ny=8, nz=8, dtype=5), gpu.GR_EINVAL),
    (dict(kernel=0, nx=8, ny=8, nz=8, dtype=gpu.GR_F64, flags=gpu.GR_FAST), gpu.GR_EINVAL),
    (dict(kernel=5, nx=8, ny=8, nz=8, dtype=gpu.GR_F32, flags=gpu.GR_FAST), gpu.GR_EUNSUPPORTED),
    (dict(kernel=1, nx=8, ny=8, nz=8, flags=gpu.GR_ROC2_SCALAR), gpu.GR_EINVAL),
    (dict(kernel=1, nx=8, ny=8, nz=8, dtype=gpu.GR_F64, flags=gpu.GR_COEF_BF16), gpu.GR_EINVAL),
])
def test_create_rejects_bad_descriptors(kw, code):
    with pytest.raises(gpu.GirihError) as ei:
        gpu.Stepper(**kw)
    assert ei.value.code == code
    assert len(str(ei.value)) > 20


def test_null_arguments_return_status():
    L = gpu.lib()
    assert L.gr_create(None, None) == gpu.GR_EINVAL
    assert L.gr_step(None, 1) == gpu.GR_EINVAL
    assert L.gr_upload(None, None, None, None, None, None) == gpu.GR_EINVAL
    assert L.gr_get_stats(None, None) == gpu.GR_EINVAL
    assert L.gr_destroy(None) == gpu.GR_OK
    assert b"null" in L.gr_last_error()


def test_sweep_dev_rejects_bad_box_without_touching_memory():
    L = gpu.lib()
    shape = (ctypes.c_int * 3)(16, 16, 16)
    c = (ctypes.c_double * 5)(0.4, 0.1, 0, 0, 0)
    bogus = ctypes.c_void_p(0x1000)
    # box touching the face for a radius-1 kernel
    assert L.gr_sweep_dev(1, gpu.GR_F64, shape, 0, 1, 1, 15, 15, 15, c, bogus, bogus, None) == gpu.GR_EINVAL
    # radius-4 kernel with a box too close to the far face
    assert L.gr_sweep_dev(0, gpu.GR_F64, shape, 4, 4, 4, 13, 12, 12, c, bogus, bogus, bogus) == gpu.GR_EINVAL
    # kernel 0 without roc2
    assert L.gr_sweep_dev(0, gpu.GR_F64, shape, 4, 4, 4, 12, 12, 12, c, bogus, bogus, None) == gpu.GR_EINVAL
    assert L.gr_sweep_dev(9, gpu.GR_F64, shape, 4, 4, 4, 12, 12, 12, c, bogus, bogus, None) == gpu.GR_EINVAL
