"""The allocation-free operator entry point (stlx_apply_operator_noalloc) through the C ABI:
its output equals stlx_apply_operator's bit for bit for every operator kind, on consecutive
calls whose second is smaller (the static scratch keeps the previous call's contents), and a
call whose staging exceeds the 64 MiB scratch fails with a clear error."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


@pytest.fixture(scope="module", autouse=True)
def need_gpu():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")


def shapes(kind, N, rows):
    if kind == 0:
        return (rows, N), (rows, 4, N, 2)
    if kind == 1:
        return (rows, 4, N, 2), (rows, N)
    if kind == 2:
        return (rows, N), (rows, N // 4)
    return (rows, N // 4), (rows, N)


@pytest.mark.parametrize("precision", ["mixed", "fp64"])
def test_noalloc_operator_equals_the_allocating_one(precision):
    from stl_port import NativeEngine
    nside, J = 64, 4
    lib = None
    with NativeEngine(nside, J, precision=precision) as eng:
        lib = eng._lib.cdll
        h = eng._h.ptr
        rng = np.random.default_rng(7)
        # (kind, level, rows): each kind twice in a row, the second call smaller
        for kind, level, rows in [(0, 1, 3), (0, 2, 1), (1, 0, 2), (1, 2, 1), (2, 0, 2), (2, 2, 1),
                                  (3, 1, 3), (3, 2, 1)]:
            N = 12 * (nside >> level) ** 2
            ins, outs = shapes(kind, N, rows)
            print(f"[gpu] noalloc kind={kind} level={level} rows={rows} {precision}", flush=True)
            x = torch.from_numpy(rng.standard_normal(ins)).cuda()
            ref = torch.full(outs, np.nan, dtype=torch.float64, device="cuda")
            got = torch.full(outs, np.nan, dtype=torch.float64, device="cuda")
            torch.cuda.synchronize()
            assert lib.stlx_apply_operator(h, kind, level, rows, x.data_ptr(), ref.data_ptr()) == 0
            assert lib.stlx_apply_operator_noalloc(h, kind, level, rows, x.data_ptr(), got.data_ptr()) == 0
            torch.cuda.synchronize()
            assert torch.isfinite(got).all()
            assert torch.equal(got, ref)


def test_noalloc_operator_refuses_staging_beyond_its_scratch():
    from stl_port import NativeEngine
    with NativeEngine(256, 2, precision="fp64") as eng:
        lib = eng._lib.cdll
        x = torch.zeros((4, 12 * 256 * 256), dtype=torch.float64, device="cuda")
        y = torch.zeros((4, 4, 12 * 256 * 256, 2), dtype=torch.float64, device="cuda")
        torch.cuda.synchronize()
        assert lib.stlx_apply_operator_noalloc(eng._h.ptr, 0, 0, 4, x.data_ptr(), y.data_ptr()) == -1
        assert "static scratch" in lib.stlp_last_error().decode()
