"""tm_op_apply_dev_noalloc against tm_op_apply_dev through the C ABI: the
no-allocation entry point runs the same kernels on its file-scope scratch
array, so its results equal the allocating entry point's bit for bit (both
with the default tiles: operators made with TM_FLAG_NO_AUTOTUNE).  Two calls
in a row, the second with a smaller input; every operator path that uses
scratch (the FP32 input scaling, the Kronecker mid buffer, the sweep aux
array, the reduced operator's workspace and permutation buffers); and an
apply that needs more than the array holds is refused before any launch."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def lt():
    import latttmat_amd as lt
    return lt


def _both(lt, op, x, dtype, compute, transpose=False, nvec=1):
    import torch
    tt = torch.float64 if dtype == lt.TM_F64 else torch.float32
    n = op.cols()
    xd = torch.from_numpy(x).to(tt).to("cuda")
    ya = torch.zeros(n * nvec, dtype=tt, device="cuda")
    yb = torch.full((n * nvec,), float("nan"), dtype=tt, device="cuda")
    torch.cuda.synchronize()
    op.apply_dev(xd.data_ptr(), ya.data_ptr(), n, nvec, dtype, compute, transpose)
    torch.cuda.synchronize()
    op.apply_dev_noalloc(xd.data_ptr(), yb.data_ptr(), n, nvec, dtype, compute, transpose)
    torch.cuda.synchronize()
    a, b = ya.cpu().numpy(), yb.cpu().numpy()
    print(f"  {op.kind()} n={n} nvec={nvec} dtype={dtype} compute={compute} tr={transpose} "
          f"max|a|={np.max(np.abs(a)):.3e} equal={np.array_equal(a, b)}", flush=True)
    return a, b


def test_noalloc_equals_apply_dev(lt):
    F = lt.TM_FLAG_NO_AUTOTUNE
    rng = np.random.default_rng(3)
    # first call: the largest input (full ANNNI, FP32 compute on FP64 vectors:
    # scaling partials, scales and the FP32 Kronecker mid buffer), two vectors
    big = lt.TransferMatrix("A", 2, 1, 18, J=0.5, kappa=0.6, flags=F)
    x = rng.uniform(-1, 1, 2 * big.cols())
    a, b = _both(lt, big, x, lt.TM_F64, lt.TM_F32, nvec=2)
    assert np.array_equal(a, b)
    # second call, smaller: the reduced operator (workspace and permutation
    # buffers), FP64
    red = lt.TransferMatrix("A", 2, 5, 16, J=0.4, kappa=0.5, neigs=4, flags=F)
    a, b = _both(lt, red, rng.uniform(-1, 1, red.cols()), lt.TM_F64, lt.TM_F64)
    assert np.array_equal(a, b)
    # the sweep circuit (aux array) with FP32 vectors, transposed
    zbt = lt.TransferMatrix("B", 2, 3, 12, J=0.5, kappa=0.5, kappa2=0.3, flags=F)
    a, b = _both(lt, zbt, rng.uniform(-1, 1, zbt.cols()), lt.TM_F32, lt.TM_F32, transpose=True)
    assert np.array_equal(a, b)
    # and the first operator again after the smaller calls
    a, b = _both(lt, big, x, lt.TM_F64, lt.TM_F32, nvec=2)
    assert np.array_equal(a, b)


def test_noalloc_refuses_beyond_scratch(lt):
    import torch
    op = lt.TransferMatrix("A", 2, 1, 24, J=0.5, kappa=0.6)  # FP32 mid buffer: 64 MiB
    n = op.cols()
    xd = torch.zeros(n, dtype=torch.float64, device="cuda")
    yd = torch.zeros(n, dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    with pytest.raises(lt.TMError) as ei:
        op.apply_dev_noalloc(xd.data_ptr(), yd.data_ptr(), n, 1, lt.TM_F64, lt.TM_F32)
    assert ei.value.status == lt.TM_EINVAL and "scratch" in str(ei.value)
