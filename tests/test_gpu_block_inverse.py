"""dgx_block_inverse (FP32 Gauss-Jordan with partial pivoting, one workgroup per block) on the GPU.

- random blocks (shift I - B) of every tile shape from n = 1 to the limit 320, including the 32-column
  tile edges (32, 33, 64, 65), against numpy's FP64 inverse: forward error within 2e-6 cond(A) (FP32
  Gauss-Jordan with partial pivoting; the measured errors are far below), residual |A X - I| <= 1e-4;
- a permuted identity and a permuted diagonal (every pivot off the diagonal, exact in FP32): the
  inverse is reproduced exactly, which pins the pivot bookkeeping and the final unscrambling;
- a singular block is reported (and its neighbours in the batch are unaffected); n > 320 and bad sizes
  are refused; zero blocks is a no-op;
- dgx_block_apply (FP32 column-major blocks times FP64 vectors) equals numpy's products of the same FP32
  blocks to FP32 accumulation accuracy, for A and A^T, across the tile edges;
- dgx_block_inverse_f32 from FP32 blocks, in place, equals the inverse of the same FP32-rounded blocks.
"""
import ctypes

import numpy as np
import pytest

from dpgsolver_amd import _ffi, implicit

pytestmark = pytest.mark.gpu


def _require_gpu():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")


def _run(B, shift):
    import torch
    Bt = torch.from_numpy(np.ascontiguousarray(B, dtype=np.float64)).cuda()
    st = torch.from_numpy(np.ascontiguousarray(shift, dtype=np.float64)).cuda()
    return implicit.block_inverse(Bt, st).cpu().numpy().astype(np.float64)


@pytest.mark.parametrize("n", [1, 5, 20, 32, 33, 64, 65, 100, 160, 250, 320])
def test_random_blocks_against_numpy(n):
    _require_gpu()
    rng = np.random.default_rng(n)
    nb = 37
    B = rng.standard_normal((nb, n, n)) / np.sqrt(n)
    shift = 1.5 + rng.random(nb)
    print(f"block inverse n={n} nb={nb}", flush=True)
    X = _run(B, shift)
    for b in range(nb):
        A = shift[b] * np.eye(n) - B[b]
        ref = np.linalg.inv(A)
        cond = np.linalg.cond(A)
        err = np.linalg.norm(X[b] - ref) / np.linalg.norm(ref)
        assert err <= 2e-6 * cond, (b, err, cond)
        assert np.abs(A @ X[b] - np.eye(n)).max() <= 1e-4


@pytest.mark.parametrize("n", [7, 33, 320])
def test_permutation_blocks_exact(n):
    """Zero diagonal everywhere: every pivot comes from another row; the inverse of P D is D^-1 P^T,
    with powers of two on D so that FP32 represents everything exactly."""
    _require_gpu()
    rng = np.random.default_rng(100 + n)
    nb = 5
    B = np.zeros((nb, n, n))
    refs = []
    for b in range(nb):
        perm = rng.permutation(n)
        while n > 1 and np.any(perm == np.arange(n)):
            perm = rng.permutation(n)
        d = 2.0 ** rng.integers(-3, 4, size=n)
        A = np.zeros((n, n))
        A[np.arange(n), perm] = d
        B[b] = -A                          # shift 0: A = -B
        refs.append(np.linalg.inv(A))
    X = _run(B, np.zeros(nb))
    for b in range(nb):
        assert np.array_equal(X[b], refs[b]), b


def test_singular_block_reported_and_neighbours_intact():
    _require_gpu()
    import torch
    rng = np.random.default_rng(7)
    n, nb = 40, 4
    B = rng.standard_normal((nb, n, n)) / np.sqrt(n)
    shift = np.full(nb, 2.0)
    B[2, 5, :] = 0.0
    B[2, 5, 5] = 2.0                       # row 5 of shift I - B is zero
    Bt = torch.from_numpy(B).cuda()
    st = torch.from_numpy(shift).cuda()
    out = torch.empty((nb, n, n), dtype=torch.float32, device="cuda")
    sing = torch.full((nb,), -1, dtype=torch.int32, device="cuda")
    lib = _ffi.load()
    status = lib.dgx_block_inverse(nb, n, ctypes.c_void_p(Bt.data_ptr()), ctypes.c_void_p(st.data_ptr()),
                                   ctypes.c_void_p(out.data_ptr()), ctypes.c_void_p(sing.data_ptr()), None)
    assert status == _ffi.DGX_OK
    torch.cuda.synchronize()
    s = sing.cpu().numpy()
    assert s[2] > 0 and s[0] == 0 and s[1] == 0 and s[3] == 0
    X = out.cpu().numpy().astype(np.float64)
    for b in (0, 1, 3):
        ref = np.linalg.inv(shift[b] * np.eye(n) - B[b])
        assert np.linalg.norm(X[b] - ref) / np.linalg.norm(ref) <= 1e-5
    with pytest.raises(FloatingPointError):
        implicit.block_inverse(Bt, st)


def test_nonfinite_and_tiny_pivot_blocks_flagged():
    """A block with one NaN entry and a block whose only pivot is an FP32 denormal (its reciprocal
    overflows) are flagged n + 1; the other blocks of the batch are unaffected."""
    _require_gpu()
    import torch
    print("case: non-finite / tiny-pivot blocks", flush=True)
    rng = np.random.default_rng(9)
    n, nb = 24, 4
    B = rng.standard_normal((nb, n, n)) / np.sqrt(n)
    shift = np.full(nb, 2.0)
    B[1, 3, 7] = np.nan
    B[2] = 0.0
    shift[2] = 1e-40                       # shift I - B = 1e-40 I: FP32 denormal pivots
    Bt = torch.from_numpy(B).cuda()
    st = torch.from_numpy(shift).cuda()
    out = torch.empty((nb, n, n), dtype=torch.float32, device="cuda")
    sing = torch.full((nb,), -1, dtype=torch.int32, device="cuda")
    lib = _ffi.load()
    status = lib.dgx_block_inverse(nb, n, ctypes.c_void_p(Bt.data_ptr()), ctypes.c_void_p(st.data_ptr()),
                                   ctypes.c_void_p(out.data_ptr()), ctypes.c_void_p(sing.data_ptr()), None)
    assert status == _ffi.DGX_OK
    torch.cuda.synchronize()
    s = sing.cpu().numpy()
    assert s[1] != 0 and s[2] != 0 and s[0] == 0 and s[3] == 0, s
    X = out.cpu().numpy().astype(np.float64)
    for b in (0, 3):
        ref = np.linalg.inv(shift[b] * np.eye(n) - B[b])
        assert np.linalg.norm(X[b] - ref) / np.linalg.norm(ref) <= 1e-5


def test_argument_checks():
    _require_gpu()
    import torch
    lib = _ffi.load()
    buf = torch.zeros(4, dtype=torch.float64, device="cuda")
    p = ctypes.c_void_p(buf.data_ptr())
    assert lib.dgx_block_inverse(1, 321, p, p, p, p, None) == _ffi.DGX_ERR_INVALID_ARG
    assert lib.dgx_block_inverse(1, 0, p, p, p, p, None) == _ffi.DGX_ERR_INVALID_ARG
    assert lib.dgx_block_inverse(-1, 4, p, p, p, p, None) == _ffi.DGX_ERR_INVALID_ARG
    assert lib.dgx_block_inverse(1, 4, None, p, p, p, None) == _ffi.DGX_ERR_INVALID_ARG
    assert lib.dgx_block_inverse(0, 4, None, None, None, None, None) == _ffi.DGX_OK


@pytest.mark.parametrize("n", [1, 7, 64, 65, 255, 256, 257, 320])
def test_block_apply_against_numpy(n):
    _require_gpu()
    import torch
    rng = np.random.default_rng(500 + n)
    nb = 23
    A = rng.standard_normal((nb, n, n)).astype(np.float32)
    x = rng.standard_normal(nb * n)
    At = torch.from_numpy(A).cuda()
    xt = torch.from_numpy(x).cuda()
    y_colT = implicit.block_apply(At, xt).cpu().numpy().reshape(nb, n)          # column-major of A^T -> A^T x
    y_col = implicit.block_apply(At.transpose(1, 2).contiguous(), xt).cpu().numpy().reshape(nb, n)
    xb = x.reshape(nb, n)
    A64 = A.astype(np.float64)
    ref = np.einsum("bij,bj->bi", A64, xb)
    refT = np.einsum("bji,bj->bi", A64, xb)
    scale = np.abs(A64).max() * np.abs(xb).max() * n
    assert np.abs(y_col - ref).max() <= 1e-6 * scale
    assert np.abs(y_colT - refT).max() <= 1e-6 * scale


@pytest.mark.parametrize("n", [9, 64, 320])
def test_block_inverse_f32_in_place(n):
    _require_gpu()
    import torch
    rng = np.random.default_rng(900 + n)
    nb = 19
    B = (rng.standard_normal((nb, n, n)) / np.sqrt(n)).astype(np.float32)
    shift = 1.5 + rng.random(nb)
    Bt = torch.from_numpy(B).cuda()
    st = torch.from_numpy(shift).cuda()
    out = implicit.block_inverse(Bt, st, out=Bt)
    assert out.data_ptr() == Bt.data_ptr()
    X = out.cpu().numpy().astype(np.float64)
    for b in range(nb):
        A = shift[b] * np.eye(n) - B[b].astype(np.float64)
        ref = np.linalg.inv(A)
        assert np.linalg.norm(X[b] - ref) / np.linalg.norm(ref) <= 2e-6 * np.linalg.cond(A)


def test_launch_split_over_2p20_blocks():
    """More blocks than one launch takes (2^20 + 5 workgroups): every block of every launch is
    inverted and applied (1 x 1 blocks keep it cheap), including those past the first launch."""
    _require_gpu()
    import torch
    nb = (1 << 20) + 5
    print(f"launch split nb={nb}", flush=True)
    g = torch.Generator(device="cuda").manual_seed(5)
    B = torch.rand((nb, 1, 1), dtype=torch.float64, device="cuda", generator=g)
    shift = 2.0 + torch.rand(nb, dtype=torch.float64, device="cuda", generator=g)
    X = implicit.block_inverse(B, shift)
    ref = (1.0 / (shift - B[:, 0, 0])).float()
    assert torch.allclose(X[:, 0, 0], ref, rtol=2e-7, atol=0)
    x = torch.rand(nb, dtype=torch.float64, device="cuda", generator=g)
    y = implicit.block_apply(X, x)
    assert torch.allclose(y, X[:, 0, 0].double() * x.float().double(), rtol=1e-6, atol=0)
    # the last blocks only (the second launch of a selection)
    sel = torch.arange(nb - 7, nb, dtype=torch.int32, device="cuda")
    sel_all = torch.cat([torch.arange(0, 1 << 20, dtype=torch.int32, device="cuda"), sel])
    y2 = torch.zeros_like(x)
    implicit.block_apply_sel(X, sel_all, x, y2)
    torch.cuda.synchronize()
    assert torch.equal(y2[nb - 7:], y[nb - 7:])
