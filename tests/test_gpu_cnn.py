"""GPU: the warm-start network's 3x3 convolution layers (matrix cores,
bfloat16 operands, FP32 accumulation) against the FP64 oracle
(oracle/cnn.py).

The oracle is evaluated on the operands rounded to bfloat16 exactly as the
kernels round them, so the only difference left is FP32 accumulation order:
relative L2 <= 1e-5.  Against the unrounded operands the bfloat16 rounding
(8-bit mantissa) shows as a relative error of a few 1e-3, printed."""
import numpy as np
import pytest

import bart_amd as ba
from oracle import cnn as R

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a, np.float32)).to("cuda")


def relerr(a, ref):
    a = np.asarray(a, np.float64)
    return float(np.linalg.norm(a - ref) / np.linalg.norm(ref))


SHAPES = [(2, 2, 48, 6, 64), (1, 48, 48, 5, 128), (2, 48, 2, 4, 64), (1, 3, 5, 3, 100), (1, 48, 48, 3, 256)]


@pytest.mark.parametrize("B,Cin,Cout,H,W", SHAPES)
def test_conv3x3_forward(B, Cin, Cout, H, W):
    rng = np.random.default_rng(Cin * 100 + Cout)
    x = rng.standard_normal((B, Cin, H, W)).astype(np.float32)
    w = (rng.standard_normal((Cout, Cin, 3, 3)) / np.sqrt(9 * Cin)).astype(np.float32)
    bias = rng.standard_normal(Cout).astype(np.float32)
    res = rng.standard_normal((B, Cout, H, W)).astype(np.float32)
    y = ba.conv3x3(dev(x), dev(w), dev(bias)).cpu().numpy()
    ref = R.conv3x3(R.bf16(x), R.bf16(w), bias)
    print(f"forward vs bf16-operand oracle {relerr(y, ref):.1e}, vs fp64 {relerr(y, R.conv3x3(x, w, bias)):.1e}")
    assert relerr(y, ref) < 1e-5
    yr = ba.conv3x3(dev(x), dev(w), dev(bias), relu=True).cpu().numpy()
    assert relerr(yr, np.maximum(ref, 0)) < 1e-5
    ys = ba.conv3x3(dev(x), dev(w), None, res=dev(res)).cpu().numpy()
    assert relerr(ys, R.conv3x3(R.bf16(x), R.bf16(w)) + res) < 1e-5


@pytest.mark.parametrize("B,Cin,Cout,H,W", SHAPES)
def test_conv3x3_backward_data(B, Cin, Cout, H, W):
    rng = np.random.default_rng(7 + Cin)
    w = (rng.standard_normal((Cout, Cin, 3, 3)) / np.sqrt(9 * Cin)).astype(np.float32)
    dy = rng.standard_normal((B, Cout, H, W)).astype(np.float32)
    dx = ba.conv3x3_bwd_data(dev(dy), dev(w)).cpu().numpy()
    ref = R.conv3x3_bwd_data(R.bf16(dy), R.bf16(w))
    assert relerr(dx, ref) < 1e-5


@pytest.mark.parametrize("B,Cin,Cout,H,W", [s for s in SHAPES if s[4] % 8 == 0] + [(3, 48, 48, 17, 64)])
def test_conv3x3_backward_weights(B, Cin, Cout, H, W):
    rng = np.random.default_rng(11 + Cout)
    x = rng.standard_normal((B, Cin, H, W)).astype(np.float32)
    dy = rng.standard_normal((B, Cout, H, W)).astype(np.float32)
    dw, db = ba.conv3x3_bwd_weights(dev(x), dev(dy))
    rdw, rdb = R.conv3x3_bwd_weights(R.bf16(x), R.bf16(dy))
    assert relerr(dw.cpu().numpy(), rdw) < 1e-5
    # the bias gradient sums dy in FP32 without rounding it
    assert relerr(db.cpu().numpy(), R.conv3x3_bwd_weights(x, dy)[1]) < 1e-5


def test_conv3x3_rejects_bad_shapes():
    x = torch.zeros((1, 65, 4, 64), device="cuda")
    w = torch.zeros((8, 65, 3, 3), device="cuda")
    with pytest.raises(ba.BartError, match="channels"):
        ba.conv3x3(x, w)
    x = torch.zeros((1, 4, 4, 100), device="cuda")
    dy = torch.zeros((1, 4, 4, 100), device="cuda")
    with pytest.raises(ba.BartError, match="multiple of 8"):
        ba.conv3x3_bwd_weights(x, dy)
