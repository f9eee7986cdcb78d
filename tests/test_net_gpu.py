"""Device kernels of the warm-start network (libpvnet.so) against the network oracle.

Inputs are bf16 values, so device and oracle see identical operands; the device accumulates in
FP32 and rounds its bf16 outputs to nearest even, the oracle accumulates in FP64.  Tolerances:
bf16 outputs within 2 units in the last place (2 * 2^-8 relative) plus an absolute floor of 1e-3
of the output's scale for sums that cancel; FP32 outputs within 1e-5 relative to the scale."""
import ctypes as C

import numpy as np
import pytest

from tests import net_oracle as O

pytestmark = pytest.mark.gpu

RNG = np.random.default_rng(99)


@pytest.fixture(scope="module")
def T():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    return torch


def _dev(T, a, dtype=None):
    dtype = dtype or T.bfloat16
    return T.from_numpy(np.ascontiguousarray(a, dtype=np.float32)).to("cuda").to(dtype)


def _np(t):
    return t.float().cpu().numpy().astype(np.float64)


def _close_bf16(got, ref, what):
    scale = max(np.abs(ref).max(), 1e-30)
    tol = 2 * 2.0 ** -8 * np.abs(ref) + 1e-3 * scale
    bad = np.abs(got - ref) > tol
    assert not bad.any(), f"{what}: {bad.sum()} of {bad.size} beyond tolerance, max err " \
                          f"{np.abs(got - ref).max():.3e} (scale {scale:.3e})"


def _rbf(shape):
    return O.bf16_round(RNG.standard_normal(shape))


def _st(T):
    return C.c_void_p(T.cuda.current_stream().cuda_stream)


def _p(t, off=0):
    return C.c_void_p(t.data_ptr() + off * t.element_size())


CONV_CASES = [
    # shape, ntaps, cin, cout, ldx, xoff, ldy, yoff, flags
    ((2, 4, 6, 8), 27, 8, 16, 8, 0, 16, 0, 1),
    ((1, 8, 8, 8), 27, 16, 16, 24, 8, 48, 32, 1),
    ((2, 4, 4, 8), 27, 48, 16, 48, 0, 16, 0, 0),
    ((1, 6, 4, 8), 27, 32, 64, 32, 0, 64, 0, 1),
    ((1, 4, 4, 4), 27, 64, 160, 64, 0, 160, 0, 1),
    ((1, 2, 2, 8), 27, 224, 64, 224, 0, 64, 0, 1),
    ((2, 4, 4, 4), 1, 16, 8, 16, 0, 8, 0, 2),
    ((1, 12, 12, 12), 27, 8, 32, 8, 0, 32, 0, 1),
]


@pytest.mark.parametrize("case", CONV_CASES, ids=[f"c{i}" for i in range(len(CONV_CASES))])
def test_conv_fwd_and_dgrad_and_wgrad(T, case):
    from pv_dec_original_amd import _native as N
    shape, ntaps, cin, cout, ldx, xoff, ldy, yoff, flags = case
    M = O.nvox(shape)
    NL = N.net()
    x = _rbf((M, cin))
    w = _rbf((cout, ntaps * cin)) / np.sqrt(ntaps * cin)
    w = O.bf16_round(w)
    b = RNG.standard_normal(cout).astype(np.float32).astype(np.float64)
    relu = bool(flags & 1)
    f32 = bool(flags & 2)
    ref = O.conv_fwd(shape, ntaps, cin, cout, x, w, b, relu)
    # device: x as a channel slice of a wider tensor, y into a slice too
    xfull = np.zeros((M, ldx))
    xfull[:, xoff:xoff + cin] = x
    xd = _dev(T, xfull)
    kpad = (ntaps * cin + 31) // 32 * 32
    kpad_d = (ntaps * cout + 31) // 32 * 32
    w32 = _dev(T, w, T.float32)
    wt = T.zeros((cout, kpad), dtype=T.bfloat16, device="cuda")
    wd = T.zeros((cin, kpad_d), dtype=T.bfloat16, device="cuda")
    N.net_check(NL.pvnet_pack_weights(_p(w32), ntaps, cin, cout, _p(wt), _p(wd), _st(T)))
    bd = _dev(T, b, T.float32)
    yd = T.full((M, ldy), 7.0, dtype=T.float32 if f32 else T.bfloat16, device="cuda")
    c = N.NetConv(N.NetShape(*shape), ntaps, cin, cout, kpad, _p(xd, xoff), ldx, flags, _p(wt), _p(bd),
                  _p(yd, yoff), ldy, 0, None)
    N.net_check(NL.pvnet_conv_fwd(C.byref(c), _st(T)))
    T.cuda.synchronize()
    got = _np(yd)
    outside = np.delete(got, np.s_[yoff:yoff + cout], axis=1)
    assert np.all(outside == 7.0), "conv wrote outside its output slice"
    got = got[:, yoff:yoff + cout]
    if f32:
        assert np.allclose(got, ref, rtol=1e-5, atol=1e-5 * np.abs(ref).max())
    else:
        _close_bf16(got, ref, "conv_fwd")

    # input gradient through the packed transposed weights, with the ReLU mask of an activation
    dy = _rbf((M, cout))
    act = _rbf((M, cin))
    refdx = O.conv_dgrad(shape, ntaps, cin, cout, dy, w) * (act > 0)
    dyd, actd = _dev(T, dy), _dev(T, act)
    dxd = T.zeros((M, cin), dtype=T.bfloat16, device="cuda")
    c = N.NetConv(N.NetShape(*shape), ntaps, cout, cin, kpad_d, _p(dyd), cout, 0, _p(wd), None,
                  _p(dxd), cin, cin, _p(actd))
    if cout % 8 == 0:
        N.net_check(NL.pvnet_conv_fwd(C.byref(c), _st(T)))
        T.cuda.synchronize()
        _close_bf16(_np(dxd), refdx, "dgrad")

    # weight and bias gradients (x read through the same slice)
    refdw, refdb = O.conv_wgrad(shape, ntaps, cin, cout, x, dy)
    ws = T.zeros(int(NL.pvnet_wgrad_ws_floats(C.byref(N.NetShape(*shape)), ntaps, cin, cout)), dtype=T.float32,
                 device="cuda")
    dw = T.zeros((cout, ntaps * cin), dtype=T.float32, device="cuda")
    db = T.zeros(cout, dtype=T.float32, device="cuda")
    g = N.NetWgrad(N.NetShape(*shape), ntaps, cin, cout, _p(xd, xoff), ldx, cout, _p(dyd), _p(ws), _p(dw), _p(db))
    N.net_check(NL.pvnet_conv_wgrad(C.byref(g), _st(T)))
    T.cuda.synchronize()
    sc = np.abs(refdw).max()
    assert np.allclose(_np(dw), refdw, rtol=1e-5, atol=1e-5 * sc), np.abs(_np(dw) - refdw).max() / sc
    assert np.allclose(_np(db), refdb, rtol=1e-5, atol=1e-5 * np.abs(refdb).max())


def test_conv_fwd_large_grid_tiles(T):
    """The 256-voxel gathering tile used for narrow outputs on large grids (PVNET_GATHER keeps the
    convolution off the brick kernel), on sampled voxels against a direct numpy sum over the 27
    taps."""
    from pv_dec_original_amd import _native as N
    NL = N.net()
    shape, cin, cout = (1, 64, 96, 96), 16, 16
    M = O.nvox(shape)
    x = _rbf((M, cin))
    w = O.bf16_round(_rbf((cout, 27 * cin)) / np.sqrt(27 * cin))
    b = RNG.standard_normal(cout).astype(np.float32).astype(np.float64)
    xd, w32, bd = _dev(T, x), _dev(T, w, T.float32), _dev(T, b, T.float32)
    kpad = (27 * cin + 31) // 32 * 32
    wt = T.zeros((cout, kpad), dtype=T.bfloat16, device="cuda")
    wd = T.zeros((cin, (27 * cout + 31) // 32 * 32), dtype=T.bfloat16, device="cuda")
    N.net_check(NL.pvnet_pack_weights(_p(w32), 27, cin, cout, _p(wt), _p(wd), _st(T)))
    yd = T.zeros((M, cout), dtype=T.bfloat16, device="cuda")
    c = N.NetConv(N.NetShape(*shape), 27, cin, cout, kpad, _p(xd), cin, 1 | N.NET_GATHER, _p(wt), _p(bd), _p(yd), cout,
                  0, None)
    N.net_check(NL.pvnet_conv_fwd(C.byref(c), _st(T)))
    T.cuda.synchronize()
    got = _np(yd)
    _, d, h, wd_ = shape
    xs = x.reshape(d, h, wd_, cin)
    ws = w.reshape(cout, 27, cin)
    for m in RNG.integers(0, M, 300):
        z, y, xx = m // (h * wd_), (m // wd_) % h, m % wd_
        acc = b.copy()
        for tap in range(27):
            dz, dy, dx = tap // 9 - 1, (tap // 3) % 3 - 1, tap % 3 - 1
            if 0 <= z + dz < d and 0 <= y + dy < h and 0 <= xx + dx < wd_:
                acc += ws[:, tap, :] @ xs[z + dz, y + dy, xx + dx]
        ref = np.maximum(acc, 0.0)
        _close_bf16(got[m], ref, f"voxel {m}")


@pytest.mark.parametrize("cin,cout,flags,ldx,xoff,ldy,yoff,mask", [
    (8, 16, 1, 8, 0, 16, 0, False), (16, 16, 1, 24, 8, 48, 32, False), (32, 16, 2, 32, 0, 16, 0, False),
    (48, 16, 0, 48, 0, 16, 0, True), (16, 32, 1, 16, 0, 64, 32, False), (32, 32, 0, 40, 8, 32, 0, True),
    (48, 32, 3, 48, 0, 32, 0, False), (16, 48, 0, 16, 0, 48, 0, True), (32, 64, 1, 32, 0, 64, 0, False),
    (16, 96, 0, 24, 8, 96, 0, True)])
def test_conv_fwd_brick_equals_gather(T, cin, cout, flags, ldx, xoff, ldy, yoff, mask):
    """3x3x3 convolutions with 16 to 96 outputs on whole 16 x 4 x 4 bricks stage each brick with
    its halo in LDS; the result must equal the per-tap gathering kernel's (PVNET_GATHER) bit for bit,
    over two boxes (zero padding at every box face), channel slices of wider tensors, ReLU, FP32
    output and the ReLU-backward mask.  Columns outside the output slice stay untouched."""
    from pv_dec_original_amd import _native as N
    NL = N.net()
    shape = (2, 32, 96, 96)
    M = O.nvox(shape)
    xfull = np.zeros((M, ldx))
    xfull[:, xoff:xoff + cin] = _rbf((M, cin))
    w = O.bf16_round(_rbf((cout, 27 * cin)) / np.sqrt(27 * cin))
    b = RNG.standard_normal(cout).astype(np.float32).astype(np.float64)
    xd, w32, bd = _dev(T, xfull), _dev(T, w, T.float32), _dev(T, b, T.float32)
    kpad = (27 * cin + 31) // 32 * 32
    wt = T.zeros((cout, kpad), dtype=T.bfloat16, device="cuda")
    wdp = T.zeros((cin, (27 * cout + 31) // 32 * 32), dtype=T.bfloat16, device="cuda")
    N.net_check(NL.pvnet_pack_weights(_p(w32), 27, cin, cout, _p(wt), _p(wdp), _st(T)))
    rel = _dev(T, _rbf((M, cout))) if mask else None
    f32 = bool(flags & 2)
    out = {}
    for gather in (False, True):
        yd = T.full((M, ldy), 7.0, dtype=T.float32 if f32 else T.bfloat16, device="cuda")
        c = N.NetConv(N.NetShape(*shape), 27, cin, cout, kpad, _p(xd, xoff), ldx, flags | (N.NET_GATHER if gather else 0),
                      _p(wt), _p(bd), _p(yd, yoff), ldy, cout if mask else 0, _p(rel) if mask else None)
        N.net_check(NL.pvnet_conv_fwd(C.byref(c), _st(T)))
        T.cuda.synchronize()
        out[gather] = yd
    a, g = out[False], out[True]
    assert T.equal(a, g), f"{(a != g).sum().item()} of {a.numel()} differ"
    assert (a[:, :yoff] == 7.0).all() and (a[:, yoff + cout:] == 7.0).all()
    assert a[:, yoff:yoff + cout].abs().max().item() > 0


@pytest.mark.parametrize("shape,cin,cout,flags,ldx,xoff,ldy,yoff,mask", [
    ((2, 8, 16, 24), 16, 8, 2, 16, 0, 8, 0, False), ((2, 8, 16, 24), 8, 16, 0, 8, 0, 16, 0, True),
    ((1, 32, 96, 96), 16, 8, 2, 24, 8, 8, 0, False), ((1, 32, 96, 96), 8, 16, 0, 8, 0, 16, 0, True),
    ((1, 12, 20, 28), 32, 40, 1, 48, 16, 64, 8, False), ((3, 4, 4, 5), 24, 64, 1, 24, 0, 64, 0, True)])
def test_conv_pointwise_equals_gather(T, shape, cin, cout, flags, ldx, xoff, ldy, yoff, mask):
    """1x1x1 convolutions with K <= 32 feed the matrix core from global memory (no LDS); the result
    must equal the tiled kernel's (PVNET_GATHER) bit for bit, with input and output channel slices,
    ReLU, FP32 output and the ReLU-backward mask, including voxel counts that leave a partial
    tile.  Columns outside the output slice stay untouched."""
    from pv_dec_original_amd import _native as N
    NL = N.net()
    M = O.nvox(shape)
    xfull = np.zeros((M, ldx))
    xfull[:, xoff:xoff + cin] = _rbf((M, cin))
    w = O.bf16_round(_rbf((cout, cin)) / np.sqrt(cin))
    b = RNG.standard_normal(cout).astype(np.float32).astype(np.float64)
    xd, w32, bd = _dev(T, xfull), _dev(T, w, T.float32), _dev(T, b, T.float32)
    kpad = (cin + 31) // 32 * 32
    wt = T.zeros((cout, kpad), dtype=T.bfloat16, device="cuda")
    wdp = T.zeros((cin, (cout + 31) // 32 * 32), dtype=T.bfloat16, device="cuda")
    N.net_check(NL.pvnet_pack_weights(_p(w32), 1, cin, cout, _p(wt), _p(wdp), _st(T)))
    rel = _dev(T, _rbf((M, cout))) if mask else None
    f32 = bool(flags & 2)
    out = {}
    for gather in (False, True):
        yd = T.full((M, ldy), 7.0, dtype=T.float32 if f32 else T.bfloat16, device="cuda")
        c = N.NetConv(N.NetShape(*shape), 1, cin, cout, kpad, _p(xd, xoff), ldx, flags | (N.NET_GATHER if gather else 0),
                      _p(wt), _p(bd), _p(yd, yoff), ldy, cout if mask else 0, _p(rel) if mask else None)
        N.net_check(NL.pvnet_conv_fwd(C.byref(c), _st(T)))
        T.cuda.synchronize()
        out[gather] = yd
    a, g = out[False], out[True]
    assert T.equal(a, g), f"{(a != g).sum().item()} of {a.numel()} differ"
    assert (a[:, :yoff] == 7.0).all() and (a[:, yoff + cout:] == 7.0).all()
    assert a[:, yoff:yoff + cout].abs().max().item() > 0


@pytest.mark.parametrize("shape,cin,cout,ldx,xoff", [
    ((2, 8, 16, 16), 8, 16, 8, 0), ((1, 12, 16, 24), 16, 16, 24, 8), ((3, 4, 8, 8), 24, 16, 24, 0),
    ((2, 8, 8, 16), 48, 16, 56, 8), ((2, 8, 16, 8), 8, 32, 8, 0), ((1, 8, 8, 16), 32, 32, 48, 16),
    ((2, 4, 8, 8), 96, 32, 96, 0)])
def test_wgrad_brick_against_oracle_and_gather(T, shape, cin, cout, ldx, xoff):
    """16- and 32-output 3x3x3 weight gradients on grids of whole 8 x 8 x 4 bricks reduce brick by
    brick with the input staged in LDS: against the FP64 oracle (FP32 accumulation, 1e-5 of the
    scale) and against the per-tap gathering kernel (PVNET_GATHER, the same sums in another order),
    over several boxes (zero padding at every face), one to twelve 8-channel groups, and an input
    slice of a wider tensor; two runs give the same bits.  The bias gradient too: the 16-output
    brick kernels sum it in the same pass (dy^T times ones on the matrix core)."""
    from pv_dec_original_amd import _native as N
    NL = N.net()
    M = O.nvox(shape)
    xfull = np.zeros((M, ldx))
    xfull[:, xoff:xoff + cin] = _rbf((M, cin))
    dy = _rbf((M, cout))
    refdw, refdb = O.conv_wgrad(shape, 27, cin, cout, xfull[:, xoff:xoff + cin], dy)
    xd, dyd = _dev(T, xfull), _dev(T, dy)
    ws = T.zeros(int(NL.pvnet_wgrad_ws_floats(C.byref(N.NetShape(*shape)), 27, cin, cout)), dtype=T.float32,
                 device="cuda")
    sc = np.abs(refdw).max()
    out = {}
    for run, flags in (("brick", 0), ("brick again", 0), ("gather", N.NET_GATHER)):
        dw = T.full((cout, 27 * cin), 7.0, dtype=T.float32, device="cuda")
        db = T.zeros(cout, dtype=T.float32, device="cuda")
        g = N.NetWgrad(N.NetShape(*shape), 27, cin, cout, _p(xd, xoff), ldx, cout, _p(dyd), _p(ws), _p(dw), _p(db),
                       flags)
        N.net_check(NL.pvnet_conv_wgrad(C.byref(g), _st(T)))
        T.cuda.synchronize()
        out[run] = _np(dw)
        out[run + " db"] = _np(db)
        assert np.allclose(out[run], refdw, rtol=1e-5, atol=1e-5 * sc), (run, np.abs(out[run] - refdw).max() / sc)
        assert np.allclose(_np(db), refdb, rtol=1e-5, atol=1e-5 * np.abs(refdb).max()), run
    assert np.array_equal(out["brick"], out["brick again"])
    assert np.array_equal(out["brick db"], out["brick again db"])
    assert np.allclose(out["brick"], out["gather"], rtol=1e-5, atol=1e-6 * sc)


def test_wgrad_deterministic(T):
    from pv_dec_original_amd import _native as N
    NL = N.net()
    shape, ntaps, cin, cout = (2, 16, 16, 16), 27, 16, 32
    M = O.nvox(shape)
    xd, dyd = _dev(T, _rbf((M, cin))), _dev(T, _rbf((M, cout)))
    ws = T.zeros(int(NL.pvnet_wgrad_ws_floats(C.byref(N.NetShape(*shape)), ntaps, cin, cout)), dtype=T.float32,
                 device="cuda")
    outs = []
    for _ in range(2):
        dw = T.zeros((cout, ntaps * cin), dtype=T.float32, device="cuda")
        g = N.NetWgrad(N.NetShape(*shape), ntaps, cin, cout, _p(xd), cin, cout, _p(dyd), _p(ws), _p(dw), None)
        N.net_check(NL.pvnet_conv_wgrad(C.byref(g), _st(T)))
        outs.append(dw.cpu().numpy())
    assert np.array_equal(outs[0], outs[1])


def test_pool_up_relu_features_loss_adam(T):
    from pv_dec_original_amd import _native as N
    NL = N.net()
    st = _st(T)
    shape, c = (2, 4, 6, 8), 16
    M = O.nvox(shape)
    Mc = O.nvox(O.coarse(shape))
    S = N.NetShape(*shape)
    x = _rbf((M, c))
    xd = _dev(T, x)
    yd = T.zeros((Mc, c), dtype=T.bfloat16, device="cuda")
    N.net_check(NL.pvnet_pool2_fwd(C.byref(S), c, _p(xd), c, _p(yd), c, st))
    T.cuda.synchronize()
    _close_bf16(_np(yd), O.pool2_fwd(shape, c, x), "pool2_fwd")
    dy = _rbf((Mc, c))
    old = _rbf((M, c))
    dxd = _dev(T, old)
    dyc = _dev(T, dy)
    N.net_check(NL.pvnet_pool2_bwd(C.byref(S), c, _p(dyc), c, _p(dxd), c, 1, st))
    T.cuda.synchronize()
    _close_bf16(_np(dxd), old + O.pool2_bwd(shape, c, dy), "pool2_bwd accumulate")
    xc = _rbf((Mc, c))
    ud = T.zeros((M, c), dtype=T.bfloat16, device="cuda")
    xcd = _dev(T, xc)
    N.net_check(NL.pvnet_up2_fwd(C.byref(S), c, _p(xcd), c, _p(ud), c, st))
    T.cuda.synchronize()
    assert np.array_equal(_np(ud), O.up2_fwd(shape, c, xc))
    dyf = _rbf((M, c))
    dcd = T.zeros((Mc, c), dtype=T.bfloat16, device="cuda")
    dyfd = _dev(T, dyf)
    N.net_check(NL.pvnet_up2_bwd(C.byref(S), c, _p(dyfd), c, _p(dcd), c, st))
    T.cuda.synchronize()
    _close_bf16(_np(dcd), O.up2_bwd(shape, c, dyf), "up2_bwd")
    a = _rbf((M, c))
    g = _rbf((M, c))
    gd = _dev(T, g)
    ad = _dev(T, a)
    N.net_check(NL.pvnet_relu_bwd(M, c, _p(ad), c, _p(gd), c, st))
    T.cuda.synchronize()
    assert np.array_equal(_np(gd), g * (a > 0))
    # the fused pooling backward with the ReLU mask equals the two calls, on channel slices of rows
    # of 3c channels
    wide_old, wide_a = _rbf((M, 3 * c)), _rbf((M, 3 * c))
    two, one, wad = _dev(T, wide_old), _dev(T, wide_old), _dev(T, wide_a)
    N.net_check(NL.pvnet_pool2_bwd(C.byref(S), c, _p(dyc), c, _p(two, c), 3 * c, 1, st))
    N.net_check(NL.pvnet_relu_bwd(M, c, _p(wad, c), 3 * c, _p(two, c), 3 * c, st))
    N.net_check(NL.pvnet_pool2_bwd_relu(C.byref(S), c, _p(dyc), c, _p(one, c), 3 * c, _p(wad, c), 3 * c, st))
    T.cuda.synchronize()
    assert T.equal(one, two)
    assert T.equal(one[:, :c], _dev(T, wide_old)[:, :c]) and T.equal(one[:, 2 * c:], _dev(T, wide_old)[:, 2 * c:])
    # wall distance and features
    code = RNG.integers(0, 4, shape).astype(np.uint8)
    code[RNG.random(shape) < 0.5] = 1
    prm = np.array([[0.03, 0.7, 12.0], [0.012, 0.58, 30.0]])
    cd = T.from_numpy(code).to("cuda")
    distd = T.zeros(M, dtype=T.float32, device="cuda")
    N.net_check(NL.pvnet_wall_distance(C.byref(S), _p(cd), _p(distd), st))
    fd = T.zeros((M, 8), dtype=T.bfloat16, device="cuda")
    prmd = _dev(T, prm, T.float32)
    N.net_check(NL.pvnet_features(C.byref(S), _p(cd), _p(prmd), _p(distd), _p(fd), st))
    T.cuda.synchronize()
    dref = O.wall_distance(shape, code)
    assert np.allclose(distd.cpu().numpy(), dref, rtol=1e-6, atol=1e-6)
    _close_bf16(_np(fd), O.bf16_round(O.features(shape, code, prm.astype(np.float32).astype(np.float64),
                                                 dref.astype(np.float32).astype(np.float64))), "features")
    # loss
    y = RNG.standard_normal((M, 8)).astype(np.float32).astype(np.float64)
    t = _rbf((M, 4))
    wc = np.array([1.0, 1.0, 2.0, 0.5])
    Lref, dyref = O.mse_loss(y, t, code, wc)
    part = T.zeros(int(NL.pvnet_loss_ws_doubles(M)), dtype=T.float64, device="cuda")
    loss = T.zeros(2, dtype=T.float64, device="cuda")
    dyd = T.zeros((M, 8), dtype=T.bfloat16, device="cuda")
    # every device operand is held by a name until the kernels have run: a temporary freed while
    # the arguments are still being built can be handed to the next temporary by the allocator
    yd, tdv, wcd = _dev(T, y, T.float32), _dev(T, t), _dev(T, wc, T.float32)
    N.net_check(NL.pvnet_mse_loss(M, 4, _p(yd), 8, _p(tdv), _p(cd), _p(wcd), _p(dyd), 8, _p(part), _p(loss), st))
    T.cuda.synchronize()
    lv = loss.cpu().numpy()
    assert np.isclose(lv[0], Lref, rtol=1e-6) and lv[1] == np.count_nonzero(code)
    _close_bf16(_np(dyd), dyref, "loss gradient")
    # adam, two steps
    n = 1000
    p0, gr = RNG.standard_normal(n), RNG.standard_normal(n)
    pd, gdv = _dev(T, p0, T.float32), _dev(T, gr, T.float32)
    md, vd = T.zeros(n, device="cuda"), T.zeros(n, device="cuda")
    pr, mr, vr = p0.astype(np.float32).astype(np.float64), np.zeros(n), np.zeros(n)
    grr = gr.astype(np.float32).astype(np.float64)
    for step in (1, 2):
        N.net_check(NL.pvnet_adam(n, _p(pd), _p(gdv), _p(md), _p(vd), 1e-3, 0.9, 0.999, 1e-8, step, st))
        pr, mr, vr = O.adam(pr, grr, mr, vr, 1e-3, 0.9, 0.999, 1e-8, step)
    T.cuda.synchronize()
    assert np.allclose(_np(pd), pr, rtol=1e-6, atol=1e-7)


def test_wall_distance_large_box(T):
    """The separable transform on a 96^3 vessel-like box against the brute-force oracle on a
    sample of voxels (the oracle is exact; every voxel would take minutes)."""
    from pv_dec_original_amd import _native as N
    NL = N.net()
    n = 96
    z, y, x = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    code = ((((y - 40) ** 2 + (z - 50) ** 2) < 14 ** 2) | (((x - 60) ** 2 + (y - 30) ** 2) < 9 ** 2)).astype(np.uint8)
    code = code.reshape(1, n, n, n)
    cd = T.from_numpy(code).to("cuda")
    distd = T.zeros(n ** 3, dtype=T.float32, device="cuda")
    N.net_check(NL.pvnet_wall_distance(C.byref(N.NetShape(1, n, n, n)), _p(cd), _p(distd), _st(T)))
    T.cuda.synchronize()
    got = distd.cpu().numpy()
    solid = np.argwhere(code[0] == 0)
    fl = np.flatnonzero(code.reshape(-1))
    for m in list(RNG.choice(fl, 150)) + list(RNG.integers(0, n ** 3, 50)):
        p = np.array([m // (n * n), (m // n) % n, m % n])
        ref = np.sqrt(((solid - p) ** 2).sum(axis=1).min())
        assert abs(got[m] - ref) < 1e-4, (p, got[m], ref)


def test_entry_points_reject_bad_arguments(T):
    from pv_dec_original_amd import _native as N
    NL = N.net()
    c = N.NetConv(N.NetShape(1, 4, 4, 4), 27, 12, 16, 352, None, 12, 0, None, None, None, 16, 0, None)
    assert NL.pvnet_conv_fwd(C.byref(c), None) == 1
    assert b"multiple of 8" in NL.pvnet_last_error()
    s = N.NetShape(1, 3, 4, 4)
    assert NL.pvnet_pool2_fwd(C.byref(s), 8, None, 8, None, 8, None) == 1


def _unet_case(T, levels, shape, seed):
    from pv_dec_original_amd.warmnet import UNet, UNetSpec
    spec = UNetSpec(levels=levels)
    params = spec.init_params(seed)
    rng = np.random.default_rng(seed)
    params = (params + 0.02 * rng.standard_normal(params.size)).astype(np.float32)
    code = rng.integers(0, 4, shape).astype(np.uint8)
    code[rng.random(shape) < 0.3] = 0
    prm = np.array([[0.01 + 0.04 * rng.random(), 0.55 + 0.35 * rng.random(), float(np.count_nonzero(code[b] == 2))]
                    for b in range(shape[0])], dtype=np.float32)
    t = O.bf16_round(rng.standard_normal((O.nvox(shape), 4)))
    return spec, params, code, prm, t, UNet(spec, params)


def test_unet_forward_backward_matches_oracle(T):
    spec, params, code, prm, t, net = _unet_case(T, (8, 16, 32), (2, 8, 8, 16), 5)
    codd = T.from_numpy(code).to("cuda")
    y = net.forward(codd, _dev(T, prm, T.float32))
    loss = net.loss(_dev(T, t), codd)
    net.backward()
    T.cuda.synchronize()
    orc = O.OracleUNet(spec, params.astype(np.float64), emulate_bf16=True)
    yref = orc.forward(code, prm.astype(np.float64))
    got = _np(y)
    err = np.linalg.norm(got - yref) / np.linalg.norm(yref)
    assert err < 5e-3, f"forward relative L2 error {err:.2e}"
    Lref, dYref = O.mse_loss(got, t, code, np.ones(4))
    assert np.isclose(loss.cpu().numpy()[0], Lref, rtol=1e-5)
    # backward from the same dL/dY so that only the backward kernels are compared
    gref = orc.backward(_np(net._bufs["dY"]))
    g = _np(net.g)
    for lay in spec.layers:
        sl = slice(lay.offset, lay.offset + lay.nparams)
        e = np.linalg.norm(g[sl] - gref[sl]) / max(np.linalg.norm(gref[sl]), 1e-30)
        assert e < 2e-2, f"{lay.name}: gradient relative L2 error {e:.2e}"


def test_unet_training_reduces_loss(T):
    spec, params, code, prm, _, net = _unet_case(T, (8, 16, 32), (2, 8, 8, 16), 11)
    codd = T.from_numpy(code).to("cuda")
    prmd = _dev(T, prm, T.float32)
    # a smooth target the network can represent: a function of position inside the fluid
    z, y, x = np.meshgrid(np.arange(8), np.arange(8), np.arange(16), indexing="ij")
    base = np.stack([np.sin(x / 5.0), np.cos(y / 4.0), z / 8.0, x * y / 128.0], -1)
    t = O.bf16_round(np.broadcast_to(base, (2, 8, 8, 16, 4)).reshape(-1, 4))
    td = _dev(T, t)
    losses = []
    for it in range(60):
        net.forward(codd, prmd)
        losses.append(float(net.loss(td, codd)[0]))
        net.backward()
        net.adam_step(3e-3)
    assert np.isfinite(losses).all()
    assert losses[-1] < 0.3 * losses[0], losses[::10]
