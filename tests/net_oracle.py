"""ctypes loader of the network oracle (oracle/libpvnet_oracle.so) and an FP64 U-Net built from it.

Test infrastructure only.  OracleUNet follows the layer table of pv_dec_original_amd.warmnet
(UNetSpec) and, with emulate_bf16=True, rounds every tensor to bf16 where the device path stores
bf16 (features, activations, activation gradients, packed weights), so that the device and the
oracle differ only by FP32-versus-FP64 accumulation.
"""
import ctypes as C
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
_L = None
_d = C.POINTER(C.c_double)
_i32 = C.POINTER(C.c_int32)


def load():
    global _L
    if _L is None:
        path = os.path.join(ROOT, "oracle", "libpvnet_oracle.so")
        if not os.path.exists(path):
            raise RuntimeError(f"{path} missing: run make -C oracle")
        L = C.CDLL(path)
        i, dp, vp = C.c_int, _d, C.c_void_p
        sig = {
            "onet_conv_fwd": (None, [_i32, i, i, i, dp, dp, dp, i, dp]),
            "onet_conv_dgrad": (None, [_i32, i, i, i, dp, dp, dp]),
            "onet_conv_wgrad": (None, [_i32, i, i, i, dp, dp, dp, dp]),
            "onet_pool2_fwd": (None, [_i32, i, dp, dp]),
            "onet_pool2_bwd": (None, [_i32, i, dp, dp]),
            "onet_up2_fwd": (None, [_i32, i, dp, dp]),
            "onet_up2_bwd": (None, [_i32, i, dp, dp]),
            "onet_wall_distance": (None, [_i32, vp, dp]),
            "onet_features": (None, [_i32, vp, dp, dp, dp]),
            "onet_mse_loss": (C.c_double, [C.c_int64, i, dp, i, dp, vp, dp, dp]),
            "onet_adam": (None, [C.c_int64, dp, dp, dp, dp, C.c_double, C.c_double, C.c_double,
                                 C.c_double, i]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype, fn.argtypes = res, args
        _L = L
    return _L


def _p(a):
    return None if a is None else a.ctypes.data_as(_d)


def _s(shape):
    return np.ascontiguousarray(shape, dtype=np.int32).ctypes.data_as(_i32)


def _c(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def bf16_round(a):
    """Nearest bf16 value (round to nearest even) of each element, via FP32 like the device."""
    f = np.ascontiguousarray(a, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u << 16).astype(np.uint32).view(np.float32).astype(np.float64)


def nvox(shape):
    return int(np.prod(shape))


def conv_fwd(shape, ntaps, cin, cout, x, w, b=None, relu=False):
    y = np.zeros((nvox(shape), cout))
    load().onet_conv_fwd(_s(shape), ntaps, cin, cout, _p(_c(x)), _p(_c(w)), _p(None if b is None else _c(b)),
                         1 if relu else 0, _p(y))
    return y


def conv_dgrad(shape, ntaps, cin, cout, dy, w):
    dx = np.zeros((nvox(shape), cin))
    load().onet_conv_dgrad(_s(shape), ntaps, cin, cout, _p(_c(dy)), _p(_c(w)), _p(dx))
    return dx


def conv_wgrad(shape, ntaps, cin, cout, x, dy):
    dw = np.zeros((cout, ntaps * cin))
    db = np.zeros(cout)
    load().onet_conv_wgrad(_s(shape), ntaps, cin, cout, _p(_c(x)), _p(_c(dy)), _p(dw), _p(db))
    return dw, db


def coarse(shape):
    nb, d, h, w = shape
    return (nb, d // 2, h // 2, w // 2)


def pool2_fwd(shape, c, x):
    y = np.zeros((nvox(coarse(shape)), c))
    load().onet_pool2_fwd(_s(shape), c, _p(_c(x)), _p(y))
    return y


def pool2_bwd(shape, c, dy):
    dx = np.zeros((nvox(shape), c))
    load().onet_pool2_bwd(_s(shape), c, _p(_c(dy)), _p(dx))
    return dx


def up2_fwd(shape, c, x):
    y = np.zeros((nvox(shape), c))
    load().onet_up2_fwd(_s(shape), c, _p(_c(x)), _p(y))
    return y


def up2_bwd(shape, c, dy):
    dx = np.zeros((nvox(coarse(shape)), c))
    load().onet_up2_bwd(_s(shape), c, _p(_c(dy)), _p(dx))
    return dx


def wall_distance(shape, code):
    out = np.zeros(nvox(shape))
    code = np.ascontiguousarray(code, dtype=np.uint8)
    load().onet_wall_distance(_s(shape), code.ctypes.data_as(C.c_void_p), _p(out))
    return out


def features(shape, code, prm, dist=None):
    """prm: (nb, 3) of (u_mean, tau, inlet_area); dist: wall distance (computed when None)."""
    out = np.zeros((nvox(shape), 8))
    code = np.ascontiguousarray(code, dtype=np.uint8)
    if dist is None:
        dist = wall_distance(shape, code)
    load().onet_features(_s(shape), code.ctypes.data_as(C.c_void_p), _p(_c(prm)), _p(_c(dist)), _p(out))
    return out


def mse_loss(y, t, code, wc, ldy=8):
    y = _c(y)
    t = _c(t)
    M = t.shape[0]
    nc = t.shape[1]
    dy = np.zeros((M, ldy))
    code = np.ascontiguousarray(code, dtype=np.uint8).reshape(-1)
    L = load().onet_mse_loss(M, nc, _p(y), ldy, _p(t), code.ctypes.data_as(C.c_void_p), _p(_c(wc)), _p(dy))
    return L, dy


def adam(p, g, m, v, lr, b1, b2, eps, t):
    p, m, v = _c(p).copy(), _c(m).copy(), _c(v).copy()
    load().onet_adam(p.size, _p(p), _p(_c(g)), _p(m), _p(v), lr, b1, b2, eps, t)
    return p, m, v


class OracleUNet:
    """FP64 U-Net over the oracle operators, same layer table and graph as warmnet.UNet."""

    def __init__(self, spec, params, emulate_bf16=True):
        self.spec = spec
        self.p = np.asarray(params, dtype=np.float64)
        self.emu = emulate_bf16

    def _r(self, a):
        return bf16_round(a) if self.emu else a

    def W(self, lay, params=None):
        p = self.p if params is None else params
        w = p[lay.offset:lay.boff].reshape(lay.cout, lay.K)
        return self._r(w), p[lay.boff:lay.boff + lay.cout].astype(np.float32).astype(np.float64) if self.emu else \
            p[lay.boff:lay.boff + lay.cout]

    def _conv(self, lay, shape, x, relu=None, f32=False):
        w, b = self.W(lay)
        y = conv_fwd(shape, lay.ntaps, lay.cin, lay.cout, x, w, b, lay.relu if relu is None else relu)
        return y if f32 else self._r(y)

    def forward(self, code, prm):
        sp = self.spec
        lv, L = sp.levels, len(sp.levels)
        nb, d, h, w = code.shape
        S = [(nb, d >> k, h >> k, w >> k) for k in range(L)]
        a = {"S": S}
        dist = wall_distance(S[0], code).astype(np.float32).astype(np.float64)
        a["X0"] = self._r(features(S[0], code, np.asarray(prm, dtype=np.float32).astype(np.float64), dist))
        x = a["X0"]
        by = sp.by_name
        for k in range(L - 1):
            a[f"A{k}"] = self._conv(by[f"e{k}a"], S[k], x)
            a[f"S{k}"] = self._conv(by[f"e{k}b"], S[k], a[f"A{k}"])
            a[f"P{k + 1}"] = self._r(pool2_fwd(S[k], lv[k], a[f"S{k}"]))
            x = a[f"P{k + 1}"]
        a["BA"] = self._conv(by["bta"], S[L - 1], x)
        a[f"DEC{L - 1}"] = self._conv(by["btb"], S[L - 1], a["BA"])
        for k in range(L - 2, -1, -1):
            up = self._r(up2_fwd(S[k], lv[k + 1], a[f"DEC{k + 1}"]))
            a[f"CAT{k}"] = np.concatenate([up, a[f"S{k}"]], axis=1)
            a[f"DA{k}"] = self._conv(by[f"d{k}a"], S[k], a[f"CAT{k}"])
            a[f"DEC{k}"] = self._conv(by[f"d{k}b"], S[k], a[f"DA{k}"])
        a["Y"] = self._conv(by["head"], S[0], a["DEC0"], f32=True)
        self.a = a
        return a["Y"]

    def backward(self, dY):
        """Gradient of every parameter (flat, spec layout) given dL/dY (M0, 8)."""
        sp, a = self.spec, self.a
        lv, L, S = sp.levels, len(sp.levels), a["S"]
        by = sp.by_name
        g = np.zeros(sp.nparams)
        r = self._r

        def wg(lay, shape, x, dy):
            dw, db = conv_wgrad(shape, lay.ntaps, lay.cin, lay.cout, x, dy)
            g[lay.offset:lay.boff] = dw.reshape(-1)
            g[lay.boff:lay.boff + lay.cout] = db

        def dg(lay, shape, dy, relu_of=None):
            dx = conv_dgrad(shape, lay.ntaps, lay.cin, lay.cout, dy, self.W(lay)[0])
            if relu_of is not None:
                dx = dx * (relu_of > 0)
            return r(dx)

        dY = r(dY)
        head = by["head"]
        wg(head, S[0], a["DEC0"], dY)
        dDEC = dg(head, S[0], dY, a["DEC0"])
        dskip = {}
        for k in range(L - 1):
            la, lb = by[f"d{k}a"], by[f"d{k}b"]
            wg(lb, S[k], a[f"DA{k}"], dDEC)
            dDA = dg(lb, S[k], dDEC, a[f"DA{k}"])
            wg(la, S[k], a[f"CAT{k}"], dDA)
            dCAT = dg(la, S[k], dDA)
            dskip[k] = dCAT[:, lv[k + 1]:]
            dDEC = r(up2_bwd(S[k], lv[k + 1], dCAT[:, :lv[k + 1]]))
            dDEC = dDEC * (a[f"DEC{k + 1}"] > 0)
        top = L - 1
        wg(by["btb"], S[top], a["BA"], dDEC)
        dBA = dg(by["btb"], S[top], dDEC, a["BA"])
        wg(by["bta"], S[top], a[f"P{top}"], dBA)
        dP = dg(by["bta"], S[top], dBA)
        for k in range(L - 2, -1, -1):
            ds = r(dskip[k] + pool2_bwd(S[k], lv[k], dP))
            ds = ds * (a[f"S{k}"] > 0)
            wg(by[f"e{k}b"], S[k], a[f"A{k}"], ds)
            dA = dg(by[f"e{k}b"], S[k], ds, a[f"A{k}"])
            xk = a["X0"] if k == 0 else a[f"P{k}"]
            wg(by[f"e{k}a"], S[k], xk, dA)
            if k > 0:
                dP = dg(by[f"e{k}a"], S[k], dA)
        return g
