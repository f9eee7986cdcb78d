"""The warm-start network: a 3D convolutional U-Net on voxel geometry, run by libpvnet.so.

The reference solver always starts from rest (Simulation.cpp:13-15: every population at its
weight).  This network maps a geometry code volume (0 solid, 1 fluid, 2 inlet face, 3 outlet face),
its wall distance and the inlet parameters (u_mean, tau, inlet area) to an initial field,
normalised as
    y[..., 0:3] = u / u_mean,   y[..., 3] = (rho - 1) / u_mean,
from which the solver runs on (pvlbm_init_fields); the solver decides the result, and starting from
rest stays the fallback.

Architecture (levels = channels per resolution, default (16, 32, 64, 160), 1.82 M parameters):
    encoder level l < L-1:  conv3(c_{l-1} -> c_l) ReLU, conv3(c_l -> c_l) ReLU = skip s_l,
                            2x2x2 average pool
    bottleneck:             conv3(c_{L-2} -> c_{L-1}) ReLU, conv3(c_{L-1} -> c_{L-1}) ReLU
    decoder level l:        nearest 2x upsampling, concat [up, s_l],
                            conv3(c_{l+1} + c_l -> c_l) ReLU, conv3(c_l -> c_l) ReLU
    head:                   conv1(c_0 -> 8), channels 0..3 used, FP32 output
Every convolution, forward and backward, is a hand-written BF16 MFMA kernel (FP32 accumulation);
activations and their gradients are bf16, master weights, gradients and Adam moments FP32.
PyTorch only provides device memory and the current stream.
"""
import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _native as N

IN_CH = 8
OUT_CH = 8
N_TARGET = 4
DEFAULT_LEVELS = (16, 32, 64, 160)


def _rup(v, m):
    return (v + m - 1) // m * m


@dataclass(frozen=True)
class Layer:
    name: str
    ntaps: int
    cin: int
    cout: int
    relu: bool
    offset: int  # of the weights in the flat parameter vector; the bias follows them

    @property
    def K(self):
        return self.ntaps * self.cin

    @property
    def kpad(self):
        return _rup(self.ntaps * self.cin, 32)

    @property
    def kpad_d(self):
        return _rup(self.ntaps * self.cout, 32)

    @property
    def boff(self):
        return self.offset + self.cout * self.K

    @property
    def nparams(self):
        return self.cout * self.K + self.cout


class UNetSpec:
    """Layer table and parameter layout of the U-Net with the given channels per level."""

    def __init__(self, levels=DEFAULT_LEVELS, cin=IN_CH, cout=OUT_CH):
        if len(levels) < 2 or any(c % 8 for c in levels) or cin % 8 or cout % 8:
            raise ValueError("need >= 2 levels and channel counts that are multiples of 8")
        self.levels = tuple(int(c) for c in levels)
        self.cin, self.cout = cin, cout
        L = len(levels)
        table = []
        cp = cin
        for lv in range(L - 1):
            table += [(f"e{lv}a", 27, cp, levels[lv], True), (f"e{lv}b", 27, levels[lv], levels[lv], True)]
            cp = levels[lv]
        table += [("bta", 27, cp, levels[L - 1], True), ("btb", 27, levels[L - 1], levels[L - 1], True)]
        for lv in range(L - 2, -1, -1):
            table += [(f"d{lv}a", 27, levels[lv + 1] + levels[lv], levels[lv], True),
                      (f"d{lv}b", 27, levels[lv], levels[lv], True)]
        table.append(("head", 1, levels[0], cout, False))
        self.layers = []
        off = 0
        for name, nt, ci, co, relu in table:
            lay = Layer(name, nt, ci, co, relu, off)
            self.layers.append(lay)
            off += lay.nparams
        self.nparams = off
        self.by_name = {lay.name: lay for lay in self.layers}
        self.divisor = 2 ** (L - 1)

    def init_params(self, seed=0):
        """He-normal weights (std sqrt(2 / fan_in)), zero biases; the head starts small."""
        rng = np.random.default_rng(seed)
        p = np.zeros(self.nparams, dtype=np.float32)
        for lay in self.layers:
            std = np.sqrt(2.0 / lay.K) if lay.relu else 0.1 / np.sqrt(lay.K)
            p[lay.offset:lay.boff] = rng.standard_normal(lay.cout * lay.K).astype(np.float32) * std
        return p

    def check_shape(self, shape):
        nb, d, h, w = shape
        if min(shape) <= 0 or d % self.divisor or h % self.divisor or w % self.divisor:
            raise ValueError(f"box sizes must be multiples of {self.divisor}, got {shape}")

    def check_train_shape(self, shape):
        """The weight-gradient kernels also need a multiple of 32 voxels at the coarsest level."""
        self.check_shape(shape)
        nb, d, h, w = shape
        if (nb * d * h * w // self.divisor ** 3) % 32:
            raise ValueError("training needs a multiple of 32 voxels at the coarsest level")


def _torch():
    import torch
    if not torch.cuda.is_available():
        raise N.NativeError("the warm-start network runs on a HIP device; none is visible")
    return torch


def _ptr(t, ch_offset=0):
    """Device address of a tensor, offset by ch_offset elements."""
    return C.c_void_p(t.data_ptr() + ch_offset * t.element_size())


class UNet:
    """Device parameters, optimizer state and activation buffers of one U-Net."""

    def __init__(self, spec=None, params=None, seed=0):
        torch = _torch()
        N.net()  # fail loudly if the kernels are not built
        self.spec = spec or UNetSpec()
        self.torch = torch
        dev = torch.device("cuda")
        p = self.spec.init_params(seed) if params is None else np.asarray(params, dtype=np.float32)
        if p.shape != (self.spec.nparams,):
            raise ValueError(f"expected {self.spec.nparams} parameters, got {p.shape}")
        self.p = torch.from_numpy(p.copy()).to(dev)
        self.g = torch.zeros_like(self.p)
        self.m = torch.zeros_like(self.p)
        self.v = torch.zeros_like(self.p)
        self.t = 0
        self.wt = {lay.name: torch.zeros((lay.cout, lay.kpad), dtype=torch.bfloat16, device=dev)
                   for lay in self.spec.layers}
        self.wd = {lay.name: torch.zeros((lay.cin, lay.kpad_d), dtype=torch.bfloat16, device=dev)
                   for lay in self.spec.layers}
        self._bufs = None
        self._shape = None
        self.pack()

    # -- plumbing ---------------------------------------------------------------------------

    def _st(self):
        return C.c_void_p(self.torch.cuda.current_stream().cuda_stream)

    def pack(self):
        L = N.net()
        for lay in self.spec.layers:
            N.net_check(L.pvnet_pack_weights(_ptr(self.p, lay.offset), lay.ntaps, lay.cin, lay.cout,
                                              _ptr(self.wt[lay.name]), _ptr(self.wd[lay.name]), self._st()))

    def params_numpy(self):
        return self.p.detach().cpu().numpy().copy()

    def buffers(self, shape):
        """Activation and gradient buffers for a batch shape (nb, d, h, w), kept between calls."""
        shape = tuple(int(v) for v in shape)
        if self._shape == shape:
            return self._bufs
        self.spec.check_shape(shape)
        torch = self.torch
        dev = torch.device("cuda")
        lv = self.spec.levels
        L = len(lv)
        nb, d, h, w = shape
        M = [nb * (d >> k) * (h >> k) * (w >> k) for k in range(L)]

        def bf(m, c):
            return torch.zeros((m, c), dtype=torch.bfloat16, device=dev)

        b = {"shape": shape, "M": M,
             "s": [N.NetShape(nb, d >> k, h >> k, w >> k) for k in range(L)],
             "X0": bf(M[0], IN_CH), "Y": torch.zeros((M[0], OUT_CH), dtype=torch.float32, device=dev),
             "dist": torch.zeros(M[0], dtype=torch.float32, device=dev),
             "dY": bf(M[0], OUT_CH)}
        for k in range(L - 1):
            ccat = lv[k + 1] + lv[k]
            b[f"A{k}"], b[f"dA{k}"] = bf(M[k], lv[k]), bf(M[k], lv[k])
            b[f"CAT{k}"], b[f"dCAT{k}"] = bf(M[k], ccat), bf(M[k], ccat)
            b[f"DA{k}"], b[f"dDA{k}"] = bf(M[k], lv[k]), bf(M[k], lv[k])
            b[f"DEC{k}"], b[f"dDEC{k}"] = bf(M[k], lv[k]), bf(M[k], lv[k])
            b[f"P{k + 1}"], b[f"dP{k + 1}"] = bf(M[k + 1], lv[k]), bf(M[k + 1], lv[k])
        b["BA"], b["dBA"] = bf(M[L - 1], lv[L - 1]), bf(M[L - 1], lv[L - 1])
        b[f"DEC{L - 1}"], b[f"dDEC{L - 1}"] = bf(M[L - 1], lv[L - 1]), bf(M[L - 1], lv[L - 1])
        ws = 0
        for lay in self.spec.layers:
            k = self._level(lay)
            ws = max(ws, N.net().pvnet_wgrad_ws_floats(C.byref(b["s"][k]), lay.ntaps, lay.cin, lay.cout))
        b["ws"] = torch.zeros(int(ws), dtype=torch.float32, device=dev)
        b["lpart"] = torch.zeros(int(N.net().pvnet_loss_ws_doubles(M[0])), dtype=torch.float64, device=dev)
        b["loss"] = torch.zeros(2, dtype=torch.float64, device=dev)
        self._bufs, self._shape = b, shape
        return b

    def _level(self, lay):
        n = lay.name
        L = len(self.spec.levels)
        if n == "head":
            return 0
        if n.startswith("bt"):
            return L - 1
        return int(n[1:-1])

    def _conv(self, lay, s, x, ldx, y, ldy, relu=None, f32=False, bias=True, xoff=0, yoff=0):
        flags = (N.NET_RELU if (lay.relu if relu is None else relu) else 0) | (N.NET_F32OUT if f32 else 0)
        c = N.NetConv(s, lay.ntaps, lay.cin, lay.cout, lay.kpad, _ptr(x, xoff), ldx, flags,
                      _ptr(self.wt[lay.name]), _ptr(self.p, lay.boff) if bias else None,
                      _ptr(y, yoff), ldy, 0, None)
        N.net_check(N.net().pvnet_conv_fwd(C.byref(c), self._st()))

    def _dgrad(self, lay, s, dy, lddy, dx, lddx, relu_of=None, ldr=0, dyoff=0, roff=0):
        c = N.NetConv(s, lay.ntaps, lay.cout, lay.cin, lay.kpad_d, _ptr(dy, dyoff), lddy, 0,
                      _ptr(self.wd[lay.name]), None, _ptr(dx), lddx, ldr,
                      _ptr(relu_of, roff) if relu_of is not None else None)
        N.net_check(N.net().pvnet_conv_fwd(C.byref(c), self._st()))

    def _wgrad(self, lay, s, x, ldx, dy, lddy, xoff=0, dyoff=0):
        b = self._bufs
        g = N.NetWgrad(s, lay.ntaps, lay.cin, lay.cout, _ptr(x, xoff), ldx, lddy, _ptr(dy, dyoff),
                       _ptr(b["ws"]), _ptr(self.g, lay.offset), _ptr(self.g, lay.boff))
        N.net_check(N.net().pvnet_conv_wgrad(C.byref(g), self._st()))

    # -- forward / backward -----------------------------------------------------------------

    def forward(self, code, prm):
        """code: uint8 device tensor (nb, d, h, w); prm: float32 device tensor (nb, 3) of
        (u_mean, tau, inlet_area).  Returns the FP32 output (nb*d*h*w, 8) (a buffer reused by
        later calls)."""
        b = self.buffers(tuple(code.shape))
        if tuple(prm.shape) != (code.shape[0], 3) or prm.dtype != self.torch.float32 or not prm.is_contiguous():
            raise ValueError("prm must be a contiguous float32 tensor (nb, 3)")
        NL = N.net()
        lv, L, S = self.spec.levels, len(self.spec.levels), b["s"]
        by = self.spec.by_name
        N.net_check(NL.pvnet_wall_distance(C.byref(S[0]), _ptr(code), _ptr(b["dist"]), self._st()))
        N.net_check(NL.pvnet_features(C.byref(S[0]), _ptr(code), _ptr(prm), _ptr(b["dist"]), _ptr(b["X0"]),
                                      self._st()))
        x, ldx = b["X0"], IN_CH
        for k in range(L - 1):
            ccat = lv[k + 1] + lv[k]
            self._conv(by[f"e{k}a"], S[k], x, ldx, b[f"A{k}"], lv[k])
            self._conv(by[f"e{k}b"], S[k], b[f"A{k}"], lv[k], b[f"CAT{k}"], ccat, yoff=lv[k + 1])
            N.net_check(NL.pvnet_pool2_fwd(C.byref(S[k]), lv[k], _ptr(b[f"CAT{k}"], lv[k + 1]), ccat,
                                           _ptr(b[f"P{k + 1}"]), lv[k], self._st()))
            x, ldx = b[f"P{k + 1}"], lv[k]
        self._conv(by["bta"], S[L - 1], x, ldx, b["BA"], lv[L - 1])
        self._conv(by["btb"], S[L - 1], b["BA"], lv[L - 1], b[f"DEC{L - 1}"], lv[L - 1])
        for k in range(L - 2, -1, -1):
            ccat = lv[k + 1] + lv[k]
            N.net_check(NL.pvnet_up2_fwd(C.byref(S[k]), lv[k + 1], _ptr(b[f"DEC{k + 1}"]), lv[k + 1],
                                         _ptr(b[f"CAT{k}"]), ccat, self._st()))
            self._conv(by[f"d{k}a"], S[k], b[f"CAT{k}"], ccat, b[f"DA{k}"], lv[k])
            self._conv(by[f"d{k}b"], S[k], b[f"DA{k}"], lv[k], b[f"DEC{k}"], lv[k])
        self._conv(by["head"], S[0], b["DEC0"], lv[0], b["Y"], OUT_CH, f32=True)
        return b["Y"]

    def loss(self, target, code, weights=(1.0, 1.0, 1.0, 1.0)):
        """Masked MSE of the last forward output against target (bf16 device tensor (M, 4));
        leaves dL/dY in the dY buffer and returns the device tensor [loss, n_fluid]."""
        b = self._bufs
        wc = self.torch.tensor(weights, dtype=self.torch.float32, device="cuda")
        M = b["M"][0]
        N.net_check(N.net().pvnet_mse_loss(M, N_TARGET, _ptr(b["Y"]), OUT_CH, _ptr(target), _ptr(code),
                                           _ptr(wc), _ptr(b["dY"]), OUT_CH, _ptr(b["lpart"]),
                                           _ptr(b["loss"]), self._st()))
        self._wc = wc  # keep alive until the stream has used it
        return b["loss"]

    def backward(self):
        """Gradients of every parameter from the dY buffer (set by loss()) into self.g."""
        b = self._bufs
        self.spec.check_train_shape(b["shape"])
        NL = N.net()
        lv, L, S = self.spec.levels, len(self.spec.levels), b["s"]
        by = self.spec.by_name
        head = by["head"]
        self._wgrad(head, S[0], b["DEC0"], lv[0], b["dY"], OUT_CH)
        self._dgrad(head, S[0], b["dY"], OUT_CH, b["dDEC0"], lv[0], relu_of=b["DEC0"], ldr=lv[0])
        for k in range(L - 1):
            ccat = lv[k + 1] + lv[k]
            la, lb = by[f"d{k}a"], by[f"d{k}b"]
            self._wgrad(lb, S[k], b[f"DA{k}"], lv[k], b[f"dDEC{k}"], lv[k])
            self._dgrad(lb, S[k], b[f"dDEC{k}"], lv[k], b[f"dDA{k}"], lv[k], relu_of=b[f"DA{k}"], ldr=lv[k])
            self._wgrad(la, S[k], b[f"CAT{k}"], ccat, b[f"dDA{k}"], lv[k])
            self._dgrad(la, S[k], b[f"dDA{k}"], lv[k], b[f"dCAT{k}"], ccat)
            N.net_check(NL.pvnet_up2_bwd(C.byref(S[k]), lv[k + 1], _ptr(b[f"dCAT{k}"]), ccat,
                                         _ptr(b[f"dDEC{k + 1}"]), lv[k + 1], self._st()))
            N.net_check(NL.pvnet_relu_bwd(b["M"][k + 1], lv[k + 1], _ptr(b[f"DEC{k + 1}"]), lv[k + 1],
                                          _ptr(b[f"dDEC{k + 1}"]), lv[k + 1], self._st()))
        ta, tb = by["bta"], by["btb"]
        top = L - 1
        self._wgrad(tb, S[top], b["BA"], lv[top], b[f"dDEC{top}"], lv[top])
        self._dgrad(tb, S[top], b[f"dDEC{top}"], lv[top], b["dBA"], lv[top], relu_of=b["BA"], ldr=lv[top])
        xin = b[f"P{top}"]
        self._wgrad(ta, S[top], xin, lv[top - 1], b["dBA"], lv[top])
        self._dgrad(ta, S[top], b["dBA"], lv[top], b[f"dP{top}"], lv[top - 1])
        for k in range(L - 2, -1, -1):
            ccat = lv[k + 1] + lv[k]
            ea, eb = by[f"e{k}a"], by[f"e{k}b"]
            N.net_check(NL.pvnet_pool2_bwd_relu(C.byref(S[k]), lv[k], _ptr(b[f"dP{k + 1}"]), lv[k],
                                                _ptr(b[f"dCAT{k}"], lv[k + 1]), ccat, _ptr(b[f"CAT{k}"], lv[k + 1]),
                                                ccat, self._st()))
            self._wgrad(eb, S[k], b[f"A{k}"], lv[k], b[f"dCAT{k}"], ccat, dyoff=lv[k + 1])
            self._dgrad(eb, S[k], b[f"dCAT{k}"], ccat, b[f"dA{k}"], lv[k], relu_of=b[f"A{k}"], ldr=lv[k],
                        dyoff=lv[k + 1])
            xk, ldk = (b["X0"], IN_CH) if k == 0 else (b[f"P{k}"], lv[k - 1])
            self._wgrad(ea, S[k], xk, ldk, b[f"dA{k}"], lv[k])
            if k > 0:
                self._dgrad(ea, S[k], b[f"dA{k}"], lv[k], b[f"dP{k}"], lv[k - 1])

    def adam_step(self, lr, b1=0.9, b2=0.999, eps=1e-8):
        self.t += 1
        N.net_check(N.net().pvnet_adam(self.spec.nparams, _ptr(self.p), _ptr(self.g), _ptr(self.m), _ptr(self.v),
                                       lr, b1, b2, eps, self.t, self._st()))
        self.pack()

