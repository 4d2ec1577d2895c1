"""The CNN M3D predictor on the GPU: the reference's predict.py with the same
classes and call, running on libeuvl's kernels.

Reference: cnn/cnn/predict/predict.py:16-151 and compmask.py:17-41.  Six
networks (CNNmodel: 1901 outputs for re0 / im0; CNNmodels: 1749 outputs for
rex, imx, rey, imy), each five conv blocks (3 x 3 circular conv, ReLU, 2 x 2
max pool, BatchNorm in eval mode; 1 -> 16 -> 32 -> 64 -> 128 -> 256
channels, 512 -> 16 pixels) and two fully connected layers (65536 -> 4096,
ReLU, -> outputs).  FP32 throughout, as the reference computes.

    mask = compmask("mask.bin")                       # device [n, 1, 512, 512]
    m = CNNmodel.load_from_checkpoint("../model/re0/model.ckpt")
    re0, im0, rex, imx, rey, imy = predict_m3d(mask, m_re0, m_im0, m_rex, m_imx, m_rey, m_imy,
                                               factor0, factorx, factory)
"""
import ctypes
import threading

import numpy as np
import torch

from . import _lib

NCUT, NCUTS = 1901, 1749      # predict.py:11-12
NDIV, NDIVM = 2048, 512       # compmask.py:14-15
CHANNELS = (1, 16, 32, 64, 128, 256)
BN_EPS = 1e-5                 # torch.nn.BatchNorm2d default

_local = threading.local()


def _handle(device):
    """A libeuvl handle per host thread, bound to torch's current stream."""
    h = getattr(_local, "h", None)
    lib = _lib.load()
    if h is None:
        h = ctypes.c_void_p()
        with torch.cuda.device(device):
            _lib.check(lib.euvl_create(ctypes.byref(h)), "euvl_create")
        _local.h = h
    s = torch.cuda.current_stream(device).cuda_stream
    _lib.check(lib.euvl_set_stream(h, ctypes.c_void_p(s)), "euvl_set_stream")
    return lib, h


def compmask(mask, ndiv=NDIV, nout=NDIVM, device="cuda:0"):
    """compmask.py:17-41 for every clip of a mask.bin (path, bytes or a device
    uint8 tensor of packed bits): device float32 [n, 1, nout, nout], the mean
    of each (ndiv/nout)^2 block, row 0 = the first mask.bin row."""
    if isinstance(mask, str):
        with open(mask, "rb") as f:
            mask = f.read()
    if isinstance(mask, (bytes, bytearray)):
        mask = torch.from_numpy(np.frombuffer(mask, dtype=np.uint8).copy()).to(device)
    per = ndiv * ndiv // 8
    n = mask.numel() // per
    if n < 1:
        raise ValueError(f"mask holds {mask.numel()} bytes, less than one {ndiv}x{ndiv} clip")
    out = torch.empty((n, 1, nout, nout), dtype=torch.float32, device=mask.device)
    lib, h = _handle(mask.device)
    _lib.check(lib.euvl_cnn_compmask(h, _lib.ptr(mask.contiguous()), ndiv, n, nout, _lib.ptr(out)),
               "euvl_cnn_compmask")
    return out


class _Network:
    """Parameters of one network on the device (names as the reference's
    state_dict); BatchNorm folded into a per-channel scale and shift."""

    n_out = NCUT

    def __init__(self, state, device="cuda:0"):
        dev = torch.device(device)

        def f32(a):
            if isinstance(a, torch.Tensor):
                return a.detach().to(device=dev, dtype=torch.float32).contiguous()
            return torch.as_tensor(np.asarray(a), dtype=torch.float32).contiguous().to(dev)

        self.device = dev
        self.conv = []
        for i in range(1, 6):
            w = f32(state[f"conv{i}.weight"])
            if tuple(w.shape) != (CHANNELS[i], CHANNELS[i - 1], 3, 3):
                raise ValueError(f"conv{i}.weight has shape {tuple(w.shape)}")
            g = np.asarray(_np(state[f"norm{i}.weight"]), dtype=np.float64)
            beta = np.asarray(_np(state[f"norm{i}.bias"]), dtype=np.float64)
            mean = np.asarray(_np(state[f"norm{i}.running_mean"]), dtype=np.float64)
            var = np.asarray(_np(state[f"norm{i}.running_var"]), dtype=np.float64)
            scale = g / np.sqrt(var + BN_EPS)
            self.conv.append((w, f32(state[f"conv{i}.bias"]), f32(scale), f32(beta - mean * scale)))
        self.fc1 = (f32(state["fc1.weight"]), f32(state["fc1.bias"]))
        self.fc2 = (f32(state["fc2.weight"]), f32(state["fc2.bias"]))
        if self.fc2[0].shape[0] != self.n_out:
            self.n_out = int(self.fc2[0].shape[0])

    @classmethod
    def load_from_checkpoint(cls, path, device="cuda:0"):
        """A PyTorch Lightning checkpoint (or a plain state_dict file), read
        with weights_only=True: nothing in the file is executed."""
        ck = torch.load(path, map_location="cpu", weights_only=True)
        return cls(ck.get("state_dict", ck), device)

    def __call__(self, x):
        return self.forward(x)

    def forward(self, x, timing=None):
        """x: device float32 [B, 1, n, n] (n a multiple of 32; 512 in the
        reference) -> device float32 [B, n_out]."""
        x = x.to(torch.float32).contiguous()
        B, _, H, W = x.shape
        lib, h = _handle(x.device)
        ev = []
        if timing is not None:  # events on the stream: start, after the conv blocks, end
            ev = [torch.cuda.Event(enable_timing=True) for _ in range(3)]
            ev[0].record()
        for i, (w, b, sc, sh) in enumerate(self.conv):
            cin, cout = CHANNELS[i], CHANNELS[i + 1]
            y = torch.empty((B, cout, H // 2, W // 2), dtype=torch.float32, device=x.device)
            _lib.check(lib.euvl_cnn_conv_block(h, B, cin, cout, H, W, _lib.ptr(x), _lib.ptr(w), _lib.ptr(b),
                                               _lib.ptr(sc), _lib.ptr(sh), _lib.ptr(y)), "euvl_cnn_conv_block")
            x, H, W = y, H // 2, W // 2
        if ev:
            ev[1].record()
        flat = x.reshape(B, -1)
        w1, b1 = self.fc1
        if flat.shape[1] != w1.shape[1]:
            raise ValueError(f"fc1 takes {w1.shape[1]} features, the conv stack gives {flat.shape[1]}")
        hid = torch.empty((B, w1.shape[0]), dtype=torch.float32, device=x.device)
        _lib.check(lib.euvl_cnn_linear(h, B, w1.shape[1], w1.shape[0], _lib.ptr(flat), _lib.ptr(w1), _lib.ptr(b1), 1,
                                       None, _lib.ptr(hid)), "euvl_cnn_linear")
        w2, b2 = self.fc2
        out = torch.empty((B, w2.shape[0]), dtype=torch.float32, device=x.device)
        _lib.check(lib.euvl_cnn_linear(h, B, w2.shape[1], w2.shape[0], _lib.ptr(hid), _lib.ptr(w2), _lib.ptr(b2), 0,
                                       None, _lib.ptr(out)), "euvl_cnn_linear")
        if ev:
            ev[2].record()
            timing.setdefault("events", []).append(ev)
        return out

    def eval(self):
        return self


def _np(a):
    return a.detach().cpu().numpy() if hasattr(a, "detach") else np.asarray(a)


class CNNmodel(_Network):
    """predict.py:16-51 (re0, im0: 1901 outputs)."""
    n_out = NCUT


class CNNmodels(_Network):
    """predict.py:53-88 (rex, imx, rey, imy: 1749 outputs)."""
    n_out = NCUTS


class Fullmodel:
    """predict.py:90-106."""

    def __init__(self, modelre0, modelim0, modelrex, modelimx, modelrey, modelimy):
        self.models = (modelre0, modelim0, modelrex, modelimx, modelrey, modelimy)

    def eval(self):
        return self

    def __call__(self, x):
        return tuple(m(x) for m in self.models)


def predict_m3d(mask, modelre0, modelim0, modelrex, modelimx, modelrey, modelimy, factor0, factorx, factory):
    """predict.py:108-120: the six outputs as float64 host arrays, each network
    output (float32) divided by its factor row as the reference divides (in
    float64)."""
    x1, x2, x3, x4, x5, x6 = Fullmodel(modelre0, modelim0, modelrex, modelimx, modelrey, modelimy).eval()(mask)
    f0, fx, fy = (np.asarray(f, dtype=np.float64) for f in (factor0, factorx, factory))
    h = lambda t: t.cpu().numpy()
    return (h(x1) / f0, h(x2) / f0, h(x3) / fx, h(x4) / fx, h(x5) / fy, h(x6) / fy)


def write_inputxx(path, arrays):
    """predict.py:145-151: np.savetxt of each array, one row per clip."""
    with open(path, "w") as f:
        for a in arrays:
            np.savetxt(f, a, delimiter=",")
