"""CNN M3D predictor, restated in NumPy float64 (TEST INFRASTRUCTURE ONLY).

Follows cnn/cnn/predict/predict.py:16-120 and compmask.py:17-41:

* ``compmask``: the 4 x 4 block means of the 2048 x 2048 mask bits
  (rowmat . mask0 . colmat, compmask.py:18-40), as float64;
* ``compmask_literal``: the same with the reference's float16 matrices and
  dots, for small sizes (the pin of ``compmask``);
* ``conv_block``: Conv2d(3x3, padding=1, padding_mode='circular') + bias,
  ReLU, MaxPool2d(2, 2), BatchNorm2d in eval mode (predict.py:36-45);
* ``forward``: five blocks, flatten, fc1 + ReLU, fc2 (predict.py:34-51);
* ``predict_m3d``: the six networks and the division by the factors
  (predict.py:108-120);
* ``torch_model``: the reference's CNNmodel class restated with torch.nn on
  the CPU (what predict.py runs), the independent pin of ``forward`` and the
  CPU baseline of the predictor.

Parameters use the reference's state_dict names (conv1.weight, norm1.running_mean,
fc1.weight, ...).
"""
import numpy as np

BN_EPS = 1e-5  # torch.nn.BatchNorm2d default


def compmask(bits, ndiv=2048, nout=512):
    """mask.bin bytes of one clip -> [nout][nout] block means (row = mask.bin row)."""
    m = np.unpackbits(np.frombuffer(bits, dtype=np.uint8)[: ndiv * ndiv // 8]).astype(np.float64)
    f = ndiv // nout
    return m.reshape(nout, f, nout, f).mean(axis=(1, 3))


def compmask_literal(bits, ndiv, nout):
    """compmask.py:17-41 with its float16 rowmat/columnmat and dots."""
    rowmat = np.zeros((nout, ndiv), dtype="float16")
    bin_size = ndiv / nout
    next_bin_break = bin_size
    which_row = 0
    which_column = 0
    while (which_row < nout) & (which_column < ndiv):
        if (next_bin_break - which_column) >= 1:
            rowmat[which_row, which_column] = 1 / bin_size
            which_column += 1
        elif abs(next_bin_break - which_column) < 0.0000001:
            which_row += 1
            next_bin_break += bin_size
    columnmat = rowmat.T.copy()
    row = np.unpackbits(np.frombuffer(bits, dtype=np.uint8)[: ndiv * ndiv // 8])
    mask0 = row.astype("float16").reshape(ndiv, ndiv)
    return np.dot(rowmat, np.dot(mask0, columnmat))


def conv3x3_circular(x, w, b):
    """x [cin][H][W], w [cout][cin][3][3] -> [cout][H][W] (cross-correlation,
    circular padding 1, as torch Conv2d)."""
    cin, H, W = x.shape
    xp = np.pad(x, ((0, 0), (1, 1), (1, 1)), mode="wrap")
    out = np.zeros((w.shape[0], H, W))
    for dy in range(3):
        for dx in range(3):
            out += np.tensordot(w[:, :, dy, dx], xp[:, dy:dy + H, dx:dx + W], axes=(1, 0))
    return out + b[:, None, None]


def conv_block(x, p, i):
    """x [cin][H][W] -> norm_i(pool(relu(conv_i(x))))."""
    y = conv3x3_circular(x, p[f"conv{i}.weight"], p[f"conv{i}.bias"])
    y = np.maximum(y, 0.0)
    c, H, W = y.shape
    y = y.reshape(c, H // 2, 2, W // 2, 2).max(axis=(2, 4))
    mean, var = p[f"norm{i}.running_mean"], p[f"norm{i}.running_var"]
    g, beta = p[f"norm{i}.weight"], p[f"norm{i}.bias"]
    return (y - mean[:, None, None]) / np.sqrt(var[:, None, None] + BN_EPS) * g[:, None, None] + beta[:, None, None]


def forward(p, x):
    """x [B][1][512][512] -> [B][n_out]."""
    outs = []
    for xb in x:
        h = xb
        for i in range(1, 6):
            h = conv_block(h, p, i)
        f = h.reshape(-1)
        f = np.maximum(p["fc1.weight"] @ f + p["fc1.bias"], 0.0)
        outs.append(p["fc2.weight"] @ f + p["fc2.bias"])
    return np.array(outs)


def predict_m3d(x, models, factor0, factorx, factory):
    """models: six parameter dicts (re0, im0, rex, imx, rey, imy)."""
    f = [factor0, factor0, factorx, factorx, factory, factory]
    return [forward(p, x) / fac for p, fac in zip(models, f)]


def random_params(n_out, seed, scale=1.0, nin=512, dtype=np.float64):
    """Seeded parameters of the reference's architecture for an nin x nin
    input (512 in the reference: fc1 takes 16 x 16 x 256), uniform like
    torch's default fan-in init, random BatchNorm statistics."""
    rng = np.random.default_rng(seed)
    p = {}
    chans = [1, 16, 32, 64, 128, 256]
    for i in range(1, 6):
        cin, cout = chans[i - 1], chans[i]
        bound = scale / np.sqrt(cin * 9)
        p[f"conv{i}.weight"] = rng.uniform(-bound, bound, (cout, cin, 3, 3))
        p[f"conv{i}.bias"] = rng.uniform(-bound, bound, cout)
        p[f"norm{i}.weight"] = rng.uniform(0.5, 1.5, cout)
        p[f"norm{i}.bias"] = rng.uniform(-0.1, 0.1, cout)
        p[f"norm{i}.running_mean"] = rng.uniform(-0.1, 0.1, cout)
        p[f"norm{i}.running_var"] = rng.uniform(0.5, 1.5, cout)
    side = nin // 32
    for name, (i, o) in (("fc1", (side * side * 256, 4096)), ("fc2", (4096, n_out))):
        bound = scale / np.sqrt(i)
        p[f"{name}.weight"] = rng.random((o, i), dtype=np.float32 if dtype == np.float32 else np.float64)
        p[f"{name}.weight"] *= 2 * bound
        p[f"{name}.weight"] -= bound
        p[f"{name}.bias"] = rng.uniform(-bound, bound, o)
    return {k: np.asarray(v, dtype=dtype) for k, v in p.items()}


def torch_model(p, n_out, nin=512, dtype="float64"):
    """The reference's CNNmodel (predict.py:16-51) built from torch.nn on the
    CPU with the parameters p, in eval mode."""
    import torch
    nn, F = torch.nn, torch.nn.functional

    class CNNmodel(nn.Module):
        def __init__(self):
            super().__init__()
            self.conv1 = nn.Conv2d(1, 16, (3, 3), padding=1, padding_mode='circular')
            self.conv2 = nn.Conv2d(16, 32, (3, 3), padding=1, padding_mode='circular')
            self.conv3 = nn.Conv2d(32, 64, (3, 3), padding=1, padding_mode='circular')
            self.conv4 = nn.Conv2d(64, 128, (3, 3), padding=1, padding_mode='circular')
            self.conv5 = nn.Conv2d(128, 256, (3, 3), padding=1, padding_mode='circular')
            self.norm1, self.norm2, self.norm3 = nn.BatchNorm2d(16), nn.BatchNorm2d(32), nn.BatchNorm2d(64)
            self.norm4, self.norm5 = nn.BatchNorm2d(128), nn.BatchNorm2d(256)
            self.pool = nn.MaxPool2d(2, 2)
            self.fc1 = nn.Linear((nin // 32) ** 2 * 256, 4096)
            self.fc2 = nn.Linear(4096, n_out)

        def forward(self, x):
            x = x.to(self.fc1.weight.dtype)
            for i in range(1, 6):
                x = getattr(self, f"norm{i}")(self.pool(F.relu(getattr(self, f"conv{i}")(x))))
            x = torch.flatten(x, 1)
            return self.fc2(F.relu(self.fc1(x)))

    m = CNNmodel().to(getattr(torch, dtype))
    m.load_state_dict({k: torch.from_numpy(np.asarray(v)).to(getattr(torch, dtype)) for k, v in p.items()},
                      strict=False)
    return m.eval()
