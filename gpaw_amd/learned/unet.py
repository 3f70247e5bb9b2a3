"""The U-Net of the learned preconditioner: torch definition (training and
FP32 reference) and the HIP inference path (libgpawhip gh_unet_*).

Architecture (periodic 3x3x3 convolutions, channels 8 / 16 / 32 on three
levels, nearest 2x upsampling with additive skips):

  level 0 (N):   e0 = relu(c0(x))                     x = (s R, v_scale v)
  level 1 (N/2): e1 = relu(c1(pool(e0)))
  level 2 (N/4): e2 = relu(c2b(relu(c2a(pool(e1)))))
  level 1:       h1 = relu(c1u(up(m2(e2)) + e1))        m2: 1x1, 32 -> 16
  level 0:       h0 = relu(c0u(up(m1(h1)) + e0))        m1: 1x1, 16 -> 8
  out = head(h0) / s                                    head: 1x1, 8 -> 1

s = 1 / rms(R) per band (the map is applied to a unit-size residual and the
correction scaled back).  In the HIP path weights and activations are BF16
and every sum FP32; torch's forward(bf16=True) rounds the same tensors at
the same places (the reference the kernels are tested against).
"""
import numpy as np
import torch
import torch.nn as nn
import torch.nn.functional as F

from .. import _lib
from .._arrays import dtype_code, stream_ptr

VSCALE = 0.5   # v (Hartree) -> the second input channel

CONV3 = (('c0', 2, 8), ('c1', 8, 16), ('c2a', 16, 32), ('c2b', 32, 32),
         ('c1u', 16, 16), ('c0u', 8, 8))
MAP1 = (('m2', 32, 16), ('m1', 16, 8), ('head', 8, 1))


def _bf(t):
    return t.to(torch.bfloat16).to(torch.float32)


class UNet3(nn.Module):
    def __init__(self):
        super().__init__()
        for name, ci, co in CONV3:
            setattr(self, name, nn.Conv3d(ci, co, 3, padding=0))
        for name, ci, co in MAP1:
            setattr(self, name, nn.Conv3d(ci, co, 1))

    @staticmethod
    def _pad(x):
        return F.pad(x, (1, 1, 1, 1, 1, 1), mode='circular')

    def forward(self, x, bf16=False):
        """x: (B, 2, n0, n1, n2) = (s R, v_scale v); returns (B, 1, ...)."""
        q = _bf if bf16 else (lambda t: t)

        def conv(name, t, relu=True):
            m = getattr(self, name)
            w = q(m.weight) if bf16 else m.weight
            b = m.bias
            y = F.conv3d(self._pad(t), w, b)
            return q(F.relu(y) if relu else y)

        def map1(name, t):
            m = getattr(self, name)
            w = q(m.weight) if bf16 else m.weight
            return F.conv3d(t, w, m.bias)

        def up(t):
            return F.interpolate(t, scale_factor=2, mode='nearest')

        x = q(x)
        e0 = conv('c0', x)
        e1 = conv('c1', q(F.avg_pool3d(e0, 2)))
        e2 = conv('c2b', conv('c2a', q(F.avg_pool3d(e1, 2))))
        u = q(map1('m2', e2))
        h1 = conv('c1u', q(up(u) + e1))
        u = q(map1('m1', h1))
        h0 = conv('c0u', q(up(u) + e0))
        return map1('head', h0)


def save_weights(model, path):
    arrs = {}
    for name, p in model.state_dict().items():
        arrs[name.replace('.', '__')] = p.detach().cpu().numpy().astype(
            np.float32)
    np.savez(path, **arrs)


def load_model(path):
    model = UNet3()
    with np.load(path, allow_pickle=False) as f:
        sd = {k.replace('__', '.'): torch.from_numpy(f[k]) for k in f.files}
    model.load_state_dict(sd)
    return model


class HIPUNet:
    """Inference with the gh_unet_* kernels: BF16 weights (packed once),
    BF16 channel-last activations, FP32 sums; real FP32 or FP64 residuals."""

    def __init__(self, model):
        self._lib = _lib.lib()
        sd = {k: v.detach().to(torch.float32).cpu().contiguous()
              for k, v in model.state_dict().items()}
        self.conv = {}
        for name, ci, co in CONV3:
            W = np.ascontiguousarray(sd[name + '.weight'].numpy(), np.float32)
            if name == 'c0':
                # the first layer stages (R, v) as 4 channels, the last two
                # zero: its weights are packed for 4 input channels
                W = np.concatenate([W, np.zeros_like(W)], axis=1)
                ci = 4
            n = int(self._lib.gh_unet_conv3_weight_elems(ci, co))
            packed = np.zeros(n, np.uint16)
            _lib.check(self._lib.gh_unet_pack_conv3(
                ci, co, W.ctypes.data, packed.ctypes.data))
            self.conv[name] = (
                torch.from_numpy(packed.view(np.int16)).cuda(),
                sd[name + '.bias'].cuda())
        self.map1 = {}
        for name, ci, co in MAP1:
            # W[ci][co], rounded to BF16 values (kept as FP32 numbers)
            W = _bf(sd[name + '.weight'].reshape(co, ci)).T.contiguous()
            self.map1[name] = (W.cuda(), sd[name + '.bias'].cuda())
        self.head_bias = float(sd['head.bias'][0])
        self._buf = {}

    def _act(self, key, shape):
        n = int(np.prod(shape))
        t = self._buf.get(key)
        if t is None or t.numel() < n:
            t = torch.empty(n, dtype=torch.int16, device='cuda')
            self._buf[key] = t
        return t[:n]

    def _conv(self, name, ci, co, mode, nb, n_c, out, inp=None, skip=None,
              R=None, v=None, scale=None, relu=True, dtype=None):
        W, b = self.conv[name]
        n = np.ascontiguousarray(n_c, np.int64)
        _lib.check(self._lib.gh_unet_conv3(
            ci, co, mode, dtype_code(dtype) if dtype is not None else 0, nb,
            n.ctypes.data_as(_lib.PI64),
            inp.data_ptr() if inp is not None else None,
            skip.data_ptr() if skip is not None else None,
            R.data_ptr() if R is not None else None,
            v.data_ptr() if v is not None else None,
            scale.data_ptr() if scale is not None else None, VSCALE,
            W.data_ptr(), b.data_ptr(), int(relu), out.data_ptr(),
            stream_ptr()))

    def _pool(self, c, nb, n_c, inp, out):
        n = np.ascontiguousarray(n_c, np.int64)
        _lib.check(self._lib.gh_unet_pool2(c, nb, n.ctypes.data_as(_lib.PI64),
                                           inp.data_ptr(), out.data_ptr(),
                                           stream_ptr()))

    def _map(self, name, ci, co, npts, inp, out):
        W, b = self.map1[name]
        _lib.check(self._lib.gh_unet_map1(ci, co, npts, inp.data_ptr(),
                                          W.data_ptr(), b.data_ptr(),
                                          out.data_ptr(), stream_ptr()))

    def __call__(self, R, v, scale, out):
        """out[b] = f(R[b] * scale[b], v) / scale[b]; R, out: (nb, n0, n1,
        n2) float32 / float64; v: (n0, n1, n2) of the same dtype; scale:
        (nb,) float32."""
        nb = R.shape[0]
        n0 = tuple(int(x) for x in R.shape[1:])
        if any(x % 4 for x in n0):
            raise ValueError('the U-Net needs grid sizes divisible by 4')
        n1 = tuple(x // 2 for x in n0)
        n2 = tuple(x // 2 for x in n1)
        p0, p1, p2 = (nb * int(np.prod(n)) for n in (n0, n1, n2))
        e0 = self._act('e0', (p0, 8))
        self._conv('c0', 2, 8, 0, nb, n0, e0, R=R, v=v, scale=scale,
                   dtype=R.dtype)
        d1 = self._act('d1', (p1, 8))
        self._pool(8, nb, n0, e0, d1)
        e1 = self._act('e1', (p1, 16))
        self._conv('c1', 8, 16, 1, nb, n1, e1, inp=d1)
        d2 = self._act('d2', (p2, 16))
        self._pool(16, nb, n1, e1, d2)
        t2 = self._act('t2', (p2, 32))
        self._conv('c2a', 16, 32, 1, nb, n2, t2, inp=d2)
        e2 = self._act('e2', (p2, 32))
        self._conv('c2b', 32, 32, 1, nb, n2, e2, inp=t2)
        u2 = self._act('u2', (p2, 16))
        self._map('m2', 32, 16, p2, e2, u2)
        h1 = self._act('h1', (p1, 16))
        self._conv('c1u', 16, 16, 2, nb, n1, h1, inp=u2, skip=e1)
        u1 = self._act('u1', (p1, 8))
        self._map('m1', 16, 8, p1, h1, u1)
        h0 = self._act('h0', (p0, 8))
        self._conv('c0u', 8, 8, 2, nb, n0, h0, inp=u1, skip=e0)
        Wh, bh = self.map1['head']
        _lib.check(self._lib.gh_unet_head(
            8, dtype_code(out.dtype), nb, p0 // nb, h0.data_ptr(),
            Wh.data_ptr(), self.head_bias, scale.data_ptr(), out.data_ptr(),
            stream_ptr()))
        return out
