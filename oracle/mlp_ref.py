"""FP64 reference of the emulator MLP (test infrastructure; the product runs
the BF16 MFMA kernels of irsounderlbl.jl_amd/csrc/k_mlp.hip).  The network has
no counterpart in the reference: its oracle is this plain numpy restatement
of the same network, loss and backpropagation, pinned against central finite
differences in tests/test_oracle_mlp.py.

Layout matches the library's packed parameters: W1[n_in_pad][H], b1[H],
W2[H][H], b2, W3[H][H], b3, W4[H][n_out], b4.  With emulate_bf16=True the
same values are rounded to bfloat16 (round to nearest even) at the points
where the GPU stores bf16: the input, the weights, each hidden activation,
the output gradient and each hidden pre-activation gradient."""
import numpy as np
from scipy.special import erf


def bf16(x):
    u = np.ascontiguousarray(x, np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def gelu(z):
    return 0.5 * z * (1.0 + erf(z / np.sqrt(2.0)))


def gelu_d(z):
    return 0.5 * (1.0 + erf(z / np.sqrt(2.0))) + z * np.exp(-0.5 * z * z) / np.sqrt(2.0 * np.pi)


class MLPRef:
    def __init__(self, params, n_in, n_in_pad, H, n_out):
        self.n_in, self.n_in_pad, self.H, self.n_out = n_in, n_in_pad, H, n_out
        dims = [n_in_pad, H, H, H, n_out]
        self.dims = dims
        p = np.asarray(params, np.float64)
        self.W, self.b, o = [], [], 0
        for l in range(4):
            self.W.append(p[o:o + dims[l] * dims[l + 1]].reshape(dims[l], dims[l + 1]))
            o += dims[l] * dims[l + 1]
            self.b.append(p[o:o + dims[l + 1]])
            o += dims[l + 1]
        self.n_params = o

    def pack(self, W, b):
        return np.concatenate([np.concatenate([W[l].ravel(), b[l]]) for l in range(4)])

    def forward(self, X, emulate_bf16=False):
        r = bf16 if emulate_bf16 else (lambda a: np.asarray(a, np.float64))
        x = np.zeros((len(X), self.n_in_pad))
        x[:, :self.n_in] = X
        x = r(x)
        W = [r(w) for w in self.W]
        acts, Z = [x], []
        h = x
        for l in range(3):
            z = h @ W[l] + self.b[l]
            Z.append(z)
            h = r(gelu(z))
            acts.append(h)
        y = h @ W[3] + self.b[3]
        self._cache = (acts, Z, W, r)
        return y

    def loss_grad(self, X, T, emulate_bf16=False):
        y = self.forward(X, emulate_bf16)
        acts, Z, W, r = self._cache
        B = len(X)
        d = y - T
        loss = np.mean(d * d)
        g = r(2.0 * d / (B * self.n_out))
        gW, gb = [None] * 4, [None] * 4
        for l in range(3, -1, -1):
            gW[l] = acts[l].T @ g
            gb[l] = g.sum(axis=0)
            if l > 0:
                g = r((g @ W[l].T) * gelu_d(Z[l - 1]))
        return loss, self.pack(gW, gb)

    def input_jacobian(self, X):
        """dy/dx [n, n_out, n_in] in FP64."""
        self.forward(X)
        _, Z, W, _ = self._cache
        J = np.broadcast_to(W[0][:self.n_in].T[None], (len(X), self.H, self.n_in)) * gelu_d(Z[0])[:, :, None]
        for l in range(1, 3):
            J = np.einsum("hk,bhi->bki", W[l], J) * gelu_d(Z[l])[:, :, None]
        return np.einsum("ho,bhi->boi", W[3], J)
