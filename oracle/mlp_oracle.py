"""TEST INFRASTRUCTURE ONLY: float64 numpy restatement of the warm-start MLP
(6 -> 64 -> 64 -> 1, SiLU) with its analytic loss gradient, used to check the HIP forward
and backward kernels.  Parameter blob layout as in include/cleo_hip.h:
W1[64][6], b1[64], W2[64][64], b2[64], W3[64], b3, xmean[6], xstd[6], ymean, ystd."""
import numpy as np

NIN, NH = 6, 64
NPARAM = NH * NIN + NH + NH * NH + NH + NH + 1


def unpack(blob):
    b = np.asarray(blob, dtype=np.float64)
    o = 0
    W1 = b[o:o + NH * NIN].reshape(NH, NIN); o += NH * NIN
    b1 = b[o:o + NH]; o += NH
    W2 = b[o:o + NH * NH].reshape(NH, NH); o += NH * NH
    b2 = b[o:o + NH]; o += NH
    W3 = b[o:o + NH]; o += NH
    b3 = b[o]; o += 1
    xm = b[o:o + NIN]; o += NIN
    xs = b[o:o + NIN]; o += NIN
    ym, ys = b[o], b[o + 1]
    return W1, b1, W2, b2, W3, b3, xm, xs, ym, ys


def bf16(a):
    """Round float32 values to bfloat16 (nearest even), returned as float64."""
    u = np.asarray(a, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def silu(z):
    return z / (1.0 + np.exp(-z))


def dsilu(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s * (1.0 + z * (1.0 - s))


def forward(blob, x, quantize=False):
    """Normalised network output for inputs x[n][6]. quantize=True rounds the weights of the two
    matrix layers, the normalised inputs and the first hidden activations to bf16, as the HIP
    kernels feed them to the matrix cores."""
    W1, b1, W2, b2, W3, b3, xm, xs, ym, ys = unpack(blob)
    xn = ((np.asarray(x, np.float64)[:, :NIN] - xm) / xs).astype(np.float32).astype(np.float64)
    if quantize:
        xn, W1, W2 = bf16(xn), bf16(W1), bf16(W2)
    z1 = xn @ W1.T + b1
    h1 = silu(z1)
    if quantize:
        h1 = bf16(h1)
    z2 = h1 @ W2.T + b2
    h2 = silu(z2)
    y = h2 @ W3 + b3
    return y, (xn, z1, h1, z2, h2)


def loss_grad(blob, x, y_target):
    """Mean squared error on normalised targets and its gradient w.r.t. the NPARAM parameters."""
    W1, b1, W2, b2, W3, b3, xm, xs, ym, ys = unpack(blob)
    t = (np.asarray(y_target, np.float64) - ym) / ys
    y, (xn, z1, h1, z2, h2) = forward(blob, x)
    n = len(t)
    dy = 2.0 * (y - t) / n
    gW3 = h2.T @ dy
    gb3 = dy.sum()
    dz2 = np.outer(dy, W3) * dsilu(z2)
    gW2 = dz2.T @ h1
    gb2 = dz2.sum(0)
    dz1 = (dz2 @ W2) * dsilu(z1)
    gW1 = dz1.T @ xn
    gb1 = dz1.sum(0)
    g = np.concatenate([gW1.ravel(), gb1, gW2.ravel(), gb2, gW3, [gb3]])
    return float(np.mean((y - t) ** 2)), g
