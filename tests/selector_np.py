"""numpy restatement of the learned precision selector's MLP (forward pass,
masked squared-error gradient, Adam step and the precision choice), the
checker of the library's HIP inference and training kernels. The weight
layout is the library's: W1[32][10], b1[32], W2[32][32], b2[32], W3[3][32],
b3[3], input mean[10], input 1/std[10]."""
import numpy as np

NF, NH, NO = 10, 32, 3
NP = NH * NF + NH + NH * NH + NH + NO * NH + NO


def unpack(w):
    w = np.asarray(w, np.float64)
    o = 0

    def take(k, shape=None):
        nonlocal o
        a = w[o:o + k]
        o += k
        return a.reshape(shape) if shape else a
    W1 = take(NH * NF, (NH, NF))
    b1 = take(NH)
    W2 = take(NH * NH, (NH, NH))
    b2 = take(NH)
    W3 = take(NO * NH, (NO, NH))
    b3 = take(NO)
    mu = take(NF)
    isd = take(NF)
    return W1, b1, W2, b2, W3, b3, mu, isd


def forward(w, X):
    W1, b1, W2, b2, W3, b3, mu, isd = unpack(w)
    x = (np.asarray(X, np.float64) - mu) * isd
    h1 = np.tanh(x @ W1.T + b1)
    h2 = np.tanh(h1 @ W2.T + b2)
    return h2 @ W3.T + b3, (x, h1, h2)


def loss_grad(w, X, Y, M=None):
    """(sum of M (y - t)^2, its gradient over the first NP weights)."""
    W1, b1, W2, b2, W3, b3, mu, isd = unpack(w)
    Y = np.asarray(Y, np.float64)
    M = np.ones_like(Y) if M is None else np.asarray(M, np.float64)
    y, (x, h1, h2) = forward(w, X)
    e = y - Y
    loss = float((M * e * e).sum())
    g3 = 2.0 * M * e
    g2 = (g3 @ W3) * (1.0 - h2 * h2)
    g1 = (g2 @ W2) * (1.0 - h1 * h1)
    grad = np.concatenate([(g1.T @ x).ravel(), g1.sum(0), (g2.T @ h1).ravel(), g2.sum(0), (g3.T @ h2).ravel(),
                           g3.sum(0)])
    return loss, grad


def adam(w, X, Y, M=None, steps=1, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8):
    """Adam (Kingma and Ba) on the mean of the masked squared error over
    the labels; the input scaling entries stay fixed."""
    w = np.array(w, np.float64, copy=True)
    Mk = np.ones_like(np.asarray(Y, np.float64)) if M is None else np.asarray(M, np.float64)
    cnt = Mk.sum()
    m = np.zeros(NP)
    v = np.zeros(NP)
    losses = []
    for t in range(1, steps + 1):
        loss, g = loss_grad(w, X, Y, Mk)
        g /= cnt
        m = b1 * m + (1 - b1) * g
        v = b2 * v + (1 - b2) * g * g
        w[:NP] -= lr * (m / (1 - b1 ** t)) / (np.sqrt(v / (1 - b2 ** t)) + eps)
        losses.append(loss / cnt)
    return w, np.array(losses)


def choose(y, tol, mask=(1, 2, 4)):
    """Precision codes (PREC_*) per row of predicted log10 errors y (FP32,
    split BF16, BF16): the cheapest allowed one whose prediction is <=
    log10(tol): BF16 (2), split BF16 (4), FP32 (1), else FP64 (0)."""
    lt = np.log10(tol)
    out = np.zeros(len(y), np.int8)
    for i, r in enumerate(np.asarray(y, np.float64)):
        if 2 in mask and r[2] <= lt:
            out[i] = 2
        elif 4 in mask and r[1] <= lt:
            out[i] = 4
        elif 1 in mask and r[0] <= lt:
            out[i] = 1
    return out
