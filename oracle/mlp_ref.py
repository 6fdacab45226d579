"""CPU reference of the row-size network (float64 numpy): forward, loss and
gradients, used only to check the GPU kernels' gradients.

Network: x (6) -> h1 = relu(W1 x + b1) (64) -> h2 = relu(W2 h1 + b2) (64)
-> z = w3 . h2 + b3; pred = ub * sigmoid(z); loss = mean((log pred - log y)^2)
with ub and y clamped below at 1.  Parameter layout (flat): W1[64][6], b1,
W2[64][64], b2, w3[64], b3.
"""
import numpy as np

IN, H = 6, 64


def unpack(w):
    w = np.asarray(w, dtype=np.float64)
    o = 0
    W1 = w[o:o + H * IN].reshape(H, IN); o += H * IN
    b1 = w[o:o + H]; o += H
    W2 = w[o:o + H * H].reshape(H, H); o += H * H
    b2 = w[o:o + H]; o += H
    w3 = w[o:o + H]; o += H
    b3 = w[o]
    return W1, b1, W2, b2, w3, b3


def loss_and_grad(w, X, ub, y):
    W1, b1, W2, b2, w3, b3 = unpack(w)
    X = np.asarray(X, np.float64)
    n = X.shape[0]
    a1 = X @ W1.T + b1
    h1 = np.maximum(a1, 0)
    a2 = h1 @ W2.T + b2
    h2 = np.maximum(a2, 0)
    z = h2 @ w3 + b3
    logsig = -np.logaddexp(0.0, -z)
    e = np.log(np.maximum(ub, 1.0)) + logsig - np.log(np.maximum(y, 1.0))
    loss = np.mean(e * e)
    dz = 2 * e * (1.0 / (1.0 + np.exp(z))) / n
    gw3 = h2.T @ dz
    gb3 = dz.sum()
    d2 = np.outer(dz, w3) * (a2 > 0)
    gW2 = d2.T @ h1
    gb2 = d2.sum(0)
    d1 = (d2 @ W2) * (a1 > 0)
    gW1 = d1.T @ X
    gb1 = d1.sum(0)
    g = np.concatenate([gW1.ravel(), gb1, gW2.ravel(), gb2, gw3, [gb3]])
    return loss, g


def predict(w, X, ub):
    W1, b1, W2, b2, w3, b3 = unpack(w)
    h1 = np.maximum(np.asarray(X, np.float64) @ W1.T + b1, 0)
    h2 = np.maximum(h1 @ W2.T + b2, 0)
    z = h2 @ w3 + b3
    return np.asarray(ub, np.float64) / (1.0 + np.exp(-z))
