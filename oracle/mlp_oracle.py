"""Float64 numpy restatement of the cost-model network (mlp.hip): forward
pass and analytic gradients of 0.5 * mean (y - t)^2 for the packed
parameters W1[nhid][nin], b1, W2[nhid][nhid], b2, W3[nhid], b3.

TEST INFRASTRUCTURE ONLY: the checker of the HIP forward/backward kernels.
The backward pass is plain backpropagation (chain rule per layer); it is
itself pinned in tests by central finite differences.
"""

import numpy as np


def unpack(p, nin, nhid):
    o = 0
    out = []
    for shape in ((nhid, nin), (nhid,), (nhid, nhid), (nhid,), (nhid,), (1,)):
        n = int(np.prod(shape))
        out.append(np.asarray(p[o:o + n], np.float64).reshape(shape))
        o += n
    return out


def forward(p, X, nin, nhid):
    W1, b1, W2, b2, W3, b3 = unpack(p, nin, nhid)
    z1 = X @ W1.T + b1
    a1 = np.maximum(z1, 0)
    z2 = a1 @ W2.T + b2
    a2 = np.maximum(z2, 0)
    y = a2 @ W3 + b3[0]
    return y, (X, z1, a1, z2, a2)


def loss_grad(p, X, t, nin, nhid):
    W1, b1, W2, b2, W3, b3 = unpack(p, nin, nhid)
    y, (X, z1, a1, z2, a2) = forward(p, X, nin, nhid)
    B = X.shape[0]
    d = y - t
    loss = 0.5 * np.mean(d ** 2)
    dy = d / B
    gW3 = a2.T @ dy
    gb3 = np.array([dy.sum()])
    d2 = np.outer(dy, W3) * (z2 > 0)
    gW2 = d2.T @ a1
    gb2 = d2.sum(0)
    d1 = (d2 @ W2) * (z1 > 0)
    gW1 = d1.T @ X
    gb1 = d1.sum(0)
    g = np.concatenate([gW1.ravel(), gb1, gW2.ravel(), gb2, gW3, gb3])
    return loss, g
