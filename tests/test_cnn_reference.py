"""Pins of the numpy restatement of the initial-guess network (cnn_reference):
the convolution against a brute-force loop, and every gradient against
central finite differences of the loss."""
import numpy as np

import cnn_reference as cnn


def test_conv_matches_brute_force():
    rng = np.random.default_rng(1)
    x = rng.standard_normal((2, 4, 3, 5))
    w = rng.standard_normal((27, 3, 2))
    b = rng.standard_normal(3)
    y = cnn.conv(x, w, b)
    nx, ny, nz = x.shape[1:]
    for co in range(3):
        for i in range(nx):
            for j in range(ny):
                for k in range(nz):
                    s = b[co]
                    for t, (dx, dy, dz) in enumerate(cnn.TAPS):
                        kk = k + dz
                        if 0 <= kk < nz:
                            s += w[t, co] @ x[:, (i + dx) % nx, (j + dy) % ny, kk]
                    assert abs(s - y[co, i, j, k]) < 1e-12


def test_gradients_match_finite_differences():
    rng = np.random.default_rng(2)
    params = cnn.init_params(seed=3)
    # a smaller network of the same form keeps the check fast
    params = [(w[:, : min(w.shape[1], 4), : min(w.shape[2], 4)].copy(), b[: min(b.shape[0], 4)].copy())
              for w, b in params]
    params[0] = (params[0][0][:, :, :3], params[0][1])
    params[-1] = (params[-1][0][:, :1, :], params[-1][1][:1])
    x = rng.standard_normal((3, 4, 3, 4))
    target = rng.standard_normal((1, 4, 3, 4))
    loss, grads = cnn.loss_and_grads(params, x, target)
    h = 1e-6
    for n, (w, b) in enumerate(params):
        for arr, g in ((w, grads[n][0]), (b, grads[n][1])):
            flat = arr.reshape(-1)
            gflat = g.reshape(-1)
            for q in rng.choice(flat.size, size=min(6, flat.size), replace=False):
                keep = flat[q]
                flat[q] = keep + h
                lp, _ = cnn.loss_and_grads(params, x, target)
                flat[q] = keep - h
                lm, _ = cnn.loss_and_grads(params, x, target)
                flat[q] = keep
                fd = (lp - lm) / (2 * h)
                assert abs(fd - gflat[q]) <= 1e-6 * max(1.0, abs(fd)), (n, q, fd, gflat[q])
