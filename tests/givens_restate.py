"""Numpy restatement of the Givens mask sweep (both precisions), shared by the CPU test that pins
it to the oracle in FP64 and the GPU test of the FP32 sweep."""
import numpy as np


def givens_numpy(x, pairs, cs, mask, is_left, dt):
    """The Givens sweep restated in numpy arithmetic of dtype dt: per pass the disjoint pairs, each
    entry a * c + b * s' with the rotation rounded to dt and every product and sum rounded to dt;
    for odd K the pair the plan appends goes after the others (mask) or before them (unmask)."""
    y = np.array(x if is_left else x.T, dt)
    K = y.shape[0]
    sg = dt(1.0 if mask else -1.0)
    for p in range(pairs.shape[0]):
        idx = np.arange(pairs.shape[1])
        groups = [idx]
        if K % 2 == 1:
            sp = pairs.shape[1] - 1 if mask else 0
            rest = idx[idx != sp]
            groups = [rest, np.array([sp])] if mask else [np.array([sp]), rest]
        for g in groups:
            r1, r2 = pairs[p, g, 0], pairs[p, g, 1]
            c = cs[p, g, 0].astype(dt)[:, None]
            s = cs[p, g, 1].astype(dt)[:, None]
            a, b = y[r1].copy(), y[r2].copy()
            y[r1] = a * c + b * (sg * s)
            y[r2] = a * (-sg * s) + b * c
    return y if is_left else y.T
