"""Learned bisection brackets on the GPU (excalibur_amd/csrc/learned.hip through the C ABI).

The pipeline generates its own training data: bidiagonals from the Golub-Kahan bidiagonalization
of matrices with spectra from four families (geometric, linear, clustered, Marchenko-Pastur), and
their singular values from the reference's bisection tree (exc_bisect_index). A 3-layer MLP
(32 -> 256 -> 256 -> 2) trained in FP32 by the library's forward/backward/Adam kernels, and run in
BF16, maps the counting function of a bidiagonal (exact Sturm counts at 29 shifts spaced half a
decade apart below its Frobenius norm) and a rank to a bracket for that singular value. Each
bracket is verified by two exact Sturm counts before the bisection uses it; one that fails falls
back to the reference's root, so the singular values are the reference's whatever the network
proposes, and only Sturm counts are saved.
"""
import ctypes
import time

import numpy as np

from . import hip

K_SHIFTS = 29          # counting-function features per bidiagonal
D_IN = 32              # K_SHIFTS + rank fraction + interpolated estimate + 1
Y_SCALE = 10.0         # target: ln(sigma / ||B||_F) / Y_SCALE
BRACKET_K = (1.0, 6.0)  # bracket ends below / above the prediction in Laplace scales (before calibration)
TOL = 1e-20            # the reference's bisection tolerance (client.cpp:984 -> bbtsvd.cpp:941)
FAMILIES = ("geometric", "linear", "clustered", "marchenko_pastur")

lib = hip.lib
DA = hip.DeviceArray


# ------------------------------------------------------------------ spectra and bidiagonals
def spectrum(family, n, rng):
    """n singular values (descending, max 1) of one of the four families."""
    i = np.arange(n)
    if family == "geometric":
        dec = rng.uniform(2.0, 8.0)
        s = 10.0 ** (-dec * i / (n - 1))
    elif family == "linear":
        lo = 10.0 ** rng.uniform(-4, -0.5)
        s = 1.0 - (1.0 - lo) * i / (n - 1)
    elif family == "clustered":
        k = int(rng.integers(2, 9))
        centers = np.sort(10.0 ** rng.uniform(-6, 0, k))[::-1]
        centers[0] = 1.0
        sizes = rng.multinomial(n - k, np.ones(k) / k) + 1
        parts = []
        for c, m in zip(centers, sizes):
            spread = 10.0 ** rng.uniform(-10, -4)
            parts.append(c * (1.0 + spread * np.linspace(-1, 1, m)))
        s = np.sort(np.concatenate(parts))[::-1]
    elif family == "marchenko_pastur":
        lam = rng.uniform(0.05, 1.0)
        a, b = (1 - np.sqrt(lam)) ** 2, (1 + np.sqrt(lam)) ** 2
        x = np.linspace(a, b, 20001)
        dens = np.sqrt(np.maximum((b - x) * (x - a), 0.0)) / np.maximum(x, 1e-300)
        cdf = np.concatenate([[0.0], np.cumsum((dens[1:] + dens[:-1]) / 2 * np.diff(x))])
        cdf /= cdf[-1]
        q = np.interp((i[::-1] + 0.5) / n, cdf, x)
        s = np.sqrt(np.maximum(q, 1e-30))
        s = s / s[0]
    else:
        raise ValueError(family)
    return np.maximum(s, 1e-12)


class BidiagSet:
    """nb bidiagonals back to back in device memory, with their roots, scales and features."""

    def __init__(self, ns, sigma, start):
        self.ns = np.asarray(ns, dtype=np.int32)
        self.off = np.concatenate([[0], np.cumsum(self.ns[:-1].astype(np.int64))]).astype(np.int64)
        self.nb = len(self.ns)
        total = int(self.ns.sum())
        self.d_n = DA.from_host(self.ns)
        self.d_off = DA.from_host(self.off)
        ds = DA.from_host(np.asarray(sigma, np.float64))
        dst = DA.from_host(np.asarray(start, np.float64))
        self.d_d = DA((total,)).zero()
        self.d_e = DA((total,)).zero()
        hip.call("exc_gk_bidiagonals", self.nb, self.d_n.ptr, self.d_off.ptr, ds.ptr, dst.ptr, self.d_d.ptr,
                 self.d_e.ptr, None)
        ds.close()
        dst.close()
        self.d_root = DA((self.nb,)).zero()
        self.d_scale = DA((self.nb,)).zero()
        hip.call("exc_bidiag_scale", self.nb, self.d_n.ptr, self.d_off.ptr, self.d_d.ptr, self.d_e.ptr,
                 self.d_root.ptr, self.d_scale.ptr, None)
        self.d_frac = DA((self.nb * K_SHIFTS,), np.float32)
        hip.call("exc_count_features", self.nb, K_SHIFTS, self.d_n.ptr, self.d_off.ptr, self.d_d.ptr, self.d_e.ptr,
                 self.d_scale.ptr, self.d_frac.ptr, None)
        hip.sync()
        self.root = self.d_root.to_host()
        self.scale = self.d_scale.to_host()
        self.frac = self.d_frac.to_host().reshape(self.nb, K_SHIFTS)

    def bidiagonal(self, b):
        d = self.d_d.to_host()
        e = self.d_e.to_host()
        o, n = int(self.off[b]), int(self.ns[b])
        return d[o:o + n].copy(), e[o:o + n - 1].copy()

    def shifts(self, tb):
        """The feature shifts scale_b * 10^(-k/2), k = 0 .. K_SHIFTS - 1, per target."""
        return self.scale[np.asarray(tb)][:, None] * 10.0 ** (-0.5 * np.arange(K_SHIFTS))[None, :]

    def cell(self, tb, ti):
        """(k, lo, hi): the shifts that bracket each target by the feature counts themselves,
        nc(lo) <= t < nc(hi) for the t-th smallest value (lo = 0 or hi = the root where no shift
        bounds it); k is the index of the first shift at or below the value (K_SHIFTS if none)."""
        tb, ti = np.asarray(tb), np.asarray(ti)
        n = self.ns[tb]
        t = n - 1 - ti
        counts = np.rint(self.frac[tb] * n[:, None]).astype(np.int64)  # nc(shift_k), decreasing in k
        below = counts <= t[:, None]
        k = np.where(below.any(1), below.argmax(1), K_SHIFTS)
        sh = self.shifts(tb)
        rows = np.arange(len(tb))
        lo = np.where(k < K_SHIFTS, sh[rows, np.minimum(k, K_SHIFTS - 1)], 0.0)
        hi = np.where(k >= 1, sh[rows, np.maximum(k - 1, 0)], self.root[tb])
        return k, lo, hi

    def features(self, tb, ti):
        """Network inputs for targets (bidiagonal tb, i-th largest value ti): the counting function
        at the shifts, the rank fraction, and the value's position interpolated linearly in log
        scale between the two shifts whose counts enclose its rank (in the target's units). All
        are free of n, so the network carries over to sizes beyond those it was trained on."""
        tb, ti = np.asarray(tb), np.asarray(ti)
        n = self.ns[tb].astype(np.float64)
        F = self.frac[tb].astype(np.float64)
        r = (n - 1 - ti + 0.5) / n
        X = np.zeros((len(tb), D_IN), np.float32)
        X[:, :K_SHIFTS] = F
        X[:, K_SHIFTS] = r
        below = F <= r[:, None]
        k = np.maximum(np.where(below.any(1), below.argmax(1), K_SHIFTS), 1)
        rows = np.arange(len(tb))
        fh = F[rows, k - 1]
        fl = np.where(k < K_SHIFTS, F[rows, np.minimum(k, K_SHIFTS - 1)], 0.0)
        den = fh - fl
        pos = np.where(den > 0, (fh - r) / np.where(den > 0, den, 1.0), 0.5)
        half_decades = (k - 1) + np.clip(pos, 0.0, 1.0)
        X[:, K_SHIFTS + 1] = -np.log(10.0) * half_decades / 2.0 / Y_SCALE
        X[:, K_SHIFTS + 2] = 1.0
        return X

    def bisect(self, tb, ti, blo=None, bhi=None, flo=None, fhi=None):
        """exc_bisect_index for the targets (bracket blo/bhi, fallback points flo/fhi); returns
        dict(sigma, leaf_lo, leaf_hi, iters, status, ms)."""
        m = len(tb)
        dtb = DA.from_host(np.asarray(tb, np.int32))
        dti = DA.from_host(np.asarray(ti, np.int32))
        dlo = DA.from_host(np.asarray(blo, np.float64)) if blo is not None else None
        dhi = DA.from_host(np.asarray(bhi, np.float64)) if bhi is not None else None
        dflo = DA.from_host(np.asarray(flo, np.float64)) if flo is not None else None
        dfhi = DA.from_host(np.asarray(fhi, np.float64)) if fhi is not None else None
        out = {k: DA((m,)) for k in ("sigma", "leaf_lo", "leaf_hi")}
        dit = DA((m,), np.int32)
        dst = DA((m,), np.int32)
        hip.sync()
        t0 = time.perf_counter()
        hip.call("exc_bisect_index", m, self.d_n.ptr, self.d_off.ptr, self.d_d.ptr, self.d_e.ptr, self.d_root.ptr,
                 TOL, dtb.ptr, dti.ptr, dlo.ptr if dlo else None, dhi.ptr if dhi else None,
                 dflo.ptr if dflo else None, dfhi.ptr if dfhi else None, out["sigma"].ptr,
                 out["leaf_lo"].ptr, out["leaf_hi"].ptr, dit.ptr, dst.ptr, None)
        hip.sync()
        ms = 1000.0 * (time.perf_counter() - t0)
        res = {k: v.to_host() for k, v in out.items()}
        res["iters"] = dit.to_host()
        res["status"] = dst.to_host()
        res["ms"] = ms
        return res


def leaf_list(res):
    """The reference's list of singular values from per-index results over every index of one
    bidiagonal (ti = 0, 1, ... in order): its tree records one value per non-empty leaf
    (bbtsvd.cpp:19-52), so indices that share a leaf (values closer than the leaf width, such as
    the copies Lanczos makes of converged values) give one entry."""
    lo, hi = res["leaf_lo"], res["leaf_hi"]
    keep = np.ones(len(lo), bool)
    keep[1:] = (lo[1:] != lo[:-1]) | (hi[1:] != hi[:-1])
    keep &= res["status"] >= 0
    return res["sigma"][keep]


def make_set(nb, n_range, seed, families=FAMILIES):
    rng = np.random.default_rng(seed)
    lo, hi = n_range
    ns = rng.integers(lo, hi + 1, nb) if hi > lo else np.full(nb, lo)
    sig, start, fam = [], [], []
    for b in range(nb):
        f = families[b % len(families)]
        sig.append(spectrum(f, int(ns[b]), rng))
        start.append(rng.standard_normal(int(ns[b])))
        fam.append(f)
    bs = BidiagSet(ns, np.concatenate(sig), np.concatenate(start))
    bs.families = fam
    return bs


# ------------------------------------------------------------------ the network
class BracketNet:
    def __init__(self, seed=0):
        n = lib.exc_mlp_param_count()
        rng = np.random.default_rng(seed)
        p = np.zeros(n, np.float32)
        o = 0
        for rows, cols in ((D_IN, 256), (256, 256), (256, 2)):
            p[o:o + rows * cols] = (rng.standard_normal(rows * cols) * np.sqrt(2.0 / rows)).astype(np.float32)
            o += rows * cols + cols
        self.n = n
        self.p = DA.from_host(p)
        self.m = DA((n,), np.float32).zero()
        self.v = DA((n,), np.float32).zero()
        self.g = DA((n,), np.float32).zero()
        self.pb = DA((n,), np.uint16)
        self.step = 0
        self.k_lo, self.k_hi = BRACKET_K

    def params(self):
        return self.p.to_host()

    def set_params(self, p):
        self.p = DA.from_host(np.asarray(p, np.float32))

    def grad(self, X, y):
        """(loss, flat gradient) of one batch through exc_mlp_forward/backward."""
        B = X.shape[0]
        dx = DA.from_host(np.ascontiguousarray(X, np.float32))
        dy = DA.from_host(np.ascontiguousarray(y, np.float32))
        acts = DA((lib.exc_mlp_acts_floats(B),), np.float32)
        out = DA((B * 2,), np.float32)
        scr = DA((lib.exc_mlp_backward_scratch_floats(B),), np.float32)
        loss = DA((1,), np.float32)
        ws = hip.workspace(hip.lib.exc_gemm_workspace(256, 256, B))
        hip.call("exc_mlp_forward", B, dx.ptr, self.p.ptr, acts.ptr, out.ptr, ws.ptr, ws.nbytes, None)
        hip.call("exc_mlp_backward", B, dx.ptr, dy.ptr, self.p.ptr, acts.ptr, out.ptr, self.g.ptr, scr.ptr, loss.ptr,
                 ws.ptr, ws.nbytes, None)
        hip.sync()
        return float(loss.to_host()[0]), self.g.to_host()

    def train(self, X, y, steps=2000, batch=4096, lr=2e-3, seed=0, log=None):
        """Adam on random minibatches; every kernel is the library's."""
        rng = np.random.default_rng(seed)
        N = X.shape[0]
        dX = DA.from_host(np.ascontiguousarray(X, np.float32))
        dY = DA.from_host(np.ascontiguousarray(y, np.float32))
        bx = DA((batch * D_IN,), np.float32)
        by = DA((batch,), np.float32)
        acts = DA((lib.exc_mlp_acts_floats(batch),), np.float32)
        out = DA((batch * 2,), np.float32)
        scr = DA((lib.exc_mlp_backward_scratch_floats(batch),), np.float32)
        loss = DA((1,), np.float32)
        ws = hip.workspace(hip.lib.exc_gemm_workspace(256, 256, batch))
        hist = []
        t0 = time.perf_counter()
        for it in range(steps):
            idx = rng.integers(0, N, batch)
            # named host batches: a pointer taken from a temporary array would dangle during the copy
            bx_h = np.ascontiguousarray(X[idx], np.float32)
            by_h = np.ascontiguousarray(y[idx], np.float32)
            hip.call("exc_memcpy_h2d", bx.ptr, bx_h.ctypes.data, bx.nbytes, None)
            hip.call("exc_memcpy_h2d", by.ptr, by_h.ctypes.data, by.nbytes, None)
            hip.call("exc_mlp_forward", batch, bx.ptr, self.p.ptr, acts.ptr, out.ptr, ws.ptr, ws.nbytes, None)
            hip.call("exc_mlp_backward", batch, bx.ptr, by.ptr, self.p.ptr, acts.ptr, out.ptr, self.g.ptr, scr.ptr,
                     loss.ptr, ws.ptr, ws.nbytes, None)
            self.step += 1
            cur_lr = lr * (0.5 * (1 + np.cos(np.pi * it / steps)))
            hip.call("exc_mlp_adam", self.p.ptr, self.g.ptr, self.m.ptr, self.v.ptr, cur_lr, 0.9, 0.999, 1e-8,
                     self.step, None)
            if it % 200 == 0 or it == steps - 1:
                hip.sync()
                hist.append((it, float(loss.to_host()[0])))
                if log:
                    log(f"step {it}: loss {hist[-1][1]:.4f}")
        hip.sync()
        del dX, dY
        return {"loss": hist, "seconds": time.perf_counter() - t0}

    def infer(self, X):
        """BF16 inference: (c, w) per row."""
        B = X.shape[0]
        hip.call("exc_mlp_to_bf16", self.p.ptr, self.pb.ptr, None)
        dx = DA.from_host(np.ascontiguousarray(X, np.float32))
        out = DA((B * 2,), np.float32)
        hip.sync()
        t0 = time.perf_counter()
        hip.call("exc_mlp_infer_bf16", B, dx.ptr, self.pb.ptr, out.ptr, None)
        hip.sync()
        self.infer_ms = 1000.0 * (time.perf_counter() - t0)
        return out.to_host().reshape(B, 2)

    def predict(self, bs, tb, ti):
        """(c, b): the BF16 network's prediction of y = ln(sigma / scale) / Y_SCALE and its Laplace
        scale, per target."""
        O = self.infer(bs.features(tb, ti)).astype(np.float64)
        return O[:, 0], np.exp(np.clip(O[:, 1], -30.0, 5.0))

    def _ends(self, bs, tb, c, b, k_lo, k_hi):
        s = bs.scale[np.asarray(tb)]
        lo = s * np.exp(np.clip(Y_SCALE * (c - k_lo * b), -700.0, 700.0))
        hi = s * np.exp(np.clip(Y_SCALE * (c + k_hi * b), -700.0, 700.0))
        return lo, hi

    def calibrate(self, bs, tb, ti, y, quantiles=(0.3, 0.5, 0.7, 0.9, 0.97, 0.99, 0.999)):
        """Bracket ends from held-out targets (bidiagonals the network was not trained on): the
        candidate ends are quantiles of the scaled error (y - c) / b on each side, and the pair that
        evaluates the fewest Sturm counts on those targets is kept. A lower end that misses still
        bounds the value from above (and an upper end that misses, from below), so the best ends
        are not simply the widest."""
        c, b = self.predict(bs, tb, ti)
        z = (np.asarray(y, np.float64) - c) / b
        _, clo, chi = bs.cell(tb, ti)
        best = None
        for k_lo in sorted(set(float(np.quantile(-z, q)) for q in quantiles)):
            for k_hi in sorted(set(float(np.quantile(z, q)) for q in quantiles)):
                if k_lo + k_hi <= 0:
                    continue
                lo, hi = self._clip(*self._ends(bs, tb, c, b, k_lo, k_hi), clo, chi)
                counts = int(bs.bisect(tb, ti, lo, hi, clo, chi)["iters"].sum())
                if best is None or counts < best[0]:
                    best = (counts, k_lo, k_hi)
        self.k_lo, self.k_hi = best[1], best[2]
        return self.k_lo, self.k_hi

    @staticmethod
    def _clip(lo, hi, clo, chi):
        lo, hi = np.maximum(lo, clo), np.minimum(hi, chi)
        off = ~(lo < hi)  # a prediction outside the enclosing shifts: keep the shifts
        lo[off], hi[off] = clo[off], chi[off]
        return lo, hi

    def brackets(self, bs, tb, ti):
        """(lo, hi, flo, fhi): the BF16 network's bracket for each target, clipped to the shifts
        that enclose it by the feature counts, and those shifts as the fallback points."""
        c, b = self.predict(bs, tb, ti)
        _, clo, chi = bs.cell(tb, ti)
        lo, hi = self._clip(*self._ends(bs, tb, c, b, self.k_lo, self.k_hi), clo, chi)
        return lo, hi, clo, chi


def targets(bs, per, seed):
    """(tb, ti, y): per random indices of every bidiagonal and their labels
    y = ln(sigma / scale) / Y_SCALE from the reference's bisection tree."""
    rng = np.random.default_rng(seed)
    tb = np.repeat(np.arange(bs.nb), per)
    ti = np.concatenate([rng.integers(0, bs.ns[b], per) for b in range(bs.nb)])
    res = bs.bisect(tb, ti)
    ok = (res["status"] >= 0) & (res["sigma"] > 0)
    y = (np.log(res["sigma"][ok] / bs.scale[tb[ok]]) / Y_SCALE).astype(np.float32)
    return tb[ok], ti[ok], y


def training_data(bs, per, seed):
    """Network inputs and labels for per random indices of every bidiagonal."""
    tb, ti, y = targets(bs, per, seed)
    return bs.features(tb, ti), y


def evaluate(net, bs, b):
    """Every singular value of bidiagonal b: plainly, from the network's brackets, and (for
    comparison) from the enclosing feature shifts alone. Identical leaves are required; returns the
    Sturm-count and time comparison."""
    n = int(bs.ns[b])
    tb = np.full(n, b)
    ti = np.arange(n)
    plain = bs.bisect(tb, ti)
    lo, hi, clo, chi = net.brackets(bs, tb, ti)
    br = bs.bisect(tb, ti, lo, hi, clo, chi)
    cell = bs.bisect(tb, ti, clo, chi)
    same = all(np.array_equal(plain[k], r[k]) for r in (br, cell) for k in ("sigma", "leaf_lo", "leaf_hi"))
    cp = plain["iters"].sum()
    return {"n": n, "identical": same, "leaves": leaf_list(plain),
            "counts_plain": int(cp), "counts_bracketed": int(br["iters"].sum()),
            "counts_shifts_only": int(cell["iters"].sum()),
            "saved_frac": 1.0 - br["iters"].sum() / cp,
            "saved_frac_shifts_only": 1.0 - cell["iters"].sum() / cp,
            "bounded_both": float(np.mean(br["status"] == 1)), "bounded_one": float(np.mean(br["status"] == 2)),
            "ms_plain": plain["ms"], "ms_bracketed": br["ms"], "ms_inference": net.infer_ms,
            "sigma": plain["sigma"]}
