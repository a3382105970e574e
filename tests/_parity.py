"""Per-batch parity of the HIP force step against the oracle (shared helpers).

Tolerance (FP32 path), per batch, on identical X and identical negative lists:
  * normwise: ||delta_gpu - delta_32|| <= 1e-5 ||delta_32||;
  * elementwise on the updated rows X' = X + delta:
      |X'_gpu - X'_32| <= 1e-5 (|X'_32| + rms(X'_32)) + 2 |X'_32 - X'_64|,
    where X'_32 / X'_64 come from the FP32 / FP64 oracle. The last term
    admits the FP32 reference's own rounding error: it sums each component
    serially in float (64 701 terms for the heaviest R-MAT-20 row), while the
    GPU sums short interleaved runs and combines them in a fixed tree, so on
    hub rows it sits closer to the FP64 result than the FP32 reference does.
The GPU reduces |v|^2 with a DPP tree, so results agree to rounding, not
bitwise. Every GPU result is bitwise reproducible.

BF16-shadow form: the forces are the FP32 forces at X rounded to bfloat16
(every gather and the row's own distance), so on identical X they meet the
FP32 bar against the oracle run on the rounded matrix.
"""
from __future__ import annotations

import numpy as np

import _oracle as orc
from distembed_amd import api

RTOL = 1e-5


def make_ctx(g, d, X, B, lists, gather=api.FP32, quirks=3, lr=0.02, tuning=None, bound=5.0):
    ctx = api.ForceEmbed(g.n, d, gather=gather, quirks=quirks)
    ctx.upload_csr(g)
    ctx.upload_x(X)
    if lists is not None:
        ctx.set_negatives(lists)
    if tuning:
        ctx.set_tuning(**tuning)
    ctx.configure(B, lr, bound)
    return ctx


def rows_of(n, B, j):
    return min((j + 1) * B, n) - j * B


def check_rows(new_gpu, new_ref, delta_gpu, delta_ref, new_ref64, what=""):
    new_ref = new_ref.astype(np.float64)
    scale = float(np.sqrt(np.mean(new_ref ** 2)))
    err = np.abs(new_gpu.astype(np.float64) - new_ref)
    bound = RTOL * (np.abs(new_ref) + scale) + 2.0 * np.abs(new_ref - new_ref64)
    assert np.all(err <= bound), f"{what}: max excess {(err - bound).max():.3e}"
    num = np.linalg.norm((delta_gpu - delta_ref).astype(np.float64))
    den = np.linalg.norm(delta_ref.astype(np.float64))
    assert num <= RTOL * den + 1e-30, f"{what}: normwise delta error {num / den:.3e}"
    return num / den if den > 0 else 0.0


def batch_parity(g, d, B, steps, ns=5, quirks=3, tuning=None, seed=0, lr=0.02, X=None,
                 bound=5.0, lists=None, apply_check=True):
    """GPU delta of every (epoch, batch) in `steps` against the FP32 and FP64
    oracle on the same X and lists; then the applied step (fe_step) writes
    exactly X + delta into the batch rows and leaves every other row alone.
    Returns the normwise delta errors."""
    if X is None:
        X = api.init_uniform(g.n, d, -0.5, 0.5, seed=seed)
    batches = -(-g.n // B)
    if lists is None and ns:
        n_lists = max(e for e, _ in steps) + batches
        lists = api.reference_negatives(g.n, n_lists, ns)[0]
    ctx = make_ctx(g, d, X, B, lists, quirks=quirks, tuning=tuning, lr=lr, bound=bound)
    X64 = X.astype(np.float64)
    errs = []
    try:
        for (epoch, j) in steps:
            rows = rows_of(g.n, B, j)
            dg = ctx.step_delta(epoch, j, rows)
            dr = orc.step(g, X, B, epoch, j, lists, lr=lr, bound=bound, quirks=quirks,
                          apply=False, want_delta=True)
            dr64 = orc.step(g, X64, B, epoch, j, lists, lr=lr, bound=bound, quirks=quirks,
                            apply=False, want_delta=True, dtype=np.float64)
            r0 = j * B
            errs.append(check_rows(X[r0:r0 + rows] + dg, X[r0:r0 + rows] + dr, dg, dr,
                                   X64[r0:r0 + rows] + dr64,
                                   f"d={d} B={B} epoch={epoch} batch={j}"))
            print(f"  parity n={g.n} d={d} B={B} ns={ns} epoch {epoch} batch {j}: normwise "
                  f"delta error {errs[-1]:.2e}", flush=True)
            if apply_check:
                # the applied step writes exactly X + delta into the batch rows
                ctx.step(epoch, j)
                Xg = ctx.download_x()
                np.testing.assert_array_equal(Xg[r0:r0 + rows], X[r0:r0 + rows] + dg)
                assert np.array_equal(Xg[:r0], X[:r0]), "rows before the batch changed"
                assert np.array_equal(Xg[r0 + rows:], X[r0 + rows:]), "rows after the batch changed"
                del Xg
                ctx.upload_x(X)
    finally:
        ctx.close()
    return errs


def bf16_batch_parity(g, d, B, steps, tuning=None, ns=5, seed=1, X=None, lists=None,
                      grad_err_vs_fp64=None):
    """BF16-shadow deltas against the oracle at the rounded rows (FP32 bar).
    grad_err_vs_fp64: when set, the per-step gradient error
    ||delta_gpu - delta_64|| / ||delta_64|| against the FP64 oracle at the
    unrounded X must stay below it. Returns [(epoch, batch, err_vs_rounded,
    grad_err_vs_fp64)]."""
    if X is None:
        X = api.init_uniform(g.n, d, -0.5, 0.5, seed=seed)
    batches = -(-g.n // B)
    if lists is None:
        lists = api.reference_negatives(g.n, max(e for e, _ in steps) + batches, ns)[0]
    ctx = make_ctx(g, d, X, B, lists, gather=api.BF16, tuning=tuning)
    out = []
    try:
        Xr = orc.bf16_round(X)
        Xr64 = Xr.astype(np.float64)
        X64 = X.astype(np.float64) if grad_err_vs_fp64 is not None else None
        for (epoch, j) in steps:
            rows = rows_of(g.n, B, j)
            dg = ctx.step_delta(epoch, j, rows)
            dr = orc.step(g, Xr, B, epoch, j, lists, apply=False, want_delta=True)
            dr64 = orc.step(g, Xr64, B, epoch, j, lists, apply=False, want_delta=True,
                            dtype=np.float64)
            r0 = j * B
            Xb64 = X[r0:r0 + rows].astype(np.float64)
            e = check_rows(X[r0:r0 + rows] + dg, X[r0:r0 + rows] + dr, dg, dr, Xb64 + dr64,
                           f"bf16 d={d} B={B} epoch={epoch} batch={j}")
            ge = None
            if grad_err_vs_fp64 is not None:
                d64 = orc.step(g, X64, B, epoch, j, lists, apply=False, want_delta=True,
                               dtype=np.float64)
                ge = float(np.linalg.norm(dg.astype(np.float64) - d64) / np.linalg.norm(d64))
                assert ge < grad_err_vs_fp64, (epoch, j, ge)
            out.append((epoch, j, e, ge))
            print(f"  bf16 parity n={g.n} d={d} B={B} epoch {epoch} batch {j}: normwise delta "
                  f"error {e:.2e}" + (f", gradient error vs FP64 {ge:.2e}" if ge is not None
                                      else ""), flush=True)
    finally:
        ctx.close()
    return out


def clamp_fractions(g, X, B, j, neg, rep, bound, quirks=3):
    """Fractions of attractive and repulsive components of batch j whose
    v * d1 exceeds the clamp bound (FP64), i.e. where scale() binds."""
    n = g.n
    r0, r1 = j * B, min((j + 1) * B, n)
    c_limit = n - 1 if quirks & 2 else n
    rows = np.repeat(np.arange(r0, r1), np.diff(g.row_ptr[r0:r1 + 1]))
    cols = g.col_idx[g.row_ptr[r0]:g.row_ptr[r1]].astype(np.int64)
    keep = cols < c_limit
    rows, cols = rows[keep], cols[keep]
    X64 = X.astype(np.float64)
    v = X64[rows] - X64[cols]
    s = (v * v).sum(axis=1, keepdims=True)
    att = np.abs(v * (-2.0 / (1.0 + s))) > bound
    rr = np.arange(r0, r0 + rep)
    vr = X64[rr][:, None, :] - X64[np.asarray(neg)][None, :, :]
    sr = (vr * vr).sum(axis=2, keepdims=True)
    rep_c = np.abs(vr * (2.0 / ((sr + 1e-6) * (1.0 + sr)))) > bound
    return (float(att.mean()) if att.size else 0.0, float(rep_c.mean()) if rep_c.size else 0.0)
