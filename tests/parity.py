"""Parity criterion between the HIP forward path and the oracle (same randoms, same trees).

Tolerance: 1e-12 relative (FP64), with a measured, counted and capped set of exceptions.

Per event, over the events both sides call ok:

* momenta, component by component: d = |dp^mu| / max(|p^mu|, E_j)  (energy-scaled relative);
* factor: |df| / |f|; volume: |dv| / |v| <= 1e-12 always;
* msq of every branch (when the HIP side returned it): |dmsq| / |msq|.

An event is *strict* when all of these are <= 1e-12.  Two correct FP64 evaluations of the
reference algorithm are not bit-identical -- the elementary functions (sqrt, exp, log, sin,
cos, tan, atan) of two libraries may round their results differently by an ulp -- and at a few
events (near a decay threshold, a massless daughter emitted against a large boost, x at the
edge of a peaked map) an ulp of difference moves the outputs far beyond 1e-12.  How far is
*measured*, not modelled: ``oracle.sensitivity`` reruns the oracle with every elementary
function result moved by -1, 0 or +1 ulp along independent seeded streams and records, per
output, the largest change (``s``).  Then:

1. ok flags are identical, except at events whose oracle ok flag itself flips under those
   one-ulp changes, or that lie within 1e-12 of a kinematic cut (``threshold_margin``);
   both kinds are counted.
2. Every non-strict event (an *exception*) has each output within 1e-12 + 8 s of the oracle:
   the HIP path is no further from the oracle than the oracle is from itself under a few
   one-ulp changes.  (Measured over the test processes, the worst ratio d / s is about 2.)
3. The number of exceptions is at most max(2, the number of ok events whose measured
   sensitivity itself exceeds 1e-12) -- the HIP path departs from 1e-12 on no more events than
   one-ulp changes of the oracle's own functions do.
4. Deviations above 1e-9 occur only at *undetermined* events, whose measured sensitivity
   exceeds 1e-10 (the reference algorithm does not fix the output to 1e-10 in FP64 there:
   x within 1e-12 of the end of a peaked map); at those the four-momentum sum and the leaves'
   masses must still hold to 1e-9 of sqrt(s).
5. Over all ok events (when there are at least 200 of them), the median strict deviation is
   below 1e-14 and the 99.9th percentile below max(1e-12, the 99.9th percentile of the measured
   sensitivity).

``assert_parity`` puts every count and maximum into its message, so a failing run shows them.
"""
from __future__ import annotations

import numpy as np

TOL = 1e-12
CAP = 1e-9
UNDETERMINED = 1e-10
SENS_FACTOR = 8.0
SENS_RUNS = 6
QUANTILE_MIN = 200


def threshold_margin(proc, msq, channels, sqrts):
    """Smallest relative distance of each event to one of the kinematic cuts of the oracle
    (phs_cpu.cpp:225-226): lambda / msq^2, (m - m1 - m2) / m and (m_max - m) / m_max, from the
    oracle's msq of every branch.  The branches are visited in the oracle's post-order up to
    and including the first one whose cut fails: the oracle returns there (phs_cpu.cpp:186-193)
    and leaves the msq of later branches at 0, so only the cuts it actually evaluated count."""
    n = msq.shape[0]
    out = np.full(n, np.inf)
    root = proc.n_prt_out - 1
    import oracle
    ms = oracle.mass_sums(proc)
    for c in range(proc.n_channels):
        sel = np.nonzero(channels == c)[0]
        if sel.size == 0:
            continue
        m2 = msq[sel]
        order = []

        def post(b):
            for k in (proc.d1[c, b] - 1, proc.d2[c, b] - 1):
                if proc.has_children[c, k]:
                    post(k)
            order.append(b)

        post(root)
        alive = np.ones(sel.size, dtype=bool)
        marg = np.full(sel.size, np.inf)
        for b in order:
            k1, k2 = proc.d1[c, b] - 1, proc.d2[c, b] - 1
            b1, b2 = m2[:, k1], m2[:, k2]
            sq = np.asarray(sqrts, np.float64)[sel] if np.ndim(sqrts) else sqrts   # per-event sqrt(s) allowed
            a = np.zeros_like(b1) + sq * sq if b == root else m2[:, b]
            lda = (a - b1 - b2) ** 2 - 4 * b1 * b2
            with np.errstate(all="ignore"):
                m, m1_, m2_ = np.sqrt(a), np.sqrt(b1), np.sqrt(b2)
                mmax = sq if b == root else sq - ms[c, root] + ms[c, b]
                cuts = [np.abs(lda) / np.maximum(a * a, 1e-300), np.abs(m - m1_ - m2_) / np.maximum(m, 1e-300)]
                if b != root:   # the root sits at m = m_max = sqrt(s) by construction
                    cuts.append(np.abs(mmax - m) / mmax)
                r = np.nan_to_num(np.minimum.reduce(cuts), nan=0.0)
                passed = (lda > 0) & (m > m1_ + m2_) & (m <= mmax)
            marg = np.where(alive, np.minimum(marg, r), marg)
            alive &= passed
        out[sel] = marg
    return out


def _rel(d, scale):
    return d / np.where(scale > 0, scale, 1.0)


def compare(gpu, ref, sens, near_threshold=None, beams=None, masses=None):
    """Statistics of the HIP outputs ``gpu`` (p, factor, volume, ok[, msq]) against the oracle's
    ``ref`` (the same keys, msq included when gpu has it), with the measured sensitivity
    ``sens`` of oracle.sensitivity (p, factor[, msq], ok_flip).  ``beams`` = (E1, E2) per event
    for the conservation check, ``masses`` the final-state masses."""
    ok_g = np.asarray(gpu["ok"]).astype(bool)
    ok_r = np.asarray(ref["ok"]).astype(bool)
    n = len(ok_g)
    mism = ok_g != ok_r
    flip = np.asarray(sens["ok_flip"], bool)
    near = np.zeros(n, bool) if near_threshold is None else np.asarray(near_threshold, bool)
    both = ok_g & ok_r
    pg, pr = np.asarray(gpu["p"])[both], np.asarray(ref["p"])[both]
    m = pg.shape[0]
    scale = np.maximum(np.abs(pr), np.abs(pr[..., :1]))
    dp = _rel(np.abs(pg - pr), scale).reshape(m, -1)
    sp = _rel(np.asarray(sens["p"])[both], scale).reshape(m, -1)
    fg, fr = np.asarray(gpu["factor"])[both], np.asarray(ref["factor"])[both]
    df = _rel(np.abs(fg - fr), np.abs(fr))
    sf = _rel(np.asarray(sens["factor"])[both], np.abs(fr))
    vg, vr = np.asarray(gpu["volume"])[both], np.asarray(ref["volume"])[both]
    dv = _rel(np.abs(vg - vr), np.abs(vr))
    # every output of an event side by side: deviations and sensitivities
    dev = [dp, df[:, None]]
    sen = [sp, sf[:, None]]
    if "msq" in gpu:
        mg, mr = np.asarray(gpu["msq"])[both], np.asarray(ref["msq"])[both]
        dm = _rel(np.abs(mg - mr), np.abs(mr))
        sm = _rel(np.asarray(sens["msq"])[both], np.abs(mr))
        dev.append(dm)
        sen.append(sm)
    dev = np.concatenate(dev, axis=1)
    sen = np.concatenate(sen, axis=1)
    finite = np.isfinite(dev).all(axis=1) & np.isfinite(pg.reshape(m, -1)).all(axis=1) & np.isfinite(fg)
    dev_e = np.where(finite, dev.max(axis=1), np.inf)
    sen_e = sen.max(axis=1)
    strict = dev_e <= TOL
    within = (dev <= TOL + SENS_FACTOR * sen).all(axis=1) & finite
    undetermined = sen_e > UNDETERMINED
    # invariants at the undetermined events: conservation and on-shell leaves (1e-9 sqrt(s))
    inv_bad = 0
    if undetermined.any():
        p_u = pg[undetermined]
        tot = p_u.sum(axis=1)
        if beams is not None:
            e1, e2 = (np.broadcast_to(np.asarray(b, np.float64), (n,))[both][undetermined] for b in beams)
        else:
            e1 = e2 = np.full(p_u.shape[0], np.nan)
        sq = 2 * np.sqrt(e1 * e2)
        want = np.stack([e1 + e2, 0 * e1, 0 * e1, e1 - e2], axis=1)
        bad = (np.abs(tot - want) > CAP * sq[:, None]).any(axis=1)
        if masses is not None:
            m2 = p_u[..., 0] ** 2 - (p_u[..., 1:] ** 2).sum(axis=-1)
            bad |= (np.abs(m2 - np.asarray(masses) ** 2) > CAP * (sq ** 2)[:, None]).any(axis=1)
        inv_bad = int(bad.sum()) if beams is not None else 0
    n_sens = int((sen_e > TOL).sum())
    exc = ~strict
    q = lambda a, p: float(np.quantile(a, p)) if a.size else 0.0
    return dict(
        n=n, n_ok=int(ok_r.sum()), n_ok_both=int(m),
        ok_mismatch=int(mism.sum()),
        ok_mismatch_explained=int((mism & (flip | near)).sum()),
        ok_mismatch_unexplained=int((mism & ~(flip | near)).sum()),
        exceptions=int(exc.sum()), exception_budget=max(2, n_sens),
        exceptions_outside_sensitivity=int((exc & ~within).sum()),
        above_cap_determined=int(((dev_e > CAP) & ~undetermined).sum()),
        undetermined=int(undetermined.sum()), invariant_failures=inv_bad,
        max_dev=float(dev_e.max()) if m else 0.0,
        max_dev_determined=float(dev_e[~undetermined].max()) if (~undetermined).any() else 0.0,
        max_dev_over_sens=float(((dev - TOL) / np.maximum(sen, 1e-300)).max()) if exc.any() else 0.0,
        median_dev=q(dev_e, 0.5), q999_dev=q(dev_e, 0.999), q999_sens=q(sen_e, 0.999),
        max_df=float(df.max()) if m else 0.0, max_dv=float(dv.max()) if m else 0.0,
        bitwise_p=float((pg == pr).all(axis=(1, 2)).mean()) if m else 1.0,
        nonfinite=int((~finite).sum()))


def summary(r):
    return (f"ok {r['n_ok']}/{r['n']}, ok mismatches {r['ok_mismatch']} "
            f"(unexplained {r['ok_mismatch_unexplained']}); exceptions to 1e-12: {r['exceptions']} "
            f"(budget {r['exception_budget']}, outside 8x sensitivity {r['exceptions_outside_sensitivity']}); "
            f"undetermined {r['undetermined']} (invariant failures {r['invariant_failures']}); "
            f"above 1e-9 while determined {r['above_cap_determined']}; max dev {r['max_dev']:.2e} "
            f"(determined {r['max_dev_determined']:.2e}, max (dev-1e-12)/sens {r['max_dev_over_sens']:.2f}); "
            f"median {r['median_dev']:.1e}, q99.9 {r['q999_dev']:.1e} (sensitivity q99.9 {r['q999_sens']:.1e}); "
            f"max df {r['max_df']:.1e}, max dv {r['max_dv']:.1e}")


def assert_parity(r):
    msg = summary(r)
    print(msg)
    assert r["nonfinite"] == 0, msg
    assert r["ok_mismatch_unexplained"] == 0, msg
    assert r["max_dv"] <= TOL, msg
    assert r["exceptions_outside_sensitivity"] == 0, msg
    assert r["exceptions"] <= r["exception_budget"], msg
    assert r["above_cap_determined"] == 0, msg
    assert r["invariant_failures"] == 0, msg
    if r["n_ok_both"] >= QUANTILE_MIN:
        assert r["median_dev"] < 1e-14, msg
        assert r["q999_dev"] <= max(TOL, r["q999_sens"]), msg


def forward_check(proc, x, channels, gpu, E1=None, E2=None, ref=None):
    """Oracle, measured sensitivity, threshold margins and the comparison of ``gpu`` in one
    call; returns the statistics (see assert_parity).  E1 / E2: per-event beam energies (None:
    the process's symmetric beams).  ``ref``: the oracle's outputs if already computed (must
    carry msq)."""
    import oracle
    want_msq = "msq" in gpu
    if ref is None or "msq" not in ref:
        ref = oracle.forward(proc, x, channels, E1=E1, E2=E2, want_msq=True)
    sens = oracle.sensitivity(proc, x, channels, E1=E1, E2=E2, runs=SENS_RUNS, base=ref, want_msq=want_msq)
    sq = proc.sqrts if E1 is None else 2 * np.sqrt(np.asarray(E1) * np.asarray(E2))
    near = threshold_margin(proc, ref["msq"], channels, sq) < TOL
    beams = (proc.sqrts / 2, proc.sqrts / 2) if E1 is None else (E1, E2)
    return compare(gpu, ref, sens, near_threshold=near, beams=beams, masses=proc.flv_masses[proc.n_in:])
