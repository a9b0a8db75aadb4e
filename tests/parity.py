"""Parity criteria between the HIP path and the oracle (same randoms, same trees).

Tolerance: 1e-12 relative (FP64).

* ok flags: identical.  A disagreement is accepted only for an event within 1e-12 of a
  kinematic threshold of the oracle (``threshold_margin`` below), and such events are
  counted (``ok_mismatch_threshold``).
* factor and volume of ok events: |d| <= 1e-12 |ref|.
* momenta of ok events, component by component:
  strict      |dp^mu| <= 1e-12 max(|p^mu|, E_j)             (energy-scaled relative)
  conditioned |dp^mu| <= tol_e max(|p^mu|, E_j, scale^mu)
  where scale^mu = (|L1_root| ... |L1_parent| |v_j|)^mu is the componentwise forward-error
  scale of the boost chain that the oracle reports (the magnitude of the terms that are
  summed into p^mu), and tol_e = max(1e-12, 64 eps cond_e) with cond_e the event's condition
  number from the oracle: max over branches of (msq + msq1 + msq2)^2 / lambda (lambda -> 0 at a
  decay threshold) and 1 / sin theta of peaked angular maps.  The factor obeys
  |df| <= tol_e |f| likewise.  Every ok event must meet the conditioned bounds; events that
  meet them but not the strict 1e-12 ones are ill-conditioned points -- near a threshold, or a
  massless daughter emitted against a large boost -- counted as ``ill_conditioned``.  Two
  correct FP64 evaluations that differ by one ulp in a transcendental can differ there by
  ~eps * cond, so no implementation can promise 1e-12 at such points.
"""
from __future__ import annotations

import numpy as np

TOL = 1e-12
EPS = 2.0 ** -52


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


def compare(gpu, ref, tol=TOL, near_threshold=None):
    ok_g = np.asarray(gpu["ok"]).astype(bool)
    ok_r = np.asarray(ref["ok"]).astype(bool)
    mism = ok_g != ok_r
    both = ok_g & ok_r
    pg, pr = np.asarray(gpu["p"])[both], np.asarray(ref["p"])[both]
    strict_scale = np.maximum(np.abs(pr), np.abs(pr[..., :1]))
    d = np.abs(pg - pr)
    dp = d / np.where(strict_scale > 0, strict_scale, 1.0)
    if "scale" in ref:
        cond_scale = np.maximum(strict_scale, np.asarray(ref["scale"])[both])
    else:
        cond_scale = strict_scale
    dpc = d / np.where(cond_scale > 0, cond_scale, 1.0)
    fg, fr = np.asarray(gpu["factor"])[both], np.asarray(ref["factor"])[both]
    df = np.abs(fg - fr) / np.where(np.abs(fr) > 0, np.abs(fr), 1.0)
    vg, vr = np.asarray(gpu["volume"])[both], np.asarray(ref["volume"])[both]
    dv = np.abs(vg - vr) / np.where(np.abs(vr) > 0, np.abs(vr), 1.0)
    cond = np.asarray(ref["cond"])[both] if "cond" in ref else np.ones(len(fg))
    tol_e = np.maximum(tol, 64 * EPS * cond)
    flat = lambda a: a.reshape(a.shape[0], int(np.prod(a.shape[1:])))
    strict_bad = (flat(dp) > tol).any(axis=1) | (df > tol)
    cond_bad = (flat(dpc) > tol_e[:, None]).any(axis=1) | (df > tol_e) | (dv > tol) | ~np.isfinite(fg) | \
        ~np.isfinite(flat(pg)).all(axis=1)
    n_thr = int((mism & near_threshold).sum()) if near_threshold is not None else 0
    return dict(n=len(ok_g), n_ok=int(ok_r.sum()), ok_mismatch=int(mism.sum()),
                ok_mismatch_threshold=n_thr, ok_mismatch_unexplained=int(mism.sum()) - n_thr,
                max_dp=float(dp.max()) if dp.size else 0.0,
                max_dp_conditioned=float((flat(dpc) / tol_e[:, None]).max() * tol) if dpc.size else 0.0,
                max_cond=float(cond.max()) if cond.size else 1.0,
                max_df=float(df.max()) if df.size else 0.0, max_dv=float(dv.max()) if dv.size else 0.0,
                ill_conditioned=int((strict_bad & ~cond_bad).sum()), n_bad=int(cond_bad.sum()),
                bitwise_p=float((pg == pr).all(axis=(1, 2)).mean()) if pg.size else 1.0)


def assert_parity(r):
    assert r["ok_mismatch_unexplained"] == 0, r
    assert r["n_bad"] == 0, r
