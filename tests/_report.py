"""Host replica of the error report's choice of targets and formula (the
err_rel_l2 of holo.h).  Of count = min(M, 256) targets, n_e = count / 2 are
drawn by the result's energy (the first target at which the running sum of
|r|^2 exceeds (i + 1/2) / n_e of the total T) and n_u = count - n_e are spread
evenly (floor(j M / n_u)); every target once (n_e = 0) when M <= 256.  The
report is sqrt(sum_i |r[m_i] - r64_i|^2 / (n_e |r[m_i]|^2 + n_u T / M)).
Without positive finite energy: the even targets floor(i M / count) and
their relative L2 error."""
import numpy as np


def rel_l2(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(np.asarray(b)))


def tmap(M, count=256):
    count = min(M, count)
    return (np.arange(count, dtype=np.int64) * M) // count


def pick(r, count=256):
    """(targets, weighted, n_e) of the report for result r (all M values)."""
    r = np.asarray(r)
    M = len(r)
    count = min(M, count)
    e = np.abs(r) ** 2
    C = np.cumsum(e)
    total = C[-1]
    if not (total > 0 and np.isfinite(total)):
        return tmap(M, count), False, 0
    ne = 0 if M <= count else count // 2
    c = (np.arange(ne) + 0.5) / max(ne, 1) * total
    m = np.searchsorted(C, c, side="right")
    last = int(np.flatnonzero(e > 0)[-1])
    drawn = np.where(m < M, m, last).astype(np.int64)
    return np.concatenate([drawn, tmap(M, count - ne)]), True, ne


def report_of(r, picked, ref):
    """The report from result r (all M values), picked = pick(r) and ref =
    r64 at its targets."""
    r = np.asarray(r)
    sub, weighted, ne = picked
    a, b = r[sub], np.asarray(ref)
    if not weighted:
        return rel_l2(a, b)
    total = float(np.sum(np.abs(r) ** 2))
    w = ne * np.abs(a) ** 2 + (len(sub) - ne) * total / len(r)
    return float(np.sqrt(np.sum(np.abs(a - b) ** 2 / w)))
