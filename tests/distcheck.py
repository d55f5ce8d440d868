"""Distribution comparison of whole-run trajectories against the reference-order fixture.

The fixture (tests/golden/ref_dist_random20k.npz) holds 200 runs of the oracle with
reference-order MT19937 randomness. A candidate is 200 runs on other seeds (the
transmission path's counter-based draws cannot reproduce the reference's streams seed
for seed, so agreement is distributional). For each of daily new_infections,
new_severe and new_deaths:
  1. on every day, the candidate's mean agrees with the reference mean within a
     two-sample interval at family-wise 95% over all days (Bonferroni: half-width
     z(1 - 0.025/n_days) * sqrt(s_r^2/n_r + s_c^2/n_c)); a pointwise quantile band is
     not used because it has zero width for rare counts (early deaths);
  2. the candidate's mean cumulative total lies inside the reference mean's 95%
     confidence interval, widened for the candidate's own sampling error
     (|m_c - m_r| <= 1.96 * sqrt(s_r^2/n_r + s_c^2/n_c));
  3. a two-sample KS test on the cumulative totals gives p > 0.01;
and for new_infections also a KS test (p > 0.01) on the day of the peak.
"""

import json
import os

import numpy as np

SERIES = ('new_infections', 'new_severe', 'new_deaths')
FIXTURE = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'golden', 'ref_dist_random20k.npz')


def load_reference():
    with np.load(FIXTURE, allow_pickle=False) as z:
        ref = {k: z[k].astype(np.float64) for k in SERIES}
        config = json.loads(str(z['config']))
    return ref, config


def compare(ref, cand):
    """Returns (failures, report): failures is a list of messages, empty when the distributions agree."""
    from scipy import stats
    failures, report = [], {}
    for k in SERIES:
        r, c = ref[k], cand[k]
        z = stats.norm.ppf(1 - 0.025 / r.shape[1])
        se = np.sqrt(r.var(0, ddof=1) / len(r) + c.var(0, ddof=1) / len(c))
        diff = np.abs(c.mean(0) - r.mean(0))
        bad_days = np.nonzero(diff > z * se)[0]
        if len(bad_days):
            failures.append(f'{k}: daily mean differs from the reference on days {bad_days.tolist()}')
        tr, tc = r.sum(1), c.sum(1)
        half = 1.96 * np.sqrt(tr.var(ddof=1) / len(tr) + tc.var(ddof=1) / len(tc))
        if abs(tc.mean() - tr.mean()) > half:
            failures.append(f'{k}: cumulative mean {tc.mean():.2f} vs reference {tr.mean():.2f} +- {half:.2f}')
        ks = stats.ks_2samp(tr, tc).pvalue
        if ks <= 0.01:
            failures.append(f'{k}: KS p = {ks:.2g} on cumulative totals')
        report[k] = dict(ref_mean=float(tr.mean()), cand_mean=float(tc.mean()), ci_half=float(half), ks_p=float(ks),
                         max_daily_z=float(np.max(np.where(se > 0, diff / np.where(se > 0, se, 1), 0))))
    pr, pc = ref['new_infections'].argmax(1), cand['new_infections'].argmax(1)
    ks = stats.ks_2samp(pr, pc).pvalue
    if ks <= 0.01:
        failures.append(f'peak day: KS p = {ks:.2g}')
    report['peak_day'] = dict(ref_mean=float(pr.mean()), cand_mean=float(pc.mean()), ks_p=float(ks))
    return failures, report
