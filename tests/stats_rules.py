"""Family-wise acceptance rule for Monte Carlo estimates over many radii.

With fixed seeds, every radius' z-score against the exact estimand must lie
within 3 sigma, and the number of radii beyond 2 sigma may not exceed the
99.9 % quantile of Binomial(n, P(|Z| > 2) = 0.0455): a biased estimator that
shifts many radii by a fraction of a sigma fails the second clause even when
no single radius reaches 3 sigma."""
import numpy as np
from scipy.stats import binom


def mc_family_rule(z):
    z = np.abs(np.asarray(z, np.float64))
    k_max = int(binom.ppf(0.999, len(z), 0.0455))
    assert np.all(z < 3.0), z
    assert int(np.sum(z > 2.0)) <= k_max, (int(np.sum(z > 2.0)), k_max, z)
