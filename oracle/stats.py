"""Linear XEB statistics (born/verify.py:36-41, analyze_challenge.py:11-16)."""
import numpy as np


def xeb_stats(z, rng_seed=424242, resamples=20000, clip=4.0, alpha=0.05):
    """z = D * p(x) per sample (D = 2^n).  XEB = mean(z) - 1, its standard error,
    a bootstrap 95% interval (seeded) and the clipped Hoeffding lower bound."""
    z = np.asarray(z, np.float64)
    rng = np.random.default_rng(rng_seed)
    boot = np.array([z[rng.integers(0, len(z), len(z))].mean() - 1 for _ in range(resamples)])
    return dict(n=int(len(z)), xeb=float(z.mean() - 1),
                standard_error=float(z.std(ddof=1) / np.sqrt(len(z))),
                bootstrap_95=[float(np.percentile(boot, 2.5)), float(np.percentile(boot, 97.5))],
                hoeffding_lower_95=float(np.minimum(z, clip).mean() - 1
                                         - clip * np.sqrt(np.log(1 / alpha) / (2 * len(z)))))
