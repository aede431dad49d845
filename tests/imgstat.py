"""Image parity statistics against the reference's high-sample estimates
(tests/golden/highspp.npz, written by tests/golden/make_highspp.py).

Two independent Monte Carlo estimates of the same image — the fixture's
n_ref samples per pixel and a test render's n samples per pixel — differ per
pixel and channel by a zero-mean error of variance var_p * (1/n + 1/n_ref),
var_p being the variance of one sample's radiance, which the fixture
records. From that:

* channel means: the image mean per channel differs by at most 3 sigma,
  sigma^2 = sum_p var_p (1/n + 1/n_ref) / P^2 (P pixels);
* RMSE: the root mean square of the per-pixel differences, against its
  expectation sqrt(mean_p var_p (1/n + 1/n_ref)); a bias anywhere in the
  image (a wrong term, a shifted or mirrored pixel) raises the ratio;
* zero-variance pixels (every reference sample the same value: flat
  surfaces in the normals integrator, background) must agree to within
  `flat_tol` on at least 99.5 % of their channels.
"""
import json
import os

import numpy as np

FIXTURE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "highspp.npz")


def load(case):
    z = np.load(FIXTURE)
    meta = json.loads(bytes(z["meta"]).decode())[case]
    return meta, z[case + "_mean"].astype(np.float64), z[case + "_var"].astype(np.float64)


def compare(mean, n, ref_mean, ref_var, n_ref, flat_tol=2e-5):
    """Statistics of `mean` (h, w, 3; n samples per pixel) against the
    fixture's estimate; returns a dict (see the module docstring)."""
    mean = np.asarray(mean, np.float64)
    k = 1.0 / n + 1.0 / n_ref
    P = mean.shape[0] * mean.shape[1]
    diff = mean - ref_mean
    d_c = diff.reshape(-1, 3).mean(0)
    sigma_c = np.sqrt(ref_var.reshape(-1, 3).sum(0) * k) / P
    rmse = float(np.sqrt((diff ** 2).mean()))
    rmse_expected = float(np.sqrt(ref_var.mean() * k))
    flat = ref_var == 0
    flat_max = float(np.abs(diff[flat]).max()) if flat.any() else 0.0
    flat_within = float((np.abs(diff[flat]) <= flat_tol).mean()) if flat.any() else 1.0
    noisy = ~flat
    z = np.abs(diff[noisy]) / np.sqrt(ref_var[noisy] * k)
    return {"channel_diff": d_c.tolist(), "channel_sigma": sigma_c.tolist(),
            "channel_z": (np.abs(d_c) / np.maximum(sigma_c, 1e-300)).tolist(),
            "rmse": rmse, "rmse_expected": rmse_expected, "rmse_ratio": rmse / max(rmse_expected, 1e-300),
            "flat_fraction": float(flat.mean()), "flat_max_abs": flat_max, "flat_within_tol": flat_within,
            "z_max": float(z.max()) if z.size else 0.0, "z_over_5": float((z > 5).mean()) if z.size else 0.0,
            # a pixel that never varied in the reference's samples may still
            # hold rare events they missed: 99.5 % of such channels must agree
            "flat_ok": flat_within >= 0.995}


def assert_parity(st, rmse_ratio_max=1.3):
    """3 sigma per channel, RMSE at most `rmse_ratio_max` x its expectation,
    zero-variance pixels within flat_tol."""
    assert all(z <= 3.0 for z in st["channel_z"]), st
    assert st["rmse_ratio"] <= rmse_ratio_max, st
    assert st["flat_ok"], st
