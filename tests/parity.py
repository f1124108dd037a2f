"""Shared parity helpers for the GPU tests.

The bar for the plain FP64 method is 1e-12 relative per conserved field:
max |gpu - ref| <= 1e-12 * max |ref| over each field (the scale of a field
that is identically zero falls back to the largest field scale of the
array). Integer work (fallback decisions, masks, step counts) must match
exactly.
"""

import numpy as np

REL = 1e-12


def field_errors(actual, expected):
    actual, expected = np.asarray(actual, dtype=float), np.asarray(expected, dtype=float)
    assert actual.shape == expected.shape, (actual.shape, expected.shape)
    if expected.ndim == 1:
        expected, actual = expected[None], actual[None]
    scales = np.array([np.max(np.abs(e)) for e in expected])
    top = scales.max() if scales.size else 1.0
    scales = np.where(scales > 0, scales, top if top > 0 else 1.0)
    return np.array([np.max(np.abs(a - e)) for a, e in zip(actual, expected)]) / scales


def assert_fields(actual, expected, rel=REL, what=""):
    err = field_errors(actual, expected)
    assert np.all(err <= rel), f"{what}: per-field relative error {err} exceeds {rel}"
    return err


def l1_rel(actual, expected):
    """Relative L1 error per field: sum|a-e| / sum|e|."""
    actual, expected = np.asarray(actual, dtype=float), np.asarray(expected, dtype=float)
    out = []
    for a, e in zip(actual, expected):
        den = np.sum(np.abs(e))
        out.append(np.sum(np.abs(a - e)) / (den if den > 0 else 1.0))
    return np.array(out)
