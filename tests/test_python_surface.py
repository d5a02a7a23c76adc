"""The Python surface on the CPU: the call signature and the settings rules.

remove_cosmics must keep L.A.Cosmic 1.4.0's signature (the reference checks
it against the installed package, test/test_reference.py:38-46; lacosmic is
not installed here, so its signature is restated below), and Cleaner rejects
the settings the reference's test rejects (test/test_cleaner.py:71-80).
"""

import inspect

import numpy as np
import pytest


def lacosmic_remove_cosmics(data, contrast, cr_threshold, neighbor_threshold, error=None, mask=None,
                            background=None, effective_gain=None, readnoise=None, maxiter=4,
                            border_mode="mirror"):
    """The signature of lacosmic.remove_cosmics in lacosmic 1.4.0."""


def test_signature_matches_lacosmic():
    from cusmic_amd import remove_cosmics
    from cusmic_amd import autograd

    assert inspect.signature(remove_cosmics) == inspect.signature(lacosmic_remove_cosmics)
    params = inspect.signature(autograd.remove_cosmics).parameters
    assert list(params) == list(inspect.signature(lacosmic_remove_cosmics).parameters)


def test_cleaner_settings():
    from cusmic_amd import Cleaner

    for settings in ({"contrast": -1}, {"cr_threshold": np.nan}, {"neighbor_threshold": np.inf},
                     {"maxiter": 1.5}, {"maxiter": True}, {"maxiter": -1}, {"border_mode": "invalid"},
                     {"contrast": True}, {"cr_threshold": "5"}):
        with pytest.raises(ValueError):
            Cleaner(**settings)
    c = Cleaner()
    assert (c.contrast, c.cr_threshold, c.neighbor_threshold, c.maxiter, c.border_mode) == (3, 5, 3, 4, "mirror")
    Cleaner(0, 0.0, 0, 0, "constant")
    Cleaner(maxiter=np.int64(7), contrast=np.float64(2.5))
    with pytest.raises(Exception):
        c.maxiter = 5  # frozen


def test_cleaner_messages():
    from cusmic_amd import Cleaner

    with pytest.raises(ValueError, match="contrast must be finite and nonnegative"):
        Cleaner(contrast=-1)
    with pytest.raises(ValueError, match="maxiter must be a nonnegative integer"):
        Cleaner(maxiter=-1)
    with pytest.raises(ValueError, match="border_mode must be one of mirror, reflect, nearest, wrap, constant"):
        Cleaner(border_mode="periodic")
