"""The object front end's host logic (no device calls): settings defaults and
validation against the reference's keyword set (python/cs_dmri/shore.py:58-78),
frame resolution (python/cs_dmri/_frame.py:13-34) and gradient-table
coercion."""

import inspect

import numpy as np
import pytest

from cs_dmri_amd.shore import FitSettings, Frame, GradientTable, ShoreModel

# the reference's ShoreModel keyword defaults (python/cs_dmri/shore.py:58-63)
REFERENCE_DEFAULTS = dict(regularization="l1", alpha_mode="l2-anchored", alpha=1.0, alpha_ratio=1e-3,
                          path_n_alphas=20, path_eps=None, slack=0.05, max_iter=1000, tol=1e-6, non_negative=False,
                          lambda_n=1e-8, lambda_l=1e-8, nonneg_max_iter=200, nonneg_tol=1e-9, nonneg_epsilon=1e-10,
                          allow_single_shell=False)


def test_settings_defaults_match_reference():
    kw = FitSettings().as_keywords()
    for k, v in REFERENCE_DEFAULTS.items():
        assert kw[k] == v, k
    assert kw["precision"] == "fp32" and kw["refresh_every"] is None
    sig = inspect.signature(ShoreModel.__init__)
    assert sig.parameters["radial_order"].default == 6 and sig.parameters["zeta"].default == 700.0
    assert sig.parameters["bvec_frame"].default == "auto"


def test_settings_validation():
    with pytest.raises(ValueError, match="regularization must be one of"):
        FitSettings(regularization="l3")
    with pytest.raises(ValueError, match="alpha_mode must be one of"):
        FitSettings(alpha_mode="grid")
    with pytest.raises(TypeError):
        FitSettings(alpha_max=1.0)  # unknown keyword, as with an explicit signature
    with pytest.raises(ValueError, match="alpha_mode"):
        ShoreModel((np.zeros(2), np.eye(3)[:2]), alpha_mode="nope")
    m = ShoreModel((np.array([0.0, 1000.0]), np.eye(3)[:2]), regularization="l2", lambda_n=1e-2)
    assert m.regularization == "l2" and m.settings.lambda_n == 1e-2 and m.radial_order == 6


def test_frame_resolution():
    aff = np.diag([-2.0, 3.0, 4.0, 1.0])
    aff[:3, 3] = 7.0
    f = Frame.resolve("auto", aff)
    assert f.name == "world" and f.sidecar_tag == "world-ras"
    np.testing.assert_array_equal(f.rotation, np.diag([-1.0, 1.0, 1.0]))
    np.testing.assert_array_equal(f.apply([[1.0, 2.0, 3.0]]), [[-1.0, 2.0, 3.0]])
    g = Frame.resolve("auto", None)
    assert g.name == "image" and g.sidecar_tag == "image-axis"
    np.testing.assert_array_equal(g.rotation, np.eye(3))
    assert Frame.resolve("image", aff).name == "image"
    with pytest.raises(ValueError, match="needs an affine"):
        Frame.resolve("world", None)
    with pytest.raises(ValueError, match="bvec_frame must be one of"):
        Frame.resolve("scanner", aff)
    # a rotation about z by 90 degrees with 2 mm voxels
    r = np.array([[0.0, -2.0, 0.0, 0.0], [2.0, 0.0, 0.0, 0.0], [0.0, 0.0, 2.0, 0.0], [0.0, 0.0, 0.0, 1.0]])
    np.testing.assert_allclose(Frame.resolve("world", r).apply([[1.0, 0.0, 0.0]]), [[0.0, 1.0, 0.0]])


def test_gradient_table_coercion():
    b = np.array([0.0, 1000.0, 2000.0, 3000.0])
    v = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]], float)
    t = GradientTable(b, v.T, big_delta=0.04)
    np.testing.assert_array_equal(t.bvecs, v)
    assert GradientTable.coerce(t) is t
    u = GradientTable.coerce((b, v))
    np.testing.assert_array_equal(u.bvals, b) and u.b0_threshold == 50.0

    class Dipy:
        bvals, bvecs, b0_threshold, big_delta, small_delta = b, v, 20.0, 0.03, 0.01

    w = GradientTable.coerce(Dipy())
    assert w.b0_threshold == 20.0 and w.big_delta == 0.03 and w.small_delta == 0.01
    with pytest.raises(ValueError, match="bvecs must have shape"):
        GradientTable(b, np.zeros((5, 3)))
