"""Failure guards of the host layer and the C ABI, on the CPU (no GPU call).

  * eps validation (basis.py:187-254): a non-finite, zero or negative eps,
    a wrong-length array, an unknown string spec and a missing dict key all
    raise ValueError before anything reaches the device;
  * the un-anchored refusal (dense.py:101-121): a calibrated jump on a model
    with no prescribed-displacement patch raises unless deflated;
  * the dense solve's condition warning (dense.py:555-562) fires above
    COND_WARN_THRESHOLD and only there;
  * the C ABI returns MBEM_EINVAL for a bad size or a missing pointer before
    it touches the device, and the Python binding turns that status into an
    MbemError naming the entry point.
"""
import ctypes
import warnings

import numpy as np
import pytest


def test_eps_validation_raises():
    from mbem_amd.kernels import basis as kb
    from mbem_amd.cases import box_fault_model
    with pytest.raises(ValueError, match="finite and > 0"):
        kb.as_eps_array(0.0, 4)
    with pytest.raises(ValueError, match="finite and > 0"):
        kb.as_eps_array(np.array([1.0, -1.0, 1.0]), 3)
    with pytest.raises(ValueError, match="finite and > 0"):
        kb.as_eps_array(np.array([1.0, np.nan]), 2)
    with pytest.raises(ValueError, match="finite and > 0"):
        kb.as_eps_array(np.inf, 2)
    with pytest.raises(ValueError, match="shape"):
        kb.as_eps_array(np.ones(3), 4)
    model, eps, _ = box_fault_model(4)
    top = model.regions[0].patches[0]
    with pytest.raises(ValueError, match="unknown eps spec"):
        kb.resolve_patch_eps("big", top)
    with pytest.raises(ValueError, match="no entry for patch"):
        kb.resolve_patch_eps({"nothing": 1.0}, top)
    # the accepted forms pass and give one positive value per source
    assert kb.resolve_patch_eps(eps, top).shape == (top.mesh.n_triangles,)
    assert np.all(kb.resolve_patch_eps("auto", top) > 0)


def _free_box():
    from mbem_amd.cases import box_meshes
    from mbem_amd.mesh import ElasticMaterial
    from mbem_amd.model import BCType, Patch, Region, RegionModel, generate_system
    top, sides, base = box_meshes(4)
    patches = [Patch("top", top, BCType.FREE_TRACTION), Patch("sides", sides, BCType.FREE_TRACTION),
               Patch("base", base, BCType.FREE_TRACTION)]
    region = Region("crust", ElasticMaterial(mu=30.0, lam=30.0), patches,
                    probe_point=np.array([1.0, 1.0, -10.0]))
    return generate_system(RegionModel([region]))


def test_unanchored_calibrated_model_is_refused():
    from mbem_amd.backends.dense import require_anchor_or_deflate
    system = _free_box()
    with pytest.raises(ValueError, match="rigid translations"):
        require_anchor_or_deflate(system, "calibrated", False)
    # deflated, or with the half jump, the same model is accepted
    require_anchor_or_deflate(system, "calibrated", True)
    require_anchor_or_deflate(system, "half", False)


def test_anchored_model_is_accepted():
    from mbem_amd.backends.dense import require_anchor_or_deflate
    from mbem_amd.cases import box_fault_model
    from mbem_amd.model import generate_system
    model, _, _ = box_fault_model(4)
    require_anchor_or_deflate(generate_system(model), "calibrated", False)


def test_condition_warning_threshold():
    from mbem_amd import defaults
    from mbem_amd.backends.dense import warn_condition
    with pytest.warns(UserWarning, match="condition estimate 3.00e\\+10"):
        assert warn_condition(3e10)
    with pytest.warns(UserWarning, match="condition estimate inf"):
        assert warn_condition(float("inf"))
    with warnings.catch_warnings():
        warnings.simplefilter("error")
        assert not warn_condition(defaults.COND_WARN_THRESHOLD)
        assert not warn_condition(6.7e3)


def test_capi_rejects_bad_arguments_without_a_device():
    from mbem_amd import _lib
    lib = _lib.load()
    null = ctypes.c_void_p(0)
    # a negative row count, then a missing matrix for a non-empty product
    assert lib.mbem_gemv(null, 0, -1, 0, null, null, 1.0, 0.0, null) == _lib.MBEM_EINVAL
    assert lib.mbem_gemv(null, 4, 4, 4, null, null, 1.0, 0.0, null) == _lib.MBEM_EINVAL
    assert lib.mbem_axpby(-5, 1.0, null, 0.0, null, null) == _lib.MBEM_EINVAL
    assert lib.mbem_multidot(null, 8, 2, 8, null, null, null, null) == _lib.MBEM_EINVAL
    # an empty product is valid and does nothing
    assert lib.mbem_gemv(null, 0, 0, 0, null, null, 1.0, 0.0, null) == _lib.MBEM_OK
    with pytest.raises(_lib.MbemError, match="mbem_gemv: invalid argument"):
        _lib.call("mbem_gemv", null, 0, -1, 0, null, null, 1.0, 0.0, null)
    with pytest.raises(_lib.MbemError, match="HIP error"):
        _lib.check(_lib.MBEM_EHIP, "probe")
    # the split far field's entry points: missing remainders, a remainder
    # pitch that breaks the 16-byte loads; an empty product is valid
    assert lib.mbem_cast_i8e(null, 4, 4, 4, null, 16, null, 1, null, 4, null) == _lib.MBEM_EINVAL
    assert lib.mbem_gemv_i8e(null, 16, null, 6, null, 1, 4, 4, null, null, 1.0, 0.0,
                             null) == _lib.MBEM_EINVAL
    assert lib.mbem_gemv_i8e(null, 16, null, 4, null, 0, 0, 0, null, null, 1.0, 0.0,
                             null) == _lib.MBEM_OK


def test_split_outer_option_is_checked():
    """The split outer operator exists for the int8 storage only, and the
    option is checked before anything touches the device."""
    from mbem_amd.la.mixed import solve_mixed
    with pytest.raises(ValueError, match="storage='int8'"):
        solve_mixed(None, storage="bf16", outer="split")
    with pytest.raises(ValueError, match="outer"):
        solve_mixed(None, storage="int8", outer="fp32")


def test_bj_inverse_descriptors_are_checked_on_the_host():
    """The chunk inversion's descriptor table (host copy) is checked before
    any pointer is used: a pitch that is not a multiple of the panel width, a
    chunk past mmax, or a working matrix past the length given are refused."""
    from mbem_amd import _lib
    lib = _lib.load()
    null = ctypes.c_void_p(0)
    nf, B = lib.mbem_bj_desc_fields(), lib.mbem_bj_panel_width()
    good = np.array([[0, 40, 0, B * 2, 0, 40, 0, 0]], dtype=np.int64)
    assert good.shape[1] == nf

    def gather(desc, mmax=64, w_len=40 * 2 * B):
        dh = ctypes.c_void_p(desc.ctypes.data)
        return lib.mbem_bj_gather_f32(null, 100, null, 40, dh, null, 1, mmax, null, w_len, null,
                                      40, null, null)

    bad_pitch = good.copy()
    bad_pitch[0, 3] = B * 2 + 1
    assert gather(bad_pitch) == _lib.MBEM_EINVAL
    assert gather(good, mmax=39) == _lib.MBEM_EINVAL
    assert gather(good, w_len=40 * 2 * B - 1) == _lib.MBEM_EINVAL
    # a valid table with no device pointers is refused too (null matrix)
    assert gather(good) == _lib.MBEM_EINVAL
    assert lib.mbem_bj_gather_f32(null, 1, null, 0, null, null, 0, 0, null, 0, null, 0, null,
                                  null) == _lib.MBEM_OK
