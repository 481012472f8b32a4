"""Volume sampling bookkeeping on the CPU (mbem_amd/volume.py, vti.py; the
reference's volume.py:104-497 and vti.py): the grid's half-cell offset and
ImageData order, the region classification of the box (no sample on a face,
the probe points inside), the clearances, the stand-off, Hooke's inverse and
the derived scalars on tensors with known answers, the difference rebuilding
its scalars, and a bitwise .vti round trip."""
import numpy as np
import pytest

import _bvp as B


@pytest.fixture(scope="module")
def box():
    model, region, sol, mat = B.exact_density_model(0.25, n_cells=8)
    return model, region, sol, mat


def test_grid_offsets_and_order(box):
    from mbem_amd import volume as vol
    model = box[0]
    g = vol.make_grid(model, 4.0)
    lo, hi = vol.model_bounds(model)
    assert np.allclose(lo, [-20, -20, -20]) and np.allclose(hi, [20, 20, 0])
    assert g.dims == (10, 10, 5)
    # half a cell in from the low corner: no plane of samples on a face
    assert np.allclose(g.origin, [-18, -18, -18])
    for d, (a, b) in enumerate(zip(lo, hi)):
        assert not np.any(np.isclose(g.points[:, d], a)) and not np.any(np.isclose(g.points[:, d], b))
    # x fastest
    assert np.allclose(g.points[1] - g.points[0], [4, 0, 0])
    f = g.fold(np.arange(g.points.shape[0]))
    assert f.shape == (5, 10, 10) and f[0, 0, 1] == 1 and f[0, 1, 0] == 10
    s = vol.section_grid(model, "y", 1.0, 4.0)
    assert s.dims == (10, 1, 5) and np.all(s.points[:, 1] == 1.0)
    with pytest.raises(ValueError):
        vol.section_grid(model, "z", 5.0, 4.0)


def test_classify_and_clearance(box):
    from mbem_amd import volume as vol
    model = box[0]
    g = vol.make_grid(model, 4.0, pad=4.0)
    code = vol.classify(model, g)
    inside = np.all((g.points > [-20, -20, -20]) & (g.points < [20, 20, 0]), axis=1)
    assert np.array_equal(code > 0, inside)
    c_h, c_eps, c_f = vol.clearance(model, g, 0.625)
    # the nearest face: distance over the element size (every element alike
    # here) and over eps = 0.625
    from mbem_amd.kernels.basis import element_sizes
    h = float(element_sizes(model.regions[0].patches[0].mesh)[0])
    d = np.min(np.abs(np.column_stack([g.points[:, 0] + 20, 20 - g.points[:, 0],
                                       g.points[:, 1] + 20, 20 - g.points[:, 1],
                                       g.points[:, 2] + 20, -g.points[:, 2]])), axis=1)
    sel = inside
    assert np.allclose(c_h[sel], d[sel] / h, rtol=1e-9)
    assert np.allclose(c_eps[sel], d[sel] / 0.625, rtol=1e-9)
    assert np.all(np.isinf(c_f))


def test_stand_off_moves_face_points_inside(box):
    from mbem_amd import volume as vol
    from mbem_amd.recover import NearField
    model, region = box[0], box[1]
    near = NearField(model)
    pts = np.array([[1.0, 2.0, 0.0], [1.0, 2.0, -0.2], [1.0, 2.0, -0.8], [19.9, 3.0, -7.0]])
    out = vol.stand_off(model, region, pts, 0.625, near)
    from mbem_amd.kernels.basis import element_sizes
    gap = near.standoff_h * float(element_sizes(region.patches[0].mesh)[0])
    assert np.allclose(out[0], [1.0, 2.0, -gap]) and np.allclose(out[1], [1.0, 2.0, -gap])
    assert np.allclose(out[2], pts[2])
    assert np.allclose(out[3], [20.0 - gap, 3.0, -7.0])


def test_hooke_and_scalars():
    from mbem_amd import volume as vol
    mu, lam = 30.0, 45.0
    e = 1e-4 * np.array([[[1.0, 0.3, 0.0], [0.3, -2.0, 0.5], [0.0, 0.5, 0.7]]])
    sig = lam * np.trace(e[0]) * np.eye(3)[None] + 2 * mu * e
    assert np.allclose(vol.strain_of(sig, mu, lam), e, rtol=1e-12, atol=0)
    uni = np.zeros((1, 3, 3))
    uni[0, 0, 0] = 7.0
    d = vol.derived(np.array([[3.0, 4.0, 0.0]]), uni, np.zeros((1, 3, 3)))
    assert d["u_mag"][0] == 5.0 and np.isclose(d["von_mises"][0], 7.0)
    assert np.isclose(d["mean_stress"][0], 7.0 / 3) and np.isclose(d["max_shear"][0], 3.5)
    fa = {"u": np.ones((2, 3)), "sigma": np.stack([uni[0], 2 * uni[0]]), "strain": np.zeros((2, 3, 3))}
    fb = {"u": np.zeros((2, 3)), "sigma": np.stack([uni[0], uni[0]]), "strain": np.zeros((2, 3, 3))}
    diff = vol.difference(fa, fb, np.array([True, False]))
    assert diff["von_mises"][0] == 0.0 and np.isnan(diff["von_mises"][1])
    assert np.isnan(diff["sigma"][1]).all()


def test_vti_round_trip(tmp_path, box):
    from mbem_amd import volume as vol
    from mbem_amd import vti
    g = vol.make_grid(box[0], 8.0)
    n = g.points.shape[0]
    rng = np.random.default_rng(0)
    fields = {"u": rng.normal(size=(n, 3)), "sigma": rng.normal(size=(n, 3, 3)),
              "von_mises": rng.normal(size=n)}
    code = np.ones(n, np.int8)
    arrays = vol.as_vti_arrays(g, fields, code, np.ones(n), np.full(n, 2.0))
    assert {"u", "sigma_xx", "sigma_xz", "von_mises", "region", "clearance_h"} <= set(arrays)
    path = vti.write(tmp_path / "v.vti", g.origin, g.spacing, g.dims, arrays)
    back = vti.read(path)
    assert back["dims"] == g.dims and np.allclose(back["origin"], g.origin)
    for k, a in arrays.items():
        assert np.array_equal(back["arrays"][k], np.asarray(a, dtype="<f4")), k
