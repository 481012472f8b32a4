"""Volume sampling on the GPU (mbem_amd.volume.sample over the device
readout) on the box carrying the exact densities of u = c + E x
(tests/_bvp.py): every grid point inside reads the field -- displacement
c + E x and elastic stress C:E -- within the Somigliana bars, with the
near-face readout and stand-off close to the faces; points outside are NaN;
the sample equals evaluate_* at the same (stood-off) points; a section
written to .vti reads back."""
import numpy as np
import pytest

import _bvp as B

pytestmark = pytest.mark.gpu


def test_sample_exact_density_box(tmp_path):
    from mbem_amd import evaluate as ev
    from mbem_amd import volume as vol
    from mbem_amd import vti
    from mbem_amd.recover import NearField
    model, region, sol, mat = B.exact_density_model(0.25, n_cells=8)
    eps = 0.625
    near = NearField(model)
    g = vol.make_grid(model, 5.0, pad=5.0)
    code = vol.classify(model, g)
    print(f"[volume] {g.points.shape[0]} points, {int((code > 0).sum())} inside", flush=True)
    out = vol.sample(model, sol, eps, g, code, chunk=97, near=near)
    inside = code > 0
    assert np.isnan(out["u"][~inside]).all() and np.isnan(out["von_mises"][~inside]).all()
    se = B.exact_sigma(mat)
    pts = vol.stand_off(model, region, g.points[inside], eps, near)
    ue = B.RIGID_C + pts @ B.STRAIN.T
    es = np.abs(out["sigma"][inside] - se).max() / np.linalg.norm(se)
    eu = np.abs(out["u"][inside] - ue).max() / np.abs(ue).max()
    print(f"[volume] stress vs C:E {es:.1e}, displacement {eu:.1e}", flush=True)
    assert es < 5e-3 and eu < 5e-3
    # the strain is Hooke's inverse of the stress; it carries the stress's error
    # scaled by (2 mu + 3 lam) / (2 mu) in its trace
    assert np.allclose(out["strain"][inside], vol.strain_of(out["sigma"][inside], mat.mu, mat.lam))
    ee = np.abs(out["strain"][inside] - B.STRAIN).max() / np.abs(B.STRAIN).max()
    print(f"[volume] strain vs E {ee:.1e}", flush=True)
    assert ee < 2e-2
    s = ev.evaluate_stress(model, region, sol, pts, eps, warn_near=False, near=near)
    assert np.abs(out["sigma"][inside] - s).max() < 1e-12 * np.abs(s).max()
    sec = vol.section_grid(model, "y", 1.0, 5.0)
    c2 = vol.classify(model, sec)
    f2 = vol.sample(model, sol, eps, sec, c2, near=near)
    ch, ce, cf = vol.clearance(model, sec, eps)
    path = vti.write(tmp_path / "sec.vti", sec.origin, sec.spacing, sec.dims,
                     vol.as_vti_arrays(sec, f2, c2, ch, ce, cf))
    back = vti.read(path)
    assert back["dims"] == sec.dims and "sigma_xz" in back["arrays"]
