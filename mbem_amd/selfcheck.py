"""Runtime guard on the fault sign convention (mbem's src/mbem/selfcheck.py).

``FAULT_ORIENTATION`` is the one statement of how a fault's slip enters the
solve and the readouts; every site reads it through
``RegionModel.orientation``.  That makes it central, not right, so it is
pinned here, once per process, against physical statements with hard-coded
directions (selfcheck.py:17-61), on a 4 x 4 x 4 km box with five clamped
faces, a free +z face and a 2 x 2 km vertical fault with normal +x carrying
slip +y, at nu = 1/4 and nu = 0.30:

  [1] evaluate_displacement: u_y > 0 at (+d, 0, 0) and < 0 at (-d, 0, 0),
      the jump a sizeable fraction of the slip;
  [2] evaluate_stress(subtract_anelastic=False): the on-fault sigma_xy is
      POSITIVE (it carries +C:eps*);
  [3] subtracting the eigenstress changes the on-fault sigma_xy by a
      NEGATIVE amount of order (3/4) mu |slip| / eps;
  [4] generate_system + the dense solve: the free face moves toward +y where
      x > 0 and toward -y where x < 0.

Stages: "core" ([1], [4]) and "stress" ([2], [3], after "core").  Runs on the
GPU kernels, the ones every caller of this guard is about to use.
"""
from __future__ import annotations

import threading

import numpy as np

_BOX_HALF = 2.0
_FAULT_HALF = 1.0
_N = 2
_SLIP = 0.01
_EPS = 0.2
_MU = 30.0
_LAMS = (30.0, 45.0)
_OBS_D = 0.5
_MIN_JUMP_FRAC = 0.05
_MAX_JUMP_FRAC = 1.00
_EIGEN_LO, _EIGEN_HI = 0.5, 2.0

STAGES = ("core", "stress")
_lock = threading.RLock()
_state = {s: "unchecked" for s in STAGES}
_report: dict = {}
_models: dict = {}


class FaultConventionError(RuntimeError):
    """The fault sign convention does not match physical reality."""


def _fail(statement: str, measured: str) -> FaultConventionError:
    from .model.core import FAULT_ORIENTATION
    return FaultConventionError(
        f"refusing to run: the fault sign convention is backwards "
        f"(FAULT_ORIENTATION = {FAULT_ORIENTATION:+d}).\n  physical statement that FAILED: "
        f"{statement}\n  measured: {measured}\n  With fault normal +x and Burgers vector "
        f"+{_SLIP} y the +x side must move toward +y and the -x side toward -y.")


def _panel(p0, e1, e2, n):
    """Parallelogram p0 + s e1 + t e2 as n x n quads, normal +e1 x e2."""
    from .mesh import TriMesh
    p0, e1, e2 = (np.asarray(v, float) for v in (p0, e1, e2))
    i, j = np.meshgrid(np.arange(n + 1), np.arange(n + 1), indexing="ij")
    verts = (p0[None] + (i.ravel() / n)[:, None] * e1[None] + (j.ravel() / n)[:, None] * e2[None])
    tris = []
    for a_i in range(n):
        for b_j in range(n):
            a = a_i * (n + 1) + b_j
            b = (a_i + 1) * (n + 1) + b_j
            tris += [[a, b, b + 1], [a, b + 1, a + 1]]
    return TriMesh(np.ascontiguousarray(verts), np.asarray(tris, dtype=np.int64))


def _model(lam: float):
    if lam in _models:
        return _models[lam]
    from .mesh import ElasticMaterial, concatenate
    from .model import BCType, Patch, Region, RegionModel
    L, n, d = _BOX_HALF, _N, 2.0 * _BOX_HALF
    top = _panel([-L, -L, +L], [d, 0, 0], [0, d, 0], n)
    rest = concatenate([_panel([-L, -L, -L], [0, d, 0], [d, 0, 0], n),
                        _panel([+L, -L, -L], [0, d, 0], [0, 0, d], n),
                        _panel([-L, -L, -L], [0, 0, d], [0, d, 0], n),
                        _panel([-L, +L, -L], [0, 0, d], [d, 0, 0], n),
                        _panel([-L, -L, -L], [d, 0, 0], [0, 0, d], n)])
    f = _FAULT_HALF
    fault = _panel([0.0, -f, -f], [0, 2 * f, 0], [0, 0, 2 * f], n)
    nt, _ = top.normals_and_areas()
    nf, _ = fault.normals_and_areas()
    assert np.allclose(nt, [0.0, 0.0, 1.0]) and np.allclose(nf, [1.0, 0.0, 0.0])
    slip = np.broadcast_to(np.array([0.0, _SLIP, 0.0]), (fault.n_triangles, 3))
    p_top = Patch("selfcheck_top", top, BCType.FREE_TRACTION)
    p_rest = Patch("selfcheck_rest", rest, BCType.PRESCRIBED_DISPLACEMENT)
    p_fault = Patch("selfcheck_fault", fault, BCType.FAULT, value=slip)
    region = Region("selfcheck_box", ElasticMaterial(mu=_MU, lam=lam), [p_top, p_rest],
                    probe_point=np.array([0.11, 0.07, -0.13]), faults=[p_fault])
    zero = {"u:selfcheck_top": np.zeros((top.n_triangles, 3)),
            "t:selfcheck_rest": np.zeros((rest.n_triangles, 3))}
    _models[lam] = (RegionModel([region]), region, top, fault, zero)
    return _models[lam]


def _check_core_at(lam: float) -> dict:
    from .backends.dense import AssembledDense
    from .evaluate import evaluate_displacement
    from .model import generate_system
    model, region, top, fault, zero = _model(lam)
    obs = np.array([[+_OBS_D, 0.0, 0.0], [-_OBS_D, 0.0, 0.0]])
    u = evaluate_displacement(model, region, zero, obs, _EPS, warn_near=False)
    up, um = float(u[0, 1]), float(u[1, 1])
    if not (um < 0.0 < up):
        raise _fail("evaluate_displacement: u_y > 0 on the +x side and < 0 on the -x side",
                    f"u_y(+x) = {up:+.4e}, u_y(-x) = {um:+.4e}")
    frac = 0.5 * (abs(up) + abs(um)) / _SLIP
    if not (_MIN_JUMP_FRAC < frac < _MAX_JUMP_FRAC):
        raise _fail("the jump across the fault is a fraction of the slip",
                    f"|u_y| / |slip| = {frac:.4f}")
    asm = AssembledDense(generate_system(model), _EPS, "direct", jump="half")
    sol = asm.solve(method="lu")
    xc = top.centroids()[:, 0]
    uy = sol["u:selfcheck_top"][:, 1]
    sp, sm = float(uy[xc > 0.0].mean()), float(uy[xc < 0.0].mean())
    if not (sm < 0.0 < sp):
        raise _fail("generate_system + solve: the free face moves toward +y where x > 0 and "
                    "toward -y where x < 0", f"u_y(x>0) = {sp:+.4e}, u_y(x<0) = {sm:+.4e}")
    return {"u_y_plus": up, "u_y_minus": um, "jump_frac": frac, "solved_u_y_plus": sp,
            "solved_u_y_minus": sm}


def _check_stress_at(lam: float) -> dict:
    from .evaluate import evaluate_stress
    model, region, top, fault, zero = _model(lam)
    fc = fault.centroids()
    s_tot = evaluate_stress(model, region, zero, fc, _EPS, subtract_anelastic=False,
                            warn_near=False)
    s_el = evaluate_stress(model, region, zero, fc, _EPS, subtract_anelastic=True,
                           warn_near=False)
    blob = 0.75 * _MU * _SLIP / _EPS
    xy_tot = float(np.mean(s_tot[:, 0, 1]))
    xy_eig = float(np.mean(s_el[:, 0, 1] - s_tot[:, 0, 1]))
    if not (xy_tot > 0.0):
        raise _fail("the TOTAL on-fault sigma_xy carries +C:eps* and is POSITIVE",
                    f"sigma_xy = {xy_tot:+.4e} (expected about {blob:+.4e})")
    if not (-_EIGEN_HI * blob < xy_eig < -_EIGEN_LO * blob):
        raise _fail("subtracting the eigenstress REMOVES +C:eps* (a negative change of order "
                    "(3/4) mu |slip| / eps)",
                    f"change = {xy_eig:+.4e}, (3/4) mu |slip| / eps = {blob:.4e}")
    return {"sigma_xy_total": xy_tot, "sigma_xy_eigen": xy_eig, "blob_peak": blob}


_RUNNERS = {"core": _check_core_at, "stress": _check_stress_at}


def ensure_fault_convention(stage: str = "core") -> None:
    """Run ``stage`` once per process (re-entrant: a check that itself reaches
    the guard passes through); raise FaultConventionError if it fails."""
    if stage not in _state:
        raise ValueError(f"unknown self-check stage {stage!r} (expected one of {STAGES})")
    with _lock:
        if stage == "stress" and _state["core"] == "unchecked":
            ensure_fault_convention("core")
        st = _state[stage]
        if st in ("ok", "running"):
            return
        if st == "failed":
            raise FaultConventionError(_report.get(stage, "self-check failed earlier"))
        _state[stage] = "running"
        try:
            _report[stage] = {f"nu={0.5 * l / (l + _MU):.2f}": _RUNNERS[stage](l) for l in _LAMS}
        except Exception as e:
            _state[stage] = "failed"
            _report[stage] = str(e)
            raise
        _state[stage] = "ok"


def report() -> dict:
    return dict(_report)
