"""Learned Nu: train the predictor on wisdom from full-Nu runs, then solve new
models at the predicted Nu and judge them against their full-Nu solve.

The reference learns Nu per point while it runs (Domain::learnWisdom,
Domain.cpp:384-440, every NU_WISDOM_LEARN_INTERVAL steps with the cutoff
NU_WISDOM_REL_AMP_CUTOFF; SolidPoint.cpp:240-272) and a later run reads the
learned field back as its Nr field (WisdomNrField.cpp:20-24).  Here the
learned field of several full-Nu runs trains a perceptron (``NuMLP``) that maps
per-point model features to the learned Nu, so that a NEW model can be solved
at a reduced Nu without its own full-Nu run first.  The full-Nu solve stays the
judge of the seismogram misfit and the fallback: elements with a point whose
features lie outside the training range keep the full Nu.
"""
import time

import numpy as np

from . import _abi
from .domain import Domain
from .mesh import ShellMesh, shell_config
from .nu_mlp import NuMLP

# a solid inner core, fluid outer core and three mantle layers (the 3-D
# solid + fluid model family), 50 radial elements
LEARN_LAYERS = [
    dict(r_bottom=1000e3, r_top=1221.5e3, nelem_radial=1, fluid=0, vp_bottom=11.3e3, vp_top=11.0e3,
         vs_bottom=3.7e3, vs_top=3.5e3, rho_bottom=13.1e3, rho_top=12.8e3, q_mu=0., q_kappa=57823.),
    dict(r_bottom=1221.5e3, r_top=3480e3, nelem_radial=10, fluid=1, vp_bottom=10.3e3, vp_top=8.0e3,
         vs_bottom=0., vs_top=0., rho_bottom=12.2e3, rho_top=9.9e3, q_mu=0., q_kappa=57823.),
    dict(r_bottom=3480e3, r_top=5701e3, nelem_radial=27, fluid=0, vp_bottom=11.0e3, vp_top=10.3e3,
         vs_bottom=6.2e3, vs_top=5.6e3, rho_bottom=5.5e3, rho_top=4.4e3, q_mu=0., q_kappa=57823.),
    dict(r_bottom=5701e3, r_top=5971e3, nelem_radial=4, fluid=0, vp_bottom=10.2e3, vp_top=9.1e3,
         vs_bottom=5.5e3, vs_top=4.9e3, rho_bottom=4.0e3, rho_top=3.7e3, q_mu=0., q_kappa=57823.),
    dict(r_bottom=5971e3, r_top=6371e3, nelem_radial=8, fluid=0, vp_bottom=8.9e3, vp_top=5.8e3,
         vs_bottom=4.8e3, vs_top=3.2e3, rho_bottom=3.5e3, rho_top=2.6e3, q_mu=0., q_kappa=57823.),
]


def model_config(seed, nstep, nelem_theta=40, nu=None, elem_nu=None, radial_scale=1, wisdom=None):
    """One seeded model of the family: TI solid, GRF dvs/dvp of +-5 % with a
    500 km correlation length, seeded moment tensor and stations; uniform Nu
    `nu`, the per-element Nu `elem_nu` or the Nr field of wisdom rows `wisdom`.  The mesh has nelem_theta x (50
    radial_scale) quads; the model, source and stations depend only on the
    seed, so meshes of one seed at several resolutions describe one model."""
    layers = [dict(l, nelem_radial=l["nelem_radial"] * radial_scale) for l in LEARN_LAYERS]
    kw = dict(layers=layers, nelem_theta=nelem_theta, material=_abi.MAT_TRANSVERSELY_ISO, dv_amp=0.05,
              dv_corr_len=500e3, seed=seed, nstep=nstep, nrec=20)
    if elem_nu is not None:
        kw["elem_nu"] = elem_nu
    elif wisdom is not None:
        kw["wisdom"] = wisdom
    else:
        kw.update(nu_mode=0, nu_const=int(nu))
    return shell_config(1, **kw)


def solve(mesh, cutoff=None, interval=5):
    """Whole run on the GPU (graph replay); wisdom learning when `cutoff` is
    given.  Returns seismograms, learned Nu per point (or None) and the
    event-timed solve in seconds."""
    d = Domain(mesh)
    try:
        if cutoff is not None:
            d.setLearning(cutoff, interval)
        d.resetZero()
        d.run(1, mesh.d.nstep, 0, use_graph=True)
        sec = d.last_run_ms() * 1e-3
        wis = d.wisdom() if cutoff is not None else None
        return d.seismograms(), wis, sec
    finally:
        d.close()


def point_nu(mesh):
    return mesh.array("point_nr") // 2


def wisdom_targets(mesh, wisdom, nu_full):
    """Learned Nu over the full Nu, 0 where nothing was learned."""
    return np.maximum(wisdom, 0).astype(np.float32) / float(nu_full)


def element_nu(mesh, nu_pred_point, flagged_point, nu_full):
    """Nu of every element: the largest predicted Nu of its points, the full
    Nu where any of its points is flagged."""
    ep = mesh.array("elem_points").reshape(mesh.d.nelem, 25)
    nu = nu_pred_point[ep].max(1)
    nu[flagged_point[ep].any(1)] = nu_full
    return np.clip(nu, 1, nu_full).astype(np.int32)


class LearnedNu:
    """Training set, trained predictor and the feature range it saw."""

    def __init__(self, nu_full=128, nstep=800, nelem_theta=40, cutoff=1e-3, interval=5, margin=1.25, pad=4,
                 range_tol=0.05, seed=1):
        self.nu_full, self.nstep, self.nelem_theta = nu_full, nstep, nelem_theta
        self.cutoff, self.interval, self.margin, self.pad, self.range_tol = cutoff, interval, margin, pad, range_tol
        self.mlp = NuMLP(seed=seed)
        self.x, self.y = [], []
        self.lo = self.hi = None

    def add_training_model(self, seed):
        m = ShellMesh(model_config(seed, self.nstep, self.nelem_theta, nu=self.nu_full))
        _, wis, sec = solve(m, self.cutoff, self.interval)
        self.x.append(m.point_features())
        self.y.append(wisdom_targets(m, wis, self.nu_full))
        return sec

    def train(self, epochs=30, batch=1024, lr=1e-3, under_weight=4.0, seed=1):
        x, y = np.concatenate(self.x), np.concatenate(self.y)
        self.lo, self.hi = x.min(0), x.max(0)
        self.mlp.set_data(x, y)
        self.mlp.train(epochs, batch, lr, under_weight, seed)
        return float(np.mean((self.mlp.predict(x) - y) ** 2))

    def predict_point_nu(self, mesh):
        """Predicted Nu of every point (with margin and pad) and the points whose
        features lie outside the training range."""
        f = mesh.point_features()
        y = np.clip(self.mlp.predict(f), 0.0, 1.0)
        nu = np.ceil(self.margin * y * self.nu_full).astype(np.int32) + self.pad
        tol = self.range_tol * np.maximum(self.hi - self.lo, 1e-6)
        flagged = ((f < self.lo - tol) | (f > self.hi + tol)).any(1)
        return nu, flagged

    def predict_element_nu(self, mesh):
        nu, flagged = self.predict_point_nu(mesh)
        return element_nu(mesh, nu, flagged, self.nu_full), int(flagged.sum())

    def write_wisdom(self, mesh, path):
        """The prediction for `mesh` as a Nu wisdom file (s, z, predicted Nu, full
        Nu per point; flagged points keep the full Nu), which the reference's
        NU_TYPE wisdom Nr field reads (WisdomNrField.cpp:20-24) and so does
        shell_config(wisdom=read_wisdom(path))."""
        from .files import write_wisdom
        nu, flagged = self.predict_point_nu(mesh)
        nu = np.where(flagged, self.nu_full, np.clip(nu, 1, self.nu_full))
        sz = mesh.array("point_sz").reshape(-1, 2)
        write_wisdom(path, np.column_stack([sz[:, 0], sz[:, 1], nu, np.full(nu.shape, self.nu_full)]))

    def evaluate(self, seed, nelem_theta=None, radial_scale=1):
        """Full-Nu judge against the predicted-Nu solve of one new model
        (optionally on a finer mesh than the training models'): wall-clock to
        solution of each (mesh set-up + solve; the learned path also builds
        the feature mesh and runs the predictor) and the misfit."""
        nt = nelem_theta or self.nelem_theta
        t0 = time.perf_counter()
        mf = ShellMesh(model_config(seed, self.nstep, nt, nu=self.nu_full, radial_scale=radial_scale))
        sf, _, solve_full = solve(mf)
        wall_full = time.perf_counter() - t0
        t0 = time.perf_counter()
        m1 = ShellMesh(model_config(seed, self.nstep, nt, nu=1, radial_scale=radial_scale))  # features only
        enu, nflag = self.predict_element_nu(m1)
        mp = ShellMesh(model_config(seed, self.nstep, nt, elem_nu=enu, radial_scale=radial_scale))
        sp, _, solve_pred = solve(mp)
        wall_pred = time.perf_counter() - t0
        misfit = float(np.linalg.norm(sp - sf) / np.linalg.norm(sf))
        return dict(seed=seed, quads=int(mf.d.nelem), misfit=misfit, solve_full_s=solve_full, solve_pred_s=solve_pred,
                    wall_full_s=wall_full, wall_pred_s=wall_pred, flagged_points=nflag,
                    mean_nu_full=float(point_nu(mf).mean()), mean_nu_pred=float(point_nu(mp).mean()),
                    updates_full=int(mf.stats["sum_nu1"]), updates_pred=int(mp.stats["sum_nu1"]))
