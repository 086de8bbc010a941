"""Learned XAS surrogate with exact-ED fallback (no reference counterpart).

A parameter vector of a NiO-like 10-bath model,

    (10Dq, F0, F2/F4 scale, Delta, V_eg, V_t2g, xi_3d, T),

maps to its thermal, polarization-summed L2,3 XAS spectrum on a fixed
1000-point grid, normalized to unit area.  The exact spectra come from the
batched GPU ED (:func:`impuritymodel_amd.batched.xas_batch`); the network is
an ensemble of MLPs trained and evaluated by the library's own BF16-MFMA
kernels (``imped_mlp_*``), and the fitting loop sends a parameter set to the
surrogate when the ensemble agrees and to exact ED otherwise.

The model family is affine in its first seven parameters (every Hamiltonian
term is a Slater integral, a crystal-field, hopping, spin-orbit or
double-counting coefficient times a fixed operator), so the operator of any
parameter set is ``O_0 + sum_i theta_i O_i`` over one shared term structure:
models are built by one matrix product instead of operator algebra.
"""

from collections import OrderedDict

import numpy as np

from . import operators as ops
from .models import get_restrictions
from .opcompile import TermTable

PARAM_NAMES = ("tenDq", "F0", "Fscale", "delta", "V_eg", "V_t2g", "xi_3d", "T")
PARAM_RANGES = np.array([
    [0.3, 1.5],    # 10Dq (eV)
    [5.0, 9.0],    # F0_dd (eV)
    [0.6, 1.0],    # scale of F2_dd = 10 eV and F4_dd = 6.2 eV
    [1.0, 8.0],    # charge-transfer correction of the double counting (eV)
    [1.0, 3.0],    # eg bath hopping (eV)
    [0.5, 2.0],    # t2g bath hopping (eV)
    [0.04, 0.12],  # 3d spin-orbit coupling (eV)
    [10.0, 600.0],  # temperature (K)
])

OMEGA = np.linspace(-20.0, 30.0, 1000)  # XAS grid (eV)
DELTA = 0.3                             # Lorentzian half width (eV)


def cubic_bath_h0(e_imp, tenDq, eps_eg, eps_t2g, v_eg, v_t2g, l=2):
    """3d crystal field plus one bath state per cubic 3d spin-orbital.

    Bath state b = 5 s + k couples to cubic orbital k (0, 1 = eg; 2, 3, 4 =
    t2g of the spherical->cubic matrix) with spin s; c_k = sum_m conj(u[m, k]) c_m.
    """
    e_eg = e_imp + 3 / 5 * tenDq
    e_t2g = e_imp - 2 / 5 * tenDq
    u = ops.get_spherical_2_cubic_matrix(spinpol=False, l=l)
    h = np.dot(u, np.dot(np.diag([e_eg, e_eg, e_t2g, e_t2g, e_t2g]).astype(complex), np.conj(u.T)))
    h = 0.5 * (h + np.conj(h.T))
    op = {}
    for i, mi in enumerate(range(-l, l + 1)):
        for j, mj in enumerate(range(-l, l + 1)):
            if h[i, j] != 0:
                for s in range(2):
                    op[(((l, s, mi), "c"), ((l, s, mj), "a"))] = h[i, j]
    for s in range(2):
        for k in range(5):
            b = 5 * s + k
            eg = k < 2
            eps, v = (eps_eg, v_eg) if eg else (eps_t2g, v_t2g)
            if eps != 0:
                op[(((l, b), "c"), ((l, b), "a"))] = complex(eps)
            for i, m in enumerate(range(-l, l + 1)):
                c = v * np.conj(u[i, k])
                if c != 0 and v != 0:
                    op[(((l, b), "c"), ((l, s, m), "a"))] = complex(c)
                    op[(((l, s, m), "c"), ((l, b), "a"))] = complex(np.conj(c))
    return op


class FamilyModel:
    """One member of :class:`XASFamily` (duck-types ImpurityModel for xas_batch)."""

    def __init__(self, family, theta, coef):
        self.family = family
        self.theta = theta
        self._coef = coef
        self.n_bath = 10
        self.nBaths = family.nBaths
        self.nValBaths = family.nValBaths
        self.n0imps = family.n0imps
        self.dnTols = family.dnTols
        self.dnValBaths = family.dnValBaths
        self.dnConBaths = family.dnConBaths
        self.restrictions = family.restrictions
        self.params = dict(zip(PARAM_NAMES, theta))
        # dipole tables are identical for the whole family
        self._dip_tables = family._dip_tables

    @property
    def n_spin_orbitals(self):
        return ops.n_spin_orbitals(self.nBaths)

    @property
    def hOp(self):
        return {k: complex(v) for k, v in zip(self.family.keys, self._coef) if v != 0}

    def term_table(self, W):
        return self.family.table(W).with_coef(self._coef)


class XASFamily:
    """Affine family hOp(theta) = O_0 + sum_i theta_i O_i of 10-bath models."""

    def __init__(self, e_imp=-1.31796, eps_eg=-2.0, eps_t2g=-1.5, Fpd=(8.5, 0.0, 6.0), Gpd=(0.0, 4.5, 0.0, 2.5),
                 xi_2p=11.5, hField=(0.0, 0.0, 1e-4)):
        self.fixed = dict(e_imp=e_imp, eps_eg=eps_eg, eps_t2g=eps_t2g, Fpd=Fpd, Gpd=Gpd, xi_2p=xi_2p, hField=hField)
        self.nBaths = OrderedDict([(1, 0), (2, 10)])
        self.nValBaths = OrderedDict([(1, 0), (2, 10)])
        self.n0imps = OrderedDict([(1, 6), (2, 8)])
        self.dnTols = OrderedDict([(1, 0), (2, 2)])
        self.dnValBaths = OrderedDict([(1, 0), (2, 2)])
        self.dnConBaths = OrderedDict([(1, 0), (2, 0)])
        self.restrictions = get_restrictions(self.nBaths, self.nValBaths, self.n0imps, self.dnTols,
                                             self.dnValBaths, self.dnConBaths)
        base = self.hop(np.zeros(7))
        units = [self.hop(np.eye(7)[i]) for i in range(7)]
        keys = list(base.keys())
        seen = set(keys)
        for u in units:
            for k in u:
                if k not in seen:
                    seen.add(k)
                    keys.append(k)
        self.keys = keys
        self.c0 = np.array([base.get(k, 0.0) for k in keys], np.complex128)
        self.cmat = np.array([[u.get(k, 0.0) for k in keys] for u in units], np.complex128) - self.c0[None, :]
        self._tables = {}
        self._dip_tables = {}
        self._bases = {}
        self._ops = {}
        self._sectors = {}

    def hop(self, theta):
        """The operator dict of one parameter set, assembled term by term."""
        tenDq, F0, fs, delta, v_eg, v_t2g, xi_3d = (float(x) for x in theta[:7])
        f = self.fixed
        h0 = cubic_bath_h0(f["e_imp"], tenDq, f["eps_eg"], f["eps_t2g"], v_eg, v_t2g)
        Fdd = (F0, 0.0, 10.0 * fs, 0.0, 6.2 * fs)
        return ops.get_hamiltonian_operator_from_h0(
            self.nBaths, [Fdd, (0.0, 0.0, 0.0), f["Fpd"], f["Gpd"]], [f["xi_2p"], xi_3d], [self.n0imps, delta],
            f["hField"], h0)

    def coef(self, thetas):
        """Term coefficients [M, n_terms] of parameter sets [M, >=7]."""
        th = np.asarray(thetas, np.float64).reshape(-1, np.asarray(thetas).shape[-1])[:, :7]
        return self.c0[None, :] + th @ self.cmat

    def table(self, W):
        if W not in self._tables:
            self._tables[W] = TermTable(dict(zip(self.keys, [1.0] * len(self.keys))), self.restrictions, W)
        return self._tables[W]

    def batch_ops(self, thetas, W=1):
        """Device operators for a batch of parameter sets: (H of every set, the
        x/y/z dipoles of every set).  The structure is compiled once per batch
        capacity (a power of two >= 16); a batch uploads only theta, and the
        term values c0 + theta @ cmat are formed on the device.  The returned
        operators are shared: their values belong to the last call."""
        from .batched import _dipole_tables
        from .engine import Ops

        th = np.ascontiguousarray(np.asarray(thetas, np.float64).reshape(len(thetas), -1)[:, :7])
        cap = max(16, 1 << (len(th) - 1).bit_length())
        tab = self.table(W)
        if (cap, W) not in self._ops:
            if len(self._ops) >= 8:
                self._ops.pop(next(iter(self._ops)))
            n = ops.n_spin_orbitals(self.nBaths)
            m0 = FamilyModel(self, np.zeros(len(PARAM_NAMES)), self.c0)
            self._ops[(cap, W)] = (Ops([tab.with_coef(self.c0)] * cap, None, n, W),
                                   Ops(_dipole_tables([m0], W) * cap, None, n, W))
        ops_h, ops_t = self._ops[(cap, W)]
        ops_h.set_affine(self.c0[tab.src], self.cmat[:, tab.src], th)
        return ops_h, ops_t

    def sectors(self, W=1):
        """The family's sector cache for W words (built on first use), or None
        when its sectors do not have the structure the cache needs."""
        from .family_sectors import FamilySectors, SectorCacheUnavailable

        if W not in self._sectors:
            try:
                self._sectors[W] = FamilySectors(self, W)
            except SectorCacheUnavailable:
                self._sectors[W] = None
        return self._sectors[W]

    def models(self, thetas):
        thetas = np.asarray(thetas, np.float64)
        cf = self.coef(thetas)
        return [FamilyModel(self, thetas[i], cf[i]) for i in range(len(thetas))]


class Ensemble:
    """Handle of an ensemble of MLPs in libimped (``imped_mlp_*``)."""

    def __init__(self, members=5, dims=(8, 512, 512, 512, 1000), max_batch=1024, seed=0):
        from ._lib import C, check, lib, ptr

        self._lib, self._check, self._ptr = lib, check, ptr
        self.E = int(members)
        self.dims = tuple(int(d) for d in dims)
        d = np.array(self.dims, np.int32)
        h = C.c_void_p()
        check(lib.imped_mlp_create(self.E, len(d) - 1, ptr(d), int(max_batch), C.byref(h)))
        self.h = h
        n = np.zeros(1, np.int64)
        check(lib.imped_mlp_param_count(self.h, ptr(n)))
        self.P = int(n[0])
        self.set_params(self.init_params(seed))

    def __del__(self):
        self._free_pinned()
        h = getattr(self, "h", None)
        lib = getattr(self, "_lib", None)
        if h and lib is not None:
            lib.imped_mlp_destroy(h)
        self.h = None

    def init_params(self, seed):
        """He-normal weights (sqrt(2/fan_in); sqrt(1/fan_in) for the output layer), zero biases."""
        rng = np.random.default_rng(seed)
        p = np.zeros((self.E, self.P), np.float32)
        L = len(self.dims) - 1
        for e in range(self.E):
            off = 0
            for l in range(L):
                i, o = self.dims[l], self.dims[l + 1]
                std = np.sqrt((2.0 if l + 1 < L else 1.0) / i)
                p[e, off: off + o * i] = rng.normal(0.0, std, o * i)
                off += o * i + o
        return p

    def set_params(self, params):
        p = np.ascontiguousarray(params, np.float32).reshape(self.E, self.P)
        self._check(self._lib.imped_mlp_set_params(self.h, self._ptr(p)))

    def get_params(self):
        p = np.zeros((self.E, self.P), np.float32)
        self._check(self._lib.imped_mlp_get_params(self.h, self._ptr(p)))
        return p

    def forward(self, x):
        """Outputs of every member [E, B, dims[-1]] for inputs x [B, dims[0]]."""
        x = np.ascontiguousarray(x, np.float32).reshape(-1, self.dims[0])
        y = np.zeros((self.E, len(x), self.dims[-1]), np.float32)
        self._check(self._lib.imped_mlp_forward(self.h, self._ptr(x), len(x), self._ptr(y)))
        return y

    def _pinned(self, B):
        """Page-locked output arrays for B samples (mean [B, D], disagreement [B], misfit [B]), grown on demand."""
        import ctypes

        D = self.dims[-1]
        if getattr(self, "_pin_rows", 0) < B:
            self._free_pinned()
            rows = max(B, 256)
            ptr = ctypes.c_void_p()
            self._check(self._lib.imped_pinned_alloc(rows * (D + 2) * 8, ctypes.byref(ptr)))
            self._pin_ptr, self._pin_rows = ptr, rows
            buf = np.ctypeslib.as_array(ctypes.cast(ptr, ctypes.POINTER(ctypes.c_double)), shape=(rows * (D + 2),))
            self._pin_mean = buf[: rows * D].reshape(rows, D)
            self._pin_dis = buf[rows * D: rows * (D + 1)]
            self._pin_l1 = buf[rows * (D + 1):]
        return self._pin_mean[:B], self._pin_dis[:B], self._pin_l1[:B]

    def _free_pinned(self):
        p = getattr(self, "_pin_ptr", None)
        lib = getattr(self, "_lib", None)
        if p is not None and lib is not None:
            lib.imped_pinned_free(p)
        self._pin_ptr, self._pin_rows = None, 0

    def predict(self, x, weights, scale):
        """Ensemble statistics on the device: (mean [B, D], disagreement [B]).

        Members m_e = max(y_e / scale, 0) normalized to sum_d weights_d m_e,d = 1;
        mean = their average renormalized the same way; disagreement = the
        members' mean of sum_d weights_d |m_e,d - mean_d|.  The results are
        read back into page-locked memory and returned as copies."""
        x = np.ascontiguousarray(x, np.float32).reshape(-1, self.dims[0])
        w = np.ascontiguousarray(weights, np.float64).reshape(self.dims[-1])
        mean, dis, _ = self._pinned(len(x))
        self._check(self._lib.imped_mlp_predict(self.h, self._ptr(x), len(x), self._ptr(w), float(scale),
                                                self._ptr(mean), self._ptr(dis)))
        return mean.copy(), dis.copy()

    def misfit(self, x, weights, scale, target, want_mean=False):
        """(disagreement [B], misfit [B]) with misfit_b = sum_d weights_d |mean_b,d - target_d|,
        all on the device; the means [B, D] come back as a third item only when asked for."""
        x = np.ascontiguousarray(x, np.float32).reshape(-1, self.dims[0])
        w = np.ascontiguousarray(weights, np.float64).reshape(self.dims[-1])
        t = np.ascontiguousarray(target, np.float64).reshape(self.dims[-1])
        mean, dis, l1 = self._pinned(len(x))
        self._check(self._lib.imped_mlp_misfit(self.h, self._ptr(x), len(x), self._ptr(w), float(scale),
                                               self._ptr(t), self._ptr(dis), self._ptr(l1),
                                               self._ptr(mean) if want_mean else None))
        if want_mean:
            return dis.copy(), l1.copy(), mean.copy()
        return dis.copy(), l1.copy()

    def gradients(self, x, t):
        x = np.ascontiguousarray(x, np.float32).reshape(-1, self.dims[0])
        t = np.ascontiguousarray(t, np.float32).reshape(len(x), self.dims[-1])
        g = np.zeros((self.E, self.P), np.float32)
        loss = np.zeros(self.E, np.float64)
        self._check(self._lib.imped_mlp_gradients(self.h, self._ptr(x), self._ptr(t), len(x), self._ptr(g),
                                                  self._ptr(loss)))
        return loss, g

    def train(self, x, t, epochs, batch=1024, lr=1e-3, lr_final=None, seed=0, chunk=200):
        """Adam on minibatches of shuffled rows; cosine decay from lr to lr_final.

        Returns the loss of every step [steps, E]."""
        x = np.ascontiguousarray(x, np.float32).reshape(-1, self.dims[0])
        t = np.ascontiguousarray(t, np.float32).reshape(len(x), self.dims[-1])
        n = len(x)
        self._check(self._lib.imped_mlp_set_data(self.h, self._ptr(x), self._ptr(t), n))
        rng = np.random.default_rng(seed)
        batch = min(batch, n)
        per_epoch = max(n // batch, 1)
        idx = np.concatenate([rng.permutation(n)[: per_epoch * batch] for _ in range(epochs)]).astype(np.int32)
        steps = per_epoch * epochs
        idx = idx.reshape(steps, batch)
        lr_final = lr if lr_final is None else lr_final
        losses = np.zeros((steps, self.E))
        for s0 in range(0, steps, chunk):
            s1 = min(steps, s0 + chunk)
            frac = (s0 + s1) / 2 / max(steps, 1)
            rate = lr_final + 0.5 * (lr - lr_final) * (1 + np.cos(np.pi * frac))
            part = np.zeros((s1 - s0, self.E))
            self._check(self._lib.imped_mlp_train(self.h, self._ptr(np.ascontiguousarray(idx[s0:s1])), s1 - s0,
                                                  batch, float(rate), 0.9, 0.999, 1e-8, self._ptr(part)))
            losses[s0:s1] = part
        return losses


def sample_params(n, seed=0):
    """Parameter sets drawn uniformly from PARAM_RANGES."""
    rng = np.random.default_rng(seed)
    lo, hi = PARAM_RANGES[:, 0], PARAM_RANGES[:, 1]
    return lo + (hi - lo) * rng.uniform(size=(n, len(lo)))


def normalize_params(p):
    """Map parameters to [-1, 1] per coordinate (network input)."""
    lo, hi = PARAM_RANGES[:, 0], PARAM_RANGES[:, 1]
    return (2.0 * (np.asarray(p) - lo) / (hi - lo) - 1.0).astype(np.float32)


def trapezoid_weights(omega):
    """w with sum_d w_d f_d = the trapezoid rule of f on the grid omega."""
    x = np.asarray(omega, np.float64)
    d = np.diff(x)
    w = np.zeros_like(x)
    w[:-1] += 0.5 * d
    w[1:] += 0.5 * d
    return w


_W_OMEGA = trapezoid_weights(OMEGA)


def _weights(omega):
    return _W_OMEGA if omega is OMEGA else trapezoid_weights(omega)


def area_normalize(spectra, omega=OMEGA):
    """Spectra scaled to unit trapezoid area on omega."""
    s = np.asarray(spectra, np.float64)
    return s / (s @ _weights(omega))[..., None]


def l1_error(a, b, omega=OMEGA):
    """L1 distance of area-normalized spectra: sum |a - b| d omega (trapezoid rule)."""
    return np.abs(area_normalize(a, omega) - area_normalize(b, omega)) @ _weights(omega)


SCALE = float(OMEGA[-1] - OMEGA[0])  # unit-area spectra times the window: mean value 1


class Surrogate:
    """Ensemble network over an :class:`XASFamily` with a calibrated trust threshold.

    ``predict`` returns the area-normalized ensemble mean and the ensemble
    disagreement: the mean L1 distance of the members' spectra from the
    mean.  Predictions with disagreement below ``threshold`` are trusted.
    """

    def __init__(self, family, ensemble):
        self.family = family
        self.net = ensemble
        self.threshold = 0.0

    def fit(self, params, spectra, epochs=200, batch=1024, lr=2e-3, lr_final=2e-5, seed=0):
        return self.net.train(normalize_params(params), np.asarray(spectra) * SCALE, epochs, batch, lr, lr_final,
                              seed)

    def predict(self, params):
        """(area-normalized ensemble mean [B, n_w], disagreement [B]), computed on the device."""
        return self.net.predict(normalize_params(params), _W_OMEGA, SCALE)

    def misfit(self, params, target_n):
        """(disagreement [B], L1 distance of the ensemble mean from the unit-area
        spectrum target_n [B]), computed on the device; the means stay there."""
        return self.net.misfit(normalize_params(params), _W_OMEGA, SCALE, target_n)

    def calibrate(self, params, exact, bar=1e-2):
        """Largest threshold under which every validation prediction meets the L1 bar."""
        pred, dis = self.predict(params)
        err = l1_error(pred, exact)
        order = np.argsort(dis)
        ok = np.cumprod(err[order] <= bar).astype(bool)
        k = int(ok.sum())
        self.threshold = float(dis[order[k - 1]]) * (1 + 1e-9) if k else 0.0
        return err, dis


def train_surrogate(family, n_train=20000, epochs=2500, n_val=2048, members=5, seed=0, bar=1e-2, margin=0.8,
                    chunk=4096, **fit_kw):
    """Training set by batched exact ED, ensemble training, threshold calibration.

    The trust threshold is calibrated to ``margin * bar`` on the validation
    set, leaving room for proposals that sit where validation is sparse.
    Returns ``(surrogate, info)``; info holds the wall times of each stage
    and the validation errors.
    """
    import time

    info = {}
    t0 = time.perf_counter()
    P = sample_params(n_train, seed=seed + 10)
    Y = np.concatenate([exact_spectra(family, P[i:i + chunk])[0] for i in range(0, n_train, chunk)])
    Pv = sample_params(n_val, seed=seed + 11)
    Yv = np.concatenate([exact_spectra(family, Pv[i:i + chunk])[0] for i in range(0, n_val, chunk)])
    info["train_set_seconds"] = time.perf_counter() - t0
    t0 = time.perf_counter()
    sur = Surrogate(family, Ensemble(members, (8, 512, 512, 512, len(OMEGA)), max_batch=1024, seed=seed))
    losses = sur.fit(P, Y, epochs=epochs, seed=seed, **fit_kw)
    info["train_seconds"] = time.perf_counter() - t0
    info["train_steps"] = int(len(losses))
    info["final_loss"] = float(losses[-50:].mean())
    err, dis = sur.calibrate(Pv, Yv, bar=bar * margin)
    info.update(n_train=n_train, epochs=epochs, n_val=n_val, threshold=sur.threshold,
                val_l1_median=float(np.median(err)), val_trusted=float((dis < sur.threshold).mean()),
                val_trusted_l1_max=float(err[dis < sur.threshold].max()) if (dis < sur.threshold).any() else None)
    return sur, info


class FittingLoop:
    """Metropolis walkers in parameter space fitting a target spectrum.

    ``step()`` lets every walker that is not waiting propose a move.  With a
    surrogate, trusted proposals are answered at once (one ensemble forward
    pass for all of them) and the walker takes its Metropolis decision; an
    untrusted proposal makes its walker wait.  Waiting proposals are evaluated
    together by one batched exact-ED call once ``flush_at`` of them have
    gathered (every step without a surrogate), so the fixed latency of an
    exact-ED launch sequence is paid once per flush instead of once per step.
    ``finish()`` evaluates what is still waiting.  Counters and the
    surrogate-served parameters are kept for throughput and for scoring
    against exact ED.

    A trusted proposal's misfit (the L1 distance of the ensemble mean from the
    target) is computed on the device, next to the prediction; only the
    disagreement and the misfit come back to the host.  With
    ``keep_spectra`` the served spectra are read back too and kept;
    otherwise ``served()`` predicts them again from the kept parameters (the
    network computes each sample's row on its own, so the rows are the same
    bit for bit).
    """

    def __init__(self, family, target, surrogate=None, n_walkers=256, step=0.05, temperature=0.02, seed=0,
                 flush_at=None, keep_spectra=False, rng=None, **xas_kw):
        self.family, self.target, self.surrogate = family, np.asarray(target), surrogate
        self.target_n = area_normalize(self.target)
        self.keep_spectra = keep_spectra
        self.step_size, self.temperature, self.xas_kw = step, temperature, xas_kw
        self.flush_at = n_walkers // 2 if flush_at is None else flush_at
        # rng: anything with normal(size=) and uniform(size=) (numpy's generator by default)
        self.rng = np.random.default_rng(seed) if rng is None else rng
        self.lo, self.hi = PARAM_RANGES[:, 0], PARAM_RANGES[:, 1]
        self.reset_counters()
        self.theta = sample_params(n_walkers, seed=seed + 1)
        self.cur = l1_error(self._exact(self.theta), self.target[None])
        self.best = float(self.cur.min())
        self.waiting = np.zeros(n_walkers, bool)
        self.proposal = self.theta.copy()

    def reset_counters(self):
        self.n_spectra = self.n_exact = self.n_surrogate = self.n_flushes = self.n_predicted = 0
        self.sur_params, self.sur_spectra, self.exact_results = [], [], []

    def _exact(self, P):
        ex, res = exact_spectra(self.family, P, **self.xas_kw)
        self.exact_results.append(res)
        self.n_spectra += len(P)
        self.n_exact += len(P)
        return ex

    def _misfit(self, spectra):
        # spectra (surrogate means and exact results) come with unit area already
        return np.abs(spectra - self.target_n[None]) @ _W_OMEGA

    def _decide(self, idx, new):
        acc = self.rng.uniform(size=len(idx)) < np.exp(-(new - self.cur[idx]) / self.temperature)
        take = idx[acc]
        self.theta[take] = self.proposal[take]
        self.cur[take] = new[acc]
        self.best = min(self.best, float(self.cur.min()))

    def flush(self):
        idx = np.nonzero(self.waiting)[0]
        if len(idx):
            self._decide(idx, self._misfit(self._exact(self.proposal[idx])))
            self.waiting[idx] = False
            self.n_flushes += 1

    def step(self):
        lo, hi = self.lo, self.hi
        free = np.nonzero(~self.waiting)[0]
        prop = self.theta[free] + self.step_size * (hi - lo) * self.rng.normal(size=(len(free), len(lo)))
        prop = np.where(prop < lo, 2 * lo - prop, prop)
        prop = np.where(prop > hi, 2 * hi - prop, prop)
        self.proposal[free] = np.clip(prop, lo, hi)
        if self.surrogate is not None and len(free):
            if self.keep_spectra:
                pred, dis = self.surrogate.predict(self.proposal[free])
                new = self._misfit(pred)
            else:
                dis, new = self.surrogate.misfit(self.proposal[free], self.target_n)
            self.n_predicted += len(free)
            use = dis < self.surrogate.threshold
            self.sur_params.append(self.proposal[free[use]])
            if self.keep_spectra:
                self.sur_spectra.append(pred[use])
            self.n_spectra += int(use.sum())
            self.n_surrogate += int(use.sum())
            self._decide(free[use], new[use])
            self.waiting[free[~use]] = True
        else:
            self.waiting[free] = True
        if self.surrogate is None or self.waiting.sum() >= self.flush_at:
            self.flush()

    def finish(self):
        self.flush()

    def served(self):
        """(params, spectra) the surrogate produced since the last counter reset."""
        if not self.sur_params:
            return np.zeros((0, len(PARAM_NAMES))), np.zeros((0, len(OMEGA)))
        params = np.concatenate(self.sur_params)
        if self.keep_spectra:
            return params, np.concatenate(self.sur_spectra)
        chunk = 1024
        return params, np.concatenate([self.surrogate.predict(params[i:i + chunk])[0]
                                       for i in range(0, len(params), chunk)])


class NativeFittingLoop:
    """:class:`FittingLoop` with a surrogate, its walkers stepped inside
    libimped (``imped_walkers_*``): one library call per step proposes for
    every walker that is not waiting, predicts all proposals with their
    misfits on the device and takes the Metropolis decisions.  Waiting
    proposals are evaluated by batched exact ED here, as in FittingLoop.

    The rules, the order in which random numbers are used and the arithmetic
    are FittingLoop's; the numbers come from two SplitMix64 streams in the
    library: standard normals (Box-Muller, two from each pair of uniforms) from the stream
    seeded with ``seed``, drawn ahead while the GPU predicts, and one uniform
    per decision from the stream seeded with ``~seed``.  FittingLoop driven
    by the same streams (its ``rng`` argument) walks the same way.
    """

    def __init__(self, family, target, surrogate, n_walkers=256, step=0.05, temperature=0.02, seed=0,
                 flush_at=None, **xas_kw):
        from ._lib import C, check, lib, ptr

        self._lib, self._check, self._ptr = lib, check, ptr
        self.family, self.target, self.surrogate = family, np.asarray(target), surrogate
        self.target_n = area_normalize(self.target)
        self.xas_kw = xas_kw
        self.n_walkers = int(n_walkers)
        self.flush_at = n_walkers // 2 if flush_at is None else flush_at
        self.n_exact = self.n_flushes = 0
        self.exact_results = []
        theta = np.ascontiguousarray(sample_params(n_walkers, seed=seed + 1))
        cur = np.ascontiguousarray(l1_error(self._exact(theta), self.target[None]))
        lo = np.ascontiguousarray(PARAM_RANGES[:, 0])
        hi = np.ascontiguousarray(PARAM_RANGES[:, 1])
        h = C.c_void_p()
        check(lib.imped_walkers_create(surrogate.net.h, self.n_walkers, len(PARAM_NAMES), ptr(theta), ptr(cur),
                                       ptr(lo), ptr(hi), float(step), float(temperature),
                                       float(surrogate.threshold), int(seed) % (1 << 64), C.byref(h)))
        self.h = h
        self._wt = np.ascontiguousarray(_W_OMEGA, np.float64)
        self._tg = np.ascontiguousarray(self.target_n, np.float64)
        self._nw = np.zeros(1, np.int32)
        self._step_args = (self.h, ptr(self._wt), float(SCALE), ptr(self._tg), ptr(self._nw))

    def __del__(self):
        h = getattr(self, "h", None)
        lib = getattr(self, "_lib", None)
        if h and lib is not None:
            lib.imped_walkers_destroy(h)
        self.h = None

    def _exact(self, P):
        ex, res = exact_spectra(self.family, P, **self.xas_kw)
        self.exact_results.append(res)
        self.n_exact += len(P)
        return ex

    def reset_counters(self):
        self.n_exact = self.n_flushes = 0
        self.exact_results = []
        self._check(self._lib.imped_walkers_reset_counters(self.h))

    def _state(self):
        theta = np.zeros((self.n_walkers, len(PARAM_NAMES)))
        cur = np.zeros(self.n_walkers)
        best = np.zeros(1)
        cnt = np.zeros(4, np.int64)
        p = self._ptr
        self._check(self._lib.imped_walkers_state(self.h, p(theta), p(cur), p(best), p(cnt)))
        return theta, cur, float(best[0]), cnt

    @property
    def theta(self):
        return self._state()[0]

    @property
    def cur(self):
        return self._state()[1]

    @property
    def best(self):
        return self._state()[2]

    @property
    def n_predicted(self):
        return int(self._state()[3][0])

    @property
    def n_surrogate(self):
        return int(self._state()[3][1])

    @property
    def n_spectra(self):
        return self.n_surrogate + self.n_exact

    def step(self):
        self._check(self._lib.imped_walkers_step(*self._step_args))
        if self._nw[0] >= self.flush_at:
            self.flush()

    def flush(self):
        W, P = self.n_walkers, len(PARAM_NAMES)
        idx = np.zeros(W, np.int32)
        props = np.zeros((W, P))
        cnt = np.zeros(1, np.int32)
        p = self._ptr
        self._check(self._lib.imped_walkers_waiting(self.h, p(idx), p(props), p(cnt)))
        k = int(cnt[0])
        if k:
            # exact spectra come with unit area
            mis = np.ascontiguousarray(np.abs(self._exact(props[:k]) - self.target_n[None]) @ _W_OMEGA)
            self._check(self._lib.imped_walkers_decide(self.h, k, p(np.ascontiguousarray(idx[:k])), p(mis)))
            self.n_flushes += 1

    def finish(self):
        self.flush()

    def set_record_cap(self, max_rows):
        """Keep at most ``max_rows`` served proposals for ``served()`` (the
        library's default is 2^20)."""
        self._check(self._lib.imped_walkers_set_record_cap(self.h, int(max_rows)))
        self._record_cap = int(max_rows)

    def served(self):
        """(params, spectra) the surrogate produced since the last counter reset
        (the first record cap of them); the spectra are predicted again (rows
        do not depend on their batch)."""
        n = min(self.n_surrogate, getattr(self, "_record_cap", 1 << 20))
        params = np.zeros((n, len(PARAM_NAMES)))
        if n:
            self._check(self._lib.imped_walkers_served(self.h, 0, n, self._ptr(params)))
            chunk = 1024
            return params, np.concatenate([self.surrogate.predict(params[i:i + chunk])[0]
                                           for i in range(0, n, chunk)])
        return params, np.zeros((0, len(OMEGA)))


def fitting_loop(family, target, surrogate=None, n_walkers=256, n_iter=8, step=0.05, temperature=0.02, seed=0,
                 **xas_kw):
    """Run ``n_iter`` steps of :class:`FittingLoop`; counts, wall time, best L1
    and the surrogate-served (params, spectra) in a dict."""
    import time

    t0 = time.perf_counter()
    loop = FittingLoop(family, target, surrogate, n_walkers, step, temperature, seed, **xas_kw)
    for _ in range(n_iter):
        loop.step()
    loop.finish()
    sp, ss = loop.served()
    return dict(n_spectra=loop.n_spectra, n_exact=loop.n_exact, n_surrogate=loop.n_surrogate,
                seconds=time.perf_counter() - t0, best_l1=loop.best, sur_params=sp, sur_spectra=ss)


def exact_spectra(family, thetas, krylovSize=200, reorth=1, sectors=True, **kw):
    """Exact XAS of parameter sets by the batched GPU ED, area-normalized.

    With ``sectors`` (the default) the family's sector cache
    (:class:`impuritymodel_amd.family_sectors.FamilySectors`) supplies every
    sector's pattern and only the values and vectors are formed per call;
    families whose sectors the cache cannot hold, and ``sectors=False``, run
    the general path (every sector rebuilt per parameter set)."""
    from .batched import ground_bases, xas_batch_ops

    from . import opcompile

    thetas = np.asarray(thetas, np.float64)
    W = opcompile.n_words(ops.n_spin_orbitals(family.nBaths))
    cache = family.sectors(W) if sectors else None
    if cache is not None:
        family._bases.setdefault(W, cache.ground_basis)
        res = cache.xas(thetas, OMEGA, delta=DELTA, krylovSize=krylovSize, T=thetas[:, 7], reorth=reorth, **kw)
        return area_normalize(res.spectra), res
    if W not in family._bases:
        family._bases[W] = ground_bases(family.models(thetas[:1]), W)[0]
    ops_h, ops_t = family.batch_ops(thetas, W)
    res = xas_batch_ops(ops_h, ops_t, [family._bases[W]] * len(thetas), OMEGA, delta=DELTA, krylovSize=krylovSize,
                        T=thetas[:, 7], reorth=reorth, **kw)
    return area_normalize(res.spectra), res
