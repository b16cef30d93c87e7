"""Python mirror of the reference HMC interface over the ghx C ABI.

`GaugeHMC` mirrors hmc.hpp:13-96 and `GaugeHiggsHMC` mirrors
gauge_higgs_hmc.hpp:15-195: the same public couplings (beta, kappa, lambda,
tau, nmd, integ, reunit_each_traj), counters (traj_count, accept_count,
last_dH) and methods (trajectory, md_evolve, kick, drift, hamiltonian,
refresh_momenta). Every call goes to the HIP kernels in libghx.so; there is no
CPU fallback. Fields are exchanged in the reference AoS layout as numpy arrays:
U (V*D, N, N) complex128, P (V*D, N*N-1), phi/pi (V, d) complex128.
"""
import ctypes as C
import os

import numpy as np

from . import _lib

LEAPFROG, OMELYAN = 0, 1
FP64, FP32MD = 0, 1
COLD, HOT = 0, 1


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double)) if a is not None else None


class GhxError(RuntimeError):
    """A negative ghx return code; `code` holds it (GHX_E_* of ghx.h)."""

    def __init__(self, msg, code=None):
        super().__init__(msg)
        self.code = code


class Lattices:
    """A ghx context: one or more replicas of the same (D, N, rep, precision)."""

    def __init__(self, D, N, extents, rep="", precision=FP64):
        lib = _lib.load()
        self._lib = lib
        ext = [list(e) if not np.isscalar(e) else [e] * D for e in extents]
        self.D, self.N, self.nrep = D, N, len(ext)
        self.extents = ext
        flat = (C.c_int * (D * self.nrep))(*[x for e in ext for x in e])
        h = C.c_void_p()
        rc = lib.ghx_create(C.byref(h), D, N, self.nrep, flat, precision)
        if rc != 0:
            raise GhxError(f"ghx_create failed ({rc})", rc)
        self.h = h
        self.precision = precision
        d = C.c_int()
        self._check(lib.ghx_set_rep(self.h, rep.encode(), C.byref(d)), GhxError)
        nal, real, dimT = C.c_int(), C.c_int(), C.c_long()
        self._check(lib.ghx_rep_info(self.h, C.byref(d), C.byref(nal), C.byref(real), C.byref(dimT)))
        self.d, self.nality, self.real, self.dimT = d.value, nal.value, bool(real.value), dimT.value
        self.ngen = N * N - 1
        self.vols = [int(np.prod(e)) for e in ext]
        self.nmd, self.integ, self.reunit = 20, OMELYAN, True

    def _check(self, rc, exc=GhxError):
        if rc != 0:
            msg = f"ghx error {rc}: {self._lib.ghx_last_error(self.h).decode()}"
            raise exc(msg, rc) if exc is GhxError else exc(msg)

    def close(self):
        if getattr(self, "h", None):
            self._lib.ghx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # --- configuration -------------------------------------------------------
    def set_couplings(self, replica, beta, kappa=0.0, lam=0.0, tau=1.0):
        self._check(self._lib.ghx_set_couplings(self.h, replica, beta, kappa, lam, tau))

    def set_integrator(self, nmd, integ=OMELYAN, reunit=True):
        self.nmd, self.integ, self.reunit = nmd, integ, reunit
        self._check(self._lib.ghx_set_integrator(self.h, nmd, integ, int(reunit)))

    def set_nested(self, n_outer, m_inner=1, inner=LEAPFROG):
        """Multi-timescale integrator: n_outer hopping-force steps (integrator of set_integrator), m_inner
        gauge/on-site steps each, leapfrog or Omelyan (inner); n_outer = 0 turns it off."""
        self._check(self._lib.ghx_set_nested(self.h, n_outer, m_inner))
        self._check(self._lib.ghx_set_nested_inner(self.h, inner))

    def set_multi_potential(self, f, mu2):
        """Multi-invariant quartic potential V = -mu2 |phi|^2 + sum_c f_c V_c (scalar_invariants.hpp);
        f=None restores the quartic potential."""
        if f is None:
            self._check(self._lib.ghx_set_multi_potential(self.h, None, 0, 0.0))
            return
        f = np.ascontiguousarray(f, dtype=np.float64)
        self._check(self._lib.ghx_set_multi_potential(self.h, _dp(f), len(f), mu2))

    def set_frozen_phi(self, on=True):
        """|phi_x| = 1 scalar: tangent momenta and geodesic drift (gauge_higgs_hmc.hpp:61-103)."""
        self._check(self._lib.ghx_set_frozen_phi(self.h, int(on)))

    def normalize_phi(self):
        self._check(self._lib.ghx_normalize_phi(self.h))

    def save_checkpoint(self, path):
        """Links, phi, couplings, seeds and counters of every replica (reference AoS, FP64)."""
        self._check(self._lib.ghx_save_checkpoint(self.h, os.fsencode(path)))

    def load_checkpoint(self, path):
        self._check(self._lib.ghx_load_checkpoint(self.h, os.fsencode(path)))

    def set_seed(self, replica, seed):
        self._check(self._lib.ghx_set_seed(self.h, replica, seed))

    def init_fields(self, replica, link_mode=HOT, sigma_U=0.8, phi_mode=1, phi_stream=12345, phi_sigma=0.3):
        self._check(self._lib.ghx_init_fields(self.h, replica, link_mode, sigma_U, phi_mode, phi_stream, phi_sigma))

    def set_counters(self, replica, traj, acc):
        self._check(self._lib.ghx_set_counters(self.h, replica, traj, acc))

    def counters(self, replica=0):
        t, a, dh = C.c_uint64(), C.c_uint64(), C.c_double()
        self._check(self._lib.ghx_get_counters(self.h, replica, C.byref(t), C.byref(a), C.byref(dh)))
        return t.value, a.value, dh.value

    # --- MD ----------------------------------------------------------------------
    def refresh_momenta(self):
        self._check(self._lib.ghx_refresh_momenta(self.h))

    def kick(self, eps):
        self._check(self._lib.ghx_kick(self.h, eps))

    def drift(self, eps):
        self._check(self._lib.ghx_drift(self.h, eps))

    def md_evolve(self):
        self._check(self._lib.ghx_md_evolve(self.h))

    def hamiltonian_parts(self):
        out = np.zeros(5 * self.nrep)
        self._check(self._lib.ghx_hamiltonian(self.h, _dp(out)))
        return out.reshape(self.nrep, 5)

    def trajectory_all(self):
        acc = (C.c_int * self.nrep)()
        dH = np.zeros(self.nrep)
        self._check(self._lib.ghx_trajectory(self.h, acc, _dp(dH)))
        return np.array(list(acc), dtype=bool), dH

    def trajectories(self, n, use_graph=True):
        self._check(self._lib.ghx_trajectories(self.h, n, int(use_graph)))

    def observables_all(self):
        out = np.zeros(6 * self.nrep)
        self._check(self._lib.ghx_observables(self.h, _dp(out)))
        return out.reshape(self.nrep, 6)

    def wilson_loops(self, probe, Ra, Rb):
        """<(1/d_R') Re Tr D^(R')(U_loop)> of the Ra x Rb loops, per replica
        (observables.hpp:50-81). probe: "fund", "adj" or Young rows "r1[,r2]"."""
        out = np.zeros(self.nrep)
        self._check(self._lib.ghx_wilson_loop(self.h, probe.encode(), Ra, Rb, _dp(out)))
        return out

    def polyakov_loops_rep(self, probe, tdir=None):
        """Polyakov loop in direction tdir (default D-1) traced in the probe
        irrep, per replica (observables.hpp:83-101)."""
        out = np.zeros(self.nrep)
        t = self.D - 1 if tdir is None else tdir
        self._check(self._lib.ghx_polyakov_loop_rep(self.h, probe.encode(), t, _dp(out)))
        return out

    def max_unitarity(self, replica=0):
        v = C.c_double()
        self._check(self._lib.ghx_max_unitarity(self.h, replica, C.byref(v)))
        return v.value

    def reunitarize(self):
        self._check(self._lib.ghx_reunitarize(self.h))

    def sync(self):
        self._check(self._lib.ghx_sync(self.h))

    def time_md_steps(self, steps):
        ms = C.c_double()
        self._check(self._lib.ghx_time_md_steps(self.h, steps, C.byref(ms)))
        return ms.value

    def set_stream_segment(self, seg):
        self._check(self._lib.ghx_set_tuning(self.h, 0, seg))

    def stream_segment(self):
        v = C.c_int()
        self._check(self._lib.ghx_get_tuning(self.h, 0, C.byref(v)))
        return v.value

    def set_block_order(self, order):
        """Block walk of the one-site-per-thread MD kernel: 0 storage order, 1 and 2
        XCD slices, 3-6 4D tiles of 128 sites (see ghx.h)."""
        self._check(self._lib.ghx_set_tuning(self.h, 1, order))

    def block_order(self):
        v = C.c_int()
        self._check(self._lib.ghx_get_tuning(self.h, 1, C.byref(v)))
        return v.value

    def set_link_lanes(self, on):
        """One lane per link (four per site) in the MD step, for small lattices."""
        self._check(self._lib.ghx_set_tuning(self.h, 2, int(on)))

    def link_lanes(self):
        v = C.c_int()
        self._check(self._lib.ghx_get_tuning(self.h, 2, C.byref(v)))
        return v.value

    def set_sym_engine(self, on):
        """Large tensor irreps: apply D(U) through symmetric powers of U (1) or
        slot by slot on the tensor (0)."""
        self._check(self._lib.ghx_set_tuning(self.h, 3, int(on)))

    def sym_engine(self):
        v = C.c_int()
        self._check(self._lib.ghx_get_tuning(self.h, 3, C.byref(v)))
        return v.value

    def set_sym_pairs(self, mode):
        """Symmetric-power engine, SU(3) with first Young row <= 6: two links per
        wave with the columns of Sym^k(U) in registers (2, default; FP32 with a
        second row <= 1, else as 1), two links per wave with Sym^k(U) built a
        column per lane into LDS (1), or one link per wave with the per-entry
        LDS program (0). sym_pairs() reports the kernel in use."""
        self._check(self._lib.ghx_set_tuning(self.h, 4, int(mode)))

    def sym_pairs(self):
        v = C.c_int()
        self._check(self._lib.ghx_get_tuning(self.h, 4, C.byref(v)))
        return v.value

    def set_x_lanes(self, on):
        """SU(3) FP32 MD where waves hold whole x-lines: each site hands its lower
        staple of plane (0, mu) to the x+0 neighbour lane (1, default) or every
        site forms its own backward staples (0)."""
        self._check(self._lib.ghx_set_tuning(self.h, 5, int(on)))

    def x_lanes(self):
        v = C.c_int()
        self._check(self._lib.ghx_get_tuning(self.h, 5, C.byref(v)))
        return v.value

    def set_fuse_inner(self, on):
        """Large irreps under the multi-timescale integrator: the inner-scale
        on-site and gauge updates in one launch (1, default) or two (0)."""
        self._check(self._lib.ghx_set_tuning(self.h, 6, int(on)))

    def fuse_inner(self):
        v = C.c_int()
        self._check(self._lib.ghx_get_tuning(self.h, 6, C.byref(v)))
        return v.value

    def set_link_split(self, on):
        """Link-parallel step without scalar components (pure gauge, or the gauge
        part of a large irrep): each link's staples split over D - 1 lanes, one
        per plane (1, default), or one lane per link (0)."""
        self._check(self._lib.ghx_set_tuning(self.h, 9, int(on)))

    def link_split(self):
        v = C.c_int()
        self._check(self._lib.ghx_get_tuning(self.h, 9, C.byref(v)))
        return v.value

    def set_store_policy(self, cached):
        """MD step stores streamed past the caches (0, default) or with the
        default cache policy (1); the autotuner tries both."""
        self._check(self._lib.ghx_set_tuning(self.h, 8, int(cached)))

    def store_policy(self):
        v = C.c_int()
        self._check(self._lib.ghx_get_tuning(self.h, 8, C.byref(v)))
        return v.value

    def set_split_inner(self, on):
        """Large irreps under the multi-timescale integrator: every on-site
        sub-step of an inner block in one launch and the links' gauge sub-steps
        alone (1, default), where the two are independent (no surrogate force,
        no plugged-in potential); 0 interleaves them."""
        self._check(self._lib.ghx_set_tuning(self.h, 7, int(on)))

    def split_inner(self):
        v = C.c_int()
        self._check(self._lib.ghx_get_tuning(self.h, 7, C.byref(v)))
        return v.value

    def autotune(self, steps=10):
        """Pick the fastest MD-step launch shape on the live fields (fields restored)."""
        best, ms = C.c_int(), C.c_double()
        self._check(self._lib.ghx_autotune(self.h, steps, C.byref(best), C.byref(ms)))
        return best.value, ms.value

    # --- learned surrogate of the matter link force ------------------------------
    def surrogate_init(self, seed=7):
        self._check(self._lib.ghx_surrogate_init(self.h, seed))
        self.nparam = self._lib.ghx_surrogate_nparams()

    def surrogate_enable(self, on=True):
        self._check(self._lib.ghx_surrogate_enable(self.h, int(on)))

    def surrogate_params(self, new=None):
        buf = np.zeros(self.nparam, dtype=np.float32)
        fp = C.POINTER(C.c_float)
        if new is None:
            self._check(self._lib.ghx_surrogate_params(self.h, 1, buf.ctypes.data_as(fp)))
            return buf
        buf[:] = new
        self._check(self._lib.ghx_surrogate_params(self.h, 0, buf.ctypes.data_as(fp)))
        return buf

    def surrogate_gradient(self):
        g = np.zeros(self.nparam, dtype=np.float32)
        loss = C.c_double()
        self._check(self._lib.ghx_surrogate_gradient(self.h, g.ctypes.data_as(C.POINTER(C.c_float)), C.byref(loss)))
        return g, loss.value

    def surrogate_train(self, iters=1, lr=1e-3):
        l0, l1 = C.c_double(), C.c_double()
        self._check(self._lib.ghx_surrogate_train(self.h, iters, lr, C.byref(l0), C.byref(l1)))
        return l0.value, l1.value

    def surrogate_snapshot(self, replica=0):
        n = self.vols[replica] * self.D
        X = np.zeros((n, 16), dtype=np.float32)
        B = np.zeros((n, 4, self.ngen), dtype=np.float32)
        S = np.zeros((n, self.ngen))
        F = np.zeros((n, self.ngen))
        fp = C.POINTER(C.c_float)
        self._check(self._lib.ghx_surrogate_snapshot(self.h, replica, X.ctypes.data_as(fp), B.ctypes.data_as(fp), _dp(S), _dp(F)))
        return X, B, S, F

    def replica_swap(self, i, j, u):
        acc, delta = C.c_int(), C.c_double()
        self._check(self._lib.ghx_replica_swap(self.h, i, j, u, C.byref(acc), C.byref(delta)))
        return bool(acc.value), delta.value

    # --- field exchange (reference AoS layout) ------------------------------------
    def get(self, name, replica=0):
        V, D, N, d = self.vols[replica], self.D, self.N, self.d
        if name == "U":
            buf = np.zeros(V * D * N * N * 2)
            self._check(self._lib.ghx_get_fields(self.h, replica, _dp(buf), None, None, None))
            return buf.view(np.complex128).reshape(V * D, N, N)
        if name == "P":
            buf = np.zeros(V * D * self.ngen)
            self._check(self._lib.ghx_get_fields(self.h, replica, None, _dp(buf), None, None))
            return buf.reshape(V * D, self.ngen)
        buf = np.zeros(V * d * 2)
        args = [None, None, _dp(buf), None] if name == "phi" else [None, None, None, _dp(buf)]
        self._check(self._lib.ghx_get_fields(self.h, replica, *args))
        return buf.view(np.complex128).reshape(V, d)

    def set(self, name, arr, replica=0):
        arr = np.ascontiguousarray(arr)
        buf = np.ascontiguousarray(arr.view(np.float64) if np.iscomplexobj(arr) else arr, dtype=np.float64).ravel()
        args = [None, None, None, None]
        args[{"U": 0, "P": 1, "phi": 2, "pi": 3}[name]] = _dp(buf)
        self._check(self._lib.ghx_set_fields(self.h, replica, *args))

    def forces(self, replica=0):
        V, D = self.vols[replica], self.D
        F = np.zeros(V * D * self.ngen)
        Fp = np.zeros(V * max(self.d, 1) * 2)
        self._check(self._lib.ghx_forces(self.h, replica, _dp(F), _dp(Fp) if self.d else None))
        return F.reshape(V * D, self.ngen), (Fp.view(np.complex128).reshape(V, self.d) if self.d else None)


class GaugeHiggsHMC:
    """Single-lattice gauge+Higgs HMC (gauge_higgs_hmc.hpp:15-195). rep='' gives GaugeHMC."""

    def __init__(self, extents, rep, seed, D=4, N=2, precision=FP64):
        if np.isscalar(extents):
            extents = [extents] * D
        self.lat = Lattices(D, N, [list(extents)], rep=rep, precision=precision)
        self.lat.set_seed(0, seed)
        self.beta, self.kappa, self.lam, self.tau = 2.3, 0.1, 0.5, 1.0
        self.nmd, self.integ, self.reunit_each_traj = 20, OMELYAN, True
        self._push()

    def _push(self):
        self.lat.set_integrator(self.nmd, self.integ, self.reunit_each_traj)
        self.lat.set_couplings(0, self.beta, self.kappa, self.lam, self.tau)

    def configure(self, **kw):
        for k, v in kw.items():
            setattr(self, k, v)
        self._push()
        return self

    @property
    def d(self):
        return self.lat.d

    @property
    def traj_count(self):
        return self.lat.counters()[0]

    @property
    def accept_count(self):
        return self.lat.counters()[1]

    @property
    def last_dH(self):
        return self.lat.counters()[2]

    def acceptance(self):
        t, a, _ = self.lat.counters()
        return a / t if t else 0.0

    def hot(self, sigma=0.8, phi_stream=12345, phi_sigma=0.3, phi_mode=1):
        self.lat.init_fields(0, HOT, sigma, phi_mode, phi_stream, phi_sigma)

    @property
    def frozen_phi(self):
        return getattr(self, "_frozen", False)

    @frozen_phi.setter
    def frozen_phi(self, on):
        self.lat.set_frozen_phi(on)
        self._frozen = bool(on)

    def normalize_phi(self):
        self.lat.normalize_phi()

    def cold(self, phi_c=1.0):
        self.lat.init_fields(0, COLD, 0.0, 0, 0, phi_c)

    def refresh_momenta(self):
        self.lat.refresh_momenta()

    def kick(self, eps):
        self.lat.kick(eps)

    def drift(self, eps):
        self.lat.drift(eps)

    def md_evolve(self):
        self.lat.md_evolve()

    def hamiltonian(self):
        return self.lat.hamiltonian_parts()[0, 4]

    def hamiltonian_parts(self):
        return self.lat.hamiltonian_parts()[0]

    def trajectory(self):
        acc, dH = self.lat.trajectory_all()
        return bool(acc[0])

    def observables(self):
        o = self.lat.observables_all()[0]
        return dict(zip(["plaq", "Sg", "SH", "Llink", "Lphi", "poly"], o))

    # measurement layer (observables.hpp:36-101)
    def wilson_loop_fund(self, Ra, Rb):
        return float(self.lat.wilson_loops("fund", Ra, Rb)[0])

    def wilson_loop_rep(self, probe, Ra, Rb):
        return float(self.lat.wilson_loops(probe, Ra, Rb)[0])

    def polyakov_loop_rep(self, probe, tdir=None):
        return float(self.lat.polyakov_loops_rep(probe, tdir)[0])

    def get(self, name):
        return self.lat.get(name)

    def set(self, name, arr):
        self.lat.set(name, arr)

    def forces(self):
        return self.lat.forces()


SERIES_KEYS = ("mean", "var", "binder", "binned_error", "tau_int", "window", "n_eff", "naive_error",
               "autocorr_error", "blocking_error")


def series_analysis(x, block=1):
    """Stats and autocorrelation analysis of one Monte Carlo series
    (observables.hpp:104-145, autocorr.hpp:55-212), computed by the library."""
    lib = _lib.load()
    x = np.ascontiguousarray(x, dtype=np.float64).ravel()
    out = np.zeros(10)
    rc = lib.ghx_series_analysis(_dp(x), x.size, int(block), _dp(out))
    if rc != 0:
        raise GhxError(f"ghx_series_analysis failed ({rc})")
    d = dict(zip(SERIES_KEYS, out))
    d["window"] = int(d["window"])
    return d


def autocorr_function(x, maxlag=-1):
    """rho(t) = C(t)/C(0) for t = 0..maxlag (autocorr.hpp:93-123)."""
    lib = _lib.load()
    x = np.ascontiguousarray(x, dtype=np.float64).ravel()
    rho = np.zeros(max(x.size, 1))
    n = C.c_long()
    rc = lib.ghx_autocorr_function(_dp(x), x.size, int(maxlag), _dp(rho), C.byref(n))
    if rc != 0:
        raise GhxError(f"ghx_autocorr_function failed ({rc})")
    return rho[: n.value].copy()


def casimir_channels(N, spec):
    """Channel Casimirs of R x Rbar for the multi-invariant potential (host only)."""
    lib = _lib.load()
    n = C.c_int()
    rc = lib.ghx_casimir_channels(N, spec.encode(), None, C.byref(n))
    if rc != 0:
        raise GhxError(f"ghx_casimir_channels failed ({rc})")
    out = np.zeros(n.value)
    lib.ghx_casimir_channels(N, spec.encode(), _dp(out), C.byref(n))
    return out


def GaugeHMC(extents, seed, D=4, N=2, precision=FP64):
    """Pure-gauge HMC (hmc.hpp:13-96): keys 0xA1 / 0xB2, H = KE_P + S_g."""
    return GaugeHiggsHMC(extents, "", seed, D=D, N=N, precision=precision)


def rng_gauss(seed, key0, n):
    """The device's Gaussian stream Rng(seed).gauss(key0 + i), i < n (rng.hpp:42-46)."""
    lib = _lib.load()
    out = np.zeros(int(n))
    rc = lib.ghx_rng_gauss(seed, key0, int(n), _dp(out))
    if rc != 0:
        raise GhxError(f"ghx_rng_gauss failed ({rc})", rc)
    return out
