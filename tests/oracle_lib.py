"""ctypes wrapper over the FP64 CPU oracle (oracle/libcpuref.so).

Test infrastructure: used by tests/ only (and by bench.py's cpu_baseline leg,
which loads the same library directly). Layouts are the reference's AoS ones:
U [V*D, N, N] complex, P [V*D, N*N-1], phi [V, d] complex.
"""
import ctypes as C
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(ROOT, "oracle", "libcpuref.so")

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            subprocess.check_call(["make", "-s", "-C", os.path.join(ROOT, "oracle")])
        L = C.CDLL(LIB_PATH)
        vp, d, i, u64, dp = C.c_void_p, C.c_double, C.c_int, C.c_uint64, C.POINTER(C.c_double)
        sig = {
            "orc_create": (vp, [i, i, C.POINTER(C.c_int), C.c_char_p, u64]),
            "orc_destroy": (None, [vp]),
            "orc_last_error": (C.c_char_p, []),
            "orc_rep_dim": (i, [vp]), "orc_rep_real": (i, [vp]), "orc_rep_dimT": (C.c_long, [vp]),
            "orc_rep_nality": (i, [vp]), "orc_vol": (C.c_int64, [vp]),
            "orc_set_couplings": (None, [vp, d, d, d, d, i, i, i]),
            "orc_hot": (None, [vp, d]), "orc_cold": (None, [vp]),
            "orc_phi_gaussian": (None, [vp, u64, d]), "orc_phi_cold": (None, [vp, d]),
            "orc_set_frozen": (None, [vp, C.c_int]), "orc_normalize_phi": (None, [vp]),
            "orc_set_multi": (C.c_int, [vp, dp, C.c_int, d, dp]), "orc_multi_eval": (C.c_int, [vp, dp, dp, dp]),
            "orc_refresh": (None, [vp]),
            "orc_get": (None, [vp, i, dp]), "orc_set": (None, [vp, i, dp]),
            "orc_forces": (None, [vp]), "orc_observables": (None, [vp, dp]), "orc_ham_parts": (None, [vp, dp]),
            "orc_kick": (None, [vp, d]), "orc_drift": (None, [vp, d]), "orc_md_evolve": (None, [vp]),
            "orc_trajectory": (i, [vp, dp]), "orc_reunitarize": (None, [vp]), "orc_max_unitarity": (d, [vp]),
            "orc_get_traj_count": (u64, [vp]), "orc_get_accept_count": (u64, [vp]),
            "orc_set_counts": (None, [vp, u64, u64]),
            "orc_rep_T": (None, [vp, dp]), "orc_rep_W": (None, [vp, dp]),
            "orc_rep_apply": (None, [vp, dp, dp, i, dp]), "orc_rep_hop_g": (None, [vp, dp, dp, dp, dp]),
            "orc_rep_matrix": (None, [vp, dp, dp]),
            "orc_wilson_loop": (d, [vp, C.c_char_p, i, i]), "orc_polyakov_rep": (d, [vp, C.c_char_p, i]),
            "orc_wilson_rect": (None, [vp, C.c_long, i, i, i, i, dp]),
            "orc_splitmix64": (u64, [u64]), "orc_key": (u64, [u64, u64, u64, u64, u64]),
            "orc_uniform": (d, [u64, u64]), "orc_gauss": (d, [u64, u64]),
            "orc_gauss_many": (None, [u64, C.POINTER(C.c_uint64), C.c_long, dp]),
            "orc_gauss_quad_many": (None, [u64, C.POINTER(C.c_uint64), C.c_long, dp]),
            "orc_gauss_range": (None, [u64, u64, C.c_long, dp, dp]),
            "orc_expi": (None, [i, dp, dp]), "orc_generators": (None, [i, dp]),
            "orc_time_md": (d, [vp, i]), "orc_time_kick_drift": (d, [vp, i, d]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


FIELD = {"U": 0, "P": 1, "phi": 2, "pi": 3, "F": 4, "Fphi": 5}


class Oracle:
    """One lattice of the FP64 oracle: GaugeHMC (rep='') or GaugeHiggsHMC."""

    def __init__(self, D, N, L, rep="", seed=1):
        self.D, self.N = D, N
        ext = (C.c_int * D)(*([L] * D if np.isscalar(L) else L))
        self.h = lib().orc_create(D, N, ext, rep.encode(), seed)
        if not self.h:
            raise ValueError(lib().orc_last_error().decode())
        self.V = lib().orc_vol(self.h)
        self.d = lib().orc_rep_dim(self.h)
        self.real = bool(lib().orc_rep_real(self.h))
        self.dimT = lib().orc_rep_dimT(self.h)
        self.nality = lib().orc_rep_nality(self.h)
        self.ngen = N * N - 1

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_destroy(self.h)
            self.h = None

    def couplings(self, beta=2.3, kappa=0.1, lam=0.5, tau=1.0, nmd=20, omelyan=True, reunit=True):
        lib().orc_set_couplings(self.h, beta, kappa, lam, tau, nmd, int(omelyan), int(reunit))

    def shape(self, name):
        V, D, N = self.V, self.D, self.N
        return {"U": (V * D, N, N), "P": (V * D, self.ngen), "F": (V * D, self.ngen)}.get(name, (V, self.d))

    def get(self, name):
        shp = self.shape(name)
        cplx = name in ("U", "phi", "pi", "Fphi")
        n = int(np.prod(shp)) * (2 if cplx else 1)
        buf = np.zeros(n)
        lib().orc_get(self.h, FIELD[name], _dp(buf))
        return buf.view(np.complex128).reshape(shp) if cplx else buf.reshape(shp)

    def set(self, name, arr):
        arr = np.ascontiguousarray(arr)
        buf = np.ascontiguousarray(arr.view(np.float64) if np.iscomplexobj(arr) else arr, dtype=np.float64).ravel()
        lib().orc_set(self.h, FIELD[name], _dp(buf))

    def hot(self, sigma=0.8):
        lib().orc_hot(self.h, sigma)

    def cold(self):
        lib().orc_cold(self.h)

    def phi_gaussian(self, stream=12345, sigma=0.3):
        lib().orc_phi_gaussian(self.h, stream, sigma)

    def phi_cold(self, c=1.0):
        lib().orc_phi_cold(self.h, c)

    def frozen(self, on=True):
        """|phi_x| = 1 scalar: tangent projections and geodesic drift (gauge_higgs_hmc.hpp:61-103)."""
        lib().orc_set_frozen(self.h, int(on))

    def normalize_phi(self):
        lib().orc_normalize_phi(self.h)

    def channels(self):
        """Distinct adjoint-Casimir eigenvalues on R x Rbar (scalar_invariants.hpp:20-52)."""
        lam = np.zeros(64)
        n = lib().orc_set_multi(self.h, None, -1, 0.0, lam.ctypes.data_as(C.POINTER(C.c_double)))
        return lam[:n]

    def multi_potential(self, f, mu2):
        f = np.ascontiguousarray(f, dtype=np.float64)
        n = lib().orc_set_multi(self.h, f.ctypes.data_as(C.POINTER(C.c_double)), len(f), mu2, None)
        assert n == len(f), "wrong number of channel couplings"

    def multi_eval(self, phi):
        x = np.ascontiguousarray(np.asarray(phi, dtype=np.complex128)).view(np.float64)
        V = C.c_double()
        dV = np.zeros(x.size)
        dp = C.POINTER(C.c_double)
        assert lib().orc_multi_eval(self.h, x.ctypes.data_as(dp), C.byref(V), dV.ctypes.data_as(dp)) == 0
        return V.value, dV.view(np.complex128)

    def refresh(self):
        lib().orc_refresh(self.h)

    def forces(self):
        lib().orc_forces(self.h)
        return self.get("F"), (self.get("Fphi") if self.d else None)

    def observables(self):
        out = np.zeros(6)
        lib().orc_observables(self.h, _dp(out))
        return dict(zip(["plaq", "Sg", "SH", "Llink", "Lphi", "poly"], out))

    def ham_parts(self):
        out = np.zeros(5)
        lib().orc_ham_parts(self.h, _dp(out))
        return out

    def kick(self, eps):
        lib().orc_kick(self.h, eps)

    def drift(self, eps):
        lib().orc_drift(self.h, eps)

    def md_evolve(self):
        lib().orc_md_evolve(self.h)

    def trajectory(self):
        out = np.zeros(3)
        acc = lib().orc_trajectory(self.h, _dp(out))
        return bool(acc), out[0], out[1], out[2]

    def reunitarize(self):
        lib().orc_reunitarize(self.h)

    def max_unitarity(self):
        return lib().orc_max_unitarity(self.h)

    @property
    def traj_count(self):
        return lib().orc_get_traj_count(self.h)

    @property
    def accept_count(self):
        return lib().orc_get_accept_count(self.h)

    def set_counts(self, traj, acc):
        lib().orc_set_counts(self.h, traj, acc)

    def rep_T(self):
        out = np.zeros(self.ngen * self.d * self.d * 2)
        lib().orc_rep_T(self.h, _dp(out))
        return out.view(np.complex128).reshape(self.ngen, self.d, self.d)

    def rep_W(self):
        out = np.zeros(self.d * self.dimT)
        lib().orc_rep_W(self.h, _dp(out))
        return out.reshape(self.d, self.dimT)

    def rep_apply(self, U, phi, dag=False):
        U = np.ascontiguousarray(U, dtype=np.complex128)
        phi = np.ascontiguousarray(phi, dtype=np.complex128)
        out = np.zeros(2 * self.d)
        lib().orc_rep_apply(self.h, _dp(U.view(np.float64)), _dp(phi.view(np.float64)), int(dag), _dp(out))
        return out.view(np.complex128)

    def rep_hop_g(self, U, px, py):
        U = np.ascontiguousarray(U, dtype=np.complex128)
        px = np.ascontiguousarray(px, dtype=np.complex128)
        py = np.ascontiguousarray(py, dtype=np.complex128)
        out = np.zeros(self.ngen)
        lib().orc_rep_hop_g(self.h, _dp(U.view(np.float64)), _dp(px.view(np.float64)), _dp(py.view(np.float64)), _dp(out))
        return out

    def rep_matrix(self, U):
        U = np.ascontiguousarray(U, dtype=np.complex128)
        out = np.zeros(2 * self.d * self.d)
        lib().orc_rep_matrix(self.h, _dp(U.view(np.float64)), _dp(out))
        return out.view(np.complex128).reshape(self.d, self.d)

    # measurement layer (observables.hpp:36-101)
    def wilson_loop(self, probe, Ra, Rb):
        v = lib().orc_wilson_loop(self.h, probe.encode(), Ra, Rb)
        if np.isnan(v):
            raise RuntimeError(lib().orc_last_error().decode())
        return v

    def polyakov_rep(self, probe, tdir):
        v = lib().orc_polyakov_rep(self.h, probe.encode(), tdir)
        if np.isnan(v):
            raise RuntimeError(lib().orc_last_error().decode())
        return v

    def wilson_rect(self, x, mu, nu, Ra, Rb):
        out = np.zeros(2 * self.N * self.N)
        lib().orc_wilson_rect(self.h, x, mu, nu, Ra, Rb, _dp(out))
        return out.view(np.complex128).reshape(self.N, self.N)


def splitmix64(x):
    return lib().orc_splitmix64(x)


def key(a, b=0, c=0, d=0, e=0):
    return lib().orc_key(a, b, c, d, e)


def uniform(seed, k):
    return lib().orc_uniform(seed, k)


def gauss(seed, k):
    return lib().orc_gauss(seed, k)


def gauss_many(seed, keys):
    keys = np.ascontiguousarray(keys, dtype=np.uint64)
    out = np.zeros(len(keys))
    lib().orc_gauss_many(seed, keys.ctypes.data_as(C.POINTER(C.c_uint64)), len(keys), _dp(out))
    return out


def gauss_quad_many(seed, keys):
    """The correctly rounded Box-Muller draws (log and cos in quad precision)."""
    keys = np.ascontiguousarray(keys, dtype=np.uint64)
    out = np.zeros(len(keys))
    lib().orc_gauss_quad_many(seed, keys.ctypes.data_as(C.POINTER(C.c_uint64)), len(keys), _dp(out))
    return out


def gauss_range(seed, key0, n, glibc=True, quad=False):
    """Draws of keys key0..key0+n-1: (glibc draws or None, correctly rounded draws or None)."""
    g = np.zeros(n) if glibc else None
    q = np.zeros(n) if quad else None
    lib().orc_gauss_range(seed, key0, n, _dp(g) if glibc else None, _dp(q) if quad else None)
    return g, q


def expi(H):
    N = H.shape[0]
    H = np.ascontiguousarray(H, dtype=np.complex128)
    out = np.zeros(2 * N * N)
    lib().orc_expi(N, _dp(H.view(np.float64)), _dp(out))
    return out.view(np.complex128).reshape(N, N)


def generators(N):
    out = np.zeros((N * N - 1) * N * N * 2)
    lib().orc_generators(N, _dp(out))
    return out.view(np.complex128).reshape(N * N - 1, N, N)
