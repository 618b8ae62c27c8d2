"""ctypes binding of libsmc (include/smc.h).

The library is built in-tree (supermc_amd/build.py) and loaded from
supermc_amd/lib/libsmc.so.  There is no fallback: if the library is missing,
loading raises.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libsmc.so")

SMC_OK = 0
SMC_ERR_ARG = 1
SMC_ERR_HIP = 2
SMC_ERR_CAPACITY = 3
SMC_ERR_NO_EVENT = 4
SMC_ERR_KLN_TABLE = 5
SMC_ERR_DNDY_ZERO = 6
SMC_ERR_UNSUPPORTED = 7
SMC_ERR_STATE = 8
SMC_ERR_MISSING_KEY = 9

POS_QUANTUM_FM = 1.0 / 16777216.0
MAX_A = 256

EV_STRIDE = 64
EV_EPS_RE, EV_EPS_IM, EV_EPSP_RE, EV_EPSP_IM, EV_RN = 0, 10, 20, 30, 40
EV_XC, EV_YC, EV_TOTAL, EV_DSDY, EV_NPART, EV_NCOLL, EV_B, EV_STATUS = 50, 51, 52, 53, 54, 55, 56, 57
EV_SYSTEM = 58

TILE_VARIANTS = 4  # SMC_TILE_VARIANTS

K_EVENT, K_SMEAR, K_MOMENTS, K_KLN = 0, 1, 2, 3


class SmcParams(C.Structure):
    _fields_ = [
        ("which_mc_model", C.c_int), ("sub_model", C.c_int), ("lambda_", C.c_double),
        ("tmax", C.c_int), ("tmax_subdivision", C.c_int), ("alpha", C.c_double),
        ("Aproj", C.c_int), ("Atarg", C.c_int), ("proj_deformed", C.c_int), ("targ_deformed", C.c_int),
        ("shape_of_nucleons", C.c_int), ("collision_criterion", C.c_int), ("shape_of_entropy", C.c_int),
        ("quark_width", C.c_double), ("gauss_nucl_width", C.c_double),
        ("ecm", C.c_double), ("bmin", C.c_double), ("bmax", C.c_double),
        ("Npmin", C.c_int), ("Npmax", C.c_int),
        ("cc_fluctuation_model", C.c_int), ("cc_fluctuation_Gamma_theta", C.c_double),
        ("maxx", C.c_double), ("maxy", C.c_double), ("dx", C.c_double), ("dy", C.c_double),
        ("finalFactor", C.c_double), ("ecc_from_order", C.c_int), ("ecc_to_order", C.c_int),
        ("sigma_nn_mb", C.c_double),
        ("proj_ws_R", C.c_double), ("proj_ws_a", C.c_double), ("targ_ws_R", C.c_double), ("targ_ws_a", C.c_double),
        ("proj_beta2", C.c_double), ("proj_beta4", C.c_double), ("targ_beta2", C.c_double), ("targ_beta4", C.c_double),
        ("nucleon_min_dist", C.c_double),
        ("kln_n_pt", C.c_int), ("kln_n_kt", C.c_int), ("kln_n_phi", C.c_int),
        ("kln_pt_min", C.c_double), ("kln_pt_max", C.c_double), ("kln_direct", C.c_int),
        ("max_events", C.c_int), ("max_pairs", C.c_int), ("max_attempts", C.c_int), ("smear_path", C.c_int),
        ("gaussian_lambda", C.c_double),
    ]


class SmcDerived(C.Structure):
    _fields_ = [
        ("siginNN", C.c_double), ("width", C.c_double), ("dsq", C.c_double),
        ("nx", C.c_int), ("ny", C.c_int), ("xmin", C.c_double), ("ymin", C.c_double),
        ("hit_n_max", C.c_int64), ("kln_dT", C.c_double), ("kln_tmax", C.c_int),
        ("ws_R", C.c_double * 2), ("ws_a", C.c_double * 2), ("beta2", C.c_double * 2),
        ("beta4", C.c_double * 2), ("dmin", C.c_double),
        ("criterion", C.c_int), ("sigma_gg", C.c_double), ("box_lim", C.c_int64),
    ]


class P3dParams(C.Structure):
    """smc_p3d_params (include/smc_p3d.h): profile_3d's constructor arguments."""
    _fields_ = [("grid_nx", C.c_int), ("grid_ny", C.c_int), ("grid_neta", C.c_int),
                ("grid_dx", C.c_double), ("grid_dy", C.c_double), ("grid_deta", C.c_double),
                ("ecm", C.c_double), ("random_flag", C.c_int)]


class P3dDerived(C.Structure):
    _fields_ = [("y_beam", C.c_double), ("sigma_inelastic", C.c_double), ("eta0", C.c_double),
                ("sigma_x0", C.c_double), ("sigma_eta0", C.c_double), ("proj_max", C.c_double),
                ("targ_max", C.c_double)]


P3D_TABLE = 1000

_P = C.POINTER
_dp = _P(C.c_double)
_ip = _P(C.c_int)
_fp = _P(C.c_float)

# name -> (restype, argtypes); every symbol include/smc.h declares
SIGNATURES = {
    "smc_params_default": (C.c_int, [_P(SmcParams)]),
    "smc_params_sizeof": (C.c_size_t, []),
    "smc_derived_sizeof": (C.c_size_t, []),
    "smc_create": (C.c_int, [_P(C.c_void_p), _P(SmcParams)]),
    "smc_destroy": (None, [C.c_void_p]),
    "smc_last_error": (C.c_char_p, [C.c_void_p]),
    "smc_get_derived": (C.c_int, [C.c_void_p, _P(SmcDerived)]),
    "smc_sample": (C.c_int, [C.c_void_p, C.c_uint64, C.c_uint64, C.c_int]),
    "smc_sample_b": (C.c_int, [C.c_void_p, C.c_uint64, C.c_uint64, C.c_int, _dp]),
    "smc_set_positions": (C.c_int, [C.c_void_p, C.c_int, _dp, _dp, _dp, C.c_uint64]),
    "smc_collide": (C.c_int, [C.c_void_p, _ip, _ip, _ip]),
    "smc_get_nucleons": (C.c_int, [C.c_void_p, C.c_int, _dp, _ip, _fp, _dp, _ip, _fp]),
    "smc_get_pairs": (C.c_int, [C.c_void_p, C.c_int, C.c_int, _ip, _ip, _fp, _ip]),
    "smc_get_sources": (C.c_int, [C.c_void_p, C.c_int, C.c_int, _dp, _ip, _fp, _ip]),
    "smc_recenter": (C.c_int, [C.c_void_p, C.c_int, C.c_int, _dp]),
    "smc_run_stream": (C.c_int, [C.c_void_p, C.c_uint64, C.c_uint64, C.c_int64, _dp]),
    "smc_run_noalloc": (C.c_int, [C.c_void_p, C.c_uint64, C.c_uint64, C.c_int, C.c_void_p]),
    "smc_thickness": (C.c_int, [C.c_void_p, _fp, _fp]),
    "smc_density": (C.c_int, [C.c_void_p, _fp, _dp]),
    "smc_density_f64": (C.c_int, [C.c_void_p, C.c_int, C.c_int, _dp]),
    "smc_set_attempts_per_pass": (C.c_int, [C.c_void_p, C.c_int]),
    "smc_get_attempts_per_pass": (C.c_int, [C.c_void_p, C.POINTER(C.c_int)]),
    "smc_autotune_attempts": (C.c_int, [C.c_void_p, C.c_int, _ip, _dp]),
    "smc_calibrate_fma": (C.c_int, [_dp, _dp]),
    "smc_moments": (C.c_int, [C.c_void_p, _dp]),
    "smc_run": (C.c_int, [C.c_void_p, C.c_uint64, C.c_uint64, C.c_int, _dp]),
    "smc_run_graph": (C.c_int, [C.c_void_p, C.c_uint64, C.c_uint64, C.c_int, _dp]),
    "smc_fetch_events": (C.c_int, [C.c_void_p, C.c_int, _dp]),
    "smc_average_profiles": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_void_p, _dp, _dp]),
    "smc_average_profiles_extra": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_void_p, _dp, _dp, C.c_void_p]),
    "smc_rho_binary": (C.c_int, [C.c_void_p, _fp]),
    "smc_spectator_density": (C.c_int, [C.c_void_p, _fp, _fp]),
    "smc_kln_dndy": (C.c_int, [C.c_void_p, C.c_int, _dp, _dp, _dp]),
    "smc_kln_make_table": (C.c_int, [C.c_void_p, _dp]),
    "smc_diff_setup": (C.c_int, [C.c_void_p, C.c_uint64, C.c_uint64, C.c_int, _dp]),
    "smc_diff_forward": (C.c_int, [C.c_void_p, _dp, _dp]),
    "smc_diff_backward": (C.c_int, [C.c_void_p, _dp, _dp]),
    "smc_diff_topology": (C.c_int, [C.c_void_p, C.c_int, _dp, _dp, _dp, _dp, C.c_int, _ip]),
    "smc_set_systems": (C.c_int, [C.c_void_p, C.c_int, _P(SmcParams), C.c_int, _ip]),
    "smc_get_event_system": (C.c_int, [C.c_void_p, C.c_int, _ip]),
    "smc_set_tile_variant": (C.c_int, [C.c_void_p, C.c_int]),
    "smc_autotune": (C.c_int, [C.c_void_p, C.c_int, _ip, _dp]),
    "smc_set_timing": (C.c_int, [C.c_void_p, C.c_int]),
    "smc_get_timing": (C.c_int, [C.c_void_p, C.c_int, _dp, _ip]),
    "smc_synchronize": (C.c_int, [C.c_void_p]),
    # include/smc_p3d.h
    "smc_p3d_create": (C.c_int, [_P(C.c_void_p), _P(P3dParams)]),
    "smc_p3d_destroy": (None, [C.c_void_p]),
    "smc_p3d_last_error": (C.c_char_p, [C.c_void_p]),
    "smc_p3d_get_derived": (C.c_int, [C.c_void_p, _P(P3dDerived), _dp, _dp, _dp]),
    "smc_p3d_sample": (C.c_int, [C.c_void_p, C.c_int, _dp, _ip, _ip, _dp, _ip, C.c_uint64, C.c_uint64, _dp, _dp]),
    "smc_p3d_smear": (C.c_int, [C.c_void_p, C.c_int, _dp, _ip, _fp]),
    "smc_p3d_generate": (C.c_int, [C.c_void_p, C.c_int, _dp, _ip, _ip, _dp, _ip, C.c_uint64, C.c_uint64, _fp, _fp,
                                   _dp, _dp]),
}

_lib = None


def load(path=LIB_PATH):
    """Loads libsmc.so; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError(f"libsmc.so not built ({path}); run `python -m supermc_amd.build`")
    lib = C.CDLL(path)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if lib.smc_params_sizeof() != C.sizeof(SmcParams):
        raise RuntimeError("smc_params layout mismatch between smc.h and the binding")
    if lib.smc_derived_sizeof() != C.sizeof(SmcDerived):
        raise RuntimeError("smc_derived layout mismatch between smc.h and the binding")
    _lib = lib
    return lib


class SmcError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"[smc status {code}] {msg}")
        self.code = code


def default_params(**kw):
    lib = load()
    p = SmcParams()
    lib.smc_params_default(C.byref(p))
    for k, v in kw.items():
        if k == "lambda":
            k = "lambda_"
        if not hasattr(p, k):
            raise KeyError(k)
        setattr(p, k, v)
    return p


def calibrate_fma():
    """smc_calibrate_fma: (TFLOP/s, ms) of a chain of independent FP32 FMAs
    on the current device -- the plain-FMA rate of this build."""
    tf, ms = C.c_double(), C.c_double()
    rc = load().smc_calibrate_fma(C.byref(tf), C.byref(ms))
    if rc:
        raise SmcError(rc, "smc_calibrate_fma failed")
    return tf.value, ms.value


def _ptr(a, ctype):
    return a.ctypes.data_as(_P(ctype)) if a is not None else None


class Context:
    """One libsmc context: a HIP stream plus device buffers for a batch."""

    def __init__(self, params):
        self.lib = load()
        self.params = params
        h = C.c_void_p()
        rc = self.lib.smc_create(C.byref(h), C.byref(params))
        if rc != SMC_OK:
            raise SmcError(rc, "smc_create failed (see stderr)")
        self.h = h
        d = SmcDerived()
        self.lib.smc_get_derived(self.h, C.byref(d))
        self.derived = d
        self.nx, self.ny = d.nx, d.ny
        self.nev = 0

    def close(self):
        if getattr(self, "h", None):
            self.lib.smc_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _check(self, rc, allow=()):
        if rc != SMC_OK and rc not in allow:
            raise SmcError(rc, self.lib.smc_last_error(self.h).decode())
        return rc

    # -- events
    def sample(self, seed, event0, nev):
        self._check(self.lib.smc_sample(self.h, seed, event0, nev))
        self.nev = nev

    def sample_b(self, seed, event0, b):
        b = np.ascontiguousarray(b, dtype=np.float64)
        self._check(self.lib.smc_sample_b(self.h, seed, event0, len(b), _ptr(b, C.c_double)))
        self.nev = len(b)

    def set_positions(self, xy_proj, xy_targ, b=None, fluct_seed=0):
        xy_proj = np.ascontiguousarray(xy_proj, dtype=np.float64)
        xy_targ = np.ascontiguousarray(xy_targ, dtype=np.float64)
        nev = xy_proj.shape[0]
        bb = None if b is None else np.ascontiguousarray(b, dtype=np.float64)
        self._check(self.lib.smc_set_positions(self.h, nev, _ptr(xy_proj, C.c_double), _ptr(xy_targ, C.c_double),
                                               _ptr(bb, C.c_double), fluct_seed))
        self.nev = nev

    def collide(self, allow=()):
        n = self.nev
        a, b, c = (np.zeros(n, np.int32) for _ in range(3))
        rc = self._check(self.lib.smc_collide(self.h, _ptr(a, C.c_int), _ptr(b, C.c_int), _ptr(c, C.c_int)), allow)
        return a, b, c, rc

    def nucleons(self, ev):
        A, B = self.params.Aproj, self.params.Atarg
        out = dict(xy_proj=np.zeros((A, 2)), wound_proj=np.zeros(A, np.int32), fluct_proj=np.zeros(A, np.float32),
                   xy_targ=np.zeros((B, 2)), wound_targ=np.zeros(B, np.int32), fluct_targ=np.zeros(B, np.float32))
        self._check(self.lib.smc_get_nucleons(
            self.h, ev, _ptr(out["xy_proj"], C.c_double), _ptr(out["wound_proj"], C.c_int),
            _ptr(out["fluct_proj"], C.c_float), _ptr(out["xy_targ"], C.c_double),
            _ptr(out["wound_targ"], C.c_int), _ptr(out["fluct_targ"], C.c_float)))
        return out

    def pairs(self, ev, cap=None):
        if cap is None:
            cap = self.params.Aproj * self.params.Atarg
        pa, pb = np.zeros(cap, np.int32), np.zeros(cap, np.int32)
        pf = np.zeros(cap, np.float32)
        n = C.c_int()
        self._check(self.lib.smc_get_pairs(self.h, ev, cap, _ptr(pa, C.c_int), _ptr(pb, C.c_int),
                                           _ptr(pf, C.c_float), C.byref(n)))
        k = n.value
        return pa[:k], pb[:k], pf[:k]

    def thickness(self):
        shape = (self.nev, self.nx, self.ny)
        ta, tb = np.zeros(shape, np.float32), np.zeros(shape, np.float32)
        self._check(self.lib.smc_thickness(self.h, _ptr(ta, C.c_float), _ptr(tb, C.c_float)))
        return ta, tb

    def density(self, want_grid=True, out=None):
        s = out if out is not None else (np.zeros((self.nev, self.nx, self.ny), np.float32) if want_grid else None)
        if s is not None:
            assert s.dtype == np.float32 and s.flags.c_contiguous and s.size >= self.nev * self.nx * self.ny
        dndy = np.zeros(self.nev)
        rc = self._check(self.lib.smc_density(self.h, _ptr(s, C.c_float), _ptr(dndy, C.c_double)),
                         allow=(SMC_ERR_DNDY_ZERO,))
        return s, dndy, rc

    def density_f64(self, ev0=0, n=None, out=None):
        """smc_density_f64: FP64 grids of events [ev0, ev0 + n) with the
        reference's double arithmetic, [n][nx][ny]."""
        if n is None:
            n = self.nev - ev0
        s = out if out is not None else np.zeros((n, self.nx, self.ny))
        assert s.dtype == np.float64 and s.flags.c_contiguous and s.size >= n * self.nx * self.ny
        self._check(self.lib.smc_density_f64(self.h, ev0, n, _ptr(s, C.c_double)))
        return s

    def moments(self):
        ev = np.zeros((self.nev, EV_STRIDE))
        self._check(self.lib.smc_moments(self.h, _ptr(ev, C.c_double)))
        return ev

    def run(self, seed, event0, nev, fetch=True):
        ev = np.zeros((nev, EV_STRIDE)) if fetch else None
        self._check(self.lib.smc_run(self.h, seed, event0, nev, _ptr(ev, C.c_double)))
        self.nev = nev
        return ev

    def run_stream(self, seed, event0, nev_total, out=None):
        """smc_run_stream: nev_total events in batches of max_events through
        the context's buffers; returns the records [nev_total][64]."""
        ev = out if out is not None else np.empty((nev_total, EV_STRIDE))
        self._check(self.lib.smc_run_stream(self.h, seed, event0, nev_total, _ptr(ev, C.c_double)))
        cap = self.params.max_events
        self.nev = nev_total - (nev_total - 1) // cap * cap  # the context holds the last batch
        return ev

    def set_systems(self, systems, pattern):
        """Mixed collision systems: systems is a list of parameter sets (only
        their nuclei may differ from this context's); sampled event g uses
        systems[pattern[g % len(pattern)]].  An empty list restores the
        context's own system."""
        arr = (SmcParams * max(1, len(systems)))(*systems) if systems else None
        pat = np.ascontiguousarray(np.asarray(pattern, np.int32)) if systems else None
        self._check(self.lib.smc_set_systems(self.h, len(systems), arr, len(pattern) if systems else 0,
                                             _ptr(pat, C.c_int)))

    def event_system(self, ev):
        s = C.c_int()
        self._check(self.lib.smc_get_event_system(self.h, ev, C.byref(s)))
        return s.value

    def set_attempts_per_pass(self, n):
        self._check(self.lib.smc_set_attempts_per_pass(self.h, n))

    def attempts_per_pass(self):
        n = C.c_int()
        self._check(self.lib.smc_get_attempts_per_pass(self.h, C.byref(n)))
        return n.value

    def autotune_attempts(self, reps=3):
        best = C.c_int()
        ms = np.zeros(2)
        self._check(self.lib.smc_autotune_attempts(self.h, reps, C.byref(best), _ptr(ms, C.c_double)))
        return best.value, ms

    def set_tile_variant(self, variant):
        self._check(self.lib.smc_set_tile_variant(self.h, variant))

    def autotune(self, reps=3):
        """Times every tile shape on the current events and keeps the
        fastest: (best, ms per shape)."""
        best = C.c_int()
        ms = np.zeros(TILE_VARIANTS)
        self._check(self.lib.smc_autotune(self.h, reps, C.byref(best), _ptr(ms, C.c_double)))
        return best.value, ms

    def run_graph(self, seed, event0, nev, fetch=True):
        """run() through a replayed hipGraph (smc_run_graph)."""
        ev = np.zeros((nev, EV_STRIDE)) if fetch else None
        self._check(self.lib.smc_run_graph(self.h, seed, event0, nev, _ptr(ev, C.c_double)))
        self.nev = nev
        return ev

    def fetch_events(self, nev):
        ev = np.zeros((nev, EV_STRIDE))
        self._check(self.lib.smc_fetch_events(self.h, nev, _ptr(ev, C.c_double)))
        return ev

    def rho_binary(self, host=True):
        """The binary-collision grids [nev][nx][ny]; host=False leaves them
        in device memory and returns None."""
        rho = np.zeros((self.nev, self.nx, self.ny), np.float32) if host else None
        self._check(self.lib.smc_rho_binary(self.h, _ptr(rho, C.c_float)))
        return rho

    def spectator_density(self, host=True):
        """Projectile and target spectator grids; host=False leaves them in
        device memory and returns (None, None)."""
        shape = (self.nev, self.nx, self.ny)
        a = np.zeros(shape, np.float32) if host else None
        b = np.zeros(shape, np.float32) if host else None
        self._check(self.lib.smc_spectator_density(self.h, _ptr(a, C.c_float), _ptr(b, C.c_float)))
        return a, b

    def sources(self, ev):
        """Event ev's sources (smc_get_sources): xy [n][2] fm, kind (0/1
        participant of projectile/target in wound order, 2 pair), weight."""
        n = C.c_int()
        self._check(self.lib.smc_get_sources(self.h, ev, 0, None, None, None, C.byref(n)), allow=(SMC_ERR_CAPACITY,))
        k = n.value
        xy, kind, w = np.zeros((max(k, 1), 2)), np.zeros(max(k, 1), np.int32), np.zeros(max(k, 1), np.float32)
        self._check(self.lib.smc_get_sources(self.h, ev, k, _ptr(xy, C.c_double), _ptr(kind, C.c_int),
                                             _ptr(w, C.c_float), C.byref(n)))
        return xy[:k], kind[:k], w[:k]

    def recenter(self, n, turn):
        """MCnucl::recenterGrid (turn False) / rotateGrid (turn True) for the
        batch (smc_recenter); returns {xc, yc, angle, sum rho} per event."""
        rot = np.zeros((self.nev, 4))
        self._check(self.lib.smc_recenter(self.h, n, 1 if turn else 0, _ptr(rot, C.c_double)))
        return rot

    EXTRA_KEYS = ("tatb", "rho_binary", "ta", "tb", "spec_a", "spec_b", "tatb_rp", "rho_binary_rp", "ta_rp", "tb_rp")

    def average_profiles_extra(self, from_order, to_order, accept=None, reaction_plane=True):
        """smc_average_profiles_extra with every optional average: returns
        (sums, rp_sums or None, dict of the optional sums)."""
        nord = to_order - from_order + 1
        shape = (nord, self.nx, self.ny)
        sums = np.zeros(shape)
        rp = np.zeros(shape) if reaction_plane else None
        keys = [k for k in self.EXTRA_KEYS if reaction_plane or not k.endswith("_rp")]
        ex = {k: np.zeros(shape) for k in keys}
        arr = (C.c_void_p * 10)(*[ex[k].ctypes.data if k in ex else None for k in self.EXTRA_KEYS])
        acc = None
        if accept is not None:
            acc = np.ascontiguousarray(np.asarray(accept, dtype=np.uint8))
        self._check(self.lib.smc_average_profiles_extra(self.h, from_order, to_order,
                                                        None if acc is None else acc.ctypes.data_as(C.c_void_p),
                                                        _ptr(sums, C.c_double), _ptr(rp, C.c_double),
                                                        C.cast(arr, C.c_void_p)))
        return sums, rp, ex

    def average_profiles(self, from_order, to_order, accept=None, reaction_plane=False):
        """Sums over the ctx's accepted events of finalFactor * rho after the
        order-n recentring and turn (and the recentred-only sums if
        reaction_plane); see smc_average_profiles."""
        nord = to_order - from_order + 1
        sums = np.zeros((nord, self.nx, self.ny))
        rp = np.zeros((nord, self.nx, self.ny)) if reaction_plane else None
        acc = None
        if accept is not None:
            acc = np.ascontiguousarray(np.asarray(accept, dtype=np.uint8))
            if acc.shape != (self.nev,):
                raise ValueError("accept must have one entry per event")
        self._check(self.lib.smc_average_profiles(self.h, from_order, to_order,
                                                  None if acc is None else acc.ctypes.data_as(C.c_void_p),
                                                  _ptr(sums, C.c_double), _ptr(rp, C.c_double)))
        return (sums, rp) if reaction_plane else sums

    def kln_dndy(self, ta, tb):
        ta = np.ascontiguousarray(ta, np.float64)
        tb = np.ascontiguousarray(tb, np.float64)
        out = np.zeros(ta.shape[0])
        self._check(self.lib.smc_kln_dndy(self.h, ta.shape[0], _ptr(ta, C.c_double), _ptr(tb, C.c_double),
                                          _ptr(out, C.c_double)))
        return out

    def kln_make_table(self):
        n = self.derived.kln_tmax
        t = np.zeros((n, n))
        self._check(self.lib.smc_kln_make_table(self.h, _ptr(t, C.c_double)))
        return t

    # differentiable batches at frozen topology: theta = (w, R, a, alpha, k)
    def diff_setup(self, seed, event0, nev):
        th = np.zeros(5)
        self._check(self.lib.smc_diff_setup(self.h, seed, event0, nev, _ptr(th, C.c_double)))
        self.nev = nev
        self.diff_nev = nev
        return th

    def diff_forward(self, theta):
        th = np.ascontiguousarray(theta, np.float64)
        out = np.zeros((self.diff_nev, 8))
        self._check(self.lib.smc_diff_forward(self.h, _ptr(th, C.c_double), _ptr(out, C.c_double)))
        self._diff_theta = th.copy()
        return out

    def diff_backward(self, cot):
        cot = np.ascontiguousarray(cot, np.float64).reshape(self.diff_nev, 3)
        g = np.zeros(5)
        self._check(self.lib.smc_diff_backward(self.h, _ptr(cot, C.c_double), _ptr(g, C.c_double)))
        return g

    def diff_topology(self, ev):
        p = self.params
        up = np.zeros((p.Aproj, 3))
        ut = np.zeros((p.Atarg, 3))
        b = C.c_double()
        cap = 2 * 256 + (p.max_pairs if p.max_pairs > 0 else min(p.Aproj * p.Atarg, 8192))
        gv = np.zeros(cap)
        n = C.c_int()
        self._check(self.lib.smc_diff_topology(self.h, ev, _ptr(up, C.c_double), _ptr(ut, C.c_double), C.byref(b),
                                               _ptr(gv, C.c_double), cap, C.byref(n)))
        return dict(uvw_proj=up, uvw_targ=ut, b=b.value, gamma_v=gv[:n.value])

    def set_timing(self, on=True):
        self._check(self.lib.smc_set_timing(self.h, 1 if on else 0))

    def timing(self, kernel):
        ms = C.c_double()
        n = C.c_int()
        self._check(self.lib.smc_get_timing(self.h, kernel, C.byref(ms), C.byref(n)))
        return ms.value, n.value

    def synchronize(self):
        self._check(self.lib.smc_synchronize(self.h))


def _events(lists, width, dtype):
    """Concatenates per-event arrays of `width` columns; returns (flat, offsets)."""
    arrs = [np.asarray(a, dtype).reshape(-1, width) if len(a) else np.zeros((0, width), dtype) for a in lists]
    off = np.zeros(len(arrs) + 1, np.int32)
    off[1:] = np.cumsum([len(a) for a in arrs])
    flat = np.ascontiguousarray(np.concatenate(arrs) if arrs else np.zeros((0, width), dtype))
    return flat, off


class Profile3D:
    """The 3D (x, y, eta) profiles of scripts/generate_3d_profiles
    (include/smc_p3d.h): one handle per grid / ecm / random_flag."""

    def __init__(self, grid_nx=261, grid_ny=261, grid_neta=101, grid_dx=0.1, grid_dy=0.1, grid_deta=0.1,
                 ecm=19.6, random_flag=1):
        self.lib = load()
        self.params = P3dParams(grid_nx, grid_ny, grid_neta, grid_dx, grid_dy, grid_deta, ecm, random_flag)
        h = C.c_void_p()
        rc = self.lib.smc_p3d_create(C.byref(h), C.byref(self.params))
        if rc != SMC_OK:
            raise SmcError(rc, "smc_p3d_create failed")
        self.h = h
        self.shape = (grid_neta, grid_nx, grid_ny)

    def close(self):
        if self.h:
            self.lib.smc_p3d_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _check(self, rc):
        if rc != SMC_OK:
            raise SmcError(rc, self.lib.smc_p3d_last_error(self.h).decode())

    def derived(self):
        d = P3dDerived()
        eta, proj, targ = (np.zeros(P3D_TABLE) for _ in range(3))
        self._check(self.lib.smc_p3d_get_derived(self.h, C.byref(d), _ptr(eta, C.c_double), _ptr(proj, C.c_double),
                                                 _ptr(targ, C.c_double)))
        return d, eta, proj, targ

    def _lists(self, parts, bins):
        pxy, poff = _events([np.asarray(p, np.float64)[:, :2] if len(p) else [] for p in parts], 2, np.float64)
        pid, _ = _events([np.asarray(p, np.float64)[:, 2].astype(np.int32) if len(p) else [] for p in parts], 1,
                         np.int32)
        bxy, boff = _events([np.asarray(b, np.float64)[:, :2] if len(b) else [] for b in bins], 2, np.float64)
        return pxy, np.ascontiguousarray(pid.ravel()), poff, bxy, boff

    def sample(self, parts, bins, seed, event0=0):
        """parts[e]: rows (x, y, id); bins[e]: rows (x, y).  Returns per-event
        source records (x, y, eta, sigma_x, sigma_y, sigma_eta)."""
        pxy, pid, poff, bxy, boff = self._lists(parts, bins)
        ps, bs = np.zeros((poff[-1], 6)), np.zeros((boff[-1], 6))
        self._check(self.lib.smc_p3d_sample(self.h, len(parts), _ptr(pxy, C.c_double), _ptr(pid, C.c_int),
                                            _ptr(poff, C.c_int), _ptr(bxy, C.c_double), _ptr(boff, C.c_int), seed,
                                            event0, _ptr(ps, C.c_double), _ptr(bs, C.c_double)))
        return ([ps[poff[e]:poff[e + 1]] for e in range(len(parts))],
                [bs[boff[e]:boff[e + 1]] for e in range(len(bins))])

    def smear(self, srcs):
        """srcs[e]: source records; returns float32 [nev, neta, nx, ny]."""
        flat, off = _events(srcs, 6, np.float64)
        rho = np.zeros((len(srcs),) + self.shape, np.float32)
        self._check(self.lib.smc_p3d_smear(self.h, len(srcs), _ptr(flat, C.c_double), _ptr(off, C.c_int),
                                           _ptr(rho, C.c_float)))
        return rho

    def generate(self, parts, bins, seed, event0=0, binary=True):
        """Sampling and both grids: (rho_part, rho_binary or None, part_src, bin_src)."""
        pxy, pid, poff, bxy, boff = self._lists(parts, bins)
        nev = len(parts)
        rp = np.zeros((nev,) + self.shape, np.float32)
        rb = np.zeros((nev,) + self.shape, np.float32) if binary else None
        ps, bs = np.zeros((poff[-1], 6)), np.zeros((boff[-1], 6))
        self._check(self.lib.smc_p3d_generate(self.h, nev, _ptr(pxy, C.c_double), _ptr(pid, C.c_int),
                                              _ptr(poff, C.c_int), _ptr(bxy, C.c_double), _ptr(boff, C.c_int), seed,
                                              event0, _ptr(rp, C.c_float), _ptr(rb, C.c_float),
                                              _ptr(ps, C.c_double), _ptr(bs, C.c_double)))
        return rp, rb, ps, bs
