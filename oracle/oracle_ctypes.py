"""ctypes bindings for the CPU oracle (liboracle.so) and the reference harness.

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py.  The product (vsmc_amd) never imports this.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "liboracle.so")
REF_SO = os.path.join(HERE, "_ref", "libvsmcref.so")

MULTINOMIAL, RESIDUAL, STRATIFIED, SYSTEMATIC, RESIDUAL_STRATIFIED, RESIDUAL_SYSTEMATIC = range(6)
SCHEME_NAMES = ["multinomial", "residual", "stratified", "systematic",
                "residual_stratified", "residual_systematic"]

u32p = np.ctypeslib.ndpointer(np.uint32, flags="C")
u64p = np.ctypeslib.ndpointer(np.uint64, flags="C")
f64p = np.ctypeslib.ndpointer(np.float64, flags="C")


class TF64Engine(C.Structure):
    _fields_ = [("ctr", C.c_uint64 * 4), ("key", C.c_uint64 * 4),
                ("buf", C.c_uint64 * 4), ("index", C.c_uint32)]


class PH32Engine(C.Structure):
    _fields_ = [("ctr", C.c_uint32 * 4), ("key", C.c_uint32 * 2),
                ("buf", C.c_uint32 * 4), ("index", C.c_uint32)]


class TF32Engine(C.Structure):
    _fields_ = [("ctr", C.c_uint32 * 4), ("key", C.c_uint32 * 4),
                ("buf", C.c_uint32 * 4), ("index", C.c_uint32)]


class CvParam(C.Structure):
    _fields_ = [(n, C.c_double) for n in
                ("dt", "sd_pos", "sd_vel", "prior_sd_pos", "prior_sd_vel",
                 "obs_scale", "nu", "llh_const")]


class StepRecord(C.Structure):
    _fields_ = [("ess", C.c_double), ("log_z", C.c_double), ("mean_x", C.c_double),
                ("mean_y", C.c_double), ("resampled", C.c_int), ("draws", C.c_uint64)]


STEP_DTYPE = np.dtype([("ess", "f8"), ("log_z", "f8"), ("mean_x", "f8"), ("mean_y", "f8"),
                       ("resampled", "i4"), ("draws", "u8")], align=True)
assert STEP_DTYPE.itemsize == C.sizeof(StepRecord)

GMM_MAX_K = 6


class GmmConfig(C.Structure):
    _fields_ = [("K", C.c_int), ("sweeps", C.c_uint32), ("data", C.c_void_p),
                ("ndata", C.c_size_t), ("threshold", C.c_double), ("cess_drop", C.c_double),
                ("scheme", C.c_int), ("seed", C.c_uint64), ("resample_seed", C.c_uint64),
                ("max_steps", C.c_size_t)]


GMM_DTYPE = np.dtype([("alpha", "f8"), ("alpha_inc", "f8"), ("ess", "f8"), ("log_z", "f8"),
                      ("path_integrand", "f8"), ("path_log_z", "f8"), ("resampled", "i4"),
                      ("bisections", "u4"), ("accept", "u8", (3,))], align=True)


def cv_param(**kw):
    """Config-0 CV model constants (dt 0.1, sd 0.1/0.5, Student-t nu=10 scale 0.1)."""
    from math import lgamma, log, pi
    p = dict(dt=0.1, sd_pos=0.1, sd_vel=0.5, prior_sd_pos=2.0, prior_sd_vel=1.0,
             obs_scale=0.1, nu=10.0)
    p.update(kw)
    nu, s = p["nu"], p["obs_scale"]
    p["llh_const"] = lgamma((nu + 1) / 2) - lgamma(nu / 2) - 0.5 * log(nu * pi) - log(s)
    return CvParam(**p)


def _load(path):
    if not os.path.exists(path):
        raise FileNotFoundError(path)
    return C.CDLL(path)


class Oracle:
    def __init__(self, path=ORACLE_SO):
        L = self.lib = _load(path)
        L.orc_philox4x32_block.argtypes = [u32p, u32p, u32p]
        L.orc_threefry4x64_block.argtypes = [u64p, u64p, u64p]
        L.orc_threefry4x32_block.argtypes = [u32p, u32p, u32p]
        L.orc_tf64_seed.argtypes = [C.POINTER(TF64Engine), C.c_uint64]
        L.orc_tf64_discard.argtypes = [C.POINTER(TF64Engine), C.c_uint64]
        L.orc_tf64_stream.argtypes = [C.POINTER(TF64Engine), C.c_size_t, u64p]
        L.orc_ph32_seed.argtypes = [C.POINTER(PH32Engine), C.c_uint32]
        L.orc_ph32_discard.argtypes = [C.POINTER(PH32Engine), C.c_uint64]
        L.orc_ph32_stream.argtypes = [C.POINTER(PH32Engine), C.c_size_t, u32p]
        L.orc_tf32_seed.argtypes = [C.POINTER(TF32Engine), C.c_uint32]
        L.orc_tf32_stream.argtypes = [C.POINTER(TF32Engine), C.c_size_t, u32p]
        L.orc_canonical_tf64.argtypes = [C.POINTER(TF64Engine)]
        L.orc_canonical_tf64.restype = C.c_double
        L.orc_canonical_ph32.argtypes = [C.POINTER(PH32Engine)]
        L.orc_canonical_ph32.restype = C.c_double
        L.orc_fold_prefix.argtypes = [C.c_size_t, f64p, f64p]
        L.orc_asum.argtypes = [C.c_size_t, f64p]
        L.orc_asum.restype = C.c_double
        L.orc_inversion_u.argtypes = [C.c_size_t, C.c_size_t, f64p, f64p, u64p]
        L.orc_inversion_u.restype = C.c_size_t
        L.orc_cfrp_trans.argtypes = [C.c_size_t, C.c_size_t, u64p, u64p]
        L.orc_resample.argtypes = [C.c_int, C.c_size_t, C.c_size_t, C.POINTER(TF64Engine),
                                   f64p, u64p]
        L.orc_resample.restype = C.c_size_t
        L.orc_u01seq.argtypes = [C.c_int, C.c_size_t, C.POINTER(TF64Engine), f64p]
        L.orc_residual_split.argtypes = [C.c_size_t, C.c_size_t, f64p, u64p, f64p]
        L.orc_residual_split.restype = C.c_size_t
        L.orc_weightset_set_log_weight.argtypes = [C.c_size_t, f64p, f64p]
        L.orc_weightset_set_log_weight.restype = C.c_double
        L.orc_weightset_add_log_weight.argtypes = [C.c_size_t, f64p, f64p, f64p]
        L.orc_weightset_add_log_weight.restype = C.c_double
        L.orc_cv_normals.argtypes = [C.c_uint64, C.c_uint64, C.c_uint64, f64p]
        L.orc_cv_llh.argtypes = [C.POINTER(CvParam)] + [C.c_double] * 4
        L.orc_cv_llh.restype = C.c_double
        L.orc_cv_move_many.argtypes = [C.POINTER(CvParam), C.c_uint64, C.c_size_t, C.c_size_t,
                                       f64p, C.c_double, C.c_double, f64p, f64p]
        L.orc_cv_llh_many.argtypes = [C.POINTER(CvParam), C.c_size_t, f64p, C.c_double,
                                      C.c_double, f64p]
        L.orc_cburng4x32_stream.argtypes = [C.c_uint32, C.c_size_t, C.c_void_p]
        L.orc_normal01_4x32_53_stream.argtypes = [C.c_uint32, C.c_size_t, f64p]
        L.orc_clpf_run.argtypes = [C.c_size_t, C.c_size_t, f64p, C.c_uint32, C.c_uint64, C.c_int,
                                   C.c_double, f64p, C.c_void_p]
        L.orc_cv_run.argtypes = [C.POINTER(CvParam), C.c_size_t, C.c_size_t, f64p, C.c_uint64,
                                 C.c_uint64, C.c_int, C.c_double, C.c_void_p, C.c_void_p,
                                 C.c_void_p, C.c_int]
        L.orc_cv_run_timed.argtypes = [C.POINTER(CvParam), C.c_size_t, C.c_size_t, f64p,
                                       C.c_uint64, C.c_uint64, C.c_int, C.c_double, C.c_void_p,
                                       C.c_int, C.c_int, f64p]
        L.orc_gmm_prior.argtypes = [f64p, C.c_size_t, f64p]
        L.orc_gmm_log_prior.argtypes = [C.c_int, f64p, f64p, f64p]
        L.orc_gmm_log_prior.restype = C.c_double
        L.orc_gmm_log_likelihood.argtypes = [C.c_int, f64p, f64p, f64p, f64p, C.c_size_t]
        L.orc_gmm_log_likelihood.restype = C.c_double
        L.orc_gmm_init_particle.argtypes = [C.c_int, C.c_uint64, C.c_uint64, f64p, f64p]
        L.orc_gmm_mh_move.argtypes = [C.c_int, C.c_int, C.c_uint64, C.c_uint64, C.c_uint32,
                                      C.c_uint32, C.c_double, f64p, f64p, f64p, C.c_size_t, f64p,
                                      C.POINTER(C.c_double), C.POINTER(C.c_double)]
        L.orc_gmm_mh_move.restype = C.c_int
        L.orc_gmm_cess.argtypes = [C.c_size_t, f64p, f64p, C.c_double]
        L.orc_gmm_cess.restype = C.c_double
        L.orc_gmm_alpha_search.argtypes = [C.c_size_t, f64p, f64p, C.c_double, C.c_double,
                                           C.POINTER(C.c_double), C.POINTER(C.c_double)]
        L.orc_gmm_alpha_search.restype = C.c_uint32
        L.orc_gmm_run.argtypes = [C.POINTER(GmmConfig), C.c_size_t, C.c_void_p, C.c_void_p,
                                  C.c_void_p, C.c_void_p, C.POINTER(C.c_size_t), C.c_int]

    # --- engines ---
    def tf64_engine(self, seed):
        e = TF64Engine()
        self.lib.orc_tf64_seed(C.byref(e), seed)
        return e

    def ph32_engine(self, seed):
        e = PH32Engine()
        self.lib.orc_ph32_seed(C.byref(e), seed)
        return e

    def philox4x32_block(self, ctr, key):
        out = np.zeros(4, np.uint32)
        self.lib.orc_philox4x32_block(np.asarray(ctr, np.uint32), np.asarray(key, np.uint32), out)
        return out

    def threefry4x64_block(self, ctr, key):
        out = np.zeros(4, np.uint64)
        self.lib.orc_threefry4x64_block(np.asarray(ctr, np.uint64), np.asarray(key, np.uint64), out)
        return out

    def threefry4x32_block(self, ctr, key):
        out = np.zeros(4, np.uint32)
        self.lib.orc_threefry4x32_block(np.asarray(ctr, np.uint32), np.asarray(key, np.uint32), out)
        return out

    def tf64_stream(self, seed, n, skip=0, ctr=None):
        e = self.tf64_engine(seed)
        if ctr is not None:
            for j in range(4):
                e.ctr[j] = int(ctr[j])
            e.index = 4
        if skip:
            self.lib.orc_tf64_discard(C.byref(e), skip)
        out = np.zeros(n, np.uint64)
        self.lib.orc_tf64_stream(C.byref(e), n, out)
        return out

    def ph32_stream(self, seed, n, skip=0, ctr=None):
        e = self.ph32_engine(seed)
        if ctr is not None:
            for j in range(4):
                e.ctr[j] = int(ctr[j])
            e.index = 4
        if skip:
            self.lib.orc_ph32_discard(C.byref(e), skip)
        out = np.zeros(n, np.uint32)
        self.lib.orc_ph32_stream(C.byref(e), n, out)
        return out

    def tf32_stream(self, seed, n):
        e = TF32Engine()
        self.lib.orc_tf32_seed(C.byref(e), seed)
        out = np.zeros(n, np.uint32)
        self.lib.orc_tf32_stream(C.byref(e), n, out)
        return out

    def canonical_tf64(self, seed, n):
        e = self.tf64_engine(seed)
        return np.array([self.lib.orc_canonical_tf64(C.byref(e)) for _ in range(n)])

    def canonical_ph32(self, seed, n):
        e = self.ph32_engine(seed)
        return np.array([self.lib.orc_canonical_ph32(C.byref(e)) for _ in range(n)])

    # --- resampling ---
    def fold_prefix(self, x):
        x = np.ascontiguousarray(x, np.float64)
        out = np.zeros_like(x)
        self.lib.orc_fold_prefix(x.size, x, out)
        return out

    def u01seq(self, kind, N, seed=None, engine=None):
        e = engine if engine is not None else self.tf64_engine(seed)
        u = np.zeros(N, np.float64)
        self.lib.orc_u01seq(kind, N, C.byref(e), u)
        return u, e

    def resample(self, scheme, w, seed=None, engine=None):
        w = np.ascontiguousarray(w, np.float64)
        N = w.size
        e = engine if engine is not None else self.tf64_engine(seed)
        rep = np.zeros(N, np.uint64)
        draws = self.lib.orc_resample(scheme, N, N, C.byref(e), w, rep)
        return rep, draws, e

    def cfrp_trans(self, rep, N=None):
        rep = np.ascontiguousarray(rep, np.uint64)
        N = rep.size if N is None else N
        cf = np.zeros(N, np.uint64)
        self.lib.orc_cfrp_trans(rep.size, N, rep, cf)
        return cf

    def inversion_u(self, w, u):
        w = np.ascontiguousarray(w, np.float64)
        u = np.ascontiguousarray(u, np.float64)
        rep = np.zeros(w.size, np.uint64)
        seen = self.lib.orc_inversion_u(w.size, u.size, w, u, rep)
        return rep, seen

    def residual_split(self, w, N=None):
        w = np.ascontiguousarray(w, np.float64)
        N = w.size if N is None else N
        ip = np.zeros(w.size, np.uint64)
        rr = np.zeros(w.size, np.float64)
        NN = self.lib.orc_residual_split(w.size, N, w, ip, rr)
        return ip, rr, NN

    def weightset(self, lw0, inc=None):
        lw = np.array(lw0, np.float64)
        w = np.zeros_like(lw)
        if inc is None:
            ess = self.lib.orc_weightset_set_log_weight(lw.size, lw, w)
        else:
            self.lib.orc_weightset_set_log_weight(lw.size, lw, w)
            ess = self.lib.orc_weightset_add_log_weight(lw.size, lw,
                                                        np.ascontiguousarray(inc, np.float64), w)
        return lw, w, ess

    # --- CV filter ---
    def cv_normals(self, seed, i, t):
        z = np.zeros(4)
        self.lib.orc_cv_normals(seed, i, t, z)
        return z

    def cv_move_many(self, seed, t, states, ox, oy, param=None):
        """One CV step (prior draw at t == 0) of every row of states (n x 4):
        the new states and the log-likelihood increments, in FP64."""
        st = np.ascontiguousarray(states, np.float64).reshape(-1, 4)
        out = np.zeros_like(st)
        inc = np.zeros(st.shape[0])
        p = param if param is not None else cv_param()
        self.lib.orc_cv_move_many(C.byref(p), seed, t, st.shape[0], st, ox, oy, out, inc)
        return out, inc

    def cv_llh_many(self, states, ox, oy, param=None):
        """FP64 log-likelihood increment of observation (ox, oy) at every row
        (x, y, vx, vy) of states."""
        st = np.ascontiguousarray(states, np.float64).reshape(-1, 4)
        inc = np.zeros(st.shape[0])
        p = param if param is not None else cv_param()
        self.lib.orc_cv_llh_many(C.byref(p), st.shape[0], st, ox, oy, inc)
        return inc

    def cburng4x32_stream(self, seed, n):
        """n draws of urng.h's cburng4x32 (Threefry4x32-20) seeded with seed."""
        out = np.zeros(n, np.uint32)
        self.lib.orc_cburng4x32_stream(seed, n, out.ctypes.data)
        return out

    def normal01_4x32_53_stream(self, seed, n):
        """n draws of normal01.h's normal01_4x32_53 over cburng4x32(seed)."""
        out = np.zeros(n)
        self.lib.orc_normal01_4x32_53_stream(seed, n, out)
        return out

    def clpf_run(self, N, obs, seed, resample_seed, scheme=SYSTEMATIC, threshold=0.5):
        """The OpenCL pf example's kernels (pf_cl.cl) on vSMC's weight set and
        resampler: final states (N x 4) and per-step records."""
        obs = np.ascontiguousarray(obs, np.float64).reshape(-1, 2)
        T = obs.shape[0]
        st = np.zeros((N, 4))
        rec = np.zeros(T, STEP_DTYPE)
        if self.lib.orc_clpf_run(N, T, obs, seed, resample_seed, scheme, threshold, st,
                                 rec.ctypes.data) != 0:
            raise MemoryError("orc_clpf_run failed")
        return st, rec

    def cv_run(self, N, obs, seed, resample_seed, scheme=SYSTEMATIC, threshold=0.5,
               param=None, threads=0, want_state=True):
        obs = np.ascontiguousarray(obs, np.float64).reshape(-1, 2)
        T = obs.shape[0]
        p = param if param is not None else cv_param()
        st = np.zeros((N, 4)) if want_state else None
        lw = np.zeros(N) if want_state else None
        rec = np.zeros(T, STEP_DTYPE)
        rc = self.lib.orc_cv_run(C.byref(p), N, T, obs, seed, resample_seed, scheme, threshold,
                                 st.ctypes.data if want_state else None,
                                 lw.ctypes.data if want_state else None,
                                 rec.ctypes.data, threads)
        if rc != 0:
            raise MemoryError("orc_cv_run failed")
        return st, lw, rec

    def cv_run_timed(self, N, obs, seed, resample_seed, scheme=SYSTEMATIC, threshold=0.5,
                     param=None, threads=0, backend="omp"):
        """The CPU path timed per phase (move, normalise, resample, copy) on
        vSMC's OpenMP backend ("omp") or std::thread backend ("threads")."""
        obs = np.ascontiguousarray(obs, np.float64).reshape(-1, 2)
        T = obs.shape[0]
        p = param if param is not None else cv_param()
        rec = np.zeros(T, STEP_DTYPE)
        phase = np.zeros(4)
        rc = self.lib.orc_cv_run_timed(C.byref(p), N, T, obs, seed, resample_seed, scheme,
                                       threshold, rec.ctypes.data, threads,
                                       1 if backend == "threads" else 0, phase)
        if rc != 0:
            raise MemoryError("orc_cv_run_timed failed")
        return rec, phase

    # --- GMM tempered SMC ---
    def gmm_prior(self, data):
        data = np.ascontiguousarray(data, np.float64)
        out = np.zeros(4)
        self.lib.orc_gmm_prior(data, data.size, out)
        return out

    def gmm_log_prior(self, K, p, prior):
        p = np.ascontiguousarray(p, np.float64)
        return self.lib.orc_gmm_log_prior(K, p[:K].copy(), p[K:2 * K].copy(),
                                          np.ascontiguousarray(prior, np.float64))

    def gmm_log_likelihood(self, K, p, data):
        p = np.ascontiguousarray(p, np.float64)
        data = np.ascontiguousarray(data, np.float64)
        return self.lib.orc_gmm_log_likelihood(K, p[:K].copy(), p[K:2 * K].copy(),
                                               p[2 * K:3 * K].copy(), data, data.size)

    def gmm_init_particle(self, K, seed, i, prior):
        out = np.zeros(3 * K)
        self.lib.orc_gmm_init_particle(K, seed, i, np.ascontiguousarray(prior, np.float64), out)
        return out

    def gmm_mh_move(self, K, move, seed, i, it, sweep, alpha, sd, prior, data, p, llh, lp):
        p = np.array(p, np.float64)
        a, b = C.c_double(llh), C.c_double(lp)
        data = np.ascontiguousarray(data, np.float64)
        acc = self.lib.orc_gmm_mh_move(K, move, seed, i, it, sweep, alpha,
                                       np.ascontiguousarray(sd, np.float64),
                                       np.ascontiguousarray(prior, np.float64), data, data.size,
                                       p, C.byref(a), C.byref(b))
        return acc, p, a.value, b.value

    def gmm_cess(self, W, llh, a_diff):
        """CESS of the reweighting by exp(a_diff * llh) of normalised weights W."""
        W = np.ascontiguousarray(W, np.float64)
        llh = np.ascontiguousarray(llh, np.float64)
        return self.lib.orc_gmm_cess(W.size, W, llh, a_diff)

    def gmm_alpha_search(self, W, llh, alpha, drop):
        """Next temperature: (alpha, alpha_inc, CESS evaluations)."""
        W = np.ascontiguousarray(W, np.float64)
        llh = np.ascontiguousarray(llh, np.float64)
        a, ai = C.c_double(), C.c_double()
        it = self.lib.orc_gmm_alpha_search(W.size, W, llh, alpha, drop, C.byref(a), C.byref(ai))
        return a.value, ai.value, int(it)

    def gmm_run(self, N, data, K=4, sweeps=5, seed=1, resample_seed=2, scheme=STRATIFIED,
                threshold=0.5, cess_drop=0.1, max_steps=1000, threads=0, want_state=True):
        data = np.ascontiguousarray(data, np.float64)
        cfg = GmmConfig(K, sweeps, data.ctypes.data, data.size, threshold, cess_drop, scheme,
                        seed, resample_seed, max_steps)
        rec = np.zeros(max_steps + 1, GMM_DTYPE)
        params = np.zeros((N, 3 * K)) if want_state else None
        llh = np.zeros(N) if want_state else None
        lw = np.zeros(N) if want_state else None
        rows = C.c_size_t()
        rc = self.lib.orc_gmm_run(C.byref(cfg), N,
                                  params.ctypes.data if want_state else None,
                                  llh.ctypes.data if want_state else None,
                                  lw.ctypes.data if want_state else None,
                                  rec.ctypes.data, C.byref(rows), threads)
        if rc != 0:
            raise RuntimeError(f"orc_gmm_run failed ({rc})")
        return params, llh, lw, rec[:rows.value]


class RefHarness:
    """The compiled vSMC headers (oracle/_ref/libvsmcref.so), when built."""

    def __init__(self, path=REF_SO):
        L = self.lib = _load(path)
        L.ref_philox4x32_block.argtypes = [u32p, u32p, u32p]
        L.ref_threefry4x64_block.argtypes = [u64p, u64p, u64p]
        L.ref_threefry4x32_block.argtypes = [u32p, u32p, u32p]
        L.ref_philox4x32_stream.argtypes = [C.c_uint32, C.c_void_p, C.c_uint64, C.c_size_t, u32p]
        L.ref_threefry4x64_stream.argtypes = [C.c_uint64, C.c_void_p, C.c_uint64, C.c_size_t,
                                              u64p]
        L.ref_threefry4x32_stream.argtypes = [C.c_uint32, C.c_size_t, u32p]
        L.ref_canonical_tf64.argtypes = [C.c_uint64, C.c_size_t, f64p]
        L.ref_canonical_ph32.argtypes = [C.c_uint32, C.c_size_t, f64p]
        L.ref_u01seq.argtypes = [C.c_int, C.c_size_t, C.c_uint64, f64p, u64p]
        L.ref_resample.argtypes = [C.c_int, C.c_size_t, C.c_uint64, f64p, u64p, u64p, u64p]
        L.ref_cfrp_trans.argtypes = [C.c_size_t, C.c_size_t, u64p, u64p]
        L.ref_weightset.argtypes = [C.c_size_t, f64p, C.c_void_p, f64p, f64p]
        L.ref_weightset.restype = C.c_double
        L.ref_cv_run.argtypes = [C.POINTER(CvParam), C.c_size_t, C.c_size_t, f64p, C.c_uint64,
                                 C.c_uint64, C.c_int, C.c_double, C.c_void_p, C.c_void_p,
                                 C.c_void_p]
        L.ref_cess.argtypes = [C.c_size_t, f64p, f64p, C.c_int, f64p]
        L.ref_cess.restype = C.c_double
        L.ref_alpha_ess.argtypes = [C.c_size_t, f64p, f64p, C.c_double, C.c_double,
                                    C.POINTER(C.c_double), C.POINTER(C.c_double),
                                    C.POINTER(C.c_uint64), f64p]
        L.ref_u01_32.argtypes = [u32p, C.c_size_t, f64p]
        L.ref_u01_64.argtypes = [u64p, C.c_size_t, f64p]

    def u01_32(self, u):
        """The eight 32-bit u01 conversions (closed_closed, closed_open,
        open_closed, open_open; _24 then _53) of every u: (n, 8)."""
        u = np.ascontiguousarray(u, np.uint32)
        out = np.zeros((u.size, 8))
        self.lib.ref_u01_32(u, u.size, out)
        return out

    def u01_64(self, u):
        """The four 64-bit u01 conversions of every u: (n, 4)."""
        u = np.ascontiguousarray(u, np.uint64)
        out = np.zeros((u.size, 4))
        self.lib.ref_u01_64(u, u.size, out)
        return out

    def philox4x32_block(self, ctr, key):
        out = np.zeros(4, np.uint32)
        self.lib.ref_philox4x32_block(np.asarray(ctr, np.uint32), np.asarray(key, np.uint32), out)
        return out

    def threefry4x64_block(self, ctr, key):
        out = np.zeros(4, np.uint64)
        self.lib.ref_threefry4x64_block(np.asarray(ctr, np.uint64), np.asarray(key, np.uint64),
                                        out)
        return out

    def threefry4x32_block(self, ctr, key):
        out = np.zeros(4, np.uint32)
        self.lib.ref_threefry4x32_block(np.asarray(ctr, np.uint32), np.asarray(key, np.uint32),
                                        out)
        return out

    def ph32_stream(self, seed, n, skip=0, ctr=None):
        out = np.zeros(n, np.uint32)
        c = None if ctr is None else np.asarray(ctr, np.uint32)
        self.lib.ref_philox4x32_stream(seed, None if c is None else c.ctypes.data, skip, n, out)
        return out

    def tf64_stream(self, seed, n, skip=0, ctr=None):
        out = np.zeros(n, np.uint64)
        c = None if ctr is None else np.asarray(ctr, np.uint64)
        self.lib.ref_threefry4x64_stream(seed, None if c is None else c.ctypes.data, skip, n, out)
        return out

    def tf32_stream(self, seed, n):
        out = np.zeros(n, np.uint32)
        self.lib.ref_threefry4x32_stream(seed, n, out)
        return out

    def canonical_tf64(self, seed, n):
        out = np.zeros(n)
        self.lib.ref_canonical_tf64(seed, n, out)
        return out

    def canonical_ph32(self, seed, n):
        out = np.zeros(n)
        self.lib.ref_canonical_ph32(seed, n, out)
        return out

    def u01seq(self, kind, N, seed):
        u = np.zeros(N)
        nxt = np.zeros(4, np.uint64)
        self.lib.ref_u01seq(kind, N, seed, u, nxt)
        return u, nxt

    def resample(self, scheme, w, seed):
        w = np.ascontiguousarray(w, np.float64)
        rep = np.zeros(w.size, np.uint64)
        cf = np.zeros(w.size, np.uint64)
        nxt = np.zeros(4, np.uint64)
        self.lib.ref_resample(scheme, w.size, seed, w, rep, cf, nxt)
        return rep, cf, nxt

    def cfrp_trans(self, rep, N=None):
        rep = np.ascontiguousarray(rep, np.uint64)
        N = rep.size if N is None else N
        cf = np.zeros(N, np.uint64)
        self.lib.ref_cfrp_trans(rep.size, N, rep, cf)
        return cf

    def weightset(self, lw0, inc=None):
        lw0 = np.ascontiguousarray(lw0, np.float64)
        lw = np.zeros_like(lw0)
        w = np.zeros_like(lw0)
        incp = None
        if inc is not None:
            inc = np.ascontiguousarray(inc, np.float64)
            incp = inc.ctypes.data
        ess = self.lib.ref_weightset(lw0.size, lw0, incp, lw, w)
        return lw, w, ess

    def cess(self, lw, mul, use_log=False):
        """WeightSet(N).set_log_weight(lw); cess(mul, use_log): (CESS, weights)."""
        lw = np.ascontiguousarray(lw, np.float64)
        mul = np.ascontiguousarray(mul, np.float64)
        w = np.zeros_like(lw)
        c = self.lib.ref_cess(lw.size, lw, mul, 1 if use_log else 0, w)
        return c, w

    def alpha_ess(self, lw, llh, alpha, drop):
        """alpha_ess<T, cess01<T>>::alpha_iter from alpha: (alpha, alpha_inc, CESS
        evaluations, weights)."""
        lw = np.ascontiguousarray(lw, np.float64)
        llh = np.ascontiguousarray(llh, np.float64)
        w = np.zeros_like(lw)
        a, ai, n = C.c_double(), C.c_double(), C.c_uint64()
        self.lib.ref_alpha_ess(lw.size, lw, llh, alpha, drop, C.byref(a), C.byref(ai),
                               C.byref(n), w)
        return a.value, ai.value, int(n.value), w

    def cv_run(self, N, obs, seed, resample_seed, scheme=SYSTEMATIC, threshold=0.5, param=None):
        obs = np.ascontiguousarray(obs, np.float64).reshape(-1, 2)
        T = obs.shape[0]
        p = param if param is not None else cv_param()
        st = np.zeros((N, 4))
        lw = np.zeros(N)
        rec = np.zeros(T, STEP_DTYPE)
        self.lib.ref_cv_run(C.byref(p), N, T, obs, seed, resample_seed, scheme, threshold,
                            st.ctypes.data, lw.ctypes.data, rec.ctypes.data)
        return st, lw, rec


def ref_available():
    return os.path.exists(REF_SO)
