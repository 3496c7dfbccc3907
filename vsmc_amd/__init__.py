"""Python bindings (ctypes) over the C ABI in include/vsmcx.h.

The compute path is the HIP library vsmc_amd/libvsmcx.so (gfx950).  There is
no CPU fallback: if the library is missing every entry point raises.
"""
import ctypes as C
import math
import os
import weakref

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libvsmcx.so")

MULTINOMIAL, RESIDUAL, STRATIFIED, SYSTEMATIC, RESIDUAL_STRATIFIED, RESIDUAL_SYSTEMATIC = range(6)
SCHEMES = {"multinomial": 0, "residual": 1, "stratified": 2, "systematic": 3,
           "residual_stratified": 4, "residual_systematic": 5}
MIXED, FP64_FUSED, FP64_PLAIN = 0, 1, 2
BATCH_PERSISTENT, BATCH_TICKS = 0, 1

ERRORS = {-1: "VSMCX_ERR_ARG", -2: "VSMCX_ERR_HIP", -3: "VSMCX_ERR_DEGENERATE",
          -4: "VSMCX_ERR_BUDGET", -5: "VSMCX_ERR_INTERNAL", -6: "VSMCX_ERR_STATE"}


class VsmcxError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{ERRORS.get(code, code)}: {msg}")
        self.code = code


class ThreefryState(C.Structure):
    """vsmcx_threefry4x64_state: ctr, key, buffer, index of a Threefry4x64 engine."""
    _fields_ = [("ctr", C.c_uint64 * 4), ("key", C.c_uint64 * 4),
                ("buffer", C.c_uint64 * 4), ("index", C.c_uint64)]

    @classmethod
    def seeded(cls, seed):
        """Threefry4x64(seed): ctr 0, key {seed, 0, 0, 0}, index 4."""
        s = cls()
        s.key[0] = seed
        s.index = 4
        return s

    def as_tuple(self):
        return (tuple(self.ctr), tuple(self.key), tuple(self.buffer), int(self.index))


class FilterDesc(C.Structure):
    _fields_ = [("n", C.c_uint64), ("horizon", C.c_uint64), ("precision", C.c_int),
                ("scheme", C.c_int), ("threshold", C.c_double), ("seed", C.c_uint64),
                ("resample_seed", C.c_uint64),
                ("dt", C.c_double), ("sd_pos", C.c_double), ("sd_vel", C.c_double),
                ("prior_sd_pos", C.c_double), ("prior_sd_vel", C.c_double),
                ("obs_scale", C.c_double), ("nu", C.c_double), ("llh_const", C.c_double)]


class BatchDesc(C.Structure):
    _fields_ = [("nfilters", C.c_uint64), ("n", C.c_void_p), ("horizon", C.c_void_p),
                ("seed", C.c_void_p), ("resample_seed", C.c_void_p), ("precision", C.c_int),
                ("scheme", C.c_int), ("threshold", C.c_double),
                ("dt", C.c_double), ("sd_pos", C.c_double), ("sd_vel", C.c_double),
                ("prior_sd_pos", C.c_double), ("prior_sd_vel", C.c_double),
                ("obs_scale", C.c_double), ("nu", C.c_double), ("llh_const", C.c_double)]


MODEL_CV, MODEL_LG = 0, 1


class ServeDesc(C.Structure):
    _fields_ = [("hbm_budget", C.c_uint64), ("max_jobs", C.c_uint32), ("pad", C.c_uint32),
                ("max_horizon", C.c_uint64), ("precision", C.c_int), ("scheme", C.c_int),
                ("threshold", C.c_double),
                ("dt", C.c_double), ("sd_pos", C.c_double), ("sd_vel", C.c_double),
                ("prior_sd_pos", C.c_double), ("prior_sd_vel", C.c_double),
                ("obs_scale", C.c_double), ("nu", C.c_double), ("llh_const", C.c_double),
                ("lg_dim", C.c_uint32), ("pad2", C.c_uint32), ("lg_a", C.c_void_p),
                ("lg_q", C.c_double), ("lg_r", C.c_double), ("lg_s0", C.c_double)]


class JobDesc(C.Structure):
    _fields_ = [("model", C.c_int), ("pad", C.c_int), ("n", C.c_uint64), ("horizon", C.c_uint64),
                ("seed", C.c_uint64), ("resample_seed", C.c_uint64)]


class ServeSummary(C.Structure):
    _fields_ = [("steps", C.c_uint64), ("particle_updates", C.c_uint64),
                ("jobs_done", C.c_uint64), ("jobs_active", C.c_uint64),
                ("jobs_waiting", C.c_uint64), ("p50_us", C.c_double), ("p99_us", C.c_double),
                ("busy_ms", C.c_double), ("elapsed_ms", C.c_double),
                ("peak_particles", C.c_uint64), ("capacity_particles", C.c_uint64)]


STEP_DTYPE = np.dtype([("ess", "f8"), ("log_z", "f8"), ("mean_x", "f8"), ("mean_y", "f8"),
                       ("resampled", "i4"), ("pad", "i4"), ("draws", "u8")])


class GmmDesc(C.Structure):
    _fields_ = [("n", C.c_uint64), ("components", C.c_uint32), ("sweeps", C.c_uint32),
                ("precision", C.c_int), ("scheme", C.c_int), ("threshold", C.c_double),
                ("cess_drop", C.c_double), ("seed", C.c_uint64), ("resample_seed", C.c_uint64),
                ("max_steps", C.c_uint64)]


GMM_DTYPE = np.dtype([("alpha", "f8"), ("alpha_inc", "f8"), ("ess", "f8"), ("log_z", "f8"),
                      ("path_integrand", "f8"), ("path_log_z", "f8"), ("resampled", "i4"),
                      ("bisections", "u4"), ("accept", "u8", (3,))])

_lib = None


def lib():
    """Load libvsmcx.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: run `python vsmc_amd/build.py` "
                           "(there is no CPU fallback)")
    L = C.CDLL(LIB_PATH)
    vp, u64, u32, i32, f64 = C.c_void_p, C.c_uint64, C.c_uint32, C.c_int, C.c_double
    P = C.POINTER
    sig = {
        "vsmcx_last_error": ([], C.c_char_p),
        "vsmcx_version": ([], C.c_char_p),
        "vsmcx_ctx_create": ([i32, u64, P(vp)], i32),
        "vsmcx_ctx_destroy": ([vp], i32),
        "vsmcx_ctx_sync": ([vp], i32),
        "vsmcx_ctx_allocated": ([vp, P(u64)], i32),
        "vsmcx_stream_probe": ([vp, u64, i32, vp], i32),
        "vsmcx_philox4x32_fill": ([vp, vp, vp, u64, vp], i32),
        "vsmcx_threefry4x64_fill": ([vp, vp, vp, u64, vp], i32),
        "vsmcx_threefry4x32_fill": ([vp, vp, vp, u64, vp], i32),
        "vsmcx_resample": ([vp, i32, u64, u64, P(ThreefryState), vp, vp, vp], i32),
        "vsmcx_resample_noalloc": ([vp, i32, u64, u64, P(ThreefryState), vp, vp, vp], i32),
        "vsmcx_fold_prefix": ([vp, u64, vp, vp], i32),
        "vsmcx_filter_create": ([vp, P(FilterDesc), P(vp)], i32),
        "vsmcx_filter_destroy": ([vp], i32),
        "vsmcx_filter_set_obs": ([vp, u64, vp], i32),
        "vsmcx_filter_run": ([vp, u64], i32),
        "vsmcx_filter_finish": ([vp], i32),
        "vsmcx_filter_steps": ([vp, P(u64)], i32),
        "vsmcx_filter_history": ([vp, u64, vp], i32),
        "vsmcx_filter_state": ([vp, vp, vp], i32),
        "vsmcx_filter_engine": ([vp, P(ThreefryState)], i32),
        "vsmcx_filter_time": ([vp, u64, P(f64), P(f64)], i32),
        "vsmcx_filter_debug_step": ([vp, vp, vp, vp, vp, P(u64), P(ThreefryState), P(i32)], i32),
        "vsmcx_timer_create": ([vp, u64, P(vp)], i32),
        "vsmcx_timer_record": ([vp, u64], i32),
        "vsmcx_timer_elapsed": ([vp, u64, u64, P(f64)], i32),
        "vsmcx_timer_destroy": ([vp], i32),
        "vsmcx_filter_profile": ([vp, i32], i32),
        "vsmcx_filter_move_times": ([vp, u64, P(u64), vp], i32),
        "vsmcx_batch_create": ([vp, P(BatchDesc), P(vp)], i32),
        "vsmcx_batch_destroy": ([vp], i32),
        "vsmcx_batch_set_obs": ([vp, u64, u64, u64, vp], i32),
        "vsmcx_batch_run": ([vp, u64, P(f64)], i32),
        "vsmcx_batch_set_launch": ([vp, C.c_uint32, i32], i32),
        "vsmcx_batch_set_strategy": ([vp, i32], i32),
        "vsmcx_batch_get_strategy": ([vp, P(i32)], i32),
        "vsmcx_batch_tune": ([vp, u64, P(C.c_uint32), P(i32)], i32),
        "vsmcx_batch_tune_report": ([vp, u64, P(u64), vp, vp, vp, vp], i32),
        "vsmcx_batch_steps": ([vp, u64, P(u64)], i32),
        "vsmcx_batch_history": ([vp, u64, u64, vp], i32),
        "vsmcx_batch_state": ([vp, u64, vp, vp], i32),
        "vsmcx_batch_particles": ([vp, P(u64)], i32),
        "vsmcx_serve_create": ([vp, P(ServeDesc), P(vp)], i32),
        "vsmcx_serve_destroy": ([vp], i32),
        "vsmcx_serve_submit": ([vp, P(JobDesc), P(u64)], i32),
        "vsmcx_serve_observe": ([vp, u64, vp], i32),
        "vsmcx_serve_poll": ([vp, i32, P(u64)], i32),
        "vsmcx_serve_observe_many": ([vp, u64, vp, vp], i32),
        "vsmcx_serve_ready": ([vp, u64, vp, P(u64)], i32),
        "vsmcx_serve_job": ([vp, u64, P(i32), P(u64)], i32),
        "vsmcx_serve_history": ([vp, u64, u64, vp], i32),
        "vsmcx_serve_forget": ([vp, u64], i32),
        "vsmcx_serve_stats": ([vp, P(ServeSummary)], i32),
        "vsmcx_ctx_stream": ([vp, P(vp)], i32),
        "vsmcx_weights_create": ([vp, u64, i32, f64, u64, u64, P(vp)], i32),
        "vsmcx_weights_destroy": ([vp], i32),
        "vsmcx_weights_buffer": ([vp, P(vp)], i32),
        "vsmcx_weights_set_log": ([vp, vp], i32),
        "vsmcx_weights_add_log": ([vp, vp], i32),
        "vsmcx_weights_resample": ([vp, vp], i32),
        "vsmcx_weights_ancestors": ([vp, P(vp)], i32),
        "vsmcx_weights_steps": ([vp, P(u64)], i32),
        "vsmcx_weights_history": ([vp, u64, vp], i32),
        "vsmcx_weights_read": ([vp, vp], i32),
        "vsmcx_filter_save": ([vp, C.c_char_p], i32),
        "vsmcx_filter_load": ([vp, C.c_char_p, P(vp)], i32),
        "vsmcx_filter_describe": ([vp, P(FilterDesc)], i32),
        "vsmcx_state_create": ([vp, u64, u32, i32, P(vp)], i32),
        "vsmcx_state_destroy": ([vp], i32),
        "vsmcx_state_clone": ([vp, P(vp)], i32),
        "vsmcx_state_size": ([vp, P(u64), P(u32)], i32),
        "vsmcx_state_write": ([vp, vp], i32),
        "vsmcx_state_read": ([vp, vp], i32),
        "vsmcx_state_copy": ([vp, u64, vp], i32),
        "vsmcx_state_cv_move": ([vp, P(FilterDesc), u64, vp, vp], i32),
        "vsmcx_gmm_create": ([vp, P(GmmDesc), vp, u64, P(vp)], i32),
        "vsmcx_gmm_destroy": ([vp], i32),
        "vsmcx_gmm_initialize": ([vp], i32),
        "vsmcx_gmm_iterate": ([vp, u64, P(u64), P(f64)], i32),
        "vsmcx_gmm_rows": ([vp, P(u64)], i32),
        "vsmcx_gmm_times": ([vp, P(f64), P(f64), P(f64)], i32),
        "vsmcx_gmm_history": ([vp, u64, vp], i32),
        "vsmcx_gmm_state": ([vp, vp, vp, vp], i32),
    }
    for name, (args, res) in sig.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = res
    _lib = L
    return L


def _check(rc):
    if rc != 0:
        raise VsmcxError(rc, lib().vsmcx_last_error().decode())


def _ptr(a):
    return None if a is None else a.ctypes.data


class Context:
    """One HIP stream on one device (vsmcx_ctx).  Objects created on it
    (filters, batches, timers, ...) are closed before it, so none outlives the
    context's device memory."""

    def __init__(self, device=0, budget=0):
        self.h = C.c_void_p()
        self._children = weakref.WeakSet()
        _check(lib().vsmcx_ctx_create(device, budget, C.byref(self.h)))

    def _adopt(self, obj):
        self._children.add(obj)

    def close(self):
        if self.h:
            for c in list(self._children):
                c.close()
            lib().vsmcx_ctx_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def allocated(self):
        v = C.c_uint64()
        _check(lib().vsmcx_ctx_allocated(self.h, C.byref(v)))
        return v.value

    def stream_probe(self, nbytes=1 << 30, reps=5):
        """Streaming rates of the device in 10^9 bytes/s: a copy of nbytes
        (read + write bytes), a read and a write, each the best of reps."""
        g = np.zeros(3)
        _check(lib().vsmcx_stream_probe(self.h, nbytes, reps, g.ctypes.data))
        return {"copy_GBps": float(g[0]), "read_GBps": float(g[1]), "write_GBps": float(g[2])}

    # ---- RNG ----
    def philox4x32_fill(self, key, ctr0, nblocks):
        k = np.asarray(key, np.uint32)
        c = np.asarray(ctr0, np.uint32)
        out = np.zeros((nblocks, 4), np.uint32)
        _check(lib().vsmcx_philox4x32_fill(self.h, k.ctypes.data, c.ctypes.data, nblocks,
                                           out.ctypes.data))
        return out

    def threefry4x64_fill(self, key, ctr0, nblocks):
        k = np.asarray(key, np.uint64)
        c = np.asarray(ctr0, np.uint64)
        out = np.zeros((nblocks, 4), np.uint64)
        _check(lib().vsmcx_threefry4x64_fill(self.h, k.ctypes.data, c.ctypes.data, nblocks,
                                             out.ctypes.data))
        return out

    def threefry4x32_fill(self, key, ctr0, nblocks):
        k = np.asarray(key, np.uint32)
        c = np.asarray(ctr0, np.uint32)
        out = np.zeros((nblocks, 4), np.uint32)
        _check(lib().vsmcx_threefry4x32_fill(self.h, k.ctypes.data, c.ctypes.data, nblocks,
                                             out.ctypes.data))
        return out

    # ---- resampling ----
    def resample(self, scheme, weight, engine, N=None, copy_from=True, noalloc=False):
        """Resample<Scheme>(M, N, rng, w, rep) (+ cfrp_trans); engine advanced in place.
        noalloc: through vsmcx_resample_noalloc (M, N <= NOALLOC_MAX; no device
        allocation in the call)."""
        w = np.ascontiguousarray(weight, np.float64)
        M = w.size
        N = M if N is None else N
        rep = np.zeros(M, np.uint64)
        cf = np.zeros(N, np.uint64) if copy_from else None
        fn = lib().vsmcx_resample_noalloc if noalloc else lib().vsmcx_resample
        _check(fn(self.h, scheme, M, N, C.byref(engine), w.ctypes.data, rep.ctypes.data,
                  _ptr(cf)))
        return rep, cf

    def fold_prefix(self, x):
        x = np.ascontiguousarray(x, np.float64)
        out = np.zeros_like(x)
        _check(lib().vsmcx_fold_prefix(self.h, x.size, x.ctypes.data, out.ctypes.data))
        return out


class Timer:
    """n HIP-event marks on a context stream."""

    def __init__(self, ctx, n):
        self.h = C.c_void_p()
        ctx._adopt(self)
        _check(lib().vsmcx_timer_create(ctx.h, n, C.byref(self.h)))

    def record(self, i):
        _check(lib().vsmcx_timer_record(self.h, i))

    def elapsed(self, i, j):
        v = C.c_double()
        _check(lib().vsmcx_timer_elapsed(self.h, i, j, C.byref(v)))
        return v.value

    def close(self):
        if self.h:
            lib().vsmcx_timer_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def cv_llh_const(nu=10.0, scale=0.1):
    """log normaliser of one Student-t(nu, scale) coordinate."""
    return (math.lgamma((nu + 1) / 2) - math.lgamma(nu / 2) - 0.5 * math.log(nu * math.pi)
            - math.log(scale))


def cv_desc(n, horizon, precision=MIXED, scheme=SYSTEMATIC, threshold=0.5, seed=42,
            resample_seed=4242, dt=0.1, sd_pos=0.1, sd_vel=0.5, prior_sd_pos=2.0,
            prior_sd_vel=1.0, obs_scale=0.1, nu=10.0):
    """Descriptor of the constant-velocity filter (dt 0.1, noise sd 0.1 / 0.5,
    Student-t(nu=10, scale 0.1) position observations)."""
    return FilterDesc(n=n, horizon=horizon, precision=precision, scheme=scheme,
                      threshold=threshold, seed=seed, resample_seed=resample_seed, dt=dt,
                      sd_pos=sd_pos, sd_vel=sd_vel, prior_sd_pos=prior_sd_pos,
                      prior_sd_vel=prior_sd_vel, obs_scale=obs_scale, nu=nu,
                      llh_const=cv_llh_const(nu, obs_scale))


class Filter:
    """Device-resident CV particle filter (vsmcx_filter)."""

    def __init__(self, ctx, desc):
        self.ctx = ctx
        self.desc = desc
        self.h = C.c_void_p()
        ctx._adopt(self)
        _check(lib().vsmcx_filter_create(ctx.h, C.byref(desc), C.byref(self.h)))

    def close(self):
        if self.h:
            lib().vsmcx_filter_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_obs(self, obs):
        o = np.ascontiguousarray(obs, np.float64).reshape(-1, 2)
        _check(lib().vsmcx_filter_set_obs(self.h, o.shape[0], o.ctypes.data))

    def run(self, nsteps):
        _check(lib().vsmcx_filter_run(self.h, nsteps))

    def save(self, path):
        """Checkpoint the whole device state to path (fused pipelines)."""
        _check(lib().vsmcx_filter_save(self.h, os.fsencode(path)))

    @classmethod
    def load(cls, ctx, path):
        """A filter resumed from a checkpoint, on ctx."""
        f = cls.__new__(cls)
        f.ctx = ctx
        f.h = C.c_void_p()
        ctx._adopt(f)
        _check(lib().vsmcx_filter_load(ctx.h, os.fsencode(path), C.byref(f.h)))
        f.desc = FilterDesc()
        _check(lib().vsmcx_filter_describe(f.h, C.byref(f.desc)))
        return f

    def finish(self):
        _check(lib().vsmcx_filter_finish(self.h))

    def steps(self):
        v = C.c_uint64()
        _check(lib().vsmcx_filter_steps(self.h, C.byref(v)))
        return v.value

    def history(self, count=None):
        count = self.steps() if count is None else count
        out = np.zeros(count, STEP_DTYPE)
        _check(lib().vsmcx_filter_history(self.h, count, out.ctypes.data))
        return out

    def state(self):
        n = self.desc.n
        st = np.zeros((n, 4))
        lw = np.zeros(n)
        _check(lib().vsmcx_filter_state(self.h, st.ctypes.data, lw.ctypes.data))
        return st, lw

    def engine(self):
        e = ThreefryState()
        _check(lib().vsmcx_filter_engine(self.h, C.byref(e)))
        return e

    def time(self, nsteps):
        el, mv = C.c_double(), C.c_double()
        _check(lib().vsmcx_filter_time(self.h, nsteps, C.byref(el), C.byref(mv)))
        return el.value, mv.value

    def profile(self, enable=True):
        _check(lib().vsmcx_filter_profile(self.h, 1 if enable else 0))

    def move_times(self, max_n=1 << 20):
        out = np.zeros(max_n)
        n = C.c_uint64()
        _check(lib().vsmcx_filter_move_times(self.h, max_n, C.byref(n), out.ctypes.data))
        return out[:n.value].copy()

    def debug_step(self):
        n = self.desc.n
        w = np.zeros(n)
        rep = np.zeros(n, np.uint64)
        cf = np.zeros(n, np.uint64)
        u = np.zeros(n)
        nn = C.c_uint64()
        eng = ThreefryState()
        res = C.c_int()
        _check(lib().vsmcx_filter_debug_step(self.h, w.ctypes.data, rep.ctypes.data,
                                             cf.ctypes.data, u.ctypes.data, C.byref(nn),
                                             C.byref(eng), C.byref(res)))
        return dict(weight=w, rep=rep, copy_from=cf, uniforms=u[:nn.value], nn=nn.value,
                    engine=eng, resampled=bool(res.value))


def cv_simulate(T, seed, dt=0.1, sd_pos=0.1, sd_vel=0.5, prior_sd_pos=2.0, prior_sd_vel=1.0,
                obs_scale=0.1, nu=10.0):
    """Synthetic observations of the CV model: a trajectory from the prior and
    dynamics, position observed with Student-t(nu) noise of scale obs_scale."""
    g = np.random.default_rng(seed)
    x = np.array([prior_sd_pos * g.standard_normal(), prior_sd_pos * g.standard_normal(),
                  prior_sd_vel * g.standard_normal(), prior_sd_vel * g.standard_normal()])
    obs = np.zeros((T, 2))
    for t in range(T):
        if t > 0:
            z = g.standard_normal(4)
            x[0] += sd_pos * z[0] + dt * x[2]
            x[1] += sd_pos * z[1] + dt * x[3]
            x[2] += sd_vel * z[2]
            x[3] += sd_vel * z[3]
        obs[t] = x[:2] + obs_scale * g.standard_t(nu, 2)
    return obs


def cv_simulate_many(horizons, seed, dt=0.1, sd_pos=0.1, sd_vel=0.5, prior_sd_pos=2.0,
                     prior_sd_vel=1.0, obs_scale=0.1, nu=10.0):
    """Observations of many independent CV trajectories (one per horizon),
    simulated together from one seed; returns a list of (T_f, 2) arrays."""
    Ts = np.asarray(horizons, np.int64)
    k, tmax = Ts.size, int(Ts.max())
    g = np.random.default_rng(seed)
    x = np.empty((k, 4))
    x[:, :2] = prior_sd_pos * g.standard_normal((k, 2))
    x[:, 2:] = prior_sd_vel * g.standard_normal((k, 2))
    obs = np.zeros((tmax, k, 2))
    for t in range(tmax):
        if t > 0:
            z = g.standard_normal((k, 4))
            x[:, 0] += sd_pos * z[:, 0] + dt * x[:, 2]
            x[:, 1] += sd_pos * z[:, 1] + dt * x[:, 3]
            x[:, 2] += sd_vel * z[:, 2]
            x[:, 3] += sd_vel * z[:, 3]
        obs[t] = x[:, :2] + obs_scale * g.standard_t(nu, (k, 2))
    return [np.ascontiguousarray(obs[:Ts[f], f]) for f in range(k)]


class Batch:
    """Many independent CV filters advanced by one persistent launch (vsmcx_batch)."""

    def __init__(self, ctx, n, horizon, seed, resample_seed, precision=MIXED, scheme=SYSTEMATIC,
                 threshold=0.5, dt=0.1, sd_pos=0.1, sd_vel=0.5, prior_sd_pos=2.0,
                 prior_sd_vel=1.0, obs_scale=0.1, nu=10.0):
        self.ctx = ctx
        self.n = np.ascontiguousarray(n, np.uint64)
        self.horizon = np.ascontiguousarray(horizon, np.uint64)
        self.seed = np.ascontiguousarray(seed, np.uint64)
        self.resample_seed = np.ascontiguousarray(resample_seed, np.uint64)
        d = BatchDesc(nfilters=self.n.size, n=self.n.ctypes.data,
                      horizon=self.horizon.ctypes.data, seed=self.seed.ctypes.data,
                      resample_seed=self.resample_seed.ctypes.data, precision=precision,
                      scheme=scheme, threshold=threshold, dt=dt, sd_pos=sd_pos, sd_vel=sd_vel,
                      prior_sd_pos=prior_sd_pos, prior_sd_vel=prior_sd_vel, obs_scale=obs_scale,
                      nu=nu, llh_const=cv_llh_const(nu, obs_scale))
        self.h = C.c_void_p()
        ctx._adopt(self)
        _check(lib().vsmcx_batch_create(ctx.h, C.byref(d), C.byref(self.h)))

    def close(self):
        if self.h:
            lib().vsmcx_batch_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_obs(self, f, obs, first=0):
        o = np.ascontiguousarray(obs, np.float64).reshape(-1, 2)
        _check(lib().vsmcx_batch_set_obs(self.h, f, first, o.shape[0], o.ctypes.data))

    def run(self, max_steps=1 << 62):
        ms = C.c_double()
        _check(lib().vsmcx_batch_run(self.h, max_steps, C.byref(ms)))
        return ms.value

    def set_launch(self, workgroups=0, order=0):
        _check(lib().vsmcx_batch_set_launch(self.h, workgroups, order))

    def tune(self, steps=1):
        """Time every launch candidate on the same `steps` steps (state restored
        after each: nothing advances) and keep the fastest."""
        wg, order = C.c_uint32(), C.c_int()
        _check(lib().vsmcx_batch_tune(self.h, steps, C.byref(wg), C.byref(order)))
        return wg.value, order.value

    def tune_report(self):
        """[(strategy, workgroups, order, ms)] of the last tuning."""
        n = C.c_uint64()
        _check(lib().vsmcx_batch_tune_report(self.h, 0, C.byref(n), None, None, None, None))
        k = n.value
        st = np.zeros(k, np.int32)
        wg = np.zeros(k, np.uint32)
        od = np.zeros(k, np.int32)
        ms = np.zeros(k)
        _check(lib().vsmcx_batch_tune_report(self.h, k, C.byref(n), st.ctypes.data,
                                             wg.ctypes.data, od.ctypes.data, ms.ctypes.data))
        return [(int(st[i]), int(wg[i]), int(od[i]), float(ms[i])) for i in range(k)]

    def set_strategy(self, strategy):
        """BATCH_PERSISTENT (one launch, workgroups take filters from a queue) or
        BATCH_TICKS (one step of every active filter per tick, grid launches)."""
        _check(lib().vsmcx_batch_set_strategy(self.h, strategy))

    def strategy(self):
        v = C.c_int()
        _check(lib().vsmcx_batch_get_strategy(self.h, C.byref(v)))
        return v.value

    def steps(self, f):
        v = C.c_uint64()
        _check(lib().vsmcx_batch_steps(self.h, f, C.byref(v)))
        return v.value

    def history(self, f, count=None):
        if count is None:
            count = self.steps(f)
        out = np.zeros(count, STEP_DTYPE)
        _check(lib().vsmcx_batch_history(self.h, f, count, out.ctypes.data))
        return out

    def state(self, f):
        n = int(self.n[f])
        st = np.zeros((n, 4))
        lw = np.zeros(n)
        _check(lib().vsmcx_batch_state(self.h, f, st.ctypes.data, lw.ctypes.data))
        return st, lw

    def particles(self):
        v = C.c_uint64()
        _check(lib().vsmcx_batch_particles(self.h, C.byref(v)))
        return v.value


def lg_matrix(dim=10, seed=0, radius=0.9):
    """A random stable transition matrix: Gaussian entries scaled to spectral
    radius `radius`."""
    g = np.random.default_rng(seed)
    m = g.standard_normal((dim, dim))
    return radius * m / np.max(np.abs(np.linalg.eigvals(m)))


def lg_simulate(A, T, seed, q=0.1, r=0.2, s0=0.25):
    """Observations of x_t = A x_{t-1} + q z, y_t = x_t + r v, x_0 ~ N(0, s0^2 I)."""
    g = np.random.default_rng(seed)
    d = A.shape[0]
    x = s0 * g.standard_normal(d)
    y = np.zeros((T, d))
    for t in range(T):
        if t > 0:
            x = A @ x + q * g.standard_normal(d)
        y[t] = x + r * g.standard_normal(d)
    return y


def lg_simulate_many(A, horizons, seed, q=0.1, r=0.2, s0=0.25):
    """lg_simulate for many independent trajectories at once (one per horizon)."""
    Ts = np.asarray(horizons, np.int64)
    k, tmax, d = Ts.size, int(Ts.max()), A.shape[0]
    g = np.random.default_rng(seed)
    x = s0 * g.standard_normal((k, d))
    y = np.zeros((tmax, k, d))
    for t in range(tmax):
        if t > 0:
            x = x @ A.T + q * g.standard_normal((k, d))
        y[t] = x + r * g.standard_normal((k, d))
    return [np.ascontiguousarray(y[:Ts[f], f]) for f in range(k)]


class Serve:
    """The serving loop (vsmcx_serve): jobs admitted under an HBM budget and
    stepped with continuous batching as their observations arrive."""

    def __init__(self, ctx, hbm_budget, max_jobs, max_horizon, lg_a, lg_q=0.1, lg_r=0.2,
                 lg_s0=0.25, precision=MIXED, scheme=SYSTEMATIC, threshold=0.5, dt=0.1,
                 sd_pos=0.1, sd_vel=0.5, prior_sd_pos=2.0, prior_sd_vel=1.0, obs_scale=0.1,
                 nu=10.0):
        self.ctx = ctx
        self.lg_a = np.ascontiguousarray(lg_a, np.float64)
        self.lg_dim = self.lg_a.shape[0]
        d = ServeDesc(hbm_budget=int(hbm_budget), max_jobs=max_jobs, max_horizon=max_horizon,
                      precision=precision, scheme=scheme, threshold=threshold, dt=dt,
                      sd_pos=sd_pos, sd_vel=sd_vel, prior_sd_pos=prior_sd_pos,
                      prior_sd_vel=prior_sd_vel, obs_scale=obs_scale, nu=nu,
                      llh_const=cv_llh_const(nu, obs_scale), lg_dim=self.lg_dim,
                      lg_a=self.lg_a.ctypes.data, lg_q=lg_q, lg_r=lg_r, lg_s0=lg_s0)
        self.h = C.c_void_p()
        ctx._adopt(self)
        _check(lib().vsmcx_serve_create(ctx.h, C.byref(d), C.byref(self.h)))

    def close(self):
        if self.h:
            lib().vsmcx_serve_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def submit(self, model, n, horizon, seed, resample_seed):
        j = JobDesc(model=model, n=n, horizon=horizon, seed=seed, resample_seed=resample_seed)
        out = C.c_uint64()
        _check(lib().vsmcx_serve_submit(self.h, C.byref(j), C.byref(out)))
        return out.value

    def observe(self, job, y):
        y = np.ascontiguousarray(y, np.float64)
        _check(lib().vsmcx_serve_observe(self.h, job, y.ctypes.data))

    def observe_many(self, jobs, ys):
        """One observation each: ys[k] (a float64 array) for jobs[k]."""
        ids = np.ascontiguousarray(jobs, np.uint64)
        ptrs = (C.c_void_p * len(ys))(*[y.ctypes.data for y in ys])
        _check(lib().vsmcx_serve_observe_many(self.h, ids.size, ids.ctypes.data, ptrs))

    def ready(self, max_ids=1 << 16):
        ids = np.zeros(max_ids, np.uint64)
        n = C.c_uint64()
        _check(lib().vsmcx_serve_ready(self.h, max_ids, ids.ctypes.data, C.byref(n)))
        return ids[:min(n.value, max_ids)]

    def poll(self, wait=False):
        n = C.c_uint64()
        _check(lib().vsmcx_serve_poll(self.h, 1 if wait else 0, C.byref(n)))
        return n.value

    def job(self, job):
        st, steps = C.c_int(), C.c_uint64()
        _check(lib().vsmcx_serve_job(self.h, job, C.byref(st), C.byref(steps)))
        return st.value, steps.value

    def history(self, job, count):
        out = np.zeros(count, STEP_DTYPE)
        _check(lib().vsmcx_serve_history(self.h, job, count, out.ctypes.data))
        return out

    def forget(self, job):
        _check(lib().vsmcx_serve_forget(self.h, job))

    def stats(self):
        s = ServeSummary()
        _check(lib().vsmcx_serve_stats(self.h, C.byref(s)))
        return {k: getattr(s, k) for k, _ in ServeSummary._fields_}


class State:
    """Device-resident particle values (vsmcx_state): n rows of dim values in
    HBM, FP32 for MIXED and FP64 otherwise; the storage of vSMC's value type
    with StateMatrix::copy as a device gather."""

    def __init__(self, ctx, n, dim, precision=MIXED):
        self.ctx = ctx
        self.n, self.dim = n, dim
        self.h = C.c_void_p()
        ctx._adopt(self)
        _check(lib().vsmcx_state_create(ctx.h, n, dim, precision, C.byref(self.h)))

    def close(self):
        if self.h:
            lib().vsmcx_state_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        self.close()

    def write(self, rows):
        r = np.ascontiguousarray(rows, np.float64).reshape(self.n, self.dim)
        _check(lib().vsmcx_state_write(self.h, r.ctypes.data))

    def read(self):
        out = np.zeros((self.n, self.dim))
        _check(lib().vsmcx_state_read(self.h, out.ctypes.data))
        return out

    def copy(self, copy_from):
        cf = np.ascontiguousarray(copy_from, np.uint64)
        _check(lib().vsmcx_state_copy(self.h, cf.size, cf.ctypes.data))

    def cv_move(self, desc, iteration, obs_xy):
        """The pf model's move of every row (the prior draw at iteration 0)
        with the constants and Philox key of desc; returns the increments."""
        o = np.ascontiguousarray(obs_xy, np.float64).reshape(2)
        inc = np.zeros(self.n)
        _check(lib().vsmcx_state_cv_move(self.h, C.byref(desc), iteration, o.ctypes.data,
                                         inc.ctypes.data))
        return inc


class Weights:
    """Device weight set (vsmcx_weights): vSMC's WeightSet in HBM, updated from
    device arrays of increments (a user kernel's output: pass its device
    address), with Particle::resample on the device.  Increments live in
    device memory, so the set / add calls take device addresses (buffer()
    gives one owned by the set)."""

    def __init__(self, ctx, n, scheme=SYSTEMATIC, threshold=0.5, resample_seed=4242,
                 max_steps=1000):
        self.ctx = ctx
        self.n = n
        self.h = C.c_void_p()
        ctx._adopt(self)
        _check(lib().vsmcx_weights_create(ctx.h, n, scheme, threshold, resample_seed, max_steps,
                                          C.byref(self.h)))

    def close(self):
        if self.h:
            lib().vsmcx_weights_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        self.close()

    def buffer(self):
        p = C.c_void_p()
        _check(lib().vsmcx_weights_buffer(self.h, C.byref(p)))
        return p.value

    def set_log(self, inc_dev):
        _check(lib().vsmcx_weights_set_log(self.h, inc_dev))

    def add_log(self, inc_dev):
        _check(lib().vsmcx_weights_add_log(self.h, inc_dev))

    def resample(self, state=None):
        _check(lib().vsmcx_weights_resample(self.h, state.h if state is not None else None))

    def steps(self):
        t = C.c_uint64()
        _check(lib().vsmcx_weights_steps(self.h, C.byref(t)))
        return t.value

    def history(self, count=None):
        count = self.steps() if count is None else count
        out = np.zeros(count, STEP_DTYPE)
        _check(lib().vsmcx_weights_history(self.h, count, out.ctypes.data))
        return out

    def read(self):
        w = np.zeros(self.n)
        _check(lib().vsmcx_weights_read(self.h, w.ctypes.data))
        return w


class Gmm:
    """Gaussian-mixture SMC sampler with adaptive (CESS) tempering (vsmcx_gmm):
    the reference's gmm_smc under the "cess" schedule.  initialize() draws from
    the prior; iterate() runs iterations until alpha reaches 1."""

    def __init__(self, ctx, data, n, components=4, sweeps=5, precision=MIXED, scheme=STRATIFIED,
                 threshold=0.5, cess_drop=0.1, seed=1, resample_seed=2, max_steps=2000):
        self.ctx = ctx
        self.K = components
        self.n = n
        self.data = np.ascontiguousarray(data, np.float64)
        d = GmmDesc(n=n, components=components, sweeps=sweeps, precision=precision,
                    scheme=scheme, threshold=threshold, cess_drop=cess_drop, seed=seed,
                    resample_seed=resample_seed, max_steps=max_steps)
        self.h = C.c_void_p()
        ctx._adopt(self)
        _check(lib().vsmcx_gmm_create(ctx.h, C.byref(d), self.data.ctypes.data, self.data.size,
                                      C.byref(self.h)))

    def close(self):
        if self.h:
            lib().vsmcx_gmm_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def initialize(self):
        _check(lib().vsmcx_gmm_initialize(self.h))

    def iterate(self, max_iter=1 << 62):
        done, alpha = C.c_uint64(), C.c_double()
        _check(lib().vsmcx_gmm_iterate(self.h, max_iter, C.byref(done), C.byref(alpha)))
        return done.value, alpha.value

    def rows(self):
        r = C.c_uint64()
        _check(lib().vsmcx_gmm_rows(self.h, C.byref(r)))
        return r.value

    def times(self):
        """Device ms so far: (tempering + reweight, resampling, MH sweeps)."""
        a, b, c = C.c_double(), C.c_double(), C.c_double()
        _check(lib().vsmcx_gmm_times(self.h, C.byref(a), C.byref(b), C.byref(c)))
        return a.value, b.value, c.value

    def history(self):
        r = self.rows()
        out = np.zeros(r, GMM_DTYPE)
        _check(lib().vsmcx_gmm_history(self.h, r, out.ctypes.data))
        return out

    def state(self):
        p = np.zeros((self.n, 3 * self.K))
        llh = np.zeros(self.n)
        lw = np.zeros(self.n)
        _check(lib().vsmcx_gmm_state(self.h, p.ctypes.data, llh.ctypes.data, lw.ctypes.data))
        return p, llh, lw


def gmm_simulate(n, seed, means=(-3.0, 0.0, 2.0, 5.0), sd_range=(0.5, 1.5)):
    """Data of the GMM workload: n points from an equal-weight mixture with the
    given means and component sds drawn uniformly in sd_range."""
    rng = np.random.default_rng(seed)
    means = np.asarray(means, np.float64)
    sds = rng.uniform(sd_range[0], sd_range[1], means.size)
    comp = rng.integers(0, means.size, n)
    return means[comp] + sds[comp] * rng.standard_normal(n)
