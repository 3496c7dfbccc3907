"""Write the golden vectors in tests/golden/ from the compiled vSMC headers.

Run in a container that has the vSMC tree (oracle/Makefile builds
oracle/_ref/libvsmcref.so from it):  python tests/golden/make_golden.py

Every output is data produced by vSMC's own templates (engines,
uniform_real_distribution, U01Sequence*, Resample<Scheme>, cfrp_trans,
WeightSet, a Particle<T>-driven CV filter, WeightSet::cess and the gmm
example's alpha_ess<T, cess01<T>> tempering search).  Inputs are seeded here so the
fixtures are reproducible.
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "..", "..", "oracle"))
from oracle_ctypes import RefHarness, cv_param  # noqa: E402


def main():
    r = RefHarness()
    rng = np.random.default_rng(20241203)
    out = {}

    # --- engine blocks on random counters/keys ---
    ph_c = rng.integers(0, 2**32, (64, 4), dtype=np.uint64).astype(np.uint32)
    ph_k = rng.integers(0, 2**32, (64, 2), dtype=np.uint64).astype(np.uint32)
    out["philox_ctr"], out["philox_key"] = ph_c, ph_k
    out["philox_out"] = np.stack([r.philox4x32_block(c, k) for c, k in zip(ph_c, ph_k)])
    tf_c = rng.integers(0, 2**64, (64, 4), dtype=np.uint64)
    tf_k = rng.integers(0, 2**64, (64, 4), dtype=np.uint64)
    out["tf64_ctr"], out["tf64_key"] = tf_c, tf_k
    out["tf64_out"] = np.stack([r.threefry4x64_block(c, k) for c, k in zip(tf_c, tf_k)])
    t3_c = rng.integers(0, 2**32, (64, 4), dtype=np.uint64).astype(np.uint32)
    t3_k = rng.integers(0, 2**32, (64, 4), dtype=np.uint64).astype(np.uint32)
    out["tf32_ctr"], out["tf32_key"] = t3_c, t3_k
    out["tf32_out"] = np.stack([r.threefry4x32_block(c, k) for c, k in zip(t3_c, t3_k)])

    # --- engine streams: seed, discard, counter carry ---
    skips = np.array([0, 1, 3, 4, 5, 8, 9, 1001], np.uint64)
    out["stream_skips"] = skips
    out["ph32_streams"] = np.stack([r.ph32_stream(2024, 23, int(s)) for s in skips])
    out["tf64_streams"] = np.stack([r.tf64_stream(2024, 23, int(s)) for s in skips])
    out["tf32_stream"] = r.tf32_stream(2024, 23)
    out["ph32_carry"] = r.ph32_stream(5, 23, 0, [2**32 - 1, 2**32 - 1, 7, 0])
    out["tf64_carry"] = r.tf64_stream(5, 23, 0, [2**64 - 2, 2**64 - 1, 5, 0])

    # --- std::uniform_real_distribution<double> ---
    out["canonical_tf64"] = r.canonical_tf64(101, 256)
    out["canonical_ph32"] = r.canonical_ph32(101, 256)

    # --- U01 sequences (sorted / stratified / systematic) + next draws ---
    for kind, name in enumerate(["sorted", "stratified", "systematic"]):
        u, nxt = r.u01seq(kind, 1000, 77)
        out[f"u01seq_{name}"] = u
        out[f"u01seq_{name}_next"] = nxt

    # --- resamplers: weights, replication, copy_from, next engine draws ---
    cases = []
    for N in (1, 2, 5, 17, 100, 1000):
        for variant in range(3):
            w = rng.random(N) ** (1 + 3 * variant)
            if variant == 2 and N > 4:
                w[rng.integers(0, N, N // 2)] = 0.0
            w = w / w.sum()
            cases.append(w)
    out["resample_ncases"] = np.array(len(cases))
    for ci, w in enumerate(cases):
        out[f"resample_w_{ci}"] = w
        for scheme in range(6):
            rep, cf, nxt = r.resample(scheme, w, 1000 + ci)
            out[f"resample_rep_{ci}_{scheme}"] = rep
            out[f"resample_cf_{ci}_{scheme}"] = cf
            out[f"resample_next_{ci}_{scheme}"] = nxt

    # --- WeightSet normalisation ---
    lw0 = rng.normal(size=777) * 20
    inc = rng.normal(size=777) * 4
    lw, w, ess = r.weightset(lw0)
    out["ws_lw0"], out["ws_inc"] = lw0, inc
    out["ws_set_lw"], out["ws_set_w"], out["ws_set_ess"] = lw, w, np.array(ess)
    lw, w, ess = r.weightset(lw0, inc)
    out["ws_add_lw"], out["ws_add_w"], out["ws_add_ess"] = lw, w, np.array(ess)

    # --- CV particle filter through vsmc::Particle<T> (N=1000, T=40) ---
    obs = np.cumsum(rng.normal(size=(40, 2)) * 0.1, axis=0) + np.array([1.0, -0.5])
    out["cv_obs"] = obs
    for scheme in range(6):
        st, lwf, rec = r.cv_run(1000, obs, 42, 4242, scheme, 0.5)
        for f in rec.dtype.names:
            out[f"cv_{scheme}_{f}"] = rec[f]
        out[f"cv_{scheme}_state"] = st
        out[f"cv_{scheme}_lw"] = lwf

    # --- GMM tempering: WeightSet::cess and alpha_ess<T, cess01<T>>::alpha_iter ---
    # (a generator of its own, so the vectors above do not move)
    trng = np.random.default_rng(20261204)
    tcases = []
    for N in (1, 7, 1000, 20000):
        for lw_kind in ("equal", "spread", "heavy"):
            for llh_sd in (0.5, 30.0, 3000.0):
                tcases.append((N, lw_kind, llh_sd))
    tcases += [(1000, "spread", "underflow"), (20000, "heavy", "underflow")]
    out["temper_ncases"] = np.array(len(tcases))
    for ci, (N, lw_kind, llh_sd) in enumerate(tcases):
        lw = (np.zeros(N) if lw_kind == "equal" else
              trng.normal(size=N) * (2.0 if lw_kind == "spread" else 40.0))
        if llh_sd == "underflow":  # every multiplier underflows once a_diff > 0.0075
            llh = -1e5 - trng.random(N)
        else:
            llh = -500.0 + trng.normal(size=N) * llh_sd
        alpha = float(trng.choice([0.0, 0.3, 0.95, 1 - 1e-7]))
        drop = float(trng.choice([0.1, 0.5, 1e-9]))
        a, ai, calls, w = r.alpha_ess(lw, llh, alpha, drop)
        a_diff = float(trng.choice([1e-4, 0.01, 0.5]))
        cess, _ = r.cess(lw, a_diff * llh, use_log=True)
        out[f"temper_lw_{ci}"], out[f"temper_llh_{ci}"] = lw, llh
        out[f"temper_w_{ci}"] = w
        out[f"temper_in_{ci}"] = np.array([alpha, drop, a_diff])
        out[f"temper_out_{ci}"] = np.array([a, ai, float(calls), cess])

    np.savez_compressed(os.path.join(HERE, "ref_vectors.npz"), **out)
    meta = {
        "source": "vSMC headers compiled by oracle/Makefile (oracle/_ref/libvsmcref.so)",
        "compiler": "g++ 11.4 -std=c++11 -O2 -ffp-contract=off, libstdc++ 11",
        "cv_param": {f: getattr(cv_param(), f) for f, _ in cv_param()._fields_},
        "cv_run": {"N": 1000, "T": 40, "seed": 42, "resample_seed": 4242, "threshold": 0.5},
    }
    with open(os.path.join(HERE, "ref_vectors.json"), "w") as f:
        json.dump(meta, f, indent=1)
    print("wrote", len(out), "arrays")


if __name__ == "__main__":
    main()
