"""Benchmark of the MPS contraction hot path on one MI355X.

Step = one two-site H_eff matvec  H_eff * theta = L W_i W_{i+1} R theta  on the isolated-
kernel workload: Heisenberg XXX MPO (D = 5, d = 2, real), L and R of shape (chi, 5, chi)
with N(0,1) entries symmetrised per MPO index, theta (chi, 2, 2, chi) N(0,1) normalised,
chi = 2048, seed 0, all inputs resident in HBM before timing. `value` is the FP32-MFMA
matvec throughput in TFLOP/s (algorithmic flops 2 (2 chi^3 D d^2 + 2 chi^2 d^3 D^2) per
matvec); the FP64 and split-BF16 runs, their errors against the FP64 result, the
truncated SVD of the 2chi x 2chi theta matrix and the FP64 CPU port (cpu_baseline) are
reported beside it.

  python bench.py [--steps K] [--warmup W] [--chi 2048] [--no-svd] [--no-cpu]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT / "tensornetworks.jl_amd"))
sys.path.insert(0, str(ROOT))

METRIC = "H_eff matvec TFLOP/s + DMRG sweep time at χ=2048, FP32-MFMA vs FP64, |ΔE|/N<1e-10"
PEAK = {"f32": 157.3, "f64": 78.6, "bf16x3": 2500.0}  # dense MFMA TFLOP/s, MI355X spec
HBM_PEAK = 8000.0  # GB/s


def matvec_flops(chi, D=5, d=2):
    return 2.0 * (2 * chi ** 3 * D * d ** 2 + 2 * chi ** 2 * d ** 3 * D ** 2)


def heisenberg_bulk_w():
    from tnhip.models import OpList, MPO, spinhalf

    H = OpList(4)
    for i in range(1, 4):
        H.add(["s+", "s-"], [i, i + 1], 0.5)
        H.add(["s-", "s+"], [i, i + 1], 0.5)
        H.add(["z", "z"], [i, i + 1], 0.25)
    return MPO(spinhalf(), H, compress=False).tensors[1]  # automaton bulk tensor (5, 2, 2, 5), real


def make_inputs(chi, D, seed, torch):
    g = torch.Generator(device="cuda").manual_seed(seed)
    dev = torch.device("cuda")

    def env():
        X = torch.randn((chi, D, chi), generator=g, device=dev, dtype=torch.float64)
        return (0.5 * (X + X.permute(2, 1, 0))).contiguous()

    L, R = env(), env()
    th = torch.randn((chi * 2 * 2 * chi,), generator=g, device=dev, dtype=torch.float64)
    th /= th.norm()
    return L, R, th


def measured_traffic(chi):
    """HBM bytes per launch of the FP32 GEMM (mean of the two stages), from the
    copy-calibrated FETCH_SIZE + WRITE_SIZE counter runs summarised in
    profiles/heff_matvec_chi2048_hbm_traffic.md (valid for the chi = 2048 workload)."""
    if chi != 2048:
        return None
    p = ROOT / "profiles" / "heff_matvec_chi2048_hbm_traffic.md"
    try:
        for line in p.read_text().splitlines():
            if line.startswith("traffic_bytes_per_launch:"):
                return float(line.split(":", 1)[1])
    except OSError:
        pass
    return None


def cpu_baseline(budget_s=20.0):
    """FP64 CPU port (oracle, optimal contraction order, numpy/OpenBLAS on all host
    cores) on a bounded sample of the same workload: matvecs at chi = 512 and 1024."""
    from oracle.projmps import product_optimal

    W = heisenberg_bulk_w()
    rng = np.random.default_rng(0)
    cores = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    best = None
    for chi in (512, 1024):
        X = rng.standard_normal((chi, 5, chi))
        L = 0.5 * (X + X.transpose(2, 1, 0))
        R = L.copy()
        th = rng.standard_normal((chi, 2, 2, chi))
        product_optimal(L, W, W, R, th)  # warm-up
        reps, t0 = 0, time.perf_counter()
        while True:
            product_optimal(L, W, W, R, th)
            reps += 1
            el = time.perf_counter() - t0
            if el > budget_s / 2 or reps >= 40:
                break
        best = {"value": matvec_flops(chi) * reps / el / 1e12, "unit": "TFLOP/s", "cores": cores,
                "kind": "port",
                "sample": f"FP64 numpy/OpenBLAS H_eff matvec (oracle/projmps.py product_optimal), "
                          f"chi={chi}, D=5, d=2, {reps} reps in {el:.1f} s"}
    return best


def heisenberg_oplist(N):
    """H = sum S.S with Pauli spin operators: (s+ s- + s- s+) / 2 + z z / 4 (real MPO, D = 5)."""
    from tnhip.models import OpList

    H = OpList(N)
    for i in range(1, N):
        H.add(["s+", "s-"], [i, i + 1], 0.5)
        H.add(["s-", "s+"], [i, i + 1], 0.5)
        H.add(["z", "z"], [i, i + 1], 0.25)
    return H


def tfim_oplist(N, g=1.0):
    """H = -sum z z - g sum x (MPO D = 3)."""
    from tnhip.models import OpList

    H = OpList(N)
    for i in range(1, N):
        H.add(["z", "z"], [i, i + 1], -1.0)
    for i in range(1, N + 1):
        H.add("x", i, -g)
    return H


def dmrg_heisenberg_n32(cpu=True):
    """Two-site DMRG, Heisenberg N = 32, random chi = 8 start (seed 0), chi_max 64, cutoff
    1e-10, 8 passes, Krylov dimension 4, FP64 on the GPU; the FP64 CPU port on the same
    initial tensors beside it."""
    from tnhip import MPO, randomMPS, spinhalf
    from tnhip.mps import dmrg

    N = 32
    H = heisenberg_oplist(N)
    psi = randomMPS(2, N, 8, rng=np.random.default_rng(0))
    start = [t.copy() for t in psi.tensors], psi.center
    M = MPO(spinhalf(), H)
    hist = []
    t0 = time.perf_counter()
    psi, e = dmrg(psi, M, maxsweeps=8, cutoff=1e-10, maxdim=64, krylovdim=4, verbose=False, history=hist)
    gpu_s = time.perf_counter() - t0
    out = {"gpu_seconds": gpu_s, "energy": e, "energy_per_site": e / N, "sweeps": len(hist),
           "maxbonddim": hist[-1]["maxbonddim"]}
    if cpu:
        from oracle.dmrg import dmrg as dmrg_cpu
        from oracle.mps import MPS

        p = MPS([t.copy() for t in start[0]], start[1])
        t0 = time.perf_counter()
        _, ec = dmrg_cpu(p, [np.asarray(w) for w in M.tensors], maxsweeps=8, cutoff=1e-10, maxdim=64,
                         krylovdim=4)
        out["cpu_port_seconds"] = time.perf_counter() - t0
        out["cpu_port_energy"] = ec
        out["abs_dE_per_site_vs_cpu"] = abs(e - ec) / N
    return out


def dmrg_tfim_n100(chi, sweeps, refine_tol=None, precisions=("f64", "f32"), cpu=False):
    """Critical TFIM N = 100 (g = 1), random chi = 16 start (seed 0), cutoff 1e-12: an
    all-FP64 run and lower-precision runs (FP32-MFMA, split-BF16x3) with FP64 refinement
    passes, the same number of passes each. The lower precision switches to FP64 when the
    relative energy change of a pass falls below refine_tol (None: the dmrg default for the
    precision; 0: only for the last two passes). Reports per-pass times and energies and
    the energy errors per site; with cpu, also the FP64 CPU port from the same initial
    tensors and its energy difference to the FP64 GPU run."""
    from tnhip import MPO, randomMPS, spinhalf
    from tnhip.mps import dmrg

    N = 100
    e_exact = 1.0 - 1.0 / np.sin(np.pi / (4 * N + 2))  # open critical TFIM closed form
    M = MPO(spinhalf(), tfim_oplist(N))
    res = {}
    start = None
    for prec in precisions:
        psi = randomMPS(2, N, 16, rng=np.random.default_rng(0))
        if start is None:
            start = [t.copy() for t in psi.tensors], psi.center
        hist = []
        t0 = time.perf_counter()
        psi, e = dmrg(psi, M, maxsweeps=sweeps, minsweeps=sweeps, tol=0.0, cutoff=1e-12, maxdim=chi,
                      verbose=False, precision=prec, refine_sweeps=2, refine_tol=refine_tol, history=hist)
        res[prec] = {"seconds": time.perf_counter() - t0, "energy": e,
                     "abs_dE_per_site_vs_exact": abs(e - e_exact) / N,
                     "pass_seconds": [round(h["seconds"], 4) for h in hist],
                     "pass_energy": [h["energy"] for h in hist],
                     "pass_precision": [h["precision"] for h in hist],
                     "maxbonddim": hist[-1]["maxbonddim"]}
    for prec in precisions:
        if prec == "f64":
            continue
        res[f"abs_dE_per_site_{prec}_vs_f64"] = abs(res[prec]["energy"] - res["f64"]["energy"]) / N
        lo = [s for s, p in zip(res[prec]["pass_seconds"], res[prec]["pass_precision"]) if p == prec]
        hi = res["f64"]["pass_seconds"][: len(lo)]
        if lo:
            res[f"{prec}_pass_speedup_vs_f64"] = sum(hi) / sum(lo)
        res[f"{prec}_run_speedup_vs_f64"] = res["f64"]["seconds"] / res[prec]["seconds"]
    if cpu:
        from oracle.dmrg import dmrg as dmrg_cpu
        from oracle.mps import MPS

        p = MPS([t.copy() for t in start[0]], start[1])
        t0 = time.perf_counter()
        _, ec = dmrg_cpu(p, [np.asarray(w) for w in M.tensors], minsweeps=sweeps, maxsweeps=sweeps, tol=0.0,
                         cutoff=1e-12, maxdim=chi)
        res["cpu_port"] = {"seconds": time.perf_counter() - t0, "energy": ec,
                           "abs_dE_per_site_vs_exact": abs(ec - e_exact) / N}
        if "f64" in res:
            res["abs_dE_per_site_f64_vs_cpu"] = abs(res["f64"]["energy"] - ec) / N
            res["f64_speedup_vs_cpu_port"] = res["cpu_port"]["seconds"] / res["f64"]["seconds"]
    res["config"] = {"N": N, "g": 1.0, "chi_max": chi, "passes": sweeps, "cutoff": 1e-12, "krylovdim": 3,
                     "refine_tol": refine_tol, "E0_exact": e_exact}
    return res


def dmrg_tfim_n100_chi2048(chi=2048, precisions=("f64", "f32", "bf16x3"), warm_passes=12, profile=False,
                           timed=None):
    """Half-sweeps of two-site DMRG on the critical TFIM, N = 100, at bond dimension chi:
    a random chi = 16 start (seed 0) is converged at chi_max 64 (cutoff 1e-12, warm_passes
    passes), then every pass keeps mindim = maxdim = chi with cutoff 0, so each bond not
    capped by the chain ends (2^min(i, N - i) < chi) runs at chi = 2048. One FP64 pass grows
    the bonds; then one timed pass per precision (matvecs and environment updates in it;
    Krylov vectors, Gram eigensolve and reductions FP64), each followed by one FP64 pass whose
    energy is compared with two all-FP64 passes from the same state (on a copy of it, so that
    the comparison does not count the sweeps' own convergence). Reports seconds per
    half-sweep and |dE|/N against the closed form and between the precisions."""
    from tnhip import MPO, Context, randomMPS, spinhalf
    from tnhip.models import GMPS
    from tnhip.mps import DMRGState, dmrg

    N = 100
    e_exact = 1.0 - 1.0 / np.sin(np.pi / (4 * N + 2))
    M = MPO(spinhalf(), tfim_oplist(N))
    psi = randomMPS(2, N, 16, rng=np.random.default_rng(0))
    t0 = time.perf_counter()
    psi, e_small = dmrg(psi, M, maxsweeps=warm_passes, minsweeps=warm_passes, tol=0.0, cutoff=1e-12, maxdim=64,
                        verbose=False)
    setup_s = time.perf_counter() - t0
    ctx = Context.default()
    st = DMRGState(psi, [M])
    state = {"direction": False}

    def one_pass(prec):
        if profile:
            ctx.profile_reset()
            ctx.profile(True)
        ctx.synchronize()
        t = time.perf_counter()
        e, D, disc = st.sweep(state["direction"], 3, 2, 1e-14, 0.0, chi, chi, prec)
        el = time.perf_counter() - t
        state["direction"] = not state["direction"]
        r = {"seconds": el, "energy": e, "abs_dE_per_site_vs_exact": abs(e - e_exact) / N, "maxbonddim": D,
             "precision": prec, "direction": "left" if not state["direction"] else "right"}
        if profile:
            r["kernel_ms"] = {f: round(ctx.profile_read(f)["ms"], 2)
                              for f in ("gemm_f64", "gemm_f32", "gemm_bf16x3", "mix", "jacobi", "jacobi_warm",
                                        "dmrg_env", "dmrg_lanczos", "dmrg_split")}
            for f in ("jacobi", "jacobi_warm"):
                j = ctx.profile_read(f)
                r[f + "_calls_sweeps"] = [int(j["launches"]), int(j["flops"])]
            ctx.profile(False)
        print(f"chi={chi} pass {prec}: {el:.2f} s, E={e:.12f}, chi_max={D}", file=sys.stderr, flush=True)
        return r

    out = {"config": {"N": N, "chi": chi, "bonds_at_chi": sum(1 for b in range(1, N) if 2 ** min(b, N - b) >= chi),
                      "krylovdim": 3, "maxiter": 2, "warm_passes_chi64": warm_passes, "E0_exact": e_exact},
           "setup_seconds_chi64": setup_s, "energy_chi64": e_small}
    # a split keeps at most twice the bond dimension it started from: grow to chi first
    out["grow_f64"] = []
    while len(out["grow_f64"]) < 8:
        r = one_pass("f64")
        out["grow_f64"].append(r)
        if r["maxbonddim"] >= chi:
            break
    # two more passes (one per direction): every split of the timed passes then starts its
    # Jacobi from the eigenvectors of the same bond's previous split in that direction
    out["grow_f64"].append(one_pass("f64"))
    out["grow_f64"].append(one_pass("f64"))
    ref = one_pass("f64")
    out["f64"] = ref
    if timed == "f64":  # diagnostics: the FP64 pass only
        st.close()
        return out

    def fp64_twin(start):
        # the all-FP64 path from the same state: two FP64 passes (right, then left) on a copy
        # of it (a new state starts with its centre at site 1, as dmrg does)
        twin = DMRGState(start, [M])
        es = []
        for dr in (False, True):
            e, _, _ = twin.sweep(dr, 3, 2, 1e-14, 0.0, chi, chi, "f64")
            es.append(e)
        twin.close()
        return es[-1]

    for prec in precisions:
        if prec == "f64":
            continue
        if state["direction"]:  # bring the centre to site 1 first, where the copy starts
            out.setdefault("align_f64", []).append(one_pass("f64"))
        start = GMPS(1, 2, st.tensors(), st.center())
        out[prec] = one_pass(prec)
        out[prec + "_then_f64"] = r2 = one_pass("f64")
        e_twin = fp64_twin(start)
        del start
        out[f"{prec}_then_f64_vs_two_f64_passes"] = {"energy_two_f64_passes": e_twin,
                                                      "abs_dE_per_site": abs(r2["energy"] - e_twin) / N}
        out[f"abs_dE_per_site_{prec}_refined_vs_f64"] = abs(r2["energy"] - e_twin) / N
        out[f"{prec}_half_sweep_speedup_vs_f64"] = ref["seconds"] / out[prec]["seconds"]
    st.close()
    return out


def gradient_tfim_n64(chi=512, block=8, cpu=True):
    """E(J, h) = <psi|H(J, h)|psi> and its 127 coupling derivatives for the inhomogeneous TFIM,
    N = 64, J_i ~ U(0.8, 1.2), h_i ~ U(0.5, 1.5) (seed 1), at its DMRG ground state held at
    bond dimension chi (converged at chi_max 128, cutoff 1e-14, then six passes at
    mindim = maxdim = chi), environments checkpointed every `block` sites. Accuracy: central
    differences of the device energy (step 1e-5) on six couplings, and Hellmann-Feynman
    against the free-fermion (BdG) ground energy's central differences for every coupling.
    CPU baseline: the same 127 <O_t> = dE/dc_t from the oracle's environment contraction
    (numpy/OpenBLAS on the host cores) on the same tensors."""
    from tnhip import MPO, randomMPS, spinhalf
    from tnhip.grad import coupling_gradient
    from tnhip.models import OpList
    from tnhip.mps import dmrg
    from tnhip.tebd import DeviceMPS

    N = 64
    r = np.random.default_rng(1)
    J, h = r.uniform(0.8, 1.2, N - 1), r.uniform(0.5, 1.5, N)
    H = OpList(N)
    for i in range(1, N):
        H.add(["z", "z"], [i, i + 1], -J[i - 1])
    for i in range(1, N + 1):
        H.add("x", i, -h[i - 1])
    st = spinhalf()
    t0 = time.perf_counter()
    psi = randomMPS(2, N, 16, rng=np.random.default_rng(2))
    kw = dict(krylovdim=10, kryloviter=6, verbose=False)
    psi, _ = dmrg(psi, MPO(st, H), maxsweeps=12, cutoff=1e-14, maxdim=128, **kw)
    psi, E_dmrg = dmrg(psi, MPO(st, H), maxsweeps=6, cutoff=0.0, maxdim=chi, mindim=chi, **kw)
    prep_s = time.perf_counter() - t0
    d = DeviceMPS.from_gmps(psi)
    coupling_gradient(st, d, H, block=block)  # warm-up
    d.ctx.synchronize()
    t0 = time.perf_counter()
    E, g = coupling_gradient(st, d, H, block=block)
    el = time.perf_counter() - t0
    step = 1e-5
    fd_err = 0.0
    for t in (0, 31, N - 2, N - 1, N - 1 + 40, 2 * N - 2):
        Hp, Hm = H * 1.0, H * 1.0
        Hp.coeffs[t] += step
        Hm.coeffs[t] -= step
        Ep, _ = coupling_gradient(st, d, Hp, block=block)
        Em, _ = coupling_gradient(st, d, Hm, block=block)
        fd_err = max(fd_err, abs((Ep - Em).real / (2 * step) - g[t].real))
    out = {"seconds": el, "energy": float(E.real), "n_couplings": len(H.coeffs), "chi": psi.maxbonddim(),
           "block": block, "state_prep_seconds": prep_s, "max_abs_fd_error": fd_err}

    def bdg(Jv, hv):  # ground energy of the open TFIM from its BdG spectrum
        A = np.diag(2 * hv) - np.diag(Jv, 1) - np.diag(Jv, -1)
        B = np.diag(Jv, 1) - np.diag(Jv, -1)
        w = np.linalg.eigvalsh(np.block([[A, B], [-B, -A]]))
        return -0.5 * np.sum(w[w > 0])

    e0 = bdg(J, h)
    out["abs_dE_per_site_vs_bdg"] = abs(float(E.real) - e0) / N
    eps = 1e-5
    errs = []
    for i in range(N - 1):
        Jp, Jm = J.copy(), J.copy()
        Jp[i] += eps
        Jm[i] -= eps
        errs.append(-g[i].real - (bdg(Jp, h) - bdg(Jm, h)) / (2 * eps))
    for i in range(N):
        hp, hm = h.copy(), h.copy()
        hp[i] += eps
        hm[i] -= eps
        errs.append(-g[N - 1 + i].real - (bdg(J, hp) - bdg(J, hm)) / (2 * eps))
    out["max_abs_hellmann_feynman_error"] = float(np.max(np.abs(errs)))
    if cpu:
        from oracle.mps import MPS as OMPS
        from oracle.mps import inner_oplist_env

        ops = {"z": np.diag([1.0, -1.0]), "x": np.array([[0.0, 1.0], [1.0, 0.0]])}
        terms = [([ops[o] for o in oo], list(ss), 1.0) for oo, ss in zip(H.ops, H.sites)]
        om = OMPS([np.asarray(t) for t in psi.tensors], psi.center)
        t0 = time.perf_counter()
        ref = inner_oplist_env(om, terms)
        cpu_s = time.perf_counter() - t0
        out["cpu_port"] = {"seconds": cpu_s, "threads": os.cpu_count(),
                           "max_abs_diff_vs_gpu": float(np.abs(ref - g).max()),
                           "sample": "all 127 <O_t> from oracle/mps.py inner_oplist_env on the same tensors"}
        out["speedup_vs_cpu_port"] = cpu_s / el
    return out


def qjmc_tfim_n40(ntraj=512, steps=200, sample_traj=16, sample_steps=20, nworkers=8, chi_max=256, cpu_traj=2):
    """Quantum-jump trajectories of the dissipative TFIM, N = 40, g = 0.5, jump operators
    s-_i at rate 0.1, dt = 0.01, second-order Trotter, cutoff 1e-8, chi_max 256, from all
    spins up, on fixed uniforms (seed 0), <sz_i> saved every step. The full workload
    (`ntraj` x `steps`) runs on the batched engine (all trajectories per launch); a bounded
    sample (the first `sample_traj` trajectories, `sample_steps` steps) also runs on the
    per-trajectory workers (`nworkers` streams), whose jump records must equal the batched
    engine's on the same uniforms, and one trajectory through the Python host loop.
    CPU baseline: the first `cpu_traj` trajectories of the workload through the oracle's MPS
    trajectory (numpy on the host cores), checked against the batched engine's rows.
    Reports trajectory steps per second for each."""
    from tnhip import productMPS, spinhalf
    from tnhip.models import OpList
    from tnhip.qjmc import qjmc_simulation, qjmc_trajectories
    from tnhip.tebd import DeviceMPS

    N, g, gamma, dt = 40, 0.5, 0.1, 0.01
    H = OpList(N)
    for i in range(1, N):
        H.add(["z", "z"], [i, i + 1], -1.0)
    for i in range(1, N + 1):
        H.add("x", i, -g)
    Jo = OpList(N)
    for i in range(1, N + 1):
        Jo.add("s-", i, gamma ** 0.5)
    zs = OpList(N)
    for i in range(1, N + 1):
        zs.add("z", i)
    st = spinhalf()
    U = np.random.default_rng(0).random((ntraj, 4 * steps + 16))
    kw = dict(observables=zs, cutoff=1e-8, maxdim=chi_max)
    # batched engine: warm-up (code paths, allocations), then the full workload
    qjmc_trajectories(st, ["up"] * N, H, Jo, 2 * dt, dt, U[:8], engine="batched", **kw)
    t0 = time.perf_counter()
    full = qjmc_trajectories(st, ["up"] * N, H, Jo, steps * dt, dt, U, engine="batched", **kw)
    el_full = time.perf_counter() - t0
    # bounded sample on both engines (sample_traj = 0: batched engine only, one host thread)
    Us = U[:max(sample_traj, 1)]
    a = qjmc_trajectories(st, ["up"] * N, H, Jo, sample_steps * dt, dt, Us, engine="batched", **kw)
    res_w = None
    if sample_traj > 0:
        t0 = time.perf_counter()
        w = qjmc_trajectories(st, ["up"] * N, H, Jo, sample_steps * dt, dt, Us, engine="workers",
                              nworkers=nworkers, **kw)
        el_w = time.perf_counter() - t0
        res_w = {"seconds": el_w, "trajectories": sample_traj, "steps": sample_steps, "workers": nworkers,
                 "trajectory_steps_per_s": sample_traj * sample_steps / el_w,
                 "jump_records_equal_batched": w["jumps"] == a["jumps"],
                 "max_abs_meas_diff_vs_batched": float(np.abs(w["measurements"] - a["measurements"]).max())}
    # host loop, one trajectory
    psi = DeviceMPS.from_gmps(productMPS(st, ["up"] * N))
    u = iter(U[0])
    t0 = time.perf_counter()
    jumps, _ = qjmc_simulation(st, psi, H, Jo, sample_steps * dt, dt, [], cutoff=1e-8, maxdim=chi_max,
                               verbose=False, rand=lambda: float(next(u)))
    psi.ctx.synchronize()
    host_el = time.perf_counter() - t0
    sz = full["measurements"].real.mean(axis=(0, 2))
    cpu = None
    if cpu_traj > 0:  # the oracle's MPS trajectory (qjmc.jl:28-167) on the same uniforms
        from oracle.qjmc import qjmc_mps

        def terms(ol):
            return [(list(o), list(s_), c) for o, s_, c in zip(ol.ops, ol.sites, ol.coeffs)]

        up = np.array([1.0, 0.0])
        t0 = time.perf_counter()
        same = True
        for t in range(cpu_traj):
            j, _, _, ms, _ = qjmc_mps([up] * N, terms(H), terms(Jo), N, dt, steps * dt, U[t],
                                     obs=[("z", i) for i in range(1, N + 1)], cutoff=1e-8, maxdim=chi_max)
            same = same and j == full["jumps"][t] and np.abs(ms.real - full["measurements"][t].real).max() <= 1e-8
        el_cpu = time.perf_counter() - t0
        cpu = {"seconds": el_cpu, "trajectories": cpu_traj, "steps": steps, "threads": os.cpu_count(),
               "trajectory_steps_per_s": cpu_traj * steps / el_cpu,
               "jump_records_and_sz_equal_batched": bool(same),
               "sample": "oracle/qjmc.py qjmc_mps, the first trajectories of the workload"}
    return {"batched": {"seconds": el_full, "trajectories": ntraj, "steps": steps,
                        "trajectory_steps_per_s": ntraj * steps / el_full, "jumps": int(sum(full["njumps"])),
                        "maxbonddim": int(max(full["maxbonddim"])), "sz_avg_t0": float(sz[0]),
                        "sz_avg_final": float(sz[-1])},
            "workers_sample": res_w,
            "host_loop": {"seconds": host_el, "steps": sample_steps,
                          "trajectory_steps_per_s": sample_steps / host_el, "jumps": len(jumps),
                          "jump_records_equal_batched": [j for j in jumps] == a["jumps"][0]},
            "cpu_port": cpu,
            "batched_speedup_vs_cpu_port": (ntraj * steps / el_full) / cpu["trajectory_steps_per_s"] if cpu else None}


def qjmc_tfim_n40_critical(ntraj=256, steps=150, sample_traj=4, nworkers=4, chi_max=256, cpu_traj=1):
    """Quantum-jump trajectories in the regime of uneven, growing bond dimensions: the
    critical dissipative TFIM, N = 40, g = 1, s-_i at rate 0.02, dt = 0.02, 150 second-order
    Trotter steps, cutoff 1e-10, chi_max 256, from all spins up, uniforms from seed 1. Bonds
    pass 16 near t = 2 and reach 31-32 by t = 3, so every
    trajectory outgrows the narrow batched slots and continues on the wide ones from the
    start of the step it outgrew them in (one 1024-thread workgroup per trajectory,
    divide-and-conquer splits); the time includes the narrow part. The first `sample_traj` trajectories also run on the per-trajectory workers
    (`nworkers` streams, one trajectory per stream), and the first `cpu_traj` through the
    oracle's MPS trajectory; both must reproduce the batched rows (jump records, <sz_i> to
    1e-8). Reports trajectory steps per second for each."""
    from tnhip import spinhalf
    from tnhip.models import OpList
    from tnhip.qjmc import qjmc_trajectories

    N, g, gamma, dt, cutoff = 40, 1.0, 0.02, 0.02, 1e-10
    H = OpList(N)
    for i in range(1, N):
        H.add(["z", "z"], [i, i + 1], -1.0)
    for i in range(1, N + 1):
        H.add("x", i, -g)
    Jo = OpList(N)
    for i in range(1, N + 1):
        Jo.add("s-", i, gamma ** 0.5)
    zs = OpList(N)
    for i in range(1, N + 1):
        zs.add("z", i)
    st = spinhalf()
    U = np.random.default_rng(1).random((ntraj, 4 * steps + 16))
    kw = dict(observables=zs, cutoff=cutoff, maxdim=chi_max)
    qjmc_trajectories(st, ["up"] * N, H, Jo, 2 * dt, dt, U[:8], engine="batched", **kw)  # warm-up
    t0 = time.perf_counter()
    full = qjmc_trajectories(st, ["up"] * N, H, Jo, steps * dt, dt, U, engine="batched", **kw)
    el_full = time.perf_counter() - t0
    print(f"qjmc critical: batched {ntraj} x {steps} in {el_full:.2f} s", file=sys.stderr, flush=True)
    res_w = None
    if sample_traj > 0:
        t0 = time.perf_counter()
        w = qjmc_trajectories(st, ["up"] * N, H, Jo, steps * dt, dt, U[:sample_traj], engine="workers",
                              nworkers=nworkers, **kw)
        el_w = time.perf_counter() - t0
        res_w = {"seconds": el_w, "trajectories": sample_traj, "steps": steps, "workers": nworkers,
                 "trajectory_steps_per_s": sample_traj * steps / el_w,
                 "jump_records_equal_batched": w["jumps"] == full["jumps"][:sample_traj],
                 "max_abs_meas_diff_vs_batched": float(np.abs(w["measurements"] -
                                                              full["measurements"][:sample_traj]).max())}
    cpu = None
    if cpu_traj > 0:
        from oracle.qjmc import qjmc_mps

        def terms(ol):
            return [(list(o), list(s_), c) for o, s_, c in zip(ol.ops, ol.sites, ol.coeffs)]

        up = np.array([1.0, 0.0])
        t0 = time.perf_counter()
        same = True
        for t in range(cpu_traj):
            j, _, _, ms, chi = qjmc_mps([up] * N, terms(H), terms(Jo), N, dt, steps * dt, U[t],
                                       obs=[("z", i) for i in range(1, N + 1)], cutoff=cutoff, maxdim=chi_max)
            same = same and j == full["jumps"][t] and chi == full["maxbonddim"][t] and \
                np.abs(ms.real - full["measurements"][t].real).max() <= 1e-8
        el_cpu = time.perf_counter() - t0
        cpu = {"seconds": el_cpu, "trajectories": cpu_traj, "steps": steps, "threads": os.cpu_count(),
               "trajectory_steps_per_s": cpu_traj * steps / el_cpu,
               "jump_records_bonds_and_sz_equal_batched": bool(same),
               "sample": "oracle/qjmc.py qjmc_mps, the first trajectories of the workload"}
    bonds = np.array(full["maxbonddim"])
    rate = ntraj * steps / el_full
    return {"batched": {"seconds": el_full, "trajectories": ntraj, "steps": steps, "engine": full["engine"],
                        "trajectory_steps_per_s": rate, "jumps": int(sum(full["njumps"])),
                        "maxbonddim_min_median_max": [int(bonds.min()), int(np.median(bonds)), int(bonds.max())]},
            "workers_sample": res_w,
            "cpu_port": cpu,
            "batched_speedup_vs_workers": rate / res_w["trajectory_steps_per_s"] if res_w else None,
            "batched_speedup_vs_cpu_port": rate / cpu["trajectory_steps_per_s"] if cpu else None}


def tebd_tfim_n40(steps=80, dt=0.05, chi_max=256, cutoff=1e-10, cpu_steps=0):
    """Real-time TEBD of the critical TFIM quench, N = 40, g = 1, from all spins up:
    second-order Trotter (exp(-i dt H), 60 two-site gates per step), cutoff 1e-10, chi_max
    256, the centre moved by truncated splits before each gate (applygates!, gatelist.jl:
    191-223) and the state renormalised every step. Complex arithmetic throughout. Reports
    seconds per Trotter step at the end of the run and the bond dimension reached. With
    cpu_steps, the final state also takes cpu_steps more steps in the CPU port (numpy/LAPACK
    on the host cores) and on the GPU, timed side by side, with the largest difference of
    the local <sz_i> after them."""
    from tnhip import productMPS, spinhalf
    from tnhip.models import OpList
    from tnhip.tebd import DeviceMPS, applygates, trotterize

    N, g = 40, 1.0
    H = OpList(N)
    for i in range(1, N):
        H.add(["z", "z"], [i, i + 1], -1.0)
    for i in range(1, N + 1):
        H.add("x", i, -g)
    st = spinhalf()
    psi = DeviceMPS.from_gmps(productMPS(st, ["up"] * N))
    gates = trotterize(st, H, dt, order=2, evol="real")
    ngates = sum(len(s) for s in gates.sites)
    per_step, bonds = [], []
    t_all = time.perf_counter()
    for _ in range(steps):
        t0 = time.perf_counter()
        applygates(psi, gates, cutoff=cutoff, maxdim=chi_max)
        psi.normalize()
        psi.ctx.synchronize()
        per_step.append(time.perf_counter() - t0)
        bonds.append(psi.maxbonddim())
    total = time.perf_counter() - t_all
    out = {"steps": steps, "dt": dt, "gates_per_step": ngates, "seconds": total,
           "final_maxbonddim": bonds[-1], "last_step_seconds": per_step[-1],
           "last_step_gates_per_s": ngates / per_step[-1],
           "step_seconds_at": {str(bonds[i]): per_step[i] for i in (9, 39, 59, steps - 1) if i < steps}}
    if cpu_steps:
        import torch

        from oracle.mps import MPS
        from oracle.mps import normalize as normalize_cpu
        from oracle.tebd import applygates_mps
        from tnhip.tebd import inner

        zs = OpList(N)
        for i in range(1, N + 1):
            zs.add("z", i)
        p = MPS([t.numpy() for t in psi.sites], psi.center)
        rows = list(zip(gates.sites, gates.gates))
        t0 = time.perf_counter()
        for _ in range(cpu_steps):
            applygates_mps(p, rows, cutoff=cutoff, maxdim=chi_max)
            normalize_cpu(p)
        cpu_s = (time.perf_counter() - t0) / cpu_steps
        t0 = time.perf_counter()
        for _ in range(cpu_steps):
            applygates(psi, gates, cutoff=cutoff, maxdim=chi_max)
            psi.normalize()
        psi.ctx.synchronize()
        gpu_s = (time.perf_counter() - t0) / cpu_steps
        # both states measured by the same (GPU) contraction of <sz_i>
        z_gpu = np.real(inner(st, psi, zs))
        z_cpu = np.real(inner(st, DeviceMPS(p.tensors, p.center, ctx=psi.ctx), zs))
        out["cpu_port_sample"] = {"steps": cpu_steps, "from_maxbonddim": bonds[-1],
                                  "cpu_seconds_per_step": cpu_s, "gpu_seconds_per_step": gpu_s,
                                  "cpu_threads": torch.get_num_threads(), "speedup_vs_cpu": cpu_s / gpu_s,
                                  "max_abs_dsz": float(np.max(np.abs(z_gpu - z_cpu)))}
    return out


def bond_step(ctx, ops, Tensor, dL, dR, dW, dx, prec, chi, torch):
    """One two-site DMRG bond step at chi on the synthetic workload: Lanczos (Krylov
    dimension 3, 2 restarts, as dmrg.jl's defaults), the truncated split to chi, and the
    left environment update. Returns seconds per phase."""
    th = dx.astype(np.float64)
    ctx.synchronize()
    t0 = time.perf_counter()
    e, nmv = ops.eigsolve_ground(dL, dW, dW, dR, th, krylovdim=3, maxiter=2, tol=1e-14, precision=prec)
    ctx.synchronize()
    t1 = time.perf_counter()
    A1, A2, s, disc = ops.replacesites2(th, 0, True, 1e-12, chi, 1)
    ctx.synchronize()
    t2 = time.perf_counter()
    Ln = ops.env_left(dL, A1, dW, precision=prec)
    ctx.synchronize()
    t3 = time.perf_counter()
    del Ln, A1, A2, th
    return {"lanczos_s": t1 - t0, "matvecs": nmv, "split_svd_s": t2 - t1, "env_update_s": t3 - t2,
            "bond_step_s": t3 - t0, "energy": e, "kept": int(s.size)}


def bond_steps(ctx, ops, Tensor, dL64, dR64, dL32, dR32, dW, dx, chi, torch):
    """The bond step with FP64, FP32-MFMA and split-BF16x3 matvecs (FP64 Lanczos vectors
    and SVD in all), from the same random theta: a worst case for the split (a dense, flat
    spectrum, cold-started Jacobi). The N = 100 TFIM half-sweeps at chi = 2048 are measured
    separately (dmrg_tfim_n100_chi2048)."""
    steps_ = {}
    for prec in ("f64", "f32", "bf16x3"):
        dL, dR = (dL64, dR64) if prec == "f64" else (dL32, dR32)
        steps_[prec] = bond_step(ctx, ops, Tensor, dL, dR, dW, dx, prec, chi, torch)
    for prec in ("f32", "bf16x3"):
        steps_[f"bond_step_speedup_{prec}_vs_f64"] = steps_["f64"]["bond_step_s"] / steps_[prec]["bond_step_s"]
        steps_[f"lanczos_speedup_{prec}_vs_f64"] = steps_["f64"]["lanczos_s"] / steps_[prec]["lanczos_s"]
    return steps_


def run_matvecs(ctx, T, ops, dL, dR, dW, dx, out, prec, steps, warmup, torch):
    for _ in range(warmup):
        ops.heff_apply(dL, dW, dW, dR, dx, out=out, precision=prec)
    ctx.synchronize()
    torch.cuda.synchronize()
    ctx.profile_reset()
    ctx.profile(True)
    t0 = time.perf_counter()
    for _ in range(steps):
        ops.heff_apply(dL, dW, dW, dR, dx, out=out, precision=prec)
    ctx.synchronize()
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    ctx.profile(False)
    fam = {"f32": "gemm_f32", "f64": "gemm_f64", "bf16x3": "gemm_bf16x3"}[prec]
    g = ctx.profile_read(fam)
    mix = ctx.profile_read("mix")
    return el, g, mix


def matvec_chi_sweep(ctx, ops, Tensor, torch, chis=(256, 512, 1024), seed=0, D=5, d=2, steps=20, warmup=3):
    """The matvec at the smaller bond dimensions of the synthetic workload, per precision,
    with the autotuned tiles and with the fixed size rule's tiles (same inputs)."""
    out = {}
    W = heisenberg_bulk_w()
    dW = Tensor.from_numpy(W, ctx=ctx)
    for chi in chis:
        Lt, Rt, tht = make_inputs(chi, D, seed, torch)
        Lt32, Rt32 = Lt.float(), Rt.float()
        o = torch.empty(chi * d * d * chi, device="cuda", dtype=torch.float64)
        dx = Tensor.wrap(tht.data_ptr(), (chi, d, d, chi), np.float64, owner=tht, ctx=ctx)
        do = Tensor.wrap(o.data_ptr(), (chi, d, d, chi), np.float64, owner=o, ctx=ctx)
        F = matvec_flops(chi, D, d)
        row = {"gflop_per_matvec": F / 1e9}
        for prec in ("f64", "f32", "bf16x3"):
            src = (Lt, Rt) if prec == "f64" else (Lt32, Rt32)
            dt = np.float64 if prec == "f64" else np.float32
            dL = Tensor.wrap(src[0].data_ptr(), (chi, D, chi), dt, owner=src[0], ctx=ctx)
            dR = Tensor.wrap(src[1].data_ptr(), (chi, D, chi), dt, owner=src[1], ctx=ctx)
            r = {}
            for mode in ("rule", "autotuned"):
                ctx.gemm_autotune(mode == "autotuned")
                el, g, _ = run_matvecs(ctx, None, ops, dL, dR, dW, dx, do, prec, steps, warmup, torch)
                r[mode] = {"ms_per_matvec": 1e3 * el / steps, "tflops": F * steps / el / 1e12,
                           "gemm_tflops": (g["flops"] / g["ms"] / 1e9) if g["ms"] > 0 else None}
            ctx.gemm_autotune(True)
            r["tflops"] = r["autotuned"]["tflops"]
            r["frac_of_peak"] = r["tflops"] / PEAK[prec]
            r["autotune_speedup_vs_rule"] = r["rule"]["ms_per_matvec"] / r["autotuned"]["ms_per_matvec"]
            row[prec] = r
        out[str(chi)] = row
        del dx, do
    return out


def svd_vs_cpu(ctx, ops, Tensor, torch, n=2048, seed=0):
    """Truncated SVD of one seeded n x n N(0,1) matrix (kept rank n / 2, as a chi = n / 2
    split) on the GPU (warm) and LAPACK gesdd through numpy on the host cores (the routine
    the reference's svd calls, src/tensors.jl:190), with the largest singular-value
    difference between them relative to the largest singular value."""
    g = torch.Generator(device="cpu").manual_seed(seed + 7)
    Mh = torch.randn(n, n, generator=g, dtype=torch.float64)
    Md = Mh.to("cuda")
    M = Tensor.wrap(Md.data_ptr(), (n, n), np.float64, owner=Md, ctx=ctx)
    times = []
    for _ in range(2):
        ctx.synchronize()
        t0 = time.perf_counter()
        U, S, Vt, disc, fb = ops.svd_trunc(M, 0.0, n // 2, 1, want=(True, True, True))
        ctx.synchronize()
        times.append(time.perf_counter() - t0)
        s_gpu = np.sort(S.numpy())[::-1]
        del U, S, Vt
    # the GPU read the buffer column-major, i.e. M^T: the same singular values
    t0 = time.perf_counter()
    s_cpu = np.linalg.svd(Mh.numpy(), compute_uv=True, full_matrices=False)[1]
    cpu_s = time.perf_counter() - t0
    k = min(len(s_gpu), n // 2)
    return {"n": n, "kept": int(k), "gpu_seconds": times[1], "cpu_seconds": cpu_s,
            "cpu_threads": torch.get_num_threads(), "speedup_vs_cpu": cpu_s / times[1],
            "max_abs_ds_rel": float(np.max(np.abs(s_gpu[:k] - s_cpu[:k])) / s_cpu[0])}


def tile_choices(ctx):
    """The autotuner's table in short form: mode, transposes, M x N x K -> chosen tile."""
    return ["%s %s %dx%dx%d -> %s (rule %s)" % (e["mode"], e["trans"], e["M"], e["N"], e["K"], e["tile"], e["rule"])
            for e in ctx.gemm_tuning()]


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--chi", type=int, default=2048)
    ap.add_argument("--no-svd", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-dmrg", action="store_true", help="skip the DMRG legs")
    ap.add_argument("--dmrg-chi", type=int, default=2048, help="chi_max of the N=100 TFIM sweep leg")
    ap.add_argument("--dmrg-passes", type=int, default=24)
    ap.add_argument("--refine-tol", type=float, default=None,
                    help="relative energy change that switches a lower-precision DMRG run to FP64")
    ap.add_argument("--dmrg-only", action="store_true",
                    help="run only the N=100 TFIM sweep leg (FP64 and FP32+refinement)")
    ap.add_argument("--no-extra-legs", action="store_true", help="skip the gradient and QJMC legs")
    ap.add_argument("--legs-only", action="store_true", help="run only the gradient and QJMC legs")
    ap.add_argument("--dmrg-small-only", action="store_true",
                    help="run only the N=32 Heisenberg DMRG leg (GPU, and the CPU port unless --no-cpu)")
    ap.add_argument("--sweep2048-only", action="store_true",
                    help="run only the N=100 TFIM half-sweeps at bond dimension 2048 (FP64, FP32, BF16x3)")
    ap.add_argument("--f64-pass-only", action="store_true",
                    help="with --sweep2048-only: stop after the first timed FP64 pass")
    ap.add_argument("--profile-passes", action="store_true",
                    help="with --sweep2048-only: kernel-family times per pass (HIP events)")
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--tebd-only", action="store_true", help="run only the real-time TEBD leg")
    ap.add_argument("--qjmc-critical-only", action="store_true",
                    help="run only the critical-chain QJMC leg (bonds above 16, wide batched slots)")
    ap.add_argument("--no-workers", action="store_true",
                    help="with --qjmc-critical-only: skip the per-trajectory workers sample")
    ap.add_argument("--qjmc-only", action="store_true",
                    help="run only the QJMC leg on the batched engine and the host loop (one host thread)")
    ap.add_argument("--bond-only", action="store_true",
                    help="run only the chi=2048 DMRG bond step (FP64 and FP32 matvecs)")
    ap.add_argument("--svd-only", action="store_true",
                    help="run only the truncated SVD of the 2chi x 2chi theta matrix (1 warm-up + 1 timed)")
    ap.add_argument("--no-sweep", action="store_true", help="skip the matvec sweep over chi = 256, 512, 1024")
    ap.add_argument("--sweep-only", action="store_true",
                    help="run only the matvec at chi = 256, 512, 1024 (autotuned vs size-rule tiles)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="run only a plain 1 GiB FP64 device copy (reformat kernel), the "
                         "known-size reference for scaling FETCH_SIZE / WRITE_SIZE counters")
    a = ap.parse_args(argv)
    if a.copy_calibration:
        import tnhip
        from tnhip import Tensor

        n = 1 << 27  # 2^27 doubles = 1 GiB read + 1 GiB written
        src = Tensor.empty((n,), np.float64)
        dst = src.astype(np.float64)
        tnhip.Context.default().synchronize()
        print(json.dumps({"copy_calibration_bytes_read": 8 * n, "copy_calibration_bytes_written": 8 * n,
                          "kernel": "reformat_kernel<double,double>"}))
        del dst
        return 0
    if a.gpus != 1:
        print("bench.py measures one GPU: the library has no multi-GPU path (the matvec "
              "would shard theta along its right bond, with a reduce-scatter of the "
              "contraction over that bond). Run with --gpus 1.", file=sys.stderr)
        return 2

    import torch
    import tnhip
    from tnhip import Tensor, ops

    chi, D, d = a.chi, 5, 2
    ctx = tnhip.Context.default()
    if a.dmrg_only:
        print(json.dumps({"dmrg_tfim_N100": dmrg_tfim_n100(a.dmrg_chi, a.dmrg_passes, a.refine_tol,
                                                           ("f64", "f32", "bf16x3"), cpu=not a.no_cpu)}))
        return 0
    if a.sweep2048_only:
        print(json.dumps({"dmrg_tfim_N100_chi%d_sweeps" % chi: dmrg_tfim_n100_chi2048(
            chi, profile=a.profile_passes, timed="f64" if a.f64_pass_only else None)}))
        return 0
    if a.dmrg_small_only:
        print(json.dumps({"dmrg_heisenberg_N32": dmrg_heisenberg_n32(cpu=not a.no_cpu)}))
        return 0
    if a.tebd_only:
        print(json.dumps({"tebd_tfim_N40": tebd_tfim_n40(cpu_steps=0 if a.no_cpu else 2)}))
        return 0
    if a.qjmc_critical_only:
        print(json.dumps({"qjmc_tfim_N40_critical": qjmc_tfim_n40_critical(
            sample_traj=0 if a.no_workers else 4, cpu_traj=0 if a.no_cpu else 1)}))
        return 0
    if a.qjmc_only:
        print(json.dumps({"qjmc_tfim_N40": qjmc_tfim_n40(sample_traj=0)}))
        return 0
    if a.legs_only:
        print(json.dumps({"gradient_tfim_N64": gradient_tfim_n64(), "qjmc_tfim_N40": qjmc_tfim_n40(),
                          "qjmc_tfim_N40_critical": qjmc_tfim_n40_critical()}))
        return 0
    if a.sweep_only:
        sw = matvec_chi_sweep(ctx, ops, Tensor, torch, seed=a.seed)
        print(json.dumps({"matvec_chi_sweep": sw, "gemm_tiles": tile_choices(ctx)}))
        return 0
    Lt, Rt, tht = make_inputs(chi, D, a.seed, torch)
    if a.svd_only:
        M = Tensor.wrap(tht.data_ptr(), (2 * chi, 2 * chi), np.float64, owner=tht, ctx=ctx)
        times = []
        for _ in range(2):
            ctx.synchronize()
            t0 = time.perf_counter()
            U, S, Vt, disc, fb = ops.svd_trunc(M, 0.0, chi, 1, want=(True, True, True))
            ctx.synchronize()
            times.append(time.perf_counter() - t0)
            del U, S, Vt
        print(json.dumps({"svd_2chi": {"n": 2 * chi, "kept": chi, "seconds_warm": times[1],
                                       "seconds_first": times[0], "discarded": disc, "fallback": fb}}))
        return 0
    Lt32, Rt32 = Lt.float(), Rt.float()
    torch.cuda.synchronize()
    W = heisenberg_bulk_w()
    dW = Tensor.from_numpy(W, ctx=ctx)
    dL64 = Tensor.wrap(Lt.data_ptr(), (chi, D, chi), np.float64, owner=Lt, ctx=ctx)
    dR64 = Tensor.wrap(Rt.data_ptr(), (chi, D, chi), np.float64, owner=Rt, ctx=ctx)
    dL32 = Tensor.wrap(Lt32.data_ptr(), (chi, D, chi), np.float32, owner=Lt32, ctx=ctx)
    dR32 = Tensor.wrap(Rt32.data_ptr(), (chi, D, chi), np.float32, owner=Rt32, ctx=ctx)
    dx = Tensor.wrap(tht.data_ptr(), (chi, d, d, chi), np.float64, owner=tht, ctx=ctx)
    if a.bond_only:  # one untimed pass first: workspaces and schedules are allocated on first use
        bond_steps(ctx, ops, Tensor, dL64, dR64, dL32, dR32, dW, dx, chi, torch)
        print(json.dumps({"dmrg_bond_step_chi%d" % chi: bond_steps(ctx, ops, Tensor, dL64, dR64, dL32, dR32,
                                                                   dW, dx, chi, torch)}))
        return 0
    outs = {p: torch.empty(chi * d * d * chi, device="cuda", dtype=torch.float64) for p in PEAK}
    douts = {p: Tensor.wrap(outs[p].data_ptr(), (chi, d, d, chi), np.float64, owner=outs[p], ctx=ctx)
             for p in PEAK}
    F = matvec_flops(chi, D, d)
    res = {}
    for prec in ("f64", "bf16x3", "f32"):
        dL, dR = (dL64, dR64) if prec == "f64" else (dL32, dR32)
        steps = a.steps
        el, g, mix = run_matvecs(ctx, None, ops, dL, dR, dW, dx, douts[prec], prec, steps, a.warmup, torch)
        res[prec] = {"ms_per_step": 1e3 * el / steps, "tflops": F * steps / el / 1e12, "gemm": g, "mix": mix,
                     "steps": steps}
    ref = outs["f64"]
    for prec in ("f32", "bf16x3"):
        res[prec]["rel_err_vs_f64"] = float((outs[prec] - ref).norm() / ref.norm())

    f32 = res["f32"]
    g = f32["gemm"]
    gemm_ms = g["ms"] / max(g["launches"], 1)
    gemm_flops = g["flops"] / max(g["launches"], 1)
    achieved = gemm_flops / (gemm_ms * 1e-3) / 1e12 if gemm_ms > 0 else 0.0
    roofline = {"bound": "mfma", "achieved": achieved, "peak": PEAK["f32"], "unit": "TFLOP/s",
                "frac": achieved / PEAK["f32"], "traffic": measured_traffic(chi),
                "traffic_source": ("profiles/heff_matvec_chi2048_hbm_traffic.md: copy-calibrated rocprofv3 "
                                   "FETCH_SIZE/WRITE_SIZE runs of this kernel, not measured in this run")
                if chi == 2048 else None,
                "kernel": "gemm_kernel<MODE_F32> (stage 1 L*theta and stage 3 T2*R^T)",
                "flops_per_launch": gemm_flops, "avg_launch_ms": gemm_ms}
    mix = f32["mix"]
    extra = {}
    for prec in ("f64", "bf16x3", "f32"):
        r = res[prec]
        gg = r["gemm"]
        extra[prec] = {"ms_per_step": r["ms_per_step"], "tflops": r["tflops"],
                       "frac_of_peak": r["tflops"] / PEAK[prec],
                       "gemm_avg_ms": gg["ms"] / max(gg["launches"], 1),
                       "gemm_tflops": (gg["flops"] / gg["ms"] / 1e9) if gg["ms"] > 0 else None,
                       "mix_avg_ms": r["mix"]["ms"] / max(r["mix"]["launches"], 1),
                       "mix_GBps": (r["mix"]["bytes"] / r["mix"]["ms"] / 1e6) if r["mix"]["ms"] > 0 else None}
        if "rel_err_vs_f64" in r:
            extra[prec]["rel_err_vs_f64"] = r["rel_err_vs_f64"]
    extra["fp32_speedup_vs_fp64"] = res["f32"]["tflops"] / res["f64"]["tflops"]
    extra["gemm_tiles_chi%d" % chi] = tile_choices(ctx)
    if not a.no_sweep:
        extra["matvec_chi_sweep"] = matvec_chi_sweep(ctx, ops, Tensor, torch, seed=a.seed)

    if not a.no_svd:
        # truncated SVD of the 2chi x 2chi theta matrix, kept rank chi (FP64 Gram + Jacobi)
        # (the first call also allocates workspaces and autotunes its GEMM size classes)
        M = Tensor.wrap(tht.data_ptr(), (2 * chi, 2 * chi), np.float64, owner=tht, ctx=ctx)
        times = []
        for _ in range(2):
            ctx.synchronize()
            t0 = time.perf_counter()
            U, S, Vt, disc, fb = ops.svd_trunc(M, 0.0, chi, 1, want=(True, True, True))
            ctx.synchronize()
            times.append(time.perf_counter() - t0)
            kept = S.shape[0]
            del U, S, Vt
        extra["svd_2chi"] = {"seconds": times[1], "seconds_first": times[0], "kept": kept, "discarded": disc,
                             "fallback": fb, "n": 2 * chi}
        if not a.no_cpu:
            extra["svd_n2048_vs_cpu"] = svd_vs_cpu(ctx, ops, Tensor, torch, n=2048, seed=a.seed)
    if not a.no_dmrg:
        # one untimed pass first: workspaces, schedules and GEMM tiles are set up on first use
        bond_steps(ctx, ops, Tensor, dL64, dR64, dL32, dR32, dW, dx, chi, torch)
        extra["dmrg_bond_step_chi%d" % chi] = bond_steps(ctx, ops, Tensor, dL64, dR64, dL32, dR32, dW, dx,
                                                          chi, torch)
        extra["dmrg_tfim_N100"] = dmrg_tfim_n100(a.dmrg_chi, a.dmrg_passes, a.refine_tol, ("f64", "f32", "bf16x3"),
                                                 cpu=not a.no_cpu)
        extra["dmrg_heisenberg_N32"] = dmrg_heisenberg_n32(cpu=not a.no_cpu)
        extra["dmrg_tfim_N100_chi2048_sweeps"] = dmrg_tfim_n100_chi2048(2048)
    if not a.no_extra_legs:
        extra["gradient_tfim_N64"] = gradient_tfim_n64()
        extra["qjmc_tfim_N40"] = qjmc_tfim_n40()
        extra["qjmc_tfim_N40_critical"] = qjmc_tfim_n40_critical()
        extra["tebd_tfim_N40"] = tebd_tfim_n40(cpu_steps=0 if a.no_cpu else 2)
    cpu = None
    if not a.no_cpu:
        cpu = cpu_baseline()
    line = {
        "metric": METRIC,
        "value": f32["tflops"],
        "unit": "TFLOP/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": f32["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"heff_matvec_heisenberg_chi{chi}_D{D}_d{d}", "chi": chi, "D": D, "d": d,
                   "precision": "fp32-mfma, fp64 krylov vectors", "seed": a.seed},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "extra": extra,
    }
    print(json.dumps(line))
    return 0


if __name__ == "__main__":
    sys.exit(main())
