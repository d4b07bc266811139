#!/usr/bin/env python3
"""Benchmark: learned-warm-start ground-state scan throughput and matrix-free H·v
bandwidth on one MI355X.

Headline workload: the transverse+longitudinal-field Ising chain
H = -Σ Z_i Z_{i+1} - hx Σ X_i - hz Σ Z_i (Pauli operators), L = 26 periodic,
momentum 0, parity +1 (N = 1,296,858), scanned on the 64×64 grid
hx ∈ [0.2, 2.0], hz ∈ [0, 0.5]. The operator is one compiled family
(couplings re-evaluated per grid point on the same sector tables). Setup,
timed and reported separately: 256 seeded training points solved by plain
Lanczos (tolerance 1e-10), the 32-dimensional snapshot space S of their
ground states (the 32 leading principal components) with the family's
projection on it, and the 2-64-64-32 FP32 network trained on the GPU by the
library's own kernels.

One step = 16 held-out grid points. The timed points are solved in one call:
16 Lanczos windows in lockstep (shared H·v launches per Lanczos step over
the windows' interleaved vectors, up to 16 FP32 and 8 FP64 vectors per
launch), in batches of 16 points that start together, each point solved to
residual
≤ 1e-8·max(|E|,1) with the FP32-matvec schedule from the learned start
S·(y(c) + MLP(c)): y(c) is the lowest eigenvector of the family projected on
S, and the network maps the couplings c to the step from y(c) to the
normalized projection of the ground state on S (`--start snapshot` makes the
plain S·y(c) start the headline instead). `value` is ground-state solves per
second over the timed steps. Beside it (extras, one compact JSON line;
`--details FILE` writes the per-point figures): the same points solved by
plain FP64 eigsh one point at a time (the baseline), from the other start,
from the headline start in FP64 and one point at a time, and plain FP64 in
lockstep; the amortized throughput of the whole 4,096-point scan including
setup; H·v bandwidth at this sector; the FP64 6-eigenpair solve of the XXZ
L=24 sector, the J1-J2 L=28 sector's H·v and ground state, the 72 XXZ L=24
momentum/magnetization sectors in one lockstep batch against one by one, and
exp(-iHt)ψ of 64 seeded states at t ∈ {0.1, 1, 10}.

The roofline object is the 16-lane FP32 batched scan H·v, the kernel the
timed steps spend most of their time in: the mean duration of its launches
inside the timed region, from HIP events the library's launch timer records
on the kernel's own stream, against its algorithmic bytes
N·(4 + 8W) + nnz·(4 + 4W) per launch (W = 16 lanes sharing one index stream).
"""

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

METRIC = "GS solves/s @ resid<1e-8, warm-start+FP32 vs plain FP64; H·v HBM GB/s"
HBM_PEAK_GBS = 8000.0
# HBM bytes per launch of the stream H·v at the headline sector (hx = 1,
# hz = 0.25): FP32 input with FP64 accumulation and output (the FP32
# schedule's product, the roofline kernel) and pure FP32, from the rocprofv3
# FETCH_SIZE / WRITE_SIZE runs of `bench.py --hv-profile` calibrated on a
# 256 MiB device copy (summary and commands in profiles/hv_stream_ising_L26.md)
HV_MIXED_TRAFFIC_BYTES = 275021548
HV_FP32_TRAFFIC_BYTES = 267536152
# HBM bytes per launch of the 16-lane FP32 batched scan H·v at 16 held-out
# grid points (the roofline kernel), from the same counter runs and
# calibration (profiles/scan_batch_hv_ising_L26.md)
SCAN_W16_TRAFFIC_BYTES = 2459658887
L_SCAN = 26
BATCH = 16
# the headline start: the lowest eigenvector of the family projected on the
# snapshot space (scan solve mode "rayleigh_ritz"); the network's start (mode
# True) is measured beside it. The space: the SNAPSHOTS leading principal
# components of the training ground states.
STARTS = {"network": True, "snapshot": "rayleigh_ritz"}
HEADLINE_TEXT = {
    "network": "learned start v0 = S(y(c)+MLP(c)) (S: 32 principal components of 256 plain-Lanczos training ground "
               "states; y(c): Rayleigh-Ritz on S; MLP: 2-64-64-32 FP32, trained on the GPU on the step from y(c) to "
               "the ground state's projection on S)",
    "snapshot": "snapshot start v0 = S y(c) (S: 32 principal components of 256 plain-Lanczos training ground states; "
                "y(c): Rayleigh-Ritz on S)",
}
EPOCHS = 4000
SPACE = "principal"
N_TRAIN = 256
SNAPSHOTS = 32
TOL = 1e-8


def ising_family(L):
    bonds = [[i, (i + 1) % L] for i in range(L)]
    static = [("zz", [(-1.0, b) for b in bonds])]
    comps = [("hx", [("x", [(-1.0, [i]) for i in range(L)])], None),
             ("hz", [("z", [(-1.0, [i]) for i in range(L)])], None)]
    return static, comps


def ising_terms(L, hx, hz):
    bonds = [[i, (i + 1) % L] for i in range(L)]
    return [("zz", [(-1.0, b) for b in bonds]), ("x", [(-hx, [i]) for i in range(L)]),
            ("z", [(-hz, [i]) for i in range(L)])]


def xxz_terms(L, delta):
    b = [[i, (i + 1) % L] for i in range(L)]
    return [("zz", [(delta, x) for x in b]), ("+-", [(0.5, x) for x in b]), ("-+", [(0.5, x) for x in b])]


def j1j2_terms(L, j2=0.5):
    out = []
    for nn, j in ((1, 1.0), (2, j2)):
        b = [[i, (i + nn) % L] for i in range(L)]
        out += [("zz", [(j, x) for x in b]), ("+-", [(0.5 * j, x) for x in b]), ("-+", [(0.5 * j, x) for x in b])]
    return out


def scan_grid():
    hx = np.linspace(0.2, 2.0, 64)
    hz = np.linspace(0.0, 0.5, 64)
    grid = np.array([(a, b) for a in hx for b in hz])
    perm = np.random.default_rng(0).permutation(len(grid))
    return grid, perm


def hv_roofline(model, dtype, vec_bytes, iters=50):
    """dtype "mixed": FP32 input gathered, FP64 output (the FP32 schedule's
    product); vec_bytes is then the input element size."""
    nnz = model.nnz()
    if dtype == "mixed":
        ms = model.time_apply_mixed(complex_=not model.is_real, iters=iters)
        out_bytes = 2 * vec_bytes
    else:
        ms = model.time_apply(dtype, iters=iters)
        out_bytes = vec_bytes
    algo = model.dim * (4 + out_bytes) + nnz * (vec_bytes + 4)
    gbs = algo / (ms * 1e-3) / 1e9
    return {"dim": model.dim, "nnz": nnz, "bytes_per_launch": algo, "ms": ms, "achieved_GBs": gbs,
            "frac": gbs / HBM_PEAK_GBS, **model.hv_config()}


def cpu_baseline(n_target):
    """The CPU oracle (single-threaded restatement of the reference algorithms)
    on a bounded sample of the scan: one ground-state eigsh each (the
    reference's non-restarted 256-vector DGKS window) of the same Hamiltonian
    at hx = 1, hz = 0.25 in the k = 0, p = +1 sectors of L = 20 and L = 22,
    sector build and assembly timed apart. Both the stored matvec and the
    reorthogonalization are linear in N at a fixed window, so the seconds per
    solve per row measured at the two sizes are compared, and the larger
    sample's rate is scaled to the L = 26 sector by N."""
    try:
        from oracle import oracle_ffi
        oracle_ffi.build()
        rows = []
        for L in (20, 22):
            t0 = time.perf_counter()
            m = oracle_ffi.Model(L, ising_terms(L, 1.0, 0.25), momentum=0, parity=1, norm="pauli_cartesian")
            build = time.perf_counter() - t0
            t0 = time.perf_counter()
            r = m.eigsh(1)
            dt = time.perf_counter() - t0
            rows.append((L, m.dim, build, dt, r.get("iterations"), r["eigenvalues"][0]))
        (l1, n1, b1, t1, i1, _), (l2, n2, b2, t2, i2, e2) = rows
        per_row = (t1 / n1, t2 / n2)
        return {"value": (1.0 / t2) * (n2 / n_target), "unit": "solves/s", "cores": 1, "kind": "port",
                "sample": f"oracle eigsh(1), window {i2}, Ising L={l2} k=0 p=+1 (N={n2}) at hx=1 hz=0.25: "
                          f"{t2:.2f} s (build {b2:.1f} s apart, E0={e2:.10f}); L={l1} (N={n1}): {t1:.2f} s; "
                          f"s/solve/row {per_row[0]:.3e} vs {per_row[1]:.3e} (ratio {per_row[1] / per_row[0]:.2f}); "
                          f"rate scaled to N={n_target} linearly in N"}
    except Exception as e:  # the baseline is reported, never required
        return {"value": None, "unit": "solves/s", "cores": 1, "kind": "port", "sample": f"unavailable: {e}"}


def hv_profile(qa):
    """Kernels for counter runs: the FP32 and FP64 H·v of the headline sector
    at hx = 1, hz = 0.25 (20 timed launches each after the autotuner's own),
    the 16-lane and 8-lane FP32 batched scan H·v at held-out grid points (2
    untimed and 20 timed launches each, in the form --scan-hv names), and 5
    device-to-device copies of 256 MiB, whose counter readings calibrate the
    bytes the H·v counters report."""
    static, comps = ising_family(L_SCAN)
    model = qa.Model({"kind": "spin", "sites": L_SCAN, "pauli": True, "momentum": 0, "parity": 1}, static,
                     components=comps)
    model.set_parameters(hx=1.0, hz=0.25)
    out = {"hv_fp32": hv_roofline(model, np.float32, 4, iters=20),
           "hv_fp64": hv_roofline(model, np.float64, 8, iters=20),
           "hv_mixed": hv_roofline(model, "mixed", 4, iters=20)}
    grid, perm = scan_grid()
    pts = grid[perm[N_TRAIN:N_TRAIN + 16]]
    nnz = model.nnz()
    for w in (16, 8):
        ms = model.time_scan_batch(pts[:w], fp32=True, iters=20)
        out[f"scan_fp32_w{w}"] = {"ms": ms, "bytes_per_launch": scan_bytes_per_launch(model.dim, nnz, w, True)}
    L = qa.library()
    nbytes = 256 << 20
    a, b = qa.DeviceBuffer(nbytes), qa.DeviceBuffer(nbytes)
    for _ in range(5):
        L.qmbed_hip_memcpy(b.ptr, a.ptr, nbytes, 3)
    L.qmbed_hip_synchronize()
    a.free()
    b.free()
    out["copy_bytes"] = nbytes
    model.close()
    print(json.dumps(out))


HV_MODE_CASES = {
    "c2_xxz_L24": ({"kind": "spin", "sites": 24, "up": 12, "momentum": 0}, lambda: xxz_terms(24, 1.5)),
    "c3_j1j2_L28": ({"kind": "spin", "sites": 28, "up": 14, "momentum": 0}, lambda: j1j2_terms(28)),
    "c5_ising_L26": ({"kind": "spin", "sites": 26, "pauli": True, "momentum": 0, "parity": 1},
                     lambda: ising_terms(26, 1.0, 0.25)),
}


def hv_modes(qa, iters=20):
    """The FP64 H·v of the XXZ L=24 up=12 k=0, J1-J2 L=28 up=14 k=0 and Ising
    L=26 k=0 p=+1 sectors in each mode on its own: the on-the-fly modes
    (orbit walk + sorted-state search; image table) and the compiled stream,
    each with the autotuner's lanes per row for that mode. Per mode: [ms,
    fraction of HBM peak for N(4+8) + nnz(4+8) bytes, lanes per row]."""
    out = {}
    for name, (basis, terms) in HV_MODE_CASES.items():
        row = {}
        for mode in ("walk", "image", "stream"):
            os.environ["QMBED_HV_MODE"] = mode
            try:
                m = qa.Model(basis, terms())
                h = hv_roofline(m, np.float64, 8, iters)
                row[mode] = [sig(h["ms"]), sig(h["frac"], 3), h["lanes_per_row"]]
                m.close()
            except qa.QmbedError as e:
                row[mode] = f"unavailable: {e}"[:80]
            finally:
                del os.environ["QMBED_HV_MODE"]
        out[name] = row
    return out


def sector_forms(qa, iters=20):
    """Timings of the forms beside the headline path, one GPU:
    * a wide sector: the Heisenberg ring L=32 up=16 k=0 p=+1 on 64-bit states
      (split stream words), its build, FP64 and FP32-schedule H·v, and the
      FP32-schedule ground state;
    * the same ring's general sector (translation k=0, even reflection, even
      spin inversion; maximum representatives, descending rows): build, H·v
      and ground state;
    * a stored CSR matrix (XXZ L=24 up=12 k=0, assembled on the GPU): its
      FP64 product beside the matrix-free one;
    * the shift target on the same sector: the 2 eigenpairs nearest a shift
      between the 4th and 5th lowest eigenvalues (MINRES inner solves).
    H·v entries are [ms, fraction of HBM peak by N·(4+s) + nnz·(4+s)]."""
    out = {}

    def hv(m, dtype, size):
        h = hv_roofline(m, dtype, size, iters)
        return [sig(h["ms"]), sig(h["frac"], 3)]

    def ground(m):
        t0 = time.perf_counter()
        r = m.eigsh(1, tolerance=1e-10, matvec_precision="fp32")
        return [sig(time.perf_counter() - t0, 3), r["iterations"], r["eigenvalues"][0]]

    L = 32
    ring = xxz_terms(L, 1.0)
    t0 = time.perf_counter()
    wide = qa.Model({"kind": "spin", "sites": L, "up": L // 2, "momentum": 0, "parity": 1}, ring)
    build = time.perf_counter() - t0
    out["wide_L32_k0p1"] = {"N": wide.dim, "nnz": wide.nnz(), "build_s": sig(build, 3),
                            "hv_fp64": hv(wide, np.float64, 8), "hv_fp32_in": hv(wide, "mixed", 4),
                            "gs_s_it_E0": ground(wide)}
    wide.close()
    maps = [{"destinations": [(i + 1) % L for i in range(L)], "sector": 0},
            {"destinations": [L - 1 - i for i in range(L)], "sector": 0},
            {"destinations": list(range(L)), "local_permutations": [[1, 0]] * L, "sector": 0}]
    t0 = time.perf_counter()
    gen = qa.Model({"kind": "spin", "sites": L, "up": L // 2, "symmetries": maps, "reverse": True}, ring)
    build = time.perf_counter() - t0
    out["general_L32_tpz"] = {"N": gen.dim, "nnz": gen.nnz(), "build_s": sig(build, 3),
                              "hv_fp64": hv(gen, np.float64, 8), "gs_s_it_E0": ground(gen)}
    gen.close()
    c2 = {"kind": "spin", "sites": 24, "up": 12, "momentum": 0}
    free = qa.Model(c2, xxz_terms(24, 1.5))
    t0 = time.perf_counter()
    csr = qa.Model(c2, xxz_terms(24, 1.5), format="csr")
    build = time.perf_counter() - t0
    out["csr_xxz_L24"] = {"build_s": sig(build, 3), "hv_fp64": hv(csr, np.float64, 8),
                          "matrix_free_hv_fp64": hv(free, np.float64, 8)}
    csr.close()
    w = free.eigsh(6, tolerance=1e-10)["eigenvalues"]
    sigma = 0.5 * (w[3] + w[4]) + 0.1 * (w[4] - w[3])
    free.close()
    req = {"basis": c2, "terms": qa.terms_from_spec(xxz_terms(24, 1.5)), "format": "matrix_free",
           "solver": {"eigenpairs": 2, "target": {"kind": "shift", "value": sigma}}}
    qa.run(req)
    t0 = time.perf_counter()
    r = qa.run(req)
    out["shift_xxz_L24"] = {"sigma": sig(sigma, 6), "s": sig(time.perf_counter() - t0, 3),
                            "iterations": r["iterations"], "eigenvalues": r["eigenvalues"]}
    return out


def sector_batch(qa):
    """XXZ Δ=1, L=24, up ∈ {11,12,13} × k ∈ 0..23: the 72 sectors' ground
    states (and 6 lowest pairs) in one lockstep batch, beside the same solves
    one sector at a time."""
    L = 24
    terms = xxz_terms(L, 1.0)
    keys = [(up, k) for up in (11, 12, 13) for k in range(L)]
    t0 = time.perf_counter()
    models = [qa.Model({"kind": "spin", "sites": L, "up": up, "momentum": k}, terms) for up, k in keys]
    build_s = time.perf_counter() - t0
    dims = [m.dim for m in models]
    out = {"c4_sectors72": {"build_s": build_s, "sectors": len(models), "dim_min": min(dims), "dim_max": max(dims),
                            "dim_sum": sum(dims)}}
    rec = out["c4_sectors72"]
    for k in (1, 6):
        qa.eigsh_batch(models, k)  # warm-up (allocations, index streams)
        t0 = time.perf_counter()
        b = qa.eigsh_batch(models, k)
        tb = time.perf_counter() - t0
        t0 = time.perf_counter()
        single = [m.eigsh(k) for m in models]
        ts = time.perf_counter() - t0
        diff = max(abs(x - y) / abs(y) for r, q in zip(b["results"], single)
                   for x, y in zip(r["eigenvalues"], q["eigenvalues"]))
        rec[f"eigenpairs_{k}"] = {
            "batch_s": tb, "batch_sectors_per_s": len(models) / tb, "steps": b["batch"]["steps"],
            "ritz_rounds": b["batch"]["ritz_rounds"], "fallback": b["batch"]["fallback"],
            "mean_iterations": mean([r["iterations"] for r in b["results"]]),
            "one_by_one_s": ts, "one_by_one_sectors_per_s": len(models) / ts, "speedup": ts / tb,
            "max_rel_diff_vs_one_by_one": diff,
            "max_residual": max(x for r in b["results"] for x in r["residuals"])}
        if k == 1:
            e = [r["eigenvalues"][0] for r in b["results"]]
            rec["lowest_sector"] = {"up": keys[int(np.argmin(e))][0], "k": keys[int(np.argmin(e))][1],
                                    "E0": float(min(e))}
    for m in models:
        m.close()
    return out


def expm_batch(qa):
    """exp(-iHt)ψ for 64 seeded normalized complex states of the XXZ Δ=1, L=24,
    up=12, k=0 sector (N = 112,720) at t ∈ {0.1, 1, 10}, Krylov 30, tolerance
    1e-10, device-resident."""
    from tests import cases

    m = qa.Model({"kind": "spin", "sites": 24, "up": 12, "momentum": 0}, xxz_terms(24, 1.0))
    states = np.array([cases.seeded_state(m.dim, c) for c in range(64)])
    times = [0.1, 1.0, 10.0]
    m.evolve(states[:4], times, krylov_dimension=30)  # warm-up
    # device-resident: states uploaded once, outputs left on the device
    bx = qa.DeviceBuffer(states.nbytes)
    bo = qa.DeviceBuffer(states.nbytes * len(times))
    bx.upload(states)
    qa.library().qmbed_hip_synchronize()
    t0 = time.perf_counter()
    diag = m.evolve_device(bx.ptr, 64, bo.ptr, times, krylov_dimension=30)
    qa.library().qmbed_hip_synchronize()
    dt = time.perf_counter() - t0
    outs = np.empty((len(times), 64, m.dim), dtype=np.complex128)
    bo.download(outs)
    bx.free()
    bo.free()
    # with the host transfers (the vectors in and out over PCIe)
    t0 = time.perf_counter()
    m.evolve(states, times, krylov_dimension=30)
    dt_host = time.perf_counter() - t0
    norm_dev = float(np.max(np.abs(np.linalg.norm(outs, axis=2) - 1.0)))
    m.close()
    return {"c4_expm64": {"seconds": dt, "states_per_s": 64 / dt, "sector_dim": int(states.shape[1]),
                          "times": times, "krylov_dimension": 30, "tolerance": 1e-10,
                          "lanczos_projections": diag["lanczos_projections"],
                          "matrix_vector_products": diag["matrix_vector_products"],
                          "rejected_trial_intervals": diag["rejected_trial_intervals"],
                          "max_norm_deviation": norm_dev,
                          "seconds_with_host_transfers": dt_host}}


def mean(xs):
    return float(np.mean(xs)) if len(xs) else None


def scan_bytes_per_launch(dim, nnz, width, fp32=True):
    """Algorithmic bytes of one batched scan H·v launch over `width` lanes:
    per row the representative's 4-byte record and `width` FP64 outputs, per
    stored entry one 4-byte stream word and the `width` inputs of its target
    row (one shared gather), N·(4 + 8W) + nnz·(4 + sW)."""
    sx = 4 if fp32 else 8
    return dim * (4 + 8 * width) + nnz * (4 + sx * width)


def sig(x, d=4):
    """x rounded to d significant digits (JSON size)."""
    if x is None or x == 0 or not np.isfinite(x):
        return x
    return float(f"{x:.{d}g}")


def mode_summary(rs, ref_energies=None, seconds=None, points=None):
    """Compact per-mode figures of a list of scan_solve results: solves/s,
    mean and max Lanczos iterations, mean FP32-matvec iterations, the largest
    relative energy difference to the plain FP64 solve."""
    def cat(key):
        return [x for r in rs for x in r[key]]
    it = cat("iterations")
    out = {"sps": sig(points / seconds) if seconds else None, "it": sig(mean(it)), "it_max": int(max(it))}
    if any(cat("fp32_iterations")):
        out["it32"] = sig(mean(cat("fp32_iterations")))
    if ref_energies is not None:
        e = np.array(cat("energies"))
        out["dE"] = sig(float(np.max(np.abs(e - ref_energies) / np.abs(ref_energies))), 2)
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=16)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-extended", action="store_true")
    ap.add_argument("--start", choices=sorted(STARTS), default="network",
                    help="the headline start (the other is measured beside it)")
    ap.add_argument("--targets", choices=("correction", "projections", "continuation"), default="correction",
                    help="what the network learns")
    ap.add_argument("--scan-hv", choices=("stream", "walk"), default="stream",
                    help="form of the batched scan H·v: the compiled index stream, or on the fly (orbit walk and "
                         "state directory per target); both give the same results bit for bit")
    ap.add_argument("--details", default=None,
                    help="write the per-point iterations, times and energy differences of every mode to this JSON file")
    ap.add_argument("--hv-profile", action="store_true",
                    help="only the headline sector's FP32 and FP64 H·v (20 launches each), 20 launches of the "
                         "16-lane and 8-lane batched scan H·v and 5 device copies of 256 MiB, for hardware-counter "
                         "runs")
    ap.add_argument("--sector-forms", action="store_true",
                    help="time the wide and general sectors, a stored CSR matrix and the shift target, and print "
                         "them as one JSON object")
    ap.add_argument("--hv-modes", action="store_true",
                    help="only the FP64 H·v of the XXZ L=24, J1-J2 L=28 and Ising L=26 sectors in the walk, image "
                         "and stream modes (20 timed launches each)")
    ap.add_argument("--batch-profile", action="store_true",
                    help="only the 72-sector XXZ L=24 ground-state batch and the 64-state evolution (one timed run "
                         "each after a warm-up), for kernel traces")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: every GPU box here has one, and there is no multi-GPU path (sectors, scan "
              "points and right-hand sides are independent units that shard without communication). Run with "
              "--gpus 1.", file=sys.stderr)
        sys.exit(2)

    import qmbed_amd as qa

    if args.scan_hv == "walk":
        os.environ["QMBED_SCAN_HV"] = "walk"
    if args.hv_profile:
        hv_profile(qa)
        return
    if args.hv_modes:
        print(json.dumps(hv_modes(qa)))
        return
    if args.sector_forms:
        print(json.dumps(sector_forms(qa)))
        return
    if args.batch_profile:
        L = 24
        keys = [(up, k) for up in (11, 12, 13) for k in range(L)]
        models = [qa.Model({"kind": "spin", "sites": L, "up": up, "momentum": k}, xxz_terms(L, 1.0))
                  for up, k in keys]
        qa.eigsh_batch(models, 1)
        b = qa.eigsh_batch(models, 1)
        for m in models:
            m.close()
        e = expm_batch(qa)
        print(json.dumps({"batch": b["batch"], "expm": e["c4_expm64"]}))
        return

    # ---- setup: family, training scan, network ----
    static, comps = ising_family(L_SCAN)
    t0 = time.perf_counter()
    model = qa.Model({"kind": "spin", "sites": L_SCAN, "pauli": True, "momentum": 0, "parity": 1}, static,
                     components=comps)
    build_s = time.perf_counter() - t0
    grid, perm = scan_grid()
    train = grid[perm[:N_TRAIN]]
    n_bench = BATCH * (args.warmup + args.steps)
    bench_pts = grid[perm[N_TRAIN:N_TRAIN + n_bench]]
    WARM = STARTS[args.start]
    OTHER = STARTS["snapshot" if args.start == "network" else "network"]
    scan = qa.Scan(model, snapshots=SNAPSHOTS, epochs=EPOCHS, learning_rate=2e-3, seed=0, space=SPACE,
                   targets=args.targets)
    t0 = time.perf_counter()
    rep = scan.train(train, tolerance=1e-10)
    train_s = time.perf_counter() - t0

    warm_kw = {"tolerance": TOL, "matvec_precision": "fp32"}
    # the baseline: plain FP64, one eigsh per point (lockstep 0), as a caller of
    # the reference solves a scan; the lockstep batching is measured beside it
    plain_kw = {"tolerance": TOL, "lockstep": 0}
    # the warm-up points, then the timed points: every mode solves all the
    # timed points in one call (lockstep modes: batches of 16 windows; the
    # refill modes hand free slots to the next points, one or 8 at a time)
    warm_pts = bench_pts[:BATCH * args.warmup]
    timed_pts = bench_pts[BATCH * args.warmup:]
    modes = {  # name: (start, solver keys); the first is the headline
        "headline": (WARM, warm_kw),
        "plain_fp64": (False, plain_kw),
        ("snapshot" if args.start == "network" else "network") + "_start_fp32": (OTHER, warm_kw),
        "headline_start_fp64": (WARM, {"tolerance": TOL}),
        "headline_one_by_one": (WARM, dict(warm_kw, lockstep=0)),
        "headline_refill": (WARM, dict(warm_kw, refill=True)),
        "headline_refill8": (WARM, dict(warm_kw, refill_group=8)),
        "plain_fp64_lockstep": (False, {"tolerance": TOL}),
    }
    # warm-up: every mode measured below, on the warm-up points
    if len(warm_pts):
        for start, kw in modes.values():
            scan.solve(warm_pts, warm=start, **kw)

    # ---- timed: the headline start + FP32 schedule; the launch timer
    # brackets every batched scan H·v launch with HIP events on its stream ----
    qa.library().qmbed_hip_synchronize()
    qa.launch_timer(True)
    t0 = time.perf_counter()
    warm = [scan.solve(timed_pts, warm=WARM, **warm_kw)]
    qa.library().qmbed_hip_synchronize()
    dt = time.perf_counter() - t0
    qa.launch_timer(False)
    launches = qa.launch_timer_report()
    value = args.steps * BATCH / dt
    ms_per_step = dt / args.steps * 1e3
    npts = args.steps * BATCH

    # ---- the same points in the other modes ----
    results = {"headline": (warm, dt)}
    for name, (start, kw) in modes.items():
        if name == "headline":
            continue
        t0 = time.perf_counter()
        rs = [scan.solve(timed_pts, warm=start, **kw)]
        results[name] = (rs, time.perf_counter() - t0)
    ep = np.array([x for r in results["plain_fp64"][0] for x in r["energies"]])
    summary = {name: mode_summary(rs, ep, sec, npts) for name, (rs, sec) in results.items()}
    plain_rate = npts / results["plain_fp64"][1]
    max_res = float(np.max([x for r in warm for x in r["residuals"]]))
    fallbacks = int(sum(r["fallbacks"] for r in warm))
    n_grid = len(grid)
    setup_s = build_s + train_s
    amort_warm = n_grid / (setup_s + n_grid / value)
    amort_plain = n_grid / (build_s + n_grid / plain_rate)

    model.set_parameters(hx=1.0, hz=0.25)
    nnz = model.nnz()
    hvmix = hv_roofline(model, "mixed", 4)
    hv32 = hv_roofline(model, np.float32, 4)
    hv64 = hv_roofline(model, np.float64, 8)

    # roofline: the 16-lane FP32 batched scan H·v, the headline's dominant kernel
    scan_kernel = "hv_scan_walk_kernel" if args.scan_hv == "walk" else "hv_scan_batch_kernel"
    key16 = [k for k in launches if k.startswith(scan_kernel + "<float, 16,")]
    rk = key16[0] if key16 else max(launches, key=lambda k: launches[k]["ms_total"]) if launches else None
    if rk is not None:
        width = int(rk.split(",")[1])
        fp32k = "<float" in rk
        algo16 = scan_bytes_per_launch(model.dim, nnz, width, fp32k)
        ach = algo16 / (launches[rk]["ms_mean"] * 1e-3) / 1e9
    else:
        algo16 = ach = None
    # the instrumented kernels of the timed region with the largest totals:
    # [launches, mean ms, total ms, GB/s of algorithmic bytes where stated]
    top = sorted(launches, key=lambda k: -launches[k]["ms_total"])[:6]
    kernel_share = {k.replace("hv_scan_batch_kernel", "scan").replace("hv_scan_mixed_kernel", "scan_mixed")
                    .replace("hv_scan_walk_kernel", "scan_walk"):
                    [launches[k]["launches"], sig(launches[k]["ms_mean"]), sig(launches[k]["ms_total"]),
                     sig(launches[k]["GBs"]) if launches[k]["bytes_total"] else None] for k in top}
    # the same kernel on its own (20 launches at 16 fixed points of the scan grid)
    alone = {}
    for w in (16, 8):
        ms = model.time_scan_batch(bench_pts[:w], fp32=True, iters=20)
        b = scan_bytes_per_launch(model.dim, nnz, w, True)
        alone[f"w{w}"] = [sig(ms), sig(b / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 3)]
    # the other form of the same action (on the fly: orbit walk and state
    # directory per target instead of the compiled stream; same results bit
    # for bit), on a second instance of the family whose first batched
    # launch fixes the form
    other = "stream" if args.scan_hv == "walk" else "walk"
    prev = os.environ.get("QMBED_SCAN_HV")
    os.environ["QMBED_SCAN_HV"] = other
    twin = qa.Model({"kind": "spin", "sites": L_SCAN, "pauli": True, "momentum": 0, "parity": 1}, static,
                    components=comps)
    for w in (16, 8):
        ms = twin.time_scan_batch(bench_pts[:w], fp32=True, iters=20)
        b = scan_bytes_per_launch(model.dim, nnz, w, True)
        alone[f"{other}_w{w}"] = [sig(ms), sig(b / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 3)]
    twin.close()
    if prev is None:
        del os.environ["QMBED_SCAN_HV"]
    else:
        os.environ["QMBED_SCAN_HV"] = prev

    def hv_short(h):
        return [sig(h["ms"]), sig(h["frac"], 3)]

    extras = {
        "scan": {
            "nnz": nnz, "targets": args.targets, "timed_points": npts, "max_residual": sig(max_res, 3),
            "fallbacks": fallbacks,
            "setup_s": [sig(build_s, 3), sig(train_s, 3), sig(rep["solve_seconds"], 3), sig(rep["train_seconds"], 3)],
            "loss": [sig(rep["loss_first"], 3), sig(rep["loss_last"], 3)],
            # per mode [solves/s, mean iterations, max iterations, mean FP32 iterations, largest relative
            # energy difference to plain FP64]
            "modes": {k: [v["sps"], v["it"], v["it_max"], v.get("it32"), v.get("dE")] for k, v in summary.items()},
            "vs_plain_fp64": [sig(value / plain_rate), sig(summary["plain_fp64"]["it"] / summary["headline"]["it"])],
            "full_scan_incl_setup": [sig(amort_warm), sig(amort_plain), sig(amort_warm / amort_plain)],
        },
        "scan_hv_launches": kernel_share,
        "scan_hv_alone": alone,
        "hv_c5": {"mixed": hv_short(hvmix), "fp32": hv_short(hv32), "fp64": hv_short(hv64)},
    }

    if not args.no_extended:
        # XXZ L=24 (N=112,720): 6 lowest eigenpairs, FP64, tolerance 1e-10
        xxz = qa.Model({"kind": "spin", "sites": 24, "up": 12, "momentum": 0}, xxz_terms(24, 1.5))
        xxz.eigsh(6, tolerance=1e-10)
        t0 = time.perf_counter()
        rs = [xxz.eigsh(6, tolerance=1e-10) for _ in range(5)]
        dtx = time.perf_counter() - t0
        extras["xxz_L24_eigsh6"] = {"sps": sig(5 / dtx), "it": rs[-1]["iterations"],
                                    "hv_fp64": hv_short(hv_roofline(xxz, np.float64, 8))}
        xxz.close()
        # J1-J2 L=28 (N=1,432,860): FP32/FP64 H·v and ground state
        big = qa.Model({"kind": "spin", "sites": 28, "up": 14, "momentum": 0}, j1j2_terms(28))
        j = {"hv_fp32": hv_short(hv_roofline(big, np.float32, 4, iters=20)),
             "hv_fp64": hv_short(hv_roofline(big, np.float64, 8, iters=20))}
        e = {}
        for prec in ("fp64", "fp32"):
            t0 = time.perf_counter()
            gs = big.eigsh(1, tolerance=1e-10, matvec_precision=prec)
            e[prec] = gs["eigenvalues"][0]
            j[f"gs_{prec}"] = [sig(time.perf_counter() - t0, 3), gs["iterations"], gs["fp32_iterations"]]
        j["E0"] = e["fp64"]
        j["fp32_rel_err"] = abs(e["fp32"] - e["fp64"]) / abs(e["fp64"])
        extras["j1j2_L28"] = j
        big.close()
        # per sector and mode [ms, fraction of peak] (the lanes per row are in --hv-modes)
        extras["hv_modes_fp64"] = {c: {m: v[:2] if isinstance(v, list) else v for m, v in row.items()}
                                   for c, row in hv_modes(qa).items()}
        sb = sector_batch(qa)["c4_sectors72"]
        extras["c4_sectors72"] = {f"k{k}": [sig(sb[f"eigenpairs_{k}"]["batch_s"], 3),
                                            sig(sb[f"eigenpairs_{k}"]["one_by_one_s"], 3),
                                            sig(sb[f"eigenpairs_{k}"]["max_rel_diff_vs_one_by_one"], 2)]
                                  for k in (1, 6)}
        eb = expm_batch(qa)["c4_expm64"]
        extras["c4_expm64"] = [sig(eb["seconds"], 3), eb["matrix_vector_products"],
                               sig(eb["max_norm_deviation"], 2)]

    if args.details:
        with open(args.details, "w") as f:
            json.dump({name: {"iterations": [x for r in rs for x in r["iterations"]],
                              "energies": [x for r in rs for x in r["energies"]],
                              "ms": [round(1e3 * x, 3) for r in rs for x in r["seconds"]]}
                       for name, (rs, _) in results.items()}, f)

    scan.close()
    line = {
        "metric": METRIC,
        "value": value,
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 matvec, f64 Lanczos/reorth/polish",
        "data": "synthetic (Hamiltonian couplings of the scan grid; seeded training and benchmark points)",
        "config": {"workload": f"Ising chain -ZZ-hx X-hz Z (Pauli), L=26 periodic, k=0, p=+1 (N={model.dim}, "
                               f"nnz={nnz}); 64x64 grid hx in [0.2,2], hz in [0,0.5]; {HEADLINE_TEXT[args.start]}; "
                               "FP32-matvec schedule; residual <= 1e-8*max(|E|,1); step = 16 held-out points, 16 "
                               "Lanczos windows in lockstep; baseline plain FP64, one eigsh per point; batched Hv "
                               + ("on the fly (orbit walk + state directory)" if args.scan_hv == "walk"
                                  else "from the compiled index stream"),
                   "global_batch": BATCH},
        "roofline": {"bound": "hbm", "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": ach / HBM_PEAK_GBS if ach else None,
                     "traffic": SCAN_W16_TRAFFIC_BYTES if rk and rk.startswith("hv_scan_batch_kernel<float, 16,")
                     else None,
                     "kernel": f"{rk}, mean of {launches[rk]['launches'] if rk else 0} launches in the timed region "
                               f"(HIP events on its stream); N(4+8W)+nnz(4+4W) = {algo16} B"},
        "cpu_baseline": None if args.no_cpu_baseline else cpu_baseline(model.dim),
        "extras": extras,
    }
    model.close()
    print(json.dumps(line))


if __name__ == "__main__":
    main()
