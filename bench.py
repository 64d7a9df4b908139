"""Benchmark: the compressed state-vector stage loop on one MI355X.

Workload (default): 32-qubit QFT-with-swaps on a seeded RY*RZ product state,
608 gates, complex128, block_size 20, inner_size 4, eps_rel 1e-3,
eps_abs 1e-7, after the QuantumCircuitOptimizer(block_size=2) pre-pass.
A step resets the compressed store to |0...0> and runs every stage
(decode -> gates -> zero check -> encode) with the state resident in HBM.
value = gates of the specified circuit / step time.

Roofline (HBM bound, 8 TB/s): every stage s moves C_in(s) + C_out(s) + 32 A(s)
algorithmic bytes (the compressed store read and written once, A(s) the
amplitudes of its active groups written and read once at complex128);
achieved = sum over all stages of those bytes / sum over all stages of the
stage time (HIP events around each stage's launches on the library stream).
`traffic` is the HBM traffic of the same launches in one step, measured in
this run by two rocprofv3 --pmc child runs of one step (FETCH_SIZE, then
WRITE_SIZE), each scaled by what the counter reads for a 4 GiB
device-to-device copy in the same child run.  The single-pass DEC->ENC
launches (the dominant kernel) are reported beside it.

Also reported: compression ratio, fidelity against the uncompressed GPU run,
the reference-restated CPU path (the oracle: compressed stage loop and the
uncompressed OpenMP sweep) timed on the host, the HEA-28 adjoint gradient and
the 512-instance QAOA batch.
"""
import argparse
import json
import os
import sys
import time

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(REPO, "compression-based-state-vector-simulator_amd"))
sys.path.insert(0, os.path.join(REPO, "oracle"))

HBM_PEAK_GBPS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def stage_bytes(s):
    """Algorithmic bytes of one compressed stage: C_in + C_out + 32 * A (the
    compressed store read and written once, the active groups' amplitudes
    written and read once at complex128)."""
    return s["bytes_in"] + s["bytes_out"] + 32.0 * s["amplitudes"]


def dominant_launches(stats):
    """Single-pass stages: each is one launch of the fused DEC->ENC wave kernel,
    the kernel that dominates the step."""
    return [s for s in stats if int(s["passes"]) == 1 and s["amplitudes"] > 0]


ROCPROF = "/opt/rocm/bin/rocprofv3"


def _pmc_child(counter, out_dir, args):
    """One rocprofv3 --pmc child run of `bench.py --traffic-probe`; returns
    {dispatch id: (kernel name, counter KiB)} in dispatch order."""
    import csv
    import subprocess

    cmd = ["timeout", "-s", "KILL", "240", ROCPROF, "--pmc", counter, "--kernel-trace", "-d", out_dir, "-o", "tp",
           "--output-format", "csv", "--", sys.executable, os.path.abspath(__file__), "--traffic-probe",
           "--qubits", str(args.qubits), "--inner-size", str(args.inner_size)]
    if args.block_size:
        cmd += ["--block-size", str(args.block_size)]
    r = subprocess.run(cmd, capture_output=True, text=True, cwd="/tmp", env=dict(os.environ, TMPDIR="/tmp"))
    if r.returncode != 0:
        raise RuntimeError(f"rocprofv3 {counter} run failed ({r.returncode}): {r.stderr[-400:]}")
    path = None
    for root, _, files in os.walk(out_dir):
        for f in files:
            if f.endswith("counter_collection.csv"):
                path = os.path.join(root, f)
    if path is None:
        raise RuntimeError(f"rocprofv3 {counter}: no counter_collection.csv under {out_dir}")
    res = {}
    with open(path) as fh:
        for row in csv.DictReader(fh):
            if row["Counter_Name"] != counter:
                continue
            did = int(row["Dispatch_Id"])
            name, v = res.get(did, (row["Kernel_Name"], 0.0))
            res[did] = (name, v + float(row["Counter_Value"]))
    return dict(sorted(res.items()))


def measure_traffic(args):
    """HBM bytes of one step's stage-loop launches from two rocprofv3 --pmc
    child runs (FETCH_SIZE, WRITE_SIZE).  The child (--traffic-probe) runs a
    warmup step, the measured step and a 4 GiB device-to-device copy; each
    counter is scaled by bytes moved / counter reading of that copy.  The
    measured step's launches are those between the last basis fill before the
    copy and the copy."""
    import shutil
    import tempfile

    if not os.path.exists(ROCPROF):
        raise RuntimeError(f"{ROCPROF} not found")
    calib = float(1 << 32)
    out = {"calibration_bytes": calib}
    tmp = tempfile.mkdtemp(prefix="qsim_traffic_", dir="/tmp")
    try:
        for counter in ("FETCH_SIZE", "WRITE_SIZE"):
            rows = _pmc_child(counter, os.path.join(tmp, counter), args)
            ids = list(rows)
            copies = [i for i in ids if "copyBuffer" in rows[i][0]]
            if not copies:
                raise RuntimeError("calibration copy not found in the trace")
            ic = copies[-1]
            fills = [i for i in ids if i < ic and "fill_basis_kernel" in rows[i][0]]
            if not fills:
                raise RuntimeError("step boundary (basis fill) not found in the trace")
            step = [i for i in ids if fills[-1] < i < ic]
            scale = calib / (rows[ic][1] * 1024.0)
            total = sum(rows[i][1] for i in step) * 1024.0 * scale
            dom = [i for i in step if "wave_kernel<1, 1" in rows[i][0]]
            out[counter] = {"scale": scale, "step_bytes": total, "launches": len(step),
                            "dec_enc_launches": len(dom),
                            "dec_enc_bytes": sum(rows[i][1] for i in dom) * 1024.0 * scale}
    finally:
        shutil.rmtree(tmp, ignore_errors=True)
    out["step_bytes"] = out["FETCH_SIZE"]["step_bytes"] + out["WRITE_SIZE"]["step_bytes"]
    n_dom = max(1, out["FETCH_SIZE"]["dec_enc_launches"])
    out["dec_enc_bytes_per_launch"] = (out["FETCH_SIZE"]["dec_enc_bytes"] + out["WRITE_SIZE"]["dec_enc_bytes"]) / n_dom
    return out


def traffic_probe(args):
    """Child run under rocprofv3 --pmc: one warmup step, one measured step,
    then a 4 GiB device-to-device copy (counter calibration)."""
    from qompsim import CompressedQuantumState, QuantumCircuitOptimizer, QuantumState, _lib
    from qompsim import workloads as W

    n = args.qubits
    B = args.block_size if args.block_size else n - 12
    src, dst = QuantumState(28), QuantumState(28)  # the copy's buffers, filled before the steps
    circ = W.qft_on_product(n, seed=2)
    QuantumCircuitOptimizer().optimize(circ, 2)
    cs = CompressedQuantumState(n, B, eps_rel=1e-3, eps_abs=1e-7, verbose=False)
    cs.partition_and_register_circuit(circ, args.inner_size)
    for _ in range(2):  # each step starts with the basis fill of its block 0
        cs.set_zero_state()
        cs.apply_circuit()
        _lib.call("qsim_synchronize")
    del cs
    dst.load(src)
    _lib.call("qsim_synchronize")
    del src, dst


def host_info():
    info = {"cpu_count": os.cpu_count()}
    try:
        info["affinity"] = len(os.sched_getaffinity(0))
    except AttributeError:
        pass
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                info["model"] = line.split(":", 1)[1].strip()
                break
        for line in open("/proc/meminfo"):
            if line.startswith("MemTotal"):
                info["ram_GiB"] = round(int(line.split()[1]) / 2**20, 1)
                break
    except OSError:
        pass
    return info


def cpu_baseline(n_cpu, inner, uncompressed_gates=64):
    """The reference-restated CPU path (the oracle, C++/OpenMP on the host
    cores), measured, not extrapolated: the compressed stage loop of the full
    QFT-n_cpu (the headline circuit at n_cpu qubits, block_size n_cpu - 12 so
    that the 2^12-block layout of the 32-qubit run is kept, the same codec
    contract; SZ3 substituted by the contract codec), and the uncompressed
    sweep (Qulacs' one-sweep-per-gate update) over the circuit's first
    `uncompressed_gates` gates at the same size.  The GPU runs the same
    QFT-n_cpu for a same-size ratio."""
    import numpy as np  # noqa: F401

    import pyoracle as O
    from qompsim import CompressedQuantumState, QuantumCircuitOptimizer, _lib
    from qompsim import workloads as W

    c = W.qft_on_product(n_cpu, seed=2)
    gates_spec = c.get_gate_count()
    QuantumCircuitOptimizer().optimize(c, 2)
    B = n_cpu - 12
    oc = O.OracleCompressedState(n_cpu, B, 1e-3, 1e-7, 1)
    nst = oc.partition_and_register_circuit(c, inner)
    t0 = time.perf_counter()
    oc.apply_circuit()
    t_cs = time.perf_counter() - t0
    del oc
    k = min(uncompressed_gates, c.get_gate_count())
    psi = O.zero_state(n_cpu)
    t0 = time.perf_counter()
    psi = O.apply_circuit(psi, c.gate_list[:k])
    t_sv = time.perf_counter() - t0
    del psi
    # the GPU on the same circuit and size
    cs = CompressedQuantumState(n_cpu, B, eps_rel=1e-3, eps_abs=1e-7, verbose=False)
    cs.partition_and_register_circuit(c, inner)
    cs.apply_circuit()
    _lib.call("qsim_synchronize")
    t0 = time.perf_counter()
    reps = 5
    for _ in range(reps):
        cs.set_zero_state()
        cs.apply_circuit()
    _lib.call("qsim_synchronize")
    t_gpu = (time.perf_counter() - t0) / reps
    del cs
    return {
        "value": gates_spec / t_cs,
        "unit": "gates/s",
        "cores": O.lib.orc_num_threads(),
        "kind": "port",
        "label": "reference-restated, SZ3 substituted",
        "sample": (f"oracle C++/OpenMP compressed stage loop of the full QFT-{n_cpu} ({gates_spec} gates, block_size "
                   f"{B}, inner {inner}, {nst} stages, eps_rel 1e-3, eps_abs 1e-7) in {t_cs:.2f} s on the host; measured "
                   f"at {n_cpu} qubits, not scaled"),
        "qubits": n_cpu,
        "compressed_s": t_cs,
        "uncompressed_gates_per_s": k / t_sv,
        "uncompressed_sample": f"first {k} gates of the optimized QFT-{n_cpu}, one OpenMP sweep per gate, {t_sv:.2f} s",
        "gpu_same_circuit_gates_per_s": gates_spec / t_gpu,
        "gpu_over_cpu_same_circuit": t_cs / t_gpu,
        "host": host_info(),
    }


def adjoint_bench(n=28, layers=8, reps=3):
    """HEA-n adjoint gradient of the TFIM energy on the uncompressed GPU state:
    times the forward pass, the forward+energy evaluation and the whole
    backprop call (forward, lambda = 2 H psi, fused reverse sweep)."""
    import numpy as np

    from qompsim import QuantumState, _lib
    from qompsim import workloads as W
    from qompsim.adjoint import adjoint_gradient

    c = W.hea(n, layers=layers, seed=3)
    obs = W.tfim_observable(n)
    npar = c.get_parameter_count()
    s = QuantumState(n)

    def timed(fn, k):
        fn()
        _lib.call("qsim_synchronize")
        t0 = time.perf_counter()
        for _ in range(k):
            fn()
        _lib.call("qsim_synchronize")
        return (time.perf_counter() - t0) / k

    def fwd():
        s.set_zero_state()
        c.update_quantum_state(s)

    def fwd_energy():
        fwd()
        obs.get_expectation_value(s)

    t_fwd = timed(fwd, reps)
    t_eval = timed(fwd_energy, reps)
    out = {}
    # one working state for every call, as the forward reuses s: the library
    # keeps its lambda and psi buffers with the state, so the timed calls do
    # not allocate (a fresh state per call allocates about 3 x 4 GiB)
    reg = QuantumState(n)

    def adj():
        out["eg"] = adjoint_gradient(c, obs, register=reg)

    t_adj = timed(adj, reps)
    energy, grads = out["eg"]
    th = c.get_parameters()
    h = 1e-5
    fd_err = 0.0
    for i in (0, npar // 3, (2 * npar) // 3, npar - 1):
        vals = []
        for d in (h, -h):
            c.set_parameter(i, th[i] + d)
            fwd()
            vals.append(obs.get_expectation_value(s))
        c.set_parameter(i, th[i])
        fd_err = max(fd_err, abs((vals[0] - vals[1]) / (2 * h) - grads[i]))
    t_fd = 2 * npar * t_eval
    # parameter shift: dE/dtheta = (E(theta + pi/2) - E(theta - pi/2)) / 2
    ps_err = 0.0
    for i in (1, npar // 2, npar - 2):
        vals = []
        for d in (np.pi / 2, -np.pi / 2):
            c.set_parameter(i, th[i] + d)
            fwd()
            vals.append(obs.get_expectation_value(s).real)
        c.set_parameter(i, th[i])
        ps_err = max(ps_err, abs((vals[0] - vals[1]) / 2 - grads[i]))
    del s
    # compressed forward checkpoints every 2 layers
    from qompsim.adjoint import CheckpointedAdjoint

    seg = 2 * (3 * n - 1)
    adj = CheckpointedAdjoint(c, seg, block_size=min(20, n))
    ck = {}

    def ck_fwd():
        adj.forward_with_checkpoints()

    def ck_bwd():
        ck["eg"] = adj.adjoint(obs)

    t_ck_fwd = timed(ck_fwd, reps)
    ck_fwd()
    t_ck_bwd = timed(ck_bwd, reps)
    nck, ck_bytes = adj.checkpoint_info()
    ck_err = float(np.max(np.abs(ck["eg"][1] - grads)))
    del adj
    return {
        "workload": f"HEA-{n}, {layers} layers (RY, RZ per qubit; CZ chain), {npar} parameters; "
                    f"H = sum Z_i Z_i+1 + 0.7 sum X_i ({len(obs.terms)} terms)",
        "energy": energy,
        "forward_ms": t_fwd * 1e3,
        "forward_plus_energy_ms": t_eval * 1e3,
        "adjoint_ms": t_adj * 1e3,
        "adjoint_over_forward": t_adj / t_fwd,
        "adjoint_over_forward_plus_energy": t_adj / t_eval,
        "grads_per_s": npar / t_adj,
        "central_fd_ms": t_fd * 1e3,
        "central_fd_basis": f"2 x {npar} evaluations x measured forward+energy time {t_eval * 1e3:.2f} ms",
        "speedup_vs_central_fd": t_fd / t_adj,
        "max_abs_grad_error_vs_fd": fd_err,
        "max_abs_grad_error_vs_parameter_shift": ps_err,
        "reverse_ms_est": (t_adj - t_fwd) * 1e3,
        "reverse_over_forward_est": (t_adj - t_fwd) / t_fwd,
        "checkpointed": {
            "segment_gates": seg,
            "checkpoints": nck,
            "checkpoint_GiB": ck_bytes / 2**30,
            "codec": "absolute bound 1e-13 per double",
            "forward_with_checkpoints_ms": t_ck_fwd * 1e3,
            "adjoint_ms": t_ck_bwd * 1e3,
            "adjoint_over_forward": t_ck_bwd / t_fwd,
            "max_abs_grad_diff_vs_plain_sweep": ck_err,
        },
    }


def tfim_bench(n=36, B=22, inner=4, steps=12, n_ref=32, store="regions"):
    """Transverse-field Ising Trotter evolution on the compressed store: X on
    qubit 1, then per step RZZ(0.2) on the open chain and RX(0.3) on every
    qubit.  Each step is registered and run as its own circuit so the
    compression ratio and the norm can be read after every step; <Z_0> is
    compared with the same circuit on n_ref qubits, uncompressed (the 12-step
    light cone of Z_0 covers qubits 0..12)."""
    from qompsim import CompressedQuantumState, QuantumCircuit, QuantumCircuitOptimizer, QuantumState, _lib
    from qompsim import workloads as W
    from qompsim.observable import Observable

    cs = CompressedQuantumState(n, B, eps_rel=1e-3, eps_abs=1e-7, store=store, verbose=False)
    raw = 16.0 * 2.0 ** n
    per_step, gates, t_run, oom = [], 0, 0.0, None
    for s in range(steps):
        c = QuantumCircuit(n)
        if s == 0:
            c.add_X_gate(1)
        for i in range(n - 1):
            c.add_multi_Pauli_rotation_gate([i, i + 1], [3, 3], 0.2)
        for q in range(n):
            c.add_RX_gate(q, 0.3)
        QuantumCircuitOptimizer().optimize(c, 2)
        cs.partition_and_register_circuit(c, inner)
        t0 = time.perf_counter()
        try:
            cs.apply_circuit()
        except _lib.QsimError as err:  # HBM exhausted: report the steps that ran
            log(f"tfim-{n} step {s + 1}: {err}")
            oom = str(err)
            break
        _lib.call("qsim_synchronize")
        gates += c.get_gate_count() + (1 if s == 0 else 0)
        dt = time.perf_counter() - t0
        t_run += dt
        hist = cs.get_bytes_history()
        nrm = cs.get_squared_norm()
        st = cs.get_store_bytes()
        per_step.append({"step": s + 1, "ms": dt * 1e3, "stages": cs.get_num_stages(),
                         "compressed_GiB": hist[-1] / 2**30, "peak_GiB": max(hist) / 2**30,
                         "ratio": raw / hist[-1], "norm_drift": nrm - 1.0,
                         "nonnull_blocks": len(cs.get_nonnull_blocks()),
                         "arena_peak_GiB": st["arena_peak"] / 2**30, "descriptors_GiB": st["descriptors"] / 2**30})
        log(f"tfim-{n} step {s + 1}: {dt * 1e3:.0f} ms, {hist[-1] / 2**30:.2f} GiB ({raw / hist[-1]:.1f}x), "
            f"arena peak {st['arena_peak'] / 2**30:.2f} GiB, norm-1 {nrm - 1:.2e}")
    z0 = Observable(n)
    z0.add_operator(1.0, "Z 0")
    zc = cs.get_expectation_value(z0) / cs.get_squared_norm() if oom is None else None
    peak = max(p["peak_GiB"] for p in per_step) if per_step else None
    done = len(per_step)
    del cs
    sv = QuantumState(n_ref)
    cref = W.tfim_trotter(n_ref, steps=done)
    QuantumCircuitOptimizer().optimize(cref, 2)
    cref.update_quantum_state(sv)
    zr = Observable(n_ref)
    zr.add_operator(1.0, "Z 0")
    zu = zr.get_expectation_value(sv)
    del sv
    return {
        "workload": f"TFIM-{n} Trotter, {steps} steps, compressed (B={B}, I={inner}, eps_rel 1e-3, eps_abs 1e-7)",
        "store": store,
        "arena_peak_GiB": max(p["arena_peak_GiB"] for p in per_step) if per_step else None,
        "descriptors_GiB": per_step[-1]["descriptors_GiB"] if per_step else None,
        "steps_completed": done,
        "stopped_by": oom,
        "gates": gates,
        "run_ms": t_run * 1e3,
        "gates_per_s": gates / t_run if t_run > 0 else None,
        "peak_compressed_GiB": peak,
        "min_compression_ratio": raw / (peak * 2**30) if peak else None,
        "per_step": per_step,
        "z0_compressed": zc,
        f"z0_uncompressed_{n_ref}q_after_{done}_steps": zu,
        "z0_abs_diff": abs(zc - zu) if zc is not None else None,
    }


def ghz_bench(n=20, B=12, inner=4, reps=20):
    """GHZ-20 (H, then a CNOT chain) after QuantumCircuitOptimizer(block_size=2):
    19 gates, 3 stages of 256 blocks.  Two codec settings: the CPU path's
    contract (pure absolute bound 1e-7, all-zero threshold 1e-6) and the GPU
    path's (eps_rel 1e-3, eps_abs 1e-7, exact zeros).  Each is timed over
    `reps` runs from |0...0>, with the final null bitmap against the expected
    non-null blocks {0, 15, 127, 255}, the fidelity against the uncompressed
    GPU run, and the oracle's stage loop timed on the host."""
    import numpy as np

    from qompsim import CompressedQuantumState, QuantumCircuitOptimizer, QuantumState, _lib
    from qompsim import workloads as W

    c = W.ghz(n)
    gates_spec = c.get_gate_count()
    QuantumCircuitOptimizer().optimize(c, 2)
    sv = QuantumState(n)
    c.update_quantum_state(sv)
    expect = [0, 15, 127, 255]
    out = {"workload": f"GHZ-{n} (B={B}, I={inner}), {gates_spec} gates, {c.get_gate_count()} after optimize",
           "gates": gates_spec}
    for name, kw in (("cpu_contract", dict(eps_rel=0.0, eps_abs=1e-7, zero_mode="threshold", zero_threshold=1e-6)),
                     ("gpu_contract", dict(eps_rel=1e-3, eps_abs=1e-7, zero_mode="exact"))):
        cs = CompressedQuantumState(n, B, verbose=False, **kw)
        cs.partition_and_register_circuit(c, inner)
        cs.apply_circuit()
        times = []
        for _ in range(reps):
            cs.set_zero_state()
            _lib.call("qsim_synchronize")
            t0 = time.perf_counter()
            cs.apply_circuit()
            _lib.call("qsim_synchronize")
            times.append(time.perf_counter() - t0)
        t = float(np.median(times))
        rec = {"stages": cs.get_num_stages(), "ms": t * 1e3, "gates_per_s": gates_spec / t,
               "nonnull_blocks": cs.get_nonnull_blocks(), "bitmap_ok": cs.get_nonnull_blocks() == expect,
               "fidelity": cs.fidelity_with(sv)}
        try:
            import pyoracle as O

            oc = O.OracleCompressedState(n, B, kw["eps_rel"], kw["eps_abs"], 1, kw["zero_mode"],
                                         kw.get("zero_threshold", 1e-6))
            oc.partition_and_register_circuit(c, inner)
            t0 = time.perf_counter()
            oc.apply_circuit()
            rec["oracle_host_ms"] = (time.perf_counter() - t0) * 1e3
            rec["oracle_bitmap_equal"] = bool(np.array_equal(cs.get_null_bitmap(), oc.null_bitmap()))
        except ImportError:
            pass
        out[name] = rec
        log(f"ghz-{n} {name}: {t * 1e3:.2f} ms, blocks {rec['nonnull_blocks']}, fidelity {rec['fidelity']:.10f}")
    return out


def qaoa_batch_bench(count=512, p=4, reps=3):
    """512 QAOA MaxCut instances on seeded random 3-regular graphs, n
    round-robin over 12..24, p = 4; energies and the 8-parameter adjoint
    gradients of all instances in one graph-captured batch, after the
    autotuner picked the tiling per qubit count.  A sample of the small
    instances is checked against the oracle's backprop."""
    import numpy as np

    from qompsim import workloads as W
    from qompsim.batch import BatchedAdjoint

    insts = W.qaoa_batch(count=count, p=p, seed=4)
    b = BatchedAdjoint()
    t0 = time.perf_counter()
    for n, edges, gm, bt in insts:
        b.add(W.qaoa_maxcut_shared(n, edges, gm, bt), W.maxcut_observable(n, edges), n_params=2 * p)
    t_add = time.perf_counter() - t0
    t0 = time.perf_counter()
    b.tune(reps=2)
    t_tune = time.perf_counter() - t0
    t0 = time.perf_counter()
    e0, g0 = b.run(use_graph=True)  # first call: plan, eager run, capture, one replay
    t_first = time.perf_counter() - t0
    times = []
    for _ in range(reps):
        t0 = time.perf_counter()
        e, g = b.run(use_graph=True)
        times.append(time.perf_counter() - t0)
        assert np.array_equal(e, e0) and np.array_equal(g, g0)
    t = min(times)
    amps = sum(2.0 ** n for n, _, _, _ in insts)
    traffic, n_persistent = b.traffic()
    err = 0.0
    try:
        import pyoracle as O

        for i, (n, edges, gm, bt) in enumerate(insts[:21]):
            if n > 16:
                continue
            er, gr = O.backprop(W.qaoa_maxcut_shared(n, edges, gm, bt), W.maxcut_observable(n, edges))
            err = max(err, abs(e0[i] - er), float(np.max(np.abs(g0[i] - gr[: 2 * p]))))
    except ImportError:
        err = None
    sizes = sorted({n for n, _, _, _ in insts})
    return {
        "workload": f"{count} QAOA MaxCut instances, random 3-regular graphs, n in {sizes}, p={p}; "
                    "energy + 8-parameter adjoint gradient each, one hipGraph launch",
        "instances": count,
        "batch_ms": t * 1e3,
        "instances_per_s": count / t,
        "gradients_per_s": count * 2 * p / t,
        "state_amplitudes_per_s": amps / t,
        "persistent_kernel_instances": n_persistent,
        "roofline": {"bound": "hbm", "achieved": traffic / t / 1e9, "peak": HBM_PEAK_GBPS, "unit": "GB/s",
                     "frac": traffic / t / 1e9 / HBM_PEAK_GBPS, "traffic": None,
                     "algorithmic_bytes_per_batch": traffic,
                     "basis": "per instance of n qubits 16*2^n for |0>, 32*2^n per forward pass and for "
                              "lambda = H psi, 64*2^n per reverse pass (qsim_batch_traffic), over the "
                              "graph replay's wall time"},
        "first_call_ms": t_first * 1e3,
        "tune_s": t_tune,
        "plan_add_s": t_add,
        "tuned": {str(n): b.config(n) for n in sizes},
        "max_abs_error_vs_oracle_n_le_16": err,
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--qubits", type=int, default=32)
    ap.add_argument("--block-size", type=int, default=None)
    ap.add_argument("--inner-size", type=int, default=4)
    ap.add_argument("--no-fidelity", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--no-reorder", action="store_true",
                    help="skip the extra QFT run with commutation-aware stage packing")
    ap.add_argument("--workload", choices=["qft32", "ghz20", "tfim36", "hea28-adjoint", "qaoa-batch"],
                    default="qft32",
                    help="qft32 is the headline bench line; the others print their parity/measurement record")
    ap.add_argument("--cpu-qubits", type=int, default=28)
    ap.add_argument("--no-batch", action="store_true", help="skip the 512-instance QAOA batch record")
    ap.add_argument("--no-traffic", action="store_true",
                    help="skip the two rocprofv3 --pmc child runs that measure roofline.traffic")
    ap.add_argument("--traffic-probe", action="store_true",
                    help="(internal) the child run profiled by the traffic measurement")
    ap.add_argument("--stage-times", action="store_true",
                    help="log every stage of the last timed step (ms, passes, compressed bytes in and out)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="end with one 4 GiB device-to-device copy (counter calibration for profiling)")
    args = ap.parse_args()
    if args.traffic_probe:
        traffic_probe(args)
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (every block group lives on one "
              "MI355X; the multi-GPU exchange is design-only)", file=sys.stderr)
        sys.exit(2)

    if args.workload == "ghz20":
        print(json.dumps({"workload_record": "ghz20", **ghz_bench()}))
        return
    if args.workload == "tfim36":
        print(json.dumps({"workload_record": "tfim36", **tfim_bench(n=args.qubits if args.qubits != 32 else 36)}))
        return
    if args.workload == "qaoa-batch":
        print(json.dumps({"workload_record": "qaoa-batch", **qaoa_batch_bench()}))
        return
    if args.workload == "hea28-adjoint":
        print(json.dumps({"workload_record": "hea28-adjoint", **adjoint_bench(n=args.qubits if args.qubits != 32 else 28)}))
        return

    # HBM traffic of one step, measured by profiled child runs before this
    # process touches the GPU
    traffic, traffic_err = None, None
    if not args.no_traffic:
        try:
            t0 = time.perf_counter()
            traffic = measure_traffic(args)
            log(f"traffic: {traffic['step_bytes'] / 1e9:.2f} GB per step over "
                f"{traffic['FETCH_SIZE']['launches']} launches ({time.perf_counter() - t0:.0f} s)")
        except Exception as err:  # a profiler failure leaves traffic unmeasured
            traffic_err = str(err)[-300:]
            log(f"traffic not measured: {traffic_err}")

    import numpy as np

    from qompsim import CompressedQuantumState, QuantumCircuitOptimizer, QuantumState, _lib
    from qompsim import workloads as W

    n = args.qubits
    B = args.block_size if args.block_size else n - 12
    inner = args.inner_size
    circ = W.qft_on_product(n, seed=2)
    gates_spec = circ.get_gate_count()
    t0 = time.perf_counter()
    QuantumCircuitOptimizer().optimize(circ, 2)
    t_opt = time.perf_counter() - t0
    log(f"QFT-{n}: {gates_spec} gates -> {circ.get_gate_count()} after optimize ({t_opt:.2f} s)")

    cs = CompressedQuantumState(n, B, eps_rel=1e-3, eps_abs=1e-7, verbose=False)
    t0 = time.perf_counter()
    cs.partition_and_register_circuit(circ, inner)
    n_stages = cs.get_num_stages()
    passes = cs.get_stage_passes()
    log(f"{n_stages} stages, passes per stage {passes} ({time.perf_counter() - t0:.2f} s to plan)")

    def step():
        cs.set_zero_state()
        cs.apply_circuit()

    for i in range(args.warmup):
        t0 = time.perf_counter()
        step()
        log(f"warmup {i}: {time.perf_counter() - t0:.3f} s")
    _lib.call("qsim_synchronize")
    t0 = time.perf_counter()
    for i in range(args.steps):
        step()
        log(f"step {i} done")
    _lib.call("qsim_synchronize")
    dt = time.perf_counter() - t0
    ms = dt / args.steps * 1e3
    stats = cs.get_stage_stats()
    gpu_s = sum(s["gpu_seconds"] for s in stats)
    all_bytes = sum(stage_bytes(s) for s in stats)
    achieved = all_bytes / gpu_s / 1e9 if gpu_s > 0 else 0.0
    dom = dominant_launches(stats)
    dom_s = sum(s["gpu_seconds"] for s in dom)
    dom_bytes = sum(stage_bytes(s) for s in dom)
    dom_gbps = dom_bytes / dom_s / 1e9 if dom_s > 0 else 0.0
    compressed_only = sum(s["bytes_in"] + s["bytes_out"] for s in dom) / dom_s / 1e9 if dom_s > 0 else 0.0
    if args.stage_times:
        for i, st in enumerate(stats[-n_stages:]):
            log(f"stage {i:2d}: {st['gpu_seconds'] * 1e3:8.2f} ms, {int(st['passes'])} pass(es), "
                f"{st['bytes_in'] / 1e9:6.2f} -> {st['bytes_out'] / 1e9:6.2f} GB compressed")
    hist = cs.get_bytes_history()
    raw = 16.0 * (1 << n)
    max_bytes = max(hist[-(n_stages + 1):])
    log(f"{ms:.1f} ms/step, stage kernels {gpu_s * 1e3:.1f} ms, {achieved:.0f} GB/s algorithmic over all stages; "
        f"DEC->ENC launches: {len(dom)}, {dom_s * 1e3:.1f} ms, {dom_gbps:.0f} GB/s algorithmic "
        f"({compressed_only:.0f} GB/s compressed bytes only); max compressed {max_bytes / 2**30:.2f} GiB "
        f"({raw / max_bytes:.2f}x)")

    result = {
        "metric": "gates/s & HBM GB/s (compressed, 32q) + adjoint grad/s vs plain, fidelity>=0.9999",
        "value": gates_spec / (ms / 1e3),
        "unit": "gates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "c128",
        "data": "synthetic",
        "config": {"workload": f"QFT-{n} compressed stage loop on a seeded RY*RZ product state (seed 2)",
                   "qubits": n, "block_size": B, "inner_size": inner, "eps_rel": 1e-3, "eps_abs": 1e-7,
                   "guard_bits": int(cs._cfg.guard_bits), "gates": gates_spec,
                   "gates_after_optimize": circ.get_gate_count(), "stages": n_stages,
                   "passes": int(sum(passes))},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBPS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBPS,
                     "traffic": traffic["step_bytes"] if traffic else None},
        "roofline_basis": "all stages of one step: sum of C_in + C_out + 32*A over the stages / sum of the stage "
                          "times (HIP events on the library stream); traffic = measured HBM bytes (FETCH_SIZE + "
                          "WRITE_SIZE, copy-calibrated) of the same step's launches",
        "roofline_dec_enc": {
            "kernel": "wave_kernel<IN_DEC, OUT_ENC> (one launch per single-pass stage)",
            "launches": len(dom), "avg_ms": dom_s * 1e3 / max(1, len(dom)),
            # the stage byte model counts 32 A bytes of decompressed amplitudes
            # that this fused kernel keeps on chip, so it can pass the HBM peak;
            # the measured rate is the counters' traffic over the launch time
            "model_gbps": dom_gbps,
            "compressed_bytes_gbps": compressed_only,
            "traffic_per_launch": traffic["dec_enc_bytes_per_launch"] if traffic else None,
            "measured_gbps": (traffic["dec_enc_bytes_per_launch"] / (dom_s / max(1, len(dom))) / 1e9
                              if traffic and dom_s > 0 else None),
        },
        "stage_kernel_ms": gpu_s * 1e3,
        "compression_ratio": raw / max_bytes,
        "max_sv_GiB": max_bytes / 2**30,
    }
    if traffic:
        result["traffic_detail"] = traffic
    elif traffic_err:
        result["traffic_error"] = traffic_err

    if not args.no_fidelity:
        sv = QuantumState(n)
        t0 = time.perf_counter()
        circ.update_quantum_state(sv)
        _lib.call("qsim_synchronize")
        t_plain = time.perf_counter() - t0
        result["fidelity"] = cs.fidelity_with(sv)
        result["uncompressed_gates_per_s"] = gates_spec / t_plain
        log(f"fidelity {result['fidelity']:.8f}; uncompressed run {t_plain * 1e3:.1f} ms")
    else:
        sv = None

    if not args.no_reorder:
        # the same circuit with commutation-aware stage packing (an extra
        # record; the headline keeps partition_circuit's stages)
        del cs
        cs2 = CompressedQuantumState(n, B, eps_rel=1e-3, eps_abs=1e-7, verbose=False)
        cs2.partition_and_register_circuit(circ, inner, reorder=True)

        def step2():
            cs2.set_zero_state()
            cs2.apply_circuit()

        step2()
        _lib.call("qsim_synchronize")
        t0 = time.perf_counter()
        for _ in range(args.steps):
            step2()
        _lib.call("qsim_synchronize")
        ms2 = (time.perf_counter() - t0) / args.steps * 1e3
        rec = {"stages": cs2.get_num_stages(), "passes": int(sum(cs2.get_stage_passes())), "ms_per_step": ms2,
               "gates_per_s": gates_spec / (ms2 / 1e3)}
        if sv is not None:
            rec["fidelity"] = cs2.fidelity_with(sv)
        result["reordered"] = rec
        log(f"commutation-aware packing: {rec['stages']} stages, {ms2:.1f} ms/step "
            f"({rec['gates_per_s']:.0f} gates/s)" + (f", fidelity {rec['fidelity']:.8f}" if sv is not None else ""))
        del cs2
    if sv is not None:
        del sv

    if not args.no_adjoint:
        result["adjoint"] = adjoint_bench()
        a = result["adjoint"]
        log(f"adjoint HEA-28: forward {a['forward_ms']:.1f} ms, backprop {a['adjoint_ms']:.1f} ms "
            f"({a['adjoint_over_forward']:.2f}x forward, {a['speedup_vs_central_fd']:.0f}x central FD), "
            f"max |g - g_fd| {a['max_abs_grad_error_vs_fd']:.2e}")

    if not args.no_batch:
        result["batch"] = qaoa_batch_bench(reps=3)
        b = result["batch"]
        log(f"QAOA batch: {b['batch_ms']:.0f} ms, {b['instances_per_s']:.0f} instances/s, "
            f"{b['gradients_per_s']:.0f} gradients/s, {b['roofline']['achieved']:.0f} GB/s algorithmic "
            f"({b['persistent_kernel_instances']} instances in the persistent kernel)")

    if not args.no_cpu:
        ncpu = min(args.cpu_qubits, n)
        result["cpu_baseline"] = cpu_baseline(ncpu, inner)
        cb = result["cpu_baseline"]
        log(f"cpu baseline: {cb['sample']}; uncompressed {cb['uncompressed_gates_per_s']:.1f} gates/s "
            f"({cb['uncompressed_sample']}); GPU on the same circuit {cb['gpu_same_circuit_gates_per_s']:.0f} gates/s")
    if args.copy_calibration:
        # a plain device-to-device copy of 2^28 complex128 (4 GiB) for scaling
        # the FETCH_SIZE / WRITE_SIZE counters
        src, dst = QuantumState(28), QuantumState(28)
        dst.load(src)
        _lib.call("qsim_synchronize")
        log("copy calibration: 4294967296 bytes device-to-device")
        del src, dst
    print(json.dumps(result))


if __name__ == "__main__":
    main()
