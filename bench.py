"""Benchmark: SOAP power spectrum + position gradients, atoms/s on one MI355X.

Workload (the batched configuration the headline metric is quoted on): 4096
synthetic structures with U{8..256} atoms, half periodic (cubic / triclinic,
0.08-0.12 atoms/A^3) and half open clusters in 15 A boxes, 4 atomic types;
cutoff 5 A (ShiftedCosine 0.5), Gaussian density 0.3 A, GTO basis with 8 radial
functions (JSON max_radial 7) and max_angular 6; FP32 device kernels.

One step is one pass of the hot path over the whole batch with positions
already resident in GPU memory: neighbour list, per-center spherical expansion,
power spectrum values and position-gradient rows. The batch is processed in
chunks of structures whose outputs fit in HBM (the full position-gradient
output is several hundred GB); each chunk's output is produced in device memory
and left there for the consumer.

Prints one JSON line: metric, value (atoms/s), ms_per_step, the roofline of the
dominant kernel (HIP-event time, algorithmic bytes, measured HBM traffic) and the
oracle's CPU rate on a bounded sample. --workload selects the other workloads
(LODE, bulk cell, learned radial integral, backward, spherical expansion by pair),
each printing its own line.
"""

import argparse
import ctypes
import gc
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "atoms/s SOAP power spectrum+grads: fast form vs plain FP32 @<=1e-4 rel err"
HBM_PEAK_GBS = 8000.0

HYPERS = {
    "cutoff": {"radius": 5.0, "smoothing": {"type": "ShiftedCosine", "width": 0.5}},
    "density": {"type": "Gaussian", "width": 0.3},
    "basis": {"type": "TensorProduct", "max_angular": 6, "radial": {"type": "Gto", "max_radial": 7}},
}


def log(*args):
    print(*args, file=sys.stderr, flush=True)


def make_chunks(systems, max_atoms):
    chunks, current, atoms = [], [], 0
    for k, (pos, t, cell) in enumerate(systems):
        if current and atoms + len(t) > max_atoms:
            chunks.append(current)
            current, atoms = [], 0
        current.append(k)
        atoms += len(t)
    if current:
        chunks.append(current)
    return chunks


def device_batch(systems, indices):
    """Positions of one chunk resident in device memory."""
    from rascaline_amd.device import DeviceSystems

    return DeviceSystems([systems[k] for k in indices])


def run_chunk(calc, batch, stream_ptr, grads):
    from rascaline_amd import _c_lib as C

    out = C.featomic_amd_device_output_t()
    C.check(C.lib().featomic_amd_compute_device(calc._as_parameter_, ctypes.byref(batch.struct), grads, 1,
                                                ctypes.c_void_p(stream_ptr), ctypes.byref(out)))
    return out


def measured_traffic(precision, gradient_rows, form="plain"):
    """HBM bytes per step of the gradient kernel measured with rocprofv3 --pmc
    (FETCH_SIZE + WRITE_SIZE, calibrated on a plain copy) for this exact workload,
    from the committed profile summary; None for any other workload."""
    path = os.path.join(ROOT, "profiles", "soap_ps_gradients_4096_structures_fp32.md")
    try:
        with open(path) as fd:
            for line in fd:
                if line.startswith("traffic_per_step:"):
                    fields = dict(item.split("=") for item in line.split()[1:])
                    if (int(fields["precision"]) == precision and int(fields["gradient_rows"]) == gradient_rows
                            and fields.get("form", "plain") == form):
                        return float(fields["fetch_bytes"]) + float(fields["write_bytes"])
    except OSError:
        pass
    return None


def cpu_baseline(systems, target_seconds=25.0):
    """The FP64 C restatement of the reference algorithm (oracle) on the host cores,
    over a bounded sample of the same structures."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import featomic_oracle as orc

    try:
        cores = len(os.sched_getaffinity(0))
    except AttributeError:
        cores = os.cpu_count() or 1
    osys = [orc.System(p, t, c) for (p, t, c) in systems]
    # calibrate on a few structures, then size the sample to ~target_seconds
    n = 2 * cores
    t0 = time.perf_counter()
    orc.soap_ps_baseline(osys[:n], HYPERS, n_threads=cores)
    dt = time.perf_counter() - t0
    per_structure = max(dt / n, 1e-4)
    n = int(min(len(osys), max(n, target_seconds / per_structure)))
    t0 = time.perf_counter()
    rows, checksum = orc.soap_ps_baseline(osys[:n], HYPERS, n_threads=cores)
    dt = time.perf_counter() - t0
    atoms = sum(len(s.types) for s in osys[:n])
    return {
        "value": atoms / dt,
        "unit": "atoms/s",
        "cores": cores,
        "kind": "port",
        "cpu_model": cpu_model(),
        "sample": f"first {n} of the benchmark's structures ({atoms} atoms, {rows} gradient rows), FP64, "
                  f"{cores} threads, {dt:.1f} s",
    }


def cpu_model():
    try:
        with open("/proc/cpuinfo") as fd:
            for line in fd:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def accuracy_sample(systems, n_structures=24):
    """The FP64 oracle's power spectrum + position gradients of the first structures of the
    workload, the reference for the accuracy of each benchmarked form."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import featomic_oracle as orc

    return orc.soap_power_spectrum([orc.System(p, t, c) for (p, t, c) in systems[:n_structures]], HYPERS,
                                   ["positions"])


def form_accuracy(out, reference, rel=1e-4):
    """Error of a device result (the first chunk, whose first systems are the reference's)
    against the FP64 oracle: per block, the smallest absolute floor (as a fraction of the
    block's largest element) with which every element passes |a - b| <= max(floor, rel
    max(|a|, |b|)), and the largest block-relative error max|a - b| / max|b|."""
    import numpy as np

    from rascaline_amd.device import DeviceResult

    dev = DeviceResult(out)
    keys = [tuple(k) for k in dev.keys]
    worst = {"values_floor": 0.0, "gradients_floor": 0.0, "values_block_relative": 0.0,
             "gradients_block_relative": 0.0}
    for key, block in zip(reference["keys"], reference["blocks"]):
        b = keys.index(tuple(key))
        ns = len(block["samples"])
        g = block["gradients"]["positions"]
        nr = len(g["samples"])
        assert np.array_equal(dev.sample_labels(b)[:ns].cpu().numpy(), block["samples"])
        assert np.array_equal(dev.gradient_labels(b)[:nr].cpu().numpy(), g["samples"])
        pairs = (("values", dev.values(b)[:ns], block["values"]), ("gradients", dev.gradients(b)[:nr], g["values"]))
        for name, got, ref in pairs:
            got = got.double().cpu().numpy().reshape(ref.shape)
            if not ref.size:
                continue
            scale = np.abs(ref).max()
            diff = np.abs(got - ref)
            fail = diff > rel * np.maximum(np.abs(got), np.abs(ref))
            floor = float(diff[fail].max() / scale) if np.any(fail) else 0.0
            worst[name + "_floor"] = max(worst[name + "_floor"], floor)
            worst[name + "_block_relative"] = max(worst[name + "_block_relative"], float(diff.max() / scale))
    return worst


def end_to_end(systems, precision, n_structures=64, repeats=2):
    """atoms/s through featomic_calculator_compute: host systems in, host TensorMap (float64
    arrays) out, on the first structures of the workload (the whole workload's float64
    gradients would not fit in host memory)."""
    import rascaline_amd as ra

    sample = systems[:n_structures]
    calc = ra.SoapPowerSpectrum(**HYPERS)
    calc.set_precision(precision)
    ra_systems = [ra.System(t, p, c) for (p, t, c) in sample]
    calc.compute(ra_systems, gradients=["positions"])
    best = float("inf")
    for _ in range(repeats):
        t0 = time.perf_counter()
        calc.compute(ra_systems, gradients=["positions"])
        best = min(best, time.perf_counter() - t0)
    atoms = sum(len(t) for (_, t, _) in sample)
    return {"value": atoms / best, "unit": "atoms/s",
            "sample": f"first {len(sample)} structures ({atoms} atoms), featomic_calculator_compute with host "
                      f"systems and a host TensorMap (float64 arrays), FP{precision} kernels, best of {repeats}"}


def main():
    parser = argparse.ArgumentParser()
    parser.add_argument("--gpus", type=int, default=1)
    parser.add_argument("--steps", type=int, default=3)
    parser.add_argument("--warmup", type=int, default=1)
    parser.add_argument("--structures", type=int, default=4096)
    parser.add_argument("--chunk-atoms", type=int, default=56000)
    parser.add_argument("--precision", type=int, default=32)
    parser.add_argument("--streams", type=int, default=4,
                        help="calculators on their own streams that consecutive chunks alternate between")
    parser.add_argument("--no-cpu-baseline", action="store_true")
    parser.add_argument("--forms", choices=["both", "plain"], default="both",
                        help="plain: leave out the fast form (profiling runs)")
    parser.add_argument("--no-graph", action="store_true", help="leave out the batch-graph timing")
    parser.add_argument("--no-end-to-end", action="store_true", help="leave out the featomic_calculator_compute line")
    parser.add_argument("--gradient-window", type=int, default=None,
                        help="use this gradient window shape instead of the autotuner's choice")
    parser.add_argument("--workload", choices=["soap", "lode", "bulk", "surrogate", "vjp", "lode-vjp", "by-pair"],
                        default="soap",
                        help="soap: the headline (SOAP power spectrum + gradients); lode: LODE spherical "
                             "expansion + gradients on periodic cells of 500-1500 atoms; bulk: power spectrum + "
                             "gradients of one 20 000-atom cell (12 radial functions, l <= 8, 5 types); "
                             "surrogate: spherical expansion + gradients of one 100 000-atom cell with the "
                             "spline and with the learned radial integral; vjp: the 4096 structures' power "
                             "spectrum values followed by the vector-Jacobian product of their position gradients "
                             "; lode-vjp: the LODE cells' values followed by the vector-Jacobian product of "
                             "their position gradients; by-pair: the spherical expansion by pair + position "
                             "gradients of the first 16 structures (each its own JSON line)")
    args = parser.parse_args()

    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); the sharding over "
              "several GPUs is a design, not built", file=sys.stderr)
        return 2

    if args.workload == "lode":
        return lode_main(args)
    if args.workload == "bulk":
        return bulk_main(args)
    if args.workload == "surrogate":
        return surrogate_main(args)
    if args.workload == "vjp":
        return vjp_main(args)
    if args.workload == "lode-vjp":
        return lode_vjp_main(args)
    if args.workload == "by-pair":
        return by_pair_main(args)

    import torch

    import rascaline_amd as ra
    from rascaline_amd.datasets import batched_uneven

    t0 = time.perf_counter()
    systems = batched_uneven(seed=0, n_structures=args.structures)
    n_atoms = sum(len(s[1]) for s in systems)
    log(f"generated {len(systems)} structures, {n_atoms} atoms in {time.perf_counter() - t0:.1f} s")

    chunks = make_chunks(systems, args.chunk_atoms)
    batches = [device_batch(systems, c) for c in chunks]
    stream = torch.cuda.current_stream().cuda_stream
    grads = (ctypes.c_char_p * 1)(b"positions")
    log(f"{len(batches)} chunks of up to {args.chunk_atoms} atoms")
    from rascaline_amd import _c_lib as C

    def make_calculator(fast):
        calc = ra.SoapPowerSpectrum(**HYPERS)
        calc.set_precision(args.precision)
        calc.set_fast_form(fast)
        if fast:
            # the fast form's learned radial integral, trained on the GPU (outside the timed
            # steps); the spline stays in use when the network misses the acceptance tolerance
            report = calc.train_radial_surrogate()
            surrogate.update({"accepted": bool(report["accepted"]), "max_value_error": report["max_value_error"],
                              "max_derivative_error": report["max_derivative_error"],
                              "train_ms": report["train_ms"]})
        return calc

    surrogate = {}

    def measure(fast, window=None):
        """Warmup, then exactly `steps` timed steps over all chunks (bracketed by device
        synchronisation), then one pass with the library's hipEvent timings for the
        roofline of the gradient kernel."""
        calc = make_calculator(fast)
        name = "fast" if fast else "plain"
        # launch shapes chosen by the autotuner on the first chunk (outside the timed steps)
        if window is not None or args.gradient_window is not None:
            # (the vector-ALU form also takes the vector-ALU expansion kernel, 8 entries per chunk)
            calc.tuning = {"expansion_entries": 8 if window is not None else 0,
                           "gradient_window": window if window is not None else args.gradient_window}
            tuning_report = {"chosen": calc.tuning}
            if window is not None:
                name += f" (gradient window {window})"
        else:
            tuning_report = calc.autotune(batches[0], repeats=2)
            log(f"{name}: autotuned {tuning_report['chosen']}")
        # consecutive chunks alternate between calculators on their own streams, so that
        # one chunk's neighbour list, expansion and values overlap the previous chunk's
        # write-bound gradient rows (each calculator's buffers are reused every
        # `streams` chunks, in its stream's order)
        lanes = [(calc, stream)]
        for _ in range(args.streams - 1):
            extra = make_calculator(fast)
            extra.tuning = calc.tuning
            lanes.append((extra, torch.cuda.Stream().cuda_stream))

        def step():
            for k, b in enumerate(batches):
                c, s = lanes[k % len(lanes)]
                run_chunk(c, b, s, grads)

        for w in range(args.warmup):
            step()
            torch.cuda.synchronize()
            log(f"{name}: warmup step {w} done")
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            step()
        torch.cuda.synchronize()
        elapsed = time.perf_counter() - t0
        ms = 1e3 * elapsed / args.steps
        rate = n_atoms * args.steps / elapsed
        log(f"{name}: timed {args.steps} steps: {ms:.1f} ms/step, {rate:.3e} atoms/s")

        # accuracy of this form on the first structures of the first chunk
        out = run_chunk(calc, batches[0], stream, grads)
        torch.cuda.synchronize()
        accuracy = form_accuracy(out, reference)
        log(f"{name}: accuracy against the FP64 oracle {accuracy}")

        total_bytes, kernel_ms, device_ms = 0, 0.0, 0.0
        rows_total, samples_total = 0, 0
        for b in batches:
            out = run_chunk(calc, b, stream, grads)
            t_ms, k_ms, k_bytes = ctypes.c_double(), ctypes.c_double(), ctypes.c_int64()
            C.check(C.lib().featomic_amd_last_timings(calc._as_parameter_, ctypes.byref(t_ms), ctypes.byref(k_ms),
                                                      ctypes.byref(k_bytes)))
            rows_total += sum(out.n_gradient_rows[k] for k in range(out.n_blocks))
            samples_total += sum(out.n_samples[k] for k in range(out.n_blocks))
            # algorithmic bytes of the gradient-row kernel, as counted by the library:
            # gradient values + labels written, per-center coefficients and self gradients
            # read once, neighbour entries read once
            total_bytes += k_bytes.value
            kernel_ms += k_ms.value
            device_ms += t_ms.value
        achieved = total_bytes / (kernel_ms * 1e-3) / 1e9
        roofline = {
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": measured_traffic(args.precision, rows_total, "vector-alu" if window is not None else
                                        "fast" if fast else "plain"),
            "kernel": (("ps_gradient_rank2_kernel (power-spectrum position-gradient rows; the tiled kernel "
                        "takes the centers with a run longer than a window)")
                       if tuning_report["chosen"]["gradient_window"] in (-1, 11, 12) and args.precision == 32 else
                       "ps_gradient_tiled_kernel (power-spectrum position-gradient rows"
                       + (", bfloat16 hi+lo MFMA contraction)" if fast else ")")),
            "kernel_ms_per_step": kernel_ms,
            "device_ms_per_step": device_ms,
            "algorithmic_bytes_per_step": total_bytes,
            "tuning": tuning_report["chosen"],
        }
        roofline["accuracy"] = accuracy
        return ms, rate, roofline, rows_total, samples_total

    t0 = time.perf_counter()
    reference = accuracy_sample(systems)
    log(f"FP64 oracle of the first 24 structures for the accuracy checks: {time.perf_counter() - t0:.1f} s")
    # the line reports the faster form among those within the metric's bar (1e-4 relative with
    # the floor the FP32 arithmetic itself needs, 3e-5 of the block's scale) and lists both
    fast_available = args.precision == 32 and args.forms == "both"
    plain_ms, plain_rate, plain_roofline, rows_total, samples_total = measure(False)
    forms = {"plain": (plain_ms, plain_rate, plain_roofline)}
    if fast_available:
        fast_ms, fast_rate, fast_roofline, _, _ = measure(True)
        forms["fast"] = (fast_ms, fast_rate, fast_roofline)
        # the plain form with its expansion and gradient rows on the vector ALUs (the tiled
        # kernels: 8 entries per expansion chunk, gradient windows of 4 runs and 8 images),
        # for the gain of the matrix-core kernels
        vec_ms, vec_rate, vec_roofline, _, _ = measure(False, window=1)
        forms["plain, vector-ALU kernels"] = (vec_ms, vec_rate, vec_roofline)

    def measure_graph(fast, tuning):
        """The same pass captured into one hipGraph (featomic_amd_batch_graph_*): the
        record pass and the capture happen outside the timed steps; each timed step is
        one graph launch that recomputes every chunk (neighbour lists to gradient rows)
        on `streams` calculators."""
        from rascaline_amd.device import BatchGraph

        gc.collect()
        torch.cuda.synchronize()
        lanes = []
        for _ in range(args.streams):
            c = make_calculator(fast)
            c.tuning = tuning
            lanes.append(c)
        t0 = time.perf_counter()
        graph = BatchGraph(lanes, batches, ["positions"])
        log(f"{'fast' if fast else 'plain'} graph: recorded and captured {len(batches)} chunks on {len(lanes)} "
            f"calculators in {time.perf_counter() - t0:.1f} s")
        try:
            for _ in range(args.warmup):
                graph.launch()
            graph.check()
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for _ in range(args.steps):
                graph.launch()
            torch.cuda.synchronize()
            elapsed = time.perf_counter() - t0
            graph.check()
        finally:
            graph.close()
        ms = 1e3 * elapsed / args.steps
        log(f"{'fast' if fast else 'plain'} graph: timed {args.steps} launches: {ms:.1f} ms/step, "
            f"{n_atoms * args.steps / elapsed:.3e} atoms/s")
        return ms, n_atoms * args.steps / elapsed

    def within_bar(rl):
        acc = rl["accuracy"]
        return acc["values_floor"] <= 3e-5 and acc["gradients_floor"] <= 3e-5

    eligible = [f for f in ("plain", "fast") if f in forms and within_bar(forms[f][2])] or ["plain"]
    best = max(eligible, key=lambda f: forms[f][1])
    graph_ms, graph_rate = (None, 0.0) if args.no_graph else measure_graph(best == "fast", forms[best][2]["tuning"])
    ms_per_step, value, roofline = forms[best]
    launch = "eager"
    if graph_rate > value:
        ms_per_step, value, launch = graph_ms, graph_rate, "graph"
    e2e = None if args.no_end_to_end else end_to_end(systems, args.precision)
    log(f"end to end: {e2e}")

    baseline = None
    if not args.no_cpu_baseline:
        baseline = cpu_baseline(systems)
        log(f"cpu baseline: {baseline}")

    result = {
        "metric": METRIC,
        "value": value,
        "unit": "atoms/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if args.precision == 32 else "f64",
        "data": "synthetic",
        "config": {
            "workload": f"{len(systems)} uneven structures (8-256 atoms; half periodic cubic/triclinic at "
                        f"0.08-0.12 A^-3, half clusters in 15 A boxes; 4 types), SOAP power spectrum + position "
                        f"gradients, rc 5.0, n 8, l<=6",
            "form": ("FP64" if args.precision == 64 else
                     "plain FP32: spline radial integral, expansion and gradient rows (rank-2 contraction) "
                     "on MFMA" if best == "plain" else
                     "fast: learned radial integral, expansion on MFMA, power-spectrum values on MFMA in "
                     "bfloat16 hi+lo, gradient rows by the rank-2 MFMA contraction"),
            "atoms": n_atoms,
            "launch": ("one hipGraph launch per step (featomic_amd_batch_graph)" if launch == "graph" else
                       "chunk by chunk (featomic_amd_compute_device)"),
            "chunks": len(batches),
            "streams": args.streams,
            "power_spectrum_samples": samples_total,
            "gradient_rows": rows_total,
        },
        "roofline": roofline,
        "cpu_baseline": baseline,
        "end_to_end": e2e,
        "forms": {name: {"value": rate, "unit": "atoms/s", "ms_per_step": ms, "roofline_frac": rl["frac"],
                         "kernel_ms_per_step": rl["kernel_ms_per_step"], "tuning": rl["tuning"],
                         "accuracy": rl["accuracy"]}
                  for name, (ms, rate, rl) in forms.items()},
    }
    if graph_ms is not None:
        result["forms"][best]["graph"] = {"ms_per_step": graph_ms, "value": graph_rate, "unit": "atoms/s"}
    if "fast" in forms:
        result["forms"]["fast"]["speedup"] = forms["fast"][1] / forms["plain"][1]
        result["forms"]["fast"]["radial_surrogate"] = surrogate
        vec = forms["plain, vector-ALU kernels"][1]
        for name in ("plain", "fast"):
            result["forms"][name]["speedup_vs_vector_alu_kernels"] = forms[name][1] / vec
    print(json.dumps(result), flush=True)
    return 0


LODE_HYPERS = {
    "density": {"type": "SmearedPowerLaw", "smearing": 1.0, "exponent": 1},
    "basis": {"type": "TensorProduct", "max_angular": 4, "radial": {"type": "Gto", "max_radial": 5, "radius": 5.0}},
}
FP32_PEAK_TFLOPS = 157.3


def lode_cpu_baseline(target_seconds=15.0):
    """The FP64 oracle's LODE values + position gradients on the host cores, on
    small cells of the same density (the cost grows as N^2 n_k, so the sample
    states its size)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import featomic_oracle as orc
    from rascaline_amd.datasets import lode_cells

    try:
        cores = len(os.sched_getaffinity(0))
    except AttributeError:
        cores = os.cpu_count() or 1
    done, atoms, t_total = 0, 0, 0.0
    cells = lode_cells(seed=7, n_structures=64, n_min=96, n_max=128)
    for pos, t, cell in cells:
        t0 = time.perf_counter()
        orc.lode_spherical_expansion([orc.System(pos, t, cell)], LODE_HYPERS, ["positions"])
        t_total += time.perf_counter() - t0
        done += 1
        atoms += len(t)
        if t_total > target_seconds:
            break
    return {
        "value": atoms / t_total,
        "unit": "atoms/s",
        "cores": 1,
        "kind": "port",
        "sample": f"{done} periodic triclinic cells of 96-128 atoms at 0.1 A^-3 ({atoms} atoms), FP64 oracle, "
                  f"one thread, {t_total:.1f} s (the GPU cells hold 500-1500 atoms; cost grows as N^2 n_k)",
    }


def by_pair_cpu_baseline(systems, target_seconds=10.0):
    """The oracle's spherical expansion by pair + position gradients (one C call per
    pair, one thread) on the first structures of the workload."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import featomic_oracle as orc

    done, samples, t_total = 0, 0, 0.0
    while t_total < target_seconds and done < len(systems):
        t0 = time.perf_counter()
        tm = orc.soap_spherical_expansion_by_pair([orc.System(*systems[done])], HYPERS, ["positions"])
        t_total += time.perf_counter() - t0
        samples += sum(len(b["samples"]) for b, k in zip(tm["blocks"], tm["keys"]) if k[0] == 0)
        done += 1
    return {"value": samples / t_total, "unit": "pair samples/s", "cores": 1, "kind": "port",
            "sample": f"the first {done} structures of the workload, FP64, 1 thread, {t_total:.1f} s"}


def by_pair_main(args):
    """SphericalExpansionByPair + position gradients (host TensorMap: the GPU writes
    every pair's rows, the host lays out the labels and copies the blocks out)."""
    import torch

    import rascaline_amd as ra
    from rascaline_amd import _c_lib as C
    from rascaline_amd.datasets import batched_uneven

    n = args.structures if args.structures != 4096 else 16
    systems = batched_uneven(seed=0, n_structures=n)
    n_atoms = sum(len(s[1]) for s in systems)
    calc = ra.SphericalExpansionByPair(**HYPERS)
    calc.set_precision(args.precision)
    ra_systems = [ra.System(t, p, c) for p, t, c in systems]
    for w in range(args.warmup):
        tm = calc.compute(ra_systems, gradients=["positions"])
        log(f"by-pair: warmup step {w} done")
        del tm
    samples = 0
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    kernel_ms, device_ms, kernel_bytes = 0.0, 0.0, 0
    for _ in range(args.steps):
        tm = calc.compute(ra_systems, gradients=["positions"])
        t_ms, k_ms, k_bytes = ctypes.c_double(), ctypes.c_double(), ctypes.c_int64()
        C.check(C.lib().featomic_amd_last_timings(calc._as_parameter_, ctypes.byref(t_ms), ctypes.byref(k_ms),
                                                  ctypes.byref(k_bytes)))
        kernel_ms += k_ms.value / args.steps
        device_ms += t_ms.value / args.steps
        kernel_bytes = k_bytes.value
        samples = sum(len(b.samples) for b, k in zip(tm.blocks(), tm.keys.values) if k[0] == 0)
        del tm
    elapsed = time.perf_counter() - t0
    ms = 1e3 * elapsed / args.steps
    rate = samples * args.steps / elapsed
    log(f"by-pair: {ms:.1f} ms/step, {rate:.3e} pair samples/s, by_pair_kernel {kernel_ms:.2f} ms")
    result = {
        "metric": "pair samples/s SOAP spherical expansion by pair + position gradients (host TensorMap)",
        "value": rate,
        "unit": "pair samples/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if args.precision == 32 else "f64",
        "data": "synthetic",
        "config": {"workload": f"the first {n} of the 4096 uneven structures, spherical expansion by pair + "
                               "position gradients, rc 5.0, n 8, l<=6",
                   "atoms": n_atoms, "pair_samples": samples},
        "roofline": hbm_roofline(kernel_ms, kernel_bytes, device_ms,
                                 "by_pair_kernel (every pair's values and position-gradient rows)"),
        "cpu_baseline": None if args.no_cpu_baseline else by_pair_cpu_baseline(systems),
    }
    print(json.dumps(result), flush=True)
    return 0


def lode_main(args):
    """LODE spherical expansion + position gradients, FP32 with FP64 structure-factor
    phases, one cell per device call (outputs of up to 4 GB per cell stay on the GPU)."""
    import torch

    import rascaline_amd as ra
    from rascaline_amd import _c_lib as C
    from rascaline_amd.datasets import lode_cells

    n = args.structures if args.structures != 4096 else 16
    t0 = time.perf_counter()
    systems = lode_cells(seed=0, n_structures=n)
    n_atoms = sum(len(s[1]) for s in systems)
    log(f"generated {len(systems)} LODE cells, {n_atoms} atoms in {time.perf_counter() - t0:.1f} s")
    batches = [device_batch(systems, [k]) for k in range(len(systems))]
    stream = torch.cuda.current_stream().cuda_stream
    grads = (ctypes.c_char_p * 1)(b"positions")
    calc = ra.LodeSphericalExpansion(**LODE_HYPERS)
    calc.set_precision(args.precision)

    def step():
        for b in batches:
            run_chunk(calc, b, stream, grads)

    for w in range(args.warmup):
        step()
        torch.cuda.synchronize()
        log(f"lode: warmup step {w} done")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    ms = 1e3 * elapsed / args.steps
    rate = n_atoms * args.steps / elapsed
    log(f"lode: timed {args.steps} steps: {ms:.1f} ms/step, {rate:.3e} atoms/s")

    flops, kernel_ms, device_ms, rows = 0, 0.0, 0.0, 0
    for b in batches:
        out = run_chunk(calc, b, stream, grads)
        t_ms, k_ms, k_flops = ctypes.c_double(), ctypes.c_double(), ctypes.c_int64()
        C.check(C.lib().featomic_amd_last_timings(calc._as_parameter_, ctypes.byref(t_ms), ctypes.byref(k_ms),
                                                  ctypes.byref(k_flops)))
        flops += k_flops.value
        kernel_ms += k_ms.value
        device_ms += t_ms.value
        rows += sum(out.n_gradient_rows[k] for k in range(out.n_blocks))
    achieved = flops / (kernel_ms * 1e-3) / 1e12
    baseline = None if args.no_cpu_baseline else lode_cpu_baseline()
    result = {
        "metric": "atoms/s LODE spherical expansion + position gradients",
        "value": rate,
        "unit": "atoms/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 (structure-factor phases f64)" if args.precision == 32 else "f64",
        "data": "synthetic",
        "config": {
            "workload": f"{len(systems)} periodic triclinic cells of 500-1500 atoms at 0.1 A^-3, 3 types, LODE p=1, "
                        "smearing 1.0, 6 radial functions, l<=4, radius 5.0, k_cutoff 1.2 pi",
            "atoms": n_atoms,
            "gradient_rows": rows,
        },
        "roofline": {
            "bound": "mfma",
            "achieved": achieved,
            "peak": FP32_PEAK_TFLOPS,
            "unit": "TFLOP/s",
            "frac": achieved / FP32_PEAK_TFLOPS,
            "traffic": None,
            "kernel": "lode_pair_mfma_kernel (pair gradient rows on FP32 MFMA: A from the phases, K = k-vectors)",
            "kernel_ms_per_step": kernel_ms,
            "device_ms_per_step": device_ms,
            "algorithmic_flops_per_step": flops,
        },
        "cpu_baseline": baseline,
    }
    print(json.dumps(result), flush=True)
    return 0

BULK_HYPERS = {
    "cutoff": {"radius": 6.0, "smoothing": {"type": "ShiftedCosine", "width": 0.5}},
    "density": {"type": "Gaussian", "width": 0.3},
    "basis": {"type": "TensorProduct", "max_angular": 8, "radial": {"type": "Gto", "max_radial": 11}},
}


def time_forms(forms, batch, n_atoms, args, grads, label):
    """For each (name, calculator): warmup, `steps` timed device computations bracketed
    by synchronisation, then one pass with the library's hipEvent timings of the
    dominant output kernel. Returns {name: (ms, atoms/s, kernel ms, kernel bytes, device ms)}."""
    import torch

    from rascaline_amd import _c_lib as C

    stream = torch.cuda.current_stream().cuda_stream
    out = {}
    for name, calc in forms:
        for w in range(args.warmup):
            run_chunk(calc, batch, stream, grads)
            torch.cuda.synchronize()
            log(f"{label} {name}: warmup step {w} done")
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            run_chunk(calc, batch, stream, grads)
        torch.cuda.synchronize()
        elapsed = time.perf_counter() - t0
        run_chunk(calc, batch, stream, grads)
        t_ms, k_ms, k_bytes = ctypes.c_double(), ctypes.c_double(), ctypes.c_int64()
        C.check(C.lib().featomic_amd_last_timings(calc._as_parameter_, ctypes.byref(t_ms), ctypes.byref(k_ms),
                                                  ctypes.byref(k_bytes)))
        out[name] = (1e3 * elapsed / args.steps, n_atoms * args.steps / elapsed, k_ms.value, k_bytes.value, t_ms.value)
        log(f"{label} {name}: {out[name][0]:.1f} ms/step, {out[name][1]:.3e} atoms/s, output kernel "
            f"{k_ms.value:.1f} ms")
    return out


def hbm_roofline(kernel_ms, kernel_bytes, device_ms, kernel):
    achieved = kernel_bytes / (kernel_ms * 1e-3) / 1e9
    return {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS, "traffic": None, "kernel": kernel, "kernel_ms_per_step": kernel_ms,
            "device_ms_per_step": device_ms, "algorithmic_bytes_per_step": kernel_bytes}


def cores_available():
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:
        return os.cpu_count() or 1


def bulk_cpu_baseline(target_seconds=15.0):
    """The oracle's threaded power spectrum + gradients on small cells of the bulk
    density and hyper-parameters (one system per thread)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import featomic_oracle as orc
    from rascaline_amd.datasets import bulk_cell

    cores = cores_available()
    cells = [orc.System(*bulk_cell(seed=100 + k, n_atoms=200, density=0.1, types=(1, 6, 7, 8, 9))[0])
             for k in range(2 * cores)]
    done, atoms, t_total = 0, 0, 0.0
    while t_total < target_seconds:
        t0 = time.perf_counter()
        orc.soap_ps_baseline(cells, BULK_HYPERS, n_threads=cores)
        t_total += time.perf_counter() - t0
        done += len(cells)
        atoms += sum(len(c.types) for c in cells)
    return {"value": atoms / t_total, "unit": "atoms/s", "cores": cores, "kind": "port",
            "sample": f"{done} periodic cells of 200 atoms at 0.1 A^-3 (same hyper-parameters, 5 types), FP64, "
                      f"{cores} threads, {t_total:.1f} s"}


def bulk_main(args):
    """Power spectrum + position gradients of one 20 000-atom periodic cell (5 types,
    rc 6, 12 radial functions, l <= 8), plain and fast FP32 forms, output (about
    145 GB) kept in device memory."""
    import rascaline_amd as ra
    from rascaline_amd.datasets import bulk_cell

    systems = bulk_cell(seed=0, n_atoms=20000, density=0.1, types=(1, 6, 7, 8, 9))
    n_atoms = len(systems[0][1])
    batch = device_batch(systems, [0])
    grads = (ctypes.c_char_p * 1)(b"positions")
    res, tuned = {}, {}
    for fast in (False, True):
        # one calculator (and its ~145 GB of outputs) at a time in device memory
        calc = ra.SoapPowerSpectrum(**BULK_HYPERS)
        calc.set_precision(32)
        calc.set_fast_form(fast)
        tuned[("fast" if fast else "plain")] = calc.autotune(batch, repeats=1)
        log(f"bulk: autotuned {tuned['fast' if fast else 'plain']}")
        res.update(time_forms([("fast" if fast else "plain", calc)], batch, n_atoms, args, grads, "bulk"))
        del calc
        gc.collect()
    best = max(res, key=lambda k: res[k][1])
    ms, rate, k_ms, k_bytes, d_ms = res[best]
    baseline = None if args.no_cpu_baseline else bulk_cpu_baseline()
    result = {
        "metric": "atoms/s SOAP power spectrum + position gradients, bulk cell",
        "value": rate, "unit": "atoms/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": ms, "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "one periodic cubic cell of 20000 atoms at 0.1 A^-3, 5 types, SOAP power spectrum + "
                               "position gradients, rc 6.0, n 12, l<=8", "form": best, "atoms": n_atoms},
        "roofline": hbm_roofline(k_ms, k_bytes, d_ms, "power-spectrum position-gradient rows"),
        "cpu_baseline": baseline,
        "forms": {k: {"value": v[1], "ms_per_step": v[0], "kernel_ms_per_step": v[2], "autotune": tuned[k]}
                  for k, v in res.items()},
    }
    print(json.dumps(result), flush=True)
    return 0


def surrogate_main(args):
    """Spherical expansion + position gradients of one 100 000-atom periodic cell
    (4 types, rc 6, 12 radial functions, l <= 8, FP32) with the spline radial
    integral and with the learned one (trained here on 1e6 distances)."""
    import rascaline_amd as ra
    from rascaline_amd.datasets import bulk_cell

    systems = bulk_cell(seed=0, n_atoms=100000, density=0.1, types=(1, 6, 7, 8))
    n_atoms = len(systems[0][1])
    batch = device_batch(systems, [0])
    grads = (ctypes.c_char_p * 1)(b"positions")
    calc = ra.SphericalExpansion(**BULK_HYPERS)
    calc.set_precision(32)
    report = calc.train_radial_surrogate()
    log(f"surrogate: accepted {report['accepted']}, value error {report['max_value_error']:.2e}, derivative error "
        f"{report['max_derivative_error']:.2e}, {report['train_ms']:.0f} ms")
    if not report["accepted"]:
        raise RuntimeError("the radial network was not accepted by the spline judge")
    res, tuned = {}, {}
    for name, use in (("spline", False), ("surrogate", True)):
        calc.use_radial_surrogate(use)
        tuned[name] = calc.autotune(batch, repeats=1)["chosen"]
        log(f"surrogate {name}: autotuned {tuned[name]}")
        res.update(time_forms([(name, calc)], batch, n_atoms, args, grads, "surrogate"))
    ms, rate, k_ms, k_bytes, d_ms = res["surrogate"]
    result = {
        "metric": "atoms/s SOAP spherical expansion + position gradients, learned radial integral",
        "value": rate, "unit": "atoms/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": ms, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32 (network fp16 hi+lo, fp32 accumulation)", "data": "synthetic",
        "config": {"workload": "one periodic cubic cell of 100000 atoms at 0.1 A^-3, 4 types, SOAP spherical "
                               "expansion + position gradients, rc 6.0, n 12, l<=8, MLP 1-64-64-216",
                   "atoms": n_atoms, "train_ms": report["train_ms"],
                   "judged_errors": [report["max_value_error"], report["max_derivative_error"]]},
        "roofline": hbm_roofline(k_ms, k_bytes, d_ms, "spherical-expansion gradient rows"),
        "cpu_baseline": None,
        "forms": {k: {"value": v[1], "ms_per_step": v[0], "kernel_ms_per_step": v[2]} for k, v in res.items()},
    }
    print(json.dumps(result), flush=True)
    return 0



def vjp_main(args):
    """Power spectrum values of the 4096 structures, then the vector-Jacobian product
    of their position gradients with a synthetic dA/dX (the backward pass a model
    trained on forces runs), without the gradient rows: a forward and a backward
    call per chunk, chunks alternating over `streams` calculators."""
    import torch

    import rascaline_amd as ra
    from rascaline_amd.datasets import batched_uneven
    from rascaline_amd.device import compute_device, positions_vjp

    systems = batched_uneven(seed=0, n_structures=args.structures)
    n_atoms = sum(len(s[1]) for s in systems)
    chunks = make_chunks(systems, args.chunk_atoms)
    batches = [device_batch(systems, c) for c in chunks]
    dtype = torch.float32 if args.precision == 32 else torch.float64
    lanes = []
    for _ in range(args.streams):
        calc = ra.SoapPowerSpectrum(**HYPERS)
        calc.set_precision(args.precision)
        lanes.append((calc, torch.cuda.Stream()))
    # dA/dX per chunk in the layout of its values (synthetic, fixed seed)
    gen = torch.Generator(device="cuda")
    gen.manual_seed(0)
    weights = []
    for b in batches:
        dev = compute_device(lanes[0][0], b)
        weights.append(torch.randn(dev.value_size, dtype=dtype, device="cuda", generator=gen))
    torch.cuda.synchronize()
    outs = [None] * len(batches)

    def step():
        for k, b in enumerate(batches):
            calc, s = lanes[k % len(lanes)]
            compute_device(calc, b, stream=s)
            outs[k] = positions_vjp(calc, b, weights[k], stream=s)

    for w in range(args.warmup):
        step()
        torch.cuda.synchronize()
        log(f"vjp: warmup step {w} done")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    ms = 1e3 * elapsed / args.steps
    rate = n_atoms * args.steps / elapsed
    log(f"vjp: timed {args.steps} steps: {ms:.1f} ms/step, {rate:.3e} atoms/s")
    # one calculator, chunk after chunk: the library's hipEvent timing of the VJP kernels
    from rascaline_amd import _c_lib as C

    calc = lanes[0][0]
    stream = torch.cuda.current_stream()
    k_ms, k_bytes, d_ms = 0.0, 0, 0.0
    for k, b in enumerate(batches):
        positions_vjp(calc, b, weights[k], stream=stream)
        t, km, kb = ctypes.c_double(), ctypes.c_double(), ctypes.c_int64()
        C.check(C.lib().featomic_amd_last_timings(calc._as_parameter_, ctypes.byref(t), ctypes.byref(km),
                                                  ctypes.byref(kb)))
        k_ms += km.value
        k_bytes += kb.value
        d_ms += t.value
    result = {
        "metric": "atoms/s SOAP power spectrum + vector-Jacobian product of the position gradients",
        "value": rate, "unit": "atoms/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": ms, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32" if args.precision == 32 else "f64", "data": "synthetic",
        "config": {"workload": f"{len(systems)} uneven structures as the default workload, power spectrum values "
                               "(forward) + dA/dr from a synthetic dA/dX (backward), no gradient rows",
                   "atoms": n_atoms, "chunks": len(batches), "streams": args.streams},
        "roofline": hbm_roofline(k_ms, k_bytes, d_ms, "vjp_pair_kernel + vjp_gather_kernel (dA/dr per entry and atom)"),
        "cpu_baseline": None,
    }
    print(json.dumps(result), flush=True)
    return 0


def lode_vjp_main(args):
    """LODE values of the 16 cells of the LODE workload, then the vector-Jacobian
    product of their position gradients with a synthetic dA/dX, without the N^2
    gradient rows (a forward and a backward call per cell)."""
    import torch

    import rascaline_amd as ra
    from rascaline_amd import _c_lib as C
    from rascaline_amd.datasets import lode_cells
    from rascaline_amd.device import compute_device, positions_vjp

    n = args.structures if args.structures != 4096 else 16
    systems = lode_cells(seed=0, n_structures=n)
    n_atoms = sum(len(s[1]) for s in systems)
    batches = [device_batch(systems, [k]) for k in range(len(systems))]
    calc = ra.LodeSphericalExpansion(**LODE_HYPERS)
    calc.set_precision(args.precision)
    dtype = torch.float32 if args.precision == 32 else torch.float64
    gen = torch.Generator(device="cuda")
    gen.manual_seed(0)
    weights = []
    for b in batches:
        dev = compute_device(calc, b)
        weights.append(torch.randn(dev.value_size, dtype=dtype, device="cuda", generator=gen))
    torch.cuda.synchronize()

    def step():
        for k, b in enumerate(batches):
            compute_device(calc, b)
            positions_vjp(calc, b, weights[k])

    for w in range(args.warmup):
        step()
        torch.cuda.synchronize()
        log(f"lode-vjp: warmup step {w} done")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    ms = 1e3 * elapsed / args.steps
    rate = n_atoms * args.steps / elapsed
    log(f"lode-vjp: timed {args.steps} steps: {ms:.1f} ms/step, {rate:.3e} atoms/s")
    k_ms, k_bytes, d_ms = 0.0, 0, 0.0
    for k, b in enumerate(batches):
        positions_vjp(calc, b, weights[k])
        t, km, kb = ctypes.c_double(), ctypes.c_double(), ctypes.c_int64()
        C.check(C.lib().featomic_amd_last_timings(calc._as_parameter_, ctypes.byref(t), ctypes.byref(km),
                                                  ctypes.byref(kb)))
        k_ms += km.value
        k_bytes += kb.value
        d_ms += t.value
    result = {
        "metric": "atoms/s LODE spherical expansion + vector-Jacobian product of the position gradients",
        "value": rate, "unit": "atoms/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": ms, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32 (structure-factor phases f64)" if args.precision == 32 else "f64", "data": "synthetic",
        "config": {"workload": f"{len(systems)} periodic triclinic cells of 500-1500 atoms as the LODE workload, "
                               "values (forward) + dA/dr from a synthetic dA/dX (backward), no gradient rows",
                   "atoms": n_atoms},
        "roofline": {"bound": "fp32", "achieved": k_bytes / (k_ms * 1e-3) / 1e12, "peak": FP32_PEAK_TFLOPS,
                     "unit": "TFLOP/s", "frac": k_bytes / (k_ms * 1e-3) / 1e12 / FP32_PEAK_TFLOPS, "traffic": None,
                     "kernel": "lode_vjp_v/u/atom kernels (V = Bg W products counted)", "kernel_ms_per_step": k_ms,
                     "device_ms_per_step": d_ms, "algorithmic_flops_per_step": k_bytes},
        "cpu_baseline": None,
    }
    print(json.dumps(result), flush=True)
    return 0

if __name__ == "__main__":
    sys.exit(main())
