"""Benchmark of the LUT simulation chain on one MI355X.

Headline workload: the batched forward of 128 events whose segment counts are log-uniform in
[2k, 400k] (9.16 M segments in a 30x30x30 cm drift volume, module0 geometry, 45x45x1950 LUT, 4
neighbour pixels, 5x5 diffusion bins, noise off).  A step = one hipGraph replay of the whole
chain (quench -> drift -> pixel rows -> waveforms -> self-trigger -> hits) over device-resident
segments.  Prints one JSON line.

Also measured in the same run (extra keys): the 200k-segment event alone, forward+backward
(7-parameter gradient) on it, the plain-forward finite-difference gradient baseline (15
forwards), the batched forward+backward of 32 uneven events in one hipGraph replay against the
per-event forward-then-finite-difference gradient, a 200-step Adam fit of the 7 parameters on
those 32 events (steps/s, parameter recovery) with the hit loss and a 100-step one with a loss
on the waveforms, the per-output parity of the 1000-segment event against the FP64 oracle (with
the threshold-tie fraction), and the CPU baselines (the float32 NumPy restatement of the
reference algorithm: whole events of the batch, one per process on every core; and one thread on
a bounded sample).

``--sdtw-only`` runs only the soft-DTW leg (value and gradient between two hit ADC sequences).
``--prob-fee-only`` runs only the probabilistic readout on the 200k-segment event's rows.
``--chop-only`` runs only the track-chopping leg (350k track rows -> ~3.5M segment rows).
``--batched-bwd-only`` runs only the 32-event forward+backward graph replays (for kernel profiles).
``--copy-calibrate`` only runs a plain 256 MiB device copy five times: profiled with the same
``rocprofv3 --pmc FETCH_SIZE`` / ``WRITE_SIZE`` passes it gives the counter-to-bytes factor
used to turn the waveform kernel's counters into HBM bytes per launch.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "larnd-sim-jax_amd"))

HBM_PEAK_GBS = 8000.0
LUT_BYTES = 45 * 45 * 1950 * 4


def algorithmic_bytes_fwd(n_seg: int, rows: int, H: int = 10, nout: int = 2000) -> int:
    """Compulsory HBM bytes of one forward: segment inputs (x, y, z, dE, dEdx, eventID as f32),
    the raw LUT read once, the waveform rows written, and per row the pixel id plus the ADC and
    tick of each of the H hit slots."""
    return 24 * n_seg + LUT_BYTES + rows * nout * 4 + rows * (8 + H * 8)


def algorithmic_bytes_fwdbwd(n_seg: int, rows: int, H: int = 10, nout: int = 2000) -> int:
    """Compulsory HBM bytes of a forward+backward: the forward's, the inputs and the raw LUT read
    again by the backward, d loss / d wfs written and read once, and the 7 partial gradients of each
    row workgroup (two adjoint passes)."""
    return algorithmic_bytes_fwd(n_seg, rows, H, nout) + 24 * n_seg + LUT_BYTES + 2 * rows * nout * 4 + 2 * rows * 56


# On-chip work of the algorithm per segment (n neighbours, nb x nb diffusion bins, L-sample windows):
# scatter-adds into pixel x tick rows and LUT gathers, counted as LDS bytes (8 per add, 4 per gather).
def onchip_per_segment(n: int = 4, nb: int = 5, L: int = 150) -> dict:
    B, K = nb * nb, (2 * n + 1) ** 2
    scatter = (B + K) * (2 * L + 2)
    gathers = 3 * B * L + K * L + 3 * B + 3 * K
    return dict(scatter_adds=scatter, lut_gathers=gathers, lds_bytes=8 * scatter + 4 * gathers)


def copy_gbs(dev, nbytes: int = 1 << 30, reps: int = 5) -> float:
    """Achievable HBM bandwidth: a plain device copy of nbytes (read + write counted)."""
    import torch

    src = torch.ones(nbytes // 4, dtype=torch.float32, device=dev)
    dst = torch.empty_like(src)
    for _ in range(2):
        dst.copy_(src)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        dst.copy_(src)
    e1.record()
    torch.cuda.synchronize()
    t = e0.elapsed_time(e1) * 1e-3 / reps
    del src, dst
    return 2 * nbytes / t / 1e9


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def cpu_baseline(params, lut, tracks, fields, n_sample: int, what: str):
    """The float32 NumPy restatement of the reference algorithm (materialise + bincount), single
    thread, on the first n_sample segments; LUT template build excluded (as on the GPU)."""
    from threadpoolctl import threadpool_limits

    sys.path.insert(0, ROOT)
    import numpy as np

    from oracle import oracle as O

    sub = tracks[:n_sample]
    with threadpool_limits(limits=1):
        tables = O.build_templates(params, lut)
        t0 = time.perf_counter()
        wfs, U, q = O.simulate_wfs(params, lut, sub, fields, tables=tables, dtype=np.float32)
        O.simulate_stochastic(params, wfs, U)
        dt = time.perf_counter() - t0
    return dict(value=n_sample / dt, unit="segments/s", cores=1, kind="port",
                sample=f"first {n_sample} segments of {what}, float32 NumPy restatement "
                       f"(materialised contributions + bincount), 1 thread, {dt:.1f} s",
                host_cpu_count=os.cpu_count())


def _cpu_event_worker(job):
    """One whole event through the float32 NumPy restatement in a fresh single-threaded process."""
    params_kw, lut, tracks, fields = job
    from threadpoolctl import threadpool_limits

    sys.path.insert(0, ROOT)
    import numpy as np

    from larndsim_amd import consts
    from oracle import oracle as O

    params = consts.default_params(**params_kw)

    with threadpool_limits(limits=1):
        tables = O.build_templates(params, lut)
        t0 = time.perf_counter()
        wfs, U, q = O.simulate_wfs(params, lut, tracks, fields, tables=tables, dtype=np.float32)
        O.simulate_stochastic(params, wfs, U)
        return time.perf_counter() - t0


def cpu_baseline_events(params_kw, lut, tracks, fields, sizes, procs: int = 16, target: int = 10_000):
    """The same restatement on every core: the `procs` events of the batch closest in size to
    `target` segments, one whole event per process (events are independent, so this is the
    reference's own parallel decomposition).  The processes run together; each builds its LUT
    templates, then times its event; the rate is the segments over the slowest event's time."""
    import multiprocessing as mp

    import numpy as np

    ev = tracks[:, fields.index("eventID")].astype(np.int64)
    picks = np.argsort(np.abs(np.asarray(sizes) - target), kind="stable")[:procs]
    jobs = []
    for e in picks:
        sub = tracks[ev == e].copy()
        sub[:, fields.index("eventID")] = 0
        jobs.append((params_kw, lut, sub, fields))
    n = int(sum(len(j[2]) for j in jobs))
    ctx = mp.get_context("spawn")
    with ctx.Pool(len(jobs)) as pool:
        t0 = time.perf_counter()
        per = pool.map(_cpu_event_worker, jobs, chunksize=1)
        wall = time.perf_counter() - t0
    return dict(value=n / max(per), unit="segments/s", cores=len(jobs), kind="port",
                sample=f"{len(jobs)} whole events of the 128-event batch ({n} segments, {min(len(j[2]) for j in jobs)}-"
                       f"{max(len(j[2]) for j in jobs)} each), one event per single-threaded process, float32 NumPy "
                       f"restatement (materialised contributions + bincount); slowest event {max(per):.1f} s "
                       f"(call {wall:.1f} s with process start and template builds)",
                host_cpu_count=os.cpu_count())


def parity_c1(simr, params, lut):
    """Forward outputs of the 1000-segment event against the FP64 oracle, per output: pixel ids and
    hit ticks (with the threshold-tie fraction), waveforms (relative to each row's peak, and
    elementwise where |ref| > 1e-3 of the row peak), ADC on rows whose ticks agree, hit x/y/z."""
    import numpy as np
    import torch

    sys.path.insert(0, ROOT)
    from larndsim_amd import synthetic
    from larndsim_amd.sim import hits_dict
    from oracle import oracle as O

    tr, f = synthetic.workload("c1")
    st = simr.forward(params, tr, f, n_events=1)
    torch.cuda.synchronize()
    ref = O.simulate(params, lut, tr, f)
    U = ref["unique_pixels"]
    valid = U >= 0
    P = st.max_rows
    pids = st.out["pixel_ids"][:P].cpu().numpy()
    w = st.out["wfs"][:P].cpu().numpy().astype(np.float64)
    rw = ref["wfs"][valid]
    peak = np.maximum(np.abs(rw).max(axis=1, keepdims=True), 1e-30)
    big = np.abs(rw) > 1e-3 * peak
    ticks = st.out["ticks"][:P].cpu().numpy()
    rt = ref["grid"]["ticks"][valid].astype(np.int64)
    same = ticks == rt
    ok = same.all(axis=1)
    adc = st.out["adc"][:P].cpu().numpy()
    ra = ref["grid"]["adcs"][valid]
    h = hits_dict(st.out)
    nh = int(h["adcs"].numel())
    coords = {}
    if nh == len(ref["adcs"]) and same.all():
        for k in ("pixel_x", "pixel_y", "pixel_z"):
            coords[k] = float(np.abs(h[k].cpu().numpy() - ref[k]).max())
    return dict(workload="1 event, 1000 segments (10 tracks x 40 cm)", reference="FP64 NumPy oracle",
                pixel_ids_exact=bool(np.array_equal(pids, U[valid])), rows=P,
                wfs_max_err_rel_to_row_peak=float((np.abs(w - rw) / peak).max()),
                wfs_max_rel_err_above_milli_peak=float((np.abs(w - rw)[big] / np.abs(rw[big])).max()),
                tick_slots=int(same.size), tick_mismatch=int((~same).sum()),
                threshold_tie_fraction=float((~same).mean()),
                adc_max_rel_err=float((np.abs(adc[ok] - ra[ok]) / np.maximum(np.abs(ra[ok]), 1e-6)).max()),
                hits=nh, hits_oracle=int(len(ref["adcs"])), hit_coordinate_max_abs_err_cm=coords)


# The sources the forward kernels (k_rows_fwd, k_records and what they call) are built from, with
# the library file that launches them; the other kernels' headers do not change them.
FORWARD_SOURCES = ("device.h", "forward.h", "scan_sort.h", "lsim.hip")


def kernel_source_stamp() -> str:
    """sha256 (first 16 hex) of the forward kernels' sources: ties a recorded counter measurement
    to the kernels it measured."""
    import hashlib

    d = os.path.join(ROOT, "larnd-sim-jax_amd", "csrc")
    h = hashlib.sha256()
    for name in FORWARD_SOURCES:
        with open(os.path.join(d, name), "rb") as f:
            h.update(name.encode() + b"\0" + f.read())
    return h.hexdigest()[:16]


TRAFFIC_JSON = os.path.join(ROOT, "profiles", "rows_fwd_traffic_128_events.json")


def _counter_rows(d, kernel_sub):
    """(dispatch id, counter value) of every dispatch of a kernel whose name contains kernel_sub,
    from the rocprofv3 counter-collection CSVs under d (one counter per run, summed over its
    instances per dispatch)."""
    import csv
    import glob

    vals, names = {}, {}
    for path in glob.glob(os.path.join(d, "**", "*counter_collection.csv"), recursive=True):
        with open(path) as fh:
            for row in csv.DictReader(fh):
                if kernel_sub in row["Kernel_Name"]:
                    k = int(row["Dispatch_Id"])
                    vals[k] = vals.get(k, 0.0) + float(row["Counter_Value"])
                    names[k] = row["Kernel_Name"].split("(")[0].replace("void ", "").replace("lsim::", "")
    _counter_rows.names = [names[k] for k in sorted(vals)]
    return [vals[k] for k in sorted(vals)]


def record_traffic(d, kernel_ms):
    """Turn rocprofv3 FETCH_SIZE / WRITE_SIZE passes (d/fetch, d/write for the waveform kernel of
    the 128-event batch; d/copy_fetch, d/copy_write for `--copy-calibrate`) into HBM bytes per
    launch, stamped with the kernel sources and the kernel time, into TRAFFIC_JSON."""
    fetch = _counter_rows(os.path.join(d, "fetch"), "k_rows_fwd")[1:]  # the first, cold dispatch excluded
    kernels = sorted(set(_counter_rows.names[1:]))
    write = _counter_rows(os.path.join(d, "write"), "k_rows_fwd")[1:]
    cf = _counter_rows(os.path.join(d, "copy_fetch"), "copyBuffer")
    cw = _counter_rows(os.path.join(d, "copy_write"), "copyBuffer")
    copy_bytes = 256 << 20
    ff = sum(cf) / len(cf) * 1024 / copy_bytes
    wf = sum(cw) / len(cw) * 1024 / copy_bytes
    fk, wk = sum(fetch) / len(fetch), sum(write) / len(write)
    rec = dict(kernel=", ".join(kernels), workload="128 events, 9 160 058 segments, 457 979 pixel rows",
               dispatches_used=f"{len(fetch)} warm dispatches of `bench.py --profile-batch` (the first excluded)",
               fetch_size_kib=fk, write_size_kib=wk, copy_fetch_kib=sum(cf) / len(cf),
               copy_write_kib=sum(cw) / len(cw), copy_bytes=copy_bytes, fetch_factor=ff, write_factor=wf,
               fetch_bytes=fk * 1024 / ff, write_bytes=wk * 1024 / wf,
               bytes_per_launch=fk * 1024 / ff + wk * 1024 / wf, kernel_ms=kernel_ms,
               source_stamp=kernel_source_stamp())
    with open(TRAFFIC_JSON, "w") as fh:
        json.dump(rec, fh, indent=1)
    return rec


COUNTERS_JSON = os.path.join(ROOT, "profiles", "kernel_counters_128_events.json")
COUNTER_KERNELS = ("k_rows_fwd", "k_records")


def _counter_means(d):
    """{kernel: {counter: mean per dispatch}} over the rocprofv3 counter CSVs under d, the first
    dispatch of each kernel excluded (cold caches), counter instances summed per dispatch."""
    import csv
    import glob

    acc = {}
    for path in glob.glob(os.path.join(d, "**", "*counter_collection.csv"), recursive=True):
        with open(path) as fh:
            for row in csv.DictReader(fh):
                name = next((k for k in COUNTER_KERNELS if k in row["Kernel_Name"]), None)
                if name is None:
                    continue
                per = acc.setdefault(name, {}).setdefault(row["Counter_Name"], {})
                k = (path, int(row["Dispatch_Id"]))
                per[k] = per.get(k, 0.0) + float(row["Counter_Value"])
    out = {}
    for name, cs in acc.items():
        out[name] = {}
        for c, v in cs.items():
            keys = sorted(v)
            vals = [v[k] for k in keys[1:]] or [v[k] for k in keys]
            out[name][c] = sum(vals) / len(vals)
    return out


def record_counters(d, cu_num=256, xcds=8):
    """Derived issue, LDS and L2 figures of the forward kernels from rocprofv3 --pmc passes of
    `bench.py --profile-batch` under d, stamped with the kernel sources, into COUNTERS_JSON.
    GRBM_GUI_ACTIVE is summed over the XCD instances, so the kernel's cycles are that sum / xcds;
    a wave64 VALU instruction occupies a SIMD for 4 cycles (16 lanes)."""
    raw = _counter_means(d)
    derived = {}
    for name, c in raw.items():
        cyc = c.get("GRBM_GUI_ACTIVE", 0.0) / xcds
        r = {}
        if cyc > 0 and "SQ_ACTIVE_INST_VALU" in c:
            r["valu_busy"] = c["SQ_ACTIVE_INST_VALU"] / cu_num / cyc
        if cyc > 0 and "SQ_INSTS_VALU" in c:
            r["valu_issue_frac"] = c["SQ_INSTS_VALU"] * 4 / (cu_num * 4 * cyc)
        if c.get("SQ_ACTIVE_INST_VALU"):
            r["valu_lane_util"] = c.get("SQ_THREAD_CYCLES_VALU", 0.0) / (c["SQ_ACTIVE_INST_VALU"] * 64)
        if c.get("SQ_LDS_IDX_ACTIVE"):
            free = max(c["SQ_LDS_IDX_ACTIVE"] - c["SQ_LDS_BANK_CONFLICT"], 1.0)
            r["lds_bank_conflict_rate"] = c["SQ_LDS_BANK_CONFLICT"] / free
        if cyc > 0 and "SQ_LDS_IDX_ACTIVE" in c:
            r["lds_busy"] = c["SQ_LDS_IDX_ACTIVE"] / cu_num / cyc
        if c.get("TCC_REQ"):
            r["l2_hit_rate"] = c["TCC_HIT"] / c["TCC_REQ"]
        if c.get("SQ_WAVE_CYCLES"):
            r["wave_wait_frac"] = c.get("SQ_WAIT_ANY", 0.0) / c["SQ_WAVE_CYCLES"]
        r["kernel_cycles_per_xcd"] = cyc
        derived[name] = r
    rec = dict(workload="128 events, 9 160 058 segments, 457 979 pixel rows (`bench.py --profile-batch`)",
               derived=derived, raw=raw, source_stamp=kernel_source_stamp(),
               definitions=dict(valu_busy="SQ_ACTIVE_INST_VALU / CU_NUM / kernel cycles",
                                valu_issue_frac="VALU instructions x 4 cycles / (4 SIMDs x CUs x kernel cycles)",
                                valu_lane_util="SQ_THREAD_CYCLES_VALU / (SQ_ACTIVE_INST_VALU x 64)",
                                lds_bank_conflict_rate="SQ_LDS_BANK_CONFLICT / (SQ_LDS_IDX_ACTIVE - "
                                                       "SQ_LDS_BANK_CONFLICT)",
                                lds_busy="SQ_LDS_IDX_ACTIVE / CU_NUM / kernel cycles",
                                l2_hit_rate="TCC_HIT / TCC_REQ", wave_wait_frac="SQ_WAIT_ANY / SQ_WAVE_CYCLES"))
    with open(COUNTERS_JSON, "w") as fh:
        json.dump(rec, fh, indent=1)
    return rec


def chop_leg(dev, steps, n_tracks=350_000, precision=0.1, cpu=True):
    """Track chopping (the step before the path): 350k edep-like rows of 0.2-1.8 cm chopped at the
    dataset's 0.1 cm into ~3.5M segment rows (the size of the 32-event batch).  The fill kernel is
    timed with events on torch's current stream (where it is launched); the whole call (count, one
    host read of the total, allocation, fill) by wall clock.  Bytes: the track rows read twice
    (count and fill), the int64 offsets written and read, the segment rows written."""
    import ctypes as C

    import numpy as np
    import torch

    from larndsim_amd import _lib, synthetic
    from larndsim_amd.dataio import _Chopper, chop_tracks

    fields = tuple(synthetic.SEGMENT_FIELDS)
    t = synthetic.make_tracks(np.random.default_rng(21), n_tracks, (0.2, 1.8), n_events=32)
    td = torch.from_numpy(t).to(dev)
    F = len(fields)
    for _ in range(2):
        out = chop_tracks(td, fields, precision)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        out = chop_tracks(td, fields, precision)
    torch.cuda.synchronize()
    call = (time.perf_counter() - t0) / steps
    M = out.shape[0]
    ch = _Chopper(td, fields, precision, dev)
    st = ch._stream()
    info = ch.info

    def fill():
        _lib.check(ch.lib.lsim_chop_fill(C.byref(ch.lay), C.c_void_p(ch.t.data_ptr()), ch.n, C.c_double(precision),
                                         C.c_void_p(ch.ws.data_ptr()), ch.ws.numel(), M, M, None, None, 0, -1,
                                         C.c_void_p(out.data_ptr()), C.c_void_p(info.data_ptr()), st))

    def count():
        _lib.check(ch.lib.lsim_chop_count(C.byref(ch.lay), C.c_void_p(ch.t.data_ptr()), ch.n, C.c_double(precision),
                                          C.c_void_p(ch.ws.data_ptr()), ch.ws.numel(), C.c_void_p(info.data_ptr()),
                                          st))

    res = {}
    for name, fn in (("fill", fill), ("count_scan", count)):
        fn()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(steps):
            fn()
        e1.record()
        torch.cuda.synchronize()
        res[name] = e0.elapsed_time(e1) / steps
    b_fill = n_tracks * F * 4 + n_tracks * 8 + M * F * 4
    b_count = n_tracks * F * 4 + 2 * n_tracks * 8
    leg = dict(workload=f"{n_tracks} track rows (0.2-1.8 cm) chopped at {precision} cm", tracks=n_tracks,
               segments=M, ms_per_call=call * 1e3, segments_per_s=M / call, fill_kernel_ms=res["fill"],
               count_scan_ms=res["count_scan"], fill_bytes=b_fill, fill_gbs=b_fill / (res["fill"] * 1e-3) / 1e9,
               fill_hbm_frac=b_fill / (res["fill"] * 1e-3) / 1e9 / HBM_PEAK_GBS,
               kernels_gbs=(b_fill + b_count) / ((res["fill"] + res["count_scan"]) * 1e-3) / 1e9)
    if cpu:
        # the vectorised NumPy restatement of chop_tracks on the host (the reference loops in Python per track)
        sys.path.insert(0, ROOT)
        from threadpoolctl import threadpool_limits

        from oracle import chop as OC

        with threadpool_limits(limits=1):
            t1 = time.perf_counter()
            ref = OC.chop_tracks(t, fields, precision)
            dtc = time.perf_counter() - t1
        leg["cpu_baseline"] = dict(value=ref.shape[0] / dtc, unit="segments/s", cores=1, kind="port",
                                   sample=f"all {n_tracks} rows, vectorised float32/float64 NumPy restatement, "
                                          f"1 thread, {dtc:.2f} s")
    del td, out, ch
    return leg


def prob_fee_leg(params, wfs, steps, cpu=True, cpu_rows=16):
    """The probabilistic readout (beam search of 20 paths, 10 hit indices, reset noise 900 e-) on
    the waveform rows of the 200k-segment event, inputs resident; timed with events on torch's
    current stream.  CPU baseline: the float64 NumPy restatement on the first cpu_rows rows."""
    import numpy as np
    import torch

    from larndsim_amd.fee import get_adc_values_average_noise_vmap

    pp = params.replace(RESET_NOISE_CHARGE=900.0)
    for _ in range(2):
        get_adc_values_average_noise_vmap(pp, wfs)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(steps):
        lp, q = get_adc_values_average_noise_vmap(pp, wfs)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / steps
    P, T = wfs.shape
    V, S = int(pp.fee_paths_scaling), int(pp.MAX_ADC_VALUES)
    active = int((lp.amax(dim=2) > -999.0).any(dim=0).sum())
    leg = dict(rows=P, ticks=T, paths=V, hit_indices=S, steps_run=active, ms=ms, rows_per_s=P / (ms * 1e-3),
               path_ticks_per_s=P * active * V * (T - 1) / (ms * 1e-3),
               out_bytes=2 * P * S * (T - 1) * 4, out_gbs=2 * P * S * (T - 1) * 4 / (ms * 1e-3) / 1e9)
    # forward + backward (lsim_prob_fee_backward, path selections held fixed) with cotangents on the
    # probabilities and on the expected charges, the shape of the reference's probabilistic losses
    wg = wfs.detach().clone().requires_grad_(True)
    gen = torch.Generator(device=wfs.device).manual_seed(11)
    Gp = torch.randn(lp.shape, generator=gen, device=wfs.device, dtype=lp.dtype)
    Gq = torch.randn(q.shape, generator=gen, device=wfs.device, dtype=q.dtype) * 1e-4

    def fwd_bwd():
        wg.grad = None
        lpg, qg = get_adc_values_average_noise_vmap(pp, wg)
        ((Gp * lpg.exp()).sum() + (Gq * qg).sum()).backward()

    fwd_bwd()
    torch.cuda.synchronize()
    e0.record()
    for _ in range(steps):
        fwd_bwd()
    e1.record()
    torch.cuda.synchronize()
    leg["ms_fwd_bwd"] = e0.elapsed_time(e1) / steps
    leg["rows_per_s_fwd_bwd"] = P / (leg["ms_fwd_bwd"] * 1e-3)
    if cpu:
        sys.path.insert(0, ROOT)
        from threadpoolctl import threadpool_limits

        from oracle import probfee as PF

        w = wfs[:cpu_rows].cpu().numpy()
        with threadpool_limits(limits=1):
            t0 = time.perf_counter()
            PF.prob_fee(w, pp.t_sampling, V, S, pp.adc_interval, pp.DISCRIMINATION_THRESHOLD, pp.RESET_NOISE_CHARGE)
            dt = time.perf_counter() - t0
        leg["cpu_baseline"] = dict(value=cpu_rows / dt, unit="rows/s", cores=1, kind="port",
                                   sample=f"first {cpu_rows} rows, float64 NumPy restatement, 1 thread, {dt:.2f} s")
    return leg


def _param_fwd_bwd(p, trk, fields, steps, ms_fwd):
    """Forward + backward (cotangents on every hit's ADC value and z) through the parametrized
    model's adjoint, against the 15 forwards (1 + 2 x 7) of a central-difference gradient."""
    import torch

    from larndsim_amd import param as GP
    from larndsim_amd.autograd import theta_of

    def fwd_bwd():
        theta = theta_of(p).cuda().requires_grad_(True)
        adcs, z, _ = GP.ParamSimFunction.apply(theta, p, trk, fields, 1, None, None)
        (adcs.double().sum() + z.double().sum()).backward()
        return theta.grad

    fwd_bwd()
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(steps):
        grad = fwd_bwd()
    e1.record()
    torch.cuda.synchronize()
    ms_fb = e0.elapsed_time(e1) / steps
    return dict(ms_fwd_bwd=ms_fb, fd_gradient_ms=15 * ms_fwd, adjoint_vs_fd=15 * ms_fwd / ms_fb,
                grad_finite=bool(torch.isfinite(grad).all()))


def param_leg(params, trk, fields, steps, cpu=True, cpu_segments=10000):
    """The parametrized current model (simulate_parametrized, sim_jax.py:339-372; mc_diff, one
    electron per segment, diffusion in the current, 4 neighbour pixels) on the 200k-segment
    event, inputs resident: the whole call (electrons, pixel set, waveforms, readout, hits) by
    events on torch's current stream, the waveform kernel alone, and forward + backward through
    the adjoint against the 15 forwards of a central-difference gradient; the same without
    diffusion in the current.  Rows that meet an overflowing pair are NaN, as in the reference
    (``rows_nan``), and carry no hits.  CPU baseline: the float64 NumPy oracle on the first
    cpu_segments segments, 1 thread."""
    import numpy as np
    import torch

    from larndsim_amd import param as GP

    p = params.replace(mc_diff=True, diffusion_in_current_sim=True)
    out = GP.simulate_parametrized(p, trk, fields, rngseed=1)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(steps):
        out = GP.simulate_parametrized(p, trk, fields, rngseed=1)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / steps
    e = GP.electrons(p, trk, fields, rngseed=1)
    GP.waveforms(p, e)
    torch.cuda.synchronize()
    e0.record()
    for _ in range(steps):
        wfs, U = GP.waveforms(p, e)
    e1.record()
    torch.cuda.synchronize()
    ms_w = e0.elapsed_time(e1) / steps
    n = int(trk.shape[0])
    K = (2 * int(p.number_pix_neighbors) + 1) ** 2
    rows_nan = int(torch.isnan(wfs).any(dim=1).sum())
    leg = dict(workload=f"200k-segment event, mc_diff, {K} pixels per electron, 51 current samples", segments=n,
               pairs=n * K, rows=int(U.numel()), rows_nan=rows_nan, hits=int(out[0].numel()), ms=ms,
               segments_per_s=n / (ms * 1e-3), waveforms_ms=ms_w, pairs_per_s=n * K / (ms_w * 1e-3))
    leg.update(_param_fwd_bwd(p, trk, fields, steps, ms))
    # without diffusion in the current (one-sided exponentials, finite everywhere): every row
    # carries hits, so the adjoint visits most pairs
    pn = p.replace(diffusion_in_current_sim=False)
    GP.simulate_parametrized(pn, trk, fields, rngseed=1)
    torch.cuda.synchronize()
    e0.record()
    for _ in range(steps):
        outn = GP.simulate_parametrized(pn, trk, fields, rngseed=1)
    e1.record()
    torch.cuda.synchronize()
    ms_n = e0.elapsed_time(e1) / steps
    leg["no_diffusion"] = dict(ms=ms_n, segments_per_s=n / (ms_n * 1e-3), hits=int(outn[0].numel()),
                               **_param_fwd_bwd(pn, trk, fields, steps, ms_n))
    if cpu:
        sys.path.insert(0, ROOT)
        from threadpoolctl import threadpool_limits

        from oracle import param as PA

        tr = trk[:cpu_segments].cpu().numpy()
        normals = np.random.default_rng(1).normal(size=(len(tr), 3))
        with threadpool_limits(limits=1), np.errstate(all="ignore"):
            t0 = time.perf_counter()
            PA.simulate_parametrized(p, tr, fields, normals=normals)
            dt = time.perf_counter() - t0
        leg["cpu_baseline"] = dict(value=cpu_segments / dt, unit="segments/s", cores=1, kind="port",
                                   sample=f"first {cpu_segments} segments, float64 NumPy oracle, 1 thread, {dt:.2f} s")
    return leg


def sdtw_leg(params, lut, trk, fields, steps, cpu=True):
    """Soft-DTW loss (value + gradient) between the hit ADC sequences of the 200k-segment event at
    nominal and at perturbed parameters, inputs resident; events around the whole lsim_sdtw call.
    CPU baseline: the float64 NumPy restatement on the first 2 000 x 2 000 values."""
    import numpy as np
    import torch

    from larndsim_amd import synthetic
    from larndsim_amd.losses import _sdtw_hip
    from larndsim_amd.strategies import HipLUTSimulation

    strat = HipLUTSimulation(lut)
    with torch.no_grad():
        a = strat.predict(params, trk, fields)["adcs"].reshape(-1)
        b = strat.predict(synthetic.perturbed_params(params, seed=4), trk, fields)["adcs"].reshape(-1)
    x, y = a[a > 0].contiguous(), b[b > 0].contiguous()
    _sdtw_hip(x, y, 1.0, True)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(steps):
        _sdtw_hip(x, y, 1.0, True)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / steps
    cells = x.numel() * y.numel()
    leg = dict(n=x.numel(), m=y.numel(), ms_value_and_grad=ms, cells_per_s=2 * cells / (ms * 1e-3))
    if cpu:
        sys.path.insert(0, ROOT)
        from oracle import sdtw as SD

        xs, ys = x[:2000].cpu().double().numpy(), y[:2000].cpu().double().numpy()
        t0 = time.perf_counter()
        SD.soft_dtw_grad(xs, ys, 1.0)
        dt = time.perf_counter() - t0
        leg["cpu_baseline"] = dict(value=2 * xs.size * ys.size / dt, unit="cells/s", cores=1, kind="port",
                                   sample=f"2000 x 2000 values, float64 NumPy by anti-diagonals, {dt:.2f} s")
    return leg


def batched_fwd_bwd(simr, params, dev, steps, fd_baseline=True):
    """32 uneven events (20k-200k segments): forward + backward replayed as one hipGraph, waveform
    rows recomputed (not stored), against the per-event forward-then-finite-difference gradient."""
    import numpy as np
    import torch

    from larndsim_amd import sim as S, synthetic

    log("[bench] generating 32 events (20k-200k segments)")
    ft, ff = synthetic.workload("c5")
    NF = ft.shape[0]
    ev = ft[:, ff.index("eventID")].astype(np.int64)
    frun = S.Runner(simr, params, torch.from_numpy(ft).to(dev), ff, n_events=32, want_wfs=False, want_tape=True)
    rng = np.random.default_rng(9)
    frun.set_cotangents(torch.from_numpy(rng.normal(size=(frun.max_rows * params.MAX_ADC_VALUES,))
                                         .astype(np.float32)).to(dev))
    frun.forward()
    frun.backward()
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):
        frun.forward()
        frun.backward()
    torch.cuda.current_stream().wait_stream(s)
    torch.cuda.synchronize()
    with torch.cuda.graph(g):
        frun.forward()
        frun.backward()
    torch.cuda.synchronize()
    for _ in range(2):
        g.replay()
    torch.cuda.synchronize()
    fsteps = max(3, steps // 4)
    t0 = time.perf_counter()
    for _ in range(fsteps):
        g.replay()
    torch.cuda.synchronize()
    dtf = (time.perf_counter() - t0) / fsteps
    del g
    bb = algorithmic_bytes_fwdbwd(NF, frun.rows)
    res = {"batched_fwd_bwd": dict(events=32, segments=NF, rows=frun.rows, ms_per_step=dtf * 1e3, segments_per_s=NF / dtf,
                                   graph=True, wfs_stored=False, algorithmic_bytes=bb,
                                   hbm_frac=bb / dtf / 1e9 / HBM_PEAK_GBS)}
    del frun
    if not fd_baseline:
        log(f"[bench] batched fwd+bwd {NF} segments {dtf * 1e3:.2f} ms -> {NF / dtf:.3e} seg/s")
        return res
    # baseline: per event, 1 + 14 plain forwards (central differences, h = 1e-3 relative)
    t_events = 0.0
    for e in range(32):
        sub = ft[ev == e].copy()
        sub[:, ff.index("eventID")] = 0
        erun = S.Runner(simr, params, torch.from_numpy(sub).to(dev), ff, n_events=1, want_wfs=True)
        erun.forward()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(3):
            erun.forward()
        torch.cuda.synchronize()
        t_events += (time.perf_counter() - t0) / 3
        del erun
    t_fd = 15 * t_events
    res["per_event_fd_baseline"] = dict(forwards_per_event=15, ms_per_gradient=t_fd * 1e3, segments_per_s=NF / t_fd,
                                        adjoint_speedup=t_fd / dtf)
    log(f"[bench] batched fwd+bwd {NF} segments {dtf * 1e3:.2f} ms -> {NF / dtf:.3e} seg/s; "
        f"per-event FD {t_fd * 1e3:.1f} ms (x{t_fd / dtf:.1f})")
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--waves", type=int, default=0, help="waves per row workgroup (0 = autotune)")
    ap.add_argument("--row-tile", type=int, default=16, help="cell tile of the row order (with --waves)")
    ap.add_argument("--row-cost", type=int, default=256,
                    help="per-row cost (pairs) of the XCD band cuts when --waves is given (else autotuned)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-bwd", action="store_true")
    ap.add_argument("--no-batched-bwd", action="store_true", help="skip the batched forward+backward of 32 events")
    ap.add_argument("--cpu-sample", type=int, default=20000)
    ap.add_argument("--cpu-procs", type=int, default=16, help="processes of the all-core CPU baseline")
    ap.add_argument("--fit-steps", type=int, default=200, help="Adam steps of the 32-event gradient fit (0: skip)")
    ap.add_argument("--fit-wfs-steps", type=int, default=100,
                    help="Adam steps of the same fit with a loss on the waveforms (0: skip)")
    ap.add_argument("--copy-calibrate", action="store_true")
    ap.add_argument("--no-chop", action="store_true", help="skip the track-chopping leg")
    ap.add_argument("--chop-only", action="store_true", help="profiling: only the track-chopping leg")
    ap.add_argument("--no-prob-fee", action="store_true", help="skip the probabilistic readout leg")
    ap.add_argument("--no-sdtw", action="store_true", help="skip the soft-DTW leg")
    ap.add_argument("--no-param", action="store_true", help="skip the parametrized current model leg")
    ap.add_argument("--param-only", action="store_true",
                    help="profiling: only the parametrized current model on the 200k-segment event")
    ap.add_argument("--sdtw-only", action="store_true", help="profiling: only the soft-DTW leg")
    ap.add_argument("--prob-fee-only", action="store_true",
                    help="profiling: only the probabilistic readout on the 200k-segment event's rows")
    ap.add_argument("--batched-bwd-only", action="store_true",
                    help="profiling: only the 32-event batched forward+backward graph replays (no baseline)")
    ap.add_argument("--single-only", action="store_true",
                    help="profiling: only the 200k-segment event's legs (forward, forward+backward, finite "
                         "differences, waveform cotangent)")
    ap.add_argument("--headline-only", action="store_true",
                    help="profiling: only the 128-event headline (no other leg, parity check or CPU baseline)")
    ap.add_argument("--profile-batch", action="store_true",
                    help="profiling: only eager forwards of the 128-event batch (for counter passes)")
    ap.add_argument("--record-traffic", metavar="DIR",
                    help="write the waveform kernel's HBM traffic from rocprofv3 counter passes under DIR "
                         "(fetch/, write/, copy_fetch/, copy_write/) into profiles/, stamped; needs --kernel-ms")
    ap.add_argument("--kernel-ms", type=float, default=None)
    ap.add_argument("--record-counters", metavar="DIR",
                    help="write issue, LDS and L2 figures of the forward kernels from rocprofv3 --pmc passes of "
                         "--profile-batch under DIR into profiles/, stamped")
    a = ap.parse_args()
    if a.record_traffic:
        print(json.dumps(record_traffic(a.record_traffic, a.kernel_ms)))
        return
    if a.record_counters:
        print(json.dumps(record_counters(a.record_counters)["derived"]))
        return
    if a.headline_only:
        a.no_bwd = a.no_batched_bwd = a.no_prob_fee = a.no_sdtw = a.no_chop = a.no_cpu = a.no_param = True
        a.fit_steps = a.fit_wfs_steps = 0
    if a.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {a.gpus} is not supported (events would shard across "
              "GPUs with no exchange in the forward and one 8-value all-reduce per gradient step)", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    if a.copy_calibrate:
        src = torch.ones(64 << 20, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(5):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": src.numel() * 4, "copy_bytes_written": dst.numel() * 4, "launches": 5}))
        return

    from larndsim_amd import consts, sim as S, synthetic

    dev = torch.device("cuda")
    params_kw = dict(number_pix_neighbors=4, RESET_NOISE_CHARGE=0.0, UNCORRELATED_NOISE_CHARGE=0.0)
    params = consts.default_params(**params_kw)
    lut = synthetic.make_lut(0)
    if a.chop_only:
        print(json.dumps(chop_leg(dev, a.steps, cpu=not a.no_cpu)))
        return
    if a.batched_bwd_only:
        if a.waves:
            S.set_row_waves(a.waves)
        print(json.dumps(batched_fwd_bwd(S.Simulator(params, lut), params, dev, a.steps, fd_baseline=False)))
        return
    if a.sdtw_only:
        tracks, fields = synthetic.workload("c2")
        print(json.dumps(sdtw_leg(params, lut, torch.from_numpy(tracks).to(dev), fields, a.steps, cpu=not a.no_cpu)))
        return
    if a.param_only:
        tracks, fields = synthetic.workload("c2")
        print(json.dumps(param_leg(params, torch.from_numpy(tracks).to(dev), fields, a.steps, cpu=not a.no_cpu)))
        return
    if a.prob_fee_only:
        tracks, fields = synthetic.workload("c2")
        st = S.Simulator(params, lut).forward(params, torch.from_numpy(tracks).to(dev), fields, n_events=1)
        print(json.dumps(prob_fee_leg(params, st.out["wfs"][: st.max_rows].contiguous(), a.steps, cpu=not a.no_cpu)))
        return
    simr = S.Simulator(params, lut)
    if a.profile_batch:
        bt, bf = synthetic.workload("c3")
        brun = S.Runner(simr, params, torch.from_numpy(bt).to(dev), bf, n_events=128, want_wfs=True)
        if a.waves:
            S.set_row_waves(a.waves)
            S.set_row_cost(a.row_cost)
            S.set_row_tile(a.row_tile)
        for i in range(a.steps):
            brun.forward()
            torch.cuda.synchronize()
            log(f"[bench] profile forward {i + 1}/{a.steps}")
        return

    if not a.single_only:
        # ---- headline: the batched forward of 128 uneven events, one graph replay per step ----
        log("[bench] generating 128 events (log-uniform 2k-400k segments)")
        bt, fields = synthetic.workload("c3")
        sizes3 = synthetic.event_sizes("c3")
        NB = bt.shape[0]
        btr = torch.from_numpy(bt).to(dev)
        brun = S.Runner(simr, params, btr, fields, n_events=128, want_wfs=True)
        PB = brun.rows
        log(f"[bench] N={NB} rows={PB}")
        if a.waves:
            S.set_row_waves(a.waves)
            S.set_row_cost(a.row_cost)
            S.set_row_tile(a.row_tile)
            waves, row_cost, row_tile, tune = a.waves, a.row_cost, a.row_tile, {}
        else:
            waves, row_cost, row_tile, tune = S.autotune_rows(brun)
        log(f"[bench] waves per row workgroup: {waves}, row cost {row_cost} pairs, row tile {row_tile} "
            f"(tuning ms: {tune})")
        g = torch.cuda.CUDAGraph()
        s = torch.cuda.Stream()
        s.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(s):
            brun.forward()
        torch.cuda.current_stream().wait_stream(s)
        torch.cuda.synchronize()
        with torch.cuda.graph(g):
            brun.forward()
        torch.cuda.synchronize()
        for _ in range(a.warmup):
            g.replay()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(a.steps):
            g.replay()
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        ms_step = dt / a.steps * 1e3
        seg_s = NB / (dt / a.steps)
        # the waveform kernel's own time: HIP events around its launches on the launch stream, over
        # eager forwards of the same batch right after the timed region
        S.timing(True)
        for _ in range(3):
            brun.forward()
        k_ms, k_n = S.timing_read()
        S.timing(False)
        k_avg = k_ms / max(k_n, 1)
        bytes_fwd = algorithmic_bytes_fwd(NB, PB)
        achieved = bytes_fwd / (k_avg * 1e-3) / 1e9
        log(f"[bench] batched {NB} segments {ms_step:.2f} ms -> {seg_s:.3e} seg/s; waveform kernel {k_avg:.2f} ms")
        del brun, btr, g

    extra = {}
    if a.single_only and a.waves:
        S.set_row_waves(a.waves)
        S.set_row_cost(a.row_cost)
        S.set_row_tile(a.row_tile)
    if a.headline_only:
        print(json.dumps({"metric": "segments/s", "value": seg_s, "ms_per_step": ms_step, "kernel_ms": k_avg,
                          "achieved_gbs": achieved, "config": {"segments": NB, "pixel_rows": PB, "row_waves": waves,
                                                                "row_cost": row_cost, "row_tile": row_tile}}),
              flush=True)
        return
    # ---- the 200k-segment event alone ----
    log("[bench] generating the 200k-segment event")
    tracks, fields = synthetic.workload("c2")
    N = tracks.shape[0]
    trk = torch.from_numpy(tracks).to(dev)
    run = S.Runner(simr, params, trk, fields, n_events=1, want_wfs=True, want_tape=True)
    P = run.rows
    for _ in range(a.warmup):
        run.forward()
    torch.cuda.synchronize()
    S.timing(True)
    t0 = time.perf_counter()
    for _ in range(a.steps):
        run.forward()
    torch.cuda.synchronize()
    dt1 = time.perf_counter() - t0
    k1_ms, k1_n = S.timing_read()
    S.timing(False)
    k1 = k1_ms / max(k1_n, 1)
    b1 = algorithmic_bytes_fwd(N, P)
    extra["single_event"] = dict(workload="1 event, 200k segments (500 tracks, 10-150 cm)", segments=N, pixel_rows=P,
                                 ms_per_step=dt1 / a.steps * 1e3, segments_per_s=N / (dt1 / a.steps),
                                 kernel="k_rows_fwd", kernel_ms=k1, algorithmic_bytes=b1,
                                 achieved_gbs=b1 / (k1 * 1e-3) / 1e9,
                                 frac=b1 / (k1 * 1e-3) / 1e9 / HBM_PEAK_GBS)
    log(f"[bench] 200k-segment event {dt1 / a.steps * 1e3:.3f} ms/step, {N / (dt1 / a.steps):.3e} seg/s; "
        f"waveform kernel {k1:.3f} ms")
    # ---- forward + backward ----
    if not a.no_bwd:
        rng = np.random.default_rng(7)
        d_adc = torch.from_numpy(rng.normal(size=(P * params.MAX_ADC_VALUES,)).astype(np.float32)).to(dev)
        run.set_cotangents(d_adc)
        for _ in range(a.warmup):
            run.forward()
            run.backward()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(a.steps):
            run.forward()
            run.backward()
        torch.cuda.synchronize()
        dtb = (time.perf_counter() - t0) / a.steps
        # finite-difference baseline: nominal + 2 x 7 perturbed plain forwards
        fd_params = [params] + [params.replace(**{n: getattr(params, n) * (1 + s * 1e-3)})
                                for n in consts.GRAD_PARAMS for s in (1, -1)]
        t0 = time.perf_counter()
        for _ in range(max(1, a.steps // 5)):
            for pp in fd_params:
                run.set_params(pp)
                run.forward()
        torch.cuda.synchronize()
        dtfd = (time.perf_counter() - t0) / max(1, a.steps // 5)
        run.set_params(params)
        bfb = algorithmic_bytes_fwdbwd(N, P)
        extra["fwd_bwd"] = dict(ms_per_step=dtb * 1e3, segments_per_s=N / dtb, algorithmic_bytes=bfb,
                                hbm_frac=bfb / dtb / 1e9 / HBM_PEAK_GBS)
        extra["fd_gradient_baseline"] = dict(forwards=len(fd_params), ms_per_gradient=dtfd * 1e3,
                                             segments_per_s=N / dtfd, adjoint_speedup=dtfd / dtb)
        log(f"[bench] fwd+bwd {dtb * 1e3:.3f} ms; FD gradient {dtfd * 1e3:.3f} ms (x{dtfd / dtb:.1f})")
        # the same with a dense cotangent on every waveform tick as well (lsim_backward_wfs)
        st_w = simr.forward(params, trk, fields, n_events=1, want_tape=True)
        d_w = torch.from_numpy(rng.normal(size=(st_w.max_rows, simr.nout)).astype(np.float32)).to(dev)
        d_a = d_adc[: st_w.max_rows * params.MAX_ADC_VALUES]
        simr.backward(st_w, d_a, None, d_w)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(a.steps):
            st_w = simr.forward(params, trk, fields, n_events=1, want_tape=True, max_rows=P)
            simr.backward(st_w, d_a, None, d_w)
        torch.cuda.synchronize()
        dtw = (time.perf_counter() - t0) / a.steps
        extra["fwd_bwd_waveform_cotangent"] = dict(ms_per_step=dtw * 1e3, segments_per_s=N / dtw,
                                                   fd_speedup=dtfd / dtw)
        log(f"[bench] fwd+bwd with a waveform cotangent {dtw * 1e3:.3f} ms")
        del st_w, d_w
    if a.single_only:
        print(json.dumps(extra), flush=True)
        return

    # ---- batched forward + backward: 32 uneven events, one graph replay, vs per-event FD ----
    if not a.no_batched_bwd:
        extra.update(batched_fwd_bwd(simr, params, dev, a.steps))

    # ---- probabilistic readout of the same event's rows ----
    if not a.no_prob_fee:
        log("[bench] probabilistic readout")
        wst = simr.forward(params, trk, fields, n_events=1)
        extra["prob_fee"] = prob_fee_leg(params, wst.out["wfs"][: wst.max_rows].contiguous(), max(3, a.steps // 4),
                                         cpu=not a.no_cpu)
        log(f"[bench] prob fee {extra['prob_fee']['rows']} rows: {extra['prob_fee']['ms']:.2f} ms")
        del wst

    # ---- the parametrized current model on the same event ----
    if not a.no_param:
        log("[bench] parametrized current model")
        extra["param"] = param_leg(params, trk, fields, max(3, a.steps // 4), cpu=not a.no_cpu)
        log(f"[bench] param {extra['param']['segments']} segments: {extra['param']['ms']:.2f} ms "
            f"(waveforms {extra['param']['waveforms_ms']:.2f} ms)")

    # ---- soft-DTW fit loss on the event's hit sequences ----
    if not a.no_sdtw:
        log("[bench] soft-DTW")
        extra["sdtw"] = sdtw_leg(params, lut, trk, fields, max(3, a.steps // 4), cpu=not a.no_cpu)
        log(f"[bench] sdtw {extra['sdtw']['n']} x {extra['sdtw']['m']}: {extra['sdtw']['ms_value_and_grad']:.2f} ms")

    # ---- track chopping (the data step before the path) ----
    if not a.no_chop:
        log("[bench] track chopping")
        extra["chop"] = chop_leg(dev, a.steps, cpu=not a.no_cpu)
        log(f"[bench] chop {extra['chop']['segments']} rows: call {extra['chop']['ms_per_call']:.3f} ms, fill kernel "
            f"{extra['chop']['fill_kernel_ms']:.3f} ms ({extra['chop']['fill_gbs']:.0f} GB/s)")

    # ---- gradient fit: 32 uneven events, Adam on the 7 parameters from nominal to a perturbed target ----
    if a.fit_steps > 0:
        from larndsim_amd.fit import GradientDescentFitter
        from larndsim_amd.losses import mse_adc_blockwise
        from larndsim_amd.strategies import GenericLossStrategy, HipLUTSimulation

        log("[bench] gradient fit on 32 events")
        ft, ff = synthetic.workload("c5")
        ftr = torch.from_numpy(ft).to(dev)
        target_params = synthetic.perturbed_params(params, seed=5)
        strat = HipLUTSimulation(lut)
        with torch.no_grad():
            target = strat.predict(target_params, ftr, ff)
        names = [n for n in consts.GRAD_PARAMS if n != "vdrift"]  # vdrift is never read: no gradient
        # the reference's fit test settings (optimize/fit_test.sh, scripts/params_test.yaml): mse_adc, Adam lr
        # 2e-2 per parameter on sigmoid-normalised values, exponential decay 0.97, clip_by_global_norm 1
        fitter = GradientDescentFitter(strat, GenericLossStrategy(mse_adc_blockwise), params, names, lr=2e-2,
                                       max_clip_norm_val=1.0, lr_scheduler="exponential_decay",
                                       lr_kw=dict(decay_rate=0.97))
        fitter.step(ftr, ff, target)  # warm-up step (allocations, kernel load)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for i in range(a.fit_steps - 1):
            fitter.step(ftr, ff, target)
            if i % 25 == 24:
                log(f"[bench] fit step {i + 2}: loss {fitter.history['loss'][-1]:.4g}")
        torch.cuda.synchronize()
        dtfit = (time.perf_counter() - t0) / max(a.fit_steps - 1, 1)
        final = fitter.physical()
        rel0 = {n: abs(float(getattr(params, n)) / float(getattr(target_params, n)) - 1) for n in names}
        rel1 = {n: abs(final[n] / float(getattr(target_params, n)) - 1) for n in names}
        extra["fit"] = dict(events=32, segments=int(ft.shape[0]), steps=a.fit_steps, steps_per_s=1.0 / dtfit,
                            segments_per_s=ft.shape[0] / dtfit, loss_first=fitter.history["loss"][0],
                            loss_last=fitter.history["loss"][-1], rel_error_start=rel0, rel_error_end=rel1,
                            optimizer="Adam lr 2e-2 on sigmoid-normalised parameters, exponential decay 0.97, "
                                      "clip_by_global_norm 1.0",
                            loss="mse_adc: hit-cloud MMD (sigma 1 cm) + relative total charge")
        log(f"[bench] fit {1.0 / dtfit:.2f} steps/s; loss {fitter.history['loss'][0]:.4g} -> "
            f"{fitter.history['loss'][-1]:.4g}; rel err start {rel0} end {rel1}")

        # the same fit on a loss over the waveforms (no hit decision in it), through the dense adjoint
        if a.fit_wfs_steps > 0:
            from larndsim_amd.strategies import LossStrategy

            class WaveformMSE(LossStrategy):
                """sum over pixels and ticks of (wfs - wfs_target)^2 / sum wfs_target^2, rows matched by
                pixel id (which carries the event)."""

                def compute(self, params, prediction, target):
                    ids, tids = prediction["unique_pixels"], target["unique_pixels"]
                    u, inv = torch.unique(torch.cat([ids, tids]), return_inverse=True)
                    A = torch.zeros(len(u), prediction["wfs"].shape[1], dtype=torch.float64, device=dev)
                    B = torch.zeros_like(A).index_add(0, inv[ids.shape[0]:], target["wfs"].double())
                    A = A.index_add(0, inv[: ids.shape[0]], prediction["wfs"].double())
                    return ((A - B) ** 2).sum() / (B ** 2).sum(), {}

            log("[bench] gradient fit on 32 events, waveform loss")
            fitter = GradientDescentFitter(strat, WaveformMSE(), params, names, lr=2e-2, max_clip_norm_val=1.0,
                                           lr_scheduler="exponential_decay", lr_kw=dict(decay_rate=0.97))
            fitter.step(ftr, ff, target)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for i in range(a.fit_wfs_steps - 1):
                fitter.step(ftr, ff, target)
                if i % 25 == 24:
                    log(f"[bench] waveform fit step {i + 2}: loss {fitter.history['loss'][-1]:.4g}")
            torch.cuda.synchronize()
            dtw = (time.perf_counter() - t0) / max(a.fit_wfs_steps - 1, 1)
            final = fitter.physical()
            rel1 = {n: abs(final[n] / float(getattr(target_params, n)) - 1) for n in names}
            extra["fit_waveform"] = dict(events=32, segments=int(ft.shape[0]), steps=a.fit_wfs_steps,
                                         steps_per_s=1.0 / dtw, segments_per_s=ft.shape[0] / dtw,
                                         loss_first=fitter.history["loss"][0], loss_last=fitter.history["loss"][-1],
                                         rel_error_start=rel0, rel_error_end=rel1,
                                         optimizer="as the fit above",
                                         loss="sum (wfs - wfs_target)^2 / sum wfs_target^2, rows matched by pixel id")
            log(f"[bench] waveform fit {1.0 / dtw:.2f} steps/s; loss {fitter.history['loss'][0]:.4g} -> "
                f"{fitter.history['loss'][-1]:.4g}; rel err end {rel1}")
        del ftr, target, fitter

    # ---- gradient check on the 50k-segment event: adjoint vs central differences of the full loss ----
    if a.fit_steps > 0:
        from larndsim_amd.losses import mse_adc_blockwise
        from larndsim_amd.strategies import GenericLossStrategy, HipLUTSimulation

        log("[bench] gradient check on the 50k-segment event")
        gt, gf = synthetic.workload("c4")
        gtr = torch.from_numpy(gt).to(dev)
        strat = HipLUTSimulation(lut)
        lossf = GenericLossStrategy(mse_adc_blockwise)
        with torch.no_grad():
            tgt = strat.predict(synthetic.perturbed_params(params, seed=4), gtr, gf)
        names = list(consts.GRAD_PARAMS)
        v0 = {n: float(getattr(params, n)) for n in names}

        def loss_at(vals, grad=False):
            th = {n: torch.tensor(vals[n], dtype=torch.float64, requires_grad=grad) for n in names}
            p = params.replace(**th)
            L, _ = lossf.compute(p, strat.predict(p, gtr, gf), tgt)
            if grad:
                L.backward()
                return {n: float(th[n].grad) if th[n].grad is not None else 0.0 for n in names}
            return float(L)

        g_adj = loss_at(v0, grad=True)
        check = {}
        for n in names:
            h = 1e-3 * abs(v0[n])
            with torch.no_grad():
                fd = (loss_at({**v0, n: v0[n] + h}) - loss_at({**v0, n: v0[n] - h})) / (2 * h)
            check[n] = dict(adjoint=g_adj[n], central_fd=fd)
        extra["gradient_check_50k"] = dict(
            loss="mse_adc vs target at perturbed parameters", h_rel=1e-3, grads=check,
            note="central differences of the full loss include the discrete self-trigger / tick decisions; "
                 "the adjoint (like reverse-mode AD through the reference) holds them fixed")

        # the same check with a loss on the waveforms, which holds no hit or tick decision
        def wfs_loss_at(vals, grad=False):
            th = {n: torch.tensor(vals[n], dtype=torch.float64, requires_grad=grad) for n in names}
            out = strat.predict(params.replace(**th), gtr, gf)
            u, inv = torch.unique(torch.cat([out["unique_pixels"], tgt["unique_pixels"]]), return_inverse=True)
            n_out = out["unique_pixels"].shape[0]
            A = torch.zeros(len(u), out["wfs"].shape[1], dtype=torch.float64, device=dev)
            B = torch.zeros_like(A)
            A = A.index_add(0, inv[:n_out], out["wfs"].double())
            B = B.index_add(0, inv[n_out:], tgt["wfs"].double())
            L = ((A - B) ** 2).sum() / (B ** 2).sum()
            if grad:
                L.backward()
                return {n: float(th[n].grad) if th[n].grad is not None else 0.0 for n in names}
            return float(L)

        g_adj = wfs_loss_at(v0, grad=True)
        check = {}
        for n in names:
            h = 1e-3 * abs(v0[n])
            with torch.no_grad():
                fd = (wfs_loss_at({**v0, n: v0[n] + h}) - wfs_loss_at({**v0, n: v0[n] - h})) / (2 * h)
            check[n] = dict(adjoint=g_adj[n], central_fd=fd,
                            rel_err=abs(g_adj[n] - fd) / abs(fd) if fd != 0 else None)
        extra["gradient_check_50k_waveform"] = dict(
            loss="sum over pixels and ticks of (wfs - wfs_target)^2 / sum wfs_target^2, target at perturbed parameters",
            h_rel=1e-3, grads=check)
        log(f"[bench] waveform-loss gradient check: {check}")
        del gtr, tgt

    log("[bench] parity of the 1000-segment event against the FP64 oracle")
    extra["parity"] = parity_c1(simr, params, lut)
    log(f"[bench] parity {extra['parity']}")

    cpu = None
    if not a.no_cpu:
        log("[bench] CPU baseline on every core (whole events of the batch)")
        cpu = cpu_baseline_events(params_kw, lut, bt, fields, sizes3, procs=a.cpu_procs)
        log(f"[bench] CPU {cpu['value']:.1f} seg/s on {cpu['cores']} processes")
        log("[bench] CPU baseline, one thread")
        extra["cpu_baseline_1thread"] = cpu_baseline(params, lut, tracks, fields, a.cpu_sample,
                                                     "the 200k-segment event")
        log(f"[bench] CPU {extra['cpu_baseline_1thread']['value']:.1f} seg/s on 1 thread")

    achievable = copy_gbs(dev)
    log(f"[bench] device copy {achievable:.0f} GB/s")
    # issue, LDS and L2 figures measured with rocprofv3 counters (valid for the sources they measured)
    onchip = {"source": os.path.relpath(COUNTERS_JSON, ROOT), "measured": None}
    if os.path.exists(COUNTERS_JSON):
        with open(COUNTERS_JSON) as f:
            crec = json.load(f)
        onchip["stale"] = crec.get("source_stamp") != kernel_source_stamp()
        if not onchip["stale"]:
            onchip["measured"] = crec.get("derived", {}).get("k_rows_fwd")
    oc = onchip_per_segment(int(params.number_pix_neighbors), int(params.nb_tran_diff_bins), int(params.signal_length))
    onchip["algorithm_per_segment"] = oc

    # HBM traffic per launch from counter passes, valid only for the kernel sources they measured
    traffic, stamp = None, None
    if os.path.exists(TRAFFIC_JSON):
        with open(TRAFFIC_JSON) as f:
            rec = json.load(f)
        stamp = dict(source_stamp=rec.get("source_stamp"), kernel_ms=rec.get("kernel_ms"),
                     current=kernel_source_stamp())
        stamp["stale"] = stamp["source_stamp"] != stamp["current"]
        if not stamp["stale"]:
            traffic = rec.get("bytes_per_launch")

    line = {
        "metric": "segments/s fwd & fwd+bwd (grad of 7 det params), 1 GPU, % of HBM roofline",
        "value": seg_s,
        "unit": "segments/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 (f64 index math)",
        "data": "synthetic (seeded tracks and 45x45x1950 response LUT)",
        "config": {"workload": "batched forward: 128 events with log-uniform 2k-400k segments (seed 3), one "
                               "hipGraph replay per step, module0 geometry, 45x45x1950 LUT, 4 neighbour pixels, "
                               "5x5 diffusion bins, noise off",
                   "segments": NB, "events": 128, "pixel_rows": PB, "row_waves": waves, "row_cost": row_cost,
                   "row_tile": row_tile, "tuning_ms": tune},
        "roofline": {"bound": "hbm", "kernel": "k_rows_fwd", "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "traffic_stamp": stamp, "algorithmic_bytes": bytes_fwd, "kernel_ms": k_avg,
                     "achievable": achievable,
                     "forward_frac": bytes_fwd / (ms_step * 1e-3) / 1e9 / HBM_PEAK_GBS},
        "onchip": onchip,
        "cpu_baseline": cpu,
    }
    line.update(extra)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
