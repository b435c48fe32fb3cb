"""Seeded synthetic inputs: muon-like track segments and a smooth bipolar response LUT.

The reference's LUT files (``response_44.npy``, ``response_44_v2a_full_tick.npz``) are not
distributed, so the workloads use a synthetic 45x45x1950 float32 LUT built here, and tracks
generated here.  Segment arrays use the reference's (N, F) float32 row-major layout with the
26 fields of the ``segments`` dataset in ``prepared_data/*.h5`` (field order below).
"""
from __future__ import annotations

import numpy as np

# Field order of the reference's segments dataset (prepared_data/input_*.h5).
SEGMENT_FIELDS = (
    "eventID", "z_end", "trackID", "tran_diff", "z_start", "x_end", "y_end", "n_electrons",
    "pdgId", "x_start", "y_start", "t_start", "t0_start", "t0_end", "t0", "dx", "long_diff",
    "pixel_plane", "t_end", "dEdx", "dE", "t", "y", "x", "z", "n_photons",
)

# Drift volume of the workloads: inside TPC 0 of module0 (anode z = 30.58975, cathode z = 0.15875).
CUBE_LO = np.array([-15.0, -15.0, 0.30])
CUBE_HI = np.array([15.0, 15.0, 30.30])


def _fold(p, lo, hi):
    """Billiard folding of an unbounded straight line into the box [lo, hi]."""
    w = hi - lo
    v = np.mod(p - lo, 2 * w)
    return lo + (w - np.abs(v - w))


def _landau_dedx(rng, n):
    from scipy.stats import moyal

    d = moyal.rvs(loc=1.95, scale=0.12, size=n, random_state=rng)
    return np.clip(d, 1.5, 20.0)


def make_event(rng, n_segments: int, n_tracks: int | None = None, length_range=(10.0, 150.0),
               track_length: float | None = None, event_id: int = 0) -> np.ndarray:
    """One event of straight tracks chopped into ``n_segments`` equal-length segments.

    Track starts are uniform in the drift cube and directions isotropic; a track longer than
    the cube folds back at the walls (billiard reflection) so every segment midpoint stays in
    the cube.  dE/dx follows a Moyal (Landau-like) law around 2.1 MeV/cm, clipped to
    [1.5, 20] MeV/cm; dE = dEdx * dx.  Times are zero.
    """
    if n_tracks is None:
        n_tracks = max(1, int(round(n_segments / 400)))
    if track_length is not None:
        lengths = np.full(n_tracks, float(track_length))
    else:
        lengths = rng.uniform(length_range[0], length_range[1], n_tracks)
    dx = lengths.sum() / n_segments
    # Largest-remainder split so the per-track counts add up to n_segments exactly.
    raw = lengths / dx
    counts = np.floor(raw).astype(np.int64)
    short = n_segments - counts.sum()
    if short > 0:
        order = np.argsort(-(raw - counts), kind="stable")
        counts[order[:short]] += 1
    counts = np.maximum(counts, 0)
    starts = rng.uniform(CUBE_LO, CUBE_HI, size=(n_tracks, 3))
    u = rng.normal(size=(n_tracks, 3))
    u /= np.linalg.norm(u, axis=1, keepdims=True)

    tid = np.repeat(np.arange(n_tracks), counts)
    k = np.concatenate([np.arange(c) for c in counts]) if n_tracks else np.zeros(0)
    s0 = k * dx
    p_start = _fold(starts[tid] + s0[:, None] * u[tid], CUBE_LO, CUBE_HI)
    p_end = _fold(starts[tid] + (s0 + dx)[:, None] * u[tid], CUBE_LO, CUBE_HI)
    p_mid = _fold(starts[tid] + (s0 + 0.5 * dx)[:, None] * u[tid], CUBE_LO, CUBE_HI)
    n = len(tid)
    dedx = _landau_dedx(rng, n)

    out = np.zeros((n, len(SEGMENT_FIELDS)), dtype=np.float32)
    col = {f: i for i, f in enumerate(SEGMENT_FIELDS)}
    out[:, col["eventID"]] = event_id
    out[:, col["trackID"]] = tid
    out[:, col["pdgId"]] = 13
    out[:, col["x_start"]], out[:, col["y_start"]], out[:, col["z_start"]] = p_start.T
    out[:, col["x_end"]], out[:, col["y_end"]], out[:, col["z_end"]] = p_end.T
    out[:, col["x"]], out[:, col["y"]], out[:, col["z"]] = p_mid.T
    out[:, col["dx"]] = dx
    out[:, col["dEdx"]] = dedx
    out[:, col["dE"]] = (dedx * dx).astype(np.float32)
    return out


def make_batch(rng, sizes, **kw) -> np.ndarray:
    """Concatenate events with local eventIDs 0..E-1 (``dataio.py:47-61`` convention)."""
    return np.concatenate([make_event(rng, int(n), event_id=e, **kw) for e, n in enumerate(sizes)], axis=0)


def make_tracks(rng, n_tracks: int, length_range=(0.2, 1.8), n_events: int = 1) -> np.ndarray:
    """Unchopped track rows (one straight piece each, as edep-sim writes them) for the chop step.

    Starts uniform in the drift cube, isotropic directions, lengths U(length_range) cm, dE/dx
    Moyal around 2.1 MeV/cm, dE = dEdx * length, dx = length, midpoint x/y/z; rows sorted by
    eventID 0 .. n_events - 1, trackID the row index within the event.
    """
    out = np.zeros((n_tracks, len(SEGMENT_FIELDS)), dtype=np.float32)
    col = {f: i for i, f in enumerate(SEGMENT_FIELDS)}
    start = rng.uniform(CUBE_LO, CUBE_HI, size=(n_tracks, 3))
    u = rng.normal(size=(n_tracks, 3))
    u /= np.linalg.norm(u, axis=1, keepdims=True)
    length = rng.uniform(length_range[0], length_range[1], n_tracks)
    end = start + u * length[:, None]
    ev = np.sort(rng.integers(0, n_events, n_tracks))
    first = np.searchsorted(ev, ev)
    dedx = _landau_dedx(rng, n_tracks)
    for k, c in enumerate("xyz"):
        out[:, col[c + "_start"]] = start[:, k]
        out[:, col[c + "_end"]] = end[:, k]
        out[:, col[c]] = 0.5 * (start[:, k] + end[:, k])
    out[:, col["eventID"]] = ev
    out[:, col["trackID"]] = np.arange(n_tracks) - first
    out[:, col["pdgId"]] = 13
    out[:, col["dx"]] = length
    out[:, col["dEdx"]] = dedx
    out[:, col["dE"]] = (dedx * length).astype(np.float32)
    return out


# --- the five workloads -------------------------------------------------------------------

def workload(name: str):
    """Segments for a named workload.  Returns (tracks (N, F) float32, fields tuple)."""
    if name == "c1":  # 1 event, 10 tracks x 40 cm at 0.4 cm = 1000 segments
        rng = np.random.default_rng(1)
        return make_event(rng, 1000, n_tracks=10, track_length=40.0), SEGMENT_FIELDS
    if name == "c2":  # 1 event, 500 tracks of U(10,150) cm, 200k segments
        rng = np.random.default_rng(2)
        return make_event(rng, 200_000, n_tracks=500), SEGMENT_FIELDS
    if name == "c3":  # 128 events, log-uniform segment counts in [2k, 400k]
        rng = np.random.default_rng(3)
        sizes = np.round(np.exp(rng.uniform(np.log(2e3), np.log(4e5), 128))).astype(np.int64)
        return make_batch(rng, sizes), SEGMENT_FIELDS
    if name == "c4":  # 1 event, 50k segments
        rng = np.random.default_rng(4)
        return make_event(rng, 50_000), SEGMENT_FIELDS
    if name == "c5":  # 32 events x U(20k, 200k)
        rng = np.random.default_rng(5)
        sizes = rng.integers(20_000, 200_001, 32)
        return make_batch(rng, sizes), SEGMENT_FIELDS
    raise ValueError(name)


def event_sizes(name: str):
    rng = np.random.default_rng(3 if name == "c3" else 5)
    if name == "c3":
        return np.round(np.exp(rng.uniform(np.log(2e3), np.log(4e5), 128))).astype(np.int64)
    return rng.integers(20_000, 200_001, 32)


def perturbed_params(params, seed: int, names=("Ab", "kb", "eField", "lifetime", "vdrift", "long_diff", "tran_diff"),
                     rel=0.05):
    """Each named parameter times (1 + U(-rel, rel)), seeded."""
    rng = np.random.default_rng(seed)
    kw = {n: float(getattr(params, n)) * (1.0 + rng.uniform(-rel, rel)) for n in names}
    return params.replace(**kw)


# --- synthetic response LUT -----------------------------------------------------------------

def make_lut(seed: int = 0, nx: int = 45, ny: int = 45, nt: int = 1950, t_sampling: float = 0.1) -> np.ndarray:
    """Smooth bipolar induced-current LUT of shape (nx, ny, nt), float32.

    Index (cx, cy) is the distance of the charge from the pixel centre in sub-bins (10 per
    pixel); arrival at the anode is at the last sample.  Bins inside the pixel (cx, cy <= 4)
    collect: a low, slow early ramp plus a sharp rise at arrival, normalised so that
    sum_t r * t_sampling = 1 (exactly 1 for the centre bin, slightly less off-centre).  Bins
    outside the pixel are bipolar (a positive lobe while the charge approaches, a negative
    lobe at arrival) with zero net integral and an amplitude that falls with distance.
    """
    rng = np.random.default_rng(seed)
    tau = (nt - 1 - np.arange(nt, dtype=np.float64))  # ticks before arrival
    lut = np.zeros((nx, ny, nt), dtype=np.float64)
    for cx in range(nx):
        for cy in range(ny):
            jit = rng.uniform(-1.0, 1.0, 4)
            if cx <= 4 and cy <= 4:
                t_fast = 10.0 + 0.8 * (cx + cy) + 1.5 * jit[0]
                t_slow = 260.0 + 40.0 * jit[1]
                eps = 0.03 + 0.01 * jit[2]
                r = eps * np.exp(-tau / t_slow) / t_slow + np.exp(-tau / t_fast) / t_fast
                r /= r.sum() * t_sampling
                if cx or cy:
                    r *= 1.0 - 0.004 * np.hypot(cx, cy) * (1.0 + 0.2 * jit[3])
            else:
                d = np.hypot(max(cx - 4.5, 0.0), max(cy - 4.5, 0.0)) / 10.0  # pitches outside the pixel
                amp = 0.35 * np.exp(-d / 0.9) * (1.0 + 0.1 * jit[0])
                mu1 = 45.0 + 25.0 * d + 5.0 * jit[1]
                s1 = 18.0 + 8.0 * d + 3.0 * jit[2]
                s2 = 7.0 + 1.0 * jit[3]
                g1 = np.exp(-0.5 * ((tau - mu1) / s1) ** 2)
                g2 = np.exp(-0.5 * (tau / s2) ** 2)
                r = g1 - g2 * (g1.sum() / g2.sum())
                r *= amp / max(np.abs(r).max(), 1e-30)
            lut[cx, cy] = r
    return lut.astype(np.float32)
