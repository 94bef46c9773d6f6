"""Autotuning of fusion width and kernel variant, measured on the device, cached on disk.

Five choices set the cost of a statevector simulation on the MI355X:

- the fusion width (qubits per dense pass): wider passes mean fewer HBM sweeps but more FP64
  work per amplitude (8 * 2^k flop per 32 B; k = 5 sits at the FP64 ridge of ~9.8 flop/B);
- the dense-kernel variant for placements with low target bits: the LDS-tiled kernel or the
  direct one;
- which operator widths (3, 4, 5 qubits) run on the FP64 matrix cores, and whether the dense
  kernels stream (non-temporal accesses) when every target is at bit 3 or above;
- dense passes or tile passes (runs of 1-2 qubit operators applied to 4096-amplitude tiles held
  in registers, one HBM sweep per run), and the tile planner's cost budget.

:func:`tune` plans a circuit at each candidate width, runs every candidate (with the tiled kernel
on and off) on a scratch state of the same size, and times it with HIP events, then times the
winner's plan for each matrix-core width mask with streaming on and off, then tile plans for
each (budget, low) candidate; the winner is stored
per (gfx architecture, qubit count) in ``~/.cache/lucivant_sim/autotune.json`` (or
``$LSIM_AUTOTUNE``) and read by :func:`apply`, which ``simulator`` calls on first use of a size;
every simulation then sets its own size's switches (:func:`set_switches`) before it launches. The LDS
program path has its own knob, the fusion width of compiled programs (0 = one op per gate), tuned
the same way by :func:`tune_program_width`.
"""

from __future__ import annotations

import json
import os
from pathlib import Path
from typing import Any

import numpy as np
import torch

from lucivant_sim import _native as nat
from lucivant_sim.circuit import Circuit
from lucivant_sim.engines import Statevector, run_plan
from lucivant_sim.plan import TILE_BITS, plan_statevector, plan_tiles

CACHE = Path(os.environ.get("LSIM_AUTOTUNE", str(Path.home() / ".cache" / "lucivant_sim" / "autotune.json")))


def _arch() -> str:
    return str(nat.device_info()["arch"]).split(":")[0]


def load_cache() -> dict[str, Any]:
    try:
        return json.loads(CACHE.read_text())
    except (OSError, ValueError):
        return {}


def save_cache(data: dict[str, Any]) -> None:
    CACHE.parent.mkdir(parents=True, exist_ok=True)
    tmp = CACHE.with_suffix(".tmp")
    tmp.write_text(json.dumps(data, indent=1, sort_keys=True))
    tmp.replace(CACHE)


def _time_passes(sv: Statevector, passes: list, reps: int) -> float:
    stream = torch.cuda.current_stream()
    run_plan(sv.buffer, sv.dtype, sv.nbits, 1, passes)  # warm-up
    start, end = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    start.record(stream)
    for _ in range(reps):
        run_plan(sv.buffer, sv.dtype, sv.nbits, 1, passes)
    end.record(stream)
    torch.cuda.synchronize()
    return start.elapsed_time(end) * 1e-3 / reps


#: (width, low_width) candidates: blocks touching qubits 0..2 may be held narrower (0: no cap)
CANDIDATES = ((3, 0), (4, 0), (5, 0), (4, 3), (5, 3), (5, 4))
LOW_BITS = 3
#: matrix-core width masks (bit k: k-qubit operators on the matrix cores)
MFMA_MASKS = (0, 1 << 5, (1 << 5) | (1 << 4), (1 << 5) | (1 << 4) | (1 << 3))
#: tile-pass (budget, low) candidates, timed after the dense winner for states of 12+ qubits
TILE_CANDIDATES = ((32.0, 4), (48.0, 4), (64.0, 4), (48.0, 5))
#: tile plans within this fraction of the fastest count as tied; among tied plans the one with
#: the fewest passes (then the largest budget) is chosen. On random layers the budgets from 48
#: up give the same passes and times; on ladder-shaped circuits (CX chains) larger budgets give
#: fewer passes: a 26-qubit, 40-layer ansatz takes 75 passes (55.3 ms) at budget 48 and 59
#: (51.6 ms) at 64.
TILE_TIE = 0.02


def tune(
    circuit: Circuit,
    candidates: tuple[tuple[int, int], ...] = CANDIDATES,
    reps: int = 2,
    store: bool = True,
    tile_candidates: tuple[tuple[float, int], ...] = TILE_CANDIDATES,
) -> dict[str, Any]:
    """Time ``circuit`` for every (width, low_width) candidate, tiled kernel on and off, then
    the winner's switch variants, then tile passes for each (budget, low) candidate; returns
    (and by default stores) the timings and the winner for this circuit's width (``tile`` is
    the winning (budget, low) when tile passes beat every dense plan, else None)."""
    lib = nat.load()
    sv = Statevector(circuit.num_qubits)
    timings: dict[str, float] = {}
    plans = {c: plan_statevector(circuit.operations, circuit.num_qubits, c[0], LOW_BITS, c[1]) for c in candidates}
    old = lib.lsim_set_dense_tiling(1)
    try:
        for tiled in (1, 0):
            lib.lsim_set_dense_tiling(tiled)
            for (w, lw), passes in plans.items():
                timings[f"w{w}_low{lw}_tiled{tiled}"] = _time_passes(sv, passes, reps)
    finally:
        lib.lsim_set_dense_tiling(old)
    best = min(timings, key=timings.__getitem__)
    w_s, lw_s, t_s = best.split("_")
    width, low_width, tiled = int(w_s[1:]), int(lw_s[3:]), int(t_s[5:])
    # the winner's plan: which operator widths run on the matrix cores, streaming accesses on/off
    mfma = int(lib.lsim_set_dense_mfma(-1))
    streaming = int(lib.lsim_set_dense_streaming(-1))
    variant_timings: dict[str, float] = {}
    winner = plans[(width, low_width)]
    old_mfma, old_streaming = mfma, streaming
    old_tiling = lib.lsim_set_dense_tiling(tiled)
    try:
        for mask in MFMA_MASKS:
            for stream_on in (1, 0):
                lib.lsim_set_dense_mfma(mask)
                lib.lsim_set_dense_streaming(stream_on)
                variant_timings[f"mfma{mask}_nt{stream_on}"] = _time_passes(sv, winner, reps)
    finally:
        lib.lsim_set_dense_mfma(old_mfma)
        lib.lsim_set_dense_streaming(old_streaming)
        lib.lsim_set_dense_tiling(old_tiling)
    choice = min(variant_timings, key=variant_timings.__getitem__)
    m_s, nt_s = choice.split("_")
    mfma, streaming = int(m_s[4:]), int(nt_s[2:])
    timings.update(variant_timings)
    # tile passes (runs of 1-2 qubit operators per sweep), under the winner's switches
    dense_seconds = min(timings[best], variant_timings[choice])
    tile_timings: dict[str, float] = {}
    tile_passes: dict[str, int] = {}
    if circuit.num_qubits >= TILE_BITS and tile_candidates:
        lib.lsim_set_dense_mfma(mfma)
        lib.lsim_set_dense_streaming(streaming)
        old_tiling = lib.lsim_set_dense_tiling(tiled)
        try:
            for budget, low in tile_candidates:
                passes = plan_tiles(circuit.operations, circuit.num_qubits, budget=budget, low=low)
                key = f"tile_b{budget:g}_l{low}"
                tile_timings[key] = _time_passes(sv, passes, reps)
                tile_passes[key] = len(passes)
        finally:
            lib.lsim_set_dense_mfma(old_mfma)
            lib.lsim_set_dense_streaming(old_streaming)
            lib.lsim_set_dense_tiling(old_tiling)
        timings.update(tile_timings)
    tile = None
    seconds = dense_seconds
    if tile_timings:
        fastest = min(tile_timings.values())
        tied = [key for key, t in tile_timings.items() if t <= fastest * (1 + TILE_TIE)]
        tbest = min(tied, key=lambda key: (tile_passes[key], -float(key.split("_")[1][1:]), tile_timings[key]))
        if tile_timings[tbest] < dense_seconds:
            b_s, l_s = tbest.split("_")[1:]
            tile, seconds = [float(b_s[1:]), int(l_s[1:])], tile_timings[tbest]
    rec = {
        "width": width,
        "low_width": low_width,
        "low_bits": LOW_BITS,
        "tiled": tiled,
        "mfma": mfma,
        "streaming": streaming,
        "tile": tile,
        "seconds": seconds,
        "timings": timings,
        "passes": {**{f"w{w}_low{lw}": len(p) for (w, lw), p in plans.items()}, **tile_passes},
    }
    if store:
        data = load_cache()
        data.setdefault(_arch(), {})[str(circuit.num_qubits)] = rec
        save_cache(data)
    return rec


#: The tuner's choices as measured on MI355X (gfx950) for random 40-layer u3 + cx/cz circuits
#: of 12 to 26 qubits (timings dropped), shipped so that a process with no cache of its own
#: plans with measured choices: tile passes from 21 qubits up, fused dense passes below.
DEFAULTS = Path(__file__).with_name("autotune_defaults.json")


def _defaults() -> dict[str, Any]:
    try:
        return json.loads(DEFAULTS.read_text())
    except (OSError, ValueError):
        return {}


def best_for(num_qubits: int) -> dict[str, Any] | None:
    """The stored choice for this device and width, or the nearest stored width; the user's
    cache first, else the shipped measurements for this device."""
    if not torch.cuda.is_available():
        return None
    arch = _arch()
    table = load_cache().get(arch) or _defaults().get(arch)
    if not table:
        return None
    if str(num_qubits) in table:
        return table[str(num_qubits)]
    nearest = min(table, key=lambda k: abs(int(k) - num_qubits))
    return table[nearest]


#: the library's own switch settings (tiled kernel on, matrix cores for 3-5 qubit operators,
#: streaming mode 1), used for sizes the autotuner has no record for
DEFAULT_SWITCHES = (1, (1 << 3) | (1 << 4) | (1 << 5), 1)


def apply(num_qubits: int) -> tuple[int, int, int, tuple[int, int, int], tuple[float, int] | None]:
    """(width, low_bits, low_width, switches, tile) to plan and launch ``num_qubits``-qubit
    circuits with; ``switches`` = (tiled, matrix-core width mask, streaming mode), set per
    launch by :func:`set_switches`; ``tile`` = (budget, low) when tile passes won for this size.
    Nothing is changed here: the switches are process-wide, so every simulation sets those of
    its own size before it launches."""
    from lucivant_sim import simulator

    rec = best_for(num_qubits)
    if rec is None:
        return simulator.FUSION_WIDTH, 0, 0, DEFAULT_SWITCHES, None
    switches = (int(rec.get("tiled", 1)), int(rec.get("mfma", DEFAULT_SWITCHES[1])), int(rec.get("streaming", 1)))
    tile = rec.get("tile")
    tile = (float(tile[0]), int(tile[1])) if tile and num_qubits >= TILE_BITS else None
    return int(rec["width"]), int(rec.get("low_bits", 0)), int(rec.get("low_width", 0)), switches, tile


def set_switches(switches: tuple[int, int, int]) -> None:
    """Set the library's tiling, matrix-core and streaming switches (bit-for-bit results don't
    depend on tiling or streaming; the matrix cores round differently from the vector FMAs)."""
    lib = nat.load()
    tiled, mfma, streaming = switches
    lib.lsim_set_dense_tiling(int(tiled))
    lib.lsim_set_dense_mfma(int(mfma))
    lib.lsim_set_dense_streaming(int(streaming))


def tune_program_width(circuits: list[Circuit], widths: tuple[int, ...] = (0, 2, 3), reps: int = 3) -> dict[str, float]:
    """Seconds per batch of ``circuits`` as LDS programs compiled at each fusion width."""
    from lucivant_sim import program as pg

    out: dict[str, float] = {}
    for w in widths:
        progs = [pg.compile_static(c, w) for c in circuits]
        pg.run_programs(progs, want_probs=True)
        torch.cuda.synchronize()
        times = []
        for _ in range(reps):
            start, end = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            start.record()
            pg.run_programs(progs, want_probs=True)
            end.record()
            torch.cuda.synchronize()
            times.append(start.elapsed_time(end) * 1e-3)
        out[f"w{w}"] = float(np.median(times))
    return out
