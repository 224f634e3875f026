"""The FP32-storage floor of the HH step, for parity bars: the FP64 oracle beside the same oracle
with FP32 storage emulated (per-segment conductances, 1/C and gate-table nodes in FP32, every gate
rounded after its update, optionally every voltage after every step).  Where a trajectory is
sensitive (a depolarisation close to threshold, repeated spikes) FP32 storage alone moves it by
more than 1e-5 relative; the bar of such a cell is three times that measured floor."""
import os

import numpy as np

from oracle import neuwon_oracle as O

from test_gpu_parity import voltage_error

BAR_REL, BAR_ABS_MV, FACTOR = 1e-5, 2e-3, 3.0


def oracle_traces(pop, P, stim, steps, round_v=False):
    """Root voltage of every cell after every step, and every voltage at the end, from the FP64
    oracle and from its FP32-storage emulation (round_v: voltages rounded after every step, as the
    stored formats keep them; the matrix-free stores keep FP64 voltages within a launch)."""
    threads = len(os.sched_getaffinity(0))
    out = []
    for emulate in (False, True):
        co = O.CHHPopulation(pop.cell_ptr, P, pop.C, pop.g_bar, dt=pop.dt, celsius=pop.celsius,
                             gate_mode=pop.gate_mode, stim=stim, fp32_storage=emulate,
                             round_v=emulate and round_v)
        tr = np.empty((pop.num_cells, steps))
        roots = pop.cell_ptr[:-1]
        for k in range(steps):
            co.advance(1, threads=threads)
            tr[:, k] = co.v[roots]
        out.append((tr, co.v.copy()))
    return out[0], out[1]


def check_cells(gpu_traces, ref_traces, floor_traces, dt, say=None, heads=None):
    """Per cell: the GPU trace's subthreshold relative error within max(1e-5, 3 x floor) and its
    error near excursions within max(2e-3 mV, 3 x floor); returns the per-cell errors.  heads:
    optional initial voltage per cell prepended to every trace (a spike in the first step)."""
    rows = []
    for c in range(len(ref_traces)):
        g, r, f = gpu_traces[c], ref_traces[c], floor_traces[c]
        if heads is not None:
            g, r, f = (np.r_[heads[c], x] for x in (g, r, f))
        e, ef = voltage_error(g, r, dt), voltage_error(f, r, dt)
        rows.append((e, ef))
    if say:
        say("per cell (GPU / FP32-storage floor): subthreshold rel "
            f"{[('%.1e/%.1e' % (e['rel_subthreshold'], f['rel_subthreshold'])) for e, f in rows]}; "
            f"near excursions mV {[('%.1e/%.1e' % (e['abs_spiking_mV'], f['abs_spiking_mV'])) for e, f in rows]}")
    for e, ef in rows:
        assert e["rel_subthreshold"] < max(BAR_REL, FACTOR * ef["rel_subthreshold"]), (e, ef)
        assert e["abs_spiking_mV"] < max(BAR_ABS_MV, FACTOR * ef["abs_spiking_mV"]), (e, ef)
    return rows
