"""A series of steps in chunks sized for the GPU, each written to the output
file as it finishes.

The reference splits a run into chunks of steps and grid rows that fit the
host's RAM with one worker process per chunk (interp/main.py:652-859), and
reorders steps so that equal variogram strings are adjacent, which lets a
worker reuse its variance arrays (interp/prepare.py:465-508).  Here one GPU
interpolates every step of a chunk in a few launches, each step with its
own variogram record, so no reordering is needed; what limits a chunk is
the device memory that its [steps x cells] outputs take next to the
library's workspaces, and the pinned host memory that receives them.

``plan_steps`` sizes the chunks from those two budgets (and an optional
cap, the reference's ``max_steps_per_chunk``); ``interpolate_series`` runs
them: the fields of chunk i are copied to pinned host memory and rounded
and written by a writer thread while the GPU interpolates chunk i + 1.
"""
from __future__ import annotations

import threading
from dataclasses import dataclass, field

import numpy as np
import torch

from . import steps as _steps


@dataclass
class Plan:
    """Chunks of steps: step_idxs[i] .. step_idxs[i + 1] - 1 (as the
    reference's step_idxs), the single grid chunk (grid_idxs, the grid's
    rows: no grid chunks on the GPU), and the budgets behind them."""
    step_idxs: np.ndarray
    grid_idxs: np.ndarray
    steps_per_chunk: int
    device_bytes_per_step: int
    host_bytes_per_step: int
    device_budget: int
    host_budget: int
    notes: list = field(default_factory=list)


def output_bytes_per_step(n_cells, methods, est_vars_flag=True):
    """Device and pinned host bytes one step's outputs take: per label the
    FP64 field and its float32 copy on the device (and the OK variance when
    wanted), the float32 fields on the host."""
    labels = [m for m in methods if m[1] != 'EST_VARS_OK']
    n_f = len(labels)
    if est_vars_flag and any(m[0] == 'OK' for m in labels):
        n_f += 1
    return n_cells * n_f * 12, n_cells * n_f * 4


def plan_steps(n_steps, n_cells, n_rows, methods, *, device_free,
               host_free, est_vars_flag=True, max_steps_per_chunk=None,
               max_mem_usage_ratio=1.0, reserve_frac=0.1):
    """Steps per chunk from the device memory free after the library's
    workspaces (device_free, bytes), a host budget (host_free x
    max_mem_usage_ratio, interp/main.py:676-677) holding two chunks of
    float32 fields in flight, and the optional cap.  A fraction
    reserve_frac of the device budget stays free for the library's
    per-call buffers.  Raises when not even one step fits, as the
    reference stops when no chunking fits (interp/main.py:796-803)."""
    assert 0 < max_mem_usage_ratio <= 1
    dev_b, host_b = output_bytes_per_step(n_cells, methods, est_vars_flag)
    dev_budget = int(device_free * (1.0 - reserve_frac))
    host_budget = int(host_free * max_mem_usage_ratio)
    by_dev = dev_budget // dev_b if dev_b else n_steps
    by_host = host_budget // (2 * host_b) if host_b else n_steps
    spc = int(min(n_steps, by_dev, by_host))
    notes = [f'device allows {by_dev} steps, host {by_host}']
    if max_steps_per_chunk is not None:
        spc = min(spc, int(max_steps_per_chunk))
        notes.append(f'capped at {max_steps_per_chunk}')
    if spc < 1:
        raise AssertionError('Not enough memory for the given grid!')
    idxs = np.arange(0, n_steps, spc)
    idxs = np.concatenate([idxs, [n_steps]]).astype(np.int64)
    return Plan(idxs, np.array([0, n_rows], dtype=np.int64), spc, dev_b,
                host_b, dev_budget, host_budget, notes)


def _host_free():
    try:
        import psutil
        return int(psutil.virtual_memory().available)
    except Exception:
        return 1 << 34


def interpolate_series(prob, vals, vgs, methods, *, writer=None,
                       nmrl_prcn=None, max_steps_per_chunk=None,
                       max_mem_usage_ratio=1.0, first_chunk=64,
                       keep_fields=False, **interp_kw):
    """Interpolate every step of vals [T x N] (host or device) in chunks.

    The first chunk (at most ``first_chunk`` steps) also sizes the
    library's workspaces; the plan for the rest comes from the device
    memory free after it.  ``writer`` (an NcWriter, or anything with
    write(label, t0, fields, nmrl_prcn)) receives every label's float32
    fields chunk by chunk, rounded to ``nmrl_prcn`` decimals.  interp_kw
    are interpolate()'s keywords (min_var_thr, cuts, neb_sel_mthd, ...).
    Returns (plan, prblm, fields) where fields holds the float32 host
    arrays when keep_fields, else None.
    """
    T = int(np.shape(vals)[0])
    C = prob.n_cells
    labels = [m[1] for m in methods if m[1] != 'EST_VARS_OK']
    want_var = interp_kw.get('est_vars_flag', True) and any(
        m[0] == 'OK' for m in methods)
    out_labels = labels + (['EST_VARS_OK'] if want_var else [])
    kept = ({lab: np.empty((T, C), dtype=np.float32) for lab in out_labels}
            if keep_fields else None)
    prblm = []
    copy_stream = torch.cuda.Stream(device=prob.device)
    pending = []  # (thread, pinned buffers) of the chunks being written
    lock = threading.Lock()

    def run_chunk(t0, t1, pin):
        f, _, prb = _steps.interpolate(prob, vals[t0:t1], vgs[t0:t1]
                                       if vgs is not None else None,
                                       methods, **interp_kw)
        prblm.extend(int(p) + t0 for p in prb)
        done = torch.cuda.Event()
        done.record()
        with torch.cuda.stream(copy_stream):
            copy_stream.wait_event(done)
            for lab in out_labels:
                pin[lab][:t1 - t0].copy_(f[lab], non_blocking=True)
                # the field's memory must outlive the copy on this stream
                f[lab].record_stream(copy_stream)
        ev = torch.cuda.Event()
        ev.record(copy_stream)
        return ev

    def write_chunk(ev, t0, t1, pin):
        ev.synchronize()
        for lab in out_labels:
            a = pin[lab][:t1 - t0].numpy()
            if kept is not None:
                kept[lab][t0:t1] = a
            if writer is not None:
                with lock:
                    writer.write(lab, t0, a.copy(), nmrl_prcn)

    def pinned(k):
        return {lab: torch.empty((k, C), dtype=torch.float32,
                                 pin_memory=True) for lab in out_labels}

    t1 = min(T, first_chunk)
    pin0 = pinned(t1)
    ev = run_chunk(0, t1, pin0)
    th = threading.Thread(target=write_chunk, args=(ev, 0, t1, pin0))
    th.start()
    pending.append((th, pin0))
    free, _ = torch.cuda.mem_get_info(prob.device)
    n_rows = getattr(prob, 'n_rows', 1)
    plan = plan_steps(T - t1, C, n_rows, methods, device_free=free,
                      host_free=_host_free(),
                      est_vars_flag=interp_kw.get('est_vars_flag', True),
                      max_steps_per_chunk=max_steps_per_chunk,
                      max_mem_usage_ratio=max_mem_usage_ratio) \
        if T > t1 else None
    if plan is not None:
        plan.step_idxs = np.concatenate([[0], plan.step_idxs + t1])
        bufs = [pinned(plan.steps_per_chunk), pinned(plan.steps_per_chunk)]
        for i in range(len(plan.step_idxs) - 2):
            a, b = int(plan.step_idxs[i + 1]), int(plan.step_idxs[i + 2])
            pin = bufs[i % 2]
            # the writer of the chunk that used this buffer must be done
            for th_, p_ in list(pending):
                if p_ is pin:
                    th_.join()
                    pending.remove((th_, p_))
            ev = run_chunk(a, b, pin)
            th = threading.Thread(target=write_chunk, args=(ev, a, b, pin))
            th.start()
            pending.append((th, pin))
    else:
        plan = Plan(np.array([0, T], dtype=np.int64),
                    np.array([0, n_rows], dtype=np.int64), T,
                    *output_bytes_per_step(C, methods,
                                           interp_kw.get('est_vars_flag',
                                                         True)),
                    int(free), _host_free(), ['one chunk'])
    for th_, _ in pending:
        th_.join()
    return plan, sorted(set(prblm)), kept
