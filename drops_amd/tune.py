"""Autotuner of the persistent solvers' block geometry and exchange, and of the batched solve.

The persistent PCG/GMRES kernels give each block a contiguous range of rows.
More rows per block make a block's gather phase longer; more blocks make every
grid synchronisation (one slot per block to poll) longer.  The blocks exchange
their rows' values either as coherent atomics behind a slot synchronisation, or
as generation-tagged records gathered once per distinct column into LDS.
`tune_pcg` times solves of every (rows per block, exchange) candidate on the
caller's matrix and leaves the fastest setting on the context
(`dhip_tune("pcg_rows_per_block", ...)`, `dhip_tune("coherent_exchange", ...)`).
"""
import time

import numpy as np

from . import dhip as D

CANDIDATES = (128, 192, 256, 320, 384, 512)
EXCHANGES = (1, 2, 3)   # coherent atomics, tagged records with the halo in LDS, pipelined over them


def tune_pcg(ctx, A, b_dev, x_dev, zero_dev, n, tol=1e-8, precision=D.FP64, candidates=CANDIDATES,
             exchanges=EXCHANGES, reps=2):
    """Returns ((rows_per_block, exchange), {(rows, exchange): seconds}).  b_dev/x_dev/zero_dev are
    device pointers; x is reset from zero_dev before each solve."""
    timings = {}
    for exch in exchanges:
        ctx.tune("coherent_exchange", exch)
        for rows in candidates:
            ctx.tune("pcg_rows_per_block", rows)
            best = np.inf
            for _ in range(reps):
                ctx.copy_d2d(x_dev, zero_dev, 8 * n)
                t0 = time.perf_counter()
                ctx.pcg(A, b_dev, x_dev, tol=tol, precision=precision, where=D.DEVICE)
                best = min(best, time.perf_counter() - t0)
            timings[(rows, exch)] = best
    choice = min(timings, key=timings.get)
    ctx.tune("pcg_rows_per_block", choice[0])
    ctx.tune("coherent_exchange", choice[1])
    return choice, timings


# (threads, workgroups per CU, grid-wide pre-pass of classification and local matrices,
#  threads per problem of a separate PCG launch or 0 for none)
BATCH_CANDIDATES = ((512, 1, 0, 0), (512, 1, 1, 0), (256, 1, 0, 0), (512, 1, 1, 1024), (512, 1, 1, 512),
                    (512, 1, 0, 1024))


def tune_batch(ctx, batch, tol=1e-8, candidates=BATCH_CANDIDATES, reps=2):
    """Workgroup size, residency, pre-pass and separate PCG launch of the batched solve: times `reps`
    graph replays per candidate (best of; the capture is untimed) on the caller's problem set, after
    one untimed launch, and leaves the fastest on the context.  Returns
    ((threads, blocks_per_cu, prepass, solve_threads), {candidate: seconds})."""
    timings = {}
    batch.solve(tol=tol)   # first launch: code object load and first-touch costs stay out of the timings
    for threads, bpc, pre, sthreads in candidates:
        ctx.tune("batch_threads", threads)
        ctx.tune("batch_blocks_per_cu", bpc)
        ctx.tune("batch_prepass", pre)
        ctx.tune("batch_split", 1 if sthreads else 0)
        if sthreads:
            ctx.tune("batch_solve_threads", sthreads)
        batch.solve(tol=tol, graph=True)   # captures this candidate's graph (untimed)
        best = np.inf
        for _ in range(reps):
            t0 = time.perf_counter()
            batch.solve(tol=tol, graph=True)
            best = min(best, time.perf_counter() - t0)
        timings[(threads, bpc, pre, sthreads)] = best
    choice = min(timings, key=timings.get)
    ctx.tune("batch_threads", choice[0])
    ctx.tune("batch_blocks_per_cu", choice[1])
    ctx.tune("batch_prepass", choice[2])
    ctx.tune("batch_split", 1 if choice[3] else 0)
    if choice[3]:
        ctx.tune("batch_solve_threads", choice[3])
    return choice, timings
