"""Autotuner and latency report for ragged multi-stream batches.

    python -m sdbe_amd.tune [--blocks 17,64,128,5,200,33] [--reps 10]

Workload: 6 streams (3 beams x 2 sidebands, every second one a spectrally
flipped lower sideband) of uneven block counts, both polarisations, per-stream
sigma ~ U[0.5, 2] seen through a unit-sigma 2-bit quantiser and a random tone,
PER_BLOCK thresholds, VDIF frames out.  To keep set-up short each stream's
spectra repeat a seeded 8-block pattern (the pipeline's cost does not depend
on the values).

The tuner times one sdbe_process_blocks call over the whole batch for every
(chunk, stage-C1 tiles per workgroup, stage-C1 interleave, HIP graph on/off)
candidate -- the knobs that set how many block-pols share a launch, how much
work each stage-C1 workgroup keeps in flight and which workgroups read
neighbouring tiles together -- and reports the fastest.  Latency is reported two ways with the best
configuration: per call for the whole batch, and per call when every stream
hands in one block at a time (the latency-bound serving case).  Prints one
JSON object per line.
"""
from __future__ import annotations

import argparse
import itertools
import json
import sys
import time

import numpy as np

from . import _lib as L
from . import synth

SAMPLES_PER_BP = 1277952


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def make_batch(seed, n_blocks, period=8):
    import torch
    lsb = [i % 2 == 1 for i in range(len(n_blocks))]
    pattern = synth.batch_streams(seed, n_blocks=[min(period, n) for n in n_blocks], lsb=lsb)
    data = []
    for d, n in zip(pattern, n_blocks):
        pb = d["n_blocks"]
        reps = -(-n // pb)
        pols = []
        for p in range(2):
            t = torch.from_numpy(d["beng"][p]).cuda()
            pols.append(t.repeat(reps, 1, 1)[:39 * n].contiguous())
        data.append(dict(beng=pols, n_blocks=n, lsb=d["lsb"], sigma=d["sigma"]))
    return data


def streams_for(stage, data, blocks=None, outs=None):
    """BatchStream objects over `data`; `blocks` selects one block per stream."""
    import torch
    from .stage import BatchStream
    out = []
    for i, d in enumerate(data):
        if blocks is None:
            b0, b1, nb = d["beng"][0], d["beng"][1], d["n_blocks"]
        else:
            j = blocks[i] % d["n_blocks"]
            b0, b1, nb = d["beng"][0][39 * j:39 * (j + 1)], d["beng"][1][39 * j:39 * (j + 1)], 1
        o = outs[i] if outs is not None else torch.empty((2, 64 * nb, stage.frame_bytes), dtype=torch.uint8,
                                                        device="cuda")
        out.append(BatchStream(b0, b1, o, flip=d["lsb"], policy=L.THRESH_PER_BLOCK, station="Sm",
                               secs=1000, df=0, ref_epoch=30, psn=1))
    return out


def time_calls(stage, make_streams, reps, warmup=2):
    import torch
    from .stage import process_blocks
    streams = make_streams()
    for _ in range(warmup):
        process_blocks(stage, streams)
    torch.cuda.synchronize()
    ms = []
    for _ in range(reps):
        t0 = time.perf_counter()
        process_blocks(stage, streams)
        torch.cuda.synchronize()
        ms.append((time.perf_counter() - t0) * 1e3)
    return np.array(ms)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--blocks", default="17,64,128,5,200,33")
    ap.add_argument("--reps", type=int, default=10)
    ap.add_argument("--seed", type=int, default=404)
    ap.add_argument("--chunks", default="64,128,256,512,1024")
    ap.add_argument("--tiles", default="4,8,16")
    ap.add_argument("--interleave", default="1,16")
    args = ap.parse_args()
    import torch
    from .stage import SdbeStage

    n_blocks = [int(x) for x in args.blocks.split(",")]
    n_bp = 2 * sum(n_blocks)
    log("[tune] generating %d streams, %d block-pols" % (len(n_blocks), n_bp))
    data = make_batch(args.seed, n_blocks)
    results = []
    for chunk, tiles, ilv, graph in itertools.product([int(x) for x in args.chunks.split(",")],
                                                      [int(x) for x in args.tiles.split(",")],
                                                      [int(x) for x in args.interleave.split(",")], (False, True)):
        st = SdbeStage(out_mode=L.OUT_FRAMES, chunk=chunk, c1_tiles_per_wg=tiles, c1_interleave=ilv,
                       use_graph=graph, max_batch=n_bp)
        ms = time_calls(st, lambda: streams_for(st, data), args.reps)
        st.close()
        torch.cuda.empty_cache()
        r = dict(chunk=chunk, c1_tiles_per_wg=tiles, c1_interleave=ilv, graph=graph,
                 batch_ms_p50=round(float(np.median(ms)), 3),
                 gsample_s=round(n_bp * SAMPLES_PER_BP / (np.median(ms) * 1e-3) / 1e9, 2))
        log("[tune] %s" % r)
        results.append(r)
    best = max(results, key=lambda r: r["gsample_s"])
    print(json.dumps(dict(kind="autotune", streams=n_blocks, block_pols=n_bp, candidates=results, best=best)))

    st = SdbeStage(out_mode=L.OUT_FRAMES, chunk=best["chunk"], c1_tiles_per_wg=best["c1_tiles_per_wg"],
                   c1_interleave=best["c1_interleave"], use_graph=best["graph"], max_batch=n_bp)
    ms = time_calls(st, lambda: streams_for(st, data), max(args.reps, 20))
    print(json.dumps(dict(kind="latency_whole_batch", block_pols=n_bp, p50_ms=round(float(np.percentile(ms, 50)), 3),
                          p99_ms=round(float(np.percentile(ms, 99)), 3),
                          per_block_us_p50=round(float(np.percentile(ms, 50)) * 1e3 / sum(n_blocks), 2))))
    # one block per stream per call: the latency of a block from hand-in to frames out
    rng = np.random.default_rng(args.seed)
    outs = [torch.empty((2, 64, st.frame_bytes), dtype=torch.uint8, device="cuda") for _ in n_blocks]
    lat = time_calls(st, lambda: streams_for(st, data, blocks=list(rng.integers(0, 1 << 30, len(n_blocks))),
                                             outs=outs), 200, warmup=5)
    print(json.dumps(dict(kind="latency_one_block_per_stream", streams=len(n_blocks),
                          block_pols_per_call=2 * len(n_blocks), p50_ms=round(float(np.percentile(lat, 50)), 3),
                          p99_ms=round(float(np.percentile(lat, 99)), 3))))
    st.close()


if __name__ == "__main__":
    main()
