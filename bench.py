"""Search-node fast-path benchmark (one GPU).

A step is one voltage block of the 96-antenna x 384-channel node stream
(16384 samples = 536.9 ms of data) through the whole path: int8-MFMA
correlator -> fused gridder + FFT imager -> image history -> fused
dedispersion + boxcar + float64 S/N statistics -> host candidate decode.
Blocks are generated on the GPU before timing (inputs resident in HBM).

Prints one JSON line.  ``value`` is the real-time factor (sum of block
durations / wall time); p50/p99 block-to-candidate latency, per-stage
times, the roofline of the dominant kernel and the CPU-oracle baseline
travel in the same line.
"""
from __future__ import annotations

import argparse
import json
import os
import platform
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for _p in (ROOT, os.path.join(ROOT, "dsa110-rt_amd")):
    if _p not in sys.path:
        sys.path.insert(0, _p)

import numpy as np
import torch

HBM_PEAK_GBPS = 8000.0
# LDS read throughput of the whole GPU, measured by dsa110-rt_amd/csrc/probe/lds_read_probe.hip
# (lane bytes per second, 256 CUs at the 2.1-2.3 GHz clock held under load): 16-byte reads of
# consecutive words 135.3 TB/s (243 B/clk/CU); ds_read2_b32 at clustered per-lane addresses --
# the dedispersion kernel's window reads, neighbouring DMs a few words apart -- 67.9 TB/s
LDS_PEAK_TBPS = 135.3
LDS_PATTERN_PEAK_TBPS = 67.9
INT8_PEAK_POPS = 5.0                           # MI355X dense int8 MFMA
MFMA_I8_PEAK_TOPS = 5000.0
FP32_PEAK_TFLOPS = 157.3


def measured_traffic(history: str):
    """HBM bytes per dedispersion launch for a ``history`` ("fp32" / "fp16")
    from the committed hardware-counter summary (rocprofv3 FETCH_SIZE /
    WRITE_SIZE, copy-calibrated), or None."""
    path = os.path.join(ROOT, "profiles", "dedisp_traffic_96ant_384ch_16384.md")
    key = f"traffic_bytes_per_launch[{history}]:"
    try:
        with open(path) as fh:
            for line in fh:
                if line.startswith(key):
                    return int(line.split(":")[1])
    except OSError:
        return None
    return None


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=12)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-host", action="store_true", help="skip the host-resident (PCIe-inclusive) leg")
    ap.add_argument("--no-cal", action="store_true", help="skip the calibrated (gains on operand load) leg")
    ap.add_argument("--history", choices=("fp16", "fp32"), default="fp32",
                    help="image history dtype of the fused form (fp32: LDS-direct window loads)")
    ap.add_argument("--form", choices=["chosen", "plain"], default="chosen")
    ap.add_argument("--dm-plan", choices=["even", "uneven", "stage1"], default="even",
                    help="uneven: 96-antenna 384-channel inputs with the three-range DM plan (1x/2x/4x, widths 1-64)")
    ap.add_argument("--serve", action="store_true",
                    help="serving variant: uneven blocks, Poisson arrivals, HBM-budget admission")
    ap.add_argument("--serve-blocks", type=int, default=200)
    ap.add_argument("--serve-rates", type=str, default="1.0,1.5,2.5")
    ap.add_argument("--calibrate-copy-mib", type=int, default=0,
                    help="run one device-to-device copy of this many MiB first (hardware-counter scale)")
    return ap.parse_args()


def workload(cfg):
    return (f"{cfg.n_ant} ant x 2 pol x {cfg.n_chan} ch (30.5 kHz, chgroup 0) x {cfg.n_samp} "
            f"samples/block, 4-bit; t_int {cfg.t_int} -> {cfg.n_steps} steps x {cfg.nbase} baselines; "
            f"{cfg.n_sub} subbands x {cfg.csf} ch gridded to {cfg.n_grid}^2 at "
            f"{cfg.grid_freq_GHz} GHz; {cfg.n_dm} DMs 0-{cfg.dm_max:g}; boxcars "
            f"{'/'.join(str(w) for w in cfg.widths)}")


def run_form(cfg, sc, blocks, steps, warmup, ring_dtype, plain=False, cal=False):
    """Time ``steps`` blocks after ``warmup``; returns wall time, per-block
    latency (block handed over -> candidate list on the host) and mean
    per-kernel HIP-event times (fused form).  ``cal``: seeded complex gains
    with 5% of the (antenna, channel, pol) gains flagged, applied on the
    correlator's operand load."""
    if plain:
        from dsx.plain import PlainNode
        node = PlainNode(cfg, sc.antpos_e, sc.antpos_n)
    else:
        from dsx.pipeline import SearchNode
        node = SearchNode(cfg, sc.antpos_e, sc.antpos_n, ring_dtype=ring_dtype)
        if cal:
            rng = np.random.default_rng(0xCA1)
            shp = (cfg.n_ant, cfg.n_chan, 2)
            g = rng.normal(1.0, 0.2, shp) + 1j * rng.normal(0.0, 0.2, shp)
            gr = torch.from_numpy(np.ascontiguousarray(g.real.transpose(1, 2, 0))).float().cuda()
            gi = torch.from_numpy(np.ascontiguousarray(g.imag.transpose(1, 2, 0))).float().cuda()
            node.set_calibration(gr, gi, cube_mask=torch.from_numpy(rng.random(shp) < 0.05).cuda())
    for i in range(warmup):
        node.process_block(blocks[i % len(blocks)])
    tune = None
    if not plain:
        # LDS layout (time tile, pair rows or not) picked by timing each candidate
        tune = node.autotune(node._check_block(cfg.n_samp))
        node.enable_timing()
    torch.cuda.synchronize()
    lat, stages, n_cands = [], {}, 0
    t_start = time.perf_counter()
    for i in range(steps):
        raw = blocks[(warmup + i) % len(blocks)]
        t0 = time.perf_counter()
        res = node.process_block(raw)
        lat.append(time.perf_counter() - t0)
        if not plain:
            for k, v in node.stage_ms().items():
                stages.setdefault(k, []).append(v)
        n_cands += len(res.candidates)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t_start
    return dict(wall=wall, lat=np.asarray(lat), stage_ms={k: float(np.mean(v)) for k, v in stages.items()},
                n_cands=n_cands, node=node, tune=tune)


def run_host_form(cfg, sc, blocks, steps, warmup, ring_dtype):
    """The chosen form fed from pinned host pages: each block's H2D copy runs
    on a copy stream while the previous block computes (``dsx.ingest``).
    Latency = page handed over (copy enqueued) -> candidate list on the host."""
    from dsx.ingest import HostIngest, PinnedPagePool, page_bytes
    from dsx.pipeline import SearchNode
    n_pages = min(4, len(blocks))
    pool = PinnedPagePool(n_pages, page_bytes(cfg, cfg.n_samp))
    for i, p in enumerate(pool.pages):
        p.copy_(blocks[i].reshape(-1).cpu())
    ingest = HostIngest(cfg)
    node = SearchNode(cfg, sc.antpos_e, sc.antpos_n, ring_dtype=ring_dtype)

    def run(n, timed):
        lat = []
        nxt = ingest.submit(pool.pages[0], cfg.n_samp, timed=timed)
        t_sub = time.perf_counter()
        for i in range(n):
            cur, t_cur = nxt, t_sub
            if i + 1 < n:
                nxt = ingest.submit(pool.pages[(i + 1) % n_pages], cfg.n_samp, timed=timed)
                t_sub = time.perf_counter()
            cur.wait()
            node.process_block(cur.raw, cfg.n_samp)
            ingest.release(cur)
            lat.append(time.perf_counter() - t_cur)
        return lat

    run(warmup, False)
    node.autotune(node._check_block(cfg.n_samp))
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    lat = np.asarray(run(steps, True))
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    out = {"value": steps * cfg.block_seconds / wall, "unit": "x real-time", "ms_per_step": wall / steps * 1e3,
           "latency_ms": {"p50": float(np.percentile(lat, 50) * 1e3), "p99": float(np.percentile(lat, 99) * 1e3)},
           "steps": steps, "warmup": warmup,
           "form": "chosen, raw pages in pinned host memory, double-buffered H2D on a copy stream overlapped "
                   "with the previous block's compute"}
    out.update(ingest.copy_rate())
    del node, ingest, pool
    return out


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or platform.machine()


def cpu_baseline(cfg, sc, raw_c2):
    """The reference's CPU path restated in PyTorch on the host's cores
    (``oracle.torch_cpu``: torch.matmul correlator, index_add_ gridder,
    torch.fft.ifft2 + fftshift imager, shift-and-add + cumsum boxcar +
    chunk maxima + top-k decode): the 16-antenna 64-channel stream in full
    (3 blocks of 2048 samples, its DM-250 burst) and a bounded sample of one
    96-antenna 384-channel block (64 steps correlated and imaged, 8 of 256 DM
    trials over 1/16 of the pixels for all 2048 centres) scaled to the block."""
    from oracle import image as oimage
    from oracle import torch_cpu as tcpu
    from dsx import config as dcfg
    from dsx import synth
    n_thr = len(os.sched_getaffinity(0))
    torch.set_num_threads(n_thr)
    # ---- 16-antenna stream, end to end
    c1 = dcfg.small_node_config()
    sc1 = synth.make_scenario(c1, synth.small_node_bursts(c1))
    rows, cols, _ = oimage.cell_map(sc1.antpos_e, sc1.antpos_n, c1.freq_out_GHz, c1.n_grid)
    mask = oimage.edge_mask(c1.n_grid, 1)
    sh1 = c1.shift_table()
    keep = int(sh1.max()) + c1.w_max // 2 + 65
    hist = torch.zeros((c1.n_sub, 0, c1.n_grid ** 2))
    tau0, st_n = 0, np.zeros(c1.n_dm)
    st_mean = torch.zeros((c1.n_dm, c1.n_grid ** 2), dtype=torch.float64)
    st_m2 = torch.zeros_like(st_mean)
    n_blk, t1, cands1 = 6, [], 0
    raws = [sc1.block(b * c1.n_samp, c1.n_samp) for b in range(n_blk)]
    for b in range(n_blk):
        t = time.perf_counter()
        vis = tcpu.correlate(raws[b], n_pkt=c1.n_pkt, n_ant=c1.n_ant, n_chan=c1.n_chan, t_int=c1.t_int, csf=c1.csf)
        img = tcpu.grid_images(vis, rows, cols, c1.n_grid, c1.n_sub, mask)
        hist = torch.cat([hist, img], dim=1)
        drop = max(0, hist.shape[1] - (keep + c1.n_steps))
        hist, tau0 = hist[:, drop:], tau0 + drop
        c_lo = b * c1.n_steps - 64
        bs, bt, bk, s1, s2, cnt = tcpu.search_block(hist, tau0, sh1, c_lo, c1.n_steps, c1.widths, st_n, st_mean,
                                                    st_m2, c1.min_stat_count)
        cands1 += len(tcpu.decode(bs, bt, bk, c1.n_grid, threshold=c1.threshold, n_top=c1.n_top,
                                  radius_t=c1.radius_t, radius_dm=c1.radius_dm, radius_lm=c1.radius_lm))
        nb = torch.from_numpy(cnt)[:, None]
        upd = nb > 0
        n_a = torch.from_numpy(st_n)[:, None]
        mean_b = torch.from_numpy(s1) / nb.clamp(min=1)
        delta = mean_b - st_mean
        n = (n_a + nb).clamp(min=1)
        st_m2 = torch.where(upd, st_m2 + (torch.from_numpy(s2) - torch.from_numpy(s1) ** 2 / nb.clamp(min=1))
                            + delta ** 2 * n_a * nb / n, st_m2)
        st_mean = torch.where(upd, st_mean + delta * nb / n, st_mean)
        st_n = st_n + cnt
        if b >= n_blk - 3:                       # the last 3 blocks (statistics burnt in) are timed
            t1.append(time.perf_counter() - t)
    c1_sec = float(np.mean(t1))
    # ---- one 96-antenna block, bounded sample
    n_st = 64
    raw = raw_c2[: n_st * cfg.t_int // 2]
    t = time.perf_counter()
    vis = tcpu.correlate(raw, n_pkt=raw.shape[0], n_ant=cfg.n_ant, n_chan=cfg.n_chan, t_int=cfg.t_int, csf=cfg.csf)
    t_corr = time.perf_counter() - t
    rows, cols, _ = oimage.cell_map(sc.antpos_e, sc.antpos_n, cfg.grid_freq_out_GHz, cfg.n_grid)
    t = time.perf_counter()
    tcpu.grid_images(vis, rows, cols, cfg.n_grid, cfg.n_sub, oimage.edge_mask(cfg.n_grid))
    t_img = time.perf_counter() - t
    sh = cfg.shift_table()
    n_pix = cfg.n_grid * cfg.n_grid // 16
    L = cfg.n_steps
    hl = int(sh.max()) + 64 + cfg.w_max + L
    histc = torch.randn((cfg.n_sub, hl, n_pix), generator=torch.Generator().manual_seed(1))
    st_n2 = np.full(cfg.n_dm, 1e4)
    dms = list(range(0, cfg.n_dm, cfg.n_dm // 8))
    t = time.perf_counter()
    tcpu.search_block(histc, 0, sh, hl - L, L, cfg.widths, st_n2, torch.zeros((cfg.n_dm, n_pix), dtype=torch.float64),
                      torch.full((cfg.n_dm, n_pix), 1e4 * cfg.n_sub, dtype=torch.float64), 0, dms=dms)
    t_dd = time.perf_counter() - t
    stages = {"correlate": t_corr * L / n_st, "image": t_img * L / n_st,
              "dedisp_detect": t_dd * (cfg.n_dm / len(dms)) * 16}
    per_block = sum(stages.values())
    return {
        "value": cfg.block_seconds / per_block,
        "unit": "x real-time",
        "cores": n_thr,
        "kind": "port",
        "sample": (f"PyTorch-CPU restatement of the reference path on {n_thr} threads: one 96-ant 384-ch "
                   f"16384-sample block from 64 correlated + imaged steps and {len(dms)} of {cfg.n_dm} DM trials "
                   f"over 1/16 of the pixels for all {L} centres, scaled to the block"),
        "seconds_per_block": per_block,
        "stage_seconds": stages,
        "cpu": cpu_model(),
        "small_node_stream": {"value": c1.block_seconds / c1_sec, "unit": "x real-time",
                              "seconds_per_block": c1_sec, "blocks_timed": len(t1), "candidates": cands1,
                              "workload": "16 ant x 64 ch x 2048-sample blocks, 64 DMs 0-500, widths 1-8, "
                                          "end to end (6 blocks, last 3 timed)"},
    }


def serve_main(args):
    """Serving variant: a stream of uneven blocks {4096, 8192, 16384, 32768}
    samples with 20 bursts, Poisson arrivals at each rate x real time,
    admission under a 200 GB HBM budget; one JSON line with p50/p99
    block-to-candidate latency and sustained real-time factor per rate."""
    from dsx import config as dcfg
    from dsx import synth
    from dsx.pipeline import SearchNode
    from dsx.serve import BlockRequest, NodeServer, poisson_arrivals, summarize
    cfg = dcfg.search_node_config()
    rng = np.random.default_rng(0xC5)
    lens = [int(x) for x in rng.choice([4096, 8192, 16384, 32768], size=args.serve_blocks)]
    starts = np.concatenate([[0], np.cumsum(lens)[:-1]]).astype(np.int64)
    total_steps = int(sum(lens) // cfg.t_int)
    bursts = synth.random_bursts(cfg, 20, 0xC5B, snr_lo=8.0, snr_hi=50.0, t_lo_step=2048,
                                 t_hi_step=total_steps - 256)
    sc = synth.make_scenario(cfg, bursts)
    gs = synth.GpuSynth(sc)
    durs = [n * cfg.t_step_us / cfg.t_int * 1e-6 for n in lens]
    out = {"metric": "serving: p50/p99 block->candidate latency and sustained real-time factor",
           "unit": "x real-time", "n_gpus": 1, "data": "synthetic (seeded 4-bit noise + 20 dispersed bursts)",
           "config": {"workload": f"{args.serve_blocks} blocks of {{4096, 8192, 16384, 32768}} samples, "
                                  f"96 ant x 384 ch, Poisson arrivals, 200 GB HBM budget",
                      "data_seconds": float(sum(durs))}, "rates": {}}
    for rate in [float(r) for r in args.serve_rates.split(",")]:
        node = SearchNode(cfg, sc.antpos_e, sc.antpos_n, max_block_steps=32768 // 8,
                          n_slots=2)
        server = NodeServer([node], hbm_budget_bytes=200 << 30)
        arr = poisson_arrivals(durs, rate, 0xA77 + int(rate * 10))
        reqs = [BlockRequest(stream_id=0, seq=k, n_samp=n, t_arrival=float(arr[k]),
                             make=lambda st, a=int(s0), b=n: gs.block(a, b))
                for k, (s0, n) in enumerate(zip(starts, lens))]
        print(f"serving {len(reqs)} blocks at {rate}x real time", file=sys.stderr, flush=True)
        done = server.serve(reqs, realtime=True)
        summ = summarize(done, cfg)
        summ["resident_bytes"] = server.resident_bytes()
        out["rates"][f"{rate:g}x"] = summ
        print(f"  {rate}x: p99 {summ['latency_ms']['p99']:.1f} ms, sustained {summ['sustained_rtf']:.2f}x",
              file=sys.stderr, flush=True)
        del server, node
        torch.cuda.empty_cache()
    r1 = out["rates"].get("1x") or next(iter(out["rates"].values()))
    out["value"] = r1["sustained_rtf"]
    out["latency_ms"] = r1["latency_ms"]
    print(json.dumps(out), flush=True)


def uneven_main(args):
    """96-antenna 384-channel inputs through the uneven DM plan (0-300: 128 trials at full
    resolution; 300-800: 96 at 2x; 800-1500: 64 at 4x; widths 1-64)."""
    from dsx import config as dcfg
    from dsx import synth
    from dsx.pipeline import MultiRangeNode
    cfg = dcfg.uneven_dm_config()
    bursts = synth.random_bursts(cfg.with_(dm_ranges=()), 8, 0xC4B, t_lo_step=cfg.n_steps,
                                 t_hi_step=(args.warmup + args.steps) * cfg.n_steps)
    sc = synth.make_scenario(cfg.with_(dm_ranges=()), bursts)
    gs = synth.GpuSynth(sc)
    blocks = [gs.block(i * cfg.n_samp, cfg.n_samp) for i in range(args.warmup + args.steps)]
    node = MultiRangeNode(cfg, sc.antpos_e, sc.antpos_n)
    for i in range(args.warmup):
        node.process_block(blocks[i])
    tune = node.autotune()
    torch.cuda.synchronize()
    lat, n_c = [], 0
    t_start = time.perf_counter()
    for i in range(args.steps):
        t0 = time.perf_counter()
        res = node.process_block(blocks[args.warmup + i])
        lat.append(time.perf_counter() - t0)
        n_c += len(res.candidates)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t_start
    lat = np.asarray(lat)
    print(json.dumps({
        "metric": "real-time factor + p99 block->candidate latency, 96-ant 384-ch node, uneven DM plan",
        "value": args.steps * cfg.block_seconds / wall, "unit": "x real-time", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": wall / args.steps * 1e3,
        "higher_is_better": True, "data": "synthetic (seeded 4-bit noise + 8 dispersed bursts)",
        "config": {"workload": "96-antenna 384-channel inputs; DM 0-300 x128 @1x, 300-800 x96 @2x, 800-1500 x64 @4x; "
                               "boxcars 1-64 samples of each range's resolution"},
        "latency_ms": {"p50": float(np.percentile(lat, 50) * 1e3), "p99": float(np.percentile(lat, 99) * 1e3),
                       "block_duration": cfg.block_seconds * 1e3},
        "dedisp_layouts": [{"tile": nd.plans[cfg.n_steps // f][0].tile, "paired": nd.plans[cfg.n_steps // f][0].paired,
                            "autotune_ms": t} for (nd, f, _, _), t in zip(node.ranges, tune)],
        "candidates": n_c}), flush=True)


def stage1_main(args):
    """96-antenna 384-channel inputs with stage 1: four coarse DMs remove the delay inside
    each subband per 8-channel group before gridding, static sky subtracted over 8
    blocks; 256 fine DMs 0-1500 searched on their nearest coarse DM's stream."""
    from dsx import config as dcfg
    from dsx import synth
    from dsx.pipeline import CoarseDMNode
    cfg = dcfg.coarse_dm_config()
    hist_dt = torch.float16 if args.history == "fp16" else torch.float32
    bursts = synth.random_bursts(cfg.with_(coarse_dms=()), 8, 0xC2B, t_lo_step=cfg.n_steps,
                                 t_hi_step=(args.warmup + args.steps) * cfg.n_steps)
    sc = synth.make_scenario(cfg.with_(coarse_dms=()), bursts)
    gs = synth.GpuSynth(sc)
    blocks = [gs.block(i * cfg.n_samp, cfg.n_samp) for i in range(args.warmup + args.steps)]
    node = CoarseDMNode(cfg, sc.antpos_e, sc.antpos_n, ring_dtype=hist_dt)
    for i in range(args.warmup):
        node.process_block(blocks[i])
    tune = node.autotune()
    node.enable_timing()
    torch.cuda.synchronize()
    lat, n_c, stages = [], 0, {}
    t_start = time.perf_counter()
    for i in range(args.steps):
        t0 = time.perf_counter()
        res = node.process_block(blocks[args.warmup + i])
        lat.append(time.perf_counter() - t0)
        for k, v in node.stage_ms().items():
            stages.setdefault(k, []).append(v)
        n_c += len(res.candidates)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t_start
    lat = np.asarray(lat)
    print(json.dumps({
        "metric": "real-time factor + p99 block->candidate latency, 96-ant 384-ch node, stage 1 + static sky",
        "value": args.steps * cfg.block_seconds / wall, "unit": "x real-time", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": wall / args.steps * 1e3,
        "higher_is_better": True, "data": "synthetic (seeded 4-bit noise + 8 dispersed bursts)",
        "dtype": f"int8->int32 (corr), fp32 FFT, {args.history} store / fp32 acc, fp64 stats",
        "config": {"workload": (workload(cfg) + f"; stage 1: coarse DMs {list(cfg.coarse_dms)}, "
                                f"{cfg.stage1_csf}-channel groups, delay {node.stage1_delay} steps; static sky "
                                f"{cfg.static_sky_blocks} blocks")},
        "latency_ms": {"p50": float(np.percentile(lat, 50) * 1e3), "p99": float(np.percentile(lat, 99) * 1e3),
                       "block_duration": cfg.block_seconds * 1e3},
        "stages_ms": {k: float(np.mean(v)) for k, v in stages.items()},
        "groups": [{"coarse_dm": nd.cfg.stage1_dm, "n_dm": nd.cfg.n_dm, "dedisp_tile": nd.plans[cfg.n_steps][0].tile,
                    "autotune_ms": t} for (nd, _), t in zip(node.groups, tune)],
        "candidates": n_c}), flush=True)


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported on this one-GPU "
              "build (the node pipeline shards by frequency and DM across GPUs)",
              file=sys.stderr)
        sys.exit(2)
    if args.serve:
        serve_main(args)
        return
    if args.dm_plan == "uneven":
        uneven_main(args)
        return
    if args.dm_plan == "stage1":
        stage1_main(args)
        return
    from dsx import config as dcfg
    from dsx import synth
    if args.calibrate_copy_mib > 0:
        n = args.calibrate_copy_mib << 20
        a = torch.ones(n, dtype=torch.uint8, device="cuda")
        b = torch.empty_like(a)
        torch.cuda.synchronize()
        b.copy_(a)
        torch.cuda.synchronize()
        del a, b
    cfg = dcfg.search_node_config()
    bursts = synth.random_bursts(cfg, 8, 0xC2B, t_lo_step=cfg.n_steps,
                                 t_hi_step=(args.warmup + args.steps) * cfg.n_steps)
    sc = synth.make_scenario(cfg, bursts)
    gs = synth.GpuSynth(sc)
    n_blocks = args.warmup + args.steps
    blocks = [gs.block(i * cfg.n_samp, cfg.n_samp) for i in range(n_blocks)]
    torch.cuda.synchronize()

    hist_dt = torch.float16 if args.history == "fp16" else torch.float32
    chosen = run_form(cfg, sc, blocks, args.steps, args.warmup, hist_dt)
    node = chosen.pop("node")
    wall = chosen["wall"]
    rtf = args.steps * cfg.block_seconds / wall
    lat = chosen["lat"]
    # dominant kernel: fused dedispersion + detection
    dd_ms = chosen["stage_ms"]["dedisp"]
    n_pix = cfg.n_grid * cfg.n_grid
    ring_bytes = cfg.n_sub * cfg.n_steps * n_pix * (2 if args.history == "fp16" else 4)
    stat_bytes = 2 * 2 * n_pix * cfg.n_dm * 8
    alg_bytes = ring_bytes + stat_bytes
    achieved = alg_bytes / (dd_ms * 1e-3) / 1e9
    plan = node.plans[cfg.n_steps][0]
    maxt = 8 if plan.n_terms <= 8 else 16 if plan.n_terms <= 16 else 0
    is_async = plan.dma
    tname = "_Float16" if args.history == "fp16" else "float"
    # template arguments <T, NW, PAIRED, MAXT, ASYNC, QB, EXACT> of the kept instance: plans with
    # exactly 16 terms (8 with pair rows) run the pipelined exact-term sums
    exact = (plan.n_terms == 8) if plan.paired else (plan.n_terms == 16)
    exact = exact and not plan.rows_bf16
    kname = (f"dedisp_kernel<{tname}, {len(cfg.widths)}, {str(plan.paired).lower()}, {maxt}, "
             f"{str(is_async).lower()}, {str(plan.rows_bf16).lower()}, {str(exact).lower()}>")
    out = {
        "metric": "real-time factor + p99 block->candidate latency, 96-ant 384-ch node; chosen vs plain",
        "value": rtf,
        "unit": "x real-time",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": f"int8->int32 (corr), fp32 FFT, {args.history} store / fp32 acc, fp64 stats",
        "data": "synthetic (seeded 4-bit noise + 8 dispersed bursts, generated on the GPU)",
        "config": {"workload": workload(cfg), "form": "chosen", "block_ms": cfg.block_seconds * 1e3},
        "latency_ms": {"p50": float(np.percentile(lat, 50) * 1e3), "p99": float(np.percentile(lat, 99) * 1e3),
                       "block_duration": cfg.block_seconds * 1e3},
        "stages_ms": chosen["stage_ms"],
        "candidates": chosen["n_cands"],
        "dedisp_layout": {"tile": plan.tile, "paired": plan.paired, "async": is_async, "lds_bytes": plan.lds_bytes,
                          "autotune_ms": chosen["tune"]},
        "roofline": {"kernel": f"dsx_dedisp_detect ({kname})", "bound": "hbm",
                     "achieved": achieved, "peak": HBM_PEAK_GBPS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBPS, "traffic": measured_traffic(args.history),
                     "alg_bytes_per_launch": alg_bytes, "avg_ms": dd_ms,
                     "note": (f"algorithmic bytes = {args.history} subband images read once "
                              f"({ring_bytes}) + float64 mean/m2 read+write ({stat_bytes}); the "
                              "kernel is not HBM-bound: per (DM, step, pixel) it reads 16 float32 window "
                              "samples from LDS at per-lane addresses (64 lane bytes) and issues ~29 VALU "
                              "instructions; it is bound by LDS read throughput for that access pattern "
                              "(roofline.lds.pattern_frac)")},
    }
    ops = cfg.n_dm * cfg.n_steps * n_pix * (cfg.n_sub + 3 * len(cfg.widths) + 4)
    out["roofline"]["valu_tops"] = ops / (dd_ms * 1e-3) / 1e12
    # LDS reads of the subband windows, 4 lane bytes per subband term of every dedispersed
    # sample, against the measured LDS read throughput (16-byte consecutive reads, and
    # ds_read2_b32 at per-lane addresses: the kernel's own pattern)
    lds_bytes = 4 * cfg.n_sub * cfg.n_dm * cfg.n_steps * n_pix
    lds_tbps = lds_bytes / (dd_ms * 1e-3) / 1e12
    out["roofline"]["lds"] = {"bytes_per_launch": lds_bytes, "achieved_TBps": lds_tbps,
                              "peak_TBps": LDS_PEAK_TBPS, "frac": lds_tbps / LDS_PEAK_TBPS,
                              "pattern_peak_TBps": LDS_PATTERN_PEAK_TBPS,
                              "pattern_frac": lds_tbps / LDS_PATTERN_PEAK_TBPS}
    # every stage against the gfx950 peaks: algorithmic HBM bytes per block / HIP-event stage time,
    # and the correlator's int8 MFMA ops / time against 5.0 POP/s dense
    esz = 2 if args.history == "fp16" else 4
    vis_bytes = cfg.n_steps * cfg.n_chan_out * cfg.nbase * 2 * 4
    raw_bytes = cfg.n_pkt * cfg.n_ant * cfg.n_chan * 4
    img_bytes = cfg.n_sub * cfg.n_steps * n_pix * esz
    hist_bytes = 2 * cfg.n_sub * n_pix * node.hist * esz
    stage_bytes = {"corr": raw_bytes + vis_bytes, "image": vis_bytes + img_bytes,
                   "ring": 2 * img_bytes + hist_bytes, "dedisp": alg_bytes}
    stages = {}
    for k, b in stage_bytes.items():
        ms = chosen["stage_ms"][k]
        gbps = b / (ms * 1e-3) / 1e9
        stages[k] = {"ms": ms, "alg_bytes": b, "GBps": gbps, "hbm_frac": gbps / HBM_PEAK_GBPS}
    corr_ops = 8 * cfg.nbase * cfg.n_chan * 2 * cfg.n_samp
    stages["corr"]["int8_pops"] = corr_ops / (chosen["stage_ms"]["corr"] * 1e-3) / 1e15
    stages["corr"]["int8_mfma_frac"] = stages["corr"]["int8_pops"] / INT8_PEAK_POPS
    out["stage_roofline"] = stages
    del node
    if not args.no_cal:
        torch.cuda.empty_cache()
        n_cal = min(args.steps, 4)
        calr = run_form(cfg, sc, blocks, n_cal, 1, hist_dt, cal=True)
        calr.pop("node")
        out["calibrated"] = {"value": n_cal * cfg.block_seconds / calr["wall"], "unit": "x real-time",
                             "ms_per_step": calr["wall"] / n_cal * 1e3, "steps": n_cal, "warmup": 1,
                             "stages_ms": calr["stage_ms"], "candidates": calr["n_cands"],
                             "form": ("chosen, complex gains + RFI mask on the correlator's operand load "
                                      "(fp16 MFMA, float32 accumulation), float32 visibilities gridded by "
                                      "float LDS atomics")}
    if not args.no_host:
        torch.cuda.empty_cache()
        out["host_resident"] = run_host_form(cfg, sc, blocks, args.steps, args.warmup, hist_dt)
    if not args.no_plain:
        torch.cuda.empty_cache()
        n_plain = 2
        plain = run_form(cfg, sc, blocks, n_plain, 1, torch.float32, plain=True)
        plain.pop("node")
        rtf_plain = n_plain * cfg.block_seconds / plain["wall"]
        out["plain"] = {"value": rtf_plain, "ms_per_step": plain["wall"] / n_plain * 1e3,
                        "latency_p99_ms": float(np.percentile(plain["lat"], 99) * 1e3),
                        "steps": n_plain, "warmup": 1, "candidates": plain["n_cands"],
                        "form": ("plain FP32, unbatched: complex64 GEMM correlator, index_add_ gridder, "
                                 "rocFFT imager, one dedispersion + boxcar pass per DM trial")}
        out["chosen_over_plain"] = rtf / rtf_plain
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(cfg, sc, blocks[0].cpu().numpy())
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
