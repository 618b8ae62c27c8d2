#!/usr/bin/env python3
"""Benchmark: batched MC-Glauber Pb+Pb on a 261^2 grid (events/s).

Workload (one "step"): one batch of --events minimum-bias Pb+Pb events through
the whole hot path on one MI355X -- impact parameter and Philox nucleon
sampling (Woods-Saxon R=6.62 fm, a=0.546 fm, hard core 0.4 fm, A=208),
acceptance (Ncoll > 0, 2 <= Npart <= 500), A x B disk test at sigma_NN=64 mb,
Gamma fluctuations (k = 1/theta = 0.8), FP32 Gaussian smearing (w=0.5 fm,
alpha=0.14) onto 261 x 261 cells (dx=0.1 fm, +-13 fm) and FP64 eps_2..eps_5 /
<r^n> moments.  Inputs are generated on the device (counter-based RNG), so the
timed region holds no host transfers.

Prints one JSON line (see the keys below).  The CPU baseline times the FP64
oracle's event loop (a restatement of the reference's MakeDensity event
loop, no file output) on a bounded sample, one process per core.
"""
import argparse
import json
import math
import multiprocessing as mp
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

FP32_PEAK = 157.3e12          # MI355X FP32 (MFMA, and VALU with packed FP32)
FP32_PEAK_NONPACKED = 78.6e12  # one FP32 FMA per lane per 4 cycles per wave64: the code is built without packed FP32
HBM_PEAK = 8.0e12
BF16_MFMA_PEAK = 2.5e15       # dense BF16 MFMA


def mfma_peak(smear_path):
    """The matrix-core peak of a smearing path's MFMA precision (FP32 MFMA
    for path 1, BF16 for the split-precision paths 3 and 4)."""
    return BF16_MFMA_PEAK if smear_path in (3, 4) else FP32_PEAK


SMEAR_PATHS = ["separable VALU tiles", "FP32 MFMA Gx.diag(w).Gy^T + exact-cut pass", "per-cell exponentials",
               "BF16x3 MFMA Gx.diag(w).Gy^T + exact-cut pass", "BF16x6 MFMA Gx.diag(w).Gy^T + exact-cut pass"]


def bench_params(nev, smear_path=0):
    from supermc_amd import default_params
    return default_params(
        which_mc_model=5, sub_model=1, shape_of_nucleons=4, gauss_nucl_width=0.5, shape_of_entropy=2,
        collision_criterion=1, Aproj=208, Atarg=208, proj_ws_R=6.62, proj_ws_a=0.546, targ_ws_R=6.62,
        targ_ws_a=0.546, nucleon_min_dist=0.4, sigma_nn_mb=64.0, alpha=0.14, cc_fluctuation_model=6,
        cc_fluctuation_Gamma_theta=1.25, bmin=0.0, bmax=20.0, Npmin=2, Npmax=500, maxx=13.0, maxy=13.0, dx=0.1,
        dy=0.1, finalFactor=1.0, ecc_from_order=2, ecc_to_order=5, max_events=nev, max_pairs=4096, smear_path=smear_path)


CPU_PBPB = dict(A=208, B=208, R_a=6.62, a_a=0.546, R_b=6.62, a_b=0.546, dmin=0.4, siginNN=64.0, w=0.5, alpha=0.14,
                gamma_theta=1.25, fluct=1, bmin=0.0, bmax=20.0, npmin=2, npmax=500, maxx=13.0, dx=0.1, from_order=2,
                to_order=5)
CPU_AUAU = dict(CPU_PBPB, A=197, B=197, R_a=6.38, a_a=0.535, R_b=6.38, a_b=0.535, siginNN=42.0, fluct=0)


def _cpu_worker(args):
    seed, nev = args[:2]
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    cfg = oracle.OrcEventCfg(**(args[2] if len(args) > 2 else CPU_PBPB))
    t = time.perf_counter()
    n, _ = oracle.run_events(cfg, seed, nev)
    return n, time.perf_counter() - t


def cpu_baseline(seconds, cfg=CPU_PBPB, system="Pb+Pb"):
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    oracle.lib()  # builds the oracle if needed, before forking
    # the host cores this process may use, at most the 16 a one-GPU box
    # grants its command
    cores = max(1, min(16, len(os.sched_getaffinity(0))))
    n1, t1 = _cpu_worker((1, 8, cfg))
    per_core = max(4, int(seconds * n1 / t1))
    t = time.perf_counter()
    with mp.get_context("fork").Pool(cores) as pool:
        res = pool.map(_cpu_worker, [(1000 + i, per_core, cfg) for i in range(cores)])
    wall = time.perf_counter() - t
    tot = sum(r[0] for r in res)
    return dict(value=tot / wall, unit="events/s", cores=cores, kind="port",
                sample=f"{tot} {system} events ({per_core} per core) through the FP64 oracle event loop "
                       f"(sample -> collide -> T_A/T_B -> density -> eps_n, no file output), {wall:.1f} s wall")


def kln_params(nev):
    """MC-KLN Pb+Pb at sqrt(s) = 2.76 TeV, lambda = 0.138, T_A/T_B from
    Gaussian participants (w = 0.5 fm), dN/dy table (tmax 71, subdivision 3)
    with the 32 x 32 x 48 Gauss-Legendre quadrature, 261^2 cells."""
    from supermc_amd import default_params
    return default_params(
        which_mc_model=1, sub_model=7, ecm=2760.0, lambda_=0.138, tmax=71, tmax_subdivision=3,
        shape_of_nucleons=4, gauss_nucl_width=0.5, shape_of_entropy=2, collision_criterion=1, Aproj=208,
        Atarg=208, proj_ws_R=6.62, proj_ws_a=0.546, targ_ws_R=6.62, targ_ws_a=0.546, nucleon_min_dist=0.4,
        sigma_nn_mb=64.0, cc_fluctuation_model=0, bmin=0.0, bmax=20.0, Npmin=2, Npmax=500, maxx=13.0,
        maxy=13.0, dx=0.1, dy=0.1, finalFactor=1.0, ecc_from_order=2, ecc_to_order=5, kln_n_pt=32,
        kln_n_kt=32, kln_n_phi=48, max_events=nev, max_pairs=4096)


def time_kln(nev, steps, seed):
    from supermc_amd import Context
    with Context(kln_params(nev)) as ctx:
        ctx.run(seed, 0, nev, fetch=False)  # warm-up, builds the dN/dy table
        ctx.synchronize()
        ctx.set_timing(True)
        t0 = time.perf_counter()
        for k in range(steps):
            ctx.run(seed, (k + 1) * nev, nev, fetch=False)
        ctx.synchronize()
        dt = time.perf_counter() - t0
        ev = ctx.fetch_events(nev)
        # the default path smears T_A, T_B and interpolates the table in one
        # launch (k_tile with the KLN epilogue), timed as smearing
        ker = {name: ctx.timing(kid)[0] / max(ctx.timing(kid)[1], 1)
               for kid, name in ((0, "k_event"), (1, "k_tile(T_A,T_B -> dN/dy table)"), (2, "k_moments"))}
        fp = footprint_cells(ctx.derived.width, ctx.params.dx, -ctx.params.maxx)
        table = (ctx.kln_make_table(), ctx.derived.kln_dT)
    # useful FP32 work of the fused launch: 2 flops per (participant, cell of
    # its 5w disk) into T_A or T_B; the FP64 table step (about 25 flops per
    # cell of the rectangle) is not counted
    work = 2.0 * float(ev[:, 54].sum()) * fp["disk"]
    tf = work / (ker["k_tile(T_A,T_B -> dN/dy table)"] * 1e-3) / 1e12
    return dict(events_per_s=nev * steps / dt, events_per_step=nev, steps=steps, kernels_avg_ms=ker,
                failed_events=int((ev[:, 57] != 0).sum()), mean_eps2=float(np.hypot(ev[:, 2], ev[:, 12]).mean()),
                roofline_smear=dict(achieved_tflops=tf, frac=tf * 1e12 / FP32_PEAK,
                                    frac_nonpacked=tf * 1e12 / FP32_PEAK_NONPACKED,
                                    work="2 flops x Npart x %.1f disk cells" % fp["disk"]),
                _table=table)


def config1_params(nev):
    """Plain MC-Glauber Au+Au: WS R 6.38 fm, a 0.535 fm, hard core 0.4 fm,
    A 197, 42 mb, b on 0-20 fm, w 0.5 fm, alpha 0.14, no fluctuations,
    261^2 cells of 0.1 fm, eps_2..eps_5."""
    from supermc_amd import default_params
    return default_params(
        which_mc_model=5, sub_model=1, shape_of_nucleons=4, gauss_nucl_width=0.5, shape_of_entropy=2,
        collision_criterion=1, Aproj=197, Atarg=197, proj_ws_R=6.38, proj_ws_a=0.535, targ_ws_R=6.38,
        targ_ws_a=0.535, nucleon_min_dist=0.4, sigma_nn_mb=42.0, alpha=0.14, cc_fluctuation_model=0, bmin=0.0,
        bmax=20.0, Npmin=2, Npmax=500, maxx=13.0, maxy=13.0, dx=0.1, dy=0.1, finalFactor=1.0, ecc_from_order=2,
        ecc_to_order=5, max_events=nev)


class PinnedHost:
    """Page-locked host arrays from hipHostMalloc of the HIP runtime that
    libsmc is linked against (loaded by its soname, so the same instance);
    freed by close()."""

    def __init__(self):
        import ctypes
        self._c = ctypes
        self._hip = ctypes.CDLL("libamdhip64.so.7")
        self._hip.hipHostMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t, ctypes.c_uint]
        self._hip.hipHostFree.argtypes = [ctypes.c_void_p]
        self._ptrs = []

    def array(self, shape, dtype):
        c = self._c
        dt = np.dtype(dtype)
        n = int(np.prod(shape))
        p = c.c_void_p()
        rc = self._hip.hipHostMalloc(c.byref(p), max(n * dt.itemsize, 1), 0)
        if rc != 0 or not p.value:
            raise RuntimeError("hipHostMalloc of %d bytes failed (%d)" % (n * dt.itemsize, rc))
        self._ptrs.append(p.value)
        buf = (c.c_char * (n * dt.itemsize)).from_address(p.value)
        return np.frombuffer(buf, dtype=dt).reshape(shape)

    def close(self):
        for p in self._ptrs:
            self._hip.hipHostFree(p)
        self._ptrs = []


def time_config1(nev, steps, seed):
    """1000 Au+Au events from a fixed seed with the grids as output: each
    step samples, collides, smears and takes eps_n (records to host), then
    copies every event's entropy grid to host memory, FP32 (smc_density)
    or FP64 with the reference's double arithmetic (smc_density_f64)."""
    from supermc_amd import Context
    out = {}
    with Context(config1_params(nev)) as ctx:
        N = ctx.nx * ctx.ny
        # page-locked host buffers from the HIP runtime libsmc runs on, so
        # the copies run at the link's rate rather than through a staging copy
        pin = PinnedHost()
        g32 = pin.array((nev, ctx.nx, ctx.ny), np.float32)
        g64 = pin.array((nev, ctx.nx, ctx.ny), np.float64)
        ctx.run(seed, 0, nev)
        ctx.density(out=g32)
        ctx.density_f64(out=g64)
        ctx.synchronize()
        for name, fetch in (("fp32_grids", lambda: ctx.density(out=g32)), ("fp64_grids", lambda: ctx.density_f64(out=g64))):
            t0 = time.perf_counter()
            for k in range(steps):
                ctx.run(seed, k * nev, nev)
                fetch()
            ctx.synchronize()
            out[name + "_events_per_s"] = nev * steps / (time.perf_counter() - t0)
        # the same work with everything left in device memory, timed per
        # kernel with HIP events (the grids' HBM roofline is 4 or 8 B per cell)
        ctx.set_timing(True)
        for k in range(steps):
            ctx.run(seed, k * nev, nev, fetch=False)
            ctx.density_f64(out=g64)
        ctx.synchronize()
        ker = {name: ctx.timing(kid)[0] / max(ctx.timing(kid)[1], 1)
               for kid, name in ((0, "k_event"), (1, "k_tile"), (2, "k_moments"), (5, "k_grid64"))}
        ev = ctx.fetch_events(nev)
        sq = square_window_cells(ctx.derived.width, ctx.params.dx, -ctx.params.maxx)
        del g32, g64
        pin.close()
    terms = float(ev[:, 54].sum()) * sq["box"] + float(ev[:, 55].sum()) * sq["win5"]
    k64 = ker["k_grid64"]
    dev32 = nev / ((ker["k_event"] + ker["k_tile"] + ker["k_moments"]) * 1e-3)
    dev64 = nev / ((ker["k_event"] + ker["k_tile"] + ker["k_moments"] + k64) * 1e-3)
    out.update(events=nev, steps=steps, grid=[261, 261], mean_npart=float(ev[:, 54].mean()),
               mean_ncoll=float(ev[:, 55].mean()), kernels_avg_ms=ker,
               device_events_per_s_fp32=dev32, device_events_per_s_fp64=dev64,
               roofline_hbm=dict(bound="hbm", peak_gbps=HBM_PEAK / 1e9,
                                 fp32_grid_gbps=dev32 * N * 4 / 1e9, fp32_frac=dev32 * N * 4 / HBM_PEAK,
                                 fp64_grid_gbps=dev64 * N * 8 / 1e9, fp64_frac=dev64 * N * 8 / HBM_PEAK,
                                 work="4 (FP32) or 8 (FP64) bytes per cell of every event's grid"),
               k_grid64_cell_terms_per_s=terms / (k64 * 1e-3),
               k_grid64_write_gbps=nev * N * 8 / (k64 * 1e-3) / 1e9,
               grid_bytes_per_event={"fp32": N * 4, "fp64": N * 8},
               note="*_grids_events_per_s: wall time incl. the device-to-host copies of every grid (pinned); "
                    "device_events_per_s_*: kernel time with the grids left in HBM; k_grid64 evaluates exp per "
                    "(source, window cell) in FP64, FP64 VALU bound")
    return out


def time_kln_direct(nev, steps, seed):
    """MC-KLN Pb+Pb with dN/dy from the k_T-factorised quadrature in every
    cell (kln_direct = 1): 32 pT x 32 kT x 48 phi Gauss-Legendre nodes, phi
    and 2 pi - phi merged, so 32 x 768 node terms per cell with T_A, T_B > 0.
    The work unit is the integrand's node term c min(1, a u) min(1, b v)
    added to the sum, 5 flops as written (four multiplies and the add); the
    kernel evaluates it as c min(1/a, u) min(1/b, v) with a b taken out per
    pT node (two minima, a multiply and an FMA)."""
    from supermc_amd import Context
    p = kln_params(nev)
    p.kln_direct = 1
    with Context(p) as ctx:
        ctx.run(seed, 0, nev, fetch=False)
        ctx.synchronize()
        ctx.set_timing(True)
        t0 = time.perf_counter()
        for k in range(steps):
            ctx.run(seed, (k + 1) * nev, nev, fetch=False)
        ctx.synchronize()
        dt = time.perf_counter() - t0
        ms, n = ctx.timing(3)
        s, _, _ = ctx.density()
        active = float((s > 0).sum())
        nodes = p.kln_n_pt * p.kln_n_kt * (p.kln_n_phi // 2)
    kms = ms / max(n, 1)
    tf = 5.0 * active * nodes / (kms * 1e-3) / 1e12
    return dict(events_per_s=nev * steps / dt, events_per_step=nev, steps=steps, k_kln_direct_ms=kms,
                active_cells_per_event=active / nev, node_terms_per_cell=nodes,
                roofline=dict(bound="valu", achieved_tflops=tf, frac=tf * 1e12 / FP32_PEAK,
                              frac_nonpacked=tf * 1e12 / FP32_PEAK_NONPACKED,
                              work="5 flops per (cell with T_A, T_B > 0, node)"))


def _cpu_kln_worker(args):
    """The oracle's MC-KLN event loop on one core: b, both nuclei, the disk
    collisions, T_A and T_B, dN/dy from the table (sixPoint2dInterp per
    cell), eps_2..5; events with Ncoll = 0 or Npart outside [2, 500] are
    redrawn (MakeDensity.cpp:2143-2162)."""
    seed, nev, table, dT = args
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    rng = oracle.Rng(seed)
    g = oracle.grid(13.0, 0.1)
    w, dsq = 0.5, oracle.lib().orc_dsq(64.0)
    t = time.perf_counter()
    done = 0
    while done < nev:
        b = 20.0 * math.sqrt(rng.uniform())
        xa = rng.populate(208, 6.62, 0.546, 0.4, b / 2)
        xb = rng.populate(208, 6.62, 0.546, 0.4, -b / 2)
        pa, pb = oracle.sort_nucleus(xa, w), oracle.sort_nucleus(xb, w)
        col = oracle.collide(xa[pa], xb[pb], w, dsq)
        npart = col["npart_a"] + col["npart_b"]
        if col["ncoll"] == 0 or not 2 <= npart <= 500:
            continue
        ia, ib = col["order_a"], col["order_b"]
        ta, tb = oracle.thickness(g, w, xa[pa][ia]), oracle.thickness(g, w, xb[pb][ib])
        s, _ = oracle.density_kln_table(g, ta, tb, table, dT)
        boxes = np.array([oracle.particle_box(x, y, w) for x, y in np.concatenate([xa[pa][ia], xb[pb][ib]])])
        oracle.moments(g, s, boxes, col["ncoll"], 2, 5)
        done += 1
    return done, time.perf_counter() - t


def cpu_kln_baseline(seconds, table, dT, seed):
    """MC-KLN Pb+Pb on the host cores through the oracle (one process per
    core), per-event work only; the one-time dN/dy table (211 x 211 points
    of the 32 x 32 x 48 quadrature) is priced from a timed sample of points."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    oracle.lib()
    cores = max(1, min(16, len(os.sched_getaffinity(0))))
    n1, t1 = _cpu_kln_worker((seed, 2, table, dT))
    per_core = max(2, int(seconds * n1 / t1))
    t = time.perf_counter()
    with mp.get_context("fork").Pool(cores) as pool:
        res = pool.map(_cpu_kln_worker, [(seed + 1 + i, per_core, table, dT) for i in range(cores)])
    wall = time.perf_counter() - t
    tot = sum(r[0] for r in res)
    k = oracle.kln(2760.0, 0.138)
    t = time.perf_counter()
    for i in range(3):
        oracle.kln_dndy_gl(k, 0.5 + i, 1.0 + 0.5 * i, 32, 32, 48)
    per_point = (time.perf_counter() - t) / 3
    return dict(value=tot / wall, unit="events/s", cores=cores, kind="port",
                sample=f"{tot} MC-KLN Pb+Pb events ({per_core} per core) through the FP64 oracle (sample -> "
                       f"collide -> T_A/T_B -> table dN/dy -> eps_n), {wall:.1f} s wall",
                table_build_s_on_these_cores=per_point * 211 * 211 / cores)


def p3d_lists(nev, seed):
    """Operation-2 participant (x, y, nucleus id) and binary-collision (x, y)
    lists of nev minimum-bias Pb+Pb events of the main workload."""
    from supermc_amd import Context
    parts, bins = [], []
    with Context(bench_params(nev)) as ctx:
        ctx.run(seed, 0, nev, fetch=False)
        for e in range(nev):
            nuc = ctx.nucleons(e)
            pa, pb, _ = ctx.pairs(e)
            wa, wb = nuc["wound_proj"] >= 0, nuc["wound_targ"] >= 0
            parts.append(np.concatenate([np.column_stack([nuc["xy_proj"][wa], np.ones(wa.sum())]),
                                         np.column_stack([nuc["xy_targ"][wb], np.full(wb.sum(), 2.0)])]))
            bins.append((nuc["xy_proj"][pa] + nuc["xy_targ"][pb]) / 2)
    return parts, bins


def time_p3d(nev, steps, seed):
    """3D (x, y, eta) profiles of scripts/generate_3d_profiles (261 x 261 x 101,
    ecm 19.6 GeV, random_flag 1) of nev events per call: eta sampling and both
    grids on the device, the grids returned to the host (2 x 27.5 MB per
    event, included in the time)."""
    from supermc_amd import Profile3D
    parts, bins = p3d_lists(nev, seed)
    with Profile3D(random_flag=1) as P:
        P.generate(parts, bins, seed, 0)  # warm-up
        t0 = time.perf_counter()
        for k in range(steps):
            P.generate(parts, bins, seed, (k + 1) * nev)
        dt = time.perf_counter() - t0
    return dict(events_per_s=nev * steps / dt, events_per_step=nev, steps=steps, grid=[101, 261, 261],
                mean_npart=float(np.mean([len(p) for p in parts])), mean_ncoll=float(np.mean([len(b) for b in bins])),
                note="wall time of smc_p3d_generate incl. the device-to-host copy of both grids")


def cpu_p3d_baseline(seed, nev=4):
    """The oracle's FP64 restatement of profile_3d::generate_3d_profile +
    generate_3d_profile_binary on the sampled sources of the first nev events
    of the GPU leg, one core."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    from supermc_amd import Profile3D
    parts, bins = p3d_lists(nev, seed)
    with Profile3D(random_flag=1) as P:
        ps, bs = P.sample(parts, bins, seed, 0)
    t = time.perf_counter()
    for e in range(nev):
        oracle.p3d_smear(261, 261, 101, 0.1, 0.1, 0.1, ps[e])
        oracle.p3d_smear(261, 261, 101, 0.1, 0.1, 0.1, bs[e])
    return nev / (time.perf_counter() - t)


MIXED = [  # (name, Aproj, Atarg, nucleus keys): the mixed-system workload
    ("p+Pb", 1, 208, dict(targ_ws_R=6.62, targ_ws_a=0.546)),
    ("d+Au", 2, 197, dict(targ_ws_R=6.38, targ_ws_a=0.535)),
    ("Cu+Cu", 63, 63, dict(proj_ws_R=4.2, proj_ws_a=0.596, targ_ws_R=4.2, targ_ws_a=0.596)),
    ("Au+Au", 197, 197, dict(proj_ws_R=6.38, proj_ws_a=0.535, targ_ws_R=6.38, targ_ws_a=0.535)),
    ("U+U", 238, 238, dict(proj_ws_R=6.81, proj_ws_a=0.6, targ_ws_R=6.81, targ_ws_a=0.6, proj_deformed=1,
                           targ_deformed=1, proj_beta2=0.28, proj_beta4=0.093, targ_beta2=0.28, targ_beta4=0.093)),
]


def mixed_params(nev, A, B, nuclei):
    from supermc_amd import default_params
    return default_params(
        which_mc_model=5, sub_model=1, shape_of_nucleons=4, gauss_nucl_width=0.5, shape_of_entropy=2,
        collision_criterion=1, Aproj=A, Atarg=B, nucleon_min_dist=0.4, sigma_nn_mb=42.0, alpha=0.14,
        cc_fluctuation_model=6, cc_fluctuation_Gamma_theta=1.25, bmin=0.0, bmax=20.0, Npmin=2, Npmax=500,
        maxx=13.0, maxy=13.0, dx=0.05, dy=0.05, finalFactor=1.0, ecc_from_order=2, ecc_to_order=5, max_events=nev,
        max_pairs=8192, **nuclei)


def smear_roofline(ev, fp, smear_ms, nev, smear_path=0):
    """Useful FP32 work of the exact-footprint smearing (2 flops per source
    and footprint cell, fp = footprint_cells) over a batch's average
    smearing time, against MI355X's FP32 figure and the plain-FMA rate
    (VALU tiles), or against the matrix-core peak of the MFMA precision."""
    work = 2.0 * (float(ev[:, 54].sum()) * fp["box"] + float(ev[:, 55].sum()) * fp["disk"]) / len(ev) * nev
    tf = work / (smear_ms * 1e-3) / 1e12
    if smear_path in (1, 3, 4):
        return dict(bound="mfma", achieved_tflops=tf, peak_tflops=mfma_peak(smear_path) / 1e12,
                    frac=tf * 1e12 / mfma_peak(smear_path))
    return dict(bound="valu", achieved_tflops=tf, frac=tf * 1e12 / FP32_PEAK,
                frac_nonpacked=tf * 1e12 / FP32_PEAK_NONPACKED)


def time_mixed(nev, total, seed, smear_path):
    """The large mixed-system workload: p+Pb, d+Au, Cu+Cu, Au+Au and deformed
    U+U cycled event by event, 521^2 cells at 0.05 fm, `total` events in one
    smc_run_stream call through a context of `nev` events (device memory of
    one batch), records streamed to host memory."""
    from supermc_amd import Context
    systems = [mixed_params(nev, A, B, k) for _, A, B, k in MIXED]
    for p in systems:
        p.smear_path = smear_path
    p0 = mixed_params(nev, 238, 238, MIXED[-1][3])
    p0.smear_path = smear_path
    with Context(p0) as ctx:
        ctx.set_systems(systems, list(range(len(systems))))
        ctx.run(seed, 0, nev, fetch=False)  # warm-up
        ctx.synchronize()
        ctx.set_timing(True)
        rec = np.empty((total, 64))
        t0 = time.perf_counter()
        ctx.run_stream(seed, nev, total, out=rec)
        dt = time.perf_counter() - t0
        ker = {name: ctx.timing(kid)[0] / max(ctx.timing(kid)[1], 1)
               for kid, name in ((0, "k_event"), (1, "k_smear"), (2, "k_moments"))}
        fp = footprint_cells(ctx.derived.width, ctx.params.dx, -ctx.params.maxx)
    npart = {MIXED[s][0]: float(rec[rec[:, 58] == s, 54].mean()) for s in range(len(MIXED))}
    return dict(events_per_s=total / dt, events=total, events_per_batch=nev, grid=[521, 521],
                smear_path=SMEAR_PATHS[smear_path], systems=[m[0] for m in MIXED], mean_npart=npart,
                failed_events=int((rec[:, 57] != 0).sum()), kernels_avg_ms_per_batch=ker,
                roofline_smear=smear_roofline(rec[:nev], fp, ker["k_smear"], nev, smear_path))


def time_systems(nev, seed):
    """k_event per batch of each mixed-workload system alone (521^2), with
    one and with two attempts per sampling pass (the events are the same)."""
    from supermc_amd import Context
    out = {}
    for name, A, B, nuc in MIXED:
        with Context(mixed_params(nev, A, B, nuc)) as ctx:
            ctx.run(seed, 0, nev, fetch=False)
            _, ms = ctx.autotune_attempts(reps=3)
        out[name] = {"k_event_ms_1_attempt_per_pass": float(ms[0]), "k_event_ms_2_attempts_per_pass": float(ms[1]),
                     "us_per_event": float(min(ms)) * 1e3 / nev}
    return out


def time_avg(nev, steps, seed):
    """Operation 3 (MakeDensity::generate_profile_average, pT-integrated) on
    the main Pb+Pb workload: per batch the events (smc_run), then
    smc_average_profiles for orders 2..5 with the reaction-plane sums --
    per order and event two recentrings (one also turned by the order-n
    angle) and two re-smearings, summed over the batch into FP64 grids on
    the device; the 8 summed grids come back to the host."""
    from supermc_amd import Context
    with Context(bench_params(nev)) as ctx:
        ctx.run(seed, 0, nev, fetch=False)
        ctx.average_profiles(2, 5, reaction_plane=True)  # warm-up (allocates the sums)
        ctx.synchronize()
        ctx.set_timing(True)
        t0 = time.perf_counter()
        for k in range(steps):
            ctx.run(seed, (k + 1) * nev, nev, fetch=False)
            sums, _ = ctx.average_profiles(2, 5, reaction_plane=True)
        ctx.synchronize()
        dt = time.perf_counter() - t0
        ms, n = ctx.timing(1)
        ctx.set_timing(False)
        # the sources are left turned: the last batch again (same counters,
        # same events) for its acceptance
        ev = ctx.run(seed, steps * nev, nev)
    acc = int((ev[:, 57] == 0).sum())
    return dict(events_per_s=nev * steps / dt, events_per_step=nev, steps=steps, orders=[2, 5], reaction_plane=True,
                smearings_per_event=1 + 2 * 4, k_smear_avg_ms=ms / max(n, 1), k_smear_launches=n,
                accepted_last_batch=acc, mean_dsdy_order2=float(sums[0].sum() * 0.01 / max(acc, 1)))


def time_jet(nev, steps, seed):
    """Operation 2 grids (MakeDensity::generate_profile_ebe_Jet) on the main
    Pb+Pb workload: per batch the events (smc_run), the binary-collision
    density (smc_rho_binary) and the projectile and target spectator
    densities (smc_spectator_density), all three grids left in device
    memory."""
    from supermc_amd import Context
    with Context(bench_params(nev)) as ctx:
        ctx.run(seed, 0, nev, fetch=False)
        ctx.rho_binary(host=False)
        ctx.spectator_density(host=False)
        ctx.synchronize()
        ctx.set_timing(True)
        t0 = time.perf_counter()
        for k in range(steps):
            ctx.run(seed, (k + 1) * nev, nev, fetch=False)
            ctx.rho_binary(host=False)
            ctx.spectator_density(host=False)
        ctx.synchronize()
        dt = time.perf_counter() - t0
        ms, n = ctx.timing(1)
    return dict(events_per_s=nev * steps / dt, events_per_step=nev, steps=steps,
                grids_per_event=["density", "rho_binary", "spectators A", "spectators B"],
                k_smear_avg_ms=ms / max(n, 1), k_smear_launches=n)


def cpu_ops_baseline(seed, nev=3):
    """The oracle's FP64 restatements of one operation-3 event (orders 2..5,
    with the reaction-plane profiles) and one event's operation-2 grids
    (collision term of the pairs, spectator thicknesses), on the first nev
    events of a device batch, one core.  Returns (op3, op2) events/s."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    import parity_helpers as H
    from supermc_amd import Context
    p = bench_params(nev)
    g = oracle.grid(p.maxx, p.dx, p.maxy, p.dy)
    inputs = []
    with Context(p) as ctx:
        ctx.run(seed, 0, nev, fetch=False)
        w = ctx.derived.width
        for e in range(nev):
            o = H.oracle_event(ctx, e)
            nuc = o["nuc"]
            pa, pb, pf = o["gpu_pairs"]
            xa, xb, ia, ib = nuc["xy_proj"], nuc["xy_targ"], o["ia"], o["ib"]
            inputs.append((xa[ia], nuc["fluct_proj"][ia].astype(float), xb[ib], nuc["fluct_targ"][ib].astype(float),
                           (xa[pa] + xb[pb]) / 2.0, pf.astype(float), xa[nuc["wound_proj"] == -1],
                           xb[nuc["wound_targ"] == -1]))
    t = time.perf_counter()
    for a, fa, b, fb, mid, fm, _, _ in inputs:
        oracle.rotated_profiles(g, w, p.alpha, a, fa, b, fb, mid, fm, 2, 5, p.finalFactor, reaction_plane=True)
    op3 = nev / (time.perf_counter() - t)
    no = np.zeros((0, 2))
    t = time.perf_counter()
    for a, fa, b, fb, mid, fm, sa, sb in inputs:
        oracle.density_glauber(g, w, p.alpha, a, fa, b, fb, mid, fm)
        oracle.density_glauber(g, w, 1.0, no, np.zeros(0), no, np.zeros(0), mid, np.ones(len(mid)))
        oracle.thickness(g, w, sa)
        oracle.thickness(g, w, sb)
    op2 = nev / (time.perf_counter() - t)
    return op3, op2


def time_defaults(nev, steps, seed):
    """The reference parameter file's defaults (smc_params_default):
    sqrt(T_A T_B) (which_mc_model 7), shape_of_nucleons 2, the Gaussian
    collision criterion, Gamma theta 0.75, eps_1..eps_9, Pb+Pb at 5020 GeV on
    301^2 cells, tile shape autotuned."""
    from supermc_amd import Context, default_params
    with Context(default_params(max_events=nev)) as ctx:
        ctx.run(seed, 0, nev, fetch=False)
        ctx.synchronize()
        ctx.autotune(reps=1)
        ctx.set_timing(True)
        t0 = time.perf_counter()
        for k in range(steps):
            ctx.run(seed, (k + 1) * nev, nev, fetch=False)
        ctx.synchronize()
        dt = time.perf_counter() - t0
        ker = {name: ctx.timing(kid)[0] / max(ctx.timing(kid)[1], 1)
               for kid, name in ((0, "k_event"), (1, "k_smear"), (2, "k_moments"))}
        ev = ctx.fetch_events(nev)
    return dict(events_per_s=nev * steps / dt, events_per_step=nev, steps=steps, grid=[301, 301],
                collision_criterion=2, orders=[1, 9], mean_npart=float(ev[:, 54].mean()),
                failed_events=int((ev[:, 57] != 0).sum()), kernels_avg_ms=ker)


def time_graph(seed):
    """Batches launched eagerly (smc_run) and replayed from a captured
    hipGraph (smc_run_graph), at 10^3 and 10^5 events per batch."""
    from supermc_amd import Context
    out = {}
    for nev, steps in ((1000, 50), (100000, 3)):
        with Context(bench_params(nev)) as ctx:
            for mode, fn in (("eager", ctx.run), ("graph", ctx.run_graph)):
                fn(seed, 0, nev, fetch=False)  # for the graph: the eager call, then the capture
                fn(seed, nev, nev, fetch=False)
                ctx.synchronize()
                t0 = time.perf_counter()
                for k in range(steps):
                    fn(seed, (k + 2) * nev, nev, fetch=False)
                ctx.synchronize()
                out[f"{mode}_events_per_s_{nev}"] = nev * steps / (time.perf_counter() - t0)
    return out


def time_adjoint(nev, steps, seed):
    """Differentiable Pb+Pb batch: forward (|eps_2|, |eps_3|, dS/dy) plus
    the adjoint d/d(w, R, a, alpha, k), per event.  The roofline figures
    count, per (source, cell of its integer window -- no disk cut on this
    path): 2 flops in the forward smearing and 10 in the backward gather
    (the Gaussian product, the cotangent product and the four weighted
    sums), against the FP32 figures."""
    from supermc_amd import Context
    with Context(bench_params(nev)) as ctx:
        th0 = ctx.diff_setup(seed, 0, nev)
        cot = np.tile([1.0 / nev, 1.0 / nev, 0.0], (nev, 1))
        ctx.diff_forward(th0)
        ctx.diff_backward(cot)
        ctx.synchronize()
        ctx.set_timing(True)
        t0 = time.perf_counter()
        for _ in range(steps):
            ctx.diff_forward(th0)
            g = ctx.diff_backward(cot)
        ctx.synchronize()
        dt = time.perf_counter() - t0
        ker = {name: ctx.timing(kid)[0] / max(ctx.timing(kid)[1], 1)
               for kid, name in ((0, "k_diff_pos+src"), (1, "k_tile (no cut)"), (2, "k_diff_moments"),
                                 (4, "k_diff_cot+gather+chain"))}
        npart = ncoll = 0
        for e in range(0, nev, max(1, nev // 256)):
            nuc = ctx.nucleons(e)
            npart += int((nuc["wound_proj"] >= 0).sum() + (nuc["wound_targ"] >= 0).sum())
            ncoll += len(ctx.pairs(e)[0])
        frac_ev = len(range(0, nev, max(1, nev // 256))) / nev
        sq = square_window_cells(ctx.derived.width, ctx.params.dx, -ctx.params.maxx)
        cells = (npart * sq["box"] + ncoll * sq["win5"]) / frac_ev
    fw = 2.0 * cells / (ker["k_tile (no cut)"] * 1e-3) / 1e12
    bw = 10.0 * cells / (ker["k_diff_cot+gather+chain"] * 1e-3) / 1e12
    return dict(events_per_s=nev * steps / dt, events_per_step=nev, steps=steps, kernels_avg_ms=ker,
                grad_eps2_plus_eps3=dict(zip(("w", "R", "a", "alpha", "k"), map(float, g))),
                roofline=dict(forward_tflops=fw, forward_frac=fw * 1e12 / FP32_PEAK,
                              backward_tflops=bw, backward_frac=bw * 1e12 / FP32_PEAK,
                              backward_frac_nonpacked=bw * 1e12 / FP32_PEAK_NONPACKED))


def footprint_cells(w, dx, xmin, seed=1, n=4000):
    """Mean number of cells a source updates under the reference's footprint
    rules (MCnucl.cpp:450-475, :732-756, Particle.cpp:122-132): the +-5w
    window with dc <= 25 w^2 (binary collisions, T_A/T_B), and the +-4w
    participant box with r <= 5w.  Source positions uniform over cells."""
    import numpy as np
    rng = np.random.default_rng(seed)
    out = {}
    for kind, h in (("disk", 5 * w), ("box", 4 * w)):
        tot = 0
        for x, y in rng.uniform(-3, 3, (n, 2)):
            xs = xmin + np.arange(int((x - h - xmin) / dx), int((x + h - xmin) / dx)) * dx
            ys = xmin + np.arange(int((y - h - xmin) / dx), int((y + h - xmin) / dx)) * dx
            tot += int((((x - xs[:, None]) ** 2 + (y - ys[None, :]) ** 2) <= 25 * w * w).sum())
        out[kind] = tot / n
    return out


def square_window_cells(w, dx, xmin, seed=1, n=4000):
    """Mean cells of the integer windows without a disk cut: the +-4w
    participant box and the +-5w window (the differentiable path)."""
    rng = np.random.default_rng(seed)
    out = {}
    for kind, h in (("win5", 5 * w), ("box", 4 * w)):
        tot = 0
        for x, y in rng.uniform(-3, 3, (n, 2)):
            nxw = int((x + h - xmin) / dx) - int((x - h - xmin) / dx)
            nyw = int((y + h - xmin) / dx) - int((y - h - xmin) / dx)
            tot += nxw * nyw
        out[kind] = tot / n
    return out


def copy_probe(nbytes=1 << 28):
    """One device-to-device copy of nbytes (a known-size kernel for scaling
    rocprofv3 FETCH_SIZE / WRITE_SIZE readings)."""
    import torch
    a = torch.ones(nbytes // 4, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    b.copy_(a)
    torch.cuda.synchronize()
    del a, b


def eps_error_vs_fp64(ctx, ev, n):
    """Largest |eps_n - eps_n(FP64 oracle)| (n = 2..5, r^2 and r^n weights)
    over n events spread evenly through the last batch, the oracle fed the
    same nucleons and Gamma factors."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import parity_helpers as H
    worst = {"eps_r2": 0.0, "eps_rn": 0.0}
    picks = np.unique(np.linspace(0, ctx.nev - 1, min(n, ctx.nev)).astype(int))
    for e in picks:
        o = H.oracle_event(ctx, e)
        m = o["mom"]
        for k in range(2, 6):
            worst["eps_r2"] = max(worst["eps_r2"], abs(ev[e, k] - m["mom_re"][k]), abs(ev[e, 10 + k] - m["mom_im"][k]))
            worst["eps_rn"] = max(worst["eps_rn"], abs(ev[e, 20 + k] - m["momp_re"][k]),
                                  abs(ev[e, 30 + k] - m["momp_im"][k]))
    worst["events"] = len(picks)
    worst["spread"] = "events evenly spaced over the batch"
    return worst


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--events", type=int, default=100000)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--seed", type=int, default=20261202)
    ap.add_argument("--smear-path", type=int, default=0, choices=range(5),
                    help="smearing path, all with the exact footprint: 0 separable VALU tiles, 1 FP32 MFMA "
                         "Gx.diag(w).Gy^T + exact-cut pass, 2 per-cell exponentials, 3 BF16x3 MFMA, 4 BF16x6 MFMA")
    ap.add_argument("--parity-events", type=int, default=100,
                    help="events of the last batch re-evaluated by the FP64 oracle for the eps_n error")
    ap.add_argument("--kln-events", type=int, default=20000, help="MC-KLN events per step (0: skip)")
    ap.add_argument("--kln-direct-events", type=int, default=1000,
                    help="MC-KLN events per step with the per-cell quadrature (0: skip)")
    ap.add_argument("--config1-events", type=int, default=1000,
                    help="Au+Au events with FP32 and FP64 grids as output (0: skip)")
    ap.add_argument("--adjoint-events", type=int, default=4096, help="differentiable batch size (0: skip)")
    ap.add_argument("--p3d-events", type=int, default=16, help="3D profile events per call (0: skip)")
    ap.add_argument("--no-graph", action="store_true", help="skip the eager vs hipGraph-replay comparison")
    ap.add_argument("--no-autotune", action="store_true", help="keep the default tile shape")
    ap.add_argument("--mixed-total", type=int, default=1000000,
                    help="events of the mixed-system workload in one smc_run_stream call")
    ap.add_argument("--mixed-events", type=int, default=50000,
                    help="mixed-system 521^2 events per batch (0: skip)")
    ap.add_argument("--avg-events", type=int, default=20000,
                    help="operation-3 averaged-profile events per batch (0: skip)")
    ap.add_argument("--jet-events", type=int, default=20000, help="operation-2 grid events per batch (0: skip)")
    ap.add_argument("--default-events", type=int, default=100000,
                    help="events per batch with the reference parameter file's defaults (0: skip)")
    ap.add_argument("--only", choices=("kln", "kln-direct", "config1", "adjoint", "p3d", "mixed", "systems", "avg",
                                       "jet", "defaults", "fma"), default=None,
                    help="time only the MC-KLN or the adjoint workload (its own JSON line), e.g. for profiling")
    ap.add_argument("--copy-probe", action="store_true",
                    help="run one 256 MiB device copy first (calibration kernel for HBM counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: measures one GPU; --gpus {args.gpus} is not supported "
              "(across GPUs the events would shard by Philox counter range, with no exchange)", file=sys.stderr)
        sys.exit(2)

    from supermc_amd import Context, EV_NPART, EV_NCOLL, EV_STATUS, load
    load()  # fails loudly when libsmc.so is missing
    if args.copy_probe:
        copy_probe()
    if args.only:
        if args.only == "kln":
            r = time_kln(args.kln_events, args.steps, args.seed)
            r.pop("_table")
            metric, wl = "events/s MC-KLN Pb+Pb 2.76 TeV 261^2 grid", "MC-KLN Pb+Pb, dN/dy table, 261x261"
        elif args.only == "kln-direct":
            r = time_kln_direct(args.kln_direct_events, args.steps, args.seed)
            metric, wl = "events/s MC-KLN Pb+Pb 2.76 TeV 261^2, per-cell quadrature", "MC-KLN Pb+Pb, kln_direct"
        elif args.only == "config1":
            r = time_config1(args.config1_events, args.steps, args.seed)
            r["events_per_s"] = r["fp64_grids_events_per_s"]
            metric, wl = "events/s Au+Au 261^2 with FP64 grids to host", "plain MC-Glauber Au+Au, 1000 events"
        elif args.only == "fma":
            from supermc_amd._native import calibrate_fma
            tf, ms = calibrate_fma()
            r = dict(events_per_s=None, fma_chain_tflops=tf, ms=ms, peak_nonpacked_tflops=FP32_PEAK_NONPACKED / 1e12)
            metric, wl = "plain FP32 FMA TFLOP/s (independent v_fma_f32 chains)", "4096 x 256 lanes x 256000 FMAs"
        elif args.only == "systems":
            r = time_systems(20000, args.seed)
            r["events_per_s"] = None
            metric, wl = "k_event ms per 2x10^4 events, each mixed-workload system alone", "521^2, 42 mb"
        elif args.only == "adjoint":
            r = time_adjoint(args.adjoint_events, args.steps, args.seed)
            metric, wl = "events/s differentiable Pb+Pb 261^2 (forward + adjoint)", "frozen-topology Pb+Pb batch"
        elif args.only == "p3d":
            r = time_p3d(args.p3d_events, args.steps, args.seed)
            metric, wl = "events/s 3D profiles 101x261x261", "eta-sampled 3D Gaussians of Pb+Pb operation-2 lists"
        elif args.only == "avg":
            r = time_avg(args.avg_events, args.steps, args.seed)
            metric, wl = "events/s averaged rotated profiles Pb+Pb 261^2", "operation 3, orders 2..5, reaction plane"
        elif args.only == "defaults":
            r = time_defaults(args.default_events, args.steps, args.seed)
            metric, wl = "events/s reference defaults Pb+Pb 5.02 TeV 301^2", "parameters.dat defaults"
        elif args.only == "jet":
            r = time_jet(args.jet_events, args.steps, args.seed)
            metric, wl = "events/s operation-2 grids Pb+Pb 261^2", "rho_binary and spectator densities"
        else:
            r = time_mixed(args.mixed_events, args.mixed_total, args.seed, args.smear_path)
            metric, wl = "events/s mixed systems 521^2", "p+Pb, d+Au, Cu+Cu, Au+Au, U+U in one stream"
        print(json.dumps({"metric": metric, "value": r["events_per_s"], "unit": "events/s", "n_gpus": 1,
                          "steps": args.steps, "warmup": 1, "higher_is_better": True, "dtype": "f32",
                          "data": "synthetic", "config": dict(workload=wl, **r)}), flush=True)
        return
    nev = args.events
    ctx = Context(bench_params(nev, args.smear_path))
    print(f"[bench] Pb+Pb 261^2, {nev} events/step, warmup {args.warmup}, steps {args.steps}", file=sys.stderr,
          flush=True)
    for w in range(args.warmup):
        ctx.run(args.seed, w * nev, nev, fetch=False)
    ctx.synchronize()
    tune = None
    if args.smear_path == 0 and not args.no_autotune:
        # tile shape chosen on the warm-up batch (smc_autotune; grids do not depend on it)
        best, ms = ctx.autotune(reps=2)
        tune = dict(variant=best, ms_per_variant=[round(float(m), 3) for m in ms],
                    variants=["32x64 chunk 8", "16x64 chunk 8", "32x64 chunk 4", "64x64 chunk 8", "64x32 rows interleaved"])
    ctx.set_timing(True)
    t0 = time.perf_counter()
    for k in range(args.steps):
        ctx.run(args.seed, (args.warmup + k) * nev, nev, fetch=False)
    ctx.synchronize()
    t1 = time.perf_counter()
    elapsed = t1 - t0
    value = nev * args.steps / elapsed
    print(f"[bench] timed region {elapsed:.3f} s -> {value:.1f} events/s", file=sys.stderr, flush=True)

    ev = ctx.fetch_events(nev)  # the last timed batch
    bad = int((ev[:, EV_STATUS] != 0).sum())
    sum_k = float((ev[:, EV_NPART] + ev[:, EV_NCOLL]).sum())
    N = ctx.nx * ctx.ny
    kernels = {}
    for kid, name in ((0, "k_event"), (1, "k_smear"), (2, "k_moments")):
        ms, n = ctx.timing(kid)
        if n:
            kernels[name] = dict(avg_ms=ms / n, launches=n)
    dom = max(kernels, key=lambda k: kernels[k]["avg_ms"])
    if dom == "k_smear":
        # useful work: one multiply-add per (source, footprint cell), summed
        # over the batch with the measured Npart and Ncoll
        fp = footprint_cells(ctx.derived.width, ctx.params.dx, -ctx.params.maxx)
        npart = float(ev[:, EV_NPART].sum())
        ncoll = float(ev[:, EV_NCOLL].sum())
        work = 2.0 * (npart * fp["box"] + ncoll * fp["disk"])
        sec = kernels[dom]["avg_ms"] * 1e-3
        achieved = work / sec
        # the VALU tiles (and the MFMA contraction's exact-cut pass) issue FP32
        # VALU; frac is against MI355X's FP32 figure, frac_nonpacked against
        # the plain-FMA rate the build can reach
        if args.smear_path in (0, 2):
            roof = dict(bound="valu", achieved=achieved / 1e12, peak=FP32_PEAK / 1e12, unit="TFLOP/s",
                        frac=achieved / FP32_PEAK, peak_nonpacked=FP32_PEAK_NONPACKED / 1e12,
                        frac_nonpacked=achieved / FP32_PEAK_NONPACKED)
        else:
            pk = mfma_peak(args.smear_path)
            roof = dict(bound="mfma", achieved=achieved / 1e12, peak=pk / 1e12, unit="TFLOP/s", frac=achieved / pk)
        roof.update(work="2 flops x (Npart x %.1f + Ncoll x %.1f) footprint cells per event" % (fp["box"], fp["disk"]),
                    dense_equivalent_tflops=2.0 * N * sum_k / sec / 1e12)
    else:
        work = 4.0 * N * nev  # one FP32 grid per event
        achieved = work / (kernels[dom]["avg_ms"] * 1e-3)
        roof = dict(bound="hbm", achieved=achieved / 1e9, peak=HBM_PEAK / 1e9, unit="GB/s", frac=achieved / HBM_PEAK)
    roof["kernel"] = dom
    try:
        from supermc_amd._native import calibrate_fma
        roof["fma_chain_tflops_measured"] = calibrate_fma()[0]
    except Exception:  # reported as missing, not guessed
        roof["fma_chain_tflops_measured"] = None
    # HBM bytes per launch from the committed rocprofv3 FETCH_SIZE /
    # WRITE_SIZE summary (calibrated against a plain copy), scaled to this
    # batch size; null for a smearing path it was not measured on
    traffic = None
    tpath = os.path.join(ROOT, "profiles", f"{dom}_pbpb261_traffic.json")
    if os.path.exists(tpath) and args.smear_path == 0:
        try:
            t = json.load(open(tpath))
            traffic = t["bytes_per_launch"] * nev / t["events_per_launch"]
        except (OSError, ValueError, KeyError):
            traffic = None
    roof["traffic"] = traffic

    eps_err = eps_error_vs_fp64(ctx, ev, args.parity_events) if args.parity_events > 0 else None
    # further batches with the event records copied to host memory (reported
    # beside `value`, which leaves them in device memory)
    ctx.set_timing(False)
    th = time.perf_counter()
    for k in range(2):
        ctx.run(args.seed, (args.warmup + args.steps + k) * nev, nev, fetch=True)
    host_rate = 2 * nev / (time.perf_counter() - th)

    kln = time_kln(args.kln_events, 2, args.seed) if args.kln_events > 0 else None
    kln_table = kln.pop("_table") if kln else None
    kdir = time_kln_direct(args.kln_direct_events, 2, args.seed) if args.kln_direct_events > 0 else None
    if kln is not None:
        kln["per_cell_quadrature"] = kdir
    cfg1 = time_config1(args.config1_events, 3, args.seed) if args.config1_events > 0 else None
    adj = time_adjoint(args.adjoint_events, 3, args.seed) if args.adjoint_events > 0 else None
    p3d = time_p3d(args.p3d_events, 2, args.seed) if args.p3d_events > 0 else None
    graph = time_graph(args.seed) if not args.no_graph else None
    mixed = None
    if args.mixed_events > 0:
        # BF16x3 MFMA, the precision the large-batch configuration names,
        # beside the exact FP32 scatter path (VALU tiles) on the same events
        mixed = dict(bf16x3_mfma=time_mixed(args.mixed_events, args.mixed_total, args.seed, 3),
                     fp32_valu_tiles=time_mixed(args.mixed_events, args.mixed_total, args.seed, 0))
    avg = time_avg(args.avg_events, 2, args.seed) if args.avg_events > 0 else None
    jet = time_jet(args.jet_events, 2, args.seed) if args.jet_events > 0 else None
    dflt = time_defaults(args.default_events, 2, args.seed) if args.default_events > 0 else None

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(args.cpu_seconds)
        if kln is not None:
            kln["cpu_baseline"] = cpu_kln_baseline(args.cpu_seconds, kln_table[0], kln_table[1], args.seed)
        if cfg1 is not None:
            cfg1["cpu_baseline"] = cpu_baseline(min(args.cpu_seconds, 10.0), CPU_AUAU, "Au+Au")
        if p3d is not None:
            p3d["cpu_oracle_events_per_s_1core"] = cpu_p3d_baseline(args.seed)
        if avg is not None or jet is not None:
            op3, op2 = cpu_ops_baseline(args.seed)
            if avg is not None:
                avg["cpu_oracle_events_per_s_1core"] = op3
            if jet is not None:
                jet["cpu_oracle_events_per_s_1core"] = op2

    line = {
        "metric": "events/s Pb+Pb 261^2 grid (Glauber, KLN); fwd+adjoint; eps_n err vs FP64",
        "value": value,
        "unit": "events/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (Philox-sampled nuclei on the device)",
        "config": {"workload": "batched MC-Glauber Pb+Pb min-bias, 261x261 grid, Gamma fluct k=0.8",
                   "events_per_step": nev, "grid": [ctx.nx, ctx.ny], "sigma_nn_mb": 64.0, "moments_dtype": "f64",
                   "mean_sources_per_event": sum_k / nev, "failed_events": bad,
                   "smear_path": SMEAR_PATHS[args.smear_path],
                   "eps_err_vs_fp64": eps_err,
                   "kernels_avg_ms": {k: v["avg_ms"] for k, v in kernels.items()},
                   "mc_kln_pbpb_2760": kln, "auau_1000_grids_fp32_fp64": cfg1, "adjoint_fwd_bwd_pbpb": adj,
                   "profile_3d_pbpb": p3d,
                   "graph_replay": graph, "tile_autotune": tune, "mixed_systems_521": mixed,
                   "averaged_profiles_op3": avg, "jet_grids_op2": jet, "reference_defaults_301": dflt,
                   "records_to_host_events_per_s": host_rate},
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)
    ctx.close()


if __name__ == "__main__":
    main()
