"""Shared helpers for the GPU parity tests: parameter sets for the
workloads, CPU-generated nucleon configurations, and the oracle
evaluation of one event from the nucleons/pairs the HIP path reports."""
import math

import numpy as np

import oracle

Q = 2.0 ** -24


def glauber_params(system="AuAu", **kw):
    """MC-Glauber parameter sets (reference keys + extension keys)."""
    from supermc_amd import default_params
    base = dict(which_mc_model=5, sub_model=1, shape_of_nucleons=4, gauss_nucl_width=0.5, shape_of_entropy=2,
                collision_criterion=1, bmin=0.0, bmax=20.0, Npmin=2, Npmax=500, maxx=13.0, maxy=13.0, dx=0.1,
                dy=0.1, finalFactor=1.0, ecc_from_order=2, ecc_to_order=5, max_events=64)
    if system == "AuAu":
        # Au+Au: R=6.38, a=0.535, d=0.4, sigma=42 mb, alpha=0.14, no fluctuations
        base.update(Aproj=197, Atarg=197, proj_ws_R=6.38, proj_ws_a=0.535, targ_ws_R=6.38, targ_ws_a=0.535,
                    nucleon_min_dist=0.4, sigma_nn_mb=42.0, alpha=0.14, cc_fluctuation_model=0)
    elif system == "PbPb":
        # Pb+Pb: R=6.62, a=0.546, sigma=64 mb, Gamma fluctuations theta=1.25 (k=0.8)
        base.update(Aproj=208, Atarg=208, proj_ws_R=6.62, proj_ws_a=0.546, targ_ws_R=6.62, targ_ws_a=0.546,
                    nucleon_min_dist=0.4, sigma_nn_mb=64.0, alpha=0.14, cc_fluctuation_model=6,
                    cc_fluctuation_Gamma_theta=1.25)
    else:
        raise ValueError(system)
    base.update(kw)
    return default_params(**base)


def cpu_configs(p, nev, seed, bvals=None):
    """Nucleon positions from the oracle's sampler, quantised to 2^-24 fm."""
    rng = oracle.Rng(seed)
    xa = np.zeros((nev, p.Aproj, 2))
    xb = np.zeros((nev, p.Atarg, 2))
    bs = np.zeros(nev)
    for e in range(nev):
        b = bvals[e] if bvals is not None else 14.0 * math.sqrt(rng.uniform())
        xa[e] = rng.populate(p.Aproj, p.proj_ws_R, p.proj_ws_a, p.nucleon_min_dist, b / 2)
        xb[e] = rng.populate(p.Atarg, p.targ_ws_R, p.targ_ws_a, p.nucleon_min_dist, -b / 2)
        bs[e] = b
    return np.round(xa / Q) * Q, np.round(xb / Q) * Q, bs


def oracle_event(ctx, ev, kln_table=None, sizes=None):
    """Runs the oracle on the nucleons (and fluct factors) of event ev as
    reported by the HIP path.  Returns a dict with the oracle collision
    result, densities and moments, plus the GPU lists for comparison.
    sizes = (A, B): the event's nucleus sizes when they are smaller than
    the context's (an event of a mixed batch)."""
    p, d = ctx.params, ctx.derived
    nuc = ctx.nucleons(ev)
    if sizes is not None:
        for side, n in zip(("proj", "targ"), sizes):
            for key in ("xy_", "wound_", "fluct_"):
                nuc[key + side] = nuc[key + side][:n]
    pa, pb, pf = ctx.pairs(ev)
    xa, xb = nuc["xy_proj"], nuc["xy_targ"]
    w = d.width
    # the HIP path returns nucleons in collision order; the oracle sorts
    # again with the reference's key and must find them already sorted
    perm_a = oracle.sort_nucleus(xa, w)
    perm_b = oracle.sort_nucleus(xb, w)
    col = oracle.collide(xa[perm_a], xb[perm_b], w, d.dsq)
    out = dict(nuc=nuc, gpu_pairs=(pa, pb, pf), col=col, perm_a=perm_a, perm_b=perm_b)
    # participants in wound order with the GPU's fluct factors
    ia = np.argsort(np.where(nuc["wound_proj"] >= 0, nuc["wound_proj"], 1 << 30))[: col["npart_a"]]
    ib = np.argsort(np.where(nuc["wound_targ"] >= 0, nuc["wound_targ"], 1 << 30))[: col["npart_b"]]
    fa = nuc["fluct_proj"][ia].astype(np.float64)
    fb = nuc["fluct_targ"][ib].astype(np.float64)
    mid = (xa[pa] + xb[pb]) / 2.0
    g = oracle.grid(p.maxx, p.dx, p.maxy, p.dy)
    out["grid"] = g
    out["TA"] = oracle.thickness(g, w, xa[ia])
    out["TB"] = oracle.thickness(g, w, xb[ib])
    if p.which_mc_model == 5:
        aw = oracle.uli_weights(pa, pb, len(xa), len(xb)) if p.sub_model == 2 else None
        s, dndy = oracle.density_glauber(g, w, p.alpha, xa[ia], fa, xb[ib], fb, mid, pf.astype(np.float64),
                                         sub_model=p.sub_model, aw=aw)
        out["aw"] = aw
    elif p.which_mc_model == 7:
        s, dndy = oracle.density_trento(g, w, xa[ia], fa, xb[ib], fb)
    else:
        # MC-KLN through the dN/dy table; the table itself is checked entry by
        # entry against the oracle's quadrature (test_kln_table_entries)
        table = kln_table if kln_table is not None else ctx.kln_make_table()
        s, dndy = oracle.density_kln_table(g, out["TA"], out["TB"], table, d.kln_dT)
    out["s"], out["dndy"] = s, dndy
    boxes = np.array([oracle.particle_box(x, y, w) for x, y in np.concatenate([xa[ia], xb[ib]])])
    out["boxes"] = boxes
    out["mom"] = oracle.moments(g, s * p.finalFactor, boxes, len(pa), p.ecc_from_order, p.ecc_to_order)
    out["ia"], out["ib"] = ia, ib
    return out


def rel_linf(a, ref):
    ref = np.asarray(ref, np.float64)
    return float(np.max(np.abs(np.asarray(a, np.float64) - ref)) / max(np.max(np.abs(ref)), 1e-300))
