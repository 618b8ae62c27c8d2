"""Writes tests/golden/reference_vectors.json from the reference translation
units that build on their own (oracle/_ref/libsmcref.so, built by
`make -C oracle ref` from the sources under /root/reference/src).

Every value here is produced by running reference code; the tests compare
the oracle restatement against these vectors, so they run without the
reference present.
"""
import ctypes as C
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import oracle  # noqa: E402


def main():
    r = oracle.ref()
    if r is None:
        raise SystemExit("oracle/_ref/libsmcref.so not built (make -C oracle ref)")
    out = {}
    ecms = [7.7, 19.6, 62.4, 200.0, 202.2113839, 2760.0, 3669.6307772, 5020.0, 5440.0]
    out["sigma_inel"] = [[e, r.ref_sigma_inel(e)] for e in ecms]

    rng = np.random.default_rng(20261202)
    pts = rng.uniform(0.0, 1.0, size=(24, 3))
    pts[0] = [0.0, 0.5, 0.5]
    pts[1] = [0.999, 0.999, 0.25]
    kfunc = []
    for ecm, lam in ((2760.0, 0.138), (200.0, 0.252)):
        for ta, tb in ((0.5, 1.5), (2.0, 0.3), (5.0, 5.0), (0.05, 8.0)):
            u = np.ascontiguousarray(pts, np.float64).ravel()
            v = np.zeros(len(pts))
            r.ref_kln_func(ecm, lam, ta, tb, len(pts), u.ctypes.data_as(C.POINTER(C.c_double)),
                           v.ctypes.data_as(C.POINTER(C.c_double)))
            kfunc.append(dict(ecm=ecm, lam=lam, ta=ta, tb=tb, u=pts.tolist(), f=v.tolist()))
    out["kln_func"] = kfunc
    kd = []
    for ta, tb in ((0.1, 0.1), (0.5, 1.5), (1.0, 1.0), (2.0, 0.3), (3.0, 3.0), (5.0, 5.0), (8.0, 2.0)):
        kd.append([ta, tb, r.ref_kln_dndy(2760.0, 0.138, ta, tb)])
    out["kln_dndy_bases"] = dict(ecm=2760.0, lam=0.138, values=kd)

    six = []
    for _ in range(10):
        a = rng.uniform(-2, 2, size=8)
        six.append([*a.tolist(), r.ref_six_point(*[float(v) for v in a])])
    out["six_point"] = six

    boxes = []
    for _ in range(20):
        x0, y0 = rng.uniform(-8, 8, size=2)
        x1, y1 = x0 + rng.uniform(-1, 1), y0 + rng.uniform(-1, 1)
        for w, qw in ((0.5, 0.3), (0.4938, 0.3), (0.5, 0.6)):
            o = np.zeros(4)
            r.ref_particle_box(x0, y0, 0, 0.0, 0.0, w, qw, o.ctypes.data_as(C.POINTER(C.c_double)))
            o2 = np.zeros(4)
            r.ref_particle_box(x0, y0, 1, x1, y1, w, qw, o2.ctypes.data_as(C.POINTER(C.c_double)))
            boxes.append(dict(x0=x0, y0=y0, x1=x1, y1=y1, w=w, qw=qw, box=o.tolist(), moved=o2.tolist()))
    out["particle_box"] = boxes

    rot = []
    for _ in range(10):
        cth, phi = rng.uniform(-1, 1), rng.uniform(0, 2 * np.pi)
        v = rng.uniform(-5, 5, size=3)
        o = np.ascontiguousarray(v.copy())
        r.ref_rotate(cth, phi, o.ctypes.data_as(C.POINTER(C.c_double)))
        rot.append([cth, phi, *v.tolist(), *o.tolist()])
    out["rotate"] = rot

    out["hulthen_cdf"] = [[x, r.ref_hulthen_cdf(x)] for x in (0.0, 0.1, 0.5, 1.0, 2.0, 4.0, 8.0, 16.0, 40.0)]
    out["hulthen_invcdf"] = [[u, r.ref_hulthen_invcdf(u)] for u in (0.01, 0.1, 0.3, 0.5, 0.7, 0.9, 0.99)]

    # parameters.dat read by the reference's ParameterReader (main.cpp:22-24)
    keys = ["which_mc_model", "sub_model", "lambda", "tmax", "tmax_subdivision", "alpha", "Aproj", "Atarg",
            "proj_deformed", "targ_deformed", "shape_of_nucleons", "collision_criterion", "shape_of_entropy",
            "quark_width", "gauss_nucl_width", "ecm", "bmin", "bmax", "Npmin", "Npmax", "cc_fluctuation_model",
            "cc_fluctuation_Gamma_theta", "maxx", "maxy", "dx", "dy", "finalFactor", "ecc_from_order",
            "ecc_to_order"]
    pd = {}
    path_pd = b"/root/reference/parameters.dat"
    for k in keys:
        found = C.c_int()
        v = r.ref_param_value(path_pd, 0, None, k.encode(), C.byref(found))
        pd[k] = v if found.value else None
    # command-line override (ParameterReader::readFromArguments, key=value)
    argv = (C.c_char_p * 2)(b"Aproj=197", b"ECM = 200  # comment")
    found = C.c_int()
    pd["_override_Aproj"] = r.ref_param_value(path_pd, 2, argv, b"aproj", C.byref(found))
    pd["_override_ecm"] = r.ref_param_value(path_pd, 2, argv, b"ecm", C.byref(found))
    out["parameters_dat"] = pd

    path = os.path.join(HERE, "reference_vectors.json")
    with open(path, "w") as f:
        json.dump(out, f, indent=1)
    print("wrote", path)


if __name__ == "__main__":
    main()
