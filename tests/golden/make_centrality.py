"""Writes tests/golden/centrality_npart.json: the Npart centrality tables the
reference ships for minimum-bias Pb+Pb at 5020 and 2760 GeV
(scripts/centrality_cut_tables/iebe_centralityCut_Npart_MCGlbPbPb{5020,2760}_
sigmaNN_gauss_d0.9_withMultFluct.dat; columns: centrality [%], Npart at the
bin's lower edge, b_min, b_max).  They come from superMC.e operation=9
bmin=0 bmax=20 with the reference's parameters.dat
(scripts/generate_minbiasEcc_jobs_WSUGrid.py:29), events sorted by Npart
(scripts/centrality_cut_h5.py:43-89).  Run here, where /root/reference is
present; the JSON is the fixture the GPU test reads."""
import json
import os

REF = "/root/reference/scripts/centrality_cut_tables"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "centrality_npart.json")


def read(ecm):
    rows = []
    name = f"iebe_centralityCut_Npart_MCGlbPbPb{ecm}_sigmaNN_gauss_d0.9_withMultFluct.dat"
    for ln in open(os.path.join(REF, name)):
        if ln.startswith("#") or not ln.strip():
            continue
        c, npart, bmin, bmax = (float(v) for v in ln.split())
        rows.append([c, npart, bmin, bmax])
    return dict(source=f"scripts/centrality_cut_tables/{name}", columns=["centrality_pct", "npart", "b_min", "b_max"],
                rows=rows)


if __name__ == "__main__":
    json.dump({"5020": read(5020), "2760": read(2760)}, open(OUT, "w"), indent=0)
    print("wrote", OUT)
