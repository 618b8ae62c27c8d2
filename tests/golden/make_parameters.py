"""Writes tests/golden/reference_parameters.json: the key = value settings
of the reference's parameters.dat (its run input; comments dropped), so a
driver test can run operation 3 with the reference's own values.  Run here,
where /root/reference is present; the JSON is the fixture the test reads."""
import json
import os

SRC = "/root/reference/parameters.dat"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "reference_parameters.json")


def read():
    vals = {}
    for ln in open(SRC):
        body = ln.split("#")[0].strip()
        if "=" not in body:
            continue
        k, v = (x.strip() for x in body.split("=", 1))
        if k:
            vals[k] = v
    return vals


if __name__ == "__main__":
    json.dump({"source": "parameters.dat", "values": read()}, open(OUT, "w"), indent=0)
    print("wrote", OUT)
