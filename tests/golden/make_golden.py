"""Writes tests/golden/reference_fixtures.json from what the reference itself holds.

Run in the build container (where /root/reference is mounted):
    python tests/golden/make_golden.py
Sources (all reference-held outputs, no reference code is executed except the
self-contained generated propagator benchmarks/test3d.cpp, built by
oracle/Makefile into oracle/_ref/test3d):
  * known-answer L2 vector: stdout of oracle/_ref/test3d
  * tests/eigenwave3d.ipynb saved outputs: 4th-order staggered weights (cell 11),
    kernel op counts (cells 18-20; 8th-order counts, see below), CFL limit (cell 3)
  * tests/eigenwave3d-higher_order.ipynb: 8th-order velocity coefficients
    (cell 31), 4th-order op counts (cells 27-28), Robertsson ghost planes
    (cells 34-35)
  * tests/eigenwave3d-read.ipynb: 4th-order read-path op counts (cells 14-15)
"""
import json
import os
import re
import subprocess

REF = os.environ.get("OPESCI_REFERENCE", "/root/reference")
HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))


def cell_text(nb, i):
    outs = []
    for o in nb["cells"][i].get("outputs", []):
        t = o.get("text") or o.get("data", {}).get("text/plain")
        if t:
            outs.append("".join(t))
    return "\n".join(outs)


def main():
    fx = {}
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "ref"], check=True)
    out = subprocess.run([os.path.join(ROOT, "oracle", "_ref", "test3d")], check=True,
                         capture_output=True, text=True,
                         env=dict(os.environ, OMP_NUM_THREADS="8", OMP_PROC_BIND="close")).stdout
    l2 = {}
    for line in out.splitlines():
        m = re.match(r"(\w+)_l2\t([0-9.]+)", line)
        if m:
            l2[m.group(1)] = float(m.group(2))
    fx["test3d_known_answer"] = {
        "source": "benchmarks/test3d.cpp built with g++ -Wall -std=c++11 -g -O3 -fno-tree-vectorize -fopenmp",
        "config": {"order": 4, "grid_size": 100, "dt": 0.002, "ntsteps": 1000, "mu": 0.5,
                   "lambda": 0.5, "rho": 1.0, "real": "float"},
        "l2": l2,
    }
    nb = json.load(open(os.path.join(REF, "tests", "eigenwave3d.ipynb")))
    t11 = cell_text(nb, 11)
    fx["fd_weights_4th"] = {"source": "tests/eigenwave3d.ipynb cell 11",
                            "text": t11.strip().splitlines()[-1].strip()}
    # eigenwave3d.ipynb's source cell 7 sets order [2,4,4,4], but its saved
    # outputs of cells 18-19 are the counts of an 8th-order kernel (3 axes x
    # 8 stencil points per update: 72 / 120); the outputs predate the source.
    # The key names the notebook, not an order.
    fx["ai_nb_eigenwave3d"] = {"source": "tests/eigenwave3d.ipynb cells 18-20",
                    "velocity": json.loads(cell_text(nb, 18).replace("(", "[").replace(")", "]")),
                    "stress": json.loads(cell_text(nb, 19).replace("(", "[").replace(")", "]")),
                    "overall": json.loads(cell_text(nb, 20).replace("(", "[").replace(")", "]"))}
    fx["cfl_4th_h001"] = {"source": "tests/eigenwave3d.ipynb cell 3", "value": float(cell_text(nb, 3))}
    nbh = json.load(open(os.path.join(REF, "tests", "eigenwave3d-higher_order.ipynb")))
    t31 = cell_text(nbh, 31)
    u_line = [l for l in t31.split("\\n") if l.startswith("U[t1][x][y][z]=")][0]
    terms = re.findall(r"([+-]?\s*[0-9.]+e[+-]?[0-9]+)\*(\w+)\[t1\]\[([^\]]+)\]\[([^\]]+)\]\[([^\]]+)\]", u_line)
    fx["velocity_8th_U_terms"] = {
        "source": "tests/eigenwave3d-higher_order.ipynb cell 31 (dt/dx = 0.2, beta = 1)",
        "terms": [[float(c.replace(" ", "")), f, a, b, c2] for c, f, a, b, c2 in terms],
    }
    # set_accuracy([1,2,2,2]) of the older API: 4th order (36 / 60)
    fx["ai_nb_higher_order"] = {"source": "tests/eigenwave3d-higher_order.ipynb cells 27-28",
                    "velocity": json.loads(cell_text(nbh, 27).replace("(", "[").replace(")", "]")),
                    "stress": json.loads(cell_text(nbh, 28).replace("(", "[").replace(")", "]"))}
    fx["robertsson_8th_U_bc"] = {"source": "tests/eigenwave3d-higher_order.ipynb cell 34",
                                 "text": cell_text(nbh, 34)}
    fx["robertsson_8th_Txx_bc"] = {"source": "tests/eigenwave3d-higher_order.ipynb cell 35",
                                   "text": cell_text(nbh, 35)}
    nbr = json.load(open(os.path.join(REF, "tests", "eigenwave3d-read.ipynb")))
    # set_accuracy([1,2,2,2]) on the read path: 4th order
    fx["ai_nb_read"] = {"source": "tests/eigenwave3d-read.ipynb cells 14-15",
                    "velocity": json.loads(cell_text(nbr, 14).replace("(", "[").replace(")", "]")),
                    "stress": json.loads(cell_text(nbr, 15).replace("(", "[").replace(")", "]"))}
    with open(os.path.join(HERE, "reference_fixtures.json"), "w") as f:
        json.dump(fx, f, indent=1, sort_keys=True)
    print("wrote", os.path.join(HERE, "reference_fixtures.json"))


if __name__ == "__main__":
    main()
