"""Extracts the numeric expectations the reference holds for the batched
path into tests/golden/golden.json.  Run on a machine that holds the
reference checkout (the tests read only the JSON it writes):

    python tests/golden/make_golden.py /root/reference

Sources (all data, no code):
  test/microscale/before-refactor-outputs/explicit_cpu_out.txt:4,27-50
      F, stress (3x3) and two 6x6 tangent blocks of Network_2_7 computed by the
      non-batched explicit solver (no division by the RVE volume).
  test/microscale/before-refactor-outputs/explicit_gpu_out.txt:27-50 (same digits)
  test/microscale/before-refactor-outputs/implicit_out.txt (last block)
  test/microscale/fiber_networks/Network_2_7.txt.summary (orientation tensor)
  test/microscale/test_reorder_mesh.cc:14-60 (ReorderMesh known answer)
"""
import json
import os
import re
import sys


def floats(line):
    return [float(t) for t in line.split()]


def parse_out(path):
    lines = open(path).read().splitlines()
    F = None
    blocks = []
    for i, l in enumerate(lines):
        m = re.search(r"F=\[(.*)\]", l)
        if m:
            rows = m.group(1).split(";")
            F = [[float(t) for t in r.split()] for r in rows]
        if l.strip() == "Stress from run":
            S = [floats(lines[i + k]) for k in (1, 2, 3)]
            Cm = [floats(lines[i + k]) for k in range(5, 11)]
            blocks.append({"stress33": S, "C": Cm, "line": i + 1})
    return F, blocks


def main(ref):
    out_dir = os.path.dirname(os.path.abspath(__file__))
    base = os.path.join(ref, "test/microscale/before-refactor-outputs")
    F, cpu = parse_out(os.path.join(base, "explicit_cpu_out.txt"))
    _, gpu = parse_out(os.path.join(base, "explicit_gpu_out.txt"))
    _, impl = parse_out(os.path.join(base, "implicit_out.txt"))
    summ = open(os.path.join(ref, "test/microscale/fiber_networks/Network_2_7.txt.summary")).read()
    orient = [[float(t) for t in r] for r in re.findall(r"\[([-\d.e]+) ([-\d.e]+) ([-\d.e]+)\]", summ)]
    src = open(os.path.join(ref, "test/microscale/test_reorder_mesh.cc")).read()

    def arr(name):
        m = re.search(name + r"\s*(?:=\s*)?\{([^}]*)\}", src)
        return [int(float(t)) for t in m.group(1).replace("\n", " ").split(",")]

    g = {
        "network_2_7": {
            "F": F,
            "explicit_cpu": cpu,
            "explicit_gpu": gpu,
            "implicit_last": impl[-1],
            "summary_orientation": orient,
        },
        "reorder_kat": {
            "nverts": 8,
            "fixed": arr(r"std::array<Ordinal,3> fixed_verts"),
            "initial_connectivity": arr(r"initial_connectivity"),
            "final_connectivity": arr(r"final_connectivity"),
            "final_coordinates": arr(r"final_coordinates"),
        },
    }
    with open(os.path.join(out_dir, "golden.json"), "w") as f:
        json.dump(g, f, indent=1)
    print("wrote", os.path.join(out_dir, "golden.json"))


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
