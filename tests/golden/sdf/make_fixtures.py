"""Writes the SDF golden fixtures here from the reference's own test data
(/root/reference/tests/fixtures/oracle-sdf: the upstream C++ oracle's field samples, optimizer
runs, curvature seeds and manifold updates), re-serialised compactly with the keys the tests read.
Run once in a container that has the reference; the outputs are data (inputs and expected outputs).
"""
import json
import os

SRC = "/root/reference/tests/fixtures/oracle-sdf"
HERE = os.path.dirname(os.path.abspath(__file__))
FIELDS = ["box", "sphere", "capsule", "cylinder", "cone", "torus", "spike-plate", "spiky-wheel", "bend", "bend-0.15",
          "bend-0.55", "bend-0.95", "twist", "twist-0.50", "twist-2.50", "twist-5.00", "displace", "displace-0.05",
          "displace-0.10", "displace-0.20"]


def load(name):
    with open(os.path.join(SRC, name)) as f:
        return json.load(f)


def dump(name, data):
    with open(os.path.join(HERE, name), "w") as f:
        json.dump(data, f, separators=(",", ":"))


def main():
    dump("fields.json", {n: {"bounds": load(n + ".json")["bounds"],
                             "samples": [[s["p"], s["phi"], s["gradient"]] for s in load(n + ".json")["samples"]]} for n in FIELDS})
    dump("seeds.json", {n: {"centerOfMass": load(f"seeds-{n}.json")["centerOfMass"],
                            "seeds": [s["p"] for s in load(f"seeds-{n}.json")["seeds"]]} for n in FIELDS})
    dump("optimizer.json", load("optimizer.json"))
    dump("optimizers.json", load("optimizers.json"))
    keep = ("name", "step", "shapeA", "shapeB", "bodyA", "bodyB", "params", "dt", "initialCache", "rng", "points", "stats")
    for name in ("manifolds.json", "manifolds-optimizers.json"):
        cases = []
        for c in load(name)["cases"]:
            c = {k: c[k] for k in keep if k in c}
            c["rng"] = {"before": c["rng"]["before"], "after": c["rng"]["after"]}
            for b in ("bodyA", "bodyB"):
                c[b] = {k: c[b][k] for k in ("position", "orientation", "scale", "static", "velocity", "angularVelocity")}
            cases.append(c)
        dump(name, {"cases": cases})


if __name__ == "__main__":
    main()
