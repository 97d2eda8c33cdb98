"""Writes the golden fixtures of tests/golden/*.npz from the compiled
reference (oracle/_ref/ref_harness: the unmodified hobywan/trinity sources
built with -O3 -ffp-contract=off, one OpenMP thread).

    make -C oracle            # builds oracle/_ref (needs the reference checkout)
    python tests/golden/make_golden.py

Each fixture holds the inputs (xy, tri, tags, solut), the parameters, and
every stage the reference produces: gradient, raw Hessian, locally
normalised metric, complexity, final metric, element qualities before
smoothing, the colouring, and xy / tensor / qualities / commit and task
counts after the smoothing sweeps.  The large GRID4 case (the reference's
own example mesh and field, examples/mesh/GRID4.mesh + examples/solut/
shock4.bb) stores SHA-256 digests of every output array plus values at 400
sampled vertices and triangles, to keep the fixture small.
"""
from __future__ import annotations

import hashlib
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402
from trinity_amd import meshgen as mg  # noqa: E402

OUT = os.path.dirname(os.path.abspath(__file__))
REF_EXAMPLES = os.environ.get("TRINITY_REF", "/root/reference") + "/examples"

FULL_KEYS = ["gradient", "hessian", "metric_local", "metric", "qualit0", "color", "xy",
             "tensor", "qualit"]


def digest(a: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def reverse_firstfit(nv, v2v_off, v2v, max_part=8):
    """A valid colouring different from the reference's: first-fit in
    descending vertex order."""
    color = np.zeros(nv, dtype=np.int32)
    for i in range(nv - 1, -1, -1):
        nb = v2v[v2v_off[i]:v2v_off[i + 1]]
        if nb.size == 0:
            continue
        used = set(color[nb].tolist())
        c = 1
        while c in used:
            c += 1
        if c >= max_part:
            raise RuntimeError("palette exceeded")
        color[i] = c
    return color


def case(name, m, tf, p, sweeps, hmin=1e-9, hmax=0.4, ext_color=False, compact=False):
    color = None
    if ext_color:
        topo = oracle.Topology(m.nv, m.tri)
        color = reverse_firstfit(m.nv, topo.v2v_off, topo.v2v)
    r = oracle.ref_dump(m.xy, m.tri, m.tags, m.solut, tf, p, hmin, hmax, sweeps, color=color,
                        threads=1)
    data = {
        "xy_in": m.xy, "tri": m.tri, "tags": m.tags, "solut": m.solut,
        "params": np.array([tf, p, hmin, hmax, sweeps], dtype=np.float64),
        "complexity": np.array([r["complexity"]]),
        "commits": r["commits"], "tasks": r["tasks"],
        "ext_color": np.array([1 if ext_color else 0], dtype=np.int32),
    }
    if compact:
        rng = np.random.default_rng(2024)
        iv = np.sort(rng.choice(m.nv, size=400, replace=False)).astype(np.int64)
        it = np.sort(rng.choice(m.nt, size=400, replace=False)).astype(np.int64)
        data["sample_v"] = iv
        data["sample_t"] = it
        for k in FULL_KEYS:
            data["sha_" + k] = np.array(digest(r[k]))
            data["s_" + k] = r[k][it] if k in ("qualit0", "qualit") else r[k][iv]
        data["ncolors"] = np.array([int(r["color"].max())], dtype=np.int32)
    else:
        for k in FULL_KEYS:
            data[k] = r[k]
    path = os.path.join(OUT, name + ".npz")
    np.savez_compressed(path, **data)
    print(f"{name}: nv={m.nv} nt={m.nt} {os.path.getsize(path) / 1024:.1f} KiB "
          f"commits={r['commits'].tolist()} colours={int(r['color'].max())}")


def main():
    if not oracle.ref_available():
        sys.exit("oracle/_ref/ref_harness is missing: run `make -C oracle` with the reference present")
    case("uniform32_tanh_p2", mg.workload("uniform32:tanh"), 0.8, 2, 2)
    case("jitter32_fronts_p1", mg.workload("jitter32:fronts"), 0.8, 1, 3)
    case("jitter40_layer_p4", mg.workload("jitter40:layer"), 0.8, 4, 3)
    case("jitter36_tanh_p0", mg.workload("jitter36:tanh"), 0.5, 0, 1)
    case("uniform24_quadratic_p2", mg.workload("uniform24:quadratic"), 1.0, 2, 2)
    case("jitter32_fronts_extcolor", mg.workload("jitter32:fronts"), 0.8, 2, 3, ext_color=True)
    g4 = mg.read_medit(REF_EXAMPLES + "/mesh/GRID4.mesh", REF_EXAMPLES + "/solut/shock4.bb")
    case("grid4_shock4", g4, 1.0, 2, 3, compact=True)


if __name__ == "__main__":
    main()
