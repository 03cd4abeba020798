"""Writes the golden fixtures of tests/golden/ from values the reference holds.

The reference is Julia and cannot run here, so each vector below is
transcribed from the reference file named beside it (data only):
  - neighbor lattice KAT          test/testCommunity.jl:361-415
  - add/remove KAT                test/testCommunity.jl:241-316
  - ODE convergence KAT           test/testCommunity.jl:420-464
  - SDE variance KAT              test/testCommunity.jl:469-512
  - Aggregation printed outputs   docs/src/Aggregation.md:146-195, :243-279
  - Philox4x32-10 known answers   Random123 kat_vectors (Salmon et al., SC'11),
    the generator the library uses for in-kernel noise (the reference's own
    Xoshiro and CUDA.jl device streams cannot be reproduced bit for bit)
Run: python tests/golden/make_golden.py
"""
import json
import os

HERE = os.path.dirname(os.path.abspath(__file__))

lattice = {
    "source": "test/testCommunity.jl:361-415",
    "simBox_1d": [0.5, 1.5],
    "radius": 1.1,
    "algorithms": [
        {"name": "Full"},
        {"name": "VerletTime", "skin": 1.1, "dtNeighborRecompute": 1.0, "nMaxNeighbors": 10},
        {"name": "VerletDisplacement", "skin": 1.1, "nMaxNeighbors": 10},
        {"name": "CellLinked", "cellEdge": 1.0},
        {"name": "CLVD", "skin": 1.1, "nMaxNeighbors": 10, "cellEdge": 10.0},
    ],
    "expected": {
        "1": [1, 2, 1],
        "2": [2, 3, 2, 3, 4, 3, 2, 3, 2],
        "3": [3, 4, 3, 4, 5, 4, 3, 4, 3,
              4, 5, 4, 5, 6, 5, 4, 5, 4,
              3, 4, 3, 4, 5, 4, 3, 4, 3],
    },
}

add_remove = {
    "source": "test/testCommunity.jl:241-316",
    "N": 100,
    "preallocateAgents": 50,
    "after_rule": {"N": 100, "NAdd_": 50, "NRemove_": 50, "idMax_": 150},
    "after_update": {"N": 100, "idMax_": 150, "lf_contains": list(range(2, 101, 2))},
}

aggregation = {
    "source": "docs/src/Aggregation.md:146-195 and :243-279",
    "model": {"rRep": 0.8, "fRep": 1.0, "rAtr": 1.0, "fAtr": 1.0, "D": 0.0, "dt": 0.1},
    "two_agents": {
        "simBox": [[-5.0, 5.0], [-2.0, 2.0]],
        "x0": [-0.1, 0.1], "y0": [0.0, 0.0],
        "x": [[-0.1, 0.1], [-0.16, 0.16], [-0.22, 0.22], [-0.268, 0.268]],
        "x_step5_agent0": -0.304,
        "fx": [[-0.6, 0.6], [-0.6, 0.6], [-0.48, 0.48], [-0.36, 0.36]],
        "fx_step5_agent1": 0.264,
        "t": [0.1, 0.2, 0.3, 0.4, 0.5],
    },
    "wall": {
        "simBox": [[-2.0, 2.0], [-2.0, 2.0]],
        "x0": [0.0], "y0": [-2.5],
        "y": [-2.5, -2.4, -2.3, -2.2, -2.1, -2.0, -1.9, -1.8],
        "fy": [1.0] * 10,
        "fx": [0.0] * 10,
    },
}

ode = {
    "source": "test/testCommunity.jl:420-464",
    "rate": 0.5, "x0": 1.0, "dt": 0.1, "steps": 100, "tol": 0.05, "N": 3,
    "simBox": [[0.0, 1.0]] * 3,
}

sde = {
    "source": "test/testCommunity.jl:469-512",
    "g": 1.0, "x0": 0.0, "dt": 0.1, "steps": 100, "tol": 0.5, "N": 10000,
    "simBox": [[0.0, 1.0]] * 3,
}

philox = {
    "source": "Random123 kat_vectors, philox4x32_10",
    "vectors": [
        {"ctr": [0, 0, 0, 0], "key": [0, 0], "out": ["6627e8d5", "e169c58d", "bc57ac4c", "9b00dbd8"]},
        {"ctr": [0xFFFFFFFF] * 4, "key": [0xFFFFFFFF] * 2, "out": ["408f276d", "41c83b0e", "a20bc7c6", "6d5451fd"]},
        {"ctr": [0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344], "key": [0xA4093822, 0x299F31D0],
         "out": ["d16cfe09", "94fdcceb", "5001e420", "24126ea1"]},
    ],
}

files = {
    "ref_neighbors_lattice.json": lattice,
    "ref_add_remove.json": add_remove,
    "ref_aggregation_lag.json": aggregation,
    "ref_ode_convergence.json": ode,
    "ref_sde_variance.json": sde,
    "philox_kat.json": philox,
}

if __name__ == "__main__":
    for name, obj in files.items():
        with open(os.path.join(HERE, name), "w") as f:
            json.dump(obj, f, indent=1)
        print("wrote", name)
