"""Writes tests/golden/reference_fixtures.json: the known answers held by the
reference package's own test suite, transcribed with their file:line.

Run from the repository root:  python tests/golden/make_golden.py
(no reference code is imported or executed; the values are data).
"""
import json
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def f32(x):
    return np.float32(x)


def search_fixture_positions():
    """Positions of test/Algorithm/NeighbourSearchTest.jl:42-59 with Julia's promotion:
    start_x, start_y, gap are Float32 fields of Domain2D(0.15, 0.1, 0.2, 0.9, 0.9);
    err = 1e-3 is Float64, so `start_x + err` is formed in Float64; gap / 2 and
    gap * 3 / 2 stay Float32; the Float32 matrix xy rounds every entry."""
    sx, sy, gap, err = f32(0.1), f32(0.2), f32(0.15), 1e-3
    half = f32(gap / f32(2))
    three_half = f32(f32(gap * f32(3)) / f32(2))
    cx = [f32(np.float64(sx) + err), f32(sx + half), f32(np.float64(f32(sx + three_half)) - err)]
    cy = [f32(np.float64(sy) + err), f32(sy + half), f32(np.float64(f32(sy + gap)) - err)]
    xy = [(cx[i % 3], cy[i // 3]) for i in range(9)]
    return [[float(x), float(y)] for x, y in xy]


def main():
    fixtures = {
        "_note": "Known answers of the reference's own tests (EtherParallelParticles.jl 0.1.0), "
                 "transcribed as data; see each entry's source.",
        "domain": {
            "source": "test/Class/DomainTest.jl:10-40",
            "args": [0.15, -1.0, -2.0, 3.0, 3.0],
            "n_x": 26, "n_y": 33, "n": 858,
            "approx": {"gap": 0.15, "gap_square": 0.0225, "first_x": -1.0, "first_y": -2.0, "last_x": 3.0,
                       "last_y": 3.0, "span_x": 4.0, "span_y": 5.0, "gap_x": 4.0 / 26, "gap_y": 5.0 / 33},
            "cartesian_to_linear": [[2, 3], 54],
            "linear_to_cartesian": [54, [2, 3]],
            "inside": [[1.5, 0.3], True],
            "cartesian_from_position": [[1.5, 0.3], [17, 16]],
            "linear_from_position": [[1.5, 0.3], 407],
        },
        "stencil": {
            "source": "test/Class/NeighbourSystem/NeighbourSystemBaseTest.jl:10-25",
            "domain": [0.15, 0.1, 0.2, 0.9, 0.9],
            "count": [4, 6, 6, 6, 4, 6, 9, 9, 9, 6, 6, 9, 9, 9, 6, 4, 6, 6, 6, 4],
            "shape": [20, 9],
        },
        "active_pair": [
            {"source": "test/Class/NeighbourSystem/ActivePairTest.jl:10-19",
             "pairs": [[1, 1], [1, 2], [2, 1], [1, 3], [3, 1]],
             "adjacency": [[1, 1, 1], [1, 0, 0], [1, 0, 0]]},
            {"source": "test/Class/NeighbourSystem/NeighbourSystemTest.jl:22-37",
             "pairs": [[1, 1], [1, 2], [2, 1]],
             "adjacency": [[1, 1], [1, 0]]},
        ],
        "neighbour_search": {
            "source": "test/Algorithm/NeighbourSearchTest.jl:10-153",
            "domain": [0.15, 0.1, 0.2, 0.9, 0.9],
            "active_pair": [[1, 1], [1, 2], [2, 1]],
            "int_named_tuple": {"Tag": 1, "IsMovable": 1, "nCount": 1, "nIndex": 50},
            "float_named_tuple": {"RVec": 2, "nRVec": 100, "nR": 50},
            "basic_index_map_dict": {"PositionVec": "RVec", "nCount": "nCount", "nIndex": "nIndex", "nR": "nR"},
            "max_neighbour_number": 50,
            "positions": search_fixture_positions(),
            "tag": 1,
            "criterion": "defaultSearchCriterion",
            "n_count": [4, 5, 3, 5, 6, 3, 4, 5, 3],
            "neighbours": [[2, 4, 5, 7], [1, 3, 4, 5, 8], [2, 6, 9], [1, 2, 5, 7, 8], [1, 2, 4, 6, 7, 8],
                           [3, 5, 9], [1, 4, 5, 8], [2, 4, 5, 7, 9], [3, 6, 8]],
        },
        "action": {
            "source": "test/Algorithm/ActionTest.jl:10-86",
            "self": "FP[I, RVec] *= 0 and FP[I, RVec + 1] *= 0",
            "interaction": "FP[I, RVec] += 1 and FP[I, RVec + 1] += 1 per neighbour",
            "order": ["search", "selfaction", "interaction", "dynamic_selfaction", "dynamic_interaction"],
            "expected_columns_1_and_2": [4, 5, 3, 5, 6, 3, 4, 5, 3],
        },
        "particle_system": {
            "source": "test/Class/ParticleSystem/ParticleSystemTest.jl:12-94",
            "int_named_tuple": {"Tag": 1, "IsMovable": 1, "nCount": 1, "nIndex": 50},
            "float_named_tuple": {"PositionVec": 2, "Mass": 1, "Density": 1, "Volume": 1, "VelocityVec": 2,
                                  "dVelocityVec": 2, "dDensity": 1, "Pressure": 1, "StrainMat": 4,
                                  "dStrainMat": 4, "StressMat": 4, "nRVec": 100, "nR": 50, "nW": 50, "nDW": 50},
            "n_particles": 100, "n_capacity": 200,
            "n_int_capacity": 53, "n_float_capacity": 273,
            "basic_index": {"Tag": 1, "IsMovable": 2, "nCount": 3, "nIndex": 4, "nR": 124, "PositionVec": 1,
                            "nRVec": 24},
        },
        "named_index": {
            "source": "test/Class/ParticleSystem/NamedIndexTest.jl:10-33",
            "int_named_tuple": {"Tag": 1, "nCount": 1, "nIndex": 50},
            "n_int_capacity": 52, "n_float_capacity": 273,
        },
        "geometry": {
            "source": "test/Geometry/GeometryTest.jl:10-25",
            "rectangle": [0.0, 0.0, 2.0, 3.0], "gap": 0.1, "count": 600,
            "inside": [[[0.1, 0.2], True], [[2.1, 0.2], False]],
        },
    }
    path = os.path.join(HERE, "reference_fixtures.json")
    with open(path, "w") as f:
        json.dump(fixtures, f, indent=1)
    print(path)


if __name__ == "__main__":
    main()
