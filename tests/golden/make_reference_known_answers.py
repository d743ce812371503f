"""Writes tests/golden/reference_known_answers.json.

Every case below is data transcribed from warp-md 0.4.30's own tests for the
RDF / RMSD / minimum-image path (inputs and the asserted outputs), cited by
file:line.  Nothing here runs or imports reference code; the reference's
assertions are restated as expected values so the CPU oracle and the HIP engine
can both be pinned against them.

Run:  python tests/golden/make_reference_known_answers.py
"""
import json
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def shell_f32(r_in, r_out):
    f = np.float32
    return float(f(f(4.0) / f(3.0)) * f(np.pi) * (f(r_out) ** 3 - f(r_in) ** 3))


TWO_CA = [[0.0, 0.0, 0.0], [1.0, 0.0, 0.0]]  # tests.rs:88-102 build_system positions0

cases = {
    "rdf": [
        {
            "name": "rdf_plan_counts",
            "source": "crates/traj-engine/src/tests/correlation_structure_tests.rs:1896-1913",
            "frames": [TWO_CA],
            "sel_a": [0, 1], "sel_b": [0, 1], "bins": 5, "r_max": 5.0, "pbc": "none",
            "options": {},
            "expect_counts": [0, 2, 0, 0, 0],
        },
        {
            "name": "rdf_plan_density_normalization_uses_common_count",
            "source": "crates/traj-engine/src/tests/correlation_structure_tests.rs:1916-1934",
            "frames": [TWO_CA],
            "sel_a": [0, 1], "sel_b": [0, 1], "bins": 5, "r_max": 5.0, "pbc": "none",
            "options": {"density": 1.0, "volume": False, "raw_rdf": False, "intrdf": True},
            "expect_counts_at": {"1": 2},
            "expect_g_at": {"1": 2.0 / shell_f32(1.0, 2.0)},
            "expect_integral_at": {"1": 1.0},
            "tol": 1e-6,
        },
        {
            "name": "rdf_plan_number_density_output_normalizes_by_reference_and_shell",
            "source": "crates/traj-engine/src/tests/correlation_structure_tests.rs:1937-1956",
            "frames": [TWO_CA],
            "sel_a": [0, 1], "sel_b": [0, 1], "bins": 5, "r_max": 5.0, "pbc": "none",
            "options": {"density": 1.0, "volume": False, "raw_rdf": False, "intrdf": True,
                        "number_density": True},
            "expect_counts_at": {"1": 2},
            "expect_g_at": {"1": 1.0 / shell_f32(1.0, 2.0)},
            "expect_integral_at": {"1": 1.0},
            "tol": 1e-6,
        },
        {
            "name": "rdf_plan_raw_output_returns_counts_from_rust",
            "source": "crates/traj-engine/src/tests/correlation_structure_tests.rs:1959-1976",
            "frames": [TWO_CA],
            "sel_a": [0, 1], "sel_b": [0, 1], "bins": 5, "r_max": 5.0, "pbc": "none",
            "options": {"density": 1.0, "volume": False, "raw_rdf": True, "intrdf": False},
            "expect_counts_at": {"1": 2},
            "expect_g_at": {"1": 2.0},
            "expect_no_integral": True,
            "tol": 0.0,
        },
        {
            "name": "rdf_plan_direct_atom_mode_uses_selected_chunks",
            "source": "crates/traj-engine/src/tests/correlation_structure_tests.rs:2012-2115",
            "frames": [
                [[0.0, 0.0, 0.0], [1000.0, 1000.0, 1000.0], [1.0, 0.0, 0.0]],
                [[0.0, 0.0, 0.0], [1000.0, 1000.0, 1000.0], [2.0, 0.0, 0.0]],
            ],
            "sel_a": [0, 2], "sel_b": [0, 2], "bins": 4, "r_max": 4.0, "pbc": "none",
            "options": {"density": 1.0, "volume": False, "raw_rdf": True, "intrdf": False},
            "expect_counts": [0, 2, 2, 0],
            "expect_g": [0.0, 2.0, 2.0, 0.0],
            "tol": 0.0,
        },
        {
            "name": "test_rdf_plan_accepts_frame_indices_with_multimodel_pdb (full)",
            "source": "python/warp_md/tests/test_rdf_plan.py:11-45",
            "frames": [[[0.0, 0.0, 0.0], [1.0, 0.0, 0.0]], [[0.0, 0.0, 0.0], [3.0, 0.0, 0.0]]],
            "sel_a": [0, 1], "sel_b": [0, 1], "bins": 4, "r_max": 4.0, "pbc": "none",
            "options": {},
            "expect_counts": [0, 2, 0, 2],
        },
        {
            "name": "test_rdf_plan_accepts_frame_indices_with_multimodel_pdb (subset)",
            "source": "python/warp_md/tests/test_rdf_plan.py:36-45",
            "frames": [[[0.0, 0.0, 0.0], [1.0, 0.0, 0.0]], [[0.0, 0.0, 0.0], [3.0, 0.0, 0.0]]],
            "frame_indices": [1, -1, 99],
            "sel_a": [0, 1], "sel_b": [0, 1], "bins": 4, "r_max": 4.0, "pbc": "none",
            "options": {},
            "expect_counts": [0, 0, 0, 4],
        },
        {
            "name": "test_rdf_wrapper_accepts_atom_index_masks",
            "source": "python/warp_md/tests/test_rdf_plan.py:267-297 (maximum=2, bin_spacing=1 -> bins=2, r_max=2)",
            "frames": [[[0.0, 0.0, 0.0], [0.5, 0.0, 0.0]]],
            "sel_a": [0], "sel_b": [1], "bins": 2, "r_max": 2.0, "pbc": "none",
            "options": {"density": 1.0, "raw_rdf": True},
            "expect_counts": [1, 0],
        },
    ],
    "rmsd": [
        {
            "name": "rmsd_plan_align_zero",
            "source": "crates/traj-engine/src/tests/analysis_smoke_tests.rs:23-37",
            "reference": TWO_CA,
            "frames": [TWO_CA],
            "sel": [0, 1],
            "align": True,
            "expect_abs_below": [1e-6],
        },
        {
            "name": "test_rotation_matrix_basic",
            "source": "python/warp_md/tests/test_rotation_matrix.py:40-59",
            "reference": [[0.0, 0.0, 0.0], [1.0, 0.0, 0.0], [0.0, 1.0, 0.0]],
            "frames": [
                [[0.0, 0.0, 0.0], [1.0, 0.0, 0.0], [0.0, 1.0, 0.0]],
                # ref @ rot.T with rot a +90 degree rotation about z
                [[0.0, 0.0, 0.0], [0.0, 1.0, 0.0], [-1.0, 0.0, 0.0]],
            ],
            "sel": [0, 1, 2],
            "align": True,
            "expect_abs_below": [1e-6, 1e-6],
            # the fitted rotation of frame 1 onto the reference is rot.T
            "expect_rotation_frame1": [[0.0, 1.0, 0.0], [-1.0, 0.0, 0.0], [0.0, 0.0, 1.0]],
        },
    ],
    "min_image": [
        {
            "name": "minimum_image_delta_wraps_orthorhombic_box",
            "source": "crates/traj-core/src/pbc_math.rs:170-174",
            "a": [9.5, 1.0, 1.0], "b": [0.5, 2.0, 1.0], "kind": "orthorhombic",
            "box": [10.0, 10.0, 10.0], "scale": 1.0, "expect": [1.0, 1.0, 0.0], "tol": 0.0,
        },
        {
            "name": "minimum_image_delta_without_box_is_raw_difference",
            "source": "crates/traj-core/src/pbc_math.rs:176-180",
            "a": [1.0, 2.0, 3.0], "b": [4.0, 8.0, 15.0], "kind": "none",
            "box": None, "scale": 1.0, "expect": [3.0, 6.0, 12.0], "tol": 0.0,
        },
        {
            "name": "minimum_image_vector_scales_orthorhombic_box",
            "source": "crates/traj-core/src/pbc_math.rs:190-205",
            "a": [0.95, 0.0, 0.0], "b": [0.05, 0.0, 0.0], "kind": "orthorhombic",
            "box": [10.0, 10.0, 10.0], "scale": 0.1, "expect": [0.1, 0.0, 0.0], "tol": 1e-12,
        },
        {
            "name": "minimum_image_displacement_scales_triclinic_box",
            "source": "crates/traj-core/src/pbc_math.rs:207-221",
            "a": [0.0, 0.0, 0.0], "b": [0.9, 0.0, 0.0], "kind": "triclinic",
            "box": [[10.0, 0.0, 0.0], [0.0, 10.0, 0.0], [0.0, 0.0, 10.0]], "scale": 0.1,
            "expect": [-0.1, 0.0, 0.0], "tol": 1e-12,
        },
    ],
}

if __name__ == "__main__":
    path = os.path.join(HERE, "reference_known_answers.json")
    with open(path, "w") as fh:
        json.dump(cases, fh, indent=1)
    print("wrote", path)
