"""Writes tests/golden/reference_fixtures.json: the known answers and inline fixtures of the
reference's own tests for this path, recorded as data (inputs and expected outputs), each with the
test it comes from. Nothing of the reference is imported or executed; the numbers are the values
those tests state.

Run: python tests/golden/make_reference_fixtures.py
"""
from __future__ import annotations

import json
import math
from pathlib import Path

OUT = Path(__file__).resolve().parent / "reference_fixtures.json"

fixtures = {
    # photara/tests/splat_test.cpp:4634-4655 (test_fused_adam_parity): first Adam step.
    "adam_first_step": {
        "source": "photara/tests/splat_test.cpp:4634-4655",
        "param": [1.0, -2.0], "grad": [0.25, -0.5], "lr": 1e-3, "step": 1, "epsilon": 1e-15,
        "beta1": 0.9, "beta2": 0.999,
        "expect_param": [0.999, -1.999], "expect_first": [0.025, -0.05], "expect_second": [6.25e-5, 2.5e-4],
        "tol_param": 1e-6, "tol_first": 1e-7, "tol_second": 1e-8,
    },
    # splat_test.cpp:4612-4632 (test_adam_rejects_non_finite_gradients).
    "adam_non_finite": {
        "source": "photara/tests/splat_test.cpp:4612-4632",
        "param": [1.0, 2.0, 3.0], "grad": ["inf", "nan", "-inf"], "lr": 1e-3, "step": 1,
        "expect_param": [1.0, 2.0, 3.0],
    },
    # splat_test.cpp:4758-4791 (test_active_sh_prefix_adam): rows of 12 (4 bases x 3), active 3.
    "adam_active_prefix": {
        "source": "photara/tests/splat_test.cpp:4758-4791",
        "rows": 2, "stride": 12, "active": 3, "lr": 1e-3, "rest_lr": 1e-4, "grad": 0.25, "epsilon": 1e-15,
        "initial": list(range(1, 25)), "expect_delta": -1e-3, "expect_first": 0.025, "expect_second": 6.25e-5,
    },
    # splat_test.cpp:3563-3645 (test_alpha_parameter_gradients): centre alpha closed forms.
    "centre_alpha": {
        "source": "photara/tests/splat_test.cpp:3615-3645",
        "mean": [0.0, 0.0, 2.0], "log_scales": [math.log(0.15)] * 3, "quaternion": [1.0, 0.0, 0.0, 0.0],
        "opacity_logit": -2.19722458, "sh": [[0.5, 0.25, 0.1]],
        "camera": {"fx": 40.0, "fy": 40.0, "cx": 16.0, "cy": 16.0, "width": 32, "height": 32},
        "centre_pixel": [16, 16],
        "expect_alpha_kernel0": 0.1, "expect_alpha_kernel01_compensated": 0.1 * 9.0 / 9.1, "tol": 1e-5,
        "fd_camera": {"fx": 40.0, "fy": 40.0, "cx": 15.5, "cy": 15.5, "width": 32, "height": 32},
        "fd_sample_pixel_offset": 2, "fd_epsilon": 1e-3, "fd_rel_tol": 2e-2,
    },
    # splat_test.cpp:2369-2462 (test_forward_backward): background identity on a 48x32 grid.
    "background_identity": {
        "source": "photara/tests/splat_test.cpp:2369-2462",
        "mean": [0.5, 0.0, 2.0], "log_scales": [math.log(0.15)] * 3, "quaternion": [1.0, 0.0, 0.0, 0.0],
        "opacity_logit": 2.0, "sh": [[0.5, 0.25, 0.1]],
        "camera": {"fx": 40.0, "fy": 40.0, "cx": 23.5, "cy": 15.5, "width": 48, "height": 32},
        "background": [0.1, 0.3, 0.7], "tol": 5e-6,
    },
    # splat_test.cpp:881-946 (test_thin_splat_rgb_backward): two captured thin splats.
    "thin_splats": {
        "source": "photara/tests/splat_test.cpp:881-946",
        "camera": {"fx": 581.1552124, "fy": 581.1552124, "cx": 319.75, "cy": 179.75, "width": 640, "height": 360,
                   "world_to_camera": [.3124826849, -.3835008442, .8690694571, 0.0,
                                       .3029498458, .9073431492, .2914615273, 0.0,
                                       -.9003199339, .1722077578, .3997105360, 0.0,
                                       -.7320452929, -1.593179226, 2.917489529, 1.0],
                   "position": [-2.917735100, .8169972301, -1.550868511]},
        "means": [[-.2691904306, 1.285746336, -.4768541157], [.87730736, 1.1177711, .52421576]],
        "scales": [[.009814070538, .001056014560, 3.274591791e-6], [3.3371716e-8, .0019804370, .0049294555]],
        "rotations": [[.8961859345, .4393746555, -.06146703660, .004758699797],
                      [.74959874, .15781423, .4023349, .5013212]],
        "opacity_logit": -.6190475, "colors_precomp": [.3, .4, .5],
        "min_alpha_sum_sample0": 0.01, "rel_tol": 2e-6,
    },
    # splat_test.cpp:948-1039 (test_pinhole_geometry_finite_differences): two overlapping Gaussians.
    "pinhole_fd": {
        "source": "photara/tests/splat_test.cpp:948-1039",
        "camera": {"fx": 38.0, "fy": 36.0, "cx": 23.2, "cy": 15.1, "width": 48, "height": 32, "angle": 0.23},
        "parameters": [-.10, .04, 2.0, .05, -.02, 2.12,
                       math.log(.22), math.log(.16), math.log(.09), math.log(.20), math.log(.14), math.log(.08),
                       .94, .13, -.18, .21, .97, -.1, .17, .05,
                       1.7, 1.2, .4, .2, .3, .1, .5, .2],
        "pixel": [30, 16], "background": [.07, .11, .03], "h": 2e-4,
        "tol_abs": 0.002, "tol_rel": 0.025,
    },
    # splat_test.cpp:8123-8218 (test_speedy_reverse_compositing_oracle): 32 Gaussians, 47x31.
    "reverse_compositing": {
        "source": "photara/tests/splat_test.cpp:8123-8218",
        "count": 32, "camera": {"fx": 40.0, "fy": 40.0, "cx": 24.0, "cy": 16.0, "width": 47, "height": 31},
        "background": [0.1, 0.3, 0.7], "sh0": 0.2820947918, "tol": 3e-5,
    },
}


def main() -> None:
    OUT.write_text(json.dumps(fixtures, indent=1) + "\n")
    print(f"wrote {OUT}")


if __name__ == "__main__":
    main()
