"""FP64 NumPy oracle for the volume-transport hot path.

TEST INFRASTRUCTURE ONLY.  Nothing in the product (the HIP library, the
``ngt`` package) imports, links or executes anything in this directory.
Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` use it, and only as the checker (or as the timed CPU baseline).

The modules restate, from the reference source text, the semantics of

* ``sampling``  - trilinear / nearest / min / max reads, data gradients and
  scatter adjoints per boundary mode
  (phitest/render/cuda/src/sampling.hpp:371-704, sampling_v2.hpp:37-559,
  resample_grid.cu.cc:1080-1151);
* ``geometry``  - OpenGL camera / grid-transform matrices and the
  IDX<->NDC<->VS<->OS mappings (camera.py:51-100, transform.py:46-86,370-384,
  transformations.hpp:70-149, transformations_v2.hpp:70-164);
* ``resample``  - the SampleGridLut / SampleGridTransform ops and their
  gradients (resample_grid.cc:23-172, resample_grid.cu.cc:255-1696);
* ``blend``     - ReduceGridBlend and its gradient (blending.hpp,
  reduce_blend.cu.cc:34-80);
* ``raymarch``  - RaymarchGridTransform and its gradient
  (raymarch_grid.cu.cc:109-744);
* ``advect``    - semi-Lagrangian / MacCormack advection with the MC_SMOOTH
  revert clamp, centered and MAC-staggered velocity
  (data_structures.py:1081-1164,1634-1737, renderer.py:559-574,922-936).

Every function computes in float64.  Where the reference's gradient differs
from the exact adjoint (count / weight normalisation, the shifted raymarch
backward lattice, the ALPHA blend backward) both forms are provided and
selected by ``semantics="exact" | "reference"``.

How each function is pinned (closed forms, invariants, brute force,
finite differences, adjoint identities) is listed in tests/test_oracle_*.py.
"""

BORDER, CLAMP, WRAP = 0, 1, 2
NEAREST, LINEAR, MIN, MAX, QUADRATIC = 0, 1, 2, 3, 4

BOUNDARY = {"BORDER": BORDER, "CLAMP": CLAMP, "WRAP": WRAP}
FILTER = {"NEAREST": NEAREST, "LINEAR": LINEAR, "MIN": MIN, "MAX": MAX}
BLEND = {"BEER_LAMBERT": 0, "ALPHA": 2, "ALPHA_ADDITIVE": 3, "ADDITIVE": 4}
