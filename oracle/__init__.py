"""CPU restatement of Simple-RF's rendering hot path.

TEST INFRASTRUCTURE ONLY.  This package is the correctness oracle for the HIP
kernels in ``simple-rf_amd``.  Only ``tests/``, ``__graft_entry__.smoke()`` and
the ``cpu_baseline`` leg of ``bench.py`` may import it, and only as the checker
(or as the timed CPU baseline), never as a product code path.

Every function is a plain PyTorch restatement (dtype-generic, so it runs in
float32 for parity and float64 as the high-precision reference) of the
reference's own algorithm, and cites the reference file:line it follows.

Pinning status (what checks each function beyond retyping; see tests/test_oracle_*.py):
  nerf.linspace_like_torch        -- bit-exact vs torch.linspace (library)
  nerf.stratified_z_vals          -- closed form from explicit u, sortedness, range
  nerf.sample_pdf                 -- brute-force searchsorted count, special cases
                                     (u=0, u=1, flat pdf -> linear, spike, denom rule)
  nerf.fine_z_vals                -- multiset + sortedness vs concatenation
  nerf.positional_encoding        -- closed form per element, dims 63/27/21
  nerf.mlp_forward                -- vectors written by the reference's own MLP class
                                     (tests/golden/mlp_golden.npz: main, points-augmented
                                     21/42 and views-augmented, FP32 to 1e-6, FP64 to 1e-12);
                                     swapped concatenation orders miss them (test_oracle_golden)
  nerf.volume_rendering           -- closed forms (sigma=0, one opaque sample,
                                     constant sigma -> geometric weights), acc<=1,
                                     telescoping identity, gradcheck
  nerf.convert_depth_from_ndc     -- round trip with convert_depth_to_ndc
  nerf.get_rays / get_ndc_rays    -- pinhole closed form; NDC maps near-plane
                                     hit to z_ndc = -1 ... and the NDC ray is a line
  tensorf.vm_*                    -- torch.nn.functional.grid_sample (FP64) on the
                                     same factors, exact texel-centre values,
                                     rank-1 separable grids, gradcheck
  tensorf.alpha_mask_sample       -- grid_sample 3D (library) + texel-centre values
  tensorf.color_mlp_forward       -- vectors written by the reference's MlpFeaturesColorPredictor
                                     (same fixture: feature/view PE 0/0 and 2/2)
"""
