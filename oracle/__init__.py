"""Oracle for the nonbonded path — TEST INFRASTRUCTURE ONLY.

CPU restatements of the reference's algorithms, used by tests/, by
__graft_entry__.smoke() and by bench.py's cpu_baseline leg as the checker.
Nothing in the product (dynamics_amd/) imports, links or runs anything here.

  reference_path  exact f32 restatement of the reference's neighbor build
                  (src/neighbors.rs:180-281), SimBox::min_image / wrap
                  (src/barostat.rs:141-170) and setup_pairs flattening
                  (src/non_bonded.rs:541-677, src/gpu_interface.rs:242-411)
  fp64            FP64 evaluation of the same Hamiltonian: pair sums,
                  smooth PME of identical discretisation, plain Ewald sums,
                  1-4 and exclusion corrections
  cpu/            C++ restatement of the reference CPU force path
                  (calc_force_cpu / f_nonbonded_cpu, src/non_bonded.rs:258-412,
                  :950-1057) used as the timed CPU baseline, plus host replays
                  of the GPU kernels' fixed-point sums (bit-exact checks)

Which function is pinned by what is listed in each module's docstring.
"""
