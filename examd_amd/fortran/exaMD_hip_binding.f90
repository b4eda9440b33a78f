!! exaMD_hip_binding.f90 -- ISO_C_BINDING view of include/examd_hip.h.
!!
!! One interface per C entry point, with the same argument meaning.  Arrays
!! of positions, velocities and forces are Real(c_double) (3, n), the layout
!! of the reference's posn(1:3, 1:natms).

Module exaMD_hip_binding

  Use, Intrinsic :: iso_c_binding

  Implicit None

  Integer(c_int), Parameter :: EXAMD_OK = 0, EXAMD_ERR_ARG = 1,          &
       EXAMD_ERR_HIP = 2, EXAMD_ERR_NOMEM = 3, EXAMD_ERR_GRID = 4,        &
       EXAMD_ERR_TILE_OVERFLOW = 5, EXAMD_ERR_NO_DEVICE = 6,              &
       EXAMD_ERR_NONFINITE = 7, EXAMD_ERR_NOT_READY = 8,                  &
       EXAMD_ERR_FIXED_RANGE = 9, EXAMD_ERR_ZERO_DISTANCE = 10,           &
       EXAMD_ERR_TABLE_RANGE = 11, EXAMD_ERR_STATIC_LIMIT = 12,          &
       EXAMD_ERR_BIN_OVERFLOW = 13, EXAMD_ERR_OUT_OF_BOX = 109

  Integer(c_int), Parameter :: EXAMD_PREC_FP64 = 0, EXAMD_PREC_MIXED = 1

  Type, Bind(C) :: examd_params
     Integer(c_int64_t) :: natoms
     Real(c_double)     :: box(3)
     Real(c_double)     :: cutoff
     Real(c_double)     :: epsilon
     Real(c_double)     :: sigma
     Real(c_double)     :: mass
     Integer(c_int32_t) :: tiles(3)
     Integer(c_int32_t) :: device
     Integer(c_int32_t) :: block_threads
     Integer(c_int32_t) :: cap_slack_pct
     Integer(c_int32_t) :: run_groups
     Integer(c_int32_t) :: graph_chunk
     Integer(c_int32_t) :: reserved(2)
  End Type examd_params

  Type, Bind(C) :: examd_tune
     Integer(c_int32_t) :: tiles(3)
     Integer(c_int32_t) :: block_threads
     Integer(c_int32_t) :: run_groups
     Integer(c_int32_t) :: lane_queue
     Integer(c_int32_t) :: tile_edge
     Integer(c_int32_t) :: workgroups
     Integer(c_int32_t) :: balance
     Integer(c_int32_t) :: rebalance_every
     Real(c_double)     :: force_ms
  End Type examd_tune

  Type, Bind(C) :: examd_system
     Integer(c_int64_t) :: natoms
     Real(c_double)     :: box(3)
  End Type examd_system

  Type, Bind(C) :: examd_stats
     Integer(c_int64_t) :: step
     Real(c_double)     :: ke, pe, pe_shifted, etot
     Integer(c_int64_t) :: pairs_cand, pairs_cut
  End Type examd_stats

  Interface

     Integer(c_int) Function examd_hip_domain(box, block_length, nblocks_xyz, &
          nblocks) Bind(C, name="examd_hip_domain")
       Import :: c_int, c_double, c_int32_t, c_int64_t
       Real(c_double), Intent(in) :: box(3)
       Real(c_double), Value :: block_length
       Integer(c_int32_t), Intent(out) :: nblocks_xyz(3)
       Integer(c_int64_t), Intent(out) :: nblocks
     End Function examd_hip_domain

     Integer(c_int) Function examd_hip_create(params, handle) &
          Bind(C, name="examd_hip_create")
       Import :: c_int, c_ptr, examd_params
       Type(examd_params), Intent(in) :: params
       Type(c_ptr), Intent(out) :: handle
     End Function examd_hip_create

     Integer(c_int) Function examd_hip_destroy(handle) &
          Bind(C, name="examd_hip_destroy")
       Import :: c_int, c_ptr
       Type(c_ptr), Value :: handle
     End Function examd_hip_destroy

     Integer(c_int) Function examd_hip_set_state(handle, pos, vel, n) &
          Bind(C, name="examd_hip_set_state")
       Import :: c_int, c_ptr, c_double, c_int64_t
       Type(c_ptr), Value :: handle
       Real(c_double), Intent(in) :: pos(*)
       Type(c_ptr), Value :: vel
       Integer(c_int64_t), Value :: n
     End Function examd_hip_set_state

     Integer(c_int) Function examd_hip_set_species(handle, ntypes, type, n, &
          eps, sigma, cutoff) Bind(C, name="examd_hip_set_species")
       Import :: c_int, c_ptr, c_int32_t, c_int64_t
       Type(c_ptr), Value :: handle
       Integer(c_int32_t), Value :: ntypes
       Type(c_ptr), Value :: type
       Integer(c_int64_t), Value :: n
       Type(c_ptr), Value :: eps, sigma, cutoff
     End Function examd_hip_set_species

     Integer(c_int) Function examd_hip_set_pair_table(handle, npoints, r2_lo, &
          g, v) Bind(C, name="examd_hip_set_pair_table")
       Import :: c_int, c_ptr, c_int32_t, c_double
       Type(c_ptr), Value :: handle
       Integer(c_int32_t), Value :: npoints
       Real(c_double), Value :: r2_lo
       Type(c_ptr), Value :: g, v
     End Function examd_hip_set_pair_table

     Integer(c_int) Function examd_hip_get_state(handle, pos, vel, force, n) &
          Bind(C, name="examd_hip_get_state")
       Import :: c_int, c_ptr, c_int64_t
       Type(c_ptr), Value :: handle, pos, vel, force
       Integer(c_int64_t), Value :: n
     End Function examd_hip_get_state

     Integer(c_int) Function examd_hip_forces(handle, precision, forces, stats) &
          Bind(C, name="examd_hip_forces")
       Import :: c_int, c_ptr, c_double, examd_stats
       Type(c_ptr), Value :: handle
       Integer(c_int), Value :: precision
       Real(c_double), Intent(out) :: forces(*)
       Type(examd_stats), Intent(out) :: stats
     End Function examd_hip_forces

     Integer(c_int) Function examd_hip_regrow_count(handle, regrows) &
          Bind(C, name="examd_hip_regrow_count")
       Import :: c_int, c_ptr, c_int64_t
       Type(c_ptr), Value :: handle
       Integer(c_int64_t), Intent(out) :: regrows
     End Function examd_hip_regrow_count

     !! pos and forces are device addresses (3, natoms); stats may be
     !! c_null_ptr.  Allocates nothing (a static device array per GPU).
     Integer(c_int) Function examd_hip_forces_static(params, precision, pos, &
          forces, stats) Bind(C, name="examd_hip_forces_static")
       Import :: c_int, c_ptr, examd_params
       Type(examd_params), Intent(in) :: params
       Integer(c_int), Value :: precision
       Type(c_ptr), Value :: pos
       Type(c_ptr), Value :: forces
       Type(c_ptr), Value :: stats
     End Function examd_hip_forces_static

     Integer(c_int) Function examd_hip_run(handle, precision, nsteps, dt, &
          energy_every, out, out_cap, n_out) Bind(C, name="examd_hip_run")
       Import :: c_int, c_ptr, c_double, c_int64_t, examd_stats
       Type(c_ptr), Value :: handle
       Integer(c_int), Value :: precision
       Integer(c_int64_t), Value :: nsteps
       Real(c_double), Value :: dt
       Integer(c_int64_t), Value :: energy_every
       Type(examd_stats), Intent(out) :: out(*)
       Integer(c_int64_t), Value :: out_cap
       Integer(c_int64_t), Intent(out) :: n_out
     End Function examd_hip_run

     Integer(c_int) Function examd_hip_block_ids(handle, blocknum, blockxyz, n) &
          Bind(C, name="examd_hip_block_ids")
       Import :: c_int, c_ptr, c_int64_t, c_int32_t
       Type(c_ptr), Value :: handle
       Integer(c_int64_t), Intent(out) :: blocknum(*)
       Integer(c_int32_t), Intent(out) :: blockxyz(*)
       Integer(c_int64_t), Value :: n
     End Function examd_hip_block_ids

     Integer(c_int) Function examd_hip_home_pair_counts(handle, precision, &
          counts, nblocks) Bind(C, name="examd_hip_home_pair_counts")
       Import :: c_int, c_ptr, c_int64_t
       Type(c_ptr), Value :: handle
       Integer(c_int), Value :: precision
       Integer(c_int64_t), Intent(out) :: counts(*)
       Integer(c_int64_t), Value :: nblocks
     End Function examd_hip_home_pair_counts

     Integer(c_int) Function examd_hip_timing(handle, enable, force_ms, &
          force_launches, reset) Bind(C, name="examd_hip_timing")
       Import :: c_int, c_ptr, c_double, c_int64_t
       Type(c_ptr), Value :: handle
       Integer(c_int), Value :: enable
       Real(c_double), Intent(out) :: force_ms
       Integer(c_int64_t), Intent(out) :: force_launches
       Integer(c_int), Value :: reset
     End Function examd_hip_timing

     Integer(c_int) Function examd_hip_generate_fcc(nc, rho, temperature, &
          mass, perturb, seed, pos, vel, box) &
          Bind(C, name="examd_hip_generate_fcc")
       Import :: c_int, c_double, c_int32_t, c_int64_t
       Integer(c_int32_t), Intent(in) :: nc(3)
       Real(c_double), Value :: rho, temperature, mass, perturb
       Integer(c_int64_t), Value :: seed
       Real(c_double), Intent(out) :: pos(*), vel(*)
       Real(c_double), Intent(out) :: box(3)
     End Function examd_hip_generate_fcc

     Integer(c_int) Function examd_hip_phase_profile(handle, precision, &
          phase_cycles) Bind(C, name="examd_hip_phase_profile")
       Import :: c_int, c_ptr, c_double
       Type(c_ptr), Value :: handle
       Integer(c_int), Value :: precision
       Real(c_double), Intent(out) :: phase_cycles(5)
     End Function examd_hip_phase_profile

     Integer(c_int) Function examd_hip_autotune(handle, precision, trials, &
          best) Bind(C, name="examd_hip_autotune")
       Import :: c_int, c_ptr, examd_tune
       Type(c_ptr), Value :: handle
       Integer(c_int), Value :: precision, trials
       Type(examd_tune), Intent(out) :: best
     End Function examd_hip_autotune

     Integer(c_int) Function examd_hip_get_config(handle, precision, out) &
          Bind(C, name="examd_hip_get_config")
       Import :: c_int, c_ptr, examd_tune
       Type(c_ptr), Value :: handle
       Integer(c_int), Value :: precision
       Type(examd_tune), Intent(out) :: out
     End Function examd_hip_get_config

     Integer(c_int) Function examd_hip_set_config(handle, precision, cfg) &
          Bind(C, name="examd_hip_set_config")
       Import :: c_int, c_ptr, examd_tune
       Type(c_ptr), Value :: handle
       Integer(c_int), Value :: precision
       Type(examd_tune), Intent(in) :: cfg
     End Function examd_hip_set_config

     Integer(c_int) Function examd_hip_get_cuts(handle, precision, system, &
          cuts, cap, n_out) Bind(C, name="examd_hip_get_cuts")
       Import :: c_int, c_ptr, c_int32_t
       Type(c_ptr), Value :: handle
       Integer(c_int), Value :: precision
       Integer(c_int32_t), Value :: system
       Integer(c_int32_t), Intent(out) :: cuts(*)
       Integer(c_int32_t), Value :: cap
       Integer(c_int32_t), Intent(out) :: n_out
     End Function examd_hip_get_cuts

     Integer(c_int) Function examd_hip_batch_create(common, systems, nsys, &
          handle) Bind(C, name="examd_hip_batch_create")
       Import :: c_int, c_ptr, c_int32_t, examd_params, examd_system
       Type(examd_params), Intent(in) :: common
       Type(examd_system), Intent(in) :: systems(*)
       Integer(c_int32_t), Value :: nsys
       Type(c_ptr), Intent(out) :: handle
     End Function examd_hip_batch_create

     Integer(c_int) Function examd_hip_batch_info(handle, nsys, natoms, &
          nblocks) Bind(C, name="examd_hip_batch_info")
       Import :: c_int, c_ptr, c_int32_t, c_int64_t
       Type(c_ptr), Value :: handle
       Integer(c_int32_t), Intent(out) :: nsys
       Integer(c_int64_t), Intent(out) :: natoms, nblocks
     End Function examd_hip_batch_info

     Integer(c_int) Function examd_hip_set_graph_chunk(handle, chunk) &
          Bind(C, name="examd_hip_set_graph_chunk")
       Import :: c_int, c_ptr, c_int64_t
       Type(c_ptr), Value :: handle
       Integer(c_int64_t), Value :: chunk
     End Function examd_hip_set_graph_chunk

     Integer(c_int) Function examd_hip_generate_vacancy_fcc(n, rho, &
          temperature, mass, perturb, seed, pos, vel, box) &
          Bind(C, name="examd_hip_generate_vacancy_fcc")
       Import :: c_int, c_int64_t, c_double
       Integer(c_int64_t), Value :: n
       Real(c_double), Value :: rho, temperature, mass, perturb
       Integer(c_int64_t), Value :: seed
       Real(c_double), Intent(out) :: pos(*), vel(*), box(3)
     End Function examd_hip_generate_vacancy_fcc

     Integer(c_int) Function examd_hip_generate_slab(box, slab_lo, slab_hi, &
          rho_slab, rho_vapour, min_sep, temperature, mass, seed, pos, vel, &
          cap, n) Bind(C, name="examd_hip_generate_slab")
       Import :: c_int, c_int64_t, c_double, c_ptr
       Real(c_double), Intent(in) :: box(3)
       Real(c_double), Value :: slab_lo, slab_hi, rho_slab, rho_vapour, &
            min_sep, temperature, mass
       Integer(c_int64_t), Value :: seed
       Type(c_ptr), Value :: pos, vel
       Integer(c_int64_t), Value :: cap
       Integer(c_int64_t), Intent(out) :: n
     End Function examd_hip_generate_slab

     Type(c_ptr) Function examd_hip_strerror(status) &
          Bind(C, name="examd_hip_strerror")
       Import :: c_int, c_ptr
       Integer(c_int), Value :: status
     End Function examd_hip_strerror

  End Interface

Contains

  !! Text of a status code as a Fortran string.
  Function examd_status_text(status) Result(text)
    Integer(c_int), Intent(in) :: status
    Character(len=80) :: text
    Character(kind=c_char), Pointer :: chars(:)
    Type(c_ptr) :: p
    Integer :: i
    text = ' '
    p = examd_hip_strerror(status)
    If (.Not. c_associated(p)) Return
    Call c_f_pointer(p, chars, (/ 80 /))
    Do i = 1, 80
       If (chars(i) == c_null_char) Exit
       text(i:i) = chars(i)
    End Do
  End Function examd_status_text

End Module exaMD_hip_binding
