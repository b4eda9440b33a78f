!! examd_md.f90 -- Fortran driver: the reference driver's setup sequence and
!! force check (exaMD.f90:46-366), then a velocity-Verlet MD loop on the GPU.
!!
!! Options:
!!   -atoms N        FCC lattice of N = 4 nc^3 atoms (default 32000)
!!   -readfile       read exaMD_particles.dat (zero initial velocities)
!!   -writefile      write exaMD_particles.dat
!!   -overlap        use exaMD_forces_midpoint_overlap
!!   -cutoff rc      cutoff (default 2.5)
!!   -rho r          FCC number density (default 0.8442)
!!   -temp T         initial temperature (default 0.72)
!!   -steps K        MD steps after the force check (default 0)
!!   -dt h           time step (default 0.005)
!!   -precision p    fp64 | mixed (default fp64)
!!   -seed s         generator seed (default 1)
!!   -restart file   start from a checkpoint (positions, velocities, box, step)
!!   -checkpoint file write the state after the MD steps as a checkpoint
!!   -rebalance K    cell blocks at equal shares of the atoms, re-partitioned
!!                   from the current positions at least every K MD steps
!!   -morse          the user function is a Morse pair (De 0.8, a 2.0,
!!                   r0 1.1), which the library tabulates for the GPU
!!   -evals N        force evaluations (default 1); the forces accumulate
!!                   in grid%forces over the calls, as the reference's
!!                   routine adds into them, and each call's time is printed
!!   -printforces    per-block sums of squared forces after the check, in the
!!                   reference's PRINTALLFORCES format (exaMD.f90:340-357)
!!   -dumpforces f   per-atom positions and forces in grid order (blocks of
!!                   locblocks, then file order), 17 significant digits
!!   -binary         two species (Kob-Andersen): every fifth atom in input
!!                   order is species 2; pair rows AA (1, 1, rc 2.5),
!!                   AB (1.5, 0.8, 2.0), BB (0.5, 0.88, 2.2) as (eps, sigma,
!!                   cutoff), set through grid%ltype, kktype, ltpvdw, prmvdw

Program examd_md

  Use, Intrinsic :: iso_c_binding
  Use exaMD_mod
  Use exaMD_partitioning
  Use exaMD_comms
  Use exaMD_forces_mod
  Use exaMD_input_formats
  Use exaMD_utils_mod
  Use user_func_mod
  Use exaMD_hip_binding

  Implicit None

  Real(kind=PRC) :: cutoff, block_length, rho, temp, dt, ftot(3), ftot2(3)
  Integer :: natms_tot, nblocks, nblocks_xyz(3), pid, noprocs, nc, i, j, k, b
  Integer :: argc, ia, nnan, nevals, ev
  Logical :: printforces
  Character(len=256) :: dump_file
  Integer(c_int64_t) :: nsteps, seed, nstats, s
  Integer, Allocatable :: ntot(:), first_particle(:), next_particle(:)
  Integer, Allocatable :: part(:,:,:)
  Real(kind=PRC), Allocatable :: particle_positions(:,:), velocities(:,:)
  Real(c_double), Allocatable :: vflat(:,:), pos_out(:,:), vel_out(:,:)
  Type(examd_stats), Allocatable :: stats(:)
  Character(len=256) :: arg, restart_file, checkpoint_file
  Integer(8) :: step0
  Logical :: inputcase, overlapcase, writeout, binary, morse
  Integer :: m, p
  Integer(8) :: c0, c1, crate
  Real(c_double) :: force_ms, e0, drift

  Call exaMD_Init(pid, noprocs, 0)

  natms_tot = 32000
  cutoff = 2.5_PRC
  rho = 0.8442_PRC
  temp = 0.72_PRC
  dt = 0.005_PRC
  nsteps = 0
  seed = 1
  inputcase = .False.
  overlapcase = .False.
  writeout = .False.
  binary = .False.
  morse = .False.
  restart_file = ''
  checkpoint_file = ''
  dump_file = ''
  nevals = 1
  printforces = .False.
  step0 = 0

  argc = command_argument_count()
  ia = 1
  Do While (ia <= argc)
     Call get_command_argument(ia, arg)
     Select Case (Trim(arg))
     Case ('-overlap')
        overlapcase = .True.
     Case ('-readfile')
        inputcase = .True.
     Case ('-writefile')
        writeout = .True.
     Case ('-binary')
        binary = .True.
     Case ('-morse')
        morse = .True.
     Case ('-printforces')
        printforces = .True.
     Case ('-dumpforces')
        ia = ia + 1
        Call get_command_argument(ia, dump_file)
     Case ('-evals')
        ia = ia + 1
        Call get_command_argument(ia, arg)
        Read (arg, *) nevals
     Case ('-atoms')
        ia = ia + 1
        Call get_command_argument(ia, arg)
        Read (arg, *) natms_tot
     Case ('-cutoff')
        ia = ia + 1
        Call get_command_argument(ia, arg)
        Read (arg, *) cutoff
     Case ('-rho')
        ia = ia + 1
        Call get_command_argument(ia, arg)
        Read (arg, *) rho
     Case ('-temp')
        ia = ia + 1
        Call get_command_argument(ia, arg)
        Read (arg, *) temp
     Case ('-steps')
        ia = ia + 1
        Call get_command_argument(ia, arg)
        Read (arg, *) nsteps
     Case ('-dt')
        ia = ia + 1
        Call get_command_argument(ia, arg)
        Read (arg, *) dt
     Case ('-seed')
        ia = ia + 1
        Call get_command_argument(ia, arg)
        Read (arg, *) seed
     Case ('-precision')
        ia = ia + 1
        Call get_command_argument(ia, arg)
        If (Trim(arg) == 'mixed') exaMD_precision = EXAMD_PREC_MIXED
     Case ('-restart')
        ia = ia + 1
        Call get_command_argument(ia, restart_file)
     Case ('-checkpoint')
        ia = ia + 1
        Call get_command_argument(ia, checkpoint_file)
     Case ('-rebalance')
        ia = ia + 1
        Call get_command_argument(ia, arg)
        Read (arg, *) exaMD_rebalance_every
     End Select
     ia = ia + 1
  End Do
  block_length = half * cutoff

  ! positions (and velocities) before the domain: the FCC box follows N, rho
  If (Len_Trim(restart_file) > 0) Then
     Call exaMD_read_checkpoint(Trim(restart_file), particle_positions, &
          velocities, natms_tot, step0)
  Else If (inputcase) Then
     Call exaMD_read_params(natms_tot)
  Else
     nc = Nint((Real(natms_tot, PRC) / 4.0_PRC)**(1.0_PRC / 3.0_PRC))
     If (4 * nc**3 /= natms_tot) Then
        Write (*, '(A,I0,A)') ' -atoms ', natms_tot, ' is not 4 nc^3'
        Call exaMD_abort(103)
     End If
  End If
  If (Len_Trim(restart_file) == 0) Then
     Allocate(particle_positions(3, natms_tot), velocities(3, natms_tot))
     velocities = zero
     If (.Not. inputcase) Call exaMD_generate_fcc(nc, rho, temp, seed, &
          particle_positions, velocities)
  End If

  Call exaMD_domain_create(block_length, nblocks, nblocks_xyz)

  Write (*, *) '-----------------------------------------------------------------------'
  Write (*, *) 'exaMD on MI355X (gfx950): midpoint-method Lennard-Jones'
  Write (*, '(A,I0)') ' Number of atoms: ', natms_tot
  Write (*, '(A,3F10.4)') ' Simulation box size:', globalbox
  Write (*, '(A,F6.2)') ' Cut-off distance:', cutoff
  Write (*, '(A,3(1X,I0))') ' Number of blocks in each dimension:', nblocks_xyz
  If (exaMD_precision == EXAMD_PREC_MIXED) Then
     Write (*, *) 'Precision: mixed (FP32 pairs, 64-bit fixed-point sums)'
  Else
     Write (*, *) 'Precision: FP64'
  End If
  Write (*, *) '-----------------------------------------------------------------------'

  Allocate(ntot(nblocks), first_particle(nblocks), next_particle(natms_tot))
  If (inputcase .And. Len_Trim(restart_file) == 0) Then
     Call exaMD_read_particles(particle_positions, natms_tot, ntot, &
          first_particle, next_particle, nblocks, nblocks_xyz, block_length)
  Else
     Call exaMD_chain_blocks(particle_positions, natms_tot, ntot, &
          first_particle, next_particle, nblocks, nblocks_xyz, block_length)
  End If
  If (writeout) Call exaMD_write_particles(particle_positions, natms_tot)

  Allocate(part(nblocks_xyz(1), nblocks_xyz(2), nblocks_xyz(3)))
  Call exaMD_partition_domain(nblocks, nblocks_xyz, ntot, pid, noprocs, part)
  Call exaMD_setup_partitioned_grid(part, nblocks, nblocks_xyz, pid)
  Call exaMD_partition_work(particle_positions, ntot, first_particle, &
       next_particle, nblocks, nblocks_xyz, pid, noprocs)

  If (binary) Then
     ! species of every atom through its block chain (input order), then
     ! the species-pair rows
     Do b = 1, nblocks
        i = Mod(b - 1, nblocks_xyz(1)) + 1
        j = Mod((b - 1) / nblocks_xyz(1), nblocks_xyz(2)) + 1
        k = (b - 1) / (nblocks_xyz(1) * nblocks_xyz(2)) + 1
        p = first_particle(b)
        Do m = 1, ntot(b)
           grid(i, j, k)%ltype(m) = 1
           If (Mod(p, 5) == 0) grid(i, j, k)%ltype(m) = 2
           p = next_particle(p)
        End Do
     End Do
     Allocate(kktype(2, 2), ltpvdw(3), prmvdw(3, 5))
     kktype = Reshape((/ 1, 2, 2, 3 /), (/ 2, 2 /))
     ltpvdw = 2
     prmvdw(1, 1:5) = (/ one, one, zero, zero, cutoff /)
     prmvdw(2, 1:5) = (/ 1.5_PRC, 0.8_PRC, zero, zero, Min(cutoff, 2.0_PRC) /)
     prmvdw(3, 1:5) = (/ half, 0.88_PRC, zero, zero, Min(cutoff, 2.2_PRC) /)
     Write (*, *) 'Species: 2 (every fifth atom is B; AA, AB, BB rows)'
  Else
     Allocate(kktype(1, 1), ltpvdw(1), prmvdw(1, 5))
     kktype = 1
     ltpvdw = 2
     prmvdw(1, 1:5) = (/ one, one, zero, zero, cutoff /)
  End If
  Call exaMD_setup_midpoint_interactions()

  ! force evaluations, timed like the reference (exaMD.f90:241-253); the
  ! calls after the first find the positions unchanged
  Do ev = 1, Max(1, nevals)
     Call system_clock(c0, crate)
     Call exaMD_start_comms_posn()
     If (morse) Then
        Call exaMD_forces_midpoint_all(morse_pair, cutoff, noprocs)
     Else If (overlapcase) Then
        Call exaMD_forces_midpoint_overlap(lennardjones_12, cutoff, pid, noprocs)
     Else
        Call exaMD_forces_midpoint_all(lennardjones_12, cutoff, noprocs)
     End If
     Call exaMD_wait_comms_forces(noprocs)
     Call system_clock(c1)
     If (nevals > 1) Write (*, '(A,I0,A,F10.4,A,I0)') '  Evaluation ', ev, &
          ' time (secs):', Real(c1 - c0, 8) / Real(crate, 8), &
          ', positions not resent: ', exaMD_positions_resent_skipped
  End Do

  ! checks of exaMD.f90:269-317
  nnan = 0
  ftot = zero
  ftot2 = zero
  Do b = 1, nblocks_loc
     i = locblocks(1, b)
     j = locblocks(2, b)
     k = locblocks(3, b)
     If (grid(i, j, k)%natms == 0) Cycle
     nnan = nnan + Count(grid(i, j, k)%forces /= grid(i, j, k)%forces)
     ftot = ftot + Sum(grid(i, j, k)%forces, dim=2)
     ftot2 = ftot2 + Sum(grid(i, j, k)%forces**2, dim=2)
  End Do
  Write (*, *) 'Checks on force calculations'
  If (nnan > 0) Write (*, '(A,I0)') ' NaN force components: ', nnan
  If (Abs(ftot(1)) > 1e-10_PRC) Write (*, *) 'WARNING! Suspected error'
  Write (*, '(A,3E10.1)') '  SUM(fxx, fyy, fzz):', ftot
  Write (*, '(A,3E10.1)') '  SUM(fxx^2, fyy^2, fzz^2):', ftot2
  Write (*, '(A,3ES25.16)') '  SUM(fxx, fyy, fzz) full:', ftot
  Write (*, '(A,3ES25.16)') '  SUM(fxx^2, fyy^2, fzz^2) full:', ftot2
  Write (*, '(A,I0,A,I0)') '  Pairs: candidate ', exaMD_last_stats%pairs_cand, &
       ', in cutoff ', exaMD_last_stats%pairs_cut
  Write (*, '(A,F16.6)') '  Potential energy:', exaMD_last_stats%pe
  Write (*, '(A,F10.4)') '  Computation time (secs):', &
       Real(c1 - c0, 8) / Real(crate, 8)

  ! exaMD.f90:340-357 (PRINTALLFORCES): i, j, k and the block's sums of
  ! squared force components, for every block with a non-zero sum
  If (printforces) Then
     Do b = 1, nblocks_loc
        i = locblocks(1, b)
        j = locblocks(2, b)
        k = locblocks(3, b)
        If (grid(i, j, k)%natms == 0) Cycle
        ftot2 = Sum(grid(i, j, k)%forces**2, dim=2)
        If (Sum(ftot2) > zero) Write (*, '(3I4.3,3(X,E20.12))') i, j, k, ftot2
     End Do
  End If
  If (Len_Trim(dump_file) > 0) Then
     Open(61, file=Trim(dump_file), status='replace')
     Do b = 1, nblocks_loc
        i = locblocks(1, b)
        j = locblocks(2, b)
        k = locblocks(3, b)
        Do m = 1, grid(i, j, k)%natms
           Write (61, '(6ES25.16)') grid(i, j, k)%posn(:, m), &
                grid(i, j, k)%forces(:, m)
        End Do
     End Do
     Close(61)
  End If

  If (nsteps > 0) Then
     ! velocities in the grid's atom order (blocks, then input order)
     Allocate(vflat(3, natms_tot), pos_out(3, natms_tot), vel_out(3, natms_tot))
     s = 0
     Do b = 1, nblocks
        j = first_particle(b)
        Do i = 1, ntot(b)
           s = s + 1
           vflat(:, s) = velocities(:, j)
           j = next_particle(j)
        End Do
     End Do
     Allocate(stats(Max(1_c_int64_t, nsteps)))
     Call system_clock(c0, crate)
     Call exaMD_md_run(cutoff, vflat, nsteps, dt, Max(1_c_int64_t, nsteps / 10), &
          stats, nstats, pos_out, vel_out, force_ms)
     Call system_clock(c1)
     Write (*, *) 'MD: step, KE, PE, E(shifted)'
     Do s = 1, nstats
        Write (*, '(I8,3F18.6)') stats(s)%step, stats(s)%ke, stats(s)%pe, &
             stats(s)%ke + stats(s)%pe_shifted
     End Do
     e0 = stats(1)%ke + stats(1)%pe_shifted
     drift = (stats(nstats)%ke + stats(nstats)%pe_shifted - e0) / Abs(e0)
     Write (*, '(A,ES12.4)') ' Relative shifted-energy change:', drift
     Write (*, '(A,ES12.4)') ' atom-steps/s:', Real(natms_tot, 8) * &
          Real(nsteps, 8) / (Real(c1 - c0, 8) / Real(crate, 8))
     Write (*, '(A,F10.4)') ' Force kernel ms/step:', force_ms / Real(nsteps, 8)
     If (Len_Trim(checkpoint_file) > 0) Then
        ! the run returns atoms in the grid's order; put them back in input
        ! order for the checkpoint
        s = 0
        Do b = 1, nblocks
           j = first_particle(b)
           Do i = 1, ntot(b)
              s = s + 1
              particle_positions(:, j) = pos_out(:, s)
              velocities(:, j) = vel_out(:, s)
              j = next_particle(j)
           End Do
        End Do
        Call exaMD_write_checkpoint(Trim(checkpoint_file), particle_positions, &
             velocities, natms_tot, step0 + nsteps)
        Write (*, '(A,A)') ' Checkpoint written: ', Trim(checkpoint_file)
     End If
  End If

  Deallocate(kktype, ltpvdw, prmvdw)
  Call exaMD_Finalize('Done')

Contains

  !! A user function in the form of user_func.f90:40-75 that is not the
  !! Lennard-Jones pair: V = De (1 - exp(-a (r - r0)))^2 - De up to the
  !! cutoff prmvdw(1,5).
  Subroutine morse_pair(dx, dy, dz, d2, f12, e12, retval)
    Real(kind=PRC), Intent(in) :: dx, dy, dz, d2
    Real(kind=PRC), Intent(out) :: f12(3), e12
    Integer, Intent(out) :: retval
    Real(kind=PRC), Parameter :: de = 0.8_PRC, a = 2.0_PRC, r0 = 1.1_PRC
    Real(kind=PRC) :: r, x, g
    retval = 0
    If (d2 > prmvdw(1, 5)**2) Then
       retval = 1
       Return
    End If
    r = Sqrt(d2)
    x = Exp(-a * (r - r0))
    g = -two * de * a * (one - x) * x / r
    f12(1) = g * dx
    f12(2) = g * dy
    f12(3) = g * dz
    e12 = de * (one - x)**2 - de
  End Subroutine morse_pair

End Program examd_md
