!! exaMD_modules.f90 -- the exaMD library call surface over the gfx950 path.
!!
!! Modules and public routines carry the reference's names and argument
!! lists (exaMD v2.0, documented in DOC/exaMD.tex:890-1061), so a driver
!! written against exaMD links against this layer instead.  State lives in
!! the same place as in the reference: module exaMD_mod's grid(:,:,:) of
!! per-block posn/forces arrays.  The force routines hand the positions to the
!! HIP library through exaMD_hip_binding and add the forces back into
!! grid%forces (accumulation, as SRC/exaMD_forces.f90:538-542 does).
!!
!! One process drives one GPU: pid = 0, noprocs = 1.  The halo exchange of
!! SRC/exaMD_comms.f90 happens on the device (ghost gather per cell block), so
!! the comms routines only keep their call sites.

Module exaMD_constants_mod
  Implicit None
#ifndef EXAMD_PRC
#define EXAMD_PRC 8
#endif
  !! Working precision (the reference's -DPRC), as a named kind.
  Integer, Parameter :: PRC = EXAMD_PRC
  Real(kind=PRC), Parameter :: zero = 0.0_PRC, one = 1.0_PRC, two = 2.0_PRC, &
       four = 4.0_PRC, half = 0.5_PRC, twentyfour = 24.0_PRC
End Module exaMD_constants_mod


Module exaMD_mod
  !! Global state (SRC/exaMD_module.f90:17-84): the block grid, block lists
  !! and the 63-row midpoint table.
  Use exaMD_constants_mod
  Implicit None

  Integer :: COMM_EXAMD = 0

  Type pframe_t
     Integer :: host = 0
     Integer, Allocatable :: neighrefs(:,:)
     Integer :: natms = 0
     Real(kind=PRC), Allocatable :: posn(:,:)
     Real(kind=PRC), Allocatable :: forces(:,:)
     Integer, Allocatable :: ltype(:)
     Integer :: partype = 0
  End Type pframe_t

  Type(pframe_t), Allocatable, Target :: grid(:,:,:)
  Real(kind=PRC) :: globalbox(3), globalboxi(3)
  Integer :: nblocks_loc = 0, n_interior = 0, n_bdry = 0, n_exterior = 0
  Integer, Allocatable :: locblocks(:,:), interior_blks(:,:)
  Integer :: n_interactions = 0
  Integer, Allocatable :: interactions(:,:)

Contains

  !! exaMD_identify_block (SRC/exaMD_module.f90:101-125): block of a
  !! position, 1-based, last block enlarged; error 109 out of range.
  Subroutine exaMD_identify_block(posxyz, blockxyz, blocknum, nblocks, &
       nblocks_xyz, nblocks_xyplane, block_length)
    Real(kind=PRC), Intent(in) :: posxyz(3), block_length
    Integer, Intent(in) :: nblocks, nblocks_xyz(3), nblocks_xyplane
    Integer, Intent(out) :: blockxyz(3), blocknum
    Integer :: a
    Do a = 1, 3
       blockxyz(a) = Min(Int(posxyz(a) / block_length + one), nblocks_xyz(a))
    End Do
    blocknum = blockxyz(1) + (blockxyz(2) - 1) * nblocks_xyz(1) &
         + (blockxyz(3) - 1) * nblocks_xyplane
    If (blocknum < 1 .Or. blocknum > nblocks) Then
       Write (*, '(A,I0,A,3(1X,I0))') ' Block number out of range: ', &
            blocknum, ' block', blockxyz
       Call exaMD_abort(109)
    End If
  End Subroutine exaMD_identify_block

  !! The 63 (home-relative stencil index) pairs of
  !! SRC/exaMD_module.f90:276-346, stencil index 14 being the home block.
  Subroutine exaMD_setup_midpoint_interactions()
    Integer, Parameter :: tbl(126) = (/ &
         14,14, 14,15, 14,16, 14,17, 14,18, 14,19, 14,20, 14,21, 14,22, &
         14,23, 14,24, 14,25, 14,26, 14,27, 10,18, 10,27, 11,17, 11,18, &
         11,25, 11,26, 11,27, 12,16, 12,17, 12,25, 13,15, 13,18, 13,24, &
         13,27, 15,16, 15,19, 15,22, 15,25, 16,21, 16,24, 17,19, 17,20, &
         17,21, 18,19, 19,5,  19,9,  20,5,  20,7,  20,8,  20,9,  21,5,  &
         21,7,  22,5,  22,6,  22,9,  23,5,  24,5,  24,4,  24,7,  25,5,  &
         25,3,  25,6,  26,5,  26,2,  26,3,  27,1,  27,2,  27,4,  27,5 /)
    n_interactions = 63
    If (Allocated(interactions)) Deallocate(interactions)
    Allocate(interactions(8, n_interactions))
    interactions = 0
    interactions(1:2, :) = Reshape(tbl, (/ 2, 63 /))
  End Subroutine exaMD_setup_midpoint_interactions

  Subroutine exaMD_abort(code)
    Integer, Intent(in) :: code
    Write (*, '(A,I0)') ' exaMD: aborting with code ', code
    Error Stop code
  End Subroutine exaMD_abort

End Module exaMD_mod


Module user_func_mod
  !! Interaction tables and the LJ pair of the example user function
  !! (user_func.f90:16-75), here with the pair energy filled in.
  Use exaMD_mod
  Implicit None
  Integer, Allocatable :: kktype(:,:)
  Integer, Allocatable :: ltpvdw(:)
  Real(kind=PRC), Allocatable :: prmvdw(:,:)
Contains
  Subroutine lennardjones_12(dx, dy, dz, d2, f12, e12, retval)
    Real(kind=PRC), Intent(in) :: dx, dy, dz, d2
    Real(kind=PRC), Intent(out) :: f12(3), e12
    Integer, Intent(out) :: retval
    Real(kind=PRC) :: s, s6, g
    retval = 0
    If (d2 > prmvdw(1, 5)**2) Then
       retval = 1
       Return
    End If
    s = prmvdw(1, 2)**2 / d2
    s6 = s * s * s
    g = -twentyfour * prmvdw(1, 1) * (two * s6 * s6 - s6) / d2
    f12(1) = g * dx
    f12(2) = g * dy
    f12(3) = g * dz
    e12 = four * prmvdw(1, 1) * (s6 * s6 - s6)
  End Subroutine lennardjones_12
End Module user_func_mod


Module exaMD_input_formats
  !! Particle file exaMD_particles.dat (SRC/exaMD_input_formats.f90:112-236)
  !! and a seeded FCC generator in place of the unseeded random one.
  Use exaMD_mod
  Implicit None
Contains
  Subroutine exaMD_read_params(natms_tot)
    Integer, Intent(inout) :: natms_tot
    Open(50, file='./exaMD_particles.dat', status='old')
    Read(50, *) natms_tot
    Read(50, *) globalbox
  End Subroutine exaMD_read_params

  !! Reads positions (after exaMD_read_params) and chains the atoms of each
  !! block in file order.
  Subroutine exaMD_read_particles(particle_positions, natms_tot, ntot, &
       first_particle, next_particle, nblocks, nblocks_xyz, block_length)
    Real(kind=PRC), Intent(inout), Allocatable :: particle_positions(:,:)
    Integer, Intent(inout), Allocatable :: ntot(:), first_particle(:), &
         next_particle(:)
    Integer, Intent(in) :: natms_tot, nblocks, nblocks_xyz(3)
    Real(kind=PRC), Intent(in) :: block_length
    Integer :: i
    Do i = 1, natms_tot
       Read(50, *) particle_positions(:, i)
    End Do
    Close(50)
    Call exaMD_chain_blocks(particle_positions, natms_tot, ntot, &
         first_particle, next_particle, nblocks, nblocks_xyz, block_length)
  End Subroutine exaMD_read_particles

  Subroutine exaMD_write_particles(particle_positions, natms_tot)
    Real(kind=PRC), Intent(in), Allocatable :: particle_positions(:,:)
    Integer, Intent(in) :: natms_tot
    Integer :: i
    Open(50, file='exaMD_particles.dat', status='replace')
    Write(50, *) natms_tot
    Write(50, *) globalbox
    Do i = 1, natms_tot
       Write(50, *) particle_positions(1:3, i)
    End Do
    Close(50)
  End Subroutine exaMD_write_particles

  !! Per-block linked lists in input order (first_particle/next_particle),
  !! as the reference builds them while reading.
  Subroutine exaMD_chain_blocks(particle_positions, natms_tot, ntot, &
       first_particle, next_particle, nblocks, nblocks_xyz, block_length)
    Real(kind=PRC), Intent(in), Allocatable :: particle_positions(:,:)
    Integer, Intent(inout), Allocatable :: ntot(:), first_particle(:), &
         next_particle(:)
    Integer, Intent(in) :: natms_tot, nblocks, nblocks_xyz(3)
    Real(kind=PRC), Intent(in) :: block_length
    Integer, Allocatable :: last(:)
    Integer :: i, bxyz(3), num
    Allocate(last(nblocks))
    ntot = 0
    first_particle = 0
    next_particle = 0
    last = 0
    Do i = 1, natms_tot
       Call exaMD_identify_block(particle_positions(:, i), bxyz, num, &
            nblocks, nblocks_xyz, nblocks_xyz(1) * nblocks_xyz(2), block_length)
       ntot(num) = ntot(num) + 1
       If (ntot(num) == 1) Then
          first_particle(num) = i
       Else
          next_particle(last(num)) = i
       End If
       last(num) = i
    End Do
    Deallocate(last)
  End Subroutine exaMD_chain_blocks

  !! Seeded FCC lattice of nc^3 cells at density rho with Maxwell-Boltzmann
  !! velocities at temperature (examd_hip_generate_fcc); sets globalbox.
  Subroutine exaMD_generate_fcc(nc, rho, temperature, seed, positions, &
       velocities)
    Use exaMD_hip_binding
    Integer, Intent(in) :: nc
    Real(kind=PRC), Intent(in) :: rho, temperature
    Integer(c_int64_t), Intent(in) :: seed
    Real(kind=PRC), Intent(out) :: positions(:,:), velocities(:,:)
    Real(c_double), Allocatable :: p(:,:), v(:,:)
    Real(c_double) :: box(3)
    Integer(c_int32_t) :: ncs(3)
    Integer(c_int) :: st
    Allocate(p(3, Size(positions, 2)), v(3, Size(positions, 2)))
    ncs = nc
    st = examd_hip_generate_fcc(ncs, Real(rho, c_double), &
         Real(temperature, c_double), 1.0_c_double, 0.0_c_double, seed, &
         p, v, box)
    If (st /= EXAMD_OK) Call exaMD_abort(103)
    positions = p
    velocities = v
    globalbox = box
  End Subroutine exaMD_generate_fcc
  !! Checkpoint of an MD state (no reference counterpart: the reference
  !! file holds positions only): line 1 the atom count and the step, line 2
  !! the box, then one line per atom with x y z vx vy vz, written with 17
  !! significant digits so that a restart continues bit for bit.
  Subroutine exaMD_write_checkpoint(filename, pos, vel, natms, step)
    Character(len=*), Intent(in) :: filename
    Real(kind=PRC), Intent(in) :: pos(:,:), vel(:,:)
    Integer, Intent(in) :: natms
    Integer(8), Intent(in) :: step
    Integer :: i
    Open(51, file=filename, status='replace')
    Write(51, '(I0,1X,I0)') natms, step
    Write(51, '(3ES25.16E3)') globalbox
    Do i = 1, natms
       Write(51, '(6ES25.16E3)') pos(1:3, i), vel(1:3, i)
    End Do
    Close(51)
  End Subroutine exaMD_write_checkpoint

  !! Reads a checkpoint written by exaMD_write_checkpoint; sets globalbox.
  Subroutine exaMD_read_checkpoint(filename, pos, vel, natms, step)
    Character(len=*), Intent(in) :: filename
    Real(kind=PRC), Allocatable, Intent(inout) :: pos(:,:), vel(:,:)
    Integer, Intent(out) :: natms
    Integer(8), Intent(out) :: step
    Integer :: i
    Open(51, file=filename, status='old')
    Read(51, *) natms, step
    Read(51, *) globalbox
    If (Allocated(pos)) Deallocate(pos)
    If (Allocated(vel)) Deallocate(vel)
    Allocate(pos(3, natms), vel(3, natms))
    Do i = 1, natms
       Read(51, *) pos(1:3, i), vel(1:3, i)
    End Do
    Close(51)
  End Subroutine exaMD_read_checkpoint

End Module exaMD_input_formats


Module exaMD_partitioning
  !! One GPU holds every block: the METIS partition of
  !! SRC/exaMD_partitioning.f90:65-455 becomes the device's cell-block map
  !! (chosen inside the HIP library), so all blocks are interior here.
  Use exaMD_mod
  Implicit None
Contains
  Subroutine exaMD_partition_domain(nblocks, nblocks_xyz, ntot, pid, noprocs, &
       vtxdist)
    Integer, Intent(in) :: nblocks, nblocks_xyz(3), pid, noprocs
    Integer, Intent(in), Allocatable :: ntot(:)
    Integer, Intent(inout), Allocatable :: vtxdist(:,:,:)
    vtxdist = 0
  End Subroutine exaMD_partition_domain

  !! Interior-block list and periodic 27-neighbour references (ii fastest,
  !! SRC/exaMD_partitioning.f90:783-798).
  Subroutine exaMD_setup_partitioned_grid(part, nblocks, nblocks_xyz, pid)
    Integer, Intent(in) :: nblocks, nblocks_xyz(3), pid
    Integer, Allocatable, Intent(in) :: part(:,:,:)
    Integer :: i, j, k, ii, jj, kk, cnt, b
    If (Allocated(locblocks)) Deallocate(locblocks)
    If (Allocated(interior_blks)) Deallocate(interior_blks)
    Allocate(locblocks(3, nblocks), interior_blks(3, nblocks))
    b = 0
    Do k = 1, nblocks_xyz(3)
       Do j = 1, nblocks_xyz(2)
          Do i = 1, nblocks_xyz(1)
             b = b + 1
             locblocks(:, b) = (/ i, j, k /)
             grid(i, j, k)%host = part(i, j, k)
             grid(i, j, k)%partype = 1
             If (Allocated(grid(i, j, k)%neighrefs)) &
                  Deallocate(grid(i, j, k)%neighrefs)
             Allocate(grid(i, j, k)%neighrefs(3, 27))
             cnt = 0
             Do kk = -1, 1
                Do jj = -1, 1
                   Do ii = -1, 1
                      cnt = cnt + 1
                      grid(i, j, k)%neighrefs(1, cnt) = &
                           Modulo(i + ii - 1, nblocks_xyz(1)) + 1
                      grid(i, j, k)%neighrefs(2, cnt) = &
                           Modulo(j + jj - 1, nblocks_xyz(2)) + 1
                      grid(i, j, k)%neighrefs(3, cnt) = &
                           Modulo(k + kk - 1, nblocks_xyz(3)) + 1
                   End Do
                End Do
             End Do
          End Do
       End Do
    End Do
    interior_blks = locblocks
    nblocks_loc = nblocks
    n_interior = nblocks
    n_bdry = 0
    n_exterior = 0
  End Subroutine exaMD_setup_partitioned_grid

  !! Copy every block's atoms (chained in input order) into grid%posn and
  !! zero grid%forces (SRC/exaMD_partitioning.f90:932-968).
  Subroutine exaMD_partition_work(particle_positions, ntot, first_particle, &
       next_particle, nblocks, nblocks_xyz, pid, noprocs)
    Integer, Intent(in) :: nblocks, nblocks_xyz(3), pid, noprocs
    Integer, Allocatable, Intent(in) :: ntot(:), first_particle(:), &
         next_particle(:)
    Real(kind=PRC), Allocatable, Intent(in) :: particle_positions(:,:)
    Integer :: i, j, k, b, m, p
    Do k = 1, nblocks_xyz(3)
       Do j = 1, nblocks_xyz(2)
          Do i = 1, nblocks_xyz(1)
             b = i + (j - 1) * nblocks_xyz(1) &
                  + (k - 1) * nblocks_xyz(1) * nblocks_xyz(2)
             Associate(blk => grid(i, j, k))
               blk%natms = ntot(b)
               If (Allocated(blk%posn)) Deallocate(blk%posn)
               If (Allocated(blk%forces)) Deallocate(blk%forces)
               If (Allocated(blk%ltype)) Deallocate(blk%ltype)
               Allocate(blk%posn(3, ntot(b)), blk%forces(3, ntot(b)), &
                    blk%ltype(ntot(b)))
               p = first_particle(b)
               Do m = 1, ntot(b)
                  blk%posn(:, m) = particle_positions(:, p)
                  p = next_particle(p)
               End Do
               blk%forces = zero
               blk%ltype = 1
             End Associate
          End Do
       End Do
    End Do
  End Subroutine exaMD_partition_work
End Module exaMD_partitioning


Module exaMD_comms
  !! The MPI halo of SRC/exaMD_comms.f90 is a ghost gather inside the force
  !! kernel on one GPU; the routines keep their call sites.
  Implicit None
Contains
  Subroutine exaMD_start_comms_posn()
  End Subroutine exaMD_start_comms_posn
  Logical Function exaMD_test_comms_posn(noprocs)
    Integer, Intent(in) :: noprocs
    exaMD_test_comms_posn = .True.
  End Function exaMD_test_comms_posn
  Subroutine exaMD_wait_comms_posn(noprocs)
    Integer, Intent(in) :: noprocs
  End Subroutine exaMD_wait_comms_posn
  Subroutine exaMD_recv_comms_posn(noprocs)
    Integer, Intent(in) :: noprocs
  End Subroutine exaMD_recv_comms_posn
  Subroutine exaMD_start_comms_forces(noprocs)
    Integer, Intent(in) :: noprocs
  End Subroutine exaMD_start_comms_forces
  Subroutine exaMD_wait_comms_forces(noprocs)
    Integer, Intent(in) :: noprocs
  End Subroutine exaMD_wait_comms_forces
End Module exaMD_comms


Module exaMD_comms_overlap
  !! SRC/exaMD_comms_overlap.f90:38-271: the partial receives and sends the
  !! overlapped force routine uses.  As at one rank in the reference, every
  !! message has arrived at once (the ghost gather and force return run
  !! inside the force kernel); the routines keep their call sites.
  Implicit None
Contains
  Integer Function exaMD_testsome_comms_posn(arrived, justin, noprocs)
    Integer, Allocatable, Intent(inout) :: arrived(:), justin(:)
    Integer, Intent(in) :: noprocs
    exaMD_testsome_comms_posn = 1
  End Function exaMD_testsome_comms_posn
  Subroutine exaMD_recvsome_comms_posn(justin, n_justin, noprocs)
    Integer, Allocatable, Intent(in) :: justin(:)
    Integer, Intent(in) :: n_justin, noprocs
  End Subroutine exaMD_recvsome_comms_posn
  Subroutine exaMD_waitrest_comms_posn(arrived, justin, noprocs)
    Integer, Allocatable :: arrived(:), justin(:)
    Integer, Intent(in) :: noprocs
  End Subroutine exaMD_waitrest_comms_posn
  Subroutine exaMD_start_comms_forcessome(sendcnt, sendlist, noprocs)
    Integer, Intent(in) :: sendcnt, noprocs
    Integer, Allocatable, Intent(in) :: sendlist(:)
  End Subroutine exaMD_start_comms_forcessome
End Module exaMD_comms_overlap


Module exaMD_forces_mod
  !! exaMD_forces_midpoint_all / _overlap (SRC/exaMD_forces.f90:41, :308)
  !! on the GPU.
  Use, Intrinsic :: iso_c_binding
  Use exaMD_mod
  Use exaMD_hip_binding
  Use user_func_mod, Only: prmvdw, kktype, ltpvdw
  Implicit None

  !! Arithmetic of the force routines: EXAMD_PREC_FP64 (default, the
  !! reference's) or EXAMD_PREC_MIXED.
  Integer(c_int) :: exaMD_precision = EXAMD_PREC_FP64
  !! Statistics of the last evaluation (pair counts, energy).
  Type(examd_stats) :: exaMD_last_stats
  !! A user function that is not the Lennard-Jones pair is sampled into a
  !! table of exaMD_table_points points uniform in r^2 from
  !! (exaMD_table_rmin * cutoff)^2 to cutoff^2 and interpolated on the GPU.
  Integer :: exaMD_table_points = 4096
  Real(kind=PRC) :: exaMD_table_rmin = 0.2_PRC
  !! exaMD_md_run with exaMD_rebalance_every > 0: the cell blocks get
  !! boundaries at equal shares of the atoms (the atom-count weights of
  !! SRC/exaMD_partitioning.f90:199-203), re-partitioned from the current
  !! positions at least every this many steps (DOC/exaMD.tex:284-291).
  Integer :: exaMD_rebalance_every = 0

  !! Force calls that found grid%posn unchanged since the positions last
  !! sent to the device, and so did not send them again.
  Integer :: exaMD_positions_resent_skipped = 0

  Type(c_ptr), Private :: handle = c_null_ptr
  Integer(c_int64_t), Private :: handle_n = -1
  !! The positions the device holds (locblocks order), while they are known.
  Real(c_double), Allocatable, Private :: sent_pos(:,:)
  Logical, Private :: handle_species = .False.
  Logical, Private :: handle_table = .False.

Contains

  Subroutine exaMD_release_device()
    Integer(c_int) :: st
    If (c_associated(handle)) st = examd_hip_destroy(handle)
    handle = c_null_ptr
    handle_n = -1
    handle_species = .False.
    handle_table = .False.
    If (Allocated(sent_pos)) Deallocate(sent_pos)
  End Subroutine exaMD_release_device

  !! Send the gathered positions unless the device already holds exactly
  !! these (the same call repeated on unchanged grid%posn).
  Subroutine send_positions(pos, n)
    Real(c_double), Intent(in), Target :: pos(:,:)
    Integer(c_int64_t), Intent(in) :: n
    Integer(c_int) :: st
    If (Allocated(sent_pos)) Then
       If (Size(sent_pos, 2) == Size(pos, 2)) Then
          If (All(sent_pos == pos)) Then
             exaMD_positions_resent_skipped = exaMD_positions_resent_skipped + 1
             Return
          End If
       End If
       Deallocate(sent_pos)
    End If
    st = examd_hip_set_state(handle, pos, c_null_ptr, n)
    Call check_status(st, 'examd_hip_set_state')
    Allocate(sent_pos, source=pos)
  End Subroutine send_positions

  !! Is user_func the Lennard-Jones pair with prmvdw(1,:)'s parameters?
  !! (checked at a few distances; the device evaluates that pair directly)
  Logical Function user_func_is_lj(user_func, cutoff)
    External :: user_func
    Real(kind=PRC), Intent(in) :: cutoff
    Real(kind=PRC) :: r(4), f12(3), e12, s6, want, d2
    Integer :: i, retval
    user_func_is_lj = .False.
    r = (/ 0.95_PRC, 1.12_PRC, 1.7_PRC, 0.99_PRC * cutoff /) * prmvdw(1, 2)
    Do i = 1, 4
       If (r(i) > cutoff) Cycle
       d2 = r(i) * r(i)
       Call user_func(r(i), zero, zero, d2, f12, e12, retval)
       s6 = (prmvdw(1, 2)**2 / d2)**3
       want = -twentyfour * prmvdw(1, 1) * (two * s6 * s6 - s6) / d2 * r(i)
       If (retval /= 0 .Or. Abs(f12(1) - want) > 1.0e-9_PRC * Max(one, &
            Abs(want))) Return
    End Do
    d2 = (1.01_PRC * cutoff)**2
    Call user_func(1.01_PRC * cutoff, zero, zero, d2, f12, e12, retval)
    If (retval <= 0) Then
       Write (*, *) 'exaMD: user_func must skip pairs beyond the cutoff'
       Call exaMD_abort(101)
    End If
    user_func_is_lj = .True.
  End Function user_func_is_lj

  !! The pair function on the device: the Lennard-Jones pair, or any other
  !! user function sampled into a table (SRC/exaMD_forces.f90:532-549: g =
  !! f12(1) / r along x, retval = 1 -> no update, retval < 0 -> error).
  Subroutine set_device_pair(user_func, cutoff)
    External :: user_func
    Real(kind=PRC), Intent(in) :: cutoff
    Real(c_double), Allocatable, Target :: g(:), v(:)
    Real(kind=PRC) :: r2lo, r2, r, f12(3), e12
    Integer :: k, n, retval
    Integer(c_int) :: st
    If (user_func_is_lj(user_func, cutoff)) Then
       If (handle_table) Then
          st = examd_hip_set_pair_table(handle, 0_c_int32_t, 0.0_c_double, &
               c_null_ptr, c_null_ptr)
          Call check_status(st, 'examd_hip_set_pair_table')
          handle_table = .False.
       End If
       Return
    End If
    n = exaMD_table_points
    r2lo = (exaMD_table_rmin * cutoff)**2
    Allocate(g(n), v(n))
    Do k = 1, n
       r2 = r2lo + (cutoff * cutoff - r2lo) * Real(k - 1, PRC) / Real(n - 1, PRC)
       r = Sqrt(r2)
       Call user_func(r, zero, zero, r2, f12, e12, retval)
       If (retval < 0) Then
          Write (*, *) 'Error from user_func: ', retval
          Call exaMD_abort(101)
       Else If (retval > 0) Then
          g(k) = 0.0_c_double
          v(k) = 0.0_c_double
       Else
          g(k) = Real(f12(1) / r, c_double)
          v(k) = Real(e12, c_double)
       End If
    End Do
    st = examd_hip_set_pair_table(handle, Int(n, c_int32_t), Real(r2lo, c_double), &
         c_loc(g), c_loc(v))
    Call check_status(st, 'examd_hip_set_pair_table')
    handle_table = .True.
  End Subroutine set_device_pair

  !! Flatten grid%posn (locblocks order) into pos(3, n).
  Subroutine gather_positions(pos, n)
    Real(c_double), Allocatable, Intent(out) :: pos(:,:)
    Integer(c_int64_t), Intent(out) :: n
    Integer :: b, m, i, j, k
    n = 0
    Do b = 1, nblocks_loc
       n = n + grid(locblocks(1, b), locblocks(2, b), locblocks(3, b))%natms
    End Do
    Allocate(pos(3, Max(n, 1_c_int64_t)))
    m = 0
    Do b = 1, nblocks_loc
       i = locblocks(1, b)
       j = locblocks(2, b)
       k = locblocks(3, b)
       If (grid(i, j, k)%natms > 0) Then
          pos(:, m + 1:m + grid(i, j, k)%natms) = grid(i, j, k)%posn
          m = m + grid(i, j, k)%natms
       End If
    End Do
  End Subroutine gather_positions

  Subroutine ensure_handle(n, cutoff)
    Integer(c_int64_t), Intent(in) :: n
    Real(kind=PRC), Intent(in) :: cutoff
    Type(examd_params) :: p
    Integer(c_int) :: st
    If (c_associated(handle) .And. handle_n == n) Return
    Call exaMD_release_device()
    p%natoms = n
    p%box = Real(globalbox, c_double)
    p%cutoff = Real(cutoff, c_double)
    p%epsilon = Real(prmvdw(1, 1), c_double)
    p%sigma = Real(prmvdw(1, 2), c_double)
    p%mass = 1.0_c_double
    p%tiles = 0
    p%device = 0
    p%block_threads = 0
    p%cap_slack_pct = 0
    p%run_groups = 0
    p%graph_chunk = 100
    p%reserved = 0
    st = examd_hip_create(p, handle)
    Call check_status(st, 'examd_hip_create')
    handle_n = n
  End Subroutine ensure_handle

  !! Several species: grid%ltype (species of each atom, 1-based), kktype
  !! (species pair -> row of prmvdw), ltpvdw (kind of each row: 2, the
  !! Lennard-Jones pair, is the only one the device evaluates) and prmvdw
  !! (row k: eps, sigma, -, -, cutoff), as exaMD.f90:199-210 sets them up for
  !! one species.  With more than one species the device evaluates each pair
  !! with its own row; with one, the handle's parameters (prmvdw(1,:)).
  Subroutine set_device_species(n)
    Integer(c_int64_t), Intent(in) :: n
    Integer(c_int32_t), Allocatable, Target :: ty(:)
    Real(c_double), Allocatable, Target :: e(:,:), s(:,:), c(:,:)
    Integer :: nt, a, b, kr, m, bl, i, j, k, na
    Integer(c_int) :: st
    nt = 0
    If (Allocated(kktype)) nt = Size(kktype, 1)
    If (nt <= 1) Then
       If (handle_species) Then
          st = examd_hip_set_species(handle, 0_c_int32_t, c_null_ptr, n, &
               c_null_ptr, c_null_ptr, c_null_ptr)
          Call check_status(st, 'examd_hip_set_species')
          handle_species = .False.
       End If
       Return
    End If
    Allocate(e(nt, nt), s(nt, nt), c(nt, nt), ty(Max(n, 1_c_int64_t)))
    Do a = 1, nt
       Do b = 1, nt
          kr = kktype(a, b)
          If (ltpvdw(kr) /= 2) Then
             Write (*, *) 'exaMD: the GPU path evaluates only the ', &
                  'Lennard-Jones pair (ltpvdw = 2)'
             Call exaMD_abort(101)
          End If
          ! row-major [a][b] on the C side is e(b, a) here
          e(b, a) = Real(prmvdw(kr, 1), c_double)
          s(b, a) = Real(prmvdw(kr, 2), c_double)
          c(b, a) = Real(prmvdw(kr, 5), c_double)
       End Do
    End Do
    m = 0
    Do bl = 1, nblocks_loc
       i = locblocks(1, bl)
       j = locblocks(2, bl)
       k = locblocks(3, bl)
       na = grid(i, j, k)%natms
       If (na > 0) Then
          ty(m + 1:m + na) = Int(grid(i, j, k)%ltype(1:na) - 1, c_int32_t)
          m = m + na
       End If
    End Do
    st = examd_hip_set_species(handle, Int(nt, c_int32_t), c_loc(ty), n, &
         c_loc(e), c_loc(s), c_loc(c))
    Call check_status(st, 'examd_hip_set_species')
    handle_species = .True.
  End Subroutine set_device_species

  Subroutine check_status(st, where)
    Integer(c_int), Intent(in) :: st
    Character(len=*), Intent(in) :: where
    If (st == EXAMD_OK) Return
    Write (*, '(1X,A,A,A,A,I0,A)') where, ': ', &
         Trim(examd_status_text(st)), ' (status ', st, ')'
    If (st == EXAMD_ERR_ZERO_DISTANCE) Return     ! the reference only prints
    If (st == EXAMD_ERR_OUT_OF_BOX) Call exaMD_abort(109)
    Call exaMD_abort(102)
  End Subroutine check_status

  !! SRC/exaMD_forces.f90:308: forces of every pair the 63-row table assigns
  !! to a home block, added into grid%forces.
  Subroutine exaMD_forces_midpoint_all(user_func, cutoff, noprocs)
    External :: user_func
    Real(kind=PRC), Intent(in) :: cutoff
    Integer, Intent(in) :: noprocs
    Real(c_double), Allocatable :: pos(:,:), f(:,:)
    Integer(c_int64_t) :: n
    Integer(c_int) :: st
    Integer :: b, m, i, j, k, na
    Call gather_positions(pos, n)
    Call ensure_handle(n, cutoff)
    Call set_device_pair(user_func, cutoff)
    Call set_device_species(n)
    Call send_positions(pos, n)
    Allocate(f(3, Max(n, 1_c_int64_t)))
    st = examd_hip_forces(handle, exaMD_precision, f, exaMD_last_stats)
    Call check_status(st, 'examd_hip_forces')
    m = 0
    Do b = 1, nblocks_loc
       i = locblocks(1, b)
       j = locblocks(2, b)
       k = locblocks(3, b)
       na = grid(i, j, k)%natms
       If (na > 0) Then
          grid(i, j, k)%forces = grid(i, j, k)%forces + f(:, m + 1:m + na)
          m = m + na
       End If
    End Do
  End Subroutine exaMD_forces_midpoint_all

  !! SRC/exaMD_forces.f90:41: on one GPU the overlap of halo messages with
  !! interior work is stream order, so this is the same evaluation.
  Subroutine exaMD_forces_midpoint_overlap(user_func, cutoff, pid, noprocs)
    External :: user_func
    Real(kind=PRC), Intent(in) :: cutoff
    Integer, Intent(in) :: pid, noprocs
    Call exaMD_forces_midpoint_all(user_func, cutoff, noprocs)
  End Subroutine exaMD_forces_midpoint_overlap

  !! Velocity-Verlet MD from the grid's positions and the given velocities
  !! (same locblocks order): nsteps of dt, energies every energy_every steps
  !! written to stats(1:nstats).  Returns the final state in pos/vel.
  Subroutine exaMD_md_run(cutoff, vel, nsteps, dt, energy_every, stats, &
       nstats, pos_out, vel_out, force_ms)
    Real(kind=PRC), Intent(in) :: cutoff, dt
    Real(c_double), Intent(in), Target :: vel(:,:)
    Integer(c_int64_t), Intent(in) :: nsteps, energy_every
    Type(examd_stats), Intent(out) :: stats(:)
    Integer(c_int64_t), Intent(out) :: nstats
    Real(c_double), Intent(out), Target :: pos_out(:,:), vel_out(:,:)
    Real(c_double), Intent(out) :: force_ms
    Real(c_double), Allocatable :: pos(:,:)
    Integer(c_int64_t) :: n, launches
    Integer(c_int) :: st
    Type(examd_tune) :: cfg
    Call gather_positions(pos, n)
    Call ensure_handle(n, cutoff)
    Call set_device_species(n)
    ! the run moves the device's positions
    If (Allocated(sent_pos)) Deallocate(sent_pos)
    st = examd_hip_set_state(handle, pos, c_loc(vel), n)
    Call check_status(st, 'examd_hip_set_state')
    If (exaMD_rebalance_every > 0) Then
       st = examd_hip_get_config(handle, exaMD_precision, cfg)
       Call check_status(st, 'examd_hip_get_config')
       cfg%balance = 1
       cfg%rebalance_every = Int(exaMD_rebalance_every, c_int32_t)
       st = examd_hip_set_config(handle, exaMD_precision, cfg)
       Call check_status(st, 'examd_hip_set_config')
    End If
    st = examd_hip_timing(handle, 1_c_int, force_ms, launches, 1_c_int)
    st = examd_hip_run(handle, exaMD_precision, nsteps, Real(dt, c_double), &
         energy_every, stats, Int(Size(stats), c_int64_t), nstats)
    Call check_status(st, 'examd_hip_run')
    st = examd_hip_timing(handle, 0_c_int, force_ms, launches, 1_c_int)
    st = examd_hip_get_state(handle, c_loc(pos_out), c_loc(vel_out), &
         c_null_ptr, n)
    Call check_status(st, 'examd_hip_get_state')
  End Subroutine exaMD_md_run

End Module exaMD_forces_mod


Module exaMD_utils_mod
  Use exaMD_mod
  Implicit None
Contains
  !! SRC/exaMD_utils.f90:38: one process, one GPU.
  Subroutine exaMD_Init(pid, noprocs, input_comm)
    Integer, Intent(out) :: pid, noprocs
    Integer, Intent(in) :: input_comm
    COMM_EXAMD = input_comm
    pid = 0
    noprocs = 1
  End Subroutine exaMD_Init

  !! SRC/exaMD_utils.f90:167-190 through examd_hip_domain.
  Subroutine exaMD_domain_create(block_length, nblocks, nblocks_xyz)
    Use exaMD_hip_binding
    Real(kind=PRC), Intent(in) :: block_length
    Integer, Intent(out) :: nblocks, nblocks_xyz(3)
    Integer(c_int32_t) :: nb(3)
    Integer(c_int64_t) :: n
    Integer(c_int) :: st
    Real(c_double) :: box(3)
    box = Real(globalbox, c_double)
    st = examd_hip_domain(box, Real(block_length, c_double), nb, n)
    If (st /= EXAMD_OK) Call exaMD_abort(107)
    nblocks_xyz = Int(nb)
    nblocks = Int(n)
    If (Allocated(grid)) Deallocate(grid)
    Allocate(grid(nblocks_xyz(1), nblocks_xyz(2), nblocks_xyz(3)))
    globalboxi = one / globalbox
  End Subroutine exaMD_domain_create

  !! SRC/exaMD_utils.f90:90: release the device handle and the grid.
  Subroutine exaMD_Finalize(text)
    Use exaMD_forces_mod, Only: exaMD_release_device
    Character(len=*), Intent(in) :: text
    Call exaMD_release_device()
    If (Allocated(grid)) Deallocate(grid)
    If (Allocated(interactions)) Deallocate(interactions)
    If (Allocated(locblocks)) Deallocate(locblocks)
    If (Allocated(interior_blks)) Deallocate(interior_blks)
    Write (*, '(1X,A)') Trim(text)
  End Subroutine exaMD_Finalize
End Module exaMD_utils_mod
