! qprobe -- golden-vector harness for the oracle (test infrastructure only).
!
! This program is this project's own driver.  It links the reference Q6
! modules compiled unmodified from src/q6 (sizes misc nrgy mpiglob prmfile
! index topo mask trj qatom md) and runs the same initialisation sequence as
! the reference main program (src/q6/qdyn.f90:83-121), then:
!   1. builds the pair lists and evaluates pot_energy at step 0
!      (md.f90:3716 make_pair_lists, md.f90:13278 pot_energy) and dumps every
!      energy term, the gradient d(:), the coordinates/velocities and all
!      pair lists at full precision;
!   2. runs md_run for the steps of the input file (md.f90:3917) and dumps
!      the final x, v, energies and gradient.
! The dump goes to the file named by the environment variable QPROBE_OUT
! (default qprobe.dump).  The input file is the last command-line argument,
! as for qdyn (md.f90:2568-2577).
program qprobe
  use md
  implicit none
  integer :: u, i, is, ip, iw, nrow
  real(8) :: Temp, Tsolute, Tsolvent, Ekinmax
  character(len=256) :: outname

  nodeid = 0
  numnodes = 1
  call md_startup

  if (.not. initialize()) call die('Invalid data in input file')
  call open_files
  call topology
  call prep_coord
  if (nstates > 0) call get_fep
  call prep_sim
  call close_input_files
  call init_shake
  call make_nbqqlist
  call shrink_topology
  if (iseed > 0) then
    call maxwell
    call initial_shaking
    call stop_cm_translation
  end if
  call distribute_nonbonds

  call get_environment_variable('QPROBE_OUT', outname)
  if (len_trim(outname) == 0) outname = 'qprobe.dump'
  u = 77
  open(unit=u, file=trim(outname), status='replace', action='write')

  write(u,'(a)') '# qprobe dump'
  write(u,'(a,9i12)') 'SIZES', natom, nat_solute, nwat, nstates, nqat, ncgp, ncgp_solute, &
       shake_constraints, shake_molecules
  write(u,'(a,6i12)') 'NDEGF', Ndegf, Ndegfree, Ndegf_solute, Ndegfree_solute, &
       Ndegf_solvent, Ndegfree_solvent
  write(u,'(a,3es25.16e3)') 'DT', dt, tau_T, Temp0

  ! initial state (after maxwell / initial shaking / cm removal)
  call dump_vec('X0', x, 3*natom)
  call dump_vec('V0', v, 3*natom)

  ! step 0 as md_run does it: temperature first (md.f90:3969-3976, it sets
  ! Tfree used by restrain_solvent), then pair lists and potential energy
  istep = 0
  Ekinmax = 1000.0*Ndegf*boltz*Temp0/2.0/real(natom)
  call temperature(Temp, Tsolute, Tsolvent, Ekinmax)
  write(u,'(a,5es25.16e3)') 'TEMP0', Temp, Tfree, Tsolute, Tsolvent, E%kinetic
  call make_pair_lists
  call pot_energy
  call dump_energies('E0')
  call dump_vec('D0', d, 3*natom)

  write(u,'(a,i12)') 'NBPP', nbpp_pair
  do ip = 1, nbpp_pair
    write(u,'(3i10)') nbpp(ip)%i, nbpp(ip)%j, nbpp(ip)%LJcod
  end do
  write(u,'(a,i12)') 'NBPW', nbpw_pair
  do ip = 1, nbpw_pair
    write(u,'(3i10)') nbpw(ip)%i, nbpw(ip)%j, nbpw(ip)%LJcod
  end do
  ! water-water list: rows of O atom numbers terminated by 0, one row per
  ! water molecule in calculation_assignment order
  nrow = 0
  do ip = 1, nbww_pair
    if (nbww(ip) /= 0) nrow = nrow + 1
  end do
  write(u,'(a,i12)') 'NBWW', nrow
  iw = calculation_assignment%ww%start
  do ip = 1, nbww_pair
    if (nbww(ip) == 0) then
      iw = iw + 1
    else
      write(u,'(2i10)') nat_solute + 3*iw - 2, nbww(ip)
    end if
  end do
  write(u,'(a,i12)') 'NBQP', nbqp_pair
  do ip = 1, nbqp_pair
    write(u,'(4i10)') nbqp(ip)%i, nbqp(ip)%j, nbqp(ip)%LJcod, nbqp(ip)%qLJcod
  end do
  write(u,'(a,i12)') 'NBQW', nbqw_pair
  do ip = 1, nbqw_pair
    write(u,'(i10)') nbqw(ip)
  end do
  do is = 1, nstates
    write(u,'(a,2i12)') 'NBQQ', is, nbqq_pair(is)
    do ip = 1, nbqq_pair(is)
      write(u,'(4i10,es25.16e3)') nbqq(ip,is)%iq, nbqq(ip,is)%j, nbqq(ip,is)%jq, &
           nbqq(ip,is)%LJcod, nbqq(ip,is)%el_scale
    end do
  end do
  if (use_LRF) then
    write(u,'(a,i12)') 'LRF', ncgp
    do i = 1, ncgp
      write(u,'(i8,43es25.16e3)') i, lrf(i)%cgp_cent, lrf(i)%phi0, lrf(i)%phi1, &
           lrf(i)%phi2, lrf(i)%phi3
    end do
  end if

  ! full run as qdyn does it; afterwards x, v, E, EQ, d hold the final state
  call md_run
  call dump_vec('XN', x, 3*natom)
  call dump_vec('VN', v, 3*natom)
  call dump_energies('EN')
  call dump_vec('DN', d, 3*natom)
  write(u,'(a)') 'END'
  close(u)
  call close_output_files

contains

  subroutine dump_vec(tag, a, n)
    character(*), intent(in) :: tag
    integer, intent(in) :: n
    real(8), intent(in) :: a(:)
    integer :: k
    write(u,'(a,i12)') tag, n
    do k = 1, n, 3
      write(u,'(3es25.16e3)') a(k:min(k+2,n))
    end do
  end subroutine dump_vec

  subroutine dump_energies(tag)
    character(*), intent(in) :: tag
    integer :: s
    write(u,'(a)') tag
    write(u,'(a,3es25.16e3)') 'pot_kin_lrf', E%potential, E%kinetic, E%LRF
    write(u,'(a,4es25.16e3)') 'p', E%p%bond, E%p%angle, E%p%torsion, E%p%improper
    write(u,'(a,4es25.16e3)') 'w', E%w%bond, E%w%angle, E%w%torsion, E%w%improper
    write(u,'(a,4es25.16e3)') 'q', E%q%bond, E%q%angle, E%q%torsion, E%q%improper
    write(u,'(a,2es25.16e3)') 'pp', E%pp%el, E%pp%vdw
    write(u,'(a,2es25.16e3)') 'pw', E%pw%el, E%pw%vdw
    write(u,'(a,2es25.16e3)') 'ww', E%ww%el, E%ww%vdw
    write(u,'(a,2es25.16e3)') 'qx', E%qx%el, E%qx%vdw
    write(u,'(a,6es25.16e3)') 'rst', E%restraint%total, E%restraint%fix, &
         E%restraint%shell, E%restraint%protein, E%restraint%solvent_radial, &
         E%restraint%water_pol
    do s = 1, nstates
      write(u,'(a,i4,15es25.16e3)') 'EQ', s, EQ(s)%lambda, EQ(s)%total, &
           EQ(s)%q%bond, EQ(s)%q%angle, EQ(s)%q%torsion, EQ(s)%q%improper, &
           EQ(s)%qx%el, EQ(s)%qx%vdw, EQ(s)%qq%el, EQ(s)%qq%vdw, &
           EQ(s)%qp%el, EQ(s)%qp%vdw, EQ(s)%qw%el, EQ(s)%qw%vdw, EQ(s)%restraint
    end do
  end subroutine dump_energies

end program qprobe
