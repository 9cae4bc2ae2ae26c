! Fortran MD loop on the device NVE driver of libccpol3b.so (ccpol3b_f module):
! the rigid-water velocity-Verlet + NO_SQUISH step of the reference's nve_1
! (evolve.f90:621-1062, 8107-8204) with the three-body forces, run from Fortran.
!
!   drive_md IN OUT
!
! IN (text): nm box / alpha3 delta33 rcut3 s2 s3 precision / SR path / LR path /
!            nm lines "x y z" (A) / nm lines "q0 q1 q2 q3" / nm lines "vx vy vz" (A/ps) /
!            nm lines "jx jy jz" (body frame, amu A^2/ps) / ns / ns lines "px py pz" (A) /
!            mass (amu) / pmi(1:3) (amu A^2) / nsteps dt (ps)
! OUT (text): nsteps lines "E3 Kt Kr E3+Kt+Kr" (kcal/mol), then nm lines each of
!             com, quat, vel, jbody after the run
program drive_md
  use, intrinsic :: iso_c_binding
  use ccpol3b_f
  implicit none
  character(len=4096) :: fin, fout, psr, plr
  integer :: nm, s2, s3, prec, i, ns, nsteps, rc
  real(8) :: box, alpha3, delta33, rcut3, mass, dt
  real(c_double), target :: bsr(17, 17), gsr(17, 17), rsr(17, 17), blr(17, 17), glr(17, 17), rlr(17, 17)
  real(c_double), target :: csr(364), clr(364)
  real(c_double) :: pmi(3)
  integer(c_int) :: nl
  real(8), allocatable :: com(:, :), quat(:, :), vel(:, :), jb(:, :), pc(:, :), rec(:, :)
  type(ccpol3b_params), target :: p
  type(c_ptr) :: ctx, md

  call get_command_argument(1, fin)
  call get_command_argument(2, fout)
  open (10, file=trim(fin), status='old', action='read')
  read (10, *) nm, box
  read (10, *) alpha3, delta33, rcut3, s2, s3, prec
  read (10, '(a)') psr
  read (10, '(a)') plr
  nl = 364
  rc = ccpol3b_read_data3b(trim(psr)//c_null_char, bsr, gsr, rsr, 17_c_int, csr, nl)
  if (rc /= 0) stop 'read SR'
  nl = 364
  rc = ccpol3b_read_data3b(trim(plr)//c_null_char, blr, glr, rlr, 17_c_int, clr, nl)
  if (rc /= 0) stop 'read LR'
  p%beta_sr = c_loc(bsr); p%gamma_sr = c_loc(gsr); p%r0_sr = c_loc(rsr)
  p%beta_lr = c_loc(blr); p%gamma_lr = c_loc(glr); p%r0_lr = c_loc(rlr)
  p%ld_pair = 17
  p%cs_sr = c_loc(csr); p%cs_lr = c_loc(clr)
  p%nlin = 364
  p%alpha3 = alpha3; p%delta33 = delta33; p%rcut3 = rcut3
  p%s2_term = s2; p%s3_term = s3; p%precision = prec; p%device = -1

  allocate (com(3, nm), quat(4, nm), vel(3, nm), jb(3, nm))
  do i = 1, nm
    read (10, *) com(:, i)
  end do
  do i = 1, nm
    read (10, *) quat(:, i)
  end do
  do i = 1, nm
    read (10, *) vel(:, i)
  end do
  do i = 1, nm
    read (10, *) jb(:, i)
  end do
  read (10, *) ns
  allocate (pc(3, ns))
  do i = 1, ns
    read (10, *) pc(:, i)
  end do
  read (10, *) mass
  read (10, *) pmi
  read (10, *) nsteps, dt
  close (10)
  allocate (rec(4, nsteps))

  rc = ccpol3b_create(p, ctx)
  if (rc /= 0) then
    print '(a,i0)', 'create failed ', rc
    stop 2
  end if
  rc = ccpol3b_md_create(ctx, int(nm, c_int), real(box, c_double), int(ns, c_int), pc, real(mass, c_double), pmi, md)
  if (rc /= 0) then
    print '(a,i0)', 'md_create failed ', rc
    stop 3
  end if
  rc = ccpol3b_md_set_state(md, com, quat, vel, jb)
  if (rc == 0) rc = ccpol3b_md_run(md, int(nsteps, c_int), real(dt, c_double), rec)
  if (rc == 0) rc = ccpol3b_md_get_state(md, com, quat, vel, jb)
  if (rc /= 0) then
    print '(a,i0)', 'md failed ', rc
    stop 4
  end if
  call ccpol3b_md_destroy(md)
  call ccpol3b_destroy(ctx)

  open (11, file=trim(fout), status='replace', action='write')
  do i = 1, nsteps
    write (11, '(4es25.16)') rec(:, i)
  end do
  do i = 1, nm
    write (11, '(3es25.16)') com(:, i)
  end do
  do i = 1, nm
    write (11, '(4es25.16)') quat(:, i)
  end do
  do i = 1, nm
    write (11, '(3es25.16)') vel(:, i)
  end do
  do i = 1, nm
    write (11, '(3es25.16)') jb(:, i)
  end do
  close (11)
  print '(a)', 'ok'
end program drive_md
